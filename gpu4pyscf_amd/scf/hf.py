'''Restricted Hartree-Fock driver on the MI355X J/K engine.

Follows the structure of gpu4pyscf/scf/hf.py:206-391 (``_kernel``): the
Fock matrix is built incrementally on the density change (get_veff,
gpu4pyscf/scf/hf.py:91-109), CDIIS extrapolation starts at cycle 1 with 8
vectors (gpu4pyscf/scf/diis.py:41-127), the generalised eigenproblem is
solved in the canonically orthogonalised basis with overlap eigenvalues below
1e-6 removed (check_linear_dependency, gpu4pyscf/scf/hf.py:625-641), and
convergence requires |dE| < conv_tol and |g| < sqrt(conv_tol)
(gpu4pyscf/scf/hf.py:309-312).  Everything is FP64 on the device.

The initial guess is 'atom' (superposition of spherically averaged atomic
RHF densities, the analogue of init_guess_by_atom, gpu4pyscf/scf/hf.py:
561-607), '1e' (core Hamiltonian), a user density, or a callable (used by the
learned guess).  The ANO-based 'minao' guess of PySCF needs basis data that is
not part of this project.
'''

import ctypes
import math
import time

import numpy as np

from gpu4pyscf_amd.lib import clib
from gpu4pyscf_amd.scf import jk as jk_mod

LINDEP_THRESHOLD = 1e-6


def _torch():
    import torch
    return torch


def int1e(mol, vhfopt):
    '''(S, T, V) in the original AO basis of ``mol`` as device tensors.'''
    torch = _torch()
    lib = clib.load()
    if vhfopt._dev is None:
        vhfopt.build()
    sm = vhfopt.sorted_mol
    d = vhfopt._dev
    dev = vhfopt.device
    n = sm.nao
    S = torch.empty((n, n), dtype=torch.float64, device=dev)
    T = torch.empty_like(S)
    V = torch.empty_like(S)
    coords = torch.as_tensor(np.asarray(mol.atom_coords(), dtype=np.float64), device=dev)
    charges = torch.as_tensor(np.asarray(mol.atom_charges(), dtype=np.float64), device=dev)
    st = lib.jk_int1e(S.data_ptr(), T.data_ptr(), V.data_ptr(), d['bas'].data_ptr(),
                      d['env'].data_ptr(), d['ao_loc'].data_ptr(), sm.nbas, n,
                      coords.data_ptr(), charges.data_ptr(), mol.natm, vhfopt._rys,
                      jk_mod._stream_ptr())
    clib.check(st, 'jk_int1e')
    C = d['coeff']
    return C.T @ S @ C, C.T @ T @ C, C.T @ V @ C


class CDIIS:
    '''Commutator DIIS on the error FDS - SDF in the orthonormal basis.
    The overlap matrix of the stored error vectors is kept incrementally:
    one update adds the dot products of the new vector with the stored ones
    (elementwise products and reductions on the device).'''

    def __init__(self, space=8):
        self.space = space
        self.fs = []
        self.es = []
        self.B = np.zeros((0, 0))

    def update(self, f, e):
        torch = _torch()
        e = e.reshape(-1)
        if len(self.fs) == self.space:
            self.fs.pop(0)
            self.es.pop(0)
            self.B = self.B[1:, 1:]
        dots = torch.stack([torch.sum(x * e) for x in self.es] + [torch.sum(e * e)])
        dots = dots.cpu().numpy()
        n = len(self.es) + 1
        B = np.empty((n, n))
        B[:n - 1, :n - 1] = self.B
        B[n - 1, :] = B[:, n - 1] = dots
        self.B = B
        self.fs.append(f)
        self.es.append(e)
        # the (n+1) x (n+1) system is solved on the host
        A = np.empty((n + 1, n + 1))
        A[:n, :n] = B
        A[n, :n] = A[:n, n] = -1.
        A[n, n] = 0.
        rhs = np.zeros(n + 1)
        rhs[n] = -1.
        try:
            c = np.linalg.solve(A, rhs)
        except np.linalg.LinAlgError:
            c = np.linalg.lstsq(A, rhs, rcond=None)[0]
        return sum(float(c[i]) * self.fs[i] for i in range(n))


class RHF:
    '''Closed-shell RHF (gpu4pyscf.scf.RHF surface: kernel, e_tot, converged,
    mo_energy, mo_coeff, mo_occ, make_rdm1, get_jk, get_veff, cycles).'''

    conv_tol = 1e-9
    conv_tol_grad = None
    max_cycle = 50
    diis_space = 8
    diis_start_cycle = 1
    direct_scf_tol = 1e-13
    init_guess = 'atom'
    # mixed precision: quartets with sqrt|(ij|ij)(kl|kl)| max|dD| below
    # fp32_tol run in FP32 (None: plain FP64); the converged energy is then
    # re-evaluated from one FP64 build of the full density (final_fp64)
    fp32_tol = None
    final_fp64 = True
    # a callable (learned) guess is projected on the closest idempotent
    # density with the right electron count before the first Fock build
    project_guess = True
    # accumulate per-phase wall times (jk, eig, diis) in self.timings; the
    # device synchronisations this needs cost a little time
    time_phases = False
    # J and K of every build: one J/K build screened on all six density
    # blocks (False), or, as the reference's RHF does
    # (gpu4pyscf/scf/hf.py:91-109), a J-only build and a K-only build, each
    # screened on its own blocks (True)
    split_jk = False
    # with split_jk, J from the McMurchie-Davidson J-engine
    # (gpu4pyscf/scf/hf.py:947-958) instead of the Rys J-only build
    use_j_engine = False

    def __init__(self, mol):
        self.mol = mol
        self.verbose = getattr(mol, 'verbose', 0)
        self._opt = None
        self.converged = False
        self.e_tot = 0.
        self.cycles = 0
        self.mo_energy = self.mo_coeff = self.mo_occ = None
        self.jk_builds = 0
        self.quartets = 0
        self.quartets_fp32 = 0
        self._fp32_now = None
        self.timings = {}
        self.history = []           # energy after the initial build and each cycle
        self.history_t = []         # seconds since kernel() started, per entry
        self.jk_hook = None
        self._jopt = None

    # -- integrals -------------------------------------------------------
    @property
    def opt(self):
        if self._opt is None:
            self._opt = jk_mod._VHFOpt(self.mol, self.direct_scf_tol).build()
        return self._opt

    @property
    def jopt(self):
        if self._jopt is None:
            from gpu4pyscf_amd.scf import j_engine
            self._jopt = j_engine._VHFOpt(self.mol, self.direct_scf_tol).build()
        return self._jopt

    def get_j(self, mol=None, dm=None, hermi=1):
        '''J of ``dm`` by the J-engine (gpu4pyscf/scf/hf.py:947-958).'''
        jopt = self.jopt
        dms = jopt.apply_coeff_C_mat_CT(dm.reshape(-1, dm.shape[-1], dm.shape[-1]))
        vj = jopt.get_j(dms, hermi)
        return jopt.apply_coeff_CT_mat_C(vj).reshape(dm.shape)

    def get_hcore_ovlp(self):
        S, T, V = int1e(self.mol, self.opt)
        return T + V, S

    def get_jk(self, mol=None, dm=None, hermi=1, with_j=True, with_k=True, screen=None):
        '''J and K of ``dm`` (device tensor in the original AO basis); screen
        as for _VHFOpt.get_jk.'''
        self.jk_builds += 1
        if self.jk_hook is not None:
            return self.jk_hook(dm, hermi)
        opt = self.opt
        dms = opt.apply_coeff_C_mat_CT(dm.reshape(-1, dm.shape[-1], dm.shape[-1]))
        vj, vk = opt.get_jk(dms, hermi, with_j=with_j, with_k=with_k,
                            fp32_tol=self._fp32_now, count_quartets=True, screen=screen)
        self.quartets += opt.last_quartets
        self.quartets_fp32 += opt.last_quartets_fp32
        shape = dm.shape
        vj = opt.apply_coeff_CT_mat_C(vj).reshape(shape) if with_j else None
        vk = opt.apply_coeff_CT_mat_C(vk).reshape(shape) if with_k else None
        return vj, vk

    def get_veff(self, dm, dm_last=None, vhf_last=None):
        '''J - K/2 built on the density change (gpu4pyscf/scf/hf.py:91-109).'''
        incremental = dm_last is not None and vhf_last is not None
        ddm = dm - dm_last if incremental else dm
        if self.split_jk:
            if self.use_j_engine:
                vj = self.get_j(self.mol, ddm, hermi=1)
            else:
                vj = self.get_jk(self.mol, ddm, hermi=1, with_k=False, screen='j')[0]
            vk = self.get_jk(self.mol, ddm, hermi=1, with_j=False, screen='k')[1]
        else:
            vj, vk = self.get_jk(self.mol, ddm, hermi=1)
        v = vj - .5 * vk
        return vhf_last + v if incremental else v

    # -- linear algebra ---------------------------------------------------
    def _orth(self, S):
        torch = _torch()
        w, U = torch.linalg.eigh(S)
        keep = w > LINDEP_THRESHOLD
        return U[:, keep] / torch.sqrt(w[keep])

    def eig(self, F, X):
        torch = _torch()
        e, c = torch.linalg.eigh(X.T @ F @ X)
        return e, X @ c

    def get_occ(self, mo_energy):
        torch = _torch()
        nocc = self.mol.nelectron // 2
        occ = torch.zeros_like(mo_energy)
        idx = torch.argsort(mo_energy)
        occ[idx[:nocc]] = 2.
        return occ

    @staticmethod
    def make_rdm1(mo_coeff, mo_occ):
        return (mo_coeff * mo_occ) @ mo_coeff.T

    def project_density(self, dm, S, X):
        '''Closest closed-shell idempotent density: the nocc leading natural
        orbitals of dm in the orthonormal basis X (X^T S X = 1), doubly
        occupied.'''
        torch = _torch()
        d_orth = X.T @ S @ dm @ S @ X
        w, v = torch.linalg.eigh(.5 * (d_orth + d_orth.T))
        nocc = self.mol.nelectron // 2
        c = X @ v[:, -nocc:]
        return 2. * (c @ c.T)

    # -- initial guess ----------------------------------------------------
    def get_init_guess(self, h1e, S, X, key=None):
        torch = _torch()
        key = self.init_guess if key is None else key
        if callable(key):
            # a learned or user guess: symmetrise and scale to Tr(DS) = N
            dm = torch.as_tensor(key(self.mol), dtype=torch.float64, device=S.device)
            dm = .5 * (dm + dm.T)
            tr = float(torch.sum(dm * S))
            if tr > 0:
                dm = dm * (self.mol.nelectron / tr)
            if self.project_guess:
                dm = self.project_density(dm, S, X)
            return dm
        if not isinstance(key, str):
            return torch.as_tensor(key, dtype=torch.float64, device=S.device)
        if key == '1e':
            e, c = self.eig(h1e, X)
            return self.make_rdm1(c, self.get_occ(e))
        if key == 'atom':
            from gpu4pyscf_amd.scf.atom_guess import atom_guess
            return torch.as_tensor(atom_guess(self.mol), dtype=torch.float64, device=S.device)
        raise ValueError(f'unknown init_guess {key!r}')

    # -- driver -------------------------------------------------------------
    def _timed(self, key, fn, *args):
        '''fn(*args), adding its wall time (device synchronised) to
        self.timings[key] when self.time_phases is set.'''
        if not self.time_phases:
            return fn(*args)
        torch = _torch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn(*args)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.timings[key] = self.timings.get(key, 0.) + dt
        self.timings.setdefault(key + '_each', []).append(dt)
        return out

    def kernel(self, dm0=None):
        torch = _torch()
        t_start = time.perf_counter()
        self.timings = {}
        mol = self.mol
        e_nuc = mol.energy_nuc()
        h1e, S = self.get_hcore_ovlp()
        X = self._orth(S)
        conv_tol_grad = self.conv_tol_grad or math.sqrt(self.conv_tol)
        dm = self.get_init_guess(h1e, S, X, dm0 if dm0 is not None else None)
        t_guess = time.perf_counter()
        self._fp32_now = None           # the full-density build is FP64
        vhf = self._timed('jk', self.get_veff, dm)
        self._fp32_now = self.fp32_tol  # incremental builds may use FP32
        e_tot = float(torch.sum(dm * (h1e + .5 * vhf))) + e_nuc
        diis = CDIIS(self.diis_space)
        self.converged = False
        self.history = [e_tot]
        self.history_t = [time.perf_counter() - t_start]
        cycle = 0
        for cycle in range(self.max_cycle):
            dm_last = dm
            last_e = e_tot
            fock = h1e + vhf
            if cycle >= self.diis_start_cycle:
                sdf = S @ dm @ fock
                err = X.T @ (sdf.T - sdf) @ X
                fock = self._timed('diis', diis.update, fock, err)
            mo_energy, mo_coeff = self._timed('eig', self.eig, fock, X)
            mo_occ = self.get_occ(mo_energy)
            dm = self.make_rdm1(mo_coeff, mo_occ)
            vhf = self._timed('jk', self.get_veff, dm, dm_last, vhf)
            e_tot = float(torch.sum(dm * (h1e + .5 * vhf))) + e_nuc
            fock = h1e + vhf
            # orbital gradient: occupied-virtual block of the MO Fock matrix
            occ = mo_occ > 0
            f_mo = mo_coeff.T @ fock @ mo_coeff
            g = f_mo[~occ][:, occ]
            norm_gorb = float(torch.linalg.norm(g)) * 2.
            self.history.append(e_tot)
            self.history_t.append(time.perf_counter() - t_start)
            if self.verbose:
                print(f'cycle {cycle+1}: E = {e_tot:.12f} dE = {e_tot-last_e:.3e} '
                      f'|g| = {norm_gorb:.3e}', flush=True)
            if abs(e_tot - last_e) < self.conv_tol and norm_gorb < conv_tol_grad:
                self.converged = True
                break
        self.cycles = cycle + 1
        if self.fp32_tol and self.final_fp64:
            self._fp32_now = None
            vhf = self.get_veff(dm)
            e_tot = float(torch.sum(dm * (h1e + .5 * vhf))) + e_nuc
        self._fp32_now = None
        self.e_tot = e_tot
        self.mo_energy, self.mo_coeff, self.mo_occ = mo_energy, mo_coeff, mo_occ
        self.dm = dm
        torch.cuda.synchronize()
        self.timings.update(guess=t_guess - t_start, total=time.perf_counter() - t_start)
        return e_tot

    scf = kernel
