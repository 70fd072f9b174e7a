'''McMurchie-Davidson J-engine on MI355X: drop-in for gpu4pyscf.scf.j_engine
(``get_j``, ``_VHFOpt``; gpu4pyscf/scf/j_engine.py:50-236).

The Coulomb matrix is built on the decontracted sorted basis (one shell per
primitive, SortedGTO(decontract=True), as the reference's J-engine does,
j_engine.py:96-97): every screened shell pair carries its Hermite density
(jk_md_dm_to_hermite), every bra pair sums the Hermite Coulomb integrals
R_{t+t',u+u',v+v'} against the Hermite densities of every ket pair whose
bound exp(q_ij + q_kl) max(|D_ij|, |D_kl|) reaches the cutoff (the J-only
predicate of the reference, create_tasks.cu:213-225; jk_md_build_j), and the
Hermite J of the bra pairs is expanded back to AO pairs
(jk_md_hermite_to_j).  Classes whose Hermite orders exceed MD_LMAX
(li + lj + lk + ll > MD_LMAX) are evaluated by the Rys J-only build
(screened on D_ij and D_kl) on the same pair lists.  Pair lists follow the
reference's J-engine: tile 1, bin 0 dropped, sorted by descending q_cond
(j_engine.py:239-259).
'''

import ctypes
import math
import os

import numpy as np

from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
from gpu4pyscf_amd.lib import clib
from gpu4pyscf_amd.scf import jk as jk_mod

__all__ = ['get_j', '_VHFOpt']

MD_LMAX = 4           # li + lj + lk + ll of the Hermite kernels
MD_PAIR_LMAX = 4      # largest l of a shell in the Hermite transforms
BOYS_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         'csrc', 'md_boys_table.npz')


def _nh(L):
    return (L + 1) * (L + 2) * (L + 3) // 6


def get_j(mol, dm, hermi=1, vhfopt=None, verbose=None):
    '''Coulomb matrix by the J-engine (gpu4pyscf/scf/j_engine.py:50-87);
    numpy in -> numpy out, torch in -> torch out.'''
    if vhfopt is None:
        vhfopt = _VHFOpt(mol).build()
    dm_t, was_torch = jk_mod._to_device(dm)
    shape = dm_t.shape
    nao0 = shape[-1]
    dms = vhfopt.apply_coeff_C_mat_CT(dm_t.reshape(-1, nao0, nao0))
    vj = vhfopt.get_j(dms, hermi)
    vj = vhfopt.apply_coeff_CT_mat_C(vj).reshape(shape)
    return vj if was_torch else vj.cpu().numpy()


class _VHFOpt(jk_mod._VHFOpt):
    '''Screening object of the J-engine (gpu4pyscf/scf/j_engine.py:89-236).'''

    def __init__(self, mol, direct_scf_tol=1e-13):
        super().__init__(mol, direct_scf_tol, tile=1)
        # the Rys J-only build covers the classes above the Hermite kernels
        self.task_filter = lambda ls: sum(ls) > MD_LMAX
        self.last_md_quartets = None
        self.last_md_hash = None

    def _make_sorted_mol(self):
        return SortedGTO(self.mol, decontract=True)

    def _order_pairs(self, q, split):
        keep = np.where(q > split[0])[0]
        return keep[np.argsort(-q[keep], kind='stable')]

    def build(self, group_size=None, verbose=None):
        super().build()
        torch = jk_mod._torch()
        mol = self.sorted_mol
        uniq = mol.uniq_l_ctr
        t = np.load(BOYS_PATH)
        boys = np.ascontiguousarray(t['boys'], dtype=np.float64)
        d = self._dev
        d['boys'] = torch.as_tensor(boys, device=self.device)
        self._boys_meta = (boys.shape[0], boys.shape[1], float(t['step']))
        ng = len(mol.l_ctr_counts)
        ls = [int(uniq[g, 0]) for g in range(ng)]
        keys = [(i, j) for i in range(ng) for j in range(i + 1)
                if len(self.bas_pair_cache[i, j][0]) and ls[i] <= MD_PAIR_LMAX and
                ls[i] + ls[j] <= MD_LMAX]
        self._md_keys = keys
        self._md_l = {k: (ls[k[0]], ls[k[1]]) for k in keys}
        # every bra group pair against every ket group pair within MD_LMAX
        self._md_tasks = [(a, b) for a in keys for b in keys
                          if sum(self._md_l[a]) + sum(self._md_l[b]) <= MD_LMAX]
        return self

    def md_tasks(self):
        return list(self._md_tasks)

    def get_j(self, dms, hermi=1, verbose=None, count_quartets=False):
        '''J of densities ``dms`` (n_dm, nao_c, nao_c) on the decontracted
        sorted basis (torch, float64, on the device).  A non-Hermitian
        density is symmetrised first (gpu4pyscf/scf/j_engine.py:79-81).
        count_quartets: leave the number of surviving canonical quartets
        (bas_ij >= bas_kl) of the Hermite classes in self.last_md_quartets and
        their set checksum (sum mod 2**64 of the SplitMix64 finaliser of
        bas_ij << 32 | bas_kl) in self.last_md_hash (last density).'''
        torch = jk_mod._torch()
        if self._dev is None:
            self.build()
        lib = clib.load()
        mol = self.sorted_mol
        d = self._dev
        nbas, nao = mol.nbas, mol.nao
        assert dms.ndim == 3 and dms.shape[-1] == nao
        dms = dms.contiguous()
        if hermi != 1:
            dms = ((dms + dms.transpose(1, 2)) * .5).contiguous()
        stream = jk_mod._stream_ptr()
        dev = dms.device
        ptr = jk_mod._ptr
        vj = torch.zeros_like(dms)
        counter = torch.zeros(2, dtype=torch.int64, device=dev) if count_quartets else None
        rows, bm, step = self._boys_meta
        dm_cond = torch.empty((nbas, nbas), dtype=torch.float32, device=dev)
        work = torch.empty((nbas, nbas), dtype=torch.float64, device=dev)
        for n in range(dms.shape[0]):
            dm = dms[n]
            clib.check(lib.jk_dm_cond(ptr(dm_cond), ptr(work), ptr(dm), 1, nao,
                                      ptr(d['ao_loc']), nbas, 1, stream), 'jk_dm_cond')
            flat = dm_cond.view(-1)
            if counter is not None:
                counter.zero_()
            dh, jh, dmax = {}, {}, {}
            for key in self._md_keys:
                li, lj = self._md_l[key]
                pairs = d['pairs'][key]
                npairs = len(pairs)
                nh = _nh(li + lj)
                dh[key] = torch.empty((npairs, nh), dtype=torch.float64, device=dev)
                jh[key] = torch.zeros((npairs, nh), dtype=torch.float64, device=dev)
                dmax[key] = flat[pairs.to(torch.int64)].contiguous()
                clib.check(lib.jk_md_dm_to_hermite(
                    ptr(dh[key]), ptr(dm), nao, ptr(pairs), npairs, ptr(d['pdata'][key]),
                    li, lj, ptr(d['bas']), ptr(d['env']), ptr(d['ao_loc']), nbas, stream),
                    'jk_md_dm_to_hermite')
            dm_penalty = float(dm_cond.max())
            args = clib.MDBuildArgs()
            args.abi_version = clib.JK_ABI_VERSION
            args.cutoff = np.float32(math.log(self.direct_scf_tol))
            args.boys = d['boys'].data_ptr()
            args.boys_rows, args.boys_m = rows, bm
            args.boys_step, args.boys_inv_step = step, 1. / step
            args.n_quartets = counter.data_ptr() if counter is not None else None
            args.quartet_hash = counter.data_ptr() + 8 if counter is not None else None
            args.dm_penalty = dm_penalty
            # the molecule's operator (screening q_cond uses the same)
            args.omega = float(getattr(self.mol, 'omega', 0.) or 0.)
            for a, b in self._md_tasks:
                args.lij, args.lkl = sum(self._md_l[a]), sum(self._md_l[b])
                args.npairs_ij, args.npairs_kl = len(d['pairs'][a]), len(d['pairs'][b])
                args.pdata_ij, args.pdata_kl = d['pdata'][a].data_ptr(), d['pdata'][b].data_ptr()
                args.q_ij, args.q_kl = d['q'][a].data_ptr(), d['q'][b].data_ptr()
                args.dmax_ij, args.dmax_kl = dmax[a].data_ptr(), dmax[b].data_ptr()
                args.pairs_ij, args.pairs_kl = d['pairs'][a].data_ptr(), d['pairs'][b].data_ptr()
                args.dh_kl = dh[b].data_ptr()
                args.jh_ij = jh[a].data_ptr()
                clib.check(lib.jk_md_build_j(ctypes.byref(args), stream), 'jk_md_build_j')
            for key in self._md_keys:
                li, lj = self._md_l[key]
                pairs = d['pairs'][key]
                clib.check(lib.jk_md_hermite_to_j(
                    ptr(vj[n]), ptr(jh[key]), nao, ptr(pairs), len(pairs),
                    ptr(d['pdata'][key]), li, lj, ptr(d['bas']), ptr(d['env']),
                    ptr(d['ao_loc']), nbas, stream), 'jk_md_hermite_to_j')
        if self.tasks():
            # classes above MD_LMAX: Rys J-only build on the same pair lists
            vj_rys = super().get_jk(dms, 1, with_k=False, screen='j')[0]
            vj += vj_rys
        if counter is not None:
            c = counter.cpu().numpy()
            self.last_md_quartets = int(c[0])
            self.last_md_hash = int(c[1]) & (2**64 - 1)
        return vj
