'''Two-electron part of the RHF/UHF nuclear gradient on the MI355X engine.

Drop-in for ``_jk_energy_per_atom`` (gpu4pyscf/grad/rhf.py:60-171): the first
derivatives, per atom, of the two-electron energy of one (closed shell) or
two (alpha, beta) densities, j_factor E_J - k_factor E_K, from the ERI
first derivatives evaluated by the runtime-l Rys kernel of libjkhip.so
(``jk_energy_ip1``).  The screening object is the J/K engine's ``_VHFOpt``;
the two-particle density and the survival predicate follow
gpu4pyscf/lib/gvhf-rys/rys_contract_jk_ip1.cu:523-577,1789-1797 and
create_tasks.cu:829-889.  Range separation follows the molecule's omega
(omega < 0: short range, computed as full range minus long range).
'''
import ctypes
import math

import numpy as np

from gpu4pyscf_amd.lib import clib
from gpu4pyscf_amd.scf import jk as jk_mod

__all__ = ['_jk_energy_per_atom', '_VHFOpt']

_VHFOpt = jk_mod._VHFOpt


def _jk_energy_per_atom(vhfopt, dm, j_factor=1., k_factor=1., omega=None, lr_factor=None,
                        sr_factor=None, verbose=None):
    '''(natm, 3) numpy array: d/dR_atom of j_factor E_J - k_factor E_K for
    ``dm`` (nao, nao) or (n_dm <= 2, nao, nao) in the original AO basis
    (numpy or torch).  Same factors as the reference: for one density the
    result is the gradient of 1/2 Tr(D (j_factor J - k_factor K/2)) with
    j_factor = k_factor = 1 the RHF two-electron gradient.'''
    assert lr_factor is None and sr_factor is None
    import torch
    lib = clib.load()
    if vhfopt._dev is None:
        vhfopt.build()
    mol = vhfopt.sorted_mol
    d = vhfopt._dev
    dm_t, _ = jk_mod._to_device(dm)
    nao0 = dm_t.shape[-1]
    dms = vhfopt.apply_coeff_C_mat_CT(dm_t.reshape(-1, nao0, nao0)).contiguous()
    n_dm, nao = dms.shape[0], dms.shape[-1]
    assert n_dm <= 2
    nbas = mol.nbas
    stream = jk_mod._stream_ptr()
    dm_cond = torch.empty((nbas, nbas), dtype=torch.float32, device=dms.device)
    work = torch.empty((nbas, nbas), dtype=torch.float64, device=dms.device)
    # log of the block maxima over the densities, not symmetrised
    clib.check(lib.jk_dm_cond(jk_mod._ptr(dm_cond), jk_mod._ptr(work), jk_mod._ptr(dms),
                              ctypes.c_int(n_dm), ctypes.c_int(nao), jk_mod._ptr(d['ao_loc']),
                              ctypes.c_int(nbas), ctypes.c_int(1), stream), 'jk_dm_cond')
    natm = vhfopt.mol.natm
    ejk = torch.zeros((natm, 3), dtype=torch.float64, device=dms.device)
    args = clib.JKBuildArgs()
    args.abi_version = clib.JK_ABI_VERSION
    args.nbas, args.nao = nbas, nao
    args.bas = d['bas'].data_ptr()
    args.env = d['env'].data_ptr()
    args.ao_loc = d['ao_loc'].data_ptr()
    args.dm_cond = dm_cond.data_ptr()
    args.cutoff = np.float32(math.log(vhfopt.direct_scf_tol))
    args.n_dm = n_dm
    args.dm = dms.data_ptr()
    args.rys = vhfopt._rys
    args.n_repl = 1
    args.repl_stride = dms.numel()
    uniq = mol.uniq_l_ctr
    passes = vhfopt.operator_passes(omega)
    for (i, j, k, l) in vhfopt.tasks():
        args.li, args.lj, args.lk, args.ll = (int(uniq[x, 0]) for x in (i, j, k, l))
        args.iprim, args.jprim, args.kprim, args.lprim = (int(uniq[x, 1]) for x in (i, j, k, l))
        pij, pkl = d['pairs'][i, j], d['pairs'][k, l]
        args.npairs_ij, args.npairs_kl = len(pij), len(pkl)
        args.pairs_ij, args.pairs_kl = pij.data_ptr(), pkl.data_ptr()
        args.q_ij, args.q_kl = d['q'][i, j].data_ptr(), d['q'][k, l].data_ptr()
        args.pdata_ij = d['pdata'][i, j].data_ptr()
        args.pdata_kl = d['pdata'][k, l].data_ptr()
        for w, f in passes:
            args.omega, args.scale = w, f
            st = lib.jk_energy_ip1(ctypes.byref(args), ctypes.c_void_p(ejk.data_ptr()),
                                   float(j_factor), float(k_factor), stream)
            clib.check(st, f'jk_energy_ip1 {tuple(int(uniq[x, 0]) for x in (i, j, k, l))}')
    return ejk.cpu().numpy()
