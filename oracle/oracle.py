'''CPU oracle for the direct-SCF J/K path -- TEST INFRASTRUCTURE ONLY.

This module is the checker the HIP path is compared against.  Only tests/,
the smoke test of the build entry point and the CPU-baseline leg of bench.py
use it; the product never imports it.

It restates, on the CPU:

* the screening setup of ``_VHFOpt.build`` -- gpu4pyscf/scf/jk.py:984-1120:
  float32 shell-pair overlap bound (lib/pbc/overlap.cu:2163-2223, REMOTE
  threshold 50), tile-ordered pair enumeration (lib/pbc/supmol_sr_estimator.cu:
  461-487, tile 12), q_cond = log(max|(ij|ij)|)/2 (lib/gvhf-rys/
  nr_sr_estimator.cu:100-397, here evaluated in FP64 then rounded to float32,
  -700 for vanishing pairs) and the descending 4-wide log bins of
  gpu4pyscf/pbc/scf/rsjk.py:2268-2277 (bin 0 dropped);
* dm_cond / dm_penalty -- gpu4pyscf/scf/jk.py:510-516 with the block absmax of
  gpu4pyscf/lib/cupy_helper.py:1068-1184;
* the kernel accumulation (rys_jk.c) and the hermi post-processing --
  gpu4pyscf/scf/jk.py:594-616;
* the AO transforms -- gpu4pyscf/scf/jk.py:122,139,146 via the dense C matrix.
'''

import ctypes
import math
import os
import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
_lib = None

LOG_CUTOFF_MARGIN = 4.
NEGLIGIBLE = -700.


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_here, 'liboracle_jk.so')
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(['make', '-s', '-C', _here])
        _lib = ctypes.CDLL(path)
        _lib.oracle_jk.restype = ctypes.c_long
        _lib.oracle_splitmix64_mix.restype = ctypes.c_uint64
        _lib.oracle_splitmix64_mix.argtypes = [ctypes.c_uint64]
        _lib.oracle_set_operator.argtypes = [ctypes.c_double] * 3
        _lib.oracle_set_screen.argtypes = [ctypes.c_int]
        _lib.oracle_ejk_ip1.argtypes = ([ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 2 +
                                        [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int] +
                                        [ctypes.c_double] * 2 + [ctypes.c_void_p])
    return _lib


def rsh_factors(omega, lr_factor=None, sr_factor=None):
    '''(omega, lr, sr) of lr erf(|omega| r)/r + sr erfc(|omega| r)/r:
    libcint's sign convention when no factor is given (omega > 0 long range,
    omega < 0 short range), gpu4pyscf/scf/jk.py:1122-1168.'''
    omega = omega or 0.
    if lr_factor is None and sr_factor is None:
        if omega == 0:
            return 0., 1., 1.
        return (abs(omega), 0., 1.) if omega < 0 else (omega, 1., 0.)
    if lr_factor is None:
        lr_factor = sr_factor if omega == 0 else 0.
    if sr_factor is None:
        sr_factor = lr_factor if omega == 0 else 0.
    return abs(omega), lr_factor, sr_factor


def set_operator(omega=0., lr_factor=None, sr_factor=None):
    '''Operator of every later ERI of the oracle (a process-wide setting).'''
    w, lr, sr = rsh_factors(omega, lr_factor, sr_factor)
    lib().oracle_set_operator(w, lr, sr)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# survival screens: the six density blocks of a J/K build, the two Coulomb
# blocks of a J-only build, the four exchange blocks of a K-only build
# (gpu4pyscf/lib/gvhf-rys/create_tasks.cu:155-163, 222-225, 86-93)
SCREENS = {'jk': 0, 'j': 1, 'k': 2}


class _screen:
    '''Context: the C oracle's survival screen.'''
    def __init__(self, screen):
        self.code = SCREENS[screen]

    def __enter__(self):
        lib().oracle_set_screen(self.code)

    def __exit__(self, *exc):
        lib().oracle_set_screen(0)


def rys_roots(n, x):
    u = np.zeros(n); w = np.zeros(n)
    lib().oracle_rys_roots(ctypes.c_int(n), ctypes.c_double(x), _p(u), _p(w))
    return u, w


def eri_shell(bas, env, i, j, k, l):
    from math import comb
    ls = bas[[i, j, k, l], 1]
    nf = [(x + 1) * (x + 2) // 2 for x in ls]
    out = np.zeros(int(np.prod(nf)))
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    lib().oracle_eri_shell(_p(bas), _p(env), ctypes.c_int(i), ctypes.c_int(j),
                           ctypes.c_int(k), ctypes.c_int(l), _p(out))
    return out.reshape(nf)


def eri_full(sorted_mol):
    '''Full Cartesian ERI tensor of the sorted basis (tiny systems only).'''
    bas, env, ao_loc = sorted_mol._bas, sorted_mol.env_kernel, sorted_mol.ao_loc
    nao = sorted_mol.nao
    eri = np.zeros((nao,)*4)
    n = sorted_mol.nbas
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    eri[ao_loc[i]:ao_loc[i+1], ao_loc[j]:ao_loc[j+1],
                        ao_loc[k]:ao_loc[k+1], ao_loc[l]:ao_loc[l+1]] = \
                        eri_shell(bas, env, i, j, k, l)
    return eri


# ---------------------------------------------------------------- screening
def overlap_mask(sorted_mol, precision):
    '''float32 log-overlap bound > log(precision), lower triangle (ish >= jsh).
    lib/pbc/overlap.cu:2163-2223 with hermi=1, single image.'''
    f32 = np.float32
    exps, cs = sorted_mol.diffuse_pgto()
    exps = exps.astype(f32)
    log_coeff = np.log(np.abs(cs.astype(f32)))
    coords = sorted_mol.bas_coords().astype(f32)
    ls = sorted_mol._bas[:, 1].astype(f32)
    n = sorted_mol.nbas
    ai = exps[:, None]; aj = exps[None, :]
    aij = ai + aj
    fi = ai / aij
    fj = aj / aij
    theta = ai * fj
    d = coords[None, :, :] - coords[:, None, :]          # xj - xi
    rr = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    fac_norm = log_coeff[:, None] + log_coeff[None, :] + f32(1.717) - f32(1.5) * np.log(aij)
    dr = np.sqrt(rr)
    dri = fj * dr
    drj = fi * dr
    li = ls[:, None]; lj = ls[None, :]
    with np.errstate(divide='ignore', invalid='ignore'):
        dri_fac = f32(.5) * li * np.log(f32(.5) * li / aij + dri * dri + f32(1e-9))
        drj_fac = f32(.5) * lj * np.log(f32(.5) * lj / aij + drj * drj + f32(1e-9))
    log_ovlp = fac_norm - theta * rr + dri_fac + drj_fac
    mask = (log_ovlp > f32(math.log(precision))) & ~(theta * rr > f32(50))
    mask = np.tril(mask)
    return mask


def tile_pairs(ish0, ish1, jsh0, jsh1, nbas, tile):
    '''Pair ids ish*nbas+jsh in the tile order of
    lib/pbc/supmol_sr_estimator.cu:461-487.'''
    ish = np.arange(ish0, ish1)
    jsh = np.arange(jsh0, jsh1)
    out = []
    for it in range(0, len(ish), tile):
        ii = ish[it:it+tile]
        for jt in range(0, len(jsh), tile):
            jj = jsh[jt:jt+tile]
            out.append((ii[:, None] * nbas + jj[None, :]).ravel())
    if not out:
        return np.zeros(0, dtype=np.int64)
    return np.concatenate(out)


def q_cond_of_pairs(sorted_mol, pairs):
    qmax = np.zeros(len(pairs))
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
    bas = np.ascontiguousarray(sorted_mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(sorted_mol.env_kernel)
    lib().oracle_pair_diag_max(_p(bas), _p(env), ctypes.c_int(sorted_mol.nbas),
                               _p(pairs), ctypes.c_int(len(pairs)), _p(qmax))
    with np.errstate(divide='ignore'):
        q = np.where(qmax > 0, np.log(qmax) * .5, NEGLIGIBLE)
    return q.astype(np.float32)


def group_by_split_points(q, split_points):
    '''Indices of q sorted into descending 4-wide log bins, bin 0 dropped
    (gpu4pyscf/pbc/scf/rsjk.py:2268-2277).'''
    bins = np.searchsorted(split_points, q.astype(np.float64))
    nb = len(split_points)
    subsets = [np.where(bins == i)[0] for i in range(1, nb + 1)]
    return np.concatenate(subsets[::-1]).astype(np.int64) if subsets else np.zeros(0, int)


def build_pairs(sorted_mol, direct_scf_tol=1e-13, tile=12, q_cond_fn=None):
    '''Per group pair (i >= j): (pair ids uint32, q_cond float32), as
    _cache_q_cond_and_non0pairs (gpu4pyscf/scf/jk.py:984-1120).'''
    nbas = sorted_mol.nbas
    mask = overlap_mask(sorted_mol, direct_scf_tol**2).ravel()
    offs = sorted_mol.group_offsets
    ng = len(offs) - 1
    split = np.arange(math.log(direct_scf_tol), 2., LOG_CUTOFF_MARGIN)
    cache = {}
    for i in range(ng):
        for j in range(i + 1):
            p = tile_pairs(offs[i], offs[i+1], offs[j], offs[j+1], nbas, tile)
            p = p[mask[p]]
            q = (q_cond_fn or q_cond_of_pairs)(sorted_mol, p) if len(p) else np.zeros(0, np.float32)
            idx = group_by_split_points(q, split)
            cache[i, j] = (p[idx].astype(np.uint32), q[idx])
    return cache


def dm_condition(dms, ao_loc, hermi):
    '''float32 log(max |D| per shell block + 1e-300) and its maximum
    (gpu4pyscf/scf/jk.py:510-516).'''
    a = np.abs(dms).max(axis=0)
    starts = np.asarray(ao_loc[:-1])
    cond = np.maximum.reduceat(np.maximum.reduceat(a, starts, axis=0), starts, axis=1)
    if hermi == 0:
        cond = cond + cond.T
    cond = np.log(cond + 1e-300).astype(np.float32)
    return cond, float(cond.max())


def flatten_pairs(cache, ngroups):
    keys = [(i, j) for i in range(ngroups) for j in range(i + 1)]
    off = np.zeros(len(keys) + 1, dtype=np.int32)
    for n, k in enumerate(keys):
        off[n+1] = off[n] + len(cache[k][0])
    pairs = np.concatenate([cache[k][0] for k in keys]).astype(np.uint32)
    q = np.concatenate([cache[k][1] for k in keys]).astype(np.float32)
    return off, pairs, q


def jk_sorted(sorted_mol, dms, hermi, cache, direct_scf_tol=1e-13,
              with_j=True, with_k=True, list_quartets=0, ij_stride=1, screen='jk'):
    '''J/K on the sorted Cartesian basis including the hermi post-processing.
    Returns (vj, vk, n_surviving, quartet_list or None).  screen 'j' (a
    J-only build) first symmetrises the densities as the reference's get_j
    does (gpu4pyscf/scf/jk.py:224-226), 'k' is a K-only build
    (gpu4pyscf/scf/jk.py:785-927).'''
    dms = np.asarray(dms, dtype=np.float64)
    if screen == 'j' and hermi != 1:
        dms = (dms + dms.transpose(0, 2, 1)) * .5
        hermi = 1
    ao_loc = sorted_mol.ao_loc
    cond, penalty = dm_condition(dms, ao_loc, hermi)
    if hermi == 0:
        dms_run = np.concatenate([dms, dms.transpose(0, 2, 1)])
    else:
        dms_run = dms
    dms_run = np.ascontiguousarray(dms_run)
    n_dm, nao = dms_run.shape[:2]
    ng = len(sorted_mol.l_ctr_counts)
    off, pairs, q = flatten_pairs(cache, ng)
    vj = np.zeros_like(dms_run)
    vk = np.zeros_like(dms_run)
    bas = np.ascontiguousarray(sorted_mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(sorted_mol.env_kernel)
    aol = np.ascontiguousarray(ao_loc, dtype=np.int32)
    cutoff = np.float32(math.log(direct_scf_tol))
    quartets = None
    qptr = None
    if list_quartets:
        quartets = np.zeros((list_quartets, 2), dtype=np.uint32)
        qptr = _p(quartets)
    L = lib()
    with _screen(screen):
        nsurv = L.oracle_jk(
            ctypes.c_int(n_dm), ctypes.c_int(nao), _p(dms_run), _p(vj), _p(vk),
            _p(bas), ctypes.c_int(sorted_mol.nbas), _p(env), _p(aol),
            ctypes.c_int(ng), _p(off), _p(pairs), _p(q), _p(cond),
            ctypes.c_float(cutoff), ctypes.c_float(penalty),
            ctypes.c_int(int(with_j)), ctypes.c_int(int(with_k)),
            qptr, ctypes.c_long(list_quartets), ctypes.c_int(ij_stride))
    if hermi == 1:
        vj *= 2.
        vk = vk + vk.transpose(0, 2, 1)
    elif hermi == 2:
        vj[:] = 0
        vk = vk - vk.transpose(0, 2, 1)
    else:
        h = n_dm // 2
        vj = vj[:h] + vj[h:].transpose(0, 2, 1)
        vk = vk[:h] + vk[h:].transpose(0, 2, 1)
    vj = vj + vj.transpose(0, 2, 1)
    return vj, vk, nsurv, quartets


def survivors(sorted_mol, dms, hermi, cache, direct_scf_tol=1e-13, screen='jk'):
    '''Sorted array of surviving quartet keys bas_ij*nbas^2 + bas_kl.'''
    dms = np.asarray(dms, dtype=np.float64)
    if screen == 'j' and hermi != 1:
        dms = (dms + dms.transpose(0, 2, 1)) * .5
        hermi = 1
    cond, penalty = dm_condition(dms, sorted_mol.ao_loc, hermi)
    ng = len(sorted_mol.l_ctr_counts)
    off, pairs, q = flatten_pairs(cache, ng)
    bas = np.ascontiguousarray(sorted_mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(sorted_mol.env_kernel)
    aol = np.ascontiguousarray(sorted_mol.ao_loc, dtype=np.int32)
    cutoff = np.float32(math.log(direct_scf_tol))
    L = lib()
    cap = 1
    while True:
        out = np.zeros((cap, 2), dtype=np.uint32)
        with _screen(screen):
            n = L.oracle_jk(ctypes.c_int(0), ctypes.c_int(sorted_mol.nao), None, None, None,
                        _p(bas), ctypes.c_int(sorted_mol.nbas), _p(env), _p(aol),
                        ctypes.c_int(ng), _p(off), _p(pairs), _p(q), _p(cond),
                        ctypes.c_float(cutoff), ctypes.c_float(penalty),
                            ctypes.c_int(0), ctypes.c_int(0), _p(out), ctypes.c_long(cap),
                            ctypes.c_int(1))
        if n <= cap:
            out = out[:n]
            break
        cap = n
    nb2 = np.uint64(sorted_mol.nbas) ** 2
    keys = out[:, 0].astype(np.uint64) * nb2 + out[:, 1].astype(np.uint64)
    return np.sort(keys)


def quartet_sets(sorted_mol, dms, hermi, cache, direct_scf_tol=1e-13, screen='jk'):
    '''{(i, j, k, l): (count, checksum)} of the surviving quartets of every
    class task with survivors: checksum = sum mod 2**64 of the SplitMix64
    finaliser of bas_ij << 32 | bas_kl (the C oracle's quartet_mix).'''
    dms = np.asarray(dms, dtype=np.float64)
    if screen == 'j' and hermi != 1:
        dms = (dms + dms.transpose(0, 2, 1)) * .5
        hermi = 1
    cond, penalty = dm_condition(dms, sorted_mol.ao_loc, hermi)
    ng = len(sorted_mol.l_ctr_counts)
    off, pairs, q = flatten_pairs(cache, ng)
    tasks = [(i, j, k, l) for i in range(ng) for j in range(i + 1)
             for k in range(i + 1) for l in range(k + 1)]
    counts = np.zeros(len(tasks), dtype=np.int64)
    hashes = np.zeros(len(tasks), dtype=np.uint64)
    with _screen(screen):
        lib().oracle_quartet_sets(
            ctypes.c_int(sorted_mol.nbas), ctypes.c_int(ng), _p(off), _p(pairs), _p(q),
            _p(cond), ctypes.c_float(np.float32(math.log(direct_scf_tol))),
            ctypes.c_float(penalty), _p(counts), _p(hashes))
    return {t: (int(c), int(h)) for t, c, h in zip(tasks, counts, hashes) if c}


def splitmix64_mix(z):
    '''The C oracle's SplitMix64 finaliser of a 64-bit integer.'''
    return int(lib().oracle_splitmix64_mix(ctypes.c_uint64(z)))


def get_jk(mol, dm, hermi=1, with_j=True, with_k=True, direct_scf_tol=1e-13,
           sorted_mol=None, cache=None, omega=None, lr_factor=None, sr_factor=None,
           screen='jk'):
    """Oracle J/K in the original AO basis of ``mol``.  The screening uses
    the molecule's operator (mol.omega); K uses (omega, lr_factor,
    sr_factor) when omega is given, J always the molecule's operator.
    screen: 'jk' (the six density blocks of gpu4pyscf's get_jk), 'j' (J only,
    get_j), 'k' (K only, get_k); with a range-separated K the two halves use
    the 'j' and 'k' screens, as gpu4pyscf/scf/jk.py:129-135 does."""
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol_omega = getattr(mol, 'omega', 0.) or 0.
    if sorted_mol is None:
        sorted_mol = SortedGTO(mol)
    try:
        set_operator(mol_omega)
        if cache is None:
            cache = build_pairs(sorted_mol, direct_scf_tol)
        dm = np.asarray(dm, dtype=np.float64)
        shape = dm.shape
        nao0 = shape[-1]
        dms = dm.reshape(-1, nao0, nao0)
        C = sorted_mol.coeff
        dms_c = C @ dms @ C.T
        if omega is None or rsh_factors(omega, lr_factor, sr_factor) == rsh_factors(mol_omega):
            vj, vk, nsurv, _ = jk_sorted(sorted_mol, dms_c, hermi, cache, direct_scf_tol,
                                         with_j, with_k, screen=screen)
        else:
            vj = vk = None
            if with_j:
                vj = jk_sorted(sorted_mol, dms_c, hermi, cache, direct_scf_tol, True, False,
                               screen='j')[0]
            if with_k:
                set_operator(omega, lr_factor, sr_factor)
                vk = jk_sorted(sorted_mol, dms_c, hermi, cache, direct_scf_tol, False, True,
                               screen='k')[1]
    finally:
        set_operator(0.)
    vj = (C.T @ vj @ C).reshape(shape) if with_j else None
    vk = (C.T @ vk @ C).reshape(shape) if with_k else None
    return vj, vk


def ejk_ip1(mol, dm, j_factor=1., k_factor=1.):
    """Per-atom first derivatives of the two-electron energy (natm x 3) for
    one (closed shell) or two (alpha, beta) densities in the original AO
    basis, with the molecule's operator; no screening."""
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    sm = SortedGTO(mol)
    dm = np.asarray(dm, dtype=np.float64)
    dms = dm.reshape(-1, dm.shape[-1], dm.shape[-1])
    C = sm.coeff
    dms_c = np.ascontiguousarray(C @ dms @ C.T)
    bas = np.ascontiguousarray(sm._bas, dtype=np.int32)
    env = np.ascontiguousarray(sm.env_kernel)
    aol = np.ascontiguousarray(sm.ao_loc, dtype=np.int32)
    atom_of = np.ascontiguousarray(sm._bas[:, 0], dtype=np.int32)
    out = np.zeros((mol.natm, 3))
    try:
        set_operator(getattr(mol, 'omega', 0.) or 0.)
        lib().oracle_ejk_ip1(dms_c.shape[0], sm.nao, _p(dms_c), _p(bas), sm.nbas, _p(env),
                             _p(aol), _p(atom_of), mol.natm, float(j_factor), float(k_factor),
                             _p(out))
    finally:
        set_operator(0.)
    return out


def fp(a):
    a = np.asarray(a)
    return float(np.dot(np.cos(np.arange(a.size)), a.ravel()))
