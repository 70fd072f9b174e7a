'''CPU tests of the host side: the C-ABI library loads and exports every
declared symbol (no GPU calls), the screening bookkeeping equals the
oracle's restatement, the basis layer's invariants and the workload
generators.'''
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'jkhip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char \*)\s*((?:jk|mlp)_\w+)\s*\(',
                                 text, re.M)))


def test_library_exports_header_symbols():
    from gpu4pyscf_amd.lib import clib
    lib = clib.load()
    syms = _declared_symbols()
    assert 'jk_build' in syms and 'jk_int1e' in syms
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(clib.EXPORTS)
    assert lib.jk_abi_version() == clib.JK_ABI_VERSION


def test_struct_layout_matches_header():
    from gpu4pyscf_amd.lib import clib
    # JKRysTable: 2 pointers + 11 int64 + 11 double + 2 int64 + int (+pad)
    assert ctypes.sizeof(clib.JKRysTable) == 16 + 88 + 88 + 16 + 8 + 8 + 16
    assert clib.JKBuildArgs.rys.offset % 8 == 0


def test_jk_build_rejects_bad_arguments():
    '''Argument validation runs on the host before any launch.'''
    from gpu4pyscf_amd.lib import clib
    lib = clib.load()
    a = clib.JKBuildArgs()
    a.abi_version = clib.JK_ABI_VERSION + 7
    assert lib.jk_build(ctypes.byref(a), None) == 1
    a.abi_version = clib.JK_ABI_VERSION
    a.li, a.lj, a.lk, a.ll = clib.JK_LMAX + 1, 0, 0, 0
    assert lib.jk_build(ctypes.byref(a), None) == 2
    assert lib.jk_energy_ip1(ctypes.byref(a), None, 1., 1., None) == 2
    a.li = 1
    a.npairs_ij = a.npairs_kl = 3
    assert lib.jk_build(ctypes.byref(a), None) == 3
    assert b'null' in lib.jk_last_error()
    # replicated accumulation with copies closer than one n_dm x nao^2 block
    # (placeholder addresses: nothing is dereferenced before the check)
    for f in ('bas', 'env', 'ao_loc', 'pairs_ij', 'pairs_kl', 'q_ij', 'q_kl', 'pdata_ij',
              'pdata_kl', 'dm_cond', 'dm', 'vj', 'vk'):
        setattr(a, f, 4096)
    a.nbas, a.nao, a.n_dm = 4, 10, 1
    a.rys.coef = 4096
    a.rys.degree = 11
    a.n_repl, a.repl_stride = 8, 99
    assert lib.jk_build(ctypes.byref(a), None) == 3
    assert b'repl_stride' in lib.jk_last_error()
    # a J-only screen builds no K, a K-only screen no J
    a.n_repl = 1
    for screen, keep in ((clib.JK_SCREEN['j'], 'vj'), (clib.JK_SCREEN['k'], 'vk'), (7, None)):
        a.screen = screen
        assert lib.jk_build(ctypes.byref(a), None) == 3
        assert b'screen' in lib.jk_last_error()
        if keep:
            saved = (a.vj, a.vk)
            a.vj, a.vk = (4096, None) if keep == 'vj' else (None, 4096)
            st = lib.jk_build(ctypes.byref(a), None)
            a.vj, a.vk = saved
            assert b'screen' not in lib.jk_last_error() or st == 0
    a.screen = 0
    # the attenuation parameter of erf(omega r)/r must not be negative
    a.n_repl, a.omega = 1, -.3
    assert lib.jk_build(ctypes.byref(a), None) == 3
    assert b'omega' in lib.jk_last_error()
    # the gradient contraction takes one or two densities and an output
    a.omega, a.n_dm = 0., 3
    assert lib.jk_energy_ip1(ctypes.byref(a), ctypes.c_void_p(4096), 1., 1., None) == 3
    a.n_dm = 1
    assert lib.jk_energy_ip1(ctypes.byref(a), None, 1., 1., None) == 3


def test_jk_build_tasks_validates_on_host():
    '''The batched entry checks its own arguments and stops at the first
    task jk_build rejects, before any launch.'''
    from gpu4pyscf_amd.lib import clib
    lib = clib.load()
    assert ctypes.sizeof(clib.JKWindows) == 4 * 8 + 2 * 4
    assert ctypes.sizeof(clib.JKTask) == 14 * 4 + 6 * 8 + 2 * ctypes.sizeof(clib.JKWindows) + 8
    a = clib.JKBuildArgs()
    a.abi_version = clib.JK_ABI_VERSION
    tasks = (clib.JKTask * 2)()
    streams = (ctypes.c_void_p * 1)(None)
    assert lib.jk_build_tasks(ctypes.byref(a), tasks, 2, streams, 0) == 1
    assert b'jk_build_tasks' in lib.jk_last_error()
    tasks[1].li = clib.JK_LMAX + 1   # launched first (last task first)
    assert lib.jk_build_tasks(ctypes.byref(a), tasks, 2, streams, 1) == 2
    assert b'angular class' in lib.jk_last_error()
    assert lib.jk_build_tasks(ctypes.byref(a), tasks, 0, streams, 1) == 0


@pytest.mark.parametrize('W', [6, 10, 16])
def test_kl_window_table_partitions_the_pairs(W):
    '''The kl column windows: a stable reordering whose windows are
    contiguous, non-empty, within W AOs per group, and hold every pair's k
    and l AOs inside their columns.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    from gpu4pyscf_amd.scf import jk
    from oracle import oracle as O
    from tests.molecules import ATOM6
    mol = SortedGTO(M(atom=ATOM6, basis='def2-tzvp', unit='B'), diffuse_cutoff=0.3)
    cache = O.build_pairs(mol)
    aol = np.asarray(mol.ao_loc)
    nbas = mol.nbas
    n_windowed = 0
    for (k, l), (p, q) in cache.items():
        w = jk.kl_window_table(p, mol.group_offsets, aol, nbas, k, l, W)
        if w is None:
            assert len(p) == 0 or (aol[mol.group_offsets[k+1]] - aol[mol.group_offsets[k]] <= W and
                                   aol[mol.group_offsets[l+1]] - aol[mol.group_offsets[l]] <= W)
            continue
        n_windowed += 1
        order, tab, width = w
        assert sorted(order.tolist()) == list(range(len(p)))
        assert tab[0, 0] == 0 and tab[-1, 1] == len(p)
        assert np.all(tab[1:, 0] == tab[:-1, 1]) and np.all(tab[:, 1] > tab[:, 0])
        nk = (aol[mol.group_offsets[k]+1] - aol[mol.group_offsets[k]])
        nl = (aol[mol.group_offsets[l]+1] - aol[mol.group_offsets[l]])
        assert np.all(tab[:, 3] <= max(W, nk)) and np.all(tab[:, 5] <= max(W, nl))
        assert width == int((tab[:, 3] + tab[:, 5]).max())
        po = p[order].astype(np.int64)
        for f, e, k0, kn, l0, ln in tab:
            ks, ls = po[f:e] // nbas, po[f:e] % nbas
            assert np.all(aol[ks] >= k0) and np.all(aol[ks + 1] <= k0 + kn)
            assert np.all(aol[ls] >= l0) and np.all(aol[ls + 1] <= l0 + ln)
            # stable: the window keeps the list's order (descending q bins)
            assert np.all(np.diff(order[f:e]) > 0)
    assert n_windowed > 0


def test_tile_pairs_and_bins_match_oracle():
    from gpu4pyscf_amd.scf import jk
    from oracle import oracle as O
    for (i0, i1, j0, j1, nb) in [(0, 30, 0, 30, 40), (5, 31, 2, 9, 40), (0, 13, 0, 25, 25)]:
        assert np.array_equal(jk.tile_pairs(i0, i1, j0, j1, nb, 12),
                              O.tile_pairs(i0, i1, j0, j1, nb, 12))
    rng = np.random.default_rng(0)
    q = rng.uniform(-40, 3, 500).astype(np.float32)
    split = np.arange(math.log(1e-13), 2., 4.)
    assert np.array_equal(jk.group_by_split_points(q, split), O.group_by_split_points(q, split))


def test_sorted_basis_invariants():
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom='O 0 0 0; H 0 1 0; C 1 0 0; H 0 0 1', basis='def2-tzvp')
    sm = SortedGTO(mol)
    l = sm._bas[:, 1]
    np_ = sm._bas[:, 2]
    key = l * 100 - np_
    assert np.all(np.diff(key) >= 0)          # sorted by (l, -nprim)
    assert sm.coeff.shape == (sm.nao, mol.nao)
    # every original AO maps to at least one sorted Cartesian AO
    assert np.all(abs(sm.coeff).sum(axis=0) > 0)


@pytest.mark.parametrize('shape,spacing,basis,nao', [
    ((2, 2, 2), 3.0, 'def2-svp', 192),
    ((4, 4, 2), 3.0, 'def2-tzvp', 1376),
    ((4, 4, 4), 3.1, 'def2-tzvp', 2752),
])
def test_water_cluster_sizes(shape, spacing, basis, nao):
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    atoms = G.water_cluster(shape, spacing, seed=0)
    assert len(atoms) == 3 * np.prod(shape)
    for k in range(0, len(atoms), 3):
        o, h1, h2 = atoms[k][1], atoms[k+1][1], atoms[k+2][1]
        assert abs(np.linalg.norm(h1 - o) - 0.9572) < 1e-12
        assert abs(np.linalg.norm(h2 - o) - 0.9572) < 1e-12
        c = np.dot(h1 - o, h2 - o) / 0.9572**2
        assert abs(np.degrees(np.arccos(c)) - 104.52) < 1e-9
    mol = M(atom=G.to_string(atoms), basis=basis)
    assert mol.nao == nao
    # deterministic
    assert G.to_string(atoms) == G.to_string(G.water_cluster(shape, spacing, seed=0))


def test_alkane_geometry():
    from gpu4pyscf_amd.gto import geometry as G
    atoms = G.alkane(30, seed=0, sigma=0.)
    sym = [a[0] for a in atoms]
    assert sym.count('C') == 30 and sym.count('H') == 62
    C = np.array([a[1] for a in atoms if a[0] == 'C'])
    H = np.array([a[1] for a in atoms if a[0] == 'H'])
    dcc = np.linalg.norm(C[1:] - C[:-1], axis=1)
    assert abs(dcc - 1.54).max() < 1e-12
    # every H is 1.09 from its nearest C and carbons are tetrahedral
    dch = np.min(np.linalg.norm(H[:, None] - C[None], axis=2), axis=1)
    assert abs(dch - 1.09).max() < 1e-12
    v1, v2 = C[0] - C[1], C[2] - C[1]
    ang = np.degrees(np.arccos(np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)))
    assert abs(ang - 109.4712206) < 1e-6
    from gpu4pyscf_amd.gto.mole import M
    mol = M(atom=G.to_string(G.alkane(30, seed=0)), basis='def2-tzvp')
    assert mol.nao == 1302


def test_cart2sph_pins():
    '''Coefficients quoted in gpu4pyscf/lib/cupy_helper/cart2sph.cu:38-95.'''
    from gpu4pyscf_amd.gto.cart2sph import c2s
    assert abs(c2s(2)[0, 1] - 1.092548430592079070) < 1e-15
    assert abs(c2s(2)[2, 5] - 0.630783130505040012) < 1e-15
    assert abs(c2s(3)[3, 9] - 0.746352665180230782) < 1e-15
    assert abs(c2s(3)[6, 3] + 1.770130769779930530) < 1e-15
    assert abs(c2s(4)[4, 14] - 0.846284375321634430) < 1e-15
    assert abs(c2s(4)[6, 10] - 0.473087347878780009) < 1e-15
    # orthonormal real harmonics: integrate products over the sphere
    from math import factorial
    def sphere_moment(a, b, c):
        if a % 2 or b % 2 or c % 2:
            return 0.
        from scipy.special import gamma
        return 2 * gamma((a+1)/2) * gamma((b+1)/2) * gamma((c+1)/2) / gamma((a+b+c+3)/2)
    from gpu4pyscf_amd.gto.cart2sph import cart_components
    for l in (2, 3, 4, 5):
        comp = cart_components(l)
        M_ = np.array([[sphere_moment(*(np.add(p, q))) for q in comp] for p in comp])
        G_ = c2s(l) @ M_ @ c2s(l).T
        assert abs(G_ - np.eye(2*l+1)).max() < 1e-12


def test_fill_task_exchanges_bra_and_ket_of_xs_xp_tasks():
    '''_VHFOpt._fill_task: the (ps|pp), (ds|dp), (fs|fp) tasks whose group
    pairs differ get the fields of (k, l, i, j) and swap_ijkl = 1; every other
    task keeps its own fields; swap_tasks = False exchanges none.'''
    import torch
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    from gpu4pyscf_amd.lib import clib
    from gpu4pyscf_amd.scf import jk
    from tests.molecules import ATOM6
    sm = SortedGTO(M(atom=ATOM6, basis='def2-tzvp', unit='B'), diffuse_cutoff=0.3)
    ng = len(sm.l_ctr_counts)
    keys = [(i, j) for i in range(ng) for j in range(i + 1)]
    d = {'pairs': {k: torch.zeros(3 + n, dtype=torch.int32) for n, k in enumerate(keys)},
         'q': {k: torch.zeros(3, dtype=torch.float32) for k in keys},
         'pdata': {k: torch.zeros(6, dtype=torch.float64) for k in keys},
         'kl_win': {}}
    opt = jk._VHFOpt.__new__(jk._VHFOpt)
    opt.sorted_mol = sm
    ls = [int(x) for x in sm.uniq_l_ctr[:, 0]]
    n_swapped = 0
    for swap_tasks in (True, False):
        opt.swap_tasks = swap_tasks
        for (i, j) in keys:
            for (k, l) in keys:
                if k > i or (k == i and l > j):
                    continue
                a = clib.JKTask()
                opt._fill_task(a, d, (i, j, k, l))
                want = (swap_tasks and (i, j) != (k, l) and ls[i] == ls[k] and ls[j] == 0 and
                        ls[l] == 1)
                assert a.swap_ijkl == int(want)
                bra, ket = ((k, l), (i, j)) if want else ((i, j), (k, l))
                assert (a.li, a.lj, a.lk, a.ll) == (ls[bra[0]], ls[bra[1]], ls[ket[0]], ls[ket[1]])
                assert a.pairs_ij == d['pairs'][bra].data_ptr()
                assert a.pairs_kl == d['pairs'][ket].data_ptr()
                assert a.npairs_ij == len(d['pairs'][bra])
                assert a.k_ao0 == sm.ao_loc[sm.group_offsets[ket[0]]]
                n_swapped += want
    assert n_swapped > 0


def test_csr_rows_restore_the_dense_matrix():
    '''The compressed rows behind the sparse AO transforms hold exactly the
    nonzeros of the coefficient matrix, row by row, for a split and
    spherical def2-TZVP basis and its transpose.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    from gpu4pyscf_amd.scf.jk import csr_rows
    from tests.molecules import ATOM6
    C = SortedGTO(M(atom=ATOM6, basis='def2-tzvp')).coeff
    for mat in (C, C.T, np.zeros((3, 4))):
        ptr, idx, coef = csr_rows(mat)
        assert ptr.dtype == np.int32 and idx.dtype == np.int32 and coef.dtype == np.float64
        assert len(ptr) == mat.shape[0] + 1 and ptr[0] == 0 and ptr[-1] == len(idx)
        back = np.zeros_like(mat)
        for i in range(mat.shape[0]):
            back[i, idx[ptr[i]:ptr[i + 1]]] = coef[ptr[i]:ptr[i + 1]]
        assert np.array_equal(back, mat)
        assert np.count_nonzero(mat) == len(idx)
