'''GPU parity of the J/K engine against the CPU oracle and the reference's
golden fingerprints (gpu4pyscf/scf/tests/test_scf_jk.py, test_rhf.py).'''
import math
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.molecules import case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GPU_CASES = ['tzvp_hermi1', 'tzvp_cart', 'tzvp_hermi0', 'sto3g_hermi2']


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


@pytest.mark.parametrize('name', GPU_CASES)
def test_jk_matches_oracle_and_golden(name):
    from gpu4pyscf_amd.scf import jk
    mol, dm, hermi, fpj, fpk, tol = case(name)
    print(f'[case {name}] nao={mol.nao}', flush=True)
    vj, vk = jk.get_jk(mol, dm, hermi=hermi)
    rj, rk = O.get_jk(mol, dm, hermi=hermi)
    # J/K to 1e-10 absolute against the oracle
    assert abs(vj - rj).max() < 1e-10
    assert abs(vk - rk).max() < 1e-10
    if fpj is not None:
        assert abs(O.fp(vj) - fpj) < tol
    if fpk is not None:
        assert abs(O.fp(vk) - fpk) < tol


def _water8_svp():
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    mol = M(atom=G.to_string(G.water_cluster((2, 2, 2), 3.0, seed=0)), basis='def2-svp')
    rng = np.random.default_rng(0)
    c = np.linalg.qr(rng.standard_normal((mol.nao, 40)))[0]
    return mol, 2 * c @ c.T


def test_water8_svp_jk_and_quartet_set():
    '''8-water def2-SVP cluster (192 AOs): J/K and the surviving-quartet
    count against the oracle's restated predicate on the same q_cond/dm_cond.'''
    import torch
    from gpu4pyscf_amd.scf import jk
    mol, dm = _water8_svp()
    assert mol.nao == 192
    opt = jk._VHFOpt(mol).build()
    sm = opt.sorted_mol
    # screening setup equals the oracle's restatement
    ref_cache = O.build_pairs(sm, opt.direct_scf_tol)
    for key, (p, q, _) in opt.bas_pair_cache.items():
        rp, rq = ref_cache[key]
        assert np.array_equal(p, rp), key
        assert abs(q - rq).max() < 1e-5 if len(q) else True
    dm_t = torch.as_tensor(dm, device='cuda')
    dms = opt.apply_coeff_C_mat_CT(dm_t[None])
    vj, vk = opt.get_jk(dms, 1, count_quartets=True)
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    rj, rk, nsurv, _ = O.jk_sorted(sm, dms.cpu().numpy(), 1, cache, opt.direct_scf_tol)
    assert opt.last_quartets == nsurv
    assert abs(vj.cpu().numpy() - rj).max() < 1e-10
    assert abs(vk.cpu().numpy() - rk).max() < 1e-10


def _gpu_task_sets(opt):
    return {t: (c, opt.last_task_hashes[t])
            for t, c in opt.last_task_counts.items() if c}


def test_water8_svp_quartet_set_checksums():
    '''Per class task, the surviving-quartet count and set checksum of the
    kernels equal the oracle's scan of the restated predicate.'''
    import torch
    from gpu4pyscf_amd.scf import jk
    mol, dm = _water8_svp()
    opt = jk._VHFOpt(mol).build()
    dms = opt.apply_coeff_C_mat_CT(torch.as_tensor(dm, device='cuda')[None])
    opt.get_jk(dms, 1, count_quartets=True, hash_quartets=True)
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    ref = O.quartet_sets(opt.sorted_mol, dms.cpu().numpy(), 1, cache, opt.direct_scf_tol)
    assert _gpu_task_sets(opt) == ref


def test_water32_tzvp_full_size_quartet_set():
    '''The bench workload at full size (32 waters, def2-TZVP, 1376 AOs, one
    incremental build on dD = 1e-3 (G + G^T)/2): screening pair lists equal
    the oracle's, and per class task the surviving quartets (count and set
    checksum, 1.39e9 quartets) equal the oracle's full scan, which has no
    early exit, in FP64 and in mixed precision.'''
    import torch
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom=G.to_string(G.water_cluster((4, 4, 2), 3.0, seed=0)), basis='def2-tzvp')
    assert mol.nao == 1376
    print('[water32 def2-TZVP] screening setup', flush=True)
    opt = jk._VHFOpt(mol).build()
    ref_cache = O.build_pairs(opt.sorted_mol, opt.direct_scf_tol)
    for key, (p, q, _) in opt.bas_pair_cache.items():
        rp, rq = ref_cache[key]
        assert np.array_equal(p, rp), key
    rng = np.random.default_rng(1)
    g = rng.standard_normal((mol.nao, mol.nao))
    ddm = torch.as_tensor(1e-3 * (g + g.T) / 2, device='cuda')
    dms = opt.apply_coeff_C_mat_CT(ddm[None])
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    print('[water32 def2-TZVP] oracle scan', flush=True)
    ref = O.quartet_sets(opt.sorted_mol, dms.cpu().numpy(), 1, cache, opt.direct_scf_tol)
    assert sum(c for c, _ in ref.values()) == 1389051040
    for tol in (None, 1e-6):
        print(f'[water32 def2-TZVP] get_jk fp32_tol={tol}', flush=True)
        opt.get_jk(dms, 1, count_quartets=True, hash_quartets=True, fp32_tol=tol)
        assert _gpu_task_sets(opt) == ref


def test_mixed_precision_incremental_jk():
    '''FP32 quartets on an incremental density change: error against the
    FP64 build stays below the bound implied by the threshold, and a large
    fraction of quartets runs in FP32.'''
    import torch
    from gpu4pyscf_amd.scf import jk
    mol, dm = _water8_svp()
    opt = jk._VHFOpt(mol).build()
    rng = np.random.default_rng(5)
    g = rng.standard_normal((mol.nao, mol.nao))
    ddm = torch.as_tensor(1e-3 * (g + g.T) / 2, device='cuda')
    dms = opt.apply_coeff_C_mat_CT(ddm[None])
    vj0, vk0 = opt.get_jk(dms, 1, count_quartets=True)
    n0 = opt.last_quartets
    assert opt.last_quartets_fp32 == 0
    vj1, vk1 = opt.get_jk(dms, 1, count_quartets=True, fp32_tol=1e-6)
    assert opt.last_quartets == n0
    frac = opt.last_quartets_fp32 / n0
    assert frac > 0.2
    err = max(float((vj1 - vj0).abs().max()), float((vk1 - vk0).abs().max()))
    print(f'fp32 fraction {frac:.3f}, max |dJ|,|dK| = {err:.2e}')
    assert err < 1e-9


def test_alkane_tzvp_jk_matches_oracle():
    '''C4H10 (all-trans, jittered) in def2-TZVP: the C and H shells of the
    alkane workload, J/K of a random symmetric density against the oracle.'''
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom=G.to_string(G.alkane(4, seed=1)), basis='def2-tzvp')
    rng = np.random.default_rng(4)
    c = np.linalg.qr(rng.standard_normal((mol.nao, mol.nelectron // 2)))[0]
    dm = 2 * c @ c.T
    print(f'[alkane C4H10] nao={mol.nao}', flush=True)
    vj, vk = jk.get_jk(mol, dm, hermi=1)
    rj, rk = O.get_jk(mol, dm, hermi=1)
    assert abs(vj - rj).max() < 1e-10
    assert abs(vk - rk).max() < 1e-10


def test_boundary_j_only_k_only_torch_and_batched():
    '''get_j / get_k (J-only and K-only launches), torch in -> torch out, and
    a batch of two symmetric densities with hermi=1, against the oracle.'''
    import torch
    from gpu4pyscf_amd.scf import jk
    mol, dm, _, _, _, _ = case('tzvp_hermi1')
    opt = jk._VHFOpt(mol).build()
    rj, rk = O.get_jk(mol, dm, hermi=1)
    vj = jk.get_j(mol, dm, hermi=1, vhfopt=opt)
    vk = jk.get_k(mol, dm, hermi=1, vhfopt=opt)
    assert abs(vj - rj).max() < 1e-10
    assert abs(vk - rk).max() < 1e-10
    vj_t, vk_t = jk.get_jk(mol, torch.as_tensor(dm, device='cuda'), hermi=1, vhfopt=opt)
    assert isinstance(vj_t, torch.Tensor) and vj_t.device.type == 'cuda'
    assert abs(vj_t.cpu().numpy() - rj).max() < 1e-10
    assert abs(vk_t.cpu().numpy() - rk).max() < 1e-10
    rng = np.random.default_rng(6)
    g = rng.standard_normal(dm.shape)
    dm2 = np.stack([dm, .1 * (g + g.T)])
    vj2, vk2 = jk.get_jk(mol, dm2, hermi=1, vhfopt=opt)
    assert vj2.shape == dm2.shape
    for n in range(2):
        rj, rk = O.get_jk(mol, dm2[n], hermi=1)
        assert abs(vj2[n] - rj).max() < 1e-10
        assert abs(vk2[n] - rk).max() < 1e-10


@pytest.mark.parametrize('screen,hermi', [('j', 1), ('k', 1), ('j', 0), ('k', 0)])
def test_one_sided_builds_match_oracle_sets_and_values(screen, hermi):
    '''J-only (get_j) and K-only (get_k) builds, screened on their own density
    blocks as the reference's RYS_build_j / RYS_build_k: per class task, the
    surviving quartets (count and set checksum) equal the oracle's
    exhaustive scan with the same screen, and J or K equal the oracle's to
    1e-10, on an 8-water def2-SVP cluster with a density that decays with
    distance (so that the screens keep different sets).'''
    import torch
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom=G.to_string(G.water_cluster((2, 2, 2), 3.0, seed=2)), basis='def2-svp')
    opt = jk._VHFOpt(mol).build()
    sm = opt.sorted_mol
    rng = np.random.default_rng(11)
    g = rng.standard_normal((mol.nao, mol.nao))
    sl = np.asarray(mol.aoslice_by_atom())
    cen = np.asarray(mol.atom_coords())[np.repeat(np.arange(mol.natm), sl[:, 3] - sl[:, 2])]
    r = np.linalg.norm(cen[:, None] - cen[None], axis=-1)
    dm = 1e-2 * g * np.exp(-.7 * r)
    if hermi == 1:
        dm = dm + dm.T
    print(f'[one-sided {screen} hermi={hermi}] nao={mol.nao}', flush=True)
    dms = opt.apply_coeff_C_mat_CT(torch.as_tensor(dm, device='cuda')[None])
    opt.get_jk(dms, hermi, with_j=screen == 'j', with_k=screen == 'k', screen=screen,
               count_quartets=True, hash_quartets=True)
    got = {t: (c, opt.last_task_hashes[t]) for t, c in opt.last_task_counts.items() if c}
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    ref = O.quartet_sets(sm, dms.cpu().numpy(), hermi, cache, opt.direct_scf_tol,
                         screen=screen)
    full = O.quartet_sets(sm, dms.cpu().numpy(), hermi, cache, opt.direct_scf_tol)
    assert got == ref
    n_one, n_full = sum(c for c, _ in ref.values()), sum(c for c, _ in full.values())
    print(f'[one-sided {screen} hermi={hermi}] {n_one} of {n_full} quartets', flush=True)
    assert n_one < n_full
    rj, rk = O.get_jk(mol, dm, hermi=hermi, with_j=screen == 'j', with_k=screen == 'k',
                      screen=screen)
    if screen == 'j':
        assert abs(jk.get_j(mol, dm, hermi=hermi, vhfopt=opt) - rj).max() < 1e-10
    else:
        assert abs(jk.get_k(mol, dm, hermi=hermi, vhfopt=opt) - rk).max() < 1e-10


@pytest.mark.parametrize('W,with_j', [(12, True), (40, True), (12, False)])
def test_kl_column_windows_match_oracle(W, with_j):
    '''kl column windows forced on every per-lane and team launch that can
    hold the window's K rows in LDS (kl_win_policy 1, windows of at most W
    AOs): the
    surviving quartets of every class task (count and set checksum) equal
    the oracle's, and J/K equal the oracle's to 1e-10, on the 8-water
    def2-SVP cluster (192 AOs; its widest groups have 64 AOs).'''
    import torch
    from gpu4pyscf_amd.scf import jk
    mol, dm = _water8_svp()
    opt = jk._VHFOpt(mol)
    opt.kl_window_ao = W
    opt.kl_window_ao_team = W
    opt.kl_win_policy = 1
    opt.build()
    assert len(opt._dev['kl_win']) > 0
    nwin = {k: len(v[0][3]) for k, v in opt._dev['kl_win'].items() if v[0] is not None}
    print(f'[kl windows W={W} with_j={with_j}] windows per group pair {nwin}', flush=True)
    sm = opt.sorted_mol
    dms = opt.apply_coeff_C_mat_CT(torch.as_tensor(dm, device='cuda')[None])
    screen = None if with_j else 'k'
    vj, vk = opt.get_jk(dms, 1, with_j=with_j, screen=screen, count_quartets=True,
                        hash_quartets=True)
    got = {t: (c, opt.last_task_hashes[t]) for t, c in opt.last_task_counts.items() if c}
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    ref = O.quartet_sets(sm, dms.cpu().numpy(), 1, cache, opt.direct_scf_tol,
                         screen=screen or 'jk')
    assert got == ref
    rj, rk = O.get_jk(mol, dm, hermi=1, with_j=with_j, screen=screen or 'jk')
    vk = opt.apply_coeff_CT_mat_C(vk)[0].cpu().numpy()
    assert abs(vk - rk).max() < 1e-10
    if with_j:
        vj = opt.apply_coeff_CT_mat_C(vj)[0].cpu().numpy()
        assert abs(vj - rj).max() < 1e-10
