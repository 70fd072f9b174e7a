'''GPU parity on the benchmark configurations at full size:

* 8-water def2-SVP (192 AOs): RHF to 1e-9 Eh against the oracle's RHF;
* 32-water def2-TZVP (1376 AOs): J and K of the incremental build against
  the oracle's values (tests/golden/water32_tzvp_incremental_jk.npz, written
  by tests/golden/make_water32_jk.py);
* all-trans C30H62 def2-TZVP (1302 AOs) and 64-water def2-TZVP (2752 AOs):
  per class task, the surviving quartets (count and SplitMix64 set checksum)
  of an incremental build against the oracle's exhaustive scan, in FP64 and
  in mixed precision;
* the launch variants of get_jk (vj/vk copies per XCD, per-launch timing,
  one stream) against the default build.
'''
import os

import numpy as np
import pytest

from tests.conftest import has_gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def _water(shape, spacing, basis, seed=0):
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    return M(atom=G.to_string(G.water_cluster(shape, spacing, seed=seed)), basis=basis)


def _incremental_dm(nao, device='cuda'):
    import torch
    rng = np.random.default_rng(1)
    g = rng.standard_normal((nao, nao))
    return torch.as_tensor(1e-3 * (g + g.T) / 2, device=device)


def test_water8_svp_rhf_energy_matches_oracle():
    '''8-water def2-SVP cluster: the GPU RHF (incremental builds, CDIIS) and
    the oracle's RHF from the same atomic guess agree to 1e-9 Eh.'''
    from gpu4pyscf_amd.scf.hf import RHF
    from gpu4pyscf_amd.scf.atom_guess import atom_guess
    from oracle import scf_oracle
    mol = _water((2, 2, 2), 3.0, 'def2-svp')
    assert mol.nao == 192
    mf = RHF(mol)
    mf.conv_tol = 1e-10
    e = mf.kernel()
    assert mf.converged
    print(f'[water8 def2-SVP] GPU RHF E = {e:.12f} in {mf.cycles} cycles', flush=True)
    e_ref, _ = scf_oracle.rhf(mol, conv_tol=1e-10, dm0=atom_guess(mol))
    print(f'[water8 def2-SVP] oracle RHF E = {e_ref:.12f}, |dE| = {abs(e - e_ref):.2e}',
          flush=True)
    assert abs(e - e_ref) < 1e-9


def test_water32_tzvp_jk_values_match_oracle():
    '''32-water def2-TZVP incremental build: ten full rows of J and K to
    1e-10 absolute and the fingerprints of the whole matrices against the
    oracle's values.'''
    from gpu4pyscf_amd.scf import jk
    ref = np.load(os.path.join(HERE, 'golden', 'water32_tzvp_incremental_jk.npz'))
    mol = _water((4, 4, 2), 3.0, 'def2-tzvp')
    ddm = _incremental_dm(mol.nao)
    vj, vk = jk.get_jk(mol, ddm, hermi=1)
    vj = vj.cpu().numpy(); vk = vk.cpu().numpy()
    rows = ref['rows']
    ej = abs(vj[rows] - ref['vj_rows']).max()
    ek = abs(vk[rows] - ref['vk_rows']).max()
    print(f'[water32 def2-TZVP] max |dJ| {ej:.2e}, |dK| {ek:.2e} on {len(rows)} rows; '
          f'fp(J) {O.fp(vj):.12f} vs {float(ref["fp_vj"]):.12f}', flush=True)
    assert ej < 1e-10 and ek < 1e-10
    # fp sums n^2 = 1.9e6 cos-weighted elements: 1e-10 per element bounds it by 2e-4
    assert abs(O.fp(vj) - float(ref['fp_vj'])) < 1e-6
    assert abs(O.fp(vk) - float(ref['fp_vk'])) < 1e-6
    assert abs(np.linalg.norm(vj) - float(ref['norm_vj'])) < 1e-9
    assert abs(np.linalg.norm(vk) - float(ref['norm_vk'])) < 1e-9


def _check_full_size_sets(mol, label):
    from gpu4pyscf_amd.scf import jk
    print(f'[{label}] screening setup, nao={mol.nao}', flush=True)
    opt = jk._VHFOpt(mol).build()
    ref_cache = O.build_pairs(opt.sorted_mol, opt.direct_scf_tol)
    for key, (p, q, _) in opt.bas_pair_cache.items():
        assert np.array_equal(p, ref_cache[key][0]), key
    dms = opt.apply_coeff_C_mat_CT(_incremental_dm(mol.nao)[None])
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    print(f'[{label}] oracle scan', flush=True)
    ref = O.quartet_sets(opt.sorted_mol, dms.cpu().numpy(), 1, cache, opt.direct_scf_tol)
    total = sum(c for c, _ in ref.values())
    for tol in (None, 1e-6):
        print(f'[{label}] get_jk fp32_tol={tol}', flush=True)
        opt.get_jk(dms, 1, count_quartets=True, hash_quartets=True, fp32_tol=tol)
        got = {t: (c, opt.last_task_hashes[t]) for t, c in opt.last_task_counts.items() if c}
        assert got == ref
    return total


def test_alkane30_tzvp_full_size_quartet_set():
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    mol = M(atom=G.to_string(G.alkane(30, seed=0)), basis='def2-tzvp')
    assert mol.nao == 1302
    total = _check_full_size_sets(mol, 'C30H62 def2-TZVP')
    print(f'[C30H62 def2-TZVP] {total} surviving quartets', flush=True)
    assert total > 5e8


def test_water64_tzvp_full_size_quartet_set():
    mol = _water((4, 4, 4), 3.1, 'def2-tzvp')
    assert mol.nao == 2752
    total = _check_full_size_sets(mol, 'water64 def2-TZVP')
    print(f'[water64 def2-TZVP] {total} surviving quartets', flush=True)
    assert total > 5e9


def test_launch_variants_agree():
    '''vj/vk copies per XCD (1, 3, 8, capped), per-launch timing, profiling
    mode and a single stream give the default build's J/K (up to the order
    of the atomic additions).'''
    import torch
    from gpu4pyscf_amd.scf import jk
    mol = _water((2, 2, 2), 3.0, 'def2-svp')
    opt = jk._VHFOpt(mol).build()
    dms = opt.apply_coeff_C_mat_CT(_incremental_dm(mol.nao)[None])
    vj0, vk0 = opt.get_jk(dms, 1)
    scale = float(vk0.abs().max())

    def same(vj, vk):
        return (float((vj - vj0).abs().max()) < 1e-13 * scale and
                float((vk - vk0).abs().max()) < 1e-13 * scale)

    for n in (1, 3, 8):
        opt.n_repl = n
        assert same(*opt.get_jk(dms, 1)), n
    opt.n_repl = 8
    vj, vk = opt.get_jk(dms, 1, time_classes=True)
    assert same(vj, vk)
    assert len(opt.last_task_ms) == len(opt.tasks())
    assert all(ms >= 0 for _, ms in opt.last_task_ms)
    opt.profile = True
    vj, vk = opt.get_jk(dms, 1)
    prof = opt.profile_results()
    opt.profile = False
    assert same(vj, vk)
    assert set(prof) == set(opt.tasks())
    opt.nstreams = 1
    assert same(*opt.get_jk(dms, 1))
    opt.nstreams = jk.NSTREAMS
    torch.cuda.synchronize()
