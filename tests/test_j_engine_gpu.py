'''GPU parity of the McMurchie-Davidson J-engine (gpu4pyscf_amd/scf/j_engine.py,
csrc/jk_md.hip) against the oracle's J and the reference's golden
fingerprints (gpu4pyscf/scf/tests/test_scf_j_engine.py):

* def2-TZVP, full range and omega = +-0.2 (test_scf_j_engine.py:22-55);
* hermi = 0 (symmetrised), several densities, Cartesian shells;
* an inline basis with s, p and f primitives on a 17-atom peptide
  fragment (test_scf_j_engine.py:132-168), where the Hermite kernels and
  the Rys J-only build of the classes above them both contribute;
* g and h shells (only the Rys part can take an h pair);
* an 8-water def2-SVP cluster with a density decaying with distance.
'''
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.molecules import ATOM6, case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PEPTIDE = '''
O  -9.2037 -0.1259  6.4262
H -11.7768  0.2184  7.9561
H -11.7819 -1.0073  7.9636
H -11.2190 -0.1224  5.3389
N  -9.2130 -0.1182  8.6103
C  -7.7662 -0.1219  8.6103
C  -7.2447 -0.1180 10.0438
O  -7.9744 -0.1125 11.0321
H  -7.4164 -1.0206  8.0911
H  -7.4110  0.2317  8.0835
H  -9.6852 -0.1162  9.5099
N  -5.9251 -0.1205 10.2766
C  -5.4305 -0.1166 11.6362
C  -3.9051 -0.1205 11.6362
O  -3.2258 -0.1262 10.6126
H  -5.7987  0.2177 12.1423
H  -5.8042 -1.0067 12.1503
'''


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def _ref_j(mol, dm, hermi=1):
    return O.get_jk(mol, dm, hermi=hermi, with_k=False, screen='j')[0]


@pytest.mark.parametrize('omega,fpj,tol', [(0., -2327.4715195591784, 1e-9),
                                           (.2, 1163.932604635460, 5e-10),
                                           (-.2, -3491.404124194866, 1e-9)])
def test_j_engine_tzvp_golden(omega, fpj, tol):
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import j_engine
    mol = M(atom=ATOM6, basis='def2-tzvp', unit='B')
    mol.omega = omega
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao)
    dm = dm.dot(dm.T)
    print(f'[j-engine tzvp omega={omega}] nao={mol.nao}', flush=True)
    opt = j_engine._VHFOpt(mol).build()
    vj = j_engine.get_j(mol, dm, vhfopt=opt)
    ref = _ref_j(mol, dm)
    print(f'[j-engine tzvp omega={omega}] max|dJ| {abs(vj - ref).max():.2e}, '
          f'fp {O.fp(vj):.12f}', flush=True)
    assert abs(vj - ref).max() < 1e-10 * max(1., abs(ref).max())
    assert abs(O.fp(vj) - fpj) < tol
    # both parts ran: Hermite classes and the Rys J-only classes above them
    assert opt.md_tasks() and opt.tasks()


def test_j_engine_hermi0_batched_cart():
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import j_engine
    mol, dm, hermi, _, _, _ = case('tzvp_hermi0')
    vj = j_engine.get_j(mol, dm, hermi=hermi)
    assert abs(vj - _ref_j(mol, dm, hermi=0)).max() < 1e-10 * max(1., abs(vj).max())
    mol, dm, _, _, _, _ = case('tzvp_g_2dm')
    vj = j_engine.get_j(mol, dm, hermi=1)
    assert vj.shape == dm.shape
    for n in range(2):
        assert abs(vj[n] - _ref_j(mol, dm[n])).max() < 1e-10 * max(1., abs(vj[n]).max())
    mol, dm, _, _, _, _ = case('tzvp_cart')
    vj = j_engine.get_j(mol, dm, hermi=1)
    assert abs(vj - _ref_j(mol, dm)).max() < 1e-10


def test_j_engine_integral_screen_peptide():
    '''test_scf_j_engine.py:132-168: even-tempered s, p, f primitives.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import j_engine
    basis = ([[0, [2.**x, 1]] for x in range(-1, 5)] +
             [[1, [2.**x, 1]] for x in range(-1, 2)] +
             [[3, [2.**x, 1]] for x in range(-1, 1)])
    mol = M(atom=PEPTIDE, basis=basis, unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao) * .1 - .05
    dm = dm.dot(dm.T)
    print(f'[j-engine peptide] nao={mol.nao}', flush=True)
    vj = j_engine.get_j(mol, dm)
    ref = _ref_j(mol, dm)
    assert abs(vj - ref).max() < 1e-10 * max(1., abs(ref).max())


def test_j_engine_g_and_h_shells():
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import j_engine
    for name in ('tzvp_g', 'mol1_hermi1'):
        mol, dm, _, _, _, _ = case(name)
        dm = dm if dm.ndim == 2 else dm[0]
        print(f'[j-engine {name}] nao={mol.nao}', flush=True)
        vj = j_engine.get_j(mol, dm)
        ref = _ref_j(mol, dm)
        assert abs(vj - ref).max() < 1e-10 * max(1., abs(ref).max())


def test_j_engine_water8_local_density():
    import torch
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import j_engine
    mol = M(atom=G.to_string(G.water_cluster((2, 2, 2), 3.0, seed=2)), basis='def2-svp')
    sl = np.asarray(mol.aoslice_by_atom())
    cen = np.asarray(mol.atom_coords())[np.repeat(np.arange(mol.natm), sl[:, 3] - sl[:, 2])]
    r = np.linalg.norm(cen[:, None] - cen[None], axis=-1)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((mol.nao, mol.nao))
    dm = 1e-2 * (g + g.T) * np.exp(-.7 * r)
    opt = j_engine._VHFOpt(mol).build()
    dms = opt.apply_coeff_C_mat_CT(torch.as_tensor(dm, device='cuda')[None])
    vj = opt.apply_coeff_CT_mat_C(opt.get_j(dms, 1, count_quartets=True))[0].cpu().numpy()
    # the Hermite kernels' survivors are the oracle's J-only set of the
    # classes they cover (count and set checksum over canonical quartets)
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    ref = O.quartet_sets(opt.sorted_mol, dms.cpu().numpy(), 1, cache, opt.direct_scf_tol,
                         screen='j')
    ls = opt.sorted_mol.uniq_l_ctr[:, 0]
    md = [v for t, v in ref.items() if sum(int(ls[x]) for x in t) <= j_engine.MD_LMAX]
    n_ref = sum(c for c, _ in md)
    h_ref = sum(h for _, h in md) & (2**64 - 1)
    print(f'[j-engine water8] {opt.last_md_quartets} Hermite-class survivors '
          f'(oracle {n_ref})', flush=True)
    assert opt.last_md_quartets == n_ref > 0
    assert opt.last_md_hash == h_ref
    assert abs(vj - _ref_j(mol, dm)).max() < 1e-10
