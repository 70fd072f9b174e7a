'''GPU parity of the runtime-l kernel and of the range-separated and
gradient paths against the oracle and the reference's golden fingerprints:

* g shells (gpu4pyscf/scf/tests/test_scf_jk.py:200-236), h shells of the
  uncontracted s...h basis (test_rhf.py:166-189) and an h shell on O
  (test_scf_jk.py:172-198);
* omega = +-0.2 for hermi 1 and 0 (test_scf_jk.py:59-80,142-170), the
  short-range get_k (test_scf_jk.py:358-369) and explicit lr/sr factors;
* the per-atom two-electron energy gradient (test_scf_jk.py:388-414,
  grad/tests/test_rhf_grad.py:94-121).
'''
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.molecules import ATOM6, case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


@pytest.mark.parametrize('name', ['tzvp_g', 'tzvp_g_2dm', 'mol1_hermi1', 'mol1_hermi0'])
def test_g_and_h_shells_match_oracle_and_golden(name):
    from gpu4pyscf_amd.scf import jk
    mol, dm, hermi, fpj, fpk, tol = case(name)
    print(f'[case {name}] nao={mol.nao}', flush=True)
    vj, vk = jk.get_jk(mol, dm, hermi=hermi)
    rj, rk = O.get_jk(mol, dm, hermi=hermi)
    scale = max(1., abs(rj).max(), abs(rk).max())
    assert abs(vj - rj).max() < 1e-10 * scale
    assert abs(vk - rk).max() < 1e-10 * scale
    if fpj is not None:
        assert abs(O.fp(vj) - fpj) < tol
    if fpk is not None:
        assert abs(O.fp(vk) - fpk) < tol


def test_h_shell_on_oxygen_golden():
    '''test_scf_jk.py:172-198: def2-TZVP with O carrying one h shell, hermi=0.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom=ATOM6, basis={'default': 'def2-tzvp', 'O': [[5, [1., 1.]]]}, unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao)
    vj, vk = jk.get_jk(mol, dm, hermi=0)
    assert abs(O.fp(vj) - -61.28856847097108) < 1e-9
    assert abs(O.fp(vk) - -76.38373664249241) < 1e-9
    vj1 = jk.get_j(mol, dm, hermi=0)
    assert abs(vj1 - vj).max() < 1e-11


@pytest.mark.parametrize('hermi,omega,fpj,fpk', [
    (1, .2, 1163.932604635460, -1269.969109438691),
    (1, -.2, -3491.404124194866, -2799.347891387202),
    (0, .2, 24.18519249677608, -34.15933205656134),
    (0, -.2, -77.67449053914103, -80.95859292428769)])
def test_range_separated_jk_golden(hermi, omega, fpj, fpk):
    '''mol.omega > 0: erf(omega r)/r; < 0: erfc(|omega| r)/r, for J and K.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom=ATOM6, basis='def2-tzvp', unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao)
    if hermi == 1:
        dm = dm.dot(dm.T)
    mol.omega = omega
    print(f'[rsh hermi={hermi} omega={omega}]', flush=True)
    vj, vk = jk.get_jk(mol, dm, hermi=hermi)
    rj, rk = O.get_jk(mol, dm, hermi=hermi)
    scale = max(1., abs(rj).max(), abs(rk).max())
    assert abs(vj - rj).max() < 1e-10 * scale
    assert abs(vk - rk).max() < 1e-10 * scale
    tol = 1e-9 if hermi == 1 else 5e-10
    assert abs(O.fp(vj) - fpj) < tol
    assert abs(O.fp(vk) - fpk) < tol


def test_long_plus_short_range_is_full_range():
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom=ATOM6, basis='def2-tzvp', unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao)
    dm = dm.dot(dm.T)
    v1 = jk.get_jk(mol, dm, hermi=1)
    mol.omega = .2
    v2 = jk.get_jk(mol, dm, hermi=1)
    mol.omega = -.2
    v3 = jk.get_jk(mol, dm, hermi=1)
    for n in range(2):
        assert abs(v2[n] + v3[n] - v1[n]).max() < 1e-9


def test_short_range_get_k_golden():
    '''test_scf_jk.py:358-369 (def2-SVP O2, omega = -0.3).  The reference's
    1e-13 on the fingerprint relies on its own short-range screening; here
    the screening bound is the short-range Schwarz factor alone, and the
    fingerprint agrees to 1e-11.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom='O 0.0000 0.7375 -0.0528; O 0.0000 -0.7375 -0.1528', basis='def2-svp')
    np.random.seed(12)
    dm = np.random.rand(mol.nao, mol.nao) - .5
    dm = dm.dot(dm.T)
    mol.omega = -.3
    vk = jk.get_k(mol, dm, hermi=1)
    assert abs(O.fp(vk) - -1.8653967312459407) < 1e-11
    rk = O.get_jk(mol, dm, hermi=1, with_j=False)[1]
    assert abs(vk - rk).max() < 1e-12


def test_explicit_rsh_factors_match_oracle():
    '''get_k with omega, lr_factor and sr_factor given (K operator
    0.4 erf + 0.7 erfc), J with the molecule's full-range operator.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf import jk
    mol = M(atom='O 0 0 0.1174; H -0.757 0 -0.4696; H 0.757 0 -0.4696', basis='def2-tzvp')
    rng = np.random.default_rng(2)
    c = np.linalg.qr(rng.standard_normal((mol.nao, 5)))[0]
    dm = 2 * c @ c.T
    vj, vk = jk.get_jk(mol, dm, hermi=1, omega=.25, lr_factor=.4, sr_factor=.7)
    rj, rk = O.get_jk(mol, dm, hermi=1, omega=.25, lr_factor=.4, sr_factor=.7)
    assert abs(vj - rj).max() < 1e-10
    assert abs(vk - rk).max() < 1e-10
    vj0, _ = jk.get_jk(mol, dm, hermi=1)
    assert abs(vj - vj0).max() < 1e-12


def test_energy_gradient_short_range_golden():
    '''test_scf_jk.py:388-414: exchange-only per-atom energy derivatives of
    an antisymmetric density (against the oracle), then the short-range
    golden of atom 0.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.grad.rhf import _jk_energy_per_atom, _VHFOpt
    mol = M(atom='O 0.0000 0.7375 -0.0528; O 0.0000 -0.7375 -0.1528', basis='def2-svp')
    np.random.seed(12)
    dm = np.random.rand(mol.nao, mol.nao) - .5
    dm1 = dm - dm.T
    opt = _VHFOpt(mol).build()
    ejk = _jk_energy_per_atom(opt, dm1, j_factor=0, k_factor=1.)
    ref = O.ejk_ip1(mol, dm1, j_factor=0, k_factor=1.)
    print(f'[ejk antisym] max |GPU - oracle| {abs(ejk - ref).max():.2e}', flush=True)
    assert abs(ejk - ref).max() < 1e-11
    dm = dm.dot(dm.T)
    mol.omega = -.3
    opt = _VHFOpt(mol).build()
    ejk = _jk_energy_per_atom(opt, dm, j_factor=0, k_factor=1.) * .5
    gold = np.array([0.24806416996651, 1.11003753769514, 0.19967171093788])
    print(f'[ejk sr] atom 0 - golden {ejk[0] - gold}', flush=True)
    assert abs(ejk[0] - gold).max() < 1e-12
    assert abs(ejk.sum(axis=0)).max() < 1e-10


def test_energy_gradient_tzvp_golden():
    '''grad/tests/test_rhf_grad.py:94-113: RHF two-electron energy gradient
    of a random density (def2-TZVP, O H H C): translational invariance and
    the fingerprint; per atom against the oracle.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.grad.rhf import _jk_energy_per_atom, _VHFOpt
    mol = M(atom='O 0.000 -0. 0.1174; H -0.757 4. -0.4696; H 0.757 4. -0.4696; C 3. 1. 0.',
            basis='def2-tzvp', unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao) - .5
    dm = dm.dot(dm.T)
    opt = _VHFOpt(mol).build()
    ejk = _jk_energy_per_atom(opt, dm) * .5
    ref = O.ejk_ip1(mol, dm) * .5
    print(f'[ejk tzvp] max |GPU - oracle| {abs(ejk - ref).max():.2e}', flush=True)
    assert abs(ejk.sum()) < 1e-9
    assert abs(O.fp(ejk) - 2710.490337642) < 1e-9
    assert abs(ejk - ref).max() < 1e-10 * max(1., abs(ref).max())
