'''Pins the CPU oracle (and the molecule/basis layer it reads) against the
golden fingerprints held by the reference's own tests.'''
import numpy as np
import pytest

from oracle import oracle as O
from tests.molecules import case, GOLDEN_CASES


@pytest.mark.parametrize('name', GOLDEN_CASES)
def test_oracle_fingerprints(name):
    mol, dm, hermi, fpj, fpk, tol = case(name)
    vj, vk = O.get_jk(mol, dm, hermi=hermi)
    if fpj is not None:
        assert abs(O.fp(vj) - fpj) < tol
    if fpk is not None:
        assert abs(O.fp(vk) - fpk) < tol
    if hermi == 2:
        assert abs(vj).max() == 0


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize('hermi,omega,fpj,fpk', [
    (1, .2, 1163.932604635460, -1269.969109438691),
    (1, -.2, -3491.404124194866, -2799.347891387202),
    (0, .2, 24.18519249677608, -34.15933205656134),
    (0, -.2, -77.67449053914103, -80.95859292428769)])
def test_oracle_range_separated_fingerprints(hermi, omega, fpj, fpk):
    '''gpu4pyscf/scf/tests/test_scf_jk.py:59-80,142-170: erf (omega > 0)
    and erfc (omega < 0) operators for J and K.'''
    from tests.molecules import ATOM6
    from gpu4pyscf_amd.gto.mole import M
    mol = M(atom=ATOM6, basis='def2-tzvp', unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao)
    if hermi == 1:
        dm = dm.dot(dm.T)
    mol.omega = omega
    vj, vk = O.get_jk(mol, dm, hermi=hermi)
    tol = 1e-9 if hermi == 1 else 5e-10
    assert abs(O.fp(vj) - fpj) < tol
    assert abs(O.fp(vk) - fpk) < tol


def test_oracle_h_shell_and_short_range_k():
    '''test_scf_jk.py:172-198 (h shell on O, hermi=0) and :358-369
    (short-range K of O2, def2-SVP; the reference's 1e-13 rests on its own
    short-range screening, the oracle screens by the short-range Schwarz
    factor and agrees to 1e-11).'''
    from tests.molecules import ATOM6
    from gpu4pyscf_amd.gto.mole import M
    mol = M(atom=ATOM6, basis={'default': 'def2-tzvp', 'O': [[5, [1., 1.]]]}, unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao)
    vj, vk = O.get_jk(mol, dm, hermi=0)
    assert abs(O.fp(vj) - -61.28856847097108) < 1e-9
    assert abs(O.fp(vk) - -76.38373664249241) < 1e-9
    mol = M(atom='O 0.0000 0.7375 -0.0528; O 0.0000 -0.7375 -0.1528', basis='def2-svp')
    np.random.seed(12)
    dm = np.random.rand(mol.nao, mol.nao) - .5
    mol.omega = -.3
    vk = O.get_jk(mol, dm.dot(dm.T), hermi=1, with_j=False)[1]
    assert abs(O.fp(vk) - -1.8653967312459407) < 1e-11


def test_oracle_energy_gradient_goldens():
    '''gpu4pyscf/scf/tests/test_scf_jk.py:404-414 (short-range exchange
    gradient of O2, atom 0) and grad/tests/test_rhf_grad.py:94-113 (RHF
    two-electron gradient, def2-TZVP): fingerprints, and the sum over atoms
    vanishes (translational invariance).'''
    from gpu4pyscf_amd.gto.mole import M
    mol = M(atom='O 0.0000 0.7375 -0.0528; O 0.0000 -0.7375 -0.1528', basis='def2-svp')
    np.random.seed(12)
    dm = np.random.rand(mol.nao, mol.nao) - .5
    mol.omega = -.3
    e = O.ejk_ip1(mol, dm.dot(dm.T), j_factor=0, k_factor=1.) * .5
    assert abs(e[0] - [0.24806416996651, 1.11003753769514, 0.19967171093788]).max() < 1e-12
    mol = M(atom='O 0.000 -0. 0.1174; H -0.757 4. -0.4696; H 0.757 4. -0.4696; C 3. 1. 0.',
            basis='def2-tzvp', unit='B')
    np.random.seed(9)
    dm = np.random.rand(mol.nao, mol.nao) - .5
    e = O.ejk_ip1(mol, dm.dot(dm.T)) * .5
    assert abs(e.sum()) < 1e-9
    assert abs(O.fp(e) - 2710.490337642) < 1e-9


def test_oracle_energy_gradient_finite_differences():
    '''The analytic per-atom derivatives against central differences of
    E = 1/2 Tr(D (J - K/2)) with the AO density held fixed while an atom
    moves (the basis functions move with it).'''
    from gpu4pyscf_amd.gto.mole import M
    atoms = [['O', (0., 0., .2217)], ['H', (-1.43, 0., -.887)], ['H', (1.43, .1, -.887)]]

    def mol_at(coords):
        return M(atom='; '.join(f'{s} {x} {y} {z}' for s, (x, y, z) in zip('OHH', coords)),
                 basis='def2-svp', unit='B')

    coords = np.array([c for _, c in atoms])
    mol = mol_at(coords)
    rng = np.random.default_rng(7)
    c = np.linalg.qr(rng.standard_normal((mol.nao, 5)))[0]
    dm = 2 * c @ c.T

    def energy(crd):
        vj, vk = O.get_jk(mol_at(crd), dm, hermi=1)
        return .5 * np.sum(dm * (vj - .5 * vk))

    g = O.ejk_ip1(mol, dm)
    h = 1e-4
    for atom, x in [(0, 2), (1, 0), (2, 1)]:
        cp, cm = coords.copy(), coords.copy()
        cp[atom, x] += h
        cm[atom, x] -= h
        fd = (energy(cp) - energy(cm)) / (2 * h)
        assert abs(fd - g[atom, x]) < 1e-7 * max(1., abs(fd)), (atom, x, fd, g[atom, x])
