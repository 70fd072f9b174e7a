'''GPU SCF driver: one-electron integrals and RHF energies against the CPU
oracle and the reference's HLi golden energy
(gpu4pyscf/scf/tests/test_diffuse_orbital.py:27-50,87-92).'''
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.molecules import HLI_BASIS

pytestmark = pytest.mark.gpu

HLI_E = -7.670162135801041


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def _water(basis='def2-svp'):
    from gpu4pyscf_amd.gto.mole import M
    return M(atom='O 0 0 0.1174; H -0.757 0 -0.4696; H 0.757 0 -0.4696', basis=basis)


def test_int1e_matches_oracle():
    from gpu4pyscf_amd.scf import hf, jk
    from oracle import scf_oracle
    mol = _water('def2-tzvp')
    opt = jk._VHFOpt(mol).build()
    S, T, V = hf.int1e(mol, opt)
    sm = opt.sorted_mol
    rS, rT, rV = scf_oracle.int1e_sorted(sm, mol.atom_coords(), mol.atom_charges().astype(float))
    C = sm.coeff
    assert abs(S.cpu().numpy() - C.T @ rS @ C).max() < 1e-12
    assert abs(T.cpu().numpy() - C.T @ rT @ C).max() < 1e-11
    assert abs(V.cpu().numpy() - C.T @ rV @ C).max() < 1e-11
    # spherical AOs are normalised
    assert abs(np.diag(S.cpu().numpy()) - 1).max() < 1e-12


def test_rhf_hli_single_p_exact():
    '''HLi with one diffuse p shell: the golden energy to 1e-9.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf.hf import RHF
    b = HLI_BASIS.replace("""            X    P
                0.010001 1.0""", "")
    mol = M(atom='H 0 0 0; Li 1.0 0.1 0', basis=b)
    mf = RHF(mol)
    mf.conv_tol = 1e-10
    e = mf.kernel()
    assert mf.converged
    assert abs(e - HLI_E) < 1e-9


def test_rhf_hli_golden():
    '''HLi with the near-duplicate p shells: the reference test's 1e-5.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf.hf import RHF
    mol = M(atom='H 0 0 0; Li 1.0 0.1 0', basis=HLI_BASIS)
    mf = RHF(mol)
    mf.conv_tol = 1e-10
    e = mf.kernel()
    assert mf.converged
    assert abs(e - HLI_E) < 1e-5


def test_rhf_water_matches_oracle():
    from gpu4pyscf_amd.scf.hf import RHF
    from oracle import scf_oracle
    mol = _water('def2-svp')
    mf = RHF(mol)
    mf.conv_tol = 1e-10
    e = mf.kernel()
    assert mf.converged
    e_ref, _ = scf_oracle.rhf(mol, conv_tol=1e-10)
    assert abs(e - e_ref) < 1e-9


def test_rhf_mixed_precision_water_cluster():
    '''Mixed FP32/FP64 incremental builds reach the FP64 energy within 1e-6.'''
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.scf.hf import RHF
    mol = M(atom=G.to_string(G.water_cluster((2, 2, 1), 3.0, seed=0)), basis='def2-svp')
    mf = RHF(mol)
    mf.conv_tol = 1e-9
    e64 = mf.kernel()
    mf2 = RHF(mol)
    mf2.conv_tol = 1e-9
    mf2.fp32_tol = 1e-6
    e32 = mf2.kernel()
    assert mf.converged and mf2.converged
    print(f'mixed: |dE| = {abs(e32 - e64):.2e}, fp32 quartets '
          f'{mf2.quartets_fp32 / mf2.quartets:.3f}')
    assert abs(e32 - e64) < 1e-6
    assert mf2.quartets_fp32 > 0
