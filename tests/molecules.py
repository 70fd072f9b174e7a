'''Molecules and density matrices used by the reference's J/K tests
(gpu4pyscf/scf/tests/test_scf_jk.py, test_rhf.py), rebuilt without PySCF.'''
import numpy as np
from gpu4pyscf_amd.gto.mole import M

ATOM6 = '''
        O   0.000   -0.    0.1174
        H  -0.757    4.   -0.4696
        H   0.757    4.   -0.4696
        C   1.      1.    0.
        H   4.      0.    3.
        H   0.      1.    .6
        '''

MOL1_BASIS = '''unc
#BASIS SET:
H    S
      1.815041   1
      0.591063   1
H    P
      2.305000   1
#BASIS SET:
C    S
      8.383976   1
      3.577015   1
      1.547118   1
C    P
      2.305000   1
      1.098827   1
      0.806750   1
      0.282362   1
C    D
      1.81900    1
      0.72760    1
      0.29104    1
C    F
      0.970109   1
C    G
      0.625000   1
C    H
      0.4        1
      '''

MOL1_ATOM = '''
C  -1.20806619, -0.34108413, -0.00755148
C   1.28636081, -0.34128013, -0.00668648
H   2.53407081,  1.81906387, -0.00736748
H   1.28693681,  3.97963587, -0.00925948
'''

HLI_BASIS = """
            X    S
                0.3425250914E+01       0.1543289673E+00
                0.6239137298E+00       0.5353281423E+00
                0.1688554040E+00       0.4446345422E+00
            X    S
                0.1611957475E+02       0.1543289673E+00
                0.2936200663E+01       0.5353281423E+00
                0.7946504870E+00       0.4446345422E+00
            X    P
                0.010000 1.0
            X    P
                0.010001 1.0
        """


def case(name):
    '''(mol, dm, hermi, fp_vj, fp_vk, tol) of a reference golden case.'''
    if name == 'tzvp_hermi1':      # test_scf_jk.py:24-49
        mol = M(atom=ATOM6, basis='def2-tzvp', unit='B')
        np.random.seed(9)
        dm = np.random.rand(mol.nao, mol.nao); dm = dm.dot(dm.T)
        return mol, dm, 1, -2327.4715195591784, -4069.3170008260583, 1e-9
    if name == 'tzvp_cart':        # test_scf_jk.py:82-108
        mol = M(atom=ATOM6, basis='def2-tzvp', unit='B', cart=True)
        np.random.seed(9)
        dm = np.random.rand(mol.nao, mol.nao)*.1 - .03; dm = dm.dot(dm.T)
        return mol, dm, 1, 88.88500592206657, 48.57434458906684, 1e-10
    if name == 'tzvp_hermi0':      # test_scf_jk.py:116-140
        mol = M(atom=ATOM6, basis='def2-tzvp', unit='B')
        np.random.seed(9)
        dm = np.random.rand(mol.nao, mol.nao)
        return mol, dm, 0, -53.489298042359046, -115.11792498085259, 5e-10
    if name == 'tzvp_g':           # test_scf_jk.py:200-225
        mol = M(atom=ATOM6, basis=('def2-tzvp', [[4, [1, 1]]]), unit='B')
        np.random.seed(9)
        dm = np.random.rand(mol.nao, mol.nao); dm = dm.dot(dm.T)
        return mol, dm, 1, None, 5580.092102968194, 1e-9
    if name == 'tzvp_g_2dm':       # test_scf_jk.py:227-236
        mol = M(atom=ATOM6, basis=('def2-tzvp', [[4, [1, 1]]]), unit='B')
        np.random.seed(9)
        dm = np.random.rand(2, mol.nao, mol.nao) - .5
        dm = np.einsum('nij,nkj->nik', dm, dm)
        return mol, dm, 1, None, 327.9485135045478, 1e-9
    if name == 'sto3g_hermi2':     # test_scf_jk.py:371-386
        mol = M(atom='''
    O  0.0000  0.7375 -0.0528
    O  0.0000 -0.7375 -0.1528
    ''', basis='sto-3g')
        np.random.seed(1)
        dm = np.random.rand(2, mol.nao, mol.nao); dm = dm - dm.transpose(0, 2, 1)
        return mol, dm, 2, None, 2.0697365797774125, 1e-12
    if name == 'mol1_hermi1':      # test_rhf.py:41-73,166-175
        mol = M(atom=MOL1_ATOM, basis=MOL1_BASIS)
        np.random.seed(1)
        dm = np.random.random((2, mol.nao, mol.nao)); dm = dm + dm.transpose(0, 2, 1)
        return mol, dm, 1, 179.14526555375858, -34.851182918643005, 1e-7
    if name == 'mol1_hermi0':      # test_rhf.py:182-189
        mol = M(atom=MOL1_ATOM, basis=MOL1_BASIS)
        np.random.seed(1)
        dm = np.random.random((2, mol.nao, mol.nao))
        return mol, dm, 0, 89.57263277687994, -26.36969769724246, 1e-7
    raise KeyError(name)


GOLDEN_CASES = ['tzvp_hermi1', 'tzvp_cart', 'tzvp_hermi0', 'tzvp_g', 'tzvp_g_2dm',
                'sto3g_hermi2', 'mol1_hermi1', 'mol1_hermi0']
