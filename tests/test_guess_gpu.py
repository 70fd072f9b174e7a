'''Learned initial guess with the trained weights shipped in the package:
on held-out 8-water def2-SVP clusters (seeds outside the training range
1000-2999) it must start closer to the converged density than the atomic
guess, converge in no more SCF cycles and to the same energy.'''
import os

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(__file__), '..', 'gpu4pyscf_amd', 'guess', 'data',
                       'water8_def2svp.npz')


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def test_learned_guess_beats_atomic_guess():
    from gpu4pyscf_amd.guess.learned import LearnedGuess, water_mol
    from gpu4pyscf_amd.scf.atom_guess import atom_guess
    from gpu4pyscf_amd.scf.hf import RHF
    guess = LearnedGuess.load(WEIGHTS)
    cyc_atom, cyc_learned = [], []
    for seed in (5000, 5001, 5002, 5003):
        mol = water_mol(seed)
        print(f'[guess seed {seed}]', flush=True)
        res = {}
        for name, key in (('atom', 'atom'), ('learned', guess)):
            mf = RHF(mol)
            mf.conv_tol = 1e-8
            mf.init_guess = key
            e = mf.kernel()
            assert mf.converged
            res[name] = (e, mf.cycles, mf.dm.cpu().numpy())
        assert not guess.used_fallback
        d_conv = res['atom'][2]
        rms_l = np.sqrt(np.mean((guess(mol) - d_conv)**2))
        rms_a = np.sqrt(np.mean((atom_guess(mol) - d_conv)**2))
        assert rms_l < 0.5 * rms_a, (rms_l, rms_a)
        assert abs(res['learned'][0] - res['atom'][0]) < 1e-8
        cyc_atom.append(res['atom'][1])
        cyc_learned.append(res['learned'][1])
    assert sum(cyc_learned) < sum(cyc_atom), (cyc_learned, cyc_atom)


def test_learned_guess_falls_back_outside_its_basis():
    '''def2-TZVP blocks do not fit the network output: atomic guess.'''
    from gpu4pyscf_amd.guess.learned import LearnedGuess, water_mol
    from gpu4pyscf_amd.scf.atom_guess import atom_guess
    guess = LearnedGuess.load(WEIGHTS)
    mol = water_mol(5000, shape=(1, 1, 2), basis='def2-tzvp')
    dm = guess(mol)
    assert guess.used_fallback
    assert np.allclose(dm, atom_guess(mol))
