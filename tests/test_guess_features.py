'''CPU checks of the learned guess's pair features and local frames
(gpu4pyscf_amd/guess/features.py): exact block round trip, rotation
invariance of the features and of the local density blocks (the density of
a rotated molecule is W D W^T with W the AO representation of the rotation),
and the harmonic rotations' representation property.'''
import numpy as np

from gpu4pyscf_amd.gto.mole import M
from gpu4pyscf_amd.guess import features as F


def _mol(coords, symbols, basis='def2-svp'):
    atom = '; '.join(f'{s} {x:.12f} {y:.12f} {z:.12f}' for s, (x, y, z) in zip(symbols, coords))
    return M(atom=atom, basis=basis, unit='Bohr')


def _two_waters():
    sym = ['O', 'H', 'H', 'O', 'H', 'H']
    xyz = np.array([[0., 0., 0.], [1.43, 1.11, 0.], [-1.43, 1.11, 0.],
                    [0.3, -1.0, 5.2], [1.1, 0.4, 5.9], [-1.2, -0.3, 6.1]])
    return xyz, sym


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def _sym(rng, n):
    g = rng.standard_normal((n, n))
    return g + g.T


def test_block_round_trip_is_exact():
    xyz, sym = _two_waters()
    mol = _mol(xyz, sym)
    rng = np.random.default_rng(0)
    dm = _sym(rng, mol.nao)
    rows, X = F.pairs_and_features(mol)
    assert len(rows) == mol.natm * (mol.natm + 1) // 2      # every pair within 6 A
    Y, Mk = F.blocks_to_local(mol, dm, rows, F.max_block(mol))
    assert np.abs(F.local_to_density(mol, rows, Y) - dm).max() < 1e-12
    assert Mk.sum() == sum((np.diff(mol.aoslice_by_atom()[[a, b], 2:], axis=1).prod())
                           for a, b, *_ in rows)


def test_harmonic_rotation_is_a_representation():
    rng = np.random.default_rng(1)
    A, B = _random_rotation(rng), _random_rotation(rng)
    for l in (1, 2, 3):
        UA, UB = F._sh_rotation(l, A), F._sh_rotation(l, B)
        assert np.abs(F._sh_rotation(l, A @ B) - UA @ UB).max() < 1e-12
        assert np.abs(UA @ UA.T - np.eye(2 * l + 1)).max() < 1e-12


def test_features_and_local_blocks_are_rotation_invariant():
    xyz, sym = _two_waters()
    rng = np.random.default_rng(2)
    R = _random_rotation(rng)
    mol = _mol(xyz, sym)
    mol_r = _mol(xyz @ R.T + np.array([0.7, -0.2, 0.4]), sym)
    dm = _sym(rng, mol.nao)
    # W: AO representation of R, atom by atom
    aos = mol.aoslice_by_atom()
    W = np.zeros((mol.nao, mol.nao))
    for ia in range(mol.natm):
        W[aos[ia, 2]:aos[ia, 3], aos[ia, 2]:aos[ia, 3]] = F.ao_rotation(mol, ia, R)
    dm_r = W @ dm @ W.T
    rows, X = F.pairs_and_features(mol)
    rows_r, X_r = F.pairs_and_features(mol_r)
    assert np.abs(X - X_r).max() < 1e-10
    out = F.max_block(mol)
    Y, _ = F.blocks_to_local(mol, dm, rows, out)
    Y_r, _ = F.blocks_to_local(mol_r, dm_r, rows_r, out)
    assert np.abs(Y - Y_r).max() < 1e-10
    # and the lab-frame reconstruction co-rotates
    assert np.abs(F.local_to_density(mol_r, rows_r, Y) - dm_r).max() < 1e-10


def test_batched_rotations_match_single():
    rng = np.random.default_rng(3)
    Qs = np.stack([_random_rotation(rng) for _ in range(5)])
    for l in (0, 1, 2, 3):
        batch = F._sh_rotations(l, Qs)
        for n in range(5):
            assert np.abs(batch[n] - F._sh_rotation(l, Qs[n])).max() < 1e-13


def test_frames_and_layout_follow_the_neighbour_rule():
    '''Frames written out by hand from the rule (same-atom row: z to the
    nearest, x to the second nearest neighbour; pair row: z from A to B, x to
    the nearest neighbour of A other than B; ties by atom index; a lab axis
    when x is parallel to z) and the feature slots of one pair row.'''
    from gpu4pyscf_amd.gto.mole import BOHR
    s2 = np.sqrt(.5)
    cases = [
        # bent: O, H1 on z, H2 on y (equal O-H: H1 is the nearest by index)
        ([[0, 0, 0], [0, 0, 1.8], [0, 1.8, 0]], ['O', 'H', 'H'],
         {(0, 0): [[0, 1, 0], [-1, 0, 0], [0, 0, 1]],
          (0, 1): [[0, 1, 0], [-1, 0, 0], [0, 0, 1]],
          (0, 2): [[0, 0, 1], [1, 0, 0], [0, 1, 0]],
          (1, 1): [[0, 1, 0], [1, 0, 0], [0, 0, -1]],
          (1, 2): [[0, -s2, -s2], [-1, 0, 0], [0, s2, -s2]],
          (2, 2): [[0, 0, 1], [-1, 0, 0], [0, -1, 0]]}),
        # collinear: x of the O row falls back to the lab x axis
        ([[0, 0, 0], [0, 0, 1.8], [0, 0, -1.8]], ['O', 'H', 'H'],
         {(0, 0): [[1, 0, 0], [0, 1, 0], [0, 0, 1]]}),
        # two atoms along x: no third atom, the lab y axis
        ([[0, 0, 0], [1.4, 0, 0]], ['H', 'H'],
         {(0, 0): [[0, 1, 0], [0, 0, 1], [1, 0, 0]],
          (0, 1): [[0, 1, 0], [0, 0, 1], [1, 0, 0]],
          (1, 1): [[0, 1, 0], [0, 0, -1], [-1, 0, 0]]}),
    ]
    for xyz, sym, want in cases:
        rows, X = F.pairs_and_features(_mol(np.array(xyz, float), sym))
        got = {(a, b): Q for a, b, t, Q in rows}
        n = len(sym)
        assert list(got) == [(a, b) for a in range(n) for b in range(a, n)]
        for key, Q in want.items():
            assert np.allclose(got[key], Q, atol=1e-12), key
    # feature slots of the O-H1 row of the bent molecule
    rows, X = F.pairs_and_features(_mol(np.array(cases[0][0], float), cases[0][1]))
    f = X[1]
    t = F.N_TYPES - len(F.PAIR_TYPES) + F.PAIR_TYPES.index(('H', 'O'))
    assert rows[1][2] == t and f[t] == 1. and f[:F.N_TYPES].sum() == 1.
    k = F.PAIR_TYPES.index(('H', 'O'))
    assert np.allclose(f[F.N_TYPES + k * F.NRBF:F.N_TYPES + (k + 1) * F.NRBF],
                       F.rbf(1.8 * BOHR))
    off = F.N_TYPES + F.NRBF * len(F.PAIR_TYPES)
    d_hh = np.sqrt(2.) * 1.8 * BOHR
    env_o = np.zeros((len(F.ELEMENTS), F.NRBF))
    env_o[F.ELEMENTS.index('H')] = 2 * F.rbf(1.8 * BOHR)
    env_h = np.zeros((len(F.ELEMENTS), F.NRBF))
    env_h[F.ELEMENTS.index('O')] = F.rbf(1.8 * BOHR)
    env_h[F.ELEMENTS.index('H')] = F.rbf(d_hh)
    assert np.allclose(f[off:off + F.NRBF * 3], env_o.ravel())
    assert np.allclose(f[off + F.NRBF * 3:off + F.NRBF * 6], env_h.ravel())
    assert np.all(f[F.FEAT_DIM:] == 0.)
