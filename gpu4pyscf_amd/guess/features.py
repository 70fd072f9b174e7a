'''Atom-pair descriptors and local frames of the learned initial guess.

Every atom pair (A, B) with R_AB < 6 Angstrom, and every atom with itself,
gets one row of features:

* one-hot of the block type (same-atom H/C/O, pair HH/HC/HO/CC/CO/OO);
* 32 Gaussian radial functions of R_AB placed in the slot of its element
  pair (centres 0.5..6 Angstrom, width 0.25 Angstrom, cosine cutoff 6);
* the same 32 radial functions summed over the neighbours of A, per
  neighbour element, and likewise for B (environment of the pair).

The density block of a pair is covariant, so it is expressed in a local frame
(z from A to B, x towards the nearest other neighbour of A; for a same-atom
block z and x point to the two nearest neighbours): with Q the frame (rows =
local axes) and U(Q) its representation on real spherical harmonics,
D_local = U_A D_AB U_B^T is invariant under rotations of the molecule, and
D_AB = U_A^T D_local U_B restores the lab frame.
'''

import numpy as np

from gpu4pyscf_amd.gto.mole import BOHR, ANG_OF
from gpu4pyscf_amd.gto.cart2sph import c2s, cart_components

ELEMENTS = ('H', 'C', 'O')
PAIR_TYPES = [(a, b) for i, a in enumerate(ELEMENTS) for b in ELEMENTS[i:]]
NRBF = 32
RCUT = 6.0
N_TYPES = len(ELEMENTS) + len(PAIR_TYPES)          # 3 same-atom + 6 pair types
FEAT_DIM = N_TYPES + NRBF * len(PAIR_TYPES) + 2 * NRBF * len(ELEMENTS)   # 393
FEAT_PAD = 400
OUT_DIM = 196                                        # default: O x O block of def2-SVP

_centres = np.linspace(0.5, RCUT, NRBF)
_width = 0.25


def rbf(r):
    r = np.asarray(r, dtype=float)
    f = np.exp(-(r[..., None] - _centres)**2 / (2 * _width**2))
    cut = np.where(r < RCUT, .5 * (np.cos(np.pi * r / RCUT) + 1), 0.)
    return f * cut[..., None]


def pair_type(ea, eb, same_atom):
    if same_atom:
        return ELEMENTS.index(ea)
    key = tuple(sorted((ea, eb), key=ELEMENTS.index))
    return len(ELEMENTS) + PAIR_TYPES.index(key)


def _sh_basis(l, _cache={}):
    '''Sample points on the unit sphere, the monomial -> real-harmonic map
    of angular momentum l, and the least-squares projector onto the
    harmonics at those points.'''
    if l not in _cache:
        rng = np.random.default_rng(7)
        pts = rng.standard_normal((64, 3))
        pts /= np.linalg.norm(pts, axis=1)[:, None]
        comps = np.asarray(cart_components(l))
        C = c2s(l)
        Y0 = (np.prod(pts[:, None, :] ** comps[None], axis=2)) @ C.T
        _cache[l] = (pts, comps, C, np.linalg.pinv(Y0))
    return _cache[l]


def _sh_rotation(l, Q):
    '''U with Y_m(Q^T r) = sum_m' Y_m'(r) U[m', m] for PySCF-ordered real
    harmonics of angular momentum l (p ordered x, y, z).'''
    if l == 0:
        return np.ones((1, 1))
    if l == 1:
        return Q.copy()
    pts, comps, C, P = _sh_basis(l)
    q = pts @ Q            # rows: (Q^T r)^T = r^T Q
    Y1 = (np.prod(q[:, None, :] ** comps[None], axis=2)) @ C.T
    return P @ Y1


def _sh_rotations(l, Qs):
    '''Batched _sh_rotation over frames Qs (R, 3, 3) -> (R, 2l+1, 2l+1).'''
    R = len(Qs)
    if l == 0:
        return np.ones((R, 1, 1))
    if l == 1:
        return np.array(Qs, dtype=float)
    pts, comps, C, P = _sh_basis(l)
    q = pts @ Qs                                               # (R, 64, 3)
    pw = [np.ones_like(q)]
    for _ in range(l):
        pw.append(pw[-1] * q)
    pw = np.stack(pw, axis=-1)                                 # (R, 64, 3, l+1)
    mono = (pw[:, :, 0, comps[:, 0]] * pw[:, :, 1, comps[:, 1]] * pw[:, :, 2, comps[:, 2]])
    return P @ (mono @ C.T)


def _shell_ls(mol, ia):
    ls = []
    for b in mol._bas[mol._bas[:, 0] == ia]:
        ls += [int(b[ANG_OF])] * int(b[3])
    return ls


def ao_rotations(mol, atoms, Qs):
    '''AO representations (spherical, PySCF order) of the frames Qs
    (R, 3, 3) on the shells of ``atoms`` (R atoms of one element).'''
    ls = _shell_ls(mol, atoms[0])
    n = sum(2 * l + 1 for l in ls)
    out = np.zeros((len(Qs), n, n))
    rot = {l: _sh_rotations(l, Qs) for l in set(ls)}
    p = 0
    for l in ls:
        k = 2 * l + 1
        out[:, p:p + k, p:p + k] = rot[l]
        p += k
    return out


def ao_rotation(mol, ia, Q):
    '''AO representation (spherical, PySCF order) of frame rotation Q on the
    shells of atom ia.'''
    return ao_rotations(mol, [ia], np.asarray(Q)[None])[0]


def _groups(mol, rows):
    '''Row indices grouped by (element of A, element of B): the rows of a
    group share block shapes and shell structures.'''
    sym = [mol.atom_symbol(i) for i in range(mol.natm)]
    g = {}
    for r, (a, b, t, Q) in enumerate(rows):
        g.setdefault((sym[a], sym[b]), []).append(r)
    return g


def _frames(z, x_hint):
    '''Local frames of many rows at once: (N, 3, 3), rows of each = local
    x, y, z axes in lab coordinates; z along z, x the part of x_hint normal
    to z (a lab axis when x_hint is parallel to z).'''
    z = z / np.linalg.norm(z, axis=1)[:, None]
    x = x_hint - np.sum(x_hint * z, axis=1)[:, None] * z
    flat = np.linalg.norm(x, axis=1) < 1e-6
    if flat.any():
        t = np.where((np.abs(z[flat, 0]) < .9)[:, None], [1., 0., 0.], [0., 1., 0.])
        x[flat] = t - np.sum(t * z[flat], axis=1)[:, None] * z[flat]
    x = x / np.linalg.norm(x, axis=1)[:, None]
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=1)


def pairs_and_features(mol):
    '''Rows (A, B, type, Q) and the (n_rows, FEAT_PAD) feature matrix.'''
    R = mol.atom_coords() * BOHR
    sym = [mol.atom_symbol(i) for i in range(mol.natm)]
    for s in sym:
        if s not in ELEMENTS:
            raise ValueError(f'learned guess has no parameters for element {s}')
    n = len(R)
    if n < 2:
        raise ValueError('learned guess needs at least two atoms for its frames')
    elem = np.array([ELEMENTS.index(s) for s in sym])
    D = np.linalg.norm(R[:, None] - R[None], axis=2)
    # environment: radial functions of every other atom within RCUT, summed
    # per neighbour element (np.add.at adds in the row-major order of the
    # neighbour list)
    near = (D < RCUT) & ~np.eye(n, dtype=bool)
    ia, ic = np.nonzero(near)
    env = np.zeros((n, len(ELEMENTS), NRBF))
    np.add.at(env, (ia, elem[ic]), rbf(D[ia, ic]))
    # rows: A <= B, same atom or R_AB < RCUT, A-major
    a, b = np.triu_indices(n)
    keep = (a == b) | (D[a, b] < RCUT)
    a, b = a[keep], b[keep]
    same = a == b
    ptype = np.array([[pair_type(ELEMENTS[i], ELEMENTS[j], False) for j in range(len(ELEMENTS))]
                      for i in range(len(ELEMENTS))])
    t = np.where(same, elem[a], ptype[elem[a], elem[b]])
    nr = len(a)
    f = np.zeros((nr, FEAT_PAD))
    f[np.arange(nr), t] = 1.
    # pair radial functions in the slot of the element pair
    ps = np.nonzero(~same)[0]
    slot = N_TYPES + (t[ps] - len(ELEMENTS)) * NRBF
    f[ps[:, None], slot[:, None] + np.arange(NRBF)] = rbf(D[a[ps], b[ps]])
    off = N_TYPES + NRBF * len(PAIR_TYPES)
    f[:, off:off + NRBF * len(ELEMENTS)] = env[a].reshape(nr, -1)
    f[:, off + NRBF * len(ELEMENTS):off + 2 * NRBF * len(ELEMENTS)] = env[b].reshape(nr, -1)
    # neighbours by distance rounded to 1e-6 Angstrom, ties broken by atom
    # index, so that equal bond lengths give the same frame after rotation;
    # each atom itself removed from its list
    Dr = np.round(D, 6)
    order = np.array([np.lexsort((np.arange(n), Dr[i])) for i in range(n)])
    others = order[order != np.arange(n)[:, None]].reshape(n, n - 1)
    # same-atom rows: z to the nearest, x to the second nearest neighbour;
    # pair rows: z from A to B, x to the nearest neighbour of A other than B
    first = others[a, 0]
    if n > 2:
        second = others[a, 1]
        xi = np.where(same, second, np.where(first == b, second, first))
        x_hint = R[xi] - R[a]
    else:
        x_hint = np.tile([1., 0., 0.], (len(a), 1))     # no third atom
    z = np.where(same[:, None], R[first] - R[a], R[b] - R[a])
    Q = _frames(z, x_hint)
    rows = [(int(i), int(j), int(k), q) for i, j, k, q in zip(a, b, t, Q)]
    return rows, f


def max_block(mol):
    '''Largest atom-pair density block of the molecule's basis.'''
    aos = mol.aoslice_by_atom()
    n = (aos[:, 3] - aos[:, 2]).max()
    return int(n * n)


def blocks_to_local(mol, dm, rows, out_dim=OUT_DIM):
    '''Targets: the density blocks of ``rows`` in their local frames,
    (n_rows, out_dim) with a mask of the valid entries.'''
    aos = mol.aoslice_by_atom()
    Y = np.zeros((len(rows), out_dim))
    M = np.zeros((len(rows), out_dim))
    for idx in _groups(mol, rows).values():
        A = [rows[r][0] for r in idx]
        B = [rows[r][1] for r in idx]
        Qs = np.stack([rows[r][3] for r in idx])
        Ua = ao_rotations(mol, A, Qs)
        Ub = ao_rotations(mol, B, Qs)
        na, nb = Ua.shape[1], Ub.shape[1]
        if na * nb > out_dim:
            raise ValueError('density block larger than the network output')
        blk = np.stack([dm[aos[a, 2]:aos[a, 3], aos[b, 2]:aos[b, 3]] for a, b in zip(A, B)])
        loc = Ua @ blk @ Ub.transpose(0, 2, 1)
        Y[idx, :na * nb] = loc.reshape(len(idx), -1)
        M[idx, :na * nb] = 1.
    return Y, M


def local_to_density(mol, rows, Y):
    '''Assemble the lab-frame density from predicted local blocks.'''
    aos = mol.aoslice_by_atom()
    nao = mol.nao
    dm = np.zeros((nao, nao))
    for idx in _groups(mol, rows).values():
        A = [rows[r][0] for r in idx]
        B = [rows[r][1] for r in idx]
        Qs = np.stack([rows[r][3] for r in idx])
        Ua = ao_rotations(mol, A, Qs)
        Ub = ao_rotations(mol, B, Qs)
        na, nb = Ua.shape[1], Ub.shape[1]
        if na * nb > Y.shape[1]:
            raise ValueError('density block larger than the network output')
        loc = np.asarray(Y)[idx, :na * nb].reshape(len(idx), na, nb)
        blk = Ua.transpose(0, 2, 1) @ loc @ Ub
        ia = aos[A, 2][:, None] + np.arange(na)                 # (R, na) AO rows
        ib = aos[B, 2][:, None] + np.arange(nb)
        # each (A, B) pair occurs once, so the scattered writes do not collide
        dm[ia[:, :, None], ib[:, None, :]] = blk
        off = np.asarray(A) != np.asarray(B)
        dm[ib[off][:, :, None], ia[off][:, None, :]] = blk[off].transpose(0, 2, 1)
    return .5 * (dm + dm.T)
