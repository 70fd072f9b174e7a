'''Seeded synthetic geometries of the J/K workloads.

Water clusters: rigid H2O (O-H 0.9572 Angstrom, H-O-H 104.52 degrees), O on an
a x b x c cubic lattice with spacing s, each O jittered uniformly by +-0.3
Angstrom per axis, each molecule rotated by a uniformly random rotation.
All-trans linear alkanes CnH2n+2: C-C 1.54, C-H 1.09 Angstrom, tetrahedral
angles, 180 degree dihedrals, Gaussian jitter sigma=0.02 Angstrom per atom.
Every random number comes from numpy.random.default_rng(seed).
'''

import numpy as np

OH = 0.9572
HOH = np.deg2rad(104.52)


def _random_rotation(rng):
    # uniform random rotation from a normalised Gaussian quaternion
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2*(y*y + z*z), 2*(x*y - z*w), 2*(x*z + y*w)],
        [2*(x*y + z*w), 1 - 2*(x*x + z*z), 2*(y*z - x*w)],
        [2*(x*z - y*w), 2*(y*z + x*w), 1 - 2*(x*x + y*y)]])


def water_cluster(shape=(2, 2, 2), spacing=3.0, seed=0, jitter=0.3):
    '''Atom list [(symbol, xyz Angstrom)] of a water cluster.'''
    rng = np.random.default_rng(seed)
    h1 = OH * np.array([np.sin(HOH/2), 0., np.cos(HOH/2)])
    h2 = OH * np.array([-np.sin(HOH/2), 0., np.cos(HOH/2)])
    atoms = []
    a, b, c = shape
    for i in range(a):
        for j in range(b):
            for k in range(c):
                o = np.array([i, j, k], dtype=float) * spacing
                o += rng.uniform(-jitter, jitter, 3)
                R = _random_rotation(rng)
                atoms.append(('O', o))
                atoms.append(('H', o + R @ h1))
                atoms.append(('H', o + R @ h2))
    return atoms


def alkane(n_carbon=30, seed=0, sigma=0.02):
    '''All-trans CnH(2n+2) zig-zag chain.'''
    rng = np.random.default_rng(seed)
    cc, ch = 1.54, 1.09
    half = np.arccos(-1./3) / 2          # half tetrahedral angle
    dx = cc * np.sin(half)
    dz = cc * np.cos(half)
    atoms = []
    carbons = []
    for i in range(n_carbon):
        z = 0. if i % 2 == 0 else dz
        carbons.append(np.array([i * dx, 0., z]))
    for i, c in enumerate(carbons):
        atoms.append(('C', c.copy()))
    # two H per carbon in the plane perpendicular to the chain, pointing away
    # from the zig-zag; terminal carbons get a third H along the chain
    for i, c in enumerate(carbons):
        out = -1. if i % 2 == 0 else 1.   # away from the chain's other side
        # bisector direction of the H-C-H angle
        b = np.array([0., 0., out])
        hh = np.arccos(-1./3) / 2
        for s in (1., -1.):
            d = np.cos(hh) * b + np.sin(hh) * np.array([0., s, 0.])
            atoms.append(('H', c + ch * d))
        if i == 0 or i == n_carbon - 1:
            nb = carbons[1] if i == 0 else carbons[-2]
            v = c - nb
            # terminal H: continue the zig-zag, tetrahedral to the C-C bond
            d = np.array([np.sign(v[0]) * np.sin(half), 0., -out * np.cos(half)])
            atoms.append(('H', c + ch * d / np.linalg.norm(d)))
    return [(s, xyz + rng.normal(0., sigma, 3)) for s, xyz in atoms]


def to_string(atoms):
    return '\n'.join(f'{s} {x[0]:.10f} {x[1]:.10f} {x[2]:.10f}' for s, x in atoms)
