'''Cartesian -> real spherical harmonic transformation (libcint convention).

A Cartesian AO of a shell with l >= 2 is N(a) x^i y^j z^k exp(-a r^2), where
N(a) is the radial normalisation; a spherical AO is N(a) r^l Y_lm exp(-a r^2)
with Y_lm the orthonormal real spherical harmonic.  ``c2s(l)[m, c]`` are the
coefficients of r^l Y_lm in the monomials x^i y^j z^k, rows ordered
m = -l..l, columns in lexical Cartesian order.  For l = 0, 1 the matrix is the
identity (PySCF orders p as x, y, z and folds sqrt((2l+1)/4pi) into the
contraction coefficients, gpu4pyscf/gto/mole.py:1120-1132).

The coefficients come from the closed form of the real regular solid
harmonics (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
Theory", eq. 6.4.47-6.4.50), scaled by sqrt((2l+1)/4pi).  Sign and order
agree with gpu4pyscf/lib/cupy_helper/cart2sph.cu:24-47 (d, f, g).
'''

from functools import lru_cache
from math import comb, factorial, sqrt, pi
import numpy as np


def cart_components(l):
    '''Lexical Cartesian powers (lx, ly, lz) of a shell with angular momentum l.'''
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


@lru_cache(maxsize=None)
def _cart_index(l):
    return {c: i for i, c in enumerate(cart_components(l))}


@lru_cache(maxsize=None)
def c2s(l):
    '''(2l+1, (l+1)(l+2)/2) matrix; see module docstring.'''
    ncart = (l + 1) * (l + 2) // 2
    if l < 2:
        return np.eye(ncart)
    idx = _cart_index(l)
    out = np.zeros((2*l+1, ncart))
    for row, m in enumerate(range(-l, l+1)):
        am = abs(m)
        norm = sqrt(2. * factorial(l+am) * factorial(l-am) / (2. if m == 0 else 1.)) \
            / (2**am * factorial(l))
        odd = 1 if m < 0 else 0
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                for vv in range(odd, am + 1, 2):
                    sign = (-1) ** (t + (vv - odd) // 2)
                    c = sign * 0.25**t * comb(l, t) * comb(l - t, am + t) \
                        * comb(t, u) * comb(am, vv)
                    lx = 2*t + am - 2*u - vv
                    ly = 2*u + vv
                    lz = l - 2*t - am
                    out[row, idx[(lx, ly, lz)]] += norm * c
    out *= sqrt((2*l + 1) / (4 * pi))
    return out
