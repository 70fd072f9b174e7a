'''Brute-force two-electron integrals by the McMurchie-Davidson scheme --
TEST INFRASTRUCTURE ONLY.  Independent of the Rys quadrature (Boys function
from 1F1), used to pin the oracle's Rys ERIs on tiny inputs.'''
import math
import numpy as np

from gpu4pyscf_amd.gto.cart2sph import cart_components
from oracle.scf_oracle import _hermite_E, _hermite_R


def eri_shell(bas, env, i, j, k, l):
    sh = [bas[x] for x in (i, j, k, l)]
    comps = [cart_components(b[1]) for b in sh]
    R = [env[b[7]:b[7]+3] for b in sh]
    out = np.zeros([len(c) for c in comps])
    for pa in range(sh[0][2]):
        for pb in range(sh[1][2]):
            a = env[sh[0][5]+pa]; b = env[sh[1][5]+pb]
            cab = env[sh[0][6]+pa] * env[sh[1][6]+pb]
            p = a + b
            P = (a * R[0] + b * R[1]) / p
            Eab = [_hermite_E(sh[0][1], sh[1][1], a, b, R[0][d] - R[1][d]) for d in range(3)]
            for pc in range(sh[2][2]):
                for pd in range(sh[3][2]):
                    c = env[sh[2][5]+pc]; d_ = env[sh[3][5]+pd]
                    ccd = env[sh[2][6]+pc] * env[sh[3][6]+pd]
                    q = c + d_
                    Q = (c * R[2] + d_ * R[3]) / q
                    Ecd = [_hermite_E(sh[2][1], sh[3][1], c, d_, R[2][x] - R[3][x]) for x in range(3)]
                    alpha = p * q / (p + q)
                    L = sum(s[1] for s in sh)
                    r = _hermite_R(L, alpha, P - Q)
                    pref = 2 * math.pi**2.5 / (p * q * math.sqrt(p + q)) * cab * ccd
                    for ia, ca in enumerate(comps[0]):
                        for ib, cb in enumerate(comps[1]):
                            for ic, cc in enumerate(comps[2]):
                                for id_, cd in enumerate(comps[3]):
                                    v = 0.
                                    for t in range(ca[0]+cb[0]+1):
                                        for u in range(ca[1]+cb[1]+1):
                                            for w in range(ca[2]+cb[2]+1):
                                                e1 = Eab[0][ca[0], cb[0], t] * Eab[1][ca[1], cb[1], u] * Eab[2][ca[2], cb[2], w]
                                                if e1 == 0:
                                                    continue
                                                for tau in range(cc[0]+cd[0]+1):
                                                    for nu in range(cc[1]+cd[1]+1):
                                                        for phi in range(cc[2]+cd[2]+1):
                                                            e2 = Ecd[0][cc[0], cd[0], tau] * Ecd[1][cc[1], cd[1], nu] * Ecd[2][cc[2], cd[2], phi]
                                                            v += e1 * e2 * (-1)**(tau+nu+phi) * r(t+tau, u+nu, w+phi, 0)
                                    out[ia, ib, ic, id_] += pref * v
    return out
