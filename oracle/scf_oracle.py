'''CPU oracle for the SCF consumer of the J/K path -- TEST INFRASTRUCTURE ONLY.

* One-electron integrals by the McMurchie-Davidson scheme (Hermite expansion
  coefficients E_t^{ij}, Hermite Coulomb integrals R_tuv with the Boys
  function F_n(T) = 1F1(n+1/2; n+3/2; -T)/(2n+1)); an algorithm independent of
  the product's Obara-Saika / Rys 1e kernel.
* RHF following gpu4pyscf/scf/hf.py:206-341 (incremental J/K on the density
  change, CDIIS from cycle 1 with 8 vectors as gpu4pyscf/scf/diis.py:41-127,
  canonical orthogonalisation dropping overlap eigenvalues < 1e-6 as
  gpu4pyscf/scf/hf.py:625-641), in numpy with the oracle J/K.
'''

import math
import numpy as np
from scipy.special import hyp1f1

from gpu4pyscf_amd.gto.cart2sph import cart_components


def _boys(n, T):
    return hyp1f1(n + .5, n + 1.5, -T) / (2 * n + 1)


def _hermite_E(la, lb, a, b, Xab):
    '''E[i][j][t] for the 1D Gaussian product, MD recursion.'''
    p = a + b
    mu = a * b / p
    XPA = -b / p * Xab        # P - A with Xab = A - B
    XPB = a / p * Xab
    E = np.zeros((la + 2, lb + 2, la + lb + 3))
    E[0, 0, 0] = math.exp(-mu * Xab * Xab)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i-1, j, t]
                    if t > 0:
                        v += E[i-1, j, t-1] / (2 * p)
                    v += (t + 1) * E[i-1, j, t+1]
                else:
                    v = XPB * E[i, j-1, t]
                    if t > 0:
                        v += E[i, j-1, t-1] / (2 * p)
                    v += (t + 1) * E[i, j-1, t+1]
                E[i, j, t] = v
    return E


def _hermite_R(tmax, p, PC):
    T = p * float(PC @ PC)
    R = {}
    F = [_boys(n, T) for n in range(tmax + 1)]

    def r(t, u, v, n):
        key = (t, u, v, n)
        if key in R:
            return R[key]
        if t == u == v == 0:
            val = (-2 * p) ** n * F[n]
        elif t > 0:
            val = PC[0] * r(t-1, u, v, n+1)
            if t > 1:
                val += (t - 1) * r(t-2, u, v, n+1)
        elif u > 0:
            val = PC[1] * r(t, u-1, v, n+1)
            if u > 1:
                val += (u - 1) * r(t, u-2, v, n+1)
        else:
            val = PC[2] * r(t, u, v-1, n+1)
            if v > 1:
                val += (v - 1) * r(t, u, v-2, n+1)
        R[key] = val
        return val
    return r


def int1e_sorted(sm, coords, charges):
    '''(S, T, V) over the sorted Cartesian basis ``sm`` (SortedGTO).'''
    n = sm.nao
    S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
    bas, env, ao_loc = sm._bas, sm.env_kernel, sm.ao_loc
    for ish in range(sm.nbas):
        for jsh in range(sm.nbas):
            li, lj = bas[ish, 1], bas[jsh, 1]
            A = env[bas[ish, 7]:bas[ish, 7]+3]
            B = env[bas[jsh, 7]:bas[jsh, 7]+3]
            ci = cart_components(li); cj = cart_components(lj)
            blkS = np.zeros((len(ci), len(cj))); blkT = np.zeros_like(blkS); blkV = np.zeros_like(blkS)
            for ip in range(bas[ish, 2]):
                for jp in range(bas[jsh, 2]):
                    a = env[bas[ish, 5] + ip]; b = env[bas[jsh, 5] + jp]
                    cc = env[bas[ish, 6] + ip] * env[bas[jsh, 6] + jp]
                    p = a + b
                    P = (a * A + b * B) / p
                    Es = [_hermite_E(li, lj + 2, a, b, A[d] - B[d]) for d in range(3)]
                    s1 = math.sqrt(math.pi / p)
                    for x, (ix, iy, iz) in enumerate(ci):
                        for y, (jx, jy, jz) in enumerate(cj):
                            pw = [(ix, jx), (iy, jy), (iz, jz)]
                            o = [Es[d][pw[d][0], pw[d][1], 0] * s1 for d in range(3)]
                            k = []
                            for d in range(3):
                                i_, j_ = pw[d]
                                kv = -2 * b * b * Es[d][i_, j_+2, 0] * s1 + b * (2*j_+1) * Es[d][i_, j_, 0] * s1
                                if j_ >= 2:
                                    kv -= .5 * j_ * (j_ - 1) * Es[d][i_, j_-2, 0] * s1
                                k.append(kv)
                            blkS[x, y] += cc * o[0] * o[1] * o[2]
                            blkT[x, y] += cc * (k[0]*o[1]*o[2] + o[0]*k[1]*o[2] + o[0]*o[1]*k[2])
                    for C, Z in zip(coords, charges):
                        r = _hermite_R(li + lj, p, P - C)
                        for x, (ix, iy, iz) in enumerate(ci):
                            for y, (jx, jy, jz) in enumerate(cj):
                                v = 0.
                                for t in range(ix + jx + 1):
                                    for u in range(iy + jy + 1):
                                        for w in range(iz + jz + 1):
                                            v += (Es[0][ix, jx, t] * Es[1][iy, jy, u] *
                                                  Es[2][iz, jz, w] * r(t, u, w, 0))
                                blkV[x, y] += -Z * 2 * math.pi / p * cc * v
            sl = (slice(ao_loc[ish], ao_loc[ish+1]), slice(ao_loc[jsh], ao_loc[jsh+1]))
            S[sl] = blkS; T[sl] = blkT; V[sl] = blkV
    return S, T, V


def rhf(mol, conv_tol=1e-10, max_cycle=100, dm0=None, jk=None, verbose=False):
    '''Oracle RHF energy (and density) of ``mol``.'''
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    from oracle import oracle as O
    sm = SortedGTO(mol)
    cache = O.build_pairs(sm)
    C = sm.coeff
    S, T, V = int1e_sorted(sm, mol.atom_coords(), mol.atom_charges().astype(float))
    S = C.T @ S @ C; h = C.T @ (T + V) @ C
    w, U = np.linalg.eigh(S)
    keep = w > 1e-6
    X = U[:, keep] / np.sqrt(w[keep])
    nocc = mol.nelectron // 2
    e_nuc = mol.energy_nuc()

    def get_jk(dm):
        dc = C @ dm @ C.T
        vj, vk, _, _ = O.jk_sorted(sm, dc[None], 1, cache)
        return C.T @ vj[0] @ C, C.T @ vk[0] @ C

    def solve(f):
        e, c = np.linalg.eigh(X.T @ f @ X)
        c = X @ c
        return e, c, 2 * c[:, :nocc] @ c[:, :nocc].T

    if dm0 is None:
        _, _, dm = solve(h)
    else:
        dm = dm0
    vj, vk = get_jk(dm)
    vhf = vj - .5 * vk
    e = np.sum(dm * (h + .5 * vhf)) + e_nuc
    fs, es = [], []
    for cycle in range(max_cycle):
        dm_last, e_last = dm, e
        f = h + vhf
        if cycle >= 1:
            sdf = S @ dm @ f
            err = X.T @ (sdf.T - sdf) @ X
            fs.append(f); es.append(err)
            fs, es = fs[-8:], es[-8:]
            n = len(fs)
            B = np.zeros((n + 1, n + 1))
            for i in range(n):
                for j in range(n):
                    B[i, j] = np.sum(es[i] * es[j])
            B[n, :n] = B[:n, n] = -1
            rhs = np.zeros(n + 1); rhs[n] = -1
            c = np.linalg.solve(B, rhs)
            f = sum(c[i] * fs[i] for i in range(n))
        mo_e, mo_c, dm = solve(f)
        vj, vk = get_jk(dm - dm_last)
        vhf = vhf + vj - .5 * vk
        e = np.sum(dm * (h + .5 * vhf)) + e_nuc
        f = h + vhf
        g = mo_c[:, nocc:].T @ f @ mo_c[:, :nocc]
        if verbose:
            print(cycle, e, e - e_last, np.linalg.norm(g) * 2)
        if abs(e - e_last) < conv_tol and np.linalg.norm(g) * 2 < math.sqrt(conv_tol):
            return e, dm
    raise RuntimeError('oracle RHF did not converge')
