'''Generate the Chebyshev tables of Rys roots and weights used by the kernels.

For n roots the Rys quadrature of the measure exp(-x t^2) dt on t in [0, 1]
is sought in the variable u = t^2: nodes u_i and weights w_i with
sum_i w_i u_i^m = F_m(x) for m < 2n, where F_m is the Boys function.

Method (all in mpmath at 40 significant digits):

* Boys values F_0..F_{2n-1}(x): the series
  F_N(x) = exp(-x) sum_i (2x)^i / ((2N+1)(2N+3)...(2N+2i+1)) for the top order,
  then the downward recursion F_{m-1} = (2x F_m + exp(-x)) / (2m-1);
* recurrence coefficients of the monic orthogonal polynomials from the
  moments by Chebyshev's algorithm (Gautschi, "Orthogonal Polynomials:
  Computation and Approximation", 2004, Alg. 2.1);
* nodes: float64 eigenvalues of the Jacobi matrix refined by Newton steps on
  the three-term recurrence; weights by the Christoffel formula
  w_i = 1 / sum_k p_k(u_i)^2 / ||p_k||^2.

For x >= X_n the quadrature is replaced by its large-x limit
u_i = s_i^2 / x, w_i = W_i / sqrt(x) with (s_i, W_i) the positive half of the
2n-point Gauss-Hermite rule (computed from the moments Gamma(k+1/2)/2 of
exp(-s^2) ds on [0, inf) in the same way); X_n is the first integer where that limit agrees
with the exact quadrature to 1e-16 relative.

On [0, X_n) each root and weight is a Chebyshev series of degree DEGREE on
unit intervals [k, k+1).  Output: rys_table.npz with
  coef[off[n-1] + (k*2n + f)*(DEGREE+1) + d]   f = 2i (root i), 2i+1 (weight i)
  xmax[n-1] = X_n, off[n-1], large_r[n*(n-1)/2 + i], large_w[...].

For one to three roots the kernels instead take the moments F_0..F_{2n-1}
from a Boys grid, boys[g, m] = F_m(g * BOYS_STEP) for m <= BOYS_M - 1, by
the BOYS_TERMS-term Taylor series F_m(x) = sum_k F_{m+k}(x_g) (x_g - x)^k / k!
about the nearest grid point, and form the Gauss rule from the moments.

Run:  python gen_rys_table.py  (about a minute on 8 cores).
'''

import os
import sys
import numpy as np
import mpmath as mp

NMAX = 11
DEGREE = 11
WIDTH = 1.0
BOYS_STEP = 0.1
BOYS_TERMS = 8
BOYS_M = 5 + BOYS_TERMS        # F_0 .. F_12: moments of up to three roots
mp.mp.dps = 40


def boys(nmax, x):
    x = mp.mpf(x)
    if x == 0:
        return [mp.mpf(1) / (2*m + 1) for m in range(nmax + 1)]
    ex = mp.exp(-x)
    term = mp.mpf(1) / (2*nmax + 1)
    s = term
    i = 0
    while True:
        i += 1
        term *= 2*x / (2*nmax + 2*i + 1)
        s += term
        if term < s * mp.mpf(10)**(-mp.mp.dps):
            break
    f = [mp.mpf(0)] * (nmax + 1)
    f[nmax] = ex * s
    for m in range(nmax, 0, -1):
        f[m-1] = (2*x*f[m] + ex) / (2*m - 1)
    return f


def recurrence(mu, n):
    '''alpha[0..n-1], beta[0..n-1] from moments mu[0..2n-1].'''
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2*n)
    sig = list(mu)
    alpha[0] = mu[1] / mu[0]
    beta[0] = mu[0]
    for k in range(1, n):
        new = [mp.mpf(0)] * (2*n)
        for l in range(k, 2*n - k):
            new[l] = sig[l+1] - alpha[k-1]*sig[l] - beta[k-1]*sig_prev[l]
        alpha[k] = new[k+1]/new[k] - sig[k]/sig[k-1]
        beta[k] = new[k] / sig[k-1]
        sig_prev, sig = sig, new
    return alpha, beta


def quadrature(n, x):
    return quadrature_from_moments(boys(2*n - 1, x), n)


def quadrature_from_moments(mu, n):
    alpha, beta = recurrence(mu, n)
    J = np.diag([float(a) for a in alpha])
    for k in range(1, n):
        J[k, k-1] = J[k-1, k] = float(mp.sqrt(beta[k]))
    u0 = np.sort(np.linalg.eigvalsh(J))
    nodes, weights = [], []
    for r in u0:
        u = mp.mpf(r)
        for _ in range(6):
            p_prev, p = mp.mpf(0), mp.mpf(1)
            dp_prev, dp = mp.mpf(0), mp.mpf(0)
            for k in range(n):
                p_next = (u - alpha[k]) * p - (beta[k] * p_prev if k > 0 else 0)
                dp_next = p + (u - alpha[k]) * dp - (beta[k] * dp_prev if k > 0 else 0)
                p_prev, p = p, p_next
                dp_prev, dp = dp, dp_next
            step = p / dp
            u -= step
            if abs(step) < abs(u) * mp.mpf(10)**(-35):
                break
        # Christoffel weight
        s = mp.mpf(0)
        norm = mp.mpf(1)
        p_prev, p = mp.mpf(0), mp.mpf(1)
        for k in range(n):
            norm *= beta[k]
            s += p * p / norm
            p_next = (u - alpha[k]) * p - (beta[k] * p_prev if k > 0 else 0)
            p_prev, p = p, p_next
        nodes.append(u)
        weights.append(1 / s)
    return nodes, weights


def large_x(n):
    '''Nodes s_i^2 and weights W_i of the measure exp(-s^2) ds on [0, inf)
    in u = s^2, from its moments Gamma(k+1/2)/2.'''
    mu = [mp.gamma(k + mp.mpf(1)/2) / 2 for k in range(2*n)]
    r, w = quadrature_from_moments(mu, n)
    return r, w


def _values(args):
    n, x = args
    r, w = quadrature(n, x)
    out = []
    for i in range(n):
        out.extend([float(r[i]), float(w[i])])
    return out


def find_xmax(n):
    lr, lw = large_x(n)
    x = 10
    while True:
        r, w = quadrature(n, x)
        ok = True
        for i in range(n):
            er = abs(r[i] - lr[i]/x) / r[i]
            ew = abs(w[i] - lw[i]/mp.sqrt(x)) / w[i]
            if er > 1e-16 or ew > 1e-16:
                ok = False
                break
        if ok or x >= 200:
            return x
        x += 1


def main(out):
    from multiprocessing import Pool
    k = np.arange(DEGREE + 1)
    cheb_nodes = np.cos(np.pi * (k + .5) / (DEGREE + 1))      # on [-1, 1]
    T = np.cos(np.outer(k, np.arccos(cheb_nodes)))            # T_j(node)
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        xmax = pool.map(find_xmax, range(1, NMAX + 1))
        coefs = []
        offs = []
        large_r = []
        large_w = []
        off = 0
        for n in range(1, NMAX + 1):
            X = xmax[n-1]
            nint = int(np.ceil(X / WIDTH))
            pts = [(n, iv*WIDTH + (c + 1) * WIDTH / 2) for iv in range(nint) for c in cheb_nodes]
            vals = np.array(pool.map(_values, pts)).reshape(nint, DEGREE + 1, 2*n)
            c = 2. / (DEGREE + 1) * np.einsum('jk,ikf->ifj', T, vals)
            c[..., 0] *= .5
            coefs.append(c.ravel())
            offs.append(off)
            off += c.size
            lr, lw = large_x(n)
            large_r.extend([float(v) for v in lr])
            large_w.extend([float(v) for v in lw])
            print(f'n={n} xmax={X} intervals={nint}', flush=True)
        # Boys grid covering x < X_3 (nearest grid point of any such x)
        rows = int(xmax[2] / BOYS_STEP) + 2
        boys_grid = np.array([[float(v) for v in boys(BOYS_M - 1, g * BOYS_STEP)]
                              for g in range(rows)])
    np.savez(out, coef=np.concatenate(coefs), off=np.asarray(offs, dtype=np.int64),
             xmax=np.asarray(xmax, dtype=np.float64), degree=DEGREE, width=WIDTH,
             nmax=NMAX, large_r=np.asarray(large_r), large_w=np.asarray(large_w),
             boys=boys_grid, boys_step=BOYS_STEP, boys_terms=BOYS_TERMS)


if __name__ == '__main__':
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(
        os.path.dirname(os.path.abspath(__file__)), 'rys_table.npz')
    main(out)
