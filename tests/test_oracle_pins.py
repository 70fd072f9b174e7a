'''Pins of the CPU oracle beyond the reference's golden vectors: closed
forms, an independent brute-force ERI algorithm, and ground-truth Rys
quadratures.'''
import math
import numpy as np
import pytest
import mpmath as mp

from oracle import oracle as O


def _boys(m, x):
    mp.mp.dps = 30
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, x) / (2 * x**(m + mp.mpf(1) / 2))


@pytest.mark.parametrize('n', [1, 2, 4, 7, 9, 11])
@pytest.mark.parametrize('x', [0., 1e-7, .3, 2.5, 11., 37., 80., 250.])
def test_oracle_rys_moments(n, x):
    u, w = O.rys_roots(n, x)
    for m in range(2 * n):
        ref = _boys(m, x)
        val = sum(w[i] * u[i]**m for i in range(n))
        assert abs(val - float(ref)) <= 2e-14 * float(ref)


def test_rys_limits():
    # x -> 0: nodes are squared positive roots of the Legendre P_2n
    u, w = O.rys_roots(1, 0.)
    assert abs(u[0] - 1. / 3) < 1e-15 and abs(w[0] - 1.) < 1e-15
    s, ws = np.polynomial.legendre.leggauss(6)
    ref = np.sort(s[s > 0]**2)
    u, w = O.rys_roots(3, 0.)
    assert abs(u - ref).max() < 1e-14
    # x -> inf: squared positive Hermite roots / x
    x = 400.
    s, ws = np.polynomial.hermite.hermgauss(8)
    u, w = O.rys_roots(4, x)
    assert abs(u * x - np.sort(s[s > 0]**2)).max() < 1e-11


def test_product_table_matches_oracle():
    '''The product's Chebyshev Rys table against the oracle's roots.'''
    t = np.load('gpu4pyscf_amd/csrc/rys_table.npz')
    coef, off, xmax, D = t['coef'], t['off'], t['xmax'], int(t['degree'])
    lr, lw = t['large_r'], t['large_w']
    rng = np.random.default_rng(0)
    for n in range(1, 12):
        for x in np.concatenate([rng.random(8) * 110, [0., xmax[n-1] - 1e-9, xmax[n-1]]]):
            if x >= xmax[n-1]:
                b = n * (n - 1) // 2
                r, w = lr[b:b+n] / x, lw[b:b+n] / math.sqrt(x)
            else:
                k = int(x)
                c = coef[off[n-1] + k*2*n*(D+1): off[n-1] + (k+1)*2*n*(D+1)].reshape(2*n, D+1)
                v = c @ np.cos(np.arange(D+1) * np.arccos(2 * (x - k) - 1))
                r, w = v[0::2], v[1::2]
            r0, w0 = O.rys_roots(n, x)
            assert abs(r - r0).max() <= 5e-14 * abs(r0).max()
            assert abs(w - w0).max() <= 5e-14 * abs(w0).max()


def test_boys_grid_matches_incomplete_gamma():
    '''Boys grid of the product table: F_m(x_g) from mpmath's incomplete
    gamma function, an evaluation independent of the generator's series.'''
    t = np.load('gpu4pyscf_amd/csrc/rys_table.npz')
    boys, step = t['boys'], float(t['boys_step'])
    assert boys.shape[1] == 13 and (boys.shape[0] - 1) * step >= t['xmax'][2]
    for g in (0, 1, 7, 100, 233, 389, boys.shape[0] - 1):
        for m in (0, 1, 3, 6, 10, 12):
            ref = float(_boys(m, g * step))
            assert abs(boys[g, m] - ref) <= 1e-15 * ref


def test_boys_grid_roots_match_oracle():
    '''One- to three-root Gauss rules as the per-lane kernels form them: the
    8-term Taylor series of the Boys moments about the nearest grid point,
    then the rule of the moments (closed form for one and two roots;
    Chebyshev's algorithm, Viete's cubic and two Newton steps for three),
    against the oracle's roots.'''
    t = np.load('gpu4pyscf_amd/csrc/rys_table.npz')
    boys, step, xmax = t['boys'], float(t['boys_step']), t['xmax']
    inv_fact = [1. / math.factorial(k) for k in range(8)]

    def moments(x, nm):
        g = min(int(x / step + .5), boys.shape[0] - 1)
        d = g * step - x
        out = []
        for m in range(nm):
            s = boys[g, m + 7] * inv_fact[7]
            for k in range(6, -1, -1):
                s = s * d + boys[g, m + k] * inv_fact[k]
            out.append(s)
        return out

    rng = np.random.default_rng(3)
    xs = np.concatenate([[0., 1e-9, .05, .0500001, 38.95, 46.95, 53.95],
                         rng.random(40) * 54., rng.random(20)])
    for x in xs:
        if x < xmax[0]:
            f0, f1 = moments(x, 2)
            u0, w0 = O.rys_roots(1, x)
            assert abs(f1 / f0 - u0[0]) <= 1e-14 * u0[0]
            assert abs(f0 - w0[0]) <= 1e-14 * w0[0]
        if x < xmax[1]:
            f0, f1, f2, f3 = moments(x, 4)
            det = f1 * f1 - f0 * f2
            c1 = (f0 * f3 - f1 * f2) / det
            c0 = (f2 * f2 - f1 * f3) / det
            disc = math.sqrt(c1 * c1 - 4 * c0)
            q = -.5 * (c1 + math.copysign(disc, c1))
            u1, u2 = sorted((q, c0 / q))
            w1 = (f1 - u2 * f0) / (u1 - u2)
            u0, w0 = O.rys_roots(2, x)
            assert abs(np.array([u1, u2]) - u0).max() <= 2e-14 * abs(u0).max()
            assert abs(np.array([w1, f0 - w1]) - w0).max() <= 2e-14 * abs(w0).max()
        if x < xmax[2]:
            mu = moments(x, 6)
            a0, b0 = mu[1] / mu[0], mu[0]
            n1, n2 = mu[2] - a0 * mu[1], mu[3] - a0 * mu[2]
            n3, n4 = mu[4] - a0 * mu[3], mu[5] - a0 * mu[4]
            a1, b1 = n2 / n1 - a0, n1 / mu[0]
            m2, m3 = n3 - a1 * n2 - b1 * mu[2], n4 - a1 * n3 - b1 * mu[3]
            a2, b2 = m3 / m2 - n2 / n1, m2 / n1
            c2 = -(a0 + a1 + a2)
            c1 = a0 * a1 + a0 * a2 + a1 * a2 - b1 - b2
            c0 = b2 * a0 + b1 * a2 - a0 * a1 * a2
            p = c1 - c2 * c2 / 3
            q = 2 * c2**3 / 27 - c2 * c1 / 3 + c0
            rr = math.sqrt(-p / 3)
            phi = math.acos(min(1., max(-1., -q / (2 * rr**3))))
            us, ws = [], []
            for k in range(3):
                v = 2 * rr * math.cos((phi - 2 * math.pi * k) / 3) - c2 / 3
                for _ in range(2):
                    p1 = v - a0
                    p2, d2 = (v - a1) * p1 - b1, p1 + (v - a1)
                    p3, d3 = (v - a2) * p2 - b2 * p1, p2 + (v - a2) * d2 - b2
                    v -= p3 / d3
                p1 = v - a0
                p2 = (v - a1) * p1 - b1
                ws.append(1 / (1 / b0 + p1 * p1 / (b0 * b1) + p2 * p2 / (b0 * b1 * b2)))
                us.append(v)
            u0, w0 = O.rys_roots(3, x)
            assert abs(np.array(us[::-1]) - u0).max() <= 3e-14 * abs(u0).max()
            assert abs(np.array(ws[::-1]) - w0).max() <= 3e-14 * abs(w0).max()


def test_ssss_closed_form():
    '''(ss|ss) of unnormalised-free s Gaussians: 2 pi^2.5/(pq sqrt(p+q)) K F0.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom='H 0 0 0; H 0 0 1.4; H 0.3 1.1 0; H 1 1 1', unit='B',
            basis=[[0, [1.3, 1.]]])
    sm = SortedGTO(mol)
    a = 1.3
    N = (2 * a / math.pi)**.75
    R = mol.atom_coords()
    eri = O.eri_shell(sm._bas, sm.env_kernel, 0, 1, 2, 3).ravel()[0]
    p = q = 2 * a
    P = (R[0] + R[1]) / 2; Q = (R[2] + R[3]) / 2
    K = math.exp(-a / 2 * np.sum((R[0] - R[1])**2)) * math.exp(-a / 2 * np.sum((R[2] - R[3])**2))
    T = p * q / (p + q) * np.sum((P - Q)**2)
    F0 = .5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    ref = N**4 * 2 * math.pi**2.5 / (p * q * math.sqrt(p + q)) * K * F0
    assert abs(eri - ref) < 1e-14


def test_rys_eri_vs_brute_force():
    '''Oracle Rys shell quartets against the McMurchie-Davidson brute force.'''
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    from oracle import brute
    mol = M(atom='O 0 0 0.2; H 0 1.3 -0.4; C 1.1 -0.2 0.5', unit='B', spin=1,
            basis=[[0, [3.1, .5], [.9, .6]], [1, [1.7, 1.]], [2, [.8, 1.]], [3, [.6, 1.]]])
    sm = SortedGTO(mol)
    rng = np.random.default_rng(3)
    for _ in range(12):
        i, j, k, l = rng.integers(0, sm.nbas, 4)
        a = O.eri_shell(sm._bas, sm.env_kernel, i, j, k, l)
        b = brute.eri_shell(sm._bas, sm.env_kernel, i, j, k, l)
        assert abs(a - b).max() < 1e-12 * max(1., abs(b).max())


def test_eri_8fold_symmetry_and_jk_identities():
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom='O 0 0 0.1174; H -0.757 0 -0.4696; H 0.757 0 -0.4696', basis='sto-3g')
    sm = SortedGTO(mol)
    eri = O.eri_full(sm)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert abs(eri - eri.transpose(perm)).max() < 1e-14
    # J/K of the oracle pipeline equal the einsum of the full tensor
    C = sm.coeff
    rng = np.random.default_rng(0)
    d = rng.random((mol.nao, mol.nao)); d = d + d.T
    vj, vk = O.get_jk(mol, d, hermi=1)
    dc = C @ d @ C.T
    rj = C.T @ np.einsum('ijkl,lk->ij', eri, dc) @ C
    rk = C.T @ np.einsum('ijkl,jk->il', eri, dc) @ C
    assert abs(vj - rj).max() < 1e-12 and abs(vk - rk).max() < 1e-12
    # linearity and Tr(D J[D]) > 0
    vj2, _ = O.get_jk(mol, 2 * d, hermi=1)
    assert abs(vj2 - 2 * vj).max() < 1e-12
    assert np.sum(d * vj) > 0


def test_q_cond_is_log_sqrt_max_diagonal():
    """q_cond of every listed pair equals log sqrt max|(ij|ij)| of the exact
    block (the Cauchy-Schwarz factor; the dropped quartets themselves are
    checked by brute force in test_oracle_pins_indep.py)."""
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom='O 0 0 0; H 0 0 1.8; O 0 0 9.; H 0 1.7 9.5', unit='B', basis='sto-3g')
    sm = SortedGTO(mol)
    cache = O.build_pairs(sm, 1e-13)
    n = sm.nbas
    npairs = 0
    for (gi, gj), (p, q) in cache.items():
        for pij, qq in zip(p, q):
            i, j = divmod(int(pij), n)
            blk = O.eri_shell(sm._bas, sm.env_kernel, i, j, i, j)
            assert abs(math.log(abs(blk).max()) / 2 - qq) < 1e-5
            npairs += 1
    assert npairs > 20


def test_splitmix64_published_sequence():
    '''The checksum mix is the SplitMix64 finaliser: the generator built on it
    (state += 0x9e3779b97f4a7c15, output mix(state)) from seed 1234567 gives
    the published sequence of Vigna's splitmix64.c.'''
    s, out = 1234567, []
    for _ in range(5):
        s = (s + 0x9e3779b97f4a7c15) % 2**64
        out.append(O.splitmix64_mix(s))
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_quartet_set_checksums_match_survivor_list():
    '''Per-task survivor counts and set checksums of the OpenMP scan equal
    those recomputed from the serial survivor list (8-water def2-SVP,
    random symmetric density).'''
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom=G.to_string(G.water_cluster((2, 2, 2), 3.0, seed=0)), basis='def2-svp')
    sm = SortedGTO(mol)
    cache = O.build_pairs(sm, 1e-13)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((1, sm.nao, sm.nao))
    d = 1e-3 * (g + g.transpose(0, 2, 1))
    sets = O.quartet_sets(sm, d, 1, cache)
    keys = O.survivors(sm, d, 1, cache)
    nb = sm.nbas
    bij, bkl = keys // np.uint64(nb * nb), keys % np.uint64(nb * nb)
    sh = np.stack([bij // np.uint64(nb), bij % np.uint64(nb),
                   bkl // np.uint64(nb), bkl % np.uint64(nb)]).astype(np.int64)
    grp = np.searchsorted(np.asarray(sm.group_offsets), sh, side='right') - 1
    z = (bij << np.uint64(32)) | bkl
    with np.errstate(over='ignore'):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    z = z ^ (z >> np.uint64(31))
    ref = {}
    for t, h in zip(map(tuple, grp.T.tolist()), z.tolist()):
        c, hh = ref.get(t, (0, 0))
        ref[t] = (c + 1, (hh + h) % 2**64)
    assert len(keys) > 10**6
    assert sets == ref


def test_flops_model():
    from gpu4pyscf_amd.scf.flops import quartet_flops, quartet_digest_bytes
    # (ss|ss), one primitive: 1 root, 3 + 9 flops + 12 digestion
    assert quartet_flops((0, 0, 0, 0), (1, 1, 1, 1)) == 24
    # (pp|pp): n=3, Nf=81, G2=9
    assert quartet_flops((1, 1, 1, 1), (1, 1, 1, 1)) == 3 * (243 + 81) + 12 * 81
    assert quartet_digest_bytes((0, 0, 0, 0)) == 96


def test_md_boys_grid_matches_incomplete_gamma():
    '''The J-engine's Boys grid (csrc/gen_md_boys.py, mpmath) against
    scipy's regularised lower incomplete gamma: F_m(x) =
    gamma(m+1/2) P(m+1/2, x) / (2 x^(m+1/2)).'''
    import os
    from scipy.special import gammainc, gamma
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        'gpu4pyscf_amd', 'csrc', 'md_boys_table.npz')
    t = np.load(path)
    b, step = t['boys'], float(t['step'])
    assert b.shape == (601, 20)
    assert np.allclose(b[0], 1. / (2 * np.arange(20) + 1), rtol=0, atol=1e-15)
    for g in (1, 7, 123, 333, 600):
        x = g * step
        for m in (0, 3, 12, 19):
            ref = gamma(m + .5) * gammainc(m + .5, x) / (2 * x**(m + .5))
            assert abs(b[g, m] - ref) <= 1e-13 * ref, (g, m)
