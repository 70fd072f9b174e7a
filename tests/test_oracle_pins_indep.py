'''Independent pins of oracle pieces that the kernels are checked against:

* the MLP backward and AdamW of oracle/mlp_oracle.py against torch autograd
  and torch.optim.AdamW (float64, CPU) and against central finite
  differences;
* the screening (overlap mask, q_cond bins, survival predicate of
  gpu4pyscf/lib/gvhf-rys/create_tasks.cu:155-170): every quartet it drops is
  checked with brute-force ERIs, and every shell pair the overlap mask drops
  with exact overlap integrals.
'''
import math

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from oracle import oracle as O


def _net(rng, dims):
    Ws = [rng.standard_normal((a, b)) / math.sqrt(a) for a, b in zip(dims[:-1], dims[1:])]
    bs = [.1 * rng.standard_normal(b) for b in dims[1:]]
    return Ws, bs


def _torch_loss(x, Ws, bs, t, mask, scale):
    import torch
    a = x
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = a @ W + b
        a = torch.nn.functional.silu(z) if l < len(Ws) - 1 else z
    return scale * torch.sum(mask * (a - t) ** 2)


def test_mlp_backward_matches_torch_autograd():
    '''Without the bf16 rounding the oracle's forward/backward is the float64
    network: its weight and bias gradients equal torch autograd's.'''
    import torch
    rng = np.random.default_rng(11)
    dims = [7, 9, 8, 9, 5]
    Ws, bs = _net(rng, dims)
    x = rng.standard_normal((13, dims[0]))
    t = rng.standard_normal((13, dims[-1]))
    mask = (rng.random((13, dims[-1])) > .3).astype(np.float64)
    scale = .37
    y, zs, acts = MO.forward(x, Ws, bs, rnd=MO.exact)
    loss, dy = MO.mse_grad(y, t, mask, scale)
    dWs, dbs = MO.backward(dy, zs, acts, Ws, rnd=MO.exact)
    tW = [torch.tensor(W, requires_grad=True) for W in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    L = _torch_loss(torch.tensor(x), tW, tb, torch.tensor(t), torch.tensor(mask), scale)
    L.backward()
    assert abs(float(L.detach()) - loss) < 1e-12 * abs(loss)
    for l in range(len(Ws)):
        gW, gb = tW[l].grad.numpy(), tb[l].grad.numpy()
        assert abs(dWs[l] - gW).max() < 1e-12 * max(1., abs(gW).max()), l
        assert abs(dbs[l] - gb).max() < 1e-12 * max(1., abs(gb).max()), l


def test_mlp_backward_central_differences():
    '''Gradient entries of every layer against central finite differences of
    the float64 loss.'''
    rng = np.random.default_rng(12)
    dims = [4, 6, 6, 3]
    Ws, bs = _net(rng, dims)
    x = rng.standard_normal((5, dims[0]))
    t = rng.standard_normal((5, dims[-1]))
    mask = np.ones((5, dims[-1]))

    def loss_of(Ws_, bs_):
        y = MO.forward(x, Ws_, bs_, rnd=MO.exact)[0]
        return MO.mse_grad(y, t, mask, 1.)[0]

    y, zs, acts = MO.forward(x, Ws, bs, rnd=MO.exact)
    dWs, dbs = MO.backward(MO.mse_grad(y, t, mask, 1.)[1], zs, acts, Ws, rnd=MO.exact)
    h = 1e-6
    for l in range(len(Ws)):
        for (a, b) in [(0, 0), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1), (1, 2)]:
            Wp = [W.copy() for W in Ws]; Wm = [W.copy() for W in Ws]
            Wp[l][a, b] += h; Wm[l][a, b] -= h
            fd = (loss_of(Wp, bs) - loss_of(Wm, bs)) / (2 * h)
            assert abs(fd - dWs[l][a, b]) < 1e-7 * max(1., abs(fd)), (l, a, b)
        bp = [b.copy() for b in bs]; bm = [b.copy() for b in bs]
        bp[l][1] += h; bm[l][1] -= h
        fd = (loss_of(Ws, bp) - loss_of(Ws, bm)) / (2 * h)
        assert abs(fd - dbs[l][1]) < 1e-7 * max(1., abs(fd)), l


def test_mlp_planted_mistakes_fail():
    '''The autograd pin catches the mistakes it is meant to catch: a SiLU
    derivative without the z(1 - s) term, and a missing transpose.'''
    import torch
    rng = np.random.default_rng(13)
    dims = [6, 6, 6, 4]
    Ws, bs = _net(rng, dims)
    x = rng.standard_normal((8, 6))
    t = rng.standard_normal((8, 4))
    mask = np.ones((8, 4))
    y, zs, acts = MO.forward(x, Ws, bs, rnd=MO.exact)
    dy = MO.mse_grad(y, t, mask, 1.)[1]
    tW = [torch.tensor(W, requires_grad=True) for W in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    _torch_loss(torch.tensor(x), tW, tb, torch.tensor(t), torch.tensor(mask), 1.).backward()
    ref = tW[0].grad.numpy()

    def wrong_silu_backward():
        d = dy
        for l in range(len(Ws) - 1, 0, -1):
            dA = d @ Ws[l].T
            s = 1. / (1. + np.exp(-zs[l - 1]))
            d = dA * s                      # planted: z s (1 - s) term dropped
        return acts[0].T @ d

    def wrong_transpose():
        d = dy
        for l in range(len(Ws) - 1, 0, -1):
            dA = d @ Ws[l].T.T.T if Ws[l].shape[0] != Ws[l].shape[1] else d @ Ws[l]   # planted
            z = zs[l - 1]
            s = 1. / (1. + np.exp(-z))
            d = dA * s * (1. + z * (1. - s))
        return acts[0].T @ d

    assert abs(MO.backward(dy, zs, acts, Ws, rnd=MO.exact)[0][0] - ref).max() < 1e-12
    assert abs(wrong_silu_backward() - ref).max() > 1e-3
    assert abs(wrong_transpose() - ref).max() > 1e-3


def test_mlp_bf16_rounding_only_perturbs():
    '''With the bf16 rounding points the gradients stay within bf16 accuracy
    of the float64 network's.'''
    rng = np.random.default_rng(14)
    dims = [16, 32, 32, 8]
    Ws, bs = _net(rng, dims)
    x = rng.standard_normal((64, 16))
    t = rng.standard_normal((64, 8))
    mask = np.ones((64, 8))
    ye, zse, actse = MO.forward(x, Ws, bs, rnd=MO.exact)
    ge = MO.backward(MO.mse_grad(ye, t, mask, 1.)[1], zse, actse, Ws, rnd=MO.exact)[0]
    yb, zsb, actsb = MO.forward(x, Ws, bs)
    gb = MO.backward(MO.mse_grad(yb, t, mask, 1.)[1], zsb, actsb, Ws)[0]
    for a, b in zip(ge, gb):
        assert abs(a - b).max() < 3e-2 * abs(a).max()


def test_adamw_matches_torch_optim():
    '''oracle adamw == torch.optim.AdamW (decoupled weight decay, bias
    correction) over several steps in float64.'''
    import torch
    rng = np.random.default_rng(15)
    w0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(6)]
    lr, wd = 1e-3, 1e-4
    p = torch.nn.Parameter(torch.tensor(w0))
    opt = torch.optim.AdamW([p], lr=lr, betas=(.9, .999), eps=1e-8, weight_decay=wd)
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    for step, g in enumerate(grads, 1):
        p.grad = torch.tensor(g)
        opt.step()
        w, m, v = MO.adamw(w, m, v, g, step, lr, wd=wd)
        assert abs(w - p.detach().numpy()).max() < 1e-15 * max(1., abs(w).max()) + 1e-16


# ---------------------------------------------------------------- screening
def _spread_waters(basis):
    from gpu4pyscf_amd.gto.mole import M
    atoms = []
    for n, x in enumerate((0., 5.5, 13.)):
        atoms += [f'O {x} 0 {0.1 * n}', f'H {x + 0.757} 0.586 {0.1 * n}',
                  f'H {x - 0.757} 0.586 {0.1 * n}']
    return M(atom='; '.join(atoms), basis=basis)


def _block_max(D, ao_loc, a, b):
    return float(abs(D[ao_loc[a]:ao_loc[a + 1], ao_loc[b]:ao_loc[b + 1]]).max())


# the density blocks each survival screen reads: (i,j), (k,l) for J;
# (i,k), (j,k), (i,l), (j,l) for K (create_tasks.cu:155-163, 222-225, 86-93)
SCREEN_BLOCKS = {'jk': ('ij', 'kl', 'ik', 'jk', 'il', 'jl'), 'j': ('ij', 'kl'),
                 'k': ('ik', 'jk', 'il', 'jl')}


@pytest.mark.parametrize('basis,screen', [('sto-3g', 'jk'), ('def2-svp', 'jk'),
                                          ('sto-3g', 'j'), ('sto-3g', 'k')])
def test_every_dropped_quartet_is_negligible(basis, screen):
    '''Brute force over every shell quartet (ij) >= (kl) of the sorted basis
    of three waters 5.5 and 7.5 Angstrom apart: a quartet that is not in the
    surviving set of the screen -- dropped by the overlap mask, by q_cond
    bin 0 or by the predicate -- contributes at most |(ij|kl)| max|D| <=
    direct_scf_tol to what the screen builds (D over the density blocks that
    build reads: all six for J/K, D_ij and D_kl for J, the four exchange
    blocks for K), up to the float32 rounding of the log-domain bound.'''
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = _spread_waters(basis)
    sm = SortedGTO(mol)
    tol = 1e-13
    cache = O.build_pairs(sm, tol)
    rng = np.random.default_rng(16)
    g = rng.standard_normal((sm.nao, sm.nao))
    # an incremental-like density: small, with decaying long-range blocks
    cen = sm.bas_coords()[np.searchsorted(sm.ao_loc, np.arange(sm.nao), side='right') - 1]
    r = np.linalg.norm(cen[:, None] - cen[None], axis=-1)
    D = 1e-3 * (g + g.T) / 2 * np.exp(-.3 * r)
    keys = set(O.survivors(sm, D[None], 1, cache, screen=screen).tolist())
    n = sm.nbas
    ao = sm.ao_loc
    bas, env = sm._bas, sm.env_kernel
    worst = 0.
    ndropped = 0
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    for a, (i, j) in enumerate(pairs):
        bij = i * n + j
        for (k, l) in pairs[:a + 1]:
            bkl = k * n + l
            if bij * n * n + bkl in keys:
                continue
            ndropped += 1
            sh = dict(i=i, j=j, k=k, l=l)
            dmax = max(_block_max(D, ao, sh[b[0]], sh[b[1]]) for b in SCREEN_BLOCKS[screen])
            eri = abs(O.eri_shell(bas, env, i, j, k, l)).max()
            worst = max(worst, eri * dmax)
    print(f'{basis} screen {screen}: {len(keys)} survivors, {ndropped} dropped, '
          f'max dropped |(ij|kl)| max|D| = {worst:.3e}')
    assert ndropped > len(keys) // 10
    assert worst <= tol * (1 + 1e-5)


def test_screens_partition_the_jk_set():
    '''The J/K set is the union of the J-only and K-only sets (the six blocks
    are the two Coulomb blocks and the four exchange blocks), each a strict
    subset of it for a density with long-range decay; J and K of the
    one-sided builds equal those of the J/K build within the cutoff.'''
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = _spread_waters('def2-svp')
    sm = SortedGTO(mol)
    cache = O.build_pairs(sm, 1e-13)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((sm.nao, sm.nao))
    cen = sm.bas_coords()[np.searchsorted(sm.ao_loc, np.arange(sm.nao), side='right') - 1]
    r = np.linalg.norm(cen[:, None] - cen[None], axis=-1)
    D = (1e-2 * (g + g.T) / 2 * np.exp(-.8 * r))[None]
    s = {sc: set(O.survivors(sm, D, 1, cache, screen=sc).tolist()) for sc in ('jk', 'j', 'k')}
    assert s['jk'] == s['j'] | s['k']
    assert s['j'] < s['jk'] and s['k'] < s['jk']
    vj, vk, n_jk, _ = O.jk_sorted(sm, D, 1, cache)
    vj1, _, n_j, _ = O.jk_sorted(sm, D, 1, cache, with_k=False, screen='j')
    _, vk1, n_k, _ = O.jk_sorted(sm, D, 1, cache, with_j=False, screen='k')
    assert (n_jk, n_j, n_k) == (len(s['jk']), len(s['j']), len(s['k']))
    assert abs(vj1 - vj).max() < 1e-11 and abs(vk1 - vk).max() < 1e-11
    # a J-only build of a non-symmetric density symmetrises it first
    A = D + 1e-3 * rng.standard_normal(D.shape)
    vj0 = O.jk_sorted(sm, A, 0, cache, with_k=False)[0]
    vj2 = O.jk_sorted(sm, A, 0, cache, with_k=False, screen='j')[0]
    assert abs(vj2 - vj0).max() < 1e-11


def test_overlap_mask_drops_only_negligible_pairs():
    """Every shell pair the float32 overlap bound (overlap.cu:2163-2223)
    drops is negligible by exact McMurchie-Davidson overlaps:
    * dropped by the log bound (theta |Ri - Rj|^2 <= 50): exact max |S| of
      the block below the mask threshold direct_scf_tol^2 (none for these
      bases, where only the remote rule drops pairs);
    * dropped by the remote rule (theta |Ri - Rj|^2 > 50, theta of the most
      diffuse primitives): exact max |S| below 1e3 exp(-50) ~ 2e-19, the
      Gaussian decay times the largest normalised prefactor of these bases."""
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    from oracle import scf_oracle as SO
    mol = _spread_waters('def2-svp')
    sm = SortedGTO(mol)
    prec = 1e-13 ** 2
    mask = O.overlap_mask(sm, prec)
    S = SO.int1e_sorted(sm, np.asarray(mol.atom_coords()), np.asarray(mol.atom_charges()))[0]
    exps, _ = sm.diffuse_pgto()
    co = sm.bas_coords()
    ao = sm.ao_loc
    n = sm.nbas
    kept = 0
    worst_log = worst_remote = 0.
    n_log = n_remote = 0
    for i in range(n):
        for j in range(i + 1):
            if mask[i, j]:
                kept += 1
                continue
            s = _block_max(S, ao, i, j)
            theta = exps[i] * exps[j] / (exps[i] + exps[j])
            if theta * float(((co[i] - co[j]) ** 2).sum()) > 50:
                n_remote += 1
                worst_remote = max(worst_remote, s)
            else:
                n_log += 1
                worst_log = max(worst_log, s)
    print(f'overlap mask: {kept} kept; log bound dropped {n_log} (max |S| {worst_log:.3e}), '
          f'remote rule dropped {n_remote} (max |S| {worst_remote:.3e})')
    # for these bases the remote rule is what drops pairs; the log bound
    # binds only for tight primitives with small coefficients
    assert n_remote > 0
    assert worst_log < prec
    assert worst_remote < 1e3 * math.exp(-50)


def test_survivors_within_float32_rounding_of_the_cutoff():
    '''How many quartets could change sides if q_cond were rounded as the
    reference's float32 estimator rounds it (nr_sr_estimator.cu:100-397)
    rather than from FP64: on the 8-water def2-SVP cluster with an
    incremental density, the quartets whose log bound lies within
    +-4e-6 (two float32 ulps at |log 1e-13| = 29.9) of the cutoff are at
    most 1e-4 of the survivors.'''
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom=G.to_string(G.water_cluster((2, 2, 2), 3.0, seed=0)), basis='def2-svp')
    sm = SortedGTO(mol, diffuse_cutoff=0.3)
    cache = O.build_pairs(sm, 1e-13)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((sm.nao, sm.nao))
    D = (1e-3 * (g + g.T) / 2)[None]
    delta = 4e-6
    n = {s: len(O.survivors(sm, D, 1, cache, direct_scf_tol=1e-13 * math.exp(s * delta)))
         for s in (-1, 0, 1)}
    band = n[-1] - n[1]
    print(f'8-water def2-SVP: {n[0]} survivors, {band} within +-{delta} of the cutoff')
    assert n[-1] >= n[0] >= n[1]
    assert band <= 1e-4 * n[0]
