'''Learned initial guess: data generation, training and the guess callable.

Training data are converged RHF densities that this project's own GPU SCF
produces on seeded water clusters; every atom pair within 6 Angstrom (and
every atom with itself) contributes one row: its descriptor and its density
block in the pair's local frame (gpu4pyscf_amd/guess/features.py).  The
network (gpu4pyscf_amd/guess/mlp.py) is trained with AdamW by the BF16 MFMA
kernels.  At SCF start the predicted blocks are rotated back to the lab
frame, symmetrised and scaled to Tr(DS) = N_electrons; the atomic guess is
the fallback when no network is available, an element is unknown, or the
prediction is not finite.
'''

import time

import numpy as np

from gpu4pyscf_amd.guess import features as F
from gpu4pyscf_amd.guess.mlp import PairMLP, to_bf16

HIDDEN = (256, 256, 256)


def _torch():
    import torch
    return torch


def water_mol(seed, shape=(2, 2, 2), spacing=3.0, basis='def2-svp'):
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    return M(atom=G.to_string(G.water_cluster(shape, spacing, seed=seed)), basis=basis)


def converged_density(mol, conv_tol=1e-9, init_guess=None):
    from gpu4pyscf_amd.scf.hf import RHF
    mf = RHF(mol)
    mf.conv_tol = conv_tol
    if init_guess is not None:
        mf.init_guess = init_guess
    mf.kernel()
    if not mf.converged:
        raise RuntimeError('SCF for training data did not converge')
    return mf.dm.cpu().numpy(), mf


def make_dataset(seeds, shape=(2, 2, 2), spacing=3.0, basis='def2-svp', log=None,
                 label_tol=1e-9, init_guess=None):
    '''Rows (X, Y, mask) from densities of seeded clusters converged to
    label_tol (from init_guess, default the atomic guess: the converged
    density does not depend on it, only the number of SCF cycles).'''
    Xs, Ys, Ms = [], [], []
    t0 = time.perf_counter()
    out_dim = F.max_block(water_mol(seeds[0], shape, spacing, basis))
    for n, s in enumerate(seeds):
        mol = water_mol(s, shape, spacing, basis)
        dm, _ = converged_density(mol, label_tol, init_guess)
        rows, X = F.pairs_and_features(mol)
        Y, Mk = F.blocks_to_local(mol, dm, rows, out_dim)
        Xs.append(X); Ys.append(Y); Ms.append(Mk)
        if log:
            log(f'[dataset] {n+1}/{len(seeds)} geometries, {time.perf_counter()-t0:.1f} s')
    return (np.concatenate(Xs).astype(np.float32), np.concatenate(Ys).astype(np.float32),
            np.concatenate(Ms).astype(np.float32))


def train(data, epochs=50, lr=1e-3, batch=4096, seed=0, device='cuda', log=None,
          wd=1e-4):
    '''Train a PairMLP on (X, Y, mask); returns (net, loss history).  The
    training throughput is left in net.train_stats (rows/s and the GEMM
    FLOP rate of forward + backward, 6 flops per weight per row).'''
    torch = _torch()
    X, Y, Mk = data
    net = PairMLP(device, seed=seed, dims=(X.shape[1],) + HIDDEN + (Y.shape[1],))
    xb = to_bf16(X, device)
    yt = torch.as_tensor(Y, device=device)
    mt = torch.as_tensor(Mk, device=device)
    n = X.shape[0]
    # per-row mask counts on the host: the loss scale of a batch needs no
    # device round trip
    row_mask = Mk.sum(axis=1).astype(np.float64)
    total_mask = float(row_mask.sum())
    gen = torch.Generator(device='cpu').manual_seed(seed)
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm_cpu = torch.randperm(n, generator=gen)
        perm = perm_cpu.to(device)
        msum = np.add.reduceat(row_mask[perm_cpu.numpy()], np.arange(0, n, batch))
        tot = torch.zeros((), dtype=torch.float64, device=device)
        for nb, b0 in enumerate(range(0, n, batch)):
            idx = perm[b0:b0 + batch]
            xi, yi, mi = xb[idx], yt[idx], mt[idx]
            y = net.forward(xi)
            m_b = max(float(msum[nb]), 1.)
            loss, dy, dyb = net.loss_grad(y, yi, mi, 1. / m_b)
            net.backward(dy, dyb)
            net.adamw(lr=lr, wd=wd)
            tot += loss[0].double() * m_b
        hist.append(float(tot) / total_mask)
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log(f'[train] epoch {ep}: mse {hist[-1]:.3e}')
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    weights = sum(a * b for a, b in zip(net.dims[:-1], net.dims[1:]))
    net.train_stats = dict(seconds=dt, rows=n, epochs=epochs, rows_per_s=n * epochs / dt,
                           gemm_tflops=6. * weights * n * epochs / dt / 1e12)
    if log:
        log(f"[train] {n} rows x {epochs} epochs in {dt:.1f} s: "
            f"{net.train_stats['rows_per_s']:.3e} rows/s, "
            f"{net.train_stats['gemm_tflops']:.2f} TFLOP/s of forward + backward GEMMs")
    return net, hist


def predict_density(net, mol, device='cuda'):
    rows, X = F.pairs_and_features(mol)
    y = net.forward(to_bf16(X, device)).cpu().numpy().astype(np.float64)
    return F.local_to_density(mol, rows, y)


class LearnedGuess:
    '''Callable initial guess for RHF(init_guess=...); falls back to the
    atomic guess when the network cannot be applied.'''

    def __init__(self, net=None, device='cuda'):
        self.net = net
        self.device = device
        self.used_fallback = False

    def __call__(self, mol):
        from gpu4pyscf_amd.scf.atom_guess import atom_guess
        self.used_fallback = False
        if self.net is None or F.max_block(mol) > self.net.dims[-1]:
            # no network, or atom-pair blocks larger than its output
            self.used_fallback = True
            return atom_guess(mol)
        try:
            dm = predict_density(self.net, mol, self.device)
        except ValueError:
            self.used_fallback = True
            return atom_guess(mol)
        if not np.all(np.isfinite(dm)):
            self.used_fallback = True
            return atom_guess(mol)
        return dm

    def save(self, path):
        np.savez(path, w=self.net.state(), dims=np.asarray(self.net.dims))

    @classmethod
    def load(cls, path, device='cuda'):
        d = np.load(path)
        net = PairMLP(device, dims=tuple(int(x) for x in d['dims']))
        net.load_state(d['w'])
        return cls(net, device)
