'''BF16 MFMA network kernels against the numpy oracle (oracle/mlp_oracle.py).'''
import numpy as np
import pytest

from tests.conftest import has_gpu
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def test_gemm_bf16_layout_and_transposes():
    import torch, ctypes
    from gpu4pyscf_amd.lib import clib
    from gpu4pyscf_amd.guess.mlp import to_bf16, _p, _stream
    lib = clib.load()
    rng = np.random.default_rng(0)
    M, N, K = 37, 45, 70
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    Ab, Bb = to_bf16(A, 'cuda'), to_bf16(B, 'cuda')
    ref = MO.bf16(A) @ MO.bf16(B)
    C = torch.empty((M, N), dtype=torch.float32, device='cuda')
    act = torch.empty((M, N), dtype=torch.int16, device='cuda')
    bt = torch.as_tensor(bias, device='cuda')
    assert lib.mlp_gemm_bf16(_p(C), _p(act), _p(Ab), _p(Bb), _p(bt), M, N, K, K, 1, N, 1, N, 1,
                             _stream()) == 0
    c = C.cpu().numpy()
    assert abs(c - (ref + bias)).max() < 1e-4 * abs(ref).max()
    a = act.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    a = a.view(np.float32)
    assert abs(a - MO.bf16(MO.silu(c))).max() <= 1e-2 * abs(a).max()
    # transposed operands: C2 = A^T (K x M) . C (M x N) with strided reads
    Cb = to_bf16(c, 'cuda')
    C2 = torch.zeros((K, N), dtype=torch.float32, device='cuda')
    assert lib.mlp_gemm_bf16(_p(C2), None, _p(Ab), _p(Cb), None, K, N, M, 1, K, N, 1, N, 0,
                             _stream()) == 0
    ref2 = MO.bf16(A).T @ MO.bf16(c)
    assert abs(C2.cpu().numpy() - ref2).max() < 1e-4 * abs(ref2).max()


def test_forward_backward_adamw_match_oracle():
    import torch
    from gpu4pyscf_amd.guess.mlp import PairMLP, to_bf16
    dims = (48, 32, 32, 20)
    net = PairMLP('cuda', seed=3, dims=dims)
    rng = np.random.default_rng(1)
    M = 53
    x = rng.standard_normal((M, dims[0])).astype(np.float32)
    t = rng.standard_normal((M, dims[-1])).astype(np.float32)
    mask = (rng.random((M, dims[-1])) > .2).astype(np.float32)
    w = net.state()
    Ws = [w[net.offsets[2*l]:net.offsets[2*l+1]].reshape(dims[l], dims[l+1]) for l in range(3)]
    bs = [w[net.offsets[2*l+1]:net.offsets[2*l+2]] for l in range(3)]
    y = net.forward(to_bf16(x, 'cuda'))
    yr, zs, acts = MO.forward(x, Ws, bs)
    assert abs(y.cpu().numpy() - yr).max() < 1e-3 * abs(yr).max()
    scale = 1. / mask.sum()
    tt = torch.as_tensor(t, device='cuda'); mm = torch.as_tensor(mask, device='cuda')
    loss, dy, dyb = net.loss_grad(y, tt, mm, scale)
    dyr = 2 * scale * (yr - t) * mask
    assert abs(float(loss) - scale * np.sum(((yr - t) * mask)**2)) < 1e-3 * float(loss)
    # backward through the kernels, oracle fed the kernels' own dL/dY
    net.backward(dy, dyb)
    dWr, dbr = MO.backward(dy.cpu().numpy().astype(np.float64), zs, acts, Ws)
    g = net.g.cpu().numpy()
    for l in range(3):
        gw = g[net.offsets[2*l]:net.offsets[2*l+1]].reshape(dims[l], dims[l+1])
        gb = g[net.offsets[2*l+1]:net.offsets[2*l+2]]
        assert abs(gw - dWr[l]).max() <= 2e-2 * abs(dWr[l]).max(), l
        assert abs(gb - dbr[l]).max() <= 2e-2 * abs(dbr[l]).max(), l
    # AdamW step
    w0 = net.state().astype(np.float64)
    net.adamw(lr=1e-3)
    w1, _, _ = MO.adamw(w0, 0., 0., g.astype(np.float64), 1, 1e-3)
    assert abs(net.state() - w1).max() < 1e-6
