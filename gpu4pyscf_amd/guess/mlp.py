'''Pair MLP of the learned initial guess, trained and evaluated with the
hand-written BF16 MFMA kernels of libjkhip.so (mlp_* entry points).

Layers: 400 -> 256 -> 256 -> 256 -> 196, SiLU on the hidden layers, linear
output.  Weights and activations are bfloat16, accumulation and the master
weights / AdamW moments are FP32.  torch provides only device memory.
'''

import ctypes

import numpy as np

from gpu4pyscf_amd.lib import clib
from gpu4pyscf_amd.guess.features import FEAT_PAD, OUT_DIM

DIMS = (FEAT_PAD, 256, 256, 256, OUT_DIM)


def _torch():
    import torch
    return torch


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


class PairMLP:
    def __init__(self, device, seed=0, dims=DIMS):
        torch = _torch()
        self.dims = dims
        self.device = device
        rng = np.random.default_rng(seed)
        shapes = []
        for fi, fo in zip(dims[:-1], dims[1:]):
            shapes.append((fi, fo))
            shapes.append((fo,))
        self.shapes = shapes
        sizes = [int(np.prod(s)) for s in shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        n = int(self.offsets[-1])
        w = np.zeros(n, dtype=np.float32)
        for k, s in enumerate(shapes):
            o = self.offsets[k]
            if len(s) == 2:
                bound = np.sqrt(6. / s[0]) * (1. if k < 2 * (len(dims) - 2) else .5)
                w[o:o + sizes[k]] = rng.uniform(-bound, bound, sizes[k])
        self.w = torch.as_tensor(w, device=device)
        self.wb = torch.empty(n, dtype=torch.int16, device=device)
        self.m = torch.zeros(n, dtype=torch.float32, device=device)
        self.v = torch.zeros(n, dtype=torch.float32, device=device)
        self.g = torch.zeros(n, dtype=torch.float32, device=device)
        self.step_count = 0
        self._refresh_bf16()

    # -- parameter views ---------------------------------------------------
    def _view(self, buf, k):
        o = int(self.offsets[k])
        return buf[o:o + int(np.prod(self.shapes[k]))]

    def W(self, l, buf=None):
        return self._view(self.w if buf is None else buf, 2 * l)

    def b(self, l, buf=None):
        return self._view(self.w if buf is None else buf, 2 * l + 1)

    def _refresh_bf16(self):
        lib = clib.load()
        clib.check(lib.mlp_to_bf16(_p(self.wb), _p(self.w), self.w.numel(), _stream()),
                   'mlp_to_bf16')

    def state(self):
        return self.w.cpu().numpy().copy()

    def load_state(self, w):
        torch = _torch()
        self.w.copy_(torch.as_tensor(np.asarray(w, dtype=np.float32)))
        self._refresh_bf16()

    # -- forward / backward ------------------------------------------------
    def forward(self, xb):
        '''xb: (M, 400) bf16 (int16 tensor).  Returns Y (M, 196) FP32 and
        caches the pre-activations and activations for backward.'''
        torch = _torch()
        lib = clib.load()
        M = xb.shape[0]
        acts = [xb]
        zs = []
        st = _stream()
        nl = len(self.dims) - 1
        for l in range(nl):
            K, N = self.dims[l], self.dims[l + 1]
            Z = torch.empty((M, N), dtype=torch.float32, device=self.device)
            last = l == nl - 1
            A = None if last else torch.empty((M, N), dtype=torch.int16, device=self.device)
            clib.check(lib.mlp_gemm_bf16(_p(Z), _p(A), _p(acts[-1]), _p(self.W(l, self.wb)),
                                         _p(self.b(l)), M, N, K, K, 1, N, 1, N,
                                         0 if last else 1, st), 'mlp_gemm_bf16')
            zs.append(Z)
            if not last:
                acts.append(A)
        self._cache = (acts, zs)
        return zs[-1]

    def backward(self, dy, dyb):
        '''Gradients of the loss w.r.t. all parameters into self.g, given
        dL/dY in FP32 (dy) and bf16 (dyb).'''
        torch = _torch()
        lib = clib.load()
        acts, zs = self._cache
        st = _stream()
        self.g.zero_()
        nl = len(self.dims) - 1
        M = dy.shape[0]
        d32, d16 = dy, dyb
        for l in range(nl - 1, -1, -1):
            K, N = self.dims[l], self.dims[l + 1]
            if l == nl - 1:
                clib.check(lib.mlp_colsum(_p(self.b(l, self.g)), _p(d32), M, N, st), 'mlp_colsum')
            # dW = A^T dZ : (K x N)
            clib.check(lib.mlp_gemm_bf16(_p(self.W(l, self.g)), None, _p(acts[l]), _p(d16), None,
                                         K, N, M, 1, K, N, 1, N, 0, st), 'mlp_gemm_bf16')
            if l == 0:
                break
            # dA = dZ W^T : (M x K)
            dA = torch.empty((M, K), dtype=torch.float32, device=self.device)
            clib.check(lib.mlp_gemm_bf16(_p(dA), None, _p(d16), _p(self.W(l, self.wb)), None,
                                         M, K, N, N, 1, 1, N, K, 0, st), 'mlp_gemm_bf16')
            dz = torch.empty((M, K), dtype=torch.int16, device=self.device)
            clib.check(lib.mlp_silu_backward(_p(dz), _p(self.b(l - 1, self.g)), _p(dA),
                                             _p(zs[l - 1]), M, K, st), 'mlp_silu_backward')
            d16 = dz
            d32 = None

    def loss_grad(self, y, t, mask, scale):
        torch = _torch()
        lib = clib.load()
        dy = torch.empty_like(y)
        dyb = torch.empty(y.shape, dtype=torch.int16, device=self.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        clib.check(lib.mlp_mse_grad(_p(dy), _p(dyb), _p(loss), _p(y), _p(t), _p(mask),
                                    y.numel(), float(scale), _stream()), 'mlp_mse_grad')
        return loss, dy, dyb

    def adamw(self, lr=1e-3, beta1=.9, beta2=.999, eps=1e-8, wd=1e-4):
        lib = clib.load()
        self.step_count += 1
        clib.check(lib.mlp_adamw(_p(self.w), _p(self.wb), _p(self.m), _p(self.v), _p(self.g),
                                 self.w.numel(), lr, beta1, beta2, eps, wd, self.step_count,
                                 _stream()), 'mlp_adamw')


def to_bf16(x, device):
    '''float array -> bf16 bit patterns (int16 tensor) on the device.'''
    torch = _torch()
    lib = clib.load()
    xf = torch.as_tensor(np.asarray(x, dtype=np.float32), device=device).contiguous()
    out = torch.empty(xf.shape, dtype=torch.int16, device=device)
    clib.check(lib.mlp_to_bf16(_p(out), _p(xf), xf.numel(), _stream()), 'mlp_to_bf16')
    return out
