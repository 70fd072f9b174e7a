'''ctypes binding of libjkhip.so (include/jkhip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C gpu4pyscf_amd/csrc``).  There is no CPU fallback: if the library
is missing, importing the J/K engine fails with a RuntimeError.
'''

import ctypes
import os

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_here, 'libjkhip.so')
TABLE_PATH = os.path.join(os.path.dirname(_here), 'csrc', 'rys_table.npz')

JK_ABI_VERSION = 13
JK_RYS_NMAX = 11
JK_BOYS_M = 13
JK_LMAX = 5
# JKBuildArgs.screen
JK_SCREEN = {'jk': 0, 'j': 1, 'k': 2}

c_void_p = ctypes.c_void_p


class JKRysTable(ctypes.Structure):
    _fields_ = [
        ('coef', c_void_p),
        ('coef32', c_void_p),
        ('off', ctypes.c_int64 * JK_RYS_NMAX),
        ('xmax', ctypes.c_double * JK_RYS_NMAX),
        ('large_r', ctypes.c_int64),
        ('large_w', ctypes.c_int64),
        ('degree', ctypes.c_int),
        ('boys_rows', ctypes.c_int),
        ('boys', ctypes.c_int64),
        ('boys_step', ctypes.c_double),
        ('boys_inv_step', ctypes.c_double),
    ]


class JKWindows(ctypes.Structure):
    '''JKWindows of include/jkhip.h: kl column windows of a pair list.'''
    _fields_ = [('pairs', c_void_p), ('q', c_void_p), ('pdata', c_void_p), ('win', c_void_p),
                ('n_win', ctypes.c_int), ('width', ctypes.c_int)]


class JKBuildArgs(ctypes.Structure):
    _fields_ = [
        ('abi_version', ctypes.c_int),
        ('nbas', ctypes.c_int),
        ('nao', ctypes.c_int),
        ('bas', c_void_p),
        ('env', c_void_p),
        ('ao_loc', c_void_p),
        ('li', ctypes.c_int), ('lj', ctypes.c_int), ('lk', ctypes.c_int), ('ll', ctypes.c_int),
        ('iprim', ctypes.c_int), ('jprim', ctypes.c_int),
        ('kprim', ctypes.c_int), ('lprim', ctypes.c_int),
        ('k_ao0', ctypes.c_int), ('k_nao', ctypes.c_int),
        ('l_ao0', ctypes.c_int), ('l_nao', ctypes.c_int),
        ('npairs_ij', ctypes.c_int), ('npairs_kl', ctypes.c_int),
        ('pairs_ij', c_void_p), ('pairs_kl', c_void_p),
        ('q_ij', c_void_p), ('q_kl', c_void_p),
        ('pdata_ij', c_void_p), ('pdata_kl', c_void_p),
        ('dm_cond', c_void_p),
        ('cutoff', ctypes.c_float),
        ('dm_penalty', ctypes.c_float),
        ('n_dm', ctypes.c_int),
        ('dm', c_void_p),
        ('vj', c_void_p),
        ('vk', c_void_p),
        ('rys', JKRysTable),
        ('n_quartets', c_void_p),
        ('fp32_cut', ctypes.c_float),
        ('n_quartets_fp32', c_void_p),
        ('quartet_hash', c_void_p),
        ('n_repl', ctypes.c_int),
        ('repl_stride', ctypes.c_int64),
        ('omega', ctypes.c_double),
        ('scale', ctypes.c_double),
        ('screen', ctypes.c_int),
        ('kl_win', JKWindows * 2),
        ('kl_win_policy', ctypes.c_int),
        ('swap_ijkl', ctypes.c_int),
    ]


class MDBuildArgs(ctypes.Structure):
    '''MDBuildArgs of include/jkhip.h (the J-engine's bra-ket task).'''
    _fields_ = [
        ('abi_version', ctypes.c_int),
        ('lij', ctypes.c_int), ('lkl', ctypes.c_int),
        ('npairs_ij', ctypes.c_int), ('npairs_kl', ctypes.c_int),
        ('pdata_ij', c_void_p), ('pdata_kl', c_void_p),
        ('q_ij', c_void_p), ('q_kl', c_void_p),
        ('dmax_ij', c_void_p), ('dmax_kl', c_void_p),
        ('cutoff', ctypes.c_float), ('dm_penalty', ctypes.c_float),
        ('dh_kl', c_void_p), ('jh_ij', c_void_p),
        ('boys', c_void_p),
        ('boys_rows', ctypes.c_int), ('boys_m', ctypes.c_int),
        ('boys_step', ctypes.c_double), ('boys_inv_step', ctypes.c_double),
        ('n_quartets', c_void_p),
        ('quartet_hash', c_void_p),
        ('pairs_ij', c_void_p), ('pairs_kl', c_void_p),
        ('omega', ctypes.c_double),
    ]


class JKTask(ctypes.Structure):
    '''JKTask of include/jkhip.h: the per-task fields of JKBuildArgs.'''
    _fields_ = [(n, ctypes.c_int) for n in ('li', 'lj', 'lk', 'll', 'iprim', 'jprim', 'kprim',
                                            'lprim', 'k_ao0', 'k_nao', 'l_ao0', 'l_nao',
                                            'npairs_ij', 'npairs_kl')] + [
        (n, c_void_p) for n in ('pairs_ij', 'pairs_kl', 'q_ij', 'q_kl', 'pdata_ij', 'pdata_kl')] + [
        ('kl_win', JKWindows * 2), ('swap_ijkl', ctypes.c_int)]


EXPORTS = ('jk_build', 'jk_build_tasks', 'jk_pair_data', 'jk_ovlp_mask', 'jk_pair_qcond', 'jk_dm_cond',
           'jk_dm_cond_noalloc', 'jk_energy_ip1', 'jk_md_dm_to_hermite', 'jk_md_hermite_to_j',
           'jk_md_build_j', 'jk_int1e', 'jk_ao_transform', 'mlp_gemm_bf16', 'mlp_silu_backward', 'mlp_colsum', 'mlp_mse_grad',
           'mlp_adamw', 'mlp_to_bf16', 'jk_last_error', 'jk_abi_version')

_lib = None


def load():
    '''Load libjkhip.so; raises RuntimeError when it has not been built.'''
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'{LIB_PATH} is not built; run __graft_entry__.build() '
                           'or make -C gpu4pyscf_amd/csrc')
    lib = ctypes.CDLL(LIB_PATH)
    lib.jk_last_error.restype = ctypes.c_char_p
    for name in ('jk_build', 'jk_build_tasks', 'jk_pair_data', 'jk_ovlp_mask', 'jk_pair_qcond', 'jk_dm_cond',
                 'jk_dm_cond_noalloc', 'jk_energy_ip1', 'jk_abi_version'):
        getattr(lib, name).restype = ctypes.c_int
    lib.jk_build.argtypes = [ctypes.POINTER(JKBuildArgs), c_void_p]
    lib.jk_md_build_j.restype = ctypes.c_int
    lib.jk_md_build_j.argtypes = [ctypes.POINTER(MDBuildArgs), c_void_p]
    for name in ('jk_md_dm_to_hermite', 'jk_md_hermite_to_j'):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = [c_void_p, c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, c_void_p,
                      ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p, ctypes.c_int,
                      c_void_p]
    lib.jk_energy_ip1.argtypes = [ctypes.POINTER(JKBuildArgs), c_void_p, ctypes.c_double,
                                  ctypes.c_double, c_void_p]
    lib.jk_pair_qcond.argtypes = [c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p,
                                  ctypes.c_int, JKRysTable, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, c_void_p]
    lib.jk_int1e.restype = ctypes.c_int
    lib.jk_int1e.argtypes = [c_void_p] * 6 + [ctypes.c_int, ctypes.c_int, c_void_p, c_void_p,
                                               ctypes.c_int, JKRysTable, c_void_p]
    L, I, F, P = ctypes.c_long, ctypes.c_int, ctypes.c_float, c_void_p
    lib.jk_ao_transform.restype = ctypes.c_int
    lib.jk_ao_transform.argtypes = [P, P, I, I, I, P, P, P, P]
    lib.mlp_gemm_bf16.argtypes = [P, P, P, P, P, I, I, I, L, L, L, L, I, I, P]
    lib.mlp_silu_backward.argtypes = [P, P, P, P, I, I, P]
    lib.mlp_colsum.argtypes = [P, P, I, I, P]
    lib.mlp_mse_grad.argtypes = [P, P, P, P, P, P, L, F, P]
    lib.mlp_adamw.argtypes = [P, P, P, P, P, L, F, F, F, F, F, I, P]
    lib.mlp_to_bf16.argtypes = [P, P, L, P]
    for name in ('mlp_gemm_bf16', 'mlp_silu_backward', 'mlp_colsum', 'mlp_mse_grad',
                 'mlp_adamw', 'mlp_to_bf16'):
        getattr(lib, name).restype = ctypes.c_int
    if lib.jk_abi_version() != JK_ABI_VERSION:
        raise RuntimeError('libjkhip.so ABI version mismatch; rebuild it')
    _lib = lib
    return lib


def check(status, what):
    if status != 0:
        msg = load().jk_last_error().decode()
        raise RuntimeError(f'{what} failed (status {status}): {msg}')


def load_rys_table_host():
    t = np.load(TABLE_PATH)
    coef = t['coef']
    lr = t['large_r']
    lw = t['large_w']
    boys = np.ascontiguousarray(t['boys'], dtype=np.float64)
    assert boys.shape[1] == JK_BOYS_M
    data = np.concatenate([coef, lr, lw, boys.ravel()]).astype(np.float64)
    meta = dict(off=t['off'].astype(np.int64), xmax=t['xmax'].astype(np.float64),
                large_r=len(coef), large_w=len(coef) + len(lr), degree=int(t['degree']),
                boys=len(coef) + len(lr) + len(lw), boys_rows=boys.shape[0],
                boys_step=float(t['boys_step']))
    return data, meta


_tables = {}


def rys_table(device):
    '''(device tensor, JKRysTable) of the Rys Chebyshev table on ``device``.'''
    import torch
    key = str(device)
    if key not in _tables:
        data, meta = load_rys_table_host()
        buf = torch.as_tensor(data, device=device)
        buf32 = buf.to(torch.float32)
        tab = JKRysTable()
        tab.coef = buf.data_ptr()
        tab.coef32 = buf32.data_ptr()
        n = len(meta['off'])
        for i in range(JK_RYS_NMAX):
            tab.off[i] = int(meta['off'][i]) if i < n else 0
            tab.xmax[i] = float(meta['xmax'][i]) if i < n else 0.
        tab.large_r = meta['large_r']
        tab.large_w = meta['large_w']
        tab.degree = meta['degree']
        tab.boys = meta['boys']
        tab.boys_rows = meta['boys_rows']
        tab.boys_step = meta['boys_step']
        tab.boys_inv_step = 1. / meta['boys_step']
        _tables[key] = ((buf, buf32), tab)
    return _tables[key]
