'''jk_dm_cond_noalloc (the entry point that allocates nothing: its block
maxima live in the library's static device scratch) against jk_dm_cond with
a caller-supplied work buffer: bit-identical dm_cond on two calls in a row,
the second with a smaller basis, for hermi 1 and 0 and two densities; and
the size limit rejected on the host.'''
import ctypes

import numpy as np
import pytest

from tests.conftest import has_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(lib, mol_atoms, basis, n_dm, hermi, seed):
    import torch
    from gpu4pyscf_amd.gto.mole import M
    from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
    mol = M(atom=mol_atoms, basis=basis)
    sm = SortedGTO(mol)
    nbas, nao = sm.nbas, int(sm.ao_loc[-1])
    rng = np.random.default_rng(seed)
    dm = rng.standard_normal((n_dm, nao, nao)) * np.exp(-rng.random((n_dm, nao, nao)) * 30)
    dm = torch.as_tensor(dm, device='cuda')
    ao_loc = torch.as_tensor(np.asarray(sm.ao_loc, dtype=np.int32), device='cuda')
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ref = torch.empty((nbas, nbas), dtype=torch.float32, device='cuda')
    work = torch.empty((nbas, nbas), dtype=torch.float64, device='cuda')
    assert lib.jk_dm_cond(_ptr(ref), _ptr(work), _ptr(dm), n_dm, nao, _ptr(ao_loc), nbas,
                          hermi, stream) == 0
    out = torch.full((nbas, nbas), 7., dtype=torch.float32, device='cuda')
    assert lib.jk_dm_cond_noalloc(_ptr(out), _ptr(dm), n_dm, nao, _ptr(ao_loc), nbas, hermi,
                                  stream) == 0, lib.jk_last_error()
    torch.cuda.synchronize()
    print(f'[noalloc] nbas={nbas} n_dm={n_dm} hermi={hermi}', flush=True)
    return ref.cpu().numpy(), out.cpu().numpy()


def test_dm_cond_noalloc_matches_dm_cond_twice():
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.lib import clib
    lib = clib.load()
    big = G.to_string(G.water_cluster((2, 2, 2), 3.0, seed=3))
    small = 'O 0 0 0.1174; H -0.757 0 -0.4696; H 0.757 0 -0.4696'
    for atoms, basis, n_dm, hermi, seed in ((big, 'def2-tzvp', 2, 1, 0),
                                            (small, 'def2-svp', 1, 0, 1)):
        ref, out = _both(lib, atoms, basis, n_dm, hermi, seed)
        assert np.array_equal(ref.view(np.uint32), out.view(np.uint32))


def test_dm_cond_noalloc_rejects_oversized_basis():
    import torch
    from gpu4pyscf_amd.lib import clib
    lib = clib.load()
    x = torch.zeros(4, dtype=torch.float64, device='cuda')
    st = lib.jk_dm_cond_noalloc(_ptr(x), _ptr(x), 1, 2, _ptr(x), 4097, 1, None)
    assert st == 3
    assert b'2048' in lib.jk_last_error()
