'''GPU checks of the sparse AO transform kernel (jk_ao_transform) behind
_VHFOpt.apply_coeff_C_mat_CT / apply_coeff_CT_mat_C against the dense
products C M C^T and C^T M C of the same coefficient matrix, in FP64 on the
device: spherical and Cartesian def2-TZVP (contracted shells split into
compact and diffuse parts), the reference's s..h test basis, and the
decontracted basis of the J-engine.'''
import numpy as np
import pytest

from tests.conftest import has_gpu
from tests.molecules import ATOM6, MOL1_ATOM, MOL1_BASIS
from gpu4pyscf_amd.gto.mole import M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    if not has_gpu():
        pytest.skip('no HIP device')


def _opts():
    from gpu4pyscf_amd.scf import jk, j_engine
    yield 'tzvp', jk._VHFOpt(M(atom=ATOM6, basis='def2-tzvp')).build()
    yield 'tzvp_cart', jk._VHFOpt(M(atom=ATOM6, basis='def2-tzvp', cart=True)).build()
    yield 'mol1', jk._VHFOpt(M(atom=MOL1_ATOM, basis=MOL1_BASIS)).build()
    yield 'tzvp_jengine', j_engine._VHFOpt(M(atom=ATOM6, basis='def2-tzvp')).build()


def test_sparse_transforms_match_dense_products():
    import torch
    rng = np.random.default_rng(5)
    for name, opt in _opts():
        C = opt._dev['coeff']
        nc, n = C.shape
        print(f'[case {name}] coeff {nc} x {n}, nnz {int((C != 0).sum())}', flush=True)
        assert nc == opt.sorted_mol.nao
        for lead in ((), (3,)):
            m = torch.as_tensor(rng.standard_normal(lead + (n, n)), device='cuda')
            got = opt.apply_coeff_C_mat_CT(m)
            want = C @ m @ C.T
            assert got.shape == lead + (nc, nc)
            assert float((got - want).abs().max()) <= 1e-13 * float(want.abs().max())
            v = torch.as_tensor(rng.standard_normal(lead + (nc, nc)), device='cuda')
            got = opt.apply_coeff_CT_mat_C(v)
            want = C.T @ v @ C
            assert got.shape == lead + (n, n)
            assert float((got - want).abs().max()) <= 1e-13 * float(want.abs().max())
        # a non-contiguous view goes through the same path
        m = torch.as_tensor(rng.standard_normal((n, n)), device='cuda').T
        assert float((opt.apply_coeff_C_mat_CT(m) - C @ m @ C.T).abs().max()) < 1e-12
        with pytest.raises(ValueError):
            opt.apply_coeff_C_mat_CT(torch.zeros((n + 1, n + 1), dtype=torch.float64,
                                                 device='cuda'))
        with pytest.raises(TypeError):
            opt.apply_coeff_C_mat_CT(torch.zeros((n, n), dtype=torch.float32, device='cuda'))
