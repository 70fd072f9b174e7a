'''Writes water32_tzvp_incremental_jk.npz: the oracle's J and K of the
32-water def2-TZVP incremental build (4x4x2 lattice, 3.0 Angstrom, seed 0;
dD = 1e-3 (G + G^T)/2 with G ~ N(0,1) from default_rng(1)), hermi=1, in the
spherical AO basis: fingerprints of the whole matrices and ten full rows.

Run in the build container (C oracle with OpenMP; ~25 minutes on 8 cores):
    python tests/golden/make_water32_jk.py
'''
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpu4pyscf_amd.gto import geometry as G          # noqa: E402
from gpu4pyscf_amd.gto.mole import M                  # noqa: E402
from oracle import oracle as O                        # noqa: E402

ROWS = np.arange(0, 1376, 137)


def main():
    mol = M(atom=G.to_string(G.water_cluster((4, 4, 2), 3.0, seed=0)), basis='def2-tzvp')
    assert mol.nao == 1376
    rng = np.random.default_rng(1)
    g = rng.standard_normal((mol.nao, mol.nao))
    ddm = 1e-3 * (g + g.T) / 2
    t0 = time.perf_counter()
    vj, vk = O.get_jk(mol, ddm, hermi=1)
    print(f'oracle J/K in {time.perf_counter() - t0:.0f} s')
    np.savez_compressed(os.path.join(HERE, 'water32_tzvp_incremental_jk.npz'),
                        rows=ROWS, vj_rows=vj[ROWS], vk_rows=vk[ROWS],
                        fp_vj=O.fp(vj), fp_vk=O.fp(vk),
                        norm_vj=np.linalg.norm(vj), norm_vk=np.linalg.norm(vk))


if __name__ == '__main__':
    main()
