'''Boys-function grid of the McMurchie-Davidson J-engine (jk_md.hip).

boys[g, m] = F_m(g * STEP) for g < ROWS and m < M, evaluated in mpmath at 40
significant digits by the series and downward recursion of gen_rys_table.py.
The kernel takes F_m(x) for m <= L from the 8-term Taylor series about the
nearest grid point, F_m(x) = sum_k F_{m+k}(x_g) (x_g - x)^k / k!, which
needs m + 7 < M, i.e. L <= M - 8 = 12; above x = ROWS * STEP it uses
F_0 = sqrt(pi / x) / 2 and F_m = (2m - 1) / (2x) F_{m-1} (the neglected
exp(-x) terms are below 1e-13 relative there).

python gen_md_boys.py [out.npz]
'''
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from gen_rys_table import boys       # noqa: E402

STEP = 0.1
ROWS = 601                           # x = 0, 0.1, ..., 60
M = 20                               # F_0 .. F_19


def main(out):
    grid = np.array([[float(v) for v in boys(M - 1, g * STEP)] for g in range(ROWS)])
    np.savez(out, boys=grid, step=STEP, rows=ROWS, m=M)


if __name__ == '__main__':
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(
        os.path.dirname(os.path.abspath(__file__)), 'md_boys_table.npz')
    main(out)
