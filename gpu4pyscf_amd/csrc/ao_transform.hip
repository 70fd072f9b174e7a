// AO basis transforms between the caller's basis and the sorted Cartesian
// basis of the J/K kernels: out = C M C^T (density in) and out = C^T M C
// (J/K out) with the sparse recontraction / spherical-to-Cartesian matrix C
// given as compressed rows.  The reference does the same with separate bra
// and ket passes per angular shell (gpu4pyscf/lib/gvhf-rys/mole_helper.cu:
// 1734-1900, called from gpu4pyscf/gto/mole.py:667-850,906-938).
//
// One thread per output element (i, j): out[i, j] = sum over the nonzeros
// (p, c_p) of row i and (q, c_q) of row j of c_p c_q M[p, q].  A row of C
// holds at most a shell's worth of nonzeros (one original shell's spherical
// functions, or the Cartesian functions of the sorted shells split from one
// original shell), so each element costs a few to a few hundred FMAs and the
// pass is bound by the reads of M, which neighbouring threads share through
// the caches.  A wavefront covers 64 consecutive columns of one row, so its
// stores are coalesced.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "../../include/jkhip.h"

void jk_set_error(const char *msg);

#define AOT_COLS 64     // columns per workgroup (one wavefront)
#define AOT_ROWS 4      // rows per workgroup (one per wavefront)

__global__ __launch_bounds__(AOT_COLS * AOT_ROWS)
void ao_transform_kernel(double *out, const double *in, int n_out, int n_in,
                         const int32_t *ptr, const int32_t *idx, const double *coef)
{
    const int j = blockIdx.x * AOT_COLS + (threadIdx.x & (AOT_COLS - 1));
    const int i = blockIdx.y * AOT_ROWS + threadIdx.x / AOT_COLS;
    if (i >= n_out || j >= n_out) return;
    const double *src = in + (size_t)blockIdx.z * n_in * n_in;
    const int pi0 = ptr[i], pi1 = ptr[i + 1];
    const int pj0 = ptr[j], pj1 = ptr[j + 1];
    double s = 0.;
    for (int p = pi0; p < pi1; ++p) {
        const double *row = src + (size_t)idx[p] * n_in;
        double t = 0.;
        for (int q = pj0; q < pj1; ++q)
            t = fma(coef[q], row[idx[q]], t);
        s = fma(coef[p], t, s);
    }
    out[((size_t)blockIdx.z * n_out + i) * n_out + j] = s;
}

extern "C" int jk_ao_transform(double *out, const double *in, int n_batch, int n_out, int n_in,
                               const int32_t *ptr, const int32_t *idx, const double *coef,
                               void *stream)
{
    if (!out || !in || !ptr || !idx || !coef || n_batch <= 0 || n_out <= 0 || n_in <= 0) {
        jk_set_error("jk_ao_transform: bad arguments");
        return 3;
    }
    const long rows = ((long)n_out + AOT_ROWS - 1) / AOT_ROWS;
    if (rows > 65535 || n_batch > 65535) {
        jk_set_error("jk_ao_transform: n_out above 262140 or n_batch above 65535");
        return 3;
    }
    dim3 grid((n_out + AOT_COLS - 1) / AOT_COLS, (unsigned)rows, n_batch);
    hipLaunchKernelGGL(ao_transform_kernel, grid, dim3(AOT_COLS * AOT_ROWS), 0,
                       (hipStream_t)stream, out, in, n_out, n_in, ptr, idx, coef);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        char msg[256];
        snprintf(msg, sizeof(msg), "jk_ao_transform: %s", hipGetErrorString(err));
        jk_set_error(msg);
        return 4;
    }
    return 0;
}
