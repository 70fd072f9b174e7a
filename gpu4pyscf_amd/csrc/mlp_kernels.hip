// BF16 MFMA kernels of the learned initial-guess network: GEMM with fused
// bias + SiLU epilogue, SiLU backward, masked MSE gradient and AdamW.
// Weights and activations are bfloat16, every accumulation is FP32
// (v_mfma_f32_16x16x32_bf16, LDS-staged 64 x 64 workgroup tiles).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "../../include/jkhip.h"

typedef float f32x4 __attribute__((ext_vector_type(4)));

void jk_set_error(const char *msg);

__device__ __forceinline__ float bf2f(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u)
        return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }

// C(m, n) = sum_k A(m, k) B(k, n), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].
// mode 0: C = acc + bias;  mode 1: C = Z = acc + bias and act = bf16(silu(Z));
// mode 2: C += acc.
// Workgroup tile GM x GN (64 x 64) over 4 wavefronts, each a 32 x 32 quarter
// as 2 x 2 tiles of v_mfma_f32_16x16x32_bf16; K advances GK = 32 at a time.
// Both operand tiles are staged in LDS with k contiguous (A as [m][k], B
// transposed as [n][k]), loaded from HBM along whichever of the operand's
// two dimensions has unit stride, so that forward (A row major), dW = X^T dZ
// (A = X^T: unit stride along m) and dA = dZ W^T (B = W^T: unit stride
// along k) all read coalesced.  Fragment layout of the 16x16x32 bf16 MFMA:
// lane l supplies A(l%16, 8(l/16)+0..7) and B(8(l/16)+0..7, l%16) (16 bytes
// of one LDS row each) and receives D(4(l/16)+0..3, l%16).
#define GM 64
#define GN 64
#define GK 32
#define GK_PAD 40           // LDS row stride in bf16 (80 bytes: 16-byte aligned, spreads banks)

typedef short bf16x8 __attribute__((ext_vector_type(8)));

// one GR x GK operand tile into LDS t[r][k] (rows r of the tile = m of A or n of
// B); element (r, k) sits at P[r*sr + k*sk]; zero outside R x K
template <int GR>
__device__ __forceinline__ void stage_tile(uint16_t (*t)[GK_PAD], const uint16_t *P, long sr,
                                           long sk, int r0, int k0, int R, int K, int tid)
{
    constexpr int PER = GR * GK / 256;
    if (sk == 1) {
        // consecutive threads along k (rows of 32 elements)
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int n = tid + e * 256;
            const int r = n / GK, k = n % GK;
            const int gr = r0 + r, gk = k0 + k;
            t[r][k] = (gr < R && gk < K) ? P[gr * sr + gk] : (uint16_t)0;
        }
    } else {
        // consecutive threads along r
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int n = tid + e * 256;
            const int k = n / GR, r = n % GR;
            const int gr = r0 + r, gk = k0 + k;
            t[r][k] = (gr < R && gk < K) ? P[gr * sr + gk * sk] : (uint16_t)0;
        }
    }
}

__global__ __launch_bounds__(256)
void gemm_bf16_kernel(float *C, uint16_t *act, const uint16_t *A, const uint16_t *B,
                      const float *bias, int M, int N, int K, long sam, long sak,
                      long sbk, long sbn, int ldc, int mode)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[GM][GK_PAD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GN][GK_PAD];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int m0 = blockIdx.x * GM, n0 = blockIdx.y * GN;
    const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
    const int r = lane & 15, g = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += GK) {
        __syncthreads();
        stage_tile<GM>(As, A, sam, sak, m0, k0, M, K, tid);
        stage_tile<GN>(Bs, B, sbn, sbk, n0, k0, N, K, tid);
        __syncthreads();
        bf16x8 av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
            av[a] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * a + r][8 * g]);
#pragma unroll
        for (int b = 0; b < 2; ++b)
            bv[b] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * b + r][8 * g]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b],
                                                                    0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int m = m0 + wm + 16 * a + 4 * g + t, n = n0 + wn + 16 * b + r;
        if (m >= M || n >= N) continue;
        float v = acc[a][b][t];
        const long idx = (long)m * ldc + n;
        if (mode == 2) {
            C[idx] += v;
        } else {
            if (bias) v += bias[n];
            C[idx] = v;
            if (mode == 1) act[idx] = f2bf(silu(v));
        }
    }
}

// dZ = dA * silu'(Z) (bf16 out), db[n] += sum_m dZ(m, n)
__global__ void silu_backward_kernel(uint16_t *dz, float *db, const float *dA, const float *Z,
                                     int M, int N)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int mb = blockIdx.y * 256;
    const int me = min(mb + 256, M);
    float s = 0.f;
    for (int m = mb; m < me; ++m) {
        const long i = (long)m * N + n;
        const float z = Z[i];
        const float sg = 1.f / (1.f + __expf(-z));
        const float d = dA[i] * sg * (1.f + z * (1.f - sg));
        dz[i] = f2bf(d);
        s += d;
    }
    atomicAdd(db + n, s);
}

// column sums db[n] += sum_m G(m, n)
__global__ void colsum_kernel(float *db, const float *G, int M, int N)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int mb = blockIdx.y * 256;
    const int me = min(mb + 256, M);
    float s = 0.f;
    for (int m = mb; m < me; ++m) s += G[(long)m * N + n];
    atomicAdd(db + n, s);
}

// masked MSE: loss += scale sum mask (Y - T)^2 ; dY = 2 scale mask (Y - T)
__global__ void mse_grad_kernel(float *dy, uint16_t *dy_bf16, float *loss, const float *Y,
                                const float *T, const float *mask, long n, float scale)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float d = (Y[i] - T[i]) * mask[i];
        const float gv = 2.f * scale * d;
        dy[i] = gv;
        dy_bf16[i] = f2bf(gv);
        l = scale * d * d;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

__global__ void adamw_kernel(float *w, uint16_t *wb, float *m, float *v, const float *g, long n,
                             float lr, float b1, float b2, float eps, float wd, float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] * (1.f - lr * wd);
    wi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(uint16_t *out, const float *in, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

static int finish(const char *what)
{
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        char msg[256];
        snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(err));
        jk_set_error(msg);
        return 4;
    }
    return 0;
}

extern "C" {

int mlp_gemm_bf16(float *C, uint16_t *act, const uint16_t *A, const uint16_t *B,
                  const float *bias, int M, int N, int K, long sam, long sak, long sbk,
                  long sbn, int ldc, int mode, void *stream)
{
    if (!C || !A || !B || M <= 0 || N <= 0 || K <= 0 || ldc < N || mode < 0 || mode > 2 ||
        (mode == 1 && !act)) {
        jk_set_error("mlp_gemm_bf16: bad arguments");
        return 3;
    }
    dim3 grid((M + GM - 1) / GM, (N + GN - 1) / GN);
    if (grid.y > 65535u) { jk_set_error("mlp_gemm_bf16: N too large"); return 3; }
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, C, act, A, B,
                       bias, M, N, K, sam, sak, sbk, sbn, ldc, mode);
    return finish("mlp_gemm_bf16");
}

int mlp_silu_backward(uint16_t *dz, float *db, const float *dA, const float *Z, int M, int N,
                      void *stream)
{
    if (!dz || !db || !dA || !Z || M <= 0 || N <= 0) {
        jk_set_error("mlp_silu_backward: bad arguments");
        return 3;
    }
    if ((M + 255) / 256 > 65535) { jk_set_error("mlp_silu_backward: M too large"); return 3; }
    dim3 grid((N + 63) / 64, (M + 255) / 256);
    hipLaunchKernelGGL(silu_backward_kernel, grid, dim3(64), 0, (hipStream_t)stream,
                       dz, db, dA, Z, M, N);
    return finish("mlp_silu_backward");
}

int mlp_colsum(float *db, const float *G, int M, int N, void *stream)
{
    if (!db || !G || M <= 0 || N <= 0) { jk_set_error("mlp_colsum: bad arguments"); return 3; }
    if ((M + 255) / 256 > 65535) { jk_set_error("mlp_colsum: M too large"); return 3; }
    dim3 grid((N + 63) / 64, (M + 255) / 256);
    hipLaunchKernelGGL(colsum_kernel, grid, dim3(64), 0, (hipStream_t)stream, db, G, M, N);
    return finish("mlp_colsum");
}

int mlp_mse_grad(float *dy, uint16_t *dy_bf16, float *loss, const float *Y, const float *T,
                 const float *mask, long n, float scale, void *stream)
{
    if (!dy || !dy_bf16 || !loss || !Y || !T || !mask || n <= 0) {
        jk_set_error("mlp_mse_grad: bad arguments");
        return 3;
    }
    hipLaunchKernelGGL(mse_grad_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, dy, dy_bf16, loss, Y, T, mask, n, scale);
    return finish("mlp_mse_grad");
}

int mlp_adamw(float *w, uint16_t *wb, float *m, float *v, const float *g, long n, float lr,
              float beta1, float beta2, float eps, float wd, int step, void *stream)
{
    if (!w || !wb || !m || !v || !g || n <= 0 || step < 1) {
        jk_set_error("mlp_adamw: bad arguments");
        return 3;
    }
    float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adamw_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       w, wb, m, v, g, n, lr, beta1, beta2, eps, wd, c1, c2);
    return finish("mlp_adamw");
}

int mlp_to_bf16(uint16_t *out, const float *in, long n, void *stream)
{
    if (!out || !in || n <= 0) { jk_set_error("mlp_to_bf16: bad arguments"); return 3; }
    hipLaunchKernelGGL(to_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       out, in, n);
    return finish("mlp_to_bf16");
}

}  // extern "C"
