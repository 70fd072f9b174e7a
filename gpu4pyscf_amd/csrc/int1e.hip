// One-electron integrals (overlap, kinetic, nuclear attraction) on the sorted
// Cartesian basis, for the SCF driver (gpu4pyscf/scf/hf.py:138-155, 880-883
// computes these with its own GPU integral code).
//
// One thread per shell pair (ish >= jsh); the block is written to both
// (ish, jsh) and (jsh, ish).  Overlap and kinetic use the Obara-Saika 1D
// recursions; the nuclear attraction uses Rys quadrature with
// (li + lj)/2 + 1 roots of x = p |P - C|^2 (the 2e recursions with an
// infinitely tight ket).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "../../include/jkhip.h"

#define L1MAX 5          // l <= 4
#define NC1MAX 15

extern "C" const char *jk_last_error(void);
void jk_set_error(const char *msg);

__device__ static void rys_roots_1e(const JKRysTable &tab, int N, double x, double *r, double *w)
{
    if (x >= tab.xmax[N-1]) {
        const double *lr = tab.coef + tab.large_r + N * (N - 1) / 2;
        const double *lw = tab.coef + tab.large_w + N * (N - 1) / 2;
        double ix = 1. / x, sx = sqrt(ix);
        for (int i = 0; i < N; ++i) { r[i] = lr[i] * ix; w[i] = lw[i] * sx; }
        return;
    }
    const int D1 = tab.degree + 1;
    int k = (int)x;
    double u = 2. * (x - k) - 1., u2 = 2. * u;
    const double *c = tab.coef + tab.off[N-1] + (int64_t)k * (2 * N) * D1;
    for (int f = 0; f < 2 * N; ++f) {
        const double *cf = c + f * D1;
        double b1 = 0., b2 = 0.;
        for (int j = D1 - 1; j > 0; --j) {
            double t = fma(u2, b1, cf[j] - b2);
            b2 = b1; b1 = t;
        }
        double v = fma(u, b1, cf[0] - b2);
        if (f & 1) w[f >> 1] = v; else r[f >> 1] = v;
    }
}

__device__ static void cart_pows(int l, int *px, int *py, int *pz)
{
    int c = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) { px[c] = x; py[c] = y; pz[c] = l - x - y; ++c; }
}

__global__ void int1e_kernel(double *S, double *T, double *V, const int32_t *bas,
                             const double *env, const int32_t *ao_loc, int nbas, int nao,
                             const double *coords, const double *charges, int natm,
                             JKRysTable rys)
{
    int jsh = blockIdx.x * blockDim.x + threadIdx.x;
    int ish = blockIdx.y;
    if (jsh >= nbas || jsh > ish) return;
    const int32_t *bi = bas + ish * 8, *bj = bas + jsh * 8;
    int li = bi[1], lj = bj[1];
    int nfi = (li + 1) * (li + 2) / 2, nfj = (lj + 1) * (lj + 2) / 2;
    const double *A = env + bi[7], *B = env + bj[7];
    double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    double rr = AB[0]*AB[0] + AB[1]*AB[1] + AB[2]*AB[2];
    int pix[NC1MAX], piy[NC1MAX], piz[NC1MAX], pjx[NC1MAX], pjy[NC1MAX], pjz[NC1MAX];
    cart_pows(li, pix, piy, piz);
    cart_pows(lj, pjx, pjy, pjz);
    double s[NC1MAX*NC1MAX], t[NC1MAX*NC1MAX], v[NC1MAX*NC1MAX];
    for (int n = 0; n < nfi * nfj; ++n) { s[n] = 0.; t[n] = 0.; v[n] = 0.; }
    const int nr = (li + lj) / 2 + 1;
    for (int ip = 0; ip < bi[2]; ++ip)
    for (int jp = 0; jp < bj[2]; ++jp) {
        double ai = env[bi[5] + ip], aj = env[bj[5] + jp];
        double p = ai + aj, rp = 1. / p;
        double cc = env[bi[6] + ip] * env[bj[6] + jp] * exp(-ai * aj * rp * rr);
        double P[3], PA[3], PB[3];
        for (int d = 0; d < 3; ++d) {
            P[d] = (ai * A[d] + aj * B[d]) * rp;
            PA[d] = P[d] - A[d];
            PB[d] = P[d] - B[d];
        }
        // 1D overlaps O[d][a][b], a <= li+1, b <= lj+1 (Obara-Saika), times sqrt(pi/p)
        double O[3][L1MAX+1][L1MAX+1];
        double s00 = sqrt(M_PI * rp);
        for (int d = 0; d < 3; ++d) {
            O[d][0][0] = s00;
            for (int a = 0; a <= li + 1; ++a) {
                if (a > 0) {
                    double val = PA[d] * O[d][a-1][0];
                    if (a > 1) val += (a - 1) * .5 * rp * O[d][a-2][0];
                    O[d][a][0] = val;
                }
                for (int b = 1; b <= lj + 1; ++b) {
                    double val = PB[d] * O[d][a][b-1];
                    if (a > 0) val += a * .5 * rp * O[d][a-1][b-1];
                    if (b > 1) val += (b - 1) * .5 * rp * O[d][a][b-2];
                    O[d][a][b] = val;
                }
            }
        }
        // 1D kinetic: K(a,b) = 1/2 a b O(a-1,b-1) + 2 ai aj O(a+1,b+1)
        //                      - ai b O(a+1,b-1) - aj a O(a-1,b+1)
        for (int i = 0; i < nfi; ++i)
        for (int j = 0; j < nfj; ++j) {
            int ax[3] = {pix[i], piy[i], piz[i]};
            int bx[3] = {pjx[j], pjy[j], pjz[j]};
            double o[3], k[3];
            for (int d = 0; d < 3; ++d) {
                int a = ax[d], b = bx[d];
                o[d] = O[d][a][b];
                double kv = 2. * ai * aj * O[d][a+1][b+1];
                if (a > 0 && b > 0) kv += .5 * a * b * O[d][a-1][b-1];
                if (b > 0) kv -= ai * b * O[d][a+1][b-1];
                if (a > 0) kv -= aj * a * O[d][a-1][b+1];
                k[d] = kv;
            }
            s[i*nfj+j] += cc * o[0] * o[1] * o[2];
            t[i*nfj+j] += cc * (k[0] * o[1] * o[2] + o[0] * k[1] * o[2] + o[0] * o[1] * k[2]);
        }
        // nuclear attraction, Rys quadrature
        for (int ia = 0; ia < natm; ++ia) {
            const double *C = coords + ia * 3;
            double PC[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
            double x = p * (PC[0]*PC[0] + PC[1]*PC[1] + PC[2]*PC[2]);
            double rt[JK_RYS_NMAX], wt[JK_RYS_NMAX];
            rys_roots_1e(rys, nr, x, rt, wt);
            double fac = -charges[ia] * 2. * M_PI * rp * cc;
            for (int r = 0; r < nr; ++r) {
                double u = rt[r];
                double b10 = .5 * rp * (1. - u);
                double G[3][2*L1MAX][L1MAX];
                for (int d = 0; d < 3; ++d) {
                    double c0 = PA[d] - u * PC[d];
                    double g[2*L1MAX];
                    g[0] = d == 2 ? wt[r] * fac : 1.;
                    if (li + lj > 0) g[1] = c0 * g[0];
                    for (int a = 1; a < li + lj; ++a) g[a+1] = c0 * g[a] + a * b10 * g[a-1];
                    // HRR: (a, b+1) = (a+1, b) + AB (a, b)
                    for (int a = 0; a <= li + lj; ++a) G[d][a][0] = g[a];
                    for (int b = 0; b < lj; ++b)
                        for (int a = 0; a < li + lj - b; ++a)
                            G[d][a][b+1] = G[d][a+1][b] + AB[d] * G[d][a][b];
                }
                for (int i = 0; i < nfi; ++i)
                for (int j = 0; j < nfj; ++j)
                    v[i*nfj+j] += G[0][pix[i]][pjx[j]] * G[1][piy[i]][pjy[j]] * G[2][piz[i]][pjz[j]];
            }
        }
    }
    int i0 = ao_loc[ish], j0 = ao_loc[jsh];
    for (int i = 0; i < nfi; ++i)
    for (int j = 0; j < nfj; ++j) {
        size_t ij = (size_t)(i0 + i) * nao + j0 + j, ji = (size_t)(j0 + j) * nao + i0 + i;
        S[ij] = s[i*nfj+j]; S[ji] = s[i*nfj+j];
        T[ij] = t[i*nfj+j]; T[ji] = t[i*nfj+j];
        V[ij] = v[i*nfj+j]; V[ji] = v[i*nfj+j];
    }
}

extern "C" int jk_int1e(double *S, double *T, double *V, const int32_t *bas, const double *env,
                        const int32_t *ao_loc, int nbas, int nao, const double *coords,
                        const double *charges, int natm, JKRysTable rys, void *stream)
{
    if (!S || !T || !V || !bas || !env || !ao_loc || !coords || !charges || nbas <= 0 ||
        nao <= 0 || natm <= 0 || !rys.coef) {
        jk_set_error("jk_int1e: bad arguments");
        return 3;
    }
    dim3 b((nbas + 63) / 64, nbas);
    hipLaunchKernelGGL(int1e_kernel, b, dim3(64), 0, (hipStream_t)stream, S, T, V, bas, env,
                       ao_loc, nbas, nao, coords, charges, natm, rys);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        char msg[256];
        snprintf(msg, sizeof(msg), "jk_int1e: %s", hipGetErrorString(err));
        jk_set_error(msg);
        return 4;
    }
    return 0;
}
