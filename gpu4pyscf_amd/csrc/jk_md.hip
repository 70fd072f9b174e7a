// McMurchie-Davidson J-engine: the Coulomb matrix from Hermite Gaussians.
//
// J_uv = sum_kl (uv|kl) D_kl for a basis of single-primitive shells (the
// decontracted sorted basis) in three steps:
//   1. jk_md_dm_to_hermite: for every screened shell pair kl the Hermite
//      density Dh_kl[tuv] = K_kl (-1)^(t+u+v) sum_{c in k, d in l} D~_cd
//      E^{cd}_{tuv}, with E the Hermite expansion coefficients of the
//      Gaussian product, K_kl the pair prefactor of jk_pair_data and
//      D~_cd = D_cd + D_dc for k != l (D_cd for k == l);
//   2. jk_md_build_j: Jh_ij[tuv] += sum_kl (a_ij + a_kl)^-1/2
//      sum_{t'u'v'} R_{t+t', u+u', v+v'}(alpha, P - Q) Dh_kl[t'u'v'] over the
//      kl pairs that survive the J-only predicate of the reference
//      (create_tasks.cu:213-225: the bound exp(q_ij + q_kl) times the larger
//      of max|D_ij| and max|D_kl| reaches the cutoff), R the Hermite Coulomb
//      integrals of alpha = a_ij a_kl / (a_ij + a_kl);
//   3. jk_md_hermite_to_j: J_uv = K_ij sum_tuv E^{uv}_{tuv} Jh_ij[tuv] for
//      u in i, v in j (and J_vu = J_uv).
// (McMurchie and Davidson, J. Comput. Phys. 26, 218 (1978); Helgaker,
// Jorgensen and Olsen, Molecular Electronic-Structure Theory, ch. 9.)
// The contraction of step 2 is not symmetrised over (ij) <-> (kl): every
// bra pair sums over every ket pair of every group, so that each work item
// adds only into its own Jh_ij and no atomics are issued per quartet.
//
// Work decomposition of step 2: one lane per ij pair (pairs sorted by
// descending q_cond, so that the lanes of a wavefront have similar bounds),
// blocks of MD_THREADS lanes over one range of the kl list; the kl records
// (Q, a_kl, q_kl, max|D_kl|, Dh_kl) are staged in LDS MD_KL_TILE at a time
// and read by every lane (broadcast).  The Hermite integrals R are built in
// registers by the downward recursion in the auxiliary index, the Boys
// values F_0..F_L from an 8-term Taylor series about the nearest point of a
// 0.1-spaced grid (gen_md_boys.py).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "../../include/jkhip.h"

void jk_set_error(const char *msg);

#define MD_THREADS      256
#define MD_KL_TILE      64
#define MD_LMAX         4           // LIJ + LKL of the templated kernels (5 and 6 spill)
#define MD_PAIR_LMAX    4           // l of a shell in the Hermite transforms

__host__ __device__ constexpr int md_nh(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int md_ncart(int l) { return (l + 1) * (l + 2) / 2; }

// position of (t, u, v) in the Hermite order: total n = t+u+v ascending,
// then t descending, then u descending
__host__ __device__ constexpr int md_hidx(int t, int u, int v)
{
    const int n = t + u + v;
    int p = n * (n + 1) * (n + 2) / 6;
    for (int tt = n; tt > t; --tt) p += n - tt + 1;
    return p + (n - t - u);
}

template <int L>
struct HermiteOrder {
    int t[md_nh(L)], u[md_nh(L)], v[md_nh(L)];
    constexpr HermiteOrder() : t(), u(), v() {
        int c = 0;
        for (int n = 0; n <= L; ++n)
            for (int a = n; a >= 0; --a)
                for (int b = n - a; b >= 0; --b) { t[c] = a; u[c] = b; v[c] = n - a - b; ++c; }
    }
};

template <int L>
struct CartOrder {
    int x[md_ncart(L)], y[md_ncart(L)], z[md_ncart(L)];
    constexpr CartOrder() : x(), y(), z() {
        int c = 0;
        for (int a = L; a >= 0; --a)
            for (int b = L - a; b >= 0; --b) { x[c] = a; y[c] = b; z[c] = L - a - b; ++c; }
    }
};

// Hermite expansion coefficients of one Cartesian direction:
// E[i][j][t] for i <= LI, j <= LJ, t <= i + j, of x_A^i x_B^j
// exp(-a_i x_A^2 - a_j x_B^2) = K sum_t E^{ij}_t Lambda_t (K outside):
// E^{i+1,j}_t = E^{ij}_{t-1} / 2p + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}, the
// same with X_PB for j.
template <int LI, int LJ>
__device__ __forceinline__ void md_ecoef(double (&E)[LI + 1][LJ + 1][LI + LJ + 1], double xpa,
                                         double xpb, double oo2p)
{
    constexpr int LT = LI + LJ;
#pragma unroll
    for (int i = 0; i <= LI; ++i)
#pragma unroll
        for (int j = 0; j <= LJ; ++j)
#pragma unroll
            for (int t = 0; t <= LT; ++t) E[i][j][t] = 0.;
    E[0][0][0] = 1.;
#pragma unroll
    for (int j = 0; j < LJ; ++j)
#pragma unroll
        for (int t = 0; t <= j + 1; ++t) {
            double e = xpb * E[0][j][t];
            if (t > 0) e += oo2p * E[0][j][t - 1];
            if (t + 1 <= j) e += (t + 1) * E[0][j][t + 1];
            E[0][j + 1][t] = e;
        }
#pragma unroll
    for (int i = 0; i < LI; ++i)
#pragma unroll
        for (int j = 0; j <= LJ; ++j)
#pragma unroll
            for (int t = 0; t <= i + j + 1; ++t) {
                double e = xpa * E[i][j][t];
                if (t > 0) e += oo2p * E[i][j][t - 1];
                if (t + 1 <= i + j) e += (t + 1) * E[i][j][t + 1];
                E[i + 1][j][t] = e;
            }
}

struct MDPairIn {
    const uint32_t *pairs;
    const double *pdata;     // jk_pair_data records, one primitive pair each
    int npairs;
    const int32_t *bas;
    const double *env;
    const int32_t *ao_loc;
    int nbas, nao;
};

// one thread per pair: Hermite density of the pair block of dm (step 1)
template <int LI, int LJ>
__global__ void md_dm_to_hermite_kernel(double *dh, const double *dm, MDPairIn a)
{
    constexpr int LT = LI + LJ, NH = md_nh(LT);
    constexpr int NFI = md_ncart(LI), NFJ = md_ncart(LJ);
    constexpr HermiteOrder<LT> ho{};
    constexpr CartOrder<LI> ci{};
    constexpr CartOrder<LJ> cj{};
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    const uint32_t bij = a.pairs[p];
    const int ish = bij / a.nbas, jsh = bij - (bij / a.nbas) * a.nbas;
    const double *rec = a.pdata + (size_t)p * 6;
    const double *ri = a.env + a.bas[ish * 8 + 7], *rj = a.env + a.bas[jsh * 8 + 7];
    const double oo2p = .5 * rec[1];
    double Ex[LI + 1][LJ + 1][LT + 1], Ey[LI + 1][LJ + 1][LT + 1], Ez[LI + 1][LJ + 1][LT + 1];
    md_ecoef<LI, LJ>(Ex, rec[2] - ri[0], rec[2] - rj[0], oo2p);
    md_ecoef<LI, LJ>(Ey, rec[3] - ri[1], rec[3] - rj[1], oo2p);
    md_ecoef<LI, LJ>(Ez, rec[4] - ri[2], rec[4] - rj[2], oo2p);
    const int i0 = a.ao_loc[ish], j0 = a.ao_loc[jsh];
    const size_t nao = a.nao;
    double out[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) out[h] = 0.;
#pragma unroll
    for (int m = 0; m < NFI; ++m)
#pragma unroll
    for (int n = 0; n < NFJ; ++n) {
        double d = dm[(i0 + m) * nao + j0 + n];
        if (ish != jsh) d += dm[(j0 + n) * nao + i0 + m];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int t = ho.t[h], u = ho.u[h], v = ho.v[h];
            if (t > ci.x[m] + cj.x[n] || u > ci.y[m] + cj.y[n] || v > ci.z[m] + cj.z[n])
                continue;
            out[h] = fma(d, Ex[ci.x[m]][cj.x[n]][t] * Ey[ci.y[m]][cj.y[n]][u] *
                            Ez[ci.z[m]][cj.z[n]][v], out[h]);
        }
    }
    const double K = rec[5];
    double *o = dh + (size_t)p * NH;
#pragma unroll
    for (int h = 0; h < NH; ++h)
        o[h] = ((ho.t[h] + ho.u[h] + ho.v[h]) & 1 ? -K : K) * out[h];
}

// one thread per pair: J of the pair block from its Hermite J (step 3)
template <int LI, int LJ>
__global__ void md_hermite_to_j_kernel(double *vj, const double *jh, MDPairIn a)
{
    constexpr int LT = LI + LJ, NH = md_nh(LT);
    constexpr int NFI = md_ncart(LI), NFJ = md_ncart(LJ);
    constexpr HermiteOrder<LT> ho{};
    constexpr CartOrder<LI> ci{};
    constexpr CartOrder<LJ> cj{};
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    const uint32_t bij = a.pairs[p];
    const int ish = bij / a.nbas, jsh = bij - (bij / a.nbas) * a.nbas;
    const double *rec = a.pdata + (size_t)p * 6;
    const double *ri = a.env + a.bas[ish * 8 + 7], *rj = a.env + a.bas[jsh * 8 + 7];
    const double oo2p = .5 * rec[1];
    double Ex[LI + 1][LJ + 1][LT + 1], Ey[LI + 1][LJ + 1][LT + 1], Ez[LI + 1][LJ + 1][LT + 1];
    md_ecoef<LI, LJ>(Ex, rec[2] - ri[0], rec[2] - rj[0], oo2p);
    md_ecoef<LI, LJ>(Ey, rec[3] - ri[1], rec[3] - rj[1], oo2p);
    md_ecoef<LI, LJ>(Ez, rec[4] - ri[2], rec[4] - rj[2], oo2p);
    double h[NH];
    const double K = rec[5];
#pragma unroll
    for (int n = 0; n < NH; ++n) h[n] = K * jh[(size_t)p * NH + n];
    const int i0 = a.ao_loc[ish], j0 = a.ao_loc[jsh];
    const size_t nao = a.nao;
#pragma unroll
    for (int m = 0; m < NFI; ++m)
#pragma unroll
    for (int n = 0; n < NFJ; ++n) {
        double s = 0.;
#pragma unroll
        for (int q = 0; q < NH; ++q) {
            const int t = ho.t[q], u = ho.u[q], v = ho.v[q];
            if (t > ci.x[m] + cj.x[n] || u > ci.y[m] + cj.y[n] || v > ci.z[m] + cj.z[n])
                continue;
            s = fma(Ex[ci.x[m]][cj.x[n]][t] * Ey[ci.y[m]][cj.y[n]][u] * Ez[ci.z[m]][cj.z[n]][v],
                    h[q], s);
        }
        vj[(i0 + m) * nao + j0 + n] += s;
        if (ish != jsh) vj[(j0 + n) * nao + i0 + m] += s;
    }
}

// Boys values F_0..F_L(x): 8-term Taylor series about the nearest grid point
// below the grid's end, F_0 = sqrt(pi/x)/2 and upward recursion above it.
template <int L>
__device__ __forceinline__ void md_boys(const MDBuildArgs &a, double x, double *f)
{
    if (x >= (a.boys_rows - 1) * a.boys_step) {
        const double ix = 1. / x;
        f[0] = .88622692545275801365 * sqrt(ix);         // sqrt(pi)/2
#pragma unroll
        for (int m = 1; m <= L; ++m) f[m] = f[m - 1] * ((m - .5) * ix);
        return;
    }
    constexpr double INV_FACT[8] = {1., 1., 1. / 2, 1. / 6, 1. / 24, 1. / 120, 1. / 720,
                                    1. / 5040};
    const int g = (int)(x * a.boys_inv_step + .5);
    const double d = g * a.boys_step - x;
    const double *F = a.boys + (size_t)g * a.boys_m;
    double Fg[L + 8];
#pragma unroll
    for (int m = 0; m < L + 8; ++m) Fg[m] = F[m];
#pragma unroll
    for (int m = 0; m <= L; ++m) {
        double s = Fg[m + 7] * INV_FACT[7];
#pragma unroll
        for (int k = 6; k >= 0; --k) s = fma(s, d, Fg[m + k] * INV_FACT[k]);
        f[m] = s;
    }
}

// Hermite Coulomb integrals R_tuv (t+u+v <= L) from the auxiliary values
// r0[n] = R^(n)_000, by R^(n)_{t+1,u,v} = t R^(n+1)_{t-1,u,v} + X R^(n+1)_{t,u,v}
// (likewise u with Y and v with Z), n from L-1 down to 0.  In place: level n
// is written over level n+1 from the highest total order down, since an
// entry of order k reads only entries of orders k-1 and k-2.
template <int L>
__device__ __forceinline__ void md_rtable(double (&R)[md_nh(L)], const double *r0, double X,
                                          double Y, double Z)
{
    constexpr HermiteOrder<L> ho{};
    R[0] = r0[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
#pragma unroll
        for (int h = md_nh(L - n) - 1; h >= 1; --h) {
            const int t = ho.t[h], u = ho.u[h], v = ho.v[h];
            double s;
            if (t > 0) {
                s = X * R[md_hidx(t - 1, u, v)];
                if (t > 1) s = fma((double)(t - 1), R[md_hidx(t - 2, u, v)], s);
            } else if (u > 0) {
                s = Y * R[md_hidx(t, u - 1, v)];
                if (u > 1) s = fma((double)(u - 1), R[md_hidx(t, u - 2, v)], s);
            } else {
                s = Z * R[md_hidx(t, u, v - 1)];
                if (v > 1) s = fma((double)(v - 1), R[md_hidx(t, u, v - 2)], s);
            }
            R[h] = s;
        }
        R[0] = r0[n];
    }
}

// kl record staged in LDS
template <int LKL>
struct MDKetTile {
    double Q[MD_KL_TILE][3];
    double akl[MD_KL_TILE];
    double dh[MD_KL_TILE][md_nh(LKL)];
    float q[MD_KL_TILE];
    float d[MD_KL_TILE];
    uint32_t bas[MD_KL_TILE];
};

// SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014; constants of
// Vigna's splitmix64.c): the term of a survivor set checksum
__device__ __forceinline__ unsigned long long md_quartet_mix(uint32_t bas_ij, uint32_t bas_kl)
{
    unsigned long long z = ((unsigned long long)bas_ij << 32) | bas_kl;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// step 2: lane = one ij pair, the block's kl range staged tile by tile
template <int LIJ, int LKL>
__global__ __launch_bounds__(MD_THREADS)
void md_j_kernel(MDBuildArgs a, int kl_chunk)
{
    constexpr int L = LIJ + LKL;
    constexpr int NHI = md_nh(LIJ), NHK = md_nh(LKL);
    constexpr HermiteOrder<LIJ> hi{};
    constexpr HermiteOrder<LKL> hk{};
    __shared__ MDKetTile<LKL> tile;
    __shared__ int live;
    const int tid = threadIdx.x;
    const int pij = blockIdx.x * MD_THREADS + tid;
    const bool has_ij = pij < a.npairs_ij;
    const int kl0 = blockIdx.y * kl_chunk;
    const int kl1 = min(kl0 + kl_chunk, a.npairs_kl);
    // the block's largest q_ij: the pairs are sorted by descending q_cond
    const float q_first = a.q_ij[blockIdx.x * MD_THREADS];
    double P[3] = {0., 0., 0.}, aij = 1.;
    float q_ij = -1e30f, d_ij = -1e30f;
    uint32_t bas_ij = 0;
    if (has_ij) {
        const double *r = a.pdata_ij + (size_t)pij * 6;
        aij = r[0]; P[0] = r[2]; P[1] = r[3]; P[2] = r[4];
        q_ij = a.q_ij[pij];
        d_ij = a.dmax_ij[pij];
        if (a.n_quartets || a.quartet_hash) bas_ij = a.pairs_ij[pij];
    }
    const float kl_cutoff = a.cutoff - q_ij;
    const bool counting = a.n_quartets || a.quartet_hash;
    double acc[NHI];
#pragma unroll
    for (int h = 0; h < NHI; ++h) acc[h] = 0.;
    unsigned long long nq = 0, hsum = 0;
    for (int t0 = kl0; t0 < kl1; t0 += MD_KL_TILE) {
        const int nt = min(MD_KL_TILE, kl1 - t0);
        __syncthreads();
        // kl sorted by descending q_cond: nothing after the first kl that no
        // ij pair of the block can reach survives (a margin of 1 covers the
        // float rounding of the lanes' own test)
        if (tid == 0) live = a.q_kl[t0] + a.dm_penalty + q_first + 1.f >= a.cutoff;
        for (int n = tid; n < nt; n += MD_THREADS) {
            const double *r = a.pdata_kl + (size_t)(t0 + n) * 6;
            tile.akl[n] = r[0];
            tile.Q[n][0] = r[2]; tile.Q[n][1] = r[3]; tile.Q[n][2] = r[4];
            tile.q[n] = a.q_kl[t0 + n];
            tile.d[n] = a.dmax_kl[t0 + n];
            if (counting) tile.bas[n] = a.pairs_kl[t0 + n];
        }
        for (int n = tid; n < nt * NHK; n += MD_THREADS)
            tile.dh[n / NHK][n % NHK] = a.dh_kl[(size_t)t0 * NHK + n];
        __syncthreads();
        if (!live) break;
        for (int k = 0; k < nt; ++k) {
            // the J-only predicate in the reference's float32 order
            const float q_kl = tile.q[k];
            if (!(q_kl + a.dm_penalty >= kl_cutoff)) continue;
            const float d_cutoff = kl_cutoff - q_kl;
            if (!(d_ij > d_cutoff || tile.d[k] > d_cutoff)) continue;
            if (counting && bas_ij >= tile.bas[k]) {
                ++nq;
                hsum += md_quartet_mix(bas_ij, tile.bas[k]);
            }
            const double akl = tile.akl[k];
            const double X = P[0] - tile.Q[k][0], Y = P[1] - tile.Q[k][1],
                         Z = P[2] - tile.Q[k][2];
            const double rs = 1. / (aij + akl);
            const double alpha = aij * akl * rs;
            const double rr = X * X + Y * Y + Z * Z;
            // R^(n)_000 = (-2 alpha)^n F_n(alpha |PQ|^2), times (a_ij + a_kl)^-1/2;
            // erf(omega r)/r: alpha -> alpha' = alpha w^2 / (alpha + w^2), the
            // prefactor times sqrt(alpha' / alpha); erfc = full - erf
            double r0[L + 1];
#pragma unroll
            for (int n = 0; n <= L; ++n) r0[n] = 0.;
            if (a.omega <= 0.) {
                double f[L + 1];
                md_boys<L>(a, alpha * rr, f);
                double c = sqrt(rs);
#pragma unroll
                for (int n = 0; n <= L; ++n) { r0[n] = c * f[n]; c *= -2. * alpha; }
            }
            if (a.omega != 0.) {
                const double w2 = a.omega * a.omega;
                const double a2 = alpha * w2 / (alpha + w2);
                double f[L + 1];
                md_boys<L>(a, a2 * rr, f);
                double c = sqrt(rs * a2 / alpha);
                if (a.omega < 0.) c = -c;
#pragma unroll
                for (int n = 0; n <= L; ++n) { r0[n] += c * f[n]; c *= -2. * a2; }
            }
            double R[md_nh(L)];
            md_rtable<L>(R, r0, X, Y, Z);
#pragma unroll
            for (int h = 0; h < NHI; ++h) {
                double s = acc[h];
#pragma unroll
                for (int g = 0; g < NHK; ++g)
                    s = fma(R[md_hidx(hi.t[h] + hk.t[g], hi.u[h] + hk.u[g], hi.v[h] + hk.v[g])],
                            tile.dh[k][g], s);
                acc[h] = s;
            }
        }
    }
    if (has_ij) {
        double *o = a.jh_ij + (size_t)pij * NHI;
#pragma unroll
        for (int h = 0; h < NHI; ++h)
            if (acc[h] != 0.) unsafeAtomicAdd(o + h, acc[h]);
    }
    if (a.n_quartets && nq) atomicAdd(a.n_quartets, nq);
    if (a.quartet_hash && hsum) atomicAdd(a.quartet_hash, hsum);
}

template <int LIJ, int LKL>
static hipError_t launch_md_j(const MDBuildArgs &a, hipStream_t stream)
{
    const int nblk_ij = (a.npairs_ij + MD_THREADS - 1) / MD_THREADS;
    // enough blocks to fill the chip: split the kl list when the ij pairs
    // give few blocks, but keep each range at least a few tiles long
    int nchunks = (4096 + nblk_ij - 1) / nblk_ij;
    const int maxchunks = (a.npairs_kl + 4 * MD_KL_TILE - 1) / (4 * MD_KL_TILE);
    if (nchunks > maxchunks) nchunks = maxchunks;
    if (nchunks < 1) nchunks = 1;
    const int kl_chunk = (a.npairs_kl + nchunks - 1) / nchunks;
    nchunks = (a.npairs_kl + kl_chunk - 1) / kl_chunk;
    if (nchunks > 65535) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((md_j_kernel<LIJ, LKL>), dim3(nblk_ij, nchunks), dim3(MD_THREADS), 0,
                       stream, a, kl_chunk);
    return hipGetLastError();
}

template <int LIJ>
static hipError_t dispatch_md_j_kl(const MDBuildArgs &a, hipStream_t stream)
{
    switch (a.lkl) {
    case 0: return launch_md_j<LIJ, 0>(a, stream);
    case 1: if constexpr (LIJ + 1 <= MD_LMAX) return launch_md_j<LIJ, 1>(a, stream); break;
    case 2: if constexpr (LIJ + 2 <= MD_LMAX) return launch_md_j<LIJ, 2>(a, stream); break;
    case 3: if constexpr (LIJ + 3 <= MD_LMAX) return launch_md_j<LIJ, 3>(a, stream); break;
    case 4: if constexpr (LIJ + 4 <= MD_LMAX) return launch_md_j<LIJ, 4>(a, stream); break;
    case 5: if constexpr (LIJ + 5 <= MD_LMAX) return launch_md_j<LIJ, 5>(a, stream); break;
    case 6: if constexpr (LIJ + 6 <= MD_LMAX) return launch_md_j<LIJ, 6>(a, stream); break;
    }
    return hipErrorNotSupported;
}

static hipError_t dispatch_md_j(const MDBuildArgs &a, hipStream_t stream)
{
    switch (a.lij) {
    case 0: return dispatch_md_j_kl<0>(a, stream);
    case 1: return dispatch_md_j_kl<1>(a, stream);
    case 2: return dispatch_md_j_kl<2>(a, stream);
    case 3: return dispatch_md_j_kl<3>(a, stream);
    case 4: return dispatch_md_j_kl<4>(a, stream);
    case 5: return dispatch_md_j_kl<5>(a, stream);
    case 6: return dispatch_md_j_kl<6>(a, stream);
    }
    return hipErrorNotSupported;
}

// pair-transform dispatch over (li, lj), li >= lj, li <= MD_PAIR_LMAX,
// li + lj <= MD_LMAX
template <template <int, int> class K>
struct PairLaunch;

#define MD_PAIR_CASES(F)                                                           \
    F(0, 0) F(1, 0) F(1, 1) F(2, 0) F(2, 1) F(2, 2) F(3, 0) F(3, 1) F(3, 2) F(3, 3) \
    F(4, 0) F(4, 1) F(4, 2)

static hipError_t launch_dm_to_hermite(int li, int lj, double *dh, const double *dm,
                                       const MDPairIn &a, hipStream_t stream)
{
    const dim3 grid((a.npairs + 127) / 128), block(128);
#define MD_CASE(I, J)                                                                   \
    if (li == I && lj == J) {                                                           \
        hipLaunchKernelGGL((md_dm_to_hermite_kernel<I, J>), grid, block, 0, stream, dh, dm, a); \
        return hipGetLastError();                                                       \
    }
    MD_PAIR_CASES(MD_CASE)
#undef MD_CASE
    return hipErrorNotSupported;
}

static hipError_t launch_hermite_to_j(int li, int lj, double *vj, const double *jh,
                                      const MDPairIn &a, hipStream_t stream)
{
    const dim3 grid((a.npairs + 127) / 128), block(128);
#define MD_CASE(I, J)                                                                   \
    if (li == I && lj == J) {                                                           \
        hipLaunchKernelGGL((md_hermite_to_j_kernel<I, J>), grid, block, 0, stream, vj, jh, a); \
        return hipGetLastError();                                                       \
    }
    MD_PAIR_CASES(MD_CASE)
#undef MD_CASE
    return hipErrorNotSupported;
}

static thread_local char md_err[256];

static int md_fail(const char *fmt, int a0, int a1, const char *what)
{
    snprintf(md_err, sizeof(md_err), fmt, a0, a1, what);
    jk_set_error(md_err);
    return 4;
}

extern "C" {

int jk_md_dm_to_hermite(double *dh, const double *dm, int nao, const uint32_t *pairs,
                        int npairs, const double *pdata, int li, int lj, const int32_t *bas,
                        const double *env, const int32_t *ao_loc, int nbas, void *stream)
{
    if (npairs <= 0) return 0;
    if (!dh || !dm || !pairs || !pdata || !bas || !env || !ao_loc || nao <= 0 || nbas <= 0) {
        jk_set_error("jk_md_dm_to_hermite: null pointer or empty dimension");
        return 3;
    }
    if (li < lj || lj < 0 || li > MD_PAIR_LMAX || li + lj > MD_LMAX) {
        jk_set_error("jk_md_dm_to_hermite: unsupported pair class");
        return 2;
    }
    const MDPairIn a = {pairs, pdata, npairs, bas, env, ao_loc, nbas, nao};
    hipError_t err = launch_dm_to_hermite(li, lj, dh, dm, a, (hipStream_t)stream);
    if (err != hipSuccess)
        return md_fail("jk_md_dm_to_hermite (%d%d): %s", li, lj, hipGetErrorString(err));
    return 0;
}

int jk_md_hermite_to_j(double *vj, const double *jh, int nao, const uint32_t *pairs,
                       int npairs, const double *pdata, int li, int lj, const int32_t *bas,
                       const double *env, const int32_t *ao_loc, int nbas, void *stream)
{
    if (npairs <= 0) return 0;
    if (!vj || !jh || !pairs || !pdata || !bas || !env || !ao_loc || nao <= 0 || nbas <= 0) {
        jk_set_error("jk_md_hermite_to_j: null pointer or empty dimension");
        return 3;
    }
    if (li < lj || lj < 0 || li > MD_PAIR_LMAX || li + lj > MD_LMAX) {
        jk_set_error("jk_md_hermite_to_j: unsupported pair class");
        return 2;
    }
    const MDPairIn a = {pairs, pdata, npairs, bas, env, ao_loc, nbas, nao};
    hipError_t err = launch_hermite_to_j(li, lj, vj, jh, a, (hipStream_t)stream);
    if (err != hipSuccess)
        return md_fail("jk_md_hermite_to_j (%d%d): %s", li, lj, hipGetErrorString(err));
    return 0;
}

int jk_md_build_j(const MDBuildArgs *args, void *stream)
{
    if (!args) { jk_set_error("jk_md_build_j: null args"); return 1; }
    const MDBuildArgs &a = *args;
    if (a.abi_version != JK_ABI_VERSION) {
        jk_set_error("jk_md_build_j: ABI version mismatch");
        return 1;
    }
    if (a.lij < 0 || a.lkl < 0 || a.lij + a.lkl > MD_LMAX) {
        jk_set_error("jk_md_build_j: LIJ + LKL above the templated maximum");
        return 2;
    }
    if (a.npairs_ij <= 0 || a.npairs_kl <= 0) return 0;
    if (!a.pdata_ij || !a.pdata_kl || !a.q_ij || !a.q_kl || !a.dmax_ij || !a.dmax_kl ||
        !a.dh_kl || ((a.n_quartets || a.quartet_hash) && (!a.pairs_ij || !a.pairs_kl)) ||
        !a.jh_ij || !a.boys || a.boys_rows < 2 || a.boys_m < a.lij + a.lkl + 8 ||
        !(a.boys_step > 0.)) {
        jk_set_error("jk_md_build_j: null pointer, or a Boys grid too narrow for the class");
        return 3;
    }
    hipError_t err = dispatch_md_j(a, (hipStream_t)stream);
    if (err != hipSuccess)
        return md_fail("jk_md_build_j (L %d + %d): %s", a.lij, a.lkl, hipGetErrorString(err));
    return 0;
}

}  // extern "C"
