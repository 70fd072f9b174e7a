// Device code of the class kernels (per-lane and team kernels) and their
// host launchers, shared by the translation units that instantiate the
// angular classes (jk_classes.hip, compiled once per part).
// Direct-SCF J/K kernels for MI355X (gfx950), FP64 Rys quadrature.
//
// One launch evaluates one angular class (li,lj,lk,ll) of screened shell
// quartets and contracts it with the density matrices into J and K.
//
// Work decomposition: a work item is (one ij shell pair, one chunk of the kl
// pair list, or one kl column window of it).  The block scans the kl
// candidates of its chunk, applies the survival predicate of
// gpu4pyscf/lib/gvhf-rys/create_tasks.cu:114-180 and compacts the survivors
// into an LDS queue with wavefront ballots; queued quartets are then
// evaluated
//   * by the per-lane kernel (small classes, and some d/f classes in chunks of
//     block rows): one quartet per lane, the 2D Rys integrals and the
//     Cartesian block (or a chunk of its rows) in VGPRs;
//   * by the team kernel (large classes): a team of lanes per quartet, the
//     2D integrals of all roots staged in LDS, each lane owning one (i,j) row
//     of the block, the block staged through LDS for the contraction.
// A launch may run its task with bra and ket exchanged (swap_ijkl); the
// survival test and checksum then follow the task's own order.
// Digestion follows gpu4pyscf/lib/gvhf-rys/rys_contract_jk.cu:291-330 (four K
// blocks and two J blocks per quartet, symmetry factors of :156-158,214-220).
// On this 8-XCD part an FP64 atomic add into HBM is resolved on the memory
// side, so the kernels reduce in registers and LDS first: J_ij (common to a
// work item) across the wavefront or in the row-owning lane, the K rows of
// the work item's i and j AOs in LDS (per-lane kernel), the density blocks
// read by the digestion staged in LDS; what remains per quartet (J_kl, and
// K in the team kernel) goes to HBM with FP64 atomics.
// Rys roots: for up to three roots the Gauss rule of Boys moments taken from
// a Boys grid by a Taylor series; otherwise Chebyshev tables; both generated
// by gen_rys_table.py.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <type_traits>
#include "../../include/jkhip.h"

#define PI_FAC          34.98683665524972497    // 2 pi^(5/2)
#define Q_COND_MARGIN   4.f
#define WAVE            64
#define RYS_DEG_MAX     15
// the team kernel's one-root-per-lane evaluation uses the Boys-grid rule
// (all roots, one kept) up to this many roots, the Chebyshev tables above
#ifndef ROOT_ONE_BOYS_NMAX
#define ROOT_ONE_BOYS_NMAX 2
#endif
// classes whose Cartesian block plus three 2D tables fit in LANE_MAX
// doubles run one quartet per lane, the others on teams of lanes.
// Measured (ms per 32-water def2-TZVP build): 120 -> 720, 150 -> 679,
// 170 -> 660 (adds the (dd|ps)-sized classes), 240 -> 682; 64-water
// def2-TZVP build: 95 -> 3261, 120 -> 3271, 170 -> 3134.
#ifndef LANE_MAX
#define LANE_MAX        170
#endif
// classes whose block does not fit LANE_MAX with its 2D tables run per lane
// in chunks of rows i when listed in LaneChunked (below), or when they need
// at most LANE_MAX_CHUNKS chunks (1: the list only); the others on teams
#ifndef LANE_MAX_CHUNKS
#define LANE_MAX_CHUNKS 1
#endif
// doubles of one chunk (its rows of the block plus three 2D tables).  220
// with up to 3 chunks: (fp|pp) 69.9 -> 60.7 ms and (ff|ps) 55.5 -> 50.5 ms on
// the 64-water build, every class chunked before slower (larger chunks spill)
#ifndef LANE_CHUNK_MAX
#define LANE_CHUNK_MAX  LANE_MAX
#endif
// a block size the per-lane kernel leaves to the team kernel (0: none);
// sending the spilling (ff|ss)/(fs|fs) (100) or (dd|ps)-sized (108) classes
// back to the team kernel measured 658 and 670 ms against 652
#ifndef LANE_SKIP_NF
#define LANE_SKIP_NF    0
#endif
// wavefronts per block of the per-lane kernel (32-water def2-TZVP build on
// three streams, ms: 1 -> 546, 2 -> 464, 4 -> 489)
#ifndef LANE_WAVES
#define LANE_WAVES      2
#endif
// launch shapes: work items of the per-lane kernel target LANE_ITEMS blocks
// per launch with at least LANE_MIN_CHUNK kl candidates each; the team
// kernel scans TEAM_KL_CHUNK candidates per work item.  Measured on the
// 32-water def2-TZVP build (ms per build): team chunk 256/128/64/48/32 ->
// 1280/1143/1109/1065/1077 with 2048 lane items; lane items 4096 -> 2048
// at team chunk 256: 1280 -> 1236; without team K rows, lane items
// 2048 -> 1024: 883 -> 873.
#ifndef LANE_ITEMS
#define LANE_ITEMS      1024
#endif
#ifndef LANE_MIN_CHUNK
#define LANE_MIN_CHUNK  256
#endif
#ifndef TEAM_KL_CHUNK
#define TEAM_KL_CHUNK   48
#endif
// LDS K rows of the team kernel (0 none, 1 K_il/K_jl) and its minimum
// waves per SIMD (register budget).  With short work items zeroing and
// flushing the rows costs more than the atomics they save: 1065 -> 884 ms
// per build without them; a minimum of 2 or 3 waves per SIMD (fewer
// registers) measured 1105 and 1197 ms.
#ifndef TEAM_KROWS
#define TEAM_KROWS      0
#endif
// team kernel over kl column windows (kl_win[1]) with all four K blocks in
// LDS, for launches of at least TEAM_KWIN_MIN_ITEMS (ij pair, window) items.
// Off (used only under kl_win_policy 1): the rows cost more workgroups per
// CU than the atomics they save.  64-water def2-TZVP build, team kernels on
// one stream: 1080 ms without; with windows of 64 / 96 / 128 AOs 1159 /
// 1406 / 1512 ms (32-water: 280 without, 325 with 64-AO windows).
#ifndef TEAM_KWIN
#define TEAM_KWIN       0
#endif
// smallest workgroup of the team kernel (threads; 64, 128 or 256).  64-water
// def2-TZVP build, team kernels on one stream: 1082 ms with 64; 256: 1264 ms,
// 1199-1210 ms with K rows over 96- or 128-AO kl windows; 128 with 96-AO
// windows: 1176 ms.
#ifndef TEAM_TB_MIN
#define TEAM_TB_MIN     64
#endif
#ifndef TEAM_KWIN_MIN_ITEMS
#define TEAM_KWIN_MIN_ITEMS 2048
#endif
#ifndef TEAM_MIN_WAVES
#define TEAM_MIN_WAVES  1
#endif
// team kernel: evaluate several primitive quartets of a quartet per pass
// (roots, 2D tables and triple products of PB primitive quartets between
// three block barriers) instead of one.  Off: the larger LDS footprint cost
// more than the barriers it saves (64-water def2-TZVP build, team kernels on
// one stream: 1082 ms without, 1122 ms with)
#ifndef TEAM_PRIM_BATCH
#define TEAM_PRIM_BATCH 0
#endif
// LDS rows of the per-lane kernel: 2 = K_ik/K_jk and K_il/K_jl, 1 = K_il/K_jl
// only; LANE_DROWS also stages the density rows (883 ms per build; 1061
// with K_il/K_jl rows only, 887 without the density rows)
#ifndef LANE_KROWS
#define LANE_KROWS      2
#endif
#ifndef LANE_DROWS
#define LANE_DROWS      1
#endif
// rank the per-lane LDS row candidates by the workgroups per CU their LDS
// allows (0: take the largest rows that fit)
#ifndef LANE_KROWS_OCC
#define LANE_KROWS_OCC  1
#endif
// K_ik / K_jk LDS rows of the per-lane kernel: 1 = the lanes of a run of
// equal k shell first sum their values across the wavefront and the run's
// last lane adds once (no same-address LDS atomics within the run), 0 = every
// lane adds its own values.  Measured on the 64-water def2-TZVP build: 3130 ms
// with 0, 3229 ms with 1 (LDS bank conflicts are 0.05-1.5 % of the time of
// these kernels; the cross-lane sums cost more than the conflicts they avoid)
#ifndef LANE_KIK_SEGSUM
#define LANE_KIK_SEGSUM 0
#endif

void jk_set_error(const char *msg);

// ---------------------------------------------------------------- helpers
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

template <int L>
struct CartPow {
    int x[ncart(L)], y[ncart(L)], z[ncart(L)];
    constexpr CartPow() : x(), y(), z() {
        int c = 0;
        for (int lx = L; lx >= 0; --lx)
            for (int ly = L - lx; ly >= 0; --ly) {
                x[c] = lx; y[c] = ly; z[c] = L - lx - ly; ++c;
            }
    }
};

__device__ __forceinline__ void atomic_add(double *p, double v)
{
    unsafeAtomicAdd(p, v);
}

// Replicated accumulation (JKBuildArgs.n_repl): this workgroup's copy of
// vj/vk.
__device__ __forceinline__ void select_replica(JKBuildArgs &a)
{
    if (a.n_repl > 1) {
        const int64_t off = (int64_t)(blockIdx.x % a.n_repl) * a.repl_stride;
        if (a.vj) a.vj += off;
        if (a.vk) a.vk += off;
    }
}

// LDS accumulation: workgroup scope, so that it cannot be merged with the
// device-scope HBM atomics of a sibling branch into one generic atomic
__device__ __forceinline__ void lds_add(double *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, WAVE);
    return v;
}

// Sums M values over the 64 lanes of a wavefront and hands each total to
// one lane (reduce-scatter by recursive halving): at the step with lane bit
// b, lanes with bit b set keep the upper half of their values and add the
// partner's upper half, the others the lower half.  About M shuffles instead
// of 6 M for M separate butterfly sums.  v holds the M values of this lane
// and is overwritten; `put(n, x)` is called for every n < M, on the lane
// that holds the total x of value n.
// Value of the partner lane at the reduce-scatter step of lane bit BIT.
// Steps inside a 16-lane row use DPP moves (no LDS crossbar): lane ^ 1 and
// lane ^ 2 by quad permutes, lane ^ 8 by a row rotation of 8, and for
// BIT = 4 the half-row mirror, lane ^ 7: that partner also has the opposite
// bit 2 and the same higher bits, and its lanes summed so far are disjoint
// from this lane's, so the later ^ 2 and ^ 1 steps still cover all eight
// low-bit patterns.  The steps across rows (lane ^ 16, lane ^ 32) use the
// row-swap moves of gfx950: permlane16_swap(x, x) leaves row r ^ 1's values
// in its second result on even rows and in its first on odd rows,
// permlane32_swap(x, x) the other half's in its second result on lanes
// 0-31 and in its first on lanes 32-63.
// 0: LDS crossbar shuffles only; 1: DPP for the in-row steps; 2: also the
// row swaps (64-water build 2017 -> 1998 -> 1989 ms)
#ifndef WAVE_DPP
#define WAVE_DPP 2
#endif
template <int BIT>
__device__ __forceinline__ double partner_value(double x)
{
    constexpr int CTRL = BIT == 1 ? 0xB1 : BIT == 2 ? 0x4E : BIT == 4 ? 0x141 :
                         BIT == 8 ? 0x128 : -1;
    if constexpr (WAVE_DPP && CTRL >= 0) {
        const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    } else if constexpr (WAVE_DPP > 1 && (BIT == 16 || BIT == 32)) {
        const bool up = (__lane_id() & BIT) != 0;
        const int lo = __double2loint(x), hi = __double2hiint(x);
        if constexpr (BIT == 16) {
            const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
            const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
            return __hiloint2double(up ? (int)b[0] : (int)b[1], up ? (int)a[0] : (int)a[1]);
        } else {
            const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
            const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
            return __hiloint2double(up ? (int)b[0] : (int)b[1], up ? (int)a[0] : (int)a[1]);
        }
    } else {
        return __shfl_xor(x, BIT, WAVE);
    }
}

template <int M, int BIT, typename F>
__device__ __forceinline__ void reduce_scatter_step(double *v, int lane, int base, int nvalid,
                                                    F &put)
{
    if constexpr (BIT == 0) {
#pragma unroll
        for (int t = 0; t < M; ++t)
            if (t < nvalid) put(base + t, v[t]);
    } else {
        constexpr int H = (M + 1) / 2;
        const bool hi = (lane & BIT) != 0;
        double w[H];
#pragma unroll
        for (int t = 0; t < H; ++t) {
            const double lo_v = v[t];
            const double hi_v = (H + t < M) ? v[H + t] : 0.;
            const double mine = hi ? hi_v : lo_v;
            const double send = hi ? lo_v : hi_v;
            w[t] = mine + partner_value<BIT>(send);
        }
        const int nv = hi ? (nvalid > H ? nvalid - H : 0) : (nvalid < H ? nvalid : H);
        reduce_scatter_step<H, BIT / 2>(w, lane, hi ? base + H : base, nv, put);
    }
}

template <int M, typename F>
__device__ __forceinline__ void wave_reduce_scatter(double *v, int lane, F &&put)
{
    reduce_scatter_step<M, WAVE / 2>(v, lane, 0, M, put);
}

// Orders LDS accesses of the lanes of one wavefront (no block barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive segmented sum over lanes with equal, contiguous segment ids (the
// last lane of a segment receives the segment total), with the segment
// comparisons done once: same[s] says whether lane - 2^s belongs to this
// lane's segment.
struct SegMask { bool same[6]; };

__device__ __forceinline__ SegMask seg_mask(int seg, int lane)
{
    SegMask m;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int off = 1 << s;
        const int o = __shfl_up(seg, off, WAVE);
        m.same[s] = lane >= off && o == seg;
    }
    return m;
}

__device__ __forceinline__ double seg_sum(double v, const SegMask &m)
{
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const double u = __shfl_up(v, 1 << s, WAVE);
        if (m.same[s]) v += u;
    }
    return v;
}

// Boys moments F_0..F_{NM-1}(x), x < xmax[2], by the 8-term Taylor series
// about the nearest point x_g of the Boys grid:
// F_m(x) = sum_k F_{m+k}(x_g) (x_g - x)^k / k!  (|x_g - x| <= step/2).
template <int NM>
__device__ __forceinline__ void boys_moments(const JKRysTable &tab, double x, double *f)
{
    constexpr double INV_FACT[8] = {1., 1., 1. / 2, 1. / 6, 1. / 24, 1. / 120, 1. / 720,
                                    1. / 5040};
    int g = (int)(x * tab.boys_inv_step + .5);
    g = g < tab.boys_rows - 1 ? g : tab.boys_rows - 1;
    const double d = g * tab.boys_step - x;
    const double *F = tab.coef + tab.boys + (int64_t)g * JK_BOYS_M;
    double Fg[NM + 7];
#pragma unroll
    for (int m = 0; m < NM + 7; ++m) Fg[m] = F[m];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        double s = Fg[m + 7] * INV_FACT[7];
#pragma unroll
        for (int k = 6; k >= 0; --k) s = fma(s, d, Fg[m + k] * INV_FACT[k]);
        f[m] = s;
    }
}

// Rys nodes u_i = t_i^2 and weights: the large-x limit above xmax[N-1];
// below it, for one to three roots the Gauss rule of the Boys moments,
// else the Chebyshev tables.
template <int N>
__device__ __forceinline__ void rys_roots(const JKRysTable &tab, double x,
                                          double *r, double *w)
{
    if (x >= tab.xmax[N-1]) {
        const double *lr = tab.coef + tab.large_r + N * (N - 1) / 2;
        const double *lw = tab.coef + tab.large_w + N * (N - 1) / 2;
        double ix = 1. / x;
        double sx = sqrt(ix);
#pragma unroll
        for (int i = 0; i < N; ++i) { r[i] = lr[i] * ix; w[i] = lw[i] * sx; }
        return;
    }
    if constexpr (N == 1) {
        double f[2];
        boys_moments<2>(tab, x, f);
        w[0] = f[0];
        r[0] = f[1] / f[0];
        return;
    } else if constexpr (N == 2) {
        // monic p(u) = u^2 + c1 u + c0 orthogonal to 1 and u under the
        // moments: F2 + c1 F1 + c0 F0 = 0, F3 + c1 F2 + c0 F1 = 0
        double f[4];
        boys_moments<4>(tab, x, f);
        const double idet = 1. / (f[1] * f[1] - f[0] * f[2]);
        const double c1 = (f[0] * f[3] - f[1] * f[2]) * idet;
        const double c0 = (f[2] * f[2] - f[1] * f[3]) * idet;
        const double disc = sqrt(c1 * c1 - 4. * c0);
        const double q = -.5 * (c1 + copysign(disc, c1));
        const double ua = q, ub = c0 / q;
        const double u1 = fmin(ua, ub), u2 = fmax(ua, ub);
        const double w1 = (f[1] - u2 * f[0]) / (u1 - u2);
        r[0] = u1; r[1] = u2;
        w[0] = w1; w[1] = f[0] - w1;
        return;
    } else if constexpr (N == 3) {
        // recurrence coefficients of the monic orthogonal polynomials from
        // the moments (Chebyshev's algorithm, Gautschi 2004, Alg. 2.1), the
        // nodes from Viete's trigonometric solution of the characteristic
        // cubic of the Jacobi matrix refined by two Newton steps on the
        // three-term recurrence, the weights by the Christoffel formula
        double mu[6];
        boys_moments<6>(tab, x, mu);
        const double a0 = mu[1] / mu[0], b0 = mu[0];
        const double n1 = mu[2] - a0 * mu[1], n2 = mu[3] - a0 * mu[2];
        const double n3 = mu[4] - a0 * mu[3], n4 = mu[5] - a0 * mu[4];
        const double a1 = n2 / n1 - a0, b1 = n1 / mu[0];
        const double m2 = n3 - a1 * n2 - b1 * mu[2], m3 = n4 - a1 * n3 - b1 * mu[3];
        const double a2 = m3 / m2 - n2 / n1, b2 = m2 / n1;
        const double c2 = -(a0 + a1 + a2);
        const double c1 = a0 * a1 + a0 * a2 + a1 * a2 - b1 - b2;
        const double c0 = b2 * a0 + b1 * a2 - a0 * a1 * a2;
        const double p = c1 - c2 * c2 * (1. / 3.);
        const double q = c2 * c2 * c2 * (2. / 27.) - c2 * c1 * (1. / 3.) + c0;
        const double rr = sqrt(-p * (1. / 3.));
        const double phi = acos(fmin(1., fmax(-1., -q / (2. * rr * rr * rr))));
        double u[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            u[k] = 2. * rr * cos((phi - 2. * M_PI * k) * (1. / 3.)) - c2 * (1. / 3.);
        const double nrm1 = b0 * b1, nrm2 = nrm1 * b2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double v = u[k];
#pragma unroll
            for (int it = 0; it < 2; ++it) {
                const double p1 = v - a0;
                const double p2 = (v - a1) * p1 - b1, d2 = p1 + (v - a1);
                const double p3 = (v - a2) * p2 - b2 * p1, d3 = p2 + (v - a2) * d2 - b2;
                v -= p3 / d3;
            }
            const double p1 = v - a0, p2 = (v - a1) * p1 - b1;
            w[k] = 1. / (1. / b0 + p1 * p1 / nrm1 + p2 * p2 / nrm2);
            u[k] = v;
        }
        // ascending order (cos is decreasing on [0, pi]: k = 0 is the largest)
        r[0] = u[2]; r[1] = u[1]; r[2] = u[0];
        const double w0 = w[0];
        w[0] = w[2]; w[2] = w0;
        return;
    }
    const int D1 = tab.degree + 1;
    int k = (int)x;
    double u = 2. * (x - k) - 1.;
    double u2 = 2. * u;
    const double *c = tab.coef + tab.off[N-1] + (int64_t)k * (2 * N) * D1;
#pragma unroll
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

// Root `i` and its weight only (one lane per root).
template <int N>
__device__ __forceinline__ void rys_root_one(const JKRysTable &tab, double x, int i,
                                             double &r, double &w)
{
    if (x >= tab.xmax[N-1]) {
        double ix = 1. / x;
        r = tab.coef[tab.large_r + N * (N - 1) / 2 + i] * ix;
        w = tab.coef[tab.large_w + N * (N - 1) / 2 + i] * sqrt(ix);
        return;
    }
    if constexpr (N <= ROOT_ONE_BOYS_NMAX) {
        double rr[N], ww[N];
        rys_roots<N>(tab, x, rr, ww);
        r = rr[0]; w = ww[0];
#pragma unroll
        for (int k = 1; k < N; ++k)
            if (i == k) { r = rr[k]; w = ww[k]; }
        return;
    }
    const int D1 = tab.degree + 1;
    int k = (int)x;
    double u = 2. * (x - k) - 1.;
    double u2 = 2. * u;
    const double *c = tab.coef + tab.off[N-1] + ((int64_t)k * (2 * N) + 2 * i) * D1;
    double r1 = 0., r2 = 0., w1 = 0., w2 = 0.;
    for (int j = D1 - 1; j > 0; --j) {
        double t = fma(u2, r1, c[j] - r2);
        r2 = r1; r1 = t;
        double s = fma(u2, w1, c[D1 + j] - w2);
        w2 = w1; w1 = s;
    }
    r = fma(u, r1, c[0] - r2);
    w = fma(u, w1, c[D1] - w2);
}

// runtime-N variant (setup kernels)
static __device__ void rys_roots_rt(const JKRysTable &tab, int N, double x, double *r, double *w)
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

// 2D Rys integrals of one Cartesian direction for one root:
// g[((i*(LJ+1) + j)*(LK+1) + k)*(LL+1) + l], by VRR in i, TRR in k and HRR
// for j and l (the recursions of gpu4pyscf/lib/gvhf-rys/build_rys_gxyz.cuh:
// 18-123, evaluated in registers).
template <int LI, int LJ, int LK, int LL, typename T>
__device__ __forceinline__ void build_g(T *g, T g00, T c0, T cp, T b10, T b01, T b00,
                                        T rij, T rkl)
{
    constexpr int LIJ = LI + LJ, LKL = LK + LL;
    T G[LIJ+1][LKL+1];
    G[0][0] = g00;
    if (LIJ > 0) G[1][0] = c0 * g00;
#pragma unroll
    for (int a = 1; a < LIJ; ++a)
        G[a+1][0] = c0 * G[a][0] + ((T)a * b10) * G[a-1][0];
#pragma unroll
    for (int a = 0; a <= LIJ; ++a) {
        if (LKL > 0) {
            T v = cp * G[a][0];
            if (a > 0) v += ((T)a * b00) * G[a-1][0];
            G[a][1] = v;
        }
#pragma unroll
        for (int c = 1; c < LKL; ++c) {
            T v = cp * G[a][c] + ((T)c * b01) * G[a][c-1];
            if (a > 0) v += ((T)a * b00) * G[a-1][c];
            G[a][c+1] = v;
        }
    }
    // bra HRR: H[a][b][c] = H[a+1][b-1][c] - rij H[a][b-1][c]
    T H[LIJ+1][LJ+1][LKL+1];
#pragma unroll
    for (int c = 0; c <= LKL; ++c) {
#pragma unroll
        for (int a = 0; a <= LIJ; ++a) H[a][0][c] = G[a][c];
#pragma unroll
        for (int b = 0; b < LJ; ++b)
#pragma unroll
            for (int a = 0; a < LIJ - b; ++a)
                H[a][b+1][c] = H[a+1][b][c] - rij * H[a][b][c];
    }
    // ket HRR
#pragma unroll
    for (int a = 0; a <= LI; ++a)
#pragma unroll
    for (int b = 0; b <= LJ; ++b) {
        T W[LKL+1][LL+1];
#pragma unroll
        for (int c = 0; c <= LKL; ++c) W[c][0] = H[a][b][c];
#pragma unroll
        for (int e = 0; e < LL; ++e)
#pragma unroll
            for (int c = 0; c < LKL - e; ++c)
                W[c][e+1] = W[c+1][e] - rkl * W[c][e];
#pragma unroll
        for (int c = 0; c <= LK; ++c)
#pragma unroll
            for (int e = 0; e <= LL; ++e)
                g[((a*(LJ+1) + b)*(LK+1) + c)*(LL+1) + e] = W[c][e];
    }
}

// ---------------------------------------------------------------- shared
struct ShellPair {
    int ish, jsh, i0, j0;
    double ri[3], rij[3];      // Ri, Rj - Ri
    int pexpi, pexpj, pci, pcj;
};

__device__ __forceinline__ void load_shell_pair(ShellPair &s, uint32_t bas_ij, int nbas,
                                                const int32_t *bas, const double *env,
                                                const int32_t *ao_loc)
{
    s.ish = bas_ij / nbas;
    s.jsh = bas_ij - s.ish * nbas;
    const int32_t *bi = bas + s.ish * 8, *bj = bas + s.jsh * 8;
    const double *ri = env + bi[7], *rj = env + bj[7];
#pragma unroll
    for (int d = 0; d < 3; ++d) { s.ri[d] = ri[d]; s.rij[d] = rj[d] - ri[d]; }
    s.pexpi = bi[5]; s.pexpj = bj[5]; s.pci = bi[6]; s.pcj = bj[6];
    s.i0 = ao_loc[s.ish]; s.j0 = ao_loc[s.jsh];
}

// survival predicate, float32 in the reference's order: the bound test and
// pair order of create_tasks.cu:150-152, then the density blocks of the
// launch's screen (JKBuildArgs.screen): all six (:155-163), the Coulomb
// blocks D_ij, D_kl of a J-only build (:222-225) or the exchange blocks
// D_ik, D_jk, D_il, D_jl of a K-only build (:86-93)
// swapped (a launch with bra and ket exchanged, JKBuildArgs.swap_ijkl): the
// test in the task's own order and float32 rounding, with q_ij and cutoff
// the launch's bra bound and log(direct_scf_tol)
__device__ __forceinline__ bool survives(float q_kl, float kl_cutoff, float dm_penalty,
                                         uint32_t bas_ij, uint32_t bas_kl, int ish, int jsh,
                                         int nbas, const float *dm_cond, int screen,
                                         bool swapped = false, float q_ij = 0.f,
                                         float cutoff = 0.f)
{
    if (swapped) {
        kl_cutoff = cutoff - q_kl;
        q_kl = q_ij;
    }
    if (!(q_kl + dm_penalty >= kl_cutoff && (swapped ? bas_kl >= bas_ij : bas_ij >= bas_kl)))
        return false;
    int ksh = bas_kl / nbas;
    int lsh = bas_kl - ksh * nbas;
    float d_cutoff = kl_cutoff - q_kl;
    if (screen != JK_SCREEN_K &&
        (dm_cond[bas_ij] > d_cutoff || dm_cond[bas_kl] > d_cutoff))
        return true;
    // the exchange blocks D_ik, D_jk, D_il, D_jl of the task's shells
    const int ti = swapped ? ksh : ish, tj = swapped ? lsh : jsh;
    const int tk = swapped ? ish : ksh, tl = swapped ? jsh : lsh;
    return screen != JK_SCREEN_J &&
           (dm_cond[ti*nbas+tk] > d_cutoff ||
            dm_cond[tj*nbas+tk] > d_cutoff ||
            dm_cond[ti*nbas+tl] > d_cutoff ||
            dm_cond[tj*nbas+tl] > d_cutoff);
}

// SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014; constants of
// Vigna's splitmix64.c): the per-quartet term of the survivor set checksum
__device__ __forceinline__ unsigned long long quartet_mix(uint32_t bas_ij, uint32_t bas_kl);

// the checksum term of a quartet in the task's own order (bra and ket of a
// launch with swap_ijkl exchanged back)
__device__ __forceinline__ unsigned long long task_mix(const JKBuildArgs &a, uint32_t bas_ij,
                                                       uint32_t bas_kl)
{
    return a.swap_ijkl ? quartet_mix(bas_kl, bas_ij) : quartet_mix(bas_ij, bas_kl);
}

__device__ __forceinline__ unsigned long long quartet_mix(uint32_t bas_ij, uint32_t bas_kl)
{
    unsigned long long z = ((unsigned long long)bas_ij << 32) | bas_kl;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Same predicate, also returning the largest dm_cond value of the screen's
// blocks (any(v > d_cutoff) == max(v) > d_cutoff).
__device__ __forceinline__ bool survives_dmax(float q_kl, float kl_cutoff, float dm_penalty,
                                              uint32_t bas_ij, uint32_t bas_kl, int ish, int jsh,
                                              int nbas, const float *dm_cond, int screen,
                                              float &dmax, bool swapped = false,
                                              float q_ij = 0.f, float cutoff = 0.f)
{
    dmax = -1e30f;
    if (swapped) {
        kl_cutoff = cutoff - q_kl;
        q_kl = q_ij;
    }
    if (!(q_kl + dm_penalty >= kl_cutoff && (swapped ? bas_kl >= bas_ij : bas_ij >= bas_kl)))
        return false;
    int ksh = bas_kl / nbas;
    int lsh = bas_kl - ksh * nbas;
    float d_cutoff = kl_cutoff - q_kl;
    if (screen != JK_SCREEN_K)
        dmax = fmaxf(dm_cond[bas_ij], dm_cond[bas_kl]);
    // the exchange blocks D_ik, D_jk, D_il, D_jl of the task's shells
    const int ti = swapped ? ksh : ish, tj = swapped ? lsh : jsh;
    const int tk = swapped ? ish : ksh, tl = swapped ? jsh : lsh;
    if (screen != JK_SCREEN_J)
        dmax = fmaxf(dmax, fmaxf(fmaxf(dm_cond[ti*nbas+tk], dm_cond[tj*nbas+tk]),
                                 fmaxf(dm_cond[ti*nbas+tl], dm_cond[tj*nbas+tl])));
    return dmax > d_cutoff;
}

// Work item of a launch: pair_ij index and the kl range of its chunk.
// One work item: ij pair pij against kl pairs [kl0, kl1), whose k and l AOs
// lie in [k_ao0, k_ao0 + k_nao) and [l_ao0, l_ao0 + l_nao) (the K-row
// columns of the item).
struct WorkItem { int pij, kl0, kl1, k_ao0, k_nao, l_ao0, l_nao; };

// use_lds bits of the per-lane and team kernels
#define KR_MODE_MASK    3       // 0 none, 1 K_il/K_jl rows, 2 all four K row blocks
#define KR_DROWS        4       // density rows staged too
#define KR_WIN          8       // work items = ij pairs x kl column windows

// Workgroups are dispatched round-robin over the XCDs (each with its own
// L2): block b runs on XCD b % NXCD.  With XCD_SWIZZLE the blocks of one XCD
// take a contiguous range of work items (same ij pairs, neighbouring kl
// chunks), a bijection of [0, gridDim.x) for any grid size.  It is off: the
// ij pair lists are sorted by descending Schwarz bound, so contiguous ranges
// hand one XCD the heaviest work items (32-water def2-TZVP build: 1056 ms
// with it, 652 ms with the round-robin order).
#ifndef XCD_SWIZZLE
#define XCD_SWIZZLE 0
#endif
#define NXCD 8
__device__ __forceinline__ int work_index()
{
    const int b = blockIdx.x;
    if (!XCD_SWIZZLE) return b;
    const int n = gridDim.x;
    const int x = b % NXCD, per = n / NXCD, rem = n % NXCD;
    return x * per + (x < rem ? x : rem) + b / NXCD;
}

__device__ __forceinline__ WorkItem work_item(int npairs_kl, int kl_chunk)
{
    int nchunks = (npairs_kl + kl_chunk - 1) / kl_chunk;
    WorkItem w;
    const int idx = work_index();
    w.pij = idx / nchunks;
    int ch = idx - w.pij * nchunks;
    w.kl0 = ch * kl_chunk;
    w.kl1 = min(w.kl0 + kl_chunk, npairs_kl);
    w.k_ao0 = w.k_nao = w.l_ao0 = w.l_nao = 0;
    return w;
}

// work items over the kl column windows of kl_win[0] (window major, ij pair
// minor; the launcher puts the team kernels' set there); the K-row columns
// are the window's
__device__ __forceinline__ WorkItem work_item_win(const JKBuildArgs &a)
{
    WorkItem w;
    const int idx = work_index();
    const int win = idx / a.npairs_ij;
    w.pij = idx - win * a.npairs_ij;
    const int32_t *d = a.kl_win[0].win + 6 * win;
    w.kl0 = d[0]; w.kl1 = d[1];
    w.k_ao0 = d[2]; w.k_nao = d[3]; w.l_ao0 = d[4]; w.l_nao = d[5];
    return w;
}

// work item of a per-lane or team launch: kl chunks over the whole groups,
// or the kl column windows
__device__ __forceinline__ WorkItem work_item(const JKBuildArgs &a, int kl_chunk, int use_lds)
{
    if (use_lds & KR_WIN) return work_item_win(a);
    WorkItem w = work_item(a.npairs_kl, kl_chunk);
    w.k_ao0 = a.k_ao0; w.k_nao = a.k_nao; w.l_ao0 = a.l_ao0; w.l_nao = a.l_nao;
    return w;
}

// K rows of one work item accumulated in LDS (one density): rows i (NFI)
// then j (NFJ); columns: the work item's k AOs (when kcols) followed by its
// l AOs (the whole k and l groups, or one kl column window).  Flushed to K with one atomic per non-zero element.
// The same row layout also holds, after the K rows, a copy of the density
// rows i and j over the k and l group columns (dstaged), from which the
// per-lane digestion reads D_ik, D_il, D_jk and D_jl.
// p and d always point into the LDS array (so that every access compiles
// to an LDS instruction); `on` and `dstaged` say whether they are used.
struct KRows {
    double *p;          // K rows
    const double *d;    // density rows
    int width;          // doubles per row
    int loff;           // column of the first l AO
    int k_ao0, l_ao0;   // AOs of columns 0 and loff
    bool on;            // K_il / K_jl (and with kcols K_ik / K_jk) accumulated here
    bool kcols;         // K_ik / K_jk accumulated here too
    bool dstaged;       // density rows present
};

// use_lds: 0 none, 1 K_il/K_jl rows, 2 K_ik/K_jk and K_il/K_jl rows,
// 2|4 the same plus the density rows
__device__ __forceinline__ KRows make_krows(double *buf, int use_lds, int nrows,
                                            const WorkItem &w)
{
    KRows kr;
    kr.on = (use_lds & KR_MODE_MASK) != 0;
    kr.kcols = (use_lds & KR_MODE_MASK) == 2;
    kr.dstaged = kr.kcols && (use_lds & KR_DROWS);
    kr.loff = kr.kcols ? w.k_nao : 0;
    kr.width = kr.loff + w.l_nao;
    kr.k_ao0 = w.k_ao0;
    kr.l_ao0 = w.l_ao0;
    kr.p = buf;
    kr.d = buf + nrows * kr.width;
    return kr;
}

// copy the density rows (after krows_zero; a barrier must follow)
__device__ __forceinline__ void krows_load_d(const KRows &kr, int nfi, int nfj, int i0, int j0,
                                             const JKBuildArgs &a, int tid, int nthreads)
{
    if (!kr.dstaged) return;
    double *d = kr.p + (nfi + nfj) * kr.width;
    const int n_all = (nfi + nfj) * kr.width;
    for (int n = tid; n < n_all; n += nthreads) {
        int r = n / kr.width, c = n - r * kr.width;
        int row = r < nfi ? i0 + r : j0 + r - nfi;
        int col = c < kr.loff ? kr.k_ao0 + c : kr.l_ao0 + c - kr.loff;
        d[n] = a.dm[(size_t)row * a.nao + col];
    }
}

// zero (before the block's first quartet; a barrier must follow)
__device__ __forceinline__ void krows_zero(const KRows &kr, int nrows, int tid, int nthreads)
{
    if (!kr.on) return;
    for (int n = tid; n < nrows * kr.width; n += nthreads) kr.p[n] = 0.;
}

// flush (after a barrier that ends the block's accumulation)
__device__ __forceinline__ void krows_flush(const KRows &kr, int nfi, int nfj, int i0, int j0,
                                            const JKBuildArgs &a, int tid, int nthreads)
{
    if (!kr.on) return;
    const int n_all = (nfi + nfj) * kr.width;
    for (int n = tid; n < n_all; n += nthreads) {
        double v = kr.p[n];
        if (v != 0.) {
            int r = n / kr.width, c = n - r * kr.width;
            int row = r < nfi ? i0 + r : j0 + r - nfi;
            int col = c < kr.loff ? kr.k_ao0 + c : kr.l_ao0 + c - kr.loff;
            atomic_add(a.vk + (size_t)row * a.nao + col, v);
        }
    }
}

// ---------------------------------------------------------------- per-lane
// Primitive-pair record, see jk_pair_data(): (a, 1/a, Px, Py, Pz, K).
#define PREC 6

// Rys roots in FP32: Chebyshev series in float on the float copy of the
// table; the large-x asymptotic branch is evaluated in FP64.
template <int N>
__device__ __forceinline__ void rys_roots(const JKRysTable &tab, float x, float *r, float *w)
{
    if (x >= (float)tab.xmax[N-1]) {
        double xd = x;
        const double *lr = tab.coef + tab.large_r + N * (N - 1) / 2;
        const double *lw = tab.coef + tab.large_w + N * (N - 1) / 2;
        double ix = 1. / xd;
        double sx = sqrt(ix);
#pragma unroll
        for (int i = 0; i < N; ++i) { r[i] = (float)(lr[i] * ix); w[i] = (float)(lw[i] * sx); }
        return;
    }
    const int D1 = tab.degree + 1;
    int k = (int)x;
    float u = 2.f * (x - k) - 1.f;
    float u2 = 2.f * u;
    const float *c = tab.coef32 + tab.off[N-1] + (int64_t)k * (2 * N) * D1;
#pragma unroll
    for (int f = 0; f < 2 * N; ++f) {
        const float *cf = c + f * D1;
        float b1 = 0.f, b2 = 0.f;
        for (int j = D1 - 1; j > 0; --j) {
            float t = fmaf(u2, b1, cf[j] - b2);
            b2 = b1; b1 = t;
        }
        float v = fmaf(u, b1, cf[0] - b2);
        if (f & 1) w[f >> 1] = v; else r[f >> 1] = v;
    }
}

// One quartet per lane in arithmetic type T (double or float): returns the
// rows I0 <= i < I1 of the Cartesian block (times the symmetry and
// normalisation prefactors) in gout[(((i-I0)*nfj+j)*nfk+k)*nfl+l].
// pij: records of the (uniform) ij pair; pkl: records of this lane's kl pair.
template <int LI, int LJ, int LK, int LL, typename T, int I0 = 0, int I1 = ncart(LI)>
__device__ __forceinline__ void eri_lane(T *gout, const double *pij, int nij,
                                         const double *pkl, int nkl,
                                         const double *ri, const double *rij,
                                         const double *rk, const double *rkl,
                                         const JKRysTable &rys, double fac_sym, double omega2)
{
    constexpr int NFJ = ncart(LJ), NFK = ncart(LK), NFL = ncart(LL);
    constexpr int NF = (I1 - I0) * NFJ * NFK * NFL;
    constexpr int N = (LI + LJ + LK + LL) / 2 + 1;
    constexpr int GS = (LI+1)*(LJ+1)*(LK+1)*(LL+1);
    constexpr CartPow<LI> ci{};
    constexpr CartPow<LJ> cj{};
    constexpr CartPow<LK> ck{};
    constexpr CartPow<LL> cl{};
    const T rij_t[3] = {(T)rij[0], (T)rij[1], (T)rij[2]};
    const T rkl_t[3] = {(T)rkl[0], (T)rkl[1], (T)rkl[2]};
#pragma unroll
    for (int n = 0; n < NF; ++n) gout[n] = 0;
    for (int kp = 0; kp < nkl; ++kp) {
        const double *q = pkl + kp * PREC;
        const T akl = (T)q[0], rakl = (T)q[1];
        const T Kkl = (T)(q[5] * fac_sym);
        T QC[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) QC[d] = (T)(q[2+d] - rk[d]);
        for (int ip = 0; ip < nij; ++ip) {
            const double *p = pij + ip * PREC;
            const T aij = (T)p[0], raij = (T)p[1];
            T PA[3], PQ[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                PA[d] = (T)(p[2+d] - ri[d]);
                PQ[d] = (T)(p[2+d] - q[2+d]);
            }
            T rr = PQ[0]*PQ[0] + PQ[1]*PQ[1] + PQ[2]*PQ[2];
            T r_aa = (T)1 / (aij + akl);
            T theta = aij * akl * r_aa;
            T fac = (T)p[5] * Kkl * sqrt(r_aa);
            T rt[N], wt[N];
            // long-range attenuation erf(omega r)/r: the rule at x a with the
            // roots scaled by a and the weights by sqrt(a)
            T aw = (T)1;
            if (omega2 > 0.) aw = (T)omega2 / ((T)omega2 + theta);
            if constexpr (sizeof(T) == sizeof(float) && N <= 3) {
                // the Boys-grid rule in FP64 is cheaper than the float
                // Chebyshev series for up to three roots
                double rd[N], wd[N];
                rys_roots<N>(rys, (double)(theta * rr * aw), rd, wd);
#pragma unroll
                for (int i = 0; i < N; ++i) { rt[i] = (T)rd[i]; wt[i] = (T)wd[i]; }
            } else {
                rys_roots<N>(rys, theta * rr * aw, rt, wt);
            }
            if (omega2 > 0.) {
                const T sa = sqrt(aw);
#pragma unroll
                for (int i = 0; i < N; ++i) { rt[i] *= aw; wt[i] *= sa; }
            }
#pragma unroll
            for (int ir = 0; ir < N; ++ir) {
                T rt_aa = rt[ir] * r_aa;
                T rt_aij = rt_aa * akl, rt_akl = rt_aa * aij;
                T b10 = (T).5 * raij * ((T)1 - rt_aij);
                T b01 = (T).5 * rakl * ((T)1 - rt_akl);
                T b00 = (T).5 * rt_aa;
                T gx[GS], gy[GS], gz[GS];
                build_g<LI, LJ, LK, LL>(gx, (T)1, PA[0] - rt_aij * PQ[0], QC[0] + rt_akl * PQ[0],
                                        b10, b01, b00, rij_t[0], rkl_t[0]);
                build_g<LI, LJ, LK, LL>(gy, (T)1, PA[1] - rt_aij * PQ[1], QC[1] + rt_akl * PQ[1],
                                        b10, b01, b00, rij_t[1], rkl_t[1]);
                build_g<LI, LJ, LK, LL>(gz, wt[ir] * fac, PA[2] - rt_aij * PQ[2],
                                        QC[2] + rt_akl * PQ[2], b10, b01, b00, rij_t[2], rkl_t[2]);
#pragma unroll
                for (int i = I0; i < I1; ++i)
#pragma unroll
                for (int j = 0; j < NFJ; ++j)
#pragma unroll
                for (int k = 0; k < NFK; ++k)
#pragma unroll
                for (int l = 0; l < NFL; ++l) {
                    constexpr int SJ = LL+1, SK = (LK+1)*(LL+1), SI = (LJ+1)*(LK+1)*(LL+1);
                    int ix = ci.x[i]*SI + cj.x[j]*SK + ck.x[k]*SJ + cl.x[l];
                    int iy = ci.y[i]*SI + cj.y[j]*SK + ck.y[k]*SJ + cl.y[l];
                    int iz = ci.z[i]*SI + cj.z[j]*SK + ck.z[k]*SJ + cl.z[l];
                    int n = (((i - I0)*NFJ + j)*NFK + k)*NFL + l;
                    gout[n] = fma(gx[ix] * gy[iy], gz[iz], gout[n]);
                }
            }
        }
    }
}

// Rows of the Cartesian block a per-lane quartet holds in registers at a
// time: all NFI when the block and three 2D tables fit in LANE_MAX doubles,
// else chunks of CH rows i (each chunk re-evaluates the roots and 2D tables).
// Doubles of one chunk (its rows of the block plus three 2D tables) per
// class: LANE_CHUNK_MAX, or more where three larger chunks measured faster
// than five smaller ones (64-water build, ms on one stream: (fp|pp) 77.2 ->
// 60.7, (ff|ps) -> 50.5 against 55.5 on teams, (fp|fs) on 32 waters 10.6
// against 11.6 on teams).
template <int LI, int LJ, int LK, int LL> struct LaneChunkBudget {
    static constexpr int D = LANE_CHUNK_MAX;
};
#define LANE_CHUNK_BUDGET(a, b, c, d, n) \
    template <> struct LaneChunkBudget<a, b, c, d> { static constexpr int D = n; };
LANE_CHUNK_BUDGET(3, 1, 1, 1, 220) LANE_CHUNK_BUDGET(3, 3, 1, 0, 220)
LANE_CHUNK_BUDGET(3, 1, 3, 0, 220)
#undef LANE_CHUNK_BUDGET

template <int LI, int LJ, int LK, int LL>
struct LaneChunk {
    static constexpr int NFI = ncart(LI);
    static constexpr int ROW = ncart(LJ) * ncart(LK) * ncart(LL);
    static constexpr int GS3 = 3 * (LI+1)*(LJ+1)*(LK+1)*(LL+1);
    static constexpr int FIT = (LaneChunkBudget<LI, LJ, LK, LL>::D - GS3) / ROW;
    static constexpr int CH = NFI * ROW + GS3 <= LANE_MAX ? NFI : (FIT > 0 ? FIT : 1);
    static constexpr int NCH = (NFI + CH - 1) / CH;
};

// Classes above LANE_MAX that run per lane in chunks: those measured faster
// than on teams (64-water def2-TZVP build, ms on one stream, team ->
// chunked lanes: (fp|ds) 91.6 -> 54.4, (fd|ps) 85.0 -> 39.1, (dp|pp) 69.9 ->
// 52.5, (dd|ds) 37.6 -> 12.7, (fp|pp) 70.1 -> 60.7 and (ff|ps) 55.5 -> 50.5
// with the larger chunks of LaneChunkBudget; 32-water alike).  Slower per
// lane and left on teams: (fs|dp) 65.9 -> 73.6, (fs|fp) 40.3 -> 68.2,
// (ds|dd) 8.4 -> 34.9.
template <int LI, int LJ, int LK, int LL> struct LaneChunked { static constexpr bool ON = false; };
#define LANE_CHUNKED(a, b, c, d) \
    template <> struct LaneChunked<a, b, c, d> { static constexpr bool ON = true; };
LANE_CHUNKED(3, 1, 2, 0) LANE_CHUNKED(3, 2, 1, 0) LANE_CHUNKED(2, 1, 1, 1) LANE_CHUNKED(2, 2, 2, 0)
LANE_CHUNKED(3, 1, 1, 1) LANE_CHUNKED(3, 3, 1, 0) LANE_CHUNKED(3, 1, 3, 0)
#undef LANE_CHUNKED

// f(std::integral_constant<int, c>()) for c = 0 .. NC-1, unrolled at compile time
template <int C, int NC>
struct LaneChunkLoop {
    template <typename F>
    __device__ __forceinline__ static void run(F &f)
    {
        f(std::integral_constant<int, C>());
        LaneChunkLoop<C + 1, NC>::run(f);
    }
};
template <int NC>
struct LaneChunkLoop<NC, NC> {
    template <typename F>
    __device__ __forceinline__ static void run(F &) {}
};
template <int NC, typename F>
__device__ __forceinline__ void lane_chunks(F &&f)
{
    LaneChunkLoop<0, NC>::run(f);
}

// Evaluates and digests one batch of up to 64 queued quartets (one per lane)
// in arithmetic type T; digestion and accumulation are FP64.
template <int LI, int LJ, int LK, int LL, typename T>
__device__ __forceinline__ void lane_batch(const JKBuildArgs &a, const ShellPair &sp,
                                           uint32_t bas_ij, int pij, const uint32_t *queue,
                                           int nb, int lane, uint64_t lt_mask, const KRows &kr)
{
    constexpr int NFI = ncart(LI), NFJ = ncart(LJ), NFK = ncart(LK), NFL = ncart(LL);
    constexpr int CH = LaneChunk<LI, LJ, LK, LL>::CH;
    const int nbas = a.nbas, nao = a.nao;
    const size_t nao2 = (size_t)nao * nao;
    const bool active = lane < nb;
    const int pkl = queue[active ? lane : 0];
    const uint32_t bas_kl = a.pairs_kl[pkl];
    const int ksh = bas_kl / nbas, lsh = bas_kl - (bas_kl / nbas) * nbas;
    double fac_sym = a.scale;
    if (sp.ish == sp.jsh) fac_sym *= .5;
    if (ksh == lsh) fac_sym *= .5;
    if (bas_ij == bas_kl) fac_sym *= .5;
    if (!active) fac_sym = 0.;
    const double *rk = a.env + a.bas[ksh * 8 + 7], *rl = a.env + a.bas[lsh * 8 + 7];
    const double rk3[3] = {rk[0], rk[1], rk[2]};
    const double rkl[3] = {rl[0] - rk[0], rl[1] - rk[1], rl[2] - rk[2]};
    const int nij = a.iprim * a.jprim, nkl = a.kprim * a.lprim;
    const int k0 = a.ao_loc[ksh], l0 = a.ao_loc[lsh];
    const int i0 = sp.i0, j0 = sp.j0;
    // runs of equal k shell are contiguous in the queue (tile order):
    // segment = run id, from the run heads
    const int key = active ? ksh : -1;
    const int key_prev = __shfl_up(key, 1, WAVE);
    const bool head = lane == 0 || key_prev != key;
    const uint64_t heads = __ballot(head);
    const int run = __popcll(heads & (lt_mask | (1ull << lane)));
    const int run_next = __shfl_down(run, 1, WAVE);
    const bool tail = active && (lane == nb - 1 || run_next != run);
    const SegMask runs = seg_mask(run, lane);
    // columns of this quartet's k and l AOs in the LDS density rows
    const int dk = k0 - kr.k_ao0, dl = kr.loff + l0 - kr.l_ao0;
    // rows i of the block in chunks of CH (one chunk unless the block and
    // its 2D tables would not fit the registers): evaluate, then digest
    lane_chunks<(NFI + CH - 1) / CH>([&](auto cc) {
        constexpr int I0 = decltype(cc)::value * CH;
        constexpr int I1 = I0 + CH < NFI ? I0 + CH : NFI;
        constexpr int NI = I1 - I0;
        T gout[NI * NFJ * NFK * NFL];
        eri_lane<LI, LJ, LK, LL, T, I0, I1>(gout, a.pdata_ij + (size_t)pij * nij * PREC, nij,
                                            a.pdata_kl + (size_t)pkl * nkl * PREC, nkl,
                                            sp.ri, sp.rij, rk3, rkl, a.rys, fac_sym,
                                            a.omega * a.omega);
        for (int id = 0; id < a.n_dm; ++id) {
            const double *D = a.dm + id * nao2;
            if (a.vk) {
                double *K = a.vk + id * nao2;
                // K_ik += (ij|kl) D_jl ; K_il += (ij|kl) D_jk
#pragma unroll
                for (int i = I0; i < I1; ++i) {
                    double kil[NFL], kik[NFK];
#pragma unroll
                    for (int l = 0; l < NFL; ++l) kil[l] = 0.;
#pragma unroll
                    for (int k = 0; k < NFK; ++k) kik[k] = 0.;
                    if (active && kr.dstaged) {
#pragma unroll
                        for (int j = 0; j < NFJ; ++j)
#pragma unroll
                        for (int k = 0; k < NFK; ++k)
#pragma unroll
                        for (int l = 0; l < NFL; ++l) {
                            double g = gout[(((i - I0)*NFJ + j)*NFK + k)*NFL + l];
                            kil[l] = fma(g, kr.d[(NFI + j) * kr.width + dk + k], kil[l]);
                            kik[k] = fma(g, kr.d[(NFI + j) * kr.width + dl + l], kik[k]);
                        }
                    } else if (active) {
#pragma unroll
                        for (int j = 0; j < NFJ; ++j)
#pragma unroll
                        for (int k = 0; k < NFK; ++k)
#pragma unroll
                        for (int l = 0; l < NFL; ++l) {
                            double g = gout[(((i - I0)*NFJ + j)*NFK + k)*NFL + l];
                            kil[l] = fma(g, D[(j0+j)*nao + k0+k], kil[l]);
                            kik[k] = fma(g, D[(j0+j)*nao + l0+l], kik[k]);
                        }
                    }
                    if (kr.kcols && LANE_KIK_SEGSUM) {
#pragma unroll
                        for (int k = 0; k < NFK; ++k) {
                            double v = seg_sum(kik[k], runs);
                            if (tail) lds_add(kr.p + i * kr.width + k0 + k - kr.k_ao0, v);
                        }
                    } else if (kr.kcols) {
                        if (active) {
#pragma unroll
                            for (int k = 0; k < NFK; ++k)
                                lds_add(kr.p + i * kr.width + k0 + k - kr.k_ao0, kik[k]);
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < NFK; ++k) {
                            double v = seg_sum(kik[k], runs);
                            if (tail) atomic_add(K + (i0+i)*nao + k0+k, v);
                        }
                    }
                    if (active) {
#pragma unroll
                        for (int l = 0; l < NFL; ++l) {
                            if (kr.on) lds_add(kr.p + i * kr.width + kr.loff + l0 + l - kr.l_ao0, kil[l]);
                            else atomic_add(K + (i0+i)*nao + l0+l, kil[l]);
                        }
                    }
                }
                // K_jk += (ij|kl) D_il ; K_jl += (ij|kl) D_ik
#pragma unroll
                for (int j = 0; j < NFJ; ++j) {
                    double kjl[NFL], kjk[NFK];
#pragma unroll
                    for (int l = 0; l < NFL; ++l) kjl[l] = 0.;
#pragma unroll
                    for (int k = 0; k < NFK; ++k) kjk[k] = 0.;
                    if (active && kr.dstaged) {
#pragma unroll
                        for (int i = I0; i < I1; ++i)
#pragma unroll
                        for (int k = 0; k < NFK; ++k)
#pragma unroll
                        for (int l = 0; l < NFL; ++l) {
                            double g = gout[(((i - I0)*NFJ + j)*NFK + k)*NFL + l];
                            kjl[l] = fma(g, kr.d[i * kr.width + dk + k], kjl[l]);
                            kjk[k] = fma(g, kr.d[i * kr.width + dl + l], kjk[k]);
                        }
                    } else if (active) {
#pragma unroll
                        for (int i = I0; i < I1; ++i)
#pragma unroll
                        for (int k = 0; k < NFK; ++k)
#pragma unroll
                        for (int l = 0; l < NFL; ++l) {
                            double g = gout[(((i - I0)*NFJ + j)*NFK + k)*NFL + l];
                            kjl[l] = fma(g, D[(i0+i)*nao + k0+k], kjl[l]);
                            kjk[k] = fma(g, D[(i0+i)*nao + l0+l], kjk[k]);
                        }
                    }
                    if (kr.kcols && LANE_KIK_SEGSUM) {
#pragma unroll
                        for (int k = 0; k < NFK; ++k) {
                            double v = seg_sum(kjk[k], runs);
                            if (tail) lds_add(kr.p + (NFI + j) * kr.width + k0 + k - kr.k_ao0, v);
                        }
                    } else if (kr.kcols) {
                        if (active) {
#pragma unroll
                            for (int k = 0; k < NFK; ++k)
                                lds_add(kr.p + (NFI + j) * kr.width + k0 + k - kr.k_ao0, kjk[k]);
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < NFK; ++k) {
                            double v = seg_sum(kjk[k], runs);
                            if (tail) atomic_add(K + (j0+j)*nao + k0+k, v);
                        }
                    }
                    if (active) {
#pragma unroll
                        for (int l = 0; l < NFL; ++l) {
                            if (kr.on) lds_add(kr.p + (NFI + j) * kr.width + kr.loff + l0 + l - kr.l_ao0, kjl[l]);
                            else atomic_add(K + (j0+j)*nao + l0+l, kjl[l]);
                        }
                    }
                }
            }
            if (a.vj) {
                double *J = a.vj + id * nao2;
                // the D_kl loads of J_ij come before the J_kl atomics: on this GPU
                // stores and atomics share the load counter, so a load issued
                // after an atomic waits for it to complete.
                // J_ij partial sums of this lane (D_kl): reduced over the wave below
                double vij[NI*NFJ];
#pragma unroll
                for (int i = I0; i < I1; ++i)
#pragma unroll
                for (int j = 0; j < NFJ; ++j) {
                    double s = 0.;
                    if (active) {
#pragma unroll
                        for (int k = 0; k < NFK; ++k)
#pragma unroll
                        for (int l = 0; l < NFL; ++l)
                            s = fma((double)gout[(((i - I0)*NFJ + j)*NFK + k)*NFL + l],
                                    D[(k0+k)*nao + l0+l], s);
                    }
                    vij[(i - I0)*NFJ + j] = s;
                }
                // J_kl += (ij|kl) D_ij   (per lane)
                if (active) {
#pragma unroll
                    for (int k = 0; k < NFK; ++k)
#pragma unroll
                    for (int l = 0; l < NFL; ++l) {
                        double s = 0.;
#pragma unroll
                        for (int i = I0; i < I1; ++i)
#pragma unroll
                        for (int j = 0; j < NFJ; ++j)
                            s = fma((double)gout[(((i - I0)*NFJ + j)*NFK + k)*NFL + l],
                                    D[(i0+i)*nao + j0+j], s);
                        atomic_add(J + (k0+k)*nao + l0+l, s);
                    }
                }
                // J_ij += (ij|kl) D_kl   (common ij: reduce over the wave)
                wave_reduce_scatter<NI*NFJ>(vij, lane, [&](int n, double x) {
                    atomic_add(J + (i0 + I0 + n / NFJ)*nao + j0 + n % NFJ, x);
                });
            }
        }
    });
}

// Per-lane kernel.  LANE_WAVES wavefronts share one work item (and its LDS K
// rows); each wavefront scans every LANE_WAVES-th batch of 64 kl candidates
// and keeps two queues: quartets whose bound q_ij + q_kl + max(dm_cond)
// reaches fp32_cut run in FP64, the others in FP32.  PM (precision mode):
// 0 both queues, 1 the FP64 queue only, 2 the FP32 queue only (a launch
// with PM 1 and one with PM 2 together evaluate what one with PM 0 does;
// each is compiled without the other's arithmetic).  FP64 builds launch PM 1
// (64-water build 2110 -> 2092 ms against PM 0); mixed builds PM 0, since
// the pair PM 1 + PM 2 scans and flushes every work item twice (2115 ->
// 2253 ms; the FP32-only kernels mostly keep the occupancy of PM 0).
// Minimum waves per SIMD a class kernel is compiled for.  By default the
// compiler may use all 512 registers (one wave per SIMD); the classes below
// need a little over 256 and run faster with two waves per SIMD despite a few
// spilled registers (32-water def2-TZVP build, ms, one wave -> two:
// (fs|ps) 23.1 -> 16.9, (ds|pp) 15.2 -> 13.2, (ds|ds) 9.9 -> 7.1,
// (dd|pp) 11.6 -> 8.8, (fp|fp) 11.2 -> 8.5, (fd|dp) 9.4 -> 6.2; the classes
// that need 350 or more registers ran slower with two waves).  One wave per
// SIMD more for the smallest classes ((ss|ss) 6, (ps|ss) 5, (ps|ps) and
// (ds|ss) 4, (pp|ss) and (ds|ps) 3) measured slower on the 64-water build
// ((ps|ss) 119.3 -> 143.7 ms, whole build 2197 -> 2249 ms on one stream).
template <int LI, int LJ, int LK, int LL> struct ClassOcc { static constexpr int W = 1; };
#define CLASS_OCC2(a, b, c, d) \
    template <> struct ClassOcc<a, b, c, d> { static constexpr int W = 2; };
CLASS_OCC2(3, 0, 1, 0) CLASS_OCC2(2, 0, 1, 1) CLASS_OCC2(2, 0, 2, 0)
CLASS_OCC2(3, 1, 3, 1) CLASS_OCC2(3, 2, 2, 1) CLASS_OCC2(2, 2, 1, 1)
CLASS_OCC2(2, 1, 2, 2) CLASS_OCC2(3, 3, 3, 0) CLASS_OCC2(2, 2, 2, 1)
CLASS_OCC2(3, 0, 3, 2) CLASS_OCC2(3, 1, 2, 2) CLASS_OCC2(3, 3, 2, 1)
CLASS_OCC2(3, 2, 3, 0)
#undef CLASS_OCC2


template <int LI, int LJ, int LK, int LL, int PM>
__global__ __launch_bounds__(LANE_WAVES * WAVE, (ClassOcc<LI, LJ, LK, LL>::W))
void jk_lane_kernel(JKBuildArgs a, int kl_chunk, int use_lds)
{
    constexpr int NFI = ncart(LI), NFJ = ncart(LJ);
    __shared__ uint32_t queue_all[LANE_WAVES][2][2 * WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    uint32_t *queue64 = queue_all[wid][0];
    uint32_t *queue32 = queue_all[wid][1];
    const int nbas = a.nbas, nao = a.nao;
    select_replica(a);
    WorkItem wi = work_item(a, kl_chunk, use_lds);
    const uint32_t bas_ij = a.pairs_ij[wi.pij];
    const float q_ij = a.q_ij[wi.pij];
    const float kl_cutoff = a.cutoff - q_ij;
    if (a.q_kl[wi.kl0] + a.dm_penalty + Q_COND_MARGIN < kl_cutoff) return;
    ShellPair sp;
    load_shell_pair(sp, bas_ij, nbas, a.bas, a.env, a.ao_loc);
    // K rows of this work item accumulated in LDS (one density)
    extern __shared__ double kacc_dyn[];
    const KRows kr = make_krows(kacc_dyn, use_lds, NFI + NFJ, wi);
    if (kr.on) {
        krows_zero(kr, NFI + NFJ, threadIdx.x, LANE_WAVES * WAVE);
        krows_load_d(kr, NFI, NFJ, sp.i0, sp.j0, a, threadIdx.x, LANE_WAVES * WAVE);
        __syncthreads();
    }
    int qn64 = 0, qn32 = 0;
    unsigned long long nsurv = 0, nsurv32 = 0, hsum = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    bool stop = false;
    for (int base = wi.kl0 + wid * WAVE; base < wi.kl1 || qn64 > 0 || qn32 > 0;
         base += LANE_WAVES * WAVE) {
        if (base < wi.kl1 && !stop) {
            int pkl = base + lane;
            bool keep = false, fail_margin = false, lowp = false;
            if (pkl < wi.kl1) {
                float q_kl = a.q_kl[pkl];
                fail_margin = q_kl + a.dm_penalty + Q_COND_MARGIN < kl_cutoff;
                float dmax;
                const uint32_t bkl = a.pairs_kl[pkl];
                keep = survives_dmax(q_kl, kl_cutoff, a.dm_penalty, bas_ij, bkl,
                                     sp.ish, sp.jsh, nbas, a.dm_cond, a.screen, dmax,
                                     a.swap_ijkl != 0, q_ij, a.cutoff);
                lowp = keep && (q_ij + q_kl + dmax < a.fp32_cut);
                // the quartets of the other precision belong to the other launch
                if (PM == 1 && lowp) keep = lowp = false;
                if (PM == 2 && !lowp) keep = false;
                if (a.quartet_hash && keep) hsum += task_mix(a, bas_ij, bkl);
            }
            const uint64_t m64 = __ballot(keep && !lowp);
            const uint64_t m32 = __ballot(lowp);
            if (keep && !lowp) queue64[qn64 + __popcll(m64 & lt_mask)] = pkl;
            if (lowp) queue32[qn32 + __popcll(m32 & lt_mask)] = pkl;
            qn64 += __popcll(m64);
            qn32 += __popcll(m32);
            nsurv += __popcll(m64) + __popcll(m32);
            nsurv32 += __popcll(m32);
            stop = __ballot(fail_margin) != 0;
            wave_sync();
        }
        const bool flush = (base + LANE_WAVES * WAVE >= wi.kl1) || stop;
        if (PM != 2 && (qn64 >= WAVE || (flush && qn64 > 0))) {
            const int nb = min(qn64, WAVE);
            if constexpr (PM != 2)
                lane_batch<LI, LJ, LK, LL, double>(a, sp, bas_ij, wi.pij, queue64, nb, lane,
                                                   lt_mask, kr);
            wave_sync();
            uint32_t rest = (nb + lane < qn64) ? queue64[nb + lane] : 0u;
            wave_sync();
            if (nb + lane < qn64) queue64[lane] = rest;
            qn64 -= nb;
            wave_sync();
        }
        if (PM != 1 && (qn32 >= WAVE || (flush && qn32 > 0))) {
            const int nb = min(qn32, WAVE);
            if constexpr (PM != 1)
                lane_batch<LI, LJ, LK, LL, float>(a, sp, bas_ij, wi.pij, queue32, nb, lane,
                                                  lt_mask, kr);
            wave_sync();
            uint32_t rest = (nb + lane < qn32) ? queue32[nb + lane] : 0u;
            wave_sync();
            if (nb + lane < qn32) queue32[lane] = rest;
            qn32 -= nb;
            wave_sync();
        }
        if (flush && qn64 == 0 && qn32 == 0) break;
    }
    if (kr.on) {
        __syncthreads();
        krows_flush(kr, NFI, NFJ, sp.i0, sp.j0, a, threadIdx.x, LANE_WAVES * WAVE);
    }
    if (lane == 0) {
        if (a.n_quartets && nsurv) atomicAdd(a.n_quartets, nsurv);
        if (a.n_quartets_fp32 && nsurv32) atomicAdd(a.n_quartets_fp32, nsurv32);
    }
    if (a.quartet_hash && hsum) atomicAdd(a.quartet_hash, hsum);
}

// ---------------------------------------------------------------- team
// Team of TS lanes per quartet, NQ quartets per block of TB threads.
template <int LI, int LJ, int LK, int LL>
struct TeamShape {
    static constexpr int NFI = ncart(LI), NFJ = ncart(LJ), NFK = ncart(LK), NFL = ncart(LL);
    static constexpr int NROWS = NFI * NFJ;
    static constexpr int NFKL = NFK * NFL;
    static constexpr int N = (LI + LJ + LK + LL) / 2 + 1;
    static constexpr int GS = (LI+1)*(LJ+1)*(LK+1)*(LL+1);
    static constexpr int TS0 = NROWS > 3 * N ? NROWS : 3 * N;
    static constexpr int TB1 = TS0 <= 64 ? 64 : (TS0 <= 128 ? 128 : 256);
    static constexpr int TB = TB1 < TEAM_TB_MIN ? TEAM_TB_MIN : TB1;
    static constexpr int G_DBL = 2 * N + N * 3 * GS;            // roots + 2D integrals
    static constexpr int S_ROWS0 = 4096 / NFKL;                 // rows staged per pass
    static constexpr int S_ROWS = S_ROWS0 < NROWS ? (S_ROWS0 > 0 ? S_ROWS0 : 1) : NROWS;
    static constexpr int S_DBL = S_ROWS * NFKL;
    static constexpr int PER_Q = G_DBL > S_DBL ? G_DBL : S_DBL;
    static constexpr int NQ0 = TB / TS0;
    static constexpr int NQ1 = (6144 / PER_Q) > 0 ? (6144 / PER_Q) : 1;   // 48 KB of LDS
    static constexpr int NQ = NQ0 < NQ1 ? NQ0 : NQ1;
    static constexpr int TS = TB / NQ;
    // primitive quartets per pass: the 3N lanes of one primitive quartet's
    // 2D tables are replicated PB times over the team, within the same
    // 48 KB of LDS (one pass of roots, tables and triple products, three
    // block barriers, per PB primitive quartets)
    static constexpr int PB_LANES = TS / (3 * N) > 0 ? TS / (3 * N) : 1;
    static constexpr int PB_LDS0 = 6144 / (NQ * G_DBL);
    static constexpr int PB_LDS = PB_LDS0 > 0 ? PB_LDS0 : 1;
    static constexpr int PB = TEAM_PRIM_BATCH ? (PB_LANES < PB_LDS ? PB_LANES : PB_LDS) : 1;
    static constexpr int G_ALL = PB * G_DBL;
    static constexpr int PER_QB = G_ALL > S_DBL ? G_ALL : S_DBL;
    // per-team staging of the density blocks of the current quartet:
    // D_kl, then rows r of i (r < NFI) and j with their k and l columns,
    // then D_ij (used when several densities are digested)
    static constexpr int DROW = NFK + NFL;
    static constexpr int D_KL = 0, D_R = NFKL, D_IJ = D_R + (NFI + NFJ) * DROW;
    static constexpr int DSZ = D_IJ + NFI * NFJ;
};

template <int LI, int LJ, int LK, int LL>
__global__ __launch_bounds__((TeamShape<LI, LJ, LK, LL>::TB),
                             (TEAM_MIN_WAVES > ClassOcc<LI, LJ, LK, LL>::W ?
                              TEAM_MIN_WAVES : ClassOcc<LI, LJ, LK, LL>::W))
void jk_team_kernel(JKBuildArgs a, int kl_chunk, int use_lds)
{
    using S = TeamShape<LI, LJ, LK, LL>;
    constexpr int NFI = S::NFI, NFJ = S::NFJ, NFK = S::NFK, NFL = S::NFL;
    constexpr int NROWS = S::NROWS, NFKL = S::NFKL, N = S::N, GS = S::GS;
    constexpr int TB = S::TB, NQ = S::NQ, TS = S::TS;
    constexpr int NW = TB / WAVE;
    constexpr CartPow<LI> ci{};
    constexpr CartPow<LJ> cj{};
    constexpr CartPow<LK> ck{};
    constexpr CartPow<LL> cl{};
    constexpr int SJ = LL+1, SK = (LK+1)*(LL+1), SI = (LJ+1)*(LK+1)*(LL+1);

    __shared__ uint32_t queue[2 * TB];
    __shared__ int wave_cnt[NW + 1];
    __shared__ int stop_flag;
    __shared__ double lds[NQ * S::PER_QB];
    __shared__ double dstage[NQ * S::DSZ];
    __shared__ double dij_blk[NFI * NFJ];     // D_ij of the work item (one density)

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wid = tid / WAVE;
    const int team = tid / TS;
    const int tl = tid - team * TS;
    const int nbas = a.nbas, nao = a.nao;
    const size_t nao2 = (size_t)nao * nao;
    select_replica(a);
    WorkItem wi = work_item(a, kl_chunk, use_lds);
    const uint32_t bas_ij = a.pairs_ij[wi.pij];
    const float kl_cutoff = a.cutoff - a.q_ij[wi.pij];
    if (a.q_kl[wi.kl0] + a.dm_penalty + Q_COND_MARGIN < kl_cutoff) return;
    ShellPair sp;
    load_shell_pair(sp, bas_ij, nbas, a.bas, a.env, a.ao_loc);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    constexpr int PB = S::PB;
    double *my = lds + (team < NQ ? team : 0) * S::PER_QB;
    // per primitive quartet p < PB of a pass: roots and weights at
    // my + p * G_DBL ([2N]), then its 2D tables ([N][3][GS])
    const int row = tl;               // (i, j) row owned by this lane
    const bool has_row = team < NQ && row < NROWS;
    const int irow = has_row ? row / NFJ : 0, jrow = has_row ? row % NFJ : 0;
    // J_ij of this lane's row, summed over the block's quartets (one density)
    double vj_row = 0.;
    // K rows of this work item accumulated in LDS (one density)
    extern __shared__ double kacc_dyn[];
    const KRows kr = make_krows(kacc_dyn, use_lds, NFI + NFJ, wi);
    krows_zero(kr, NFI + NFJ, tid, TB);
    for (int n = tid; n < NFI * NFJ; n += TB)
        dij_blk[n] = a.dm[(size_t)(sp.i0 + n / NFJ) * nao + sp.j0 + n % NFJ];
    double *dsh = dstage + (team < NQ ? team : 0) * S::DSZ;
    __syncthreads();
    int qn = 0;
    unsigned long long nsurv = 0, hsum = 0;
    bool stop = false;
    for (int base = wi.kl0; ; base += TB) {
        if (base < wi.kl1 && !stop) {
            int pkl = base + tid;
            bool keep = false, fail_margin = false;
            uint32_t bkl = 0;
            if (pkl < wi.kl1) {
                float q_kl = a.q_kl[pkl];
                bkl = a.pairs_kl[pkl];
                fail_margin = q_kl + a.dm_penalty + Q_COND_MARGIN < kl_cutoff;
                keep = survives(q_kl, kl_cutoff, a.dm_penalty, bas_ij, bkl,
                                sp.ish, sp.jsh, nbas, a.dm_cond, a.screen, a.swap_ijkl != 0,
                                a.q_ij[wi.pij], a.cutoff);
                if (a.quartet_hash && keep) hsum += task_mix(a, bas_ij, bkl);
            }
            uint64_t m = __ballot(keep);
            uint64_t mf = __ballot(fail_margin);
            if (tid == 0) stop_flag = 0;
            __syncthreads();
            if (lane == 0) {
                wave_cnt[wid] = __popcll(m);
                if (mf) stop_flag = 1;
            }
            __syncthreads();
            int off = qn;
            for (int w = 0; w < wid; ++w) off += wave_cnt[w];
            if (keep) queue[off + __popcll(m & lt_mask)] = pkl;
            int tot = 0;
            for (int w = 0; w < NW; ++w) tot += wave_cnt[w];
            qn += tot;
            nsurv += tot;
            stop = stop_flag != 0;
            __syncthreads();
        }
        bool flush = base + TB >= wi.kl1 || stop;
        int head = 0;
        while (qn - head >= NQ || (flush && head < qn)) {
            // team `team` evaluates queue[head + team]
            const int qidx = head + team;
            const bool active = team < NQ && qidx < qn;
            const int pkl = queue[active ? qidx : head];
            uint32_t bas_kl = a.pairs_kl[pkl];
            int ksh = bas_kl / nbas, lsh = bas_kl - (bas_kl / nbas) * nbas;
            double fac_sym = a.scale;
            if (sp.ish == sp.jsh) fac_sym *= .5;
            if (ksh == lsh) fac_sym *= .5;
            if (bas_ij == bas_kl) fac_sym *= .5;
            const double *rk = a.env + a.bas[ksh * 8 + 7], *rl = a.env + a.bas[lsh * 8 + 7];
            double rkl[3] = {rl[0] - rk[0], rl[1] - rk[1], rl[2] - rk[2]};
            const double omega2 = a.omega * a.omega;
            const int nij = a.iprim * a.jprim, nkl = a.kprim * a.lprim;
            const double *pdij = a.pdata_ij + (size_t)wi.pij * nij * PREC;
            const double *pdkl = a.pdata_kl + (size_t)pkl * nkl * PREC;
            double acc[NFKL];
#pragma unroll
            for (int n = 0; n < NFKL; ++n) acc[n] = 0.;
            const int npq = nij * nkl;
            // primitive-pair quantities of primitive quartet pq (ij fastest)
            auto prim = [&](int pq, double &aij, double &raij, double &akl, double &rakl,
                            double *PA, double *PQ, double *QC, double &r_aa, double &theta,
                            double &rr, double &fac) {
                const int kp = pq / nij, ip = pq - kp * nij;
                const double *qr = pdkl + kp * PREC;
                const double *pr = pdij + ip * PREC;
                akl = qr[0]; rakl = qr[1];
                aij = pr[0]; raij = pr[1];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    PA[d] = pr[2+d] - sp.ri[d];
                    PQ[d] = pr[2+d] - qr[2+d];
                    QC[d] = qr[2+d] - rk[d];
                }
                rr = PQ[0]*PQ[0] + PQ[1]*PQ[1] + PQ[2]*PQ[2];
                r_aa = 1. / (aij + akl);
                theta = aij * akl * r_aa;
                fac = pr[5] * qr[5] * fac_sym * sqrt(r_aa);
            };
            for (int pq0 = 0; pq0 < npq; pq0 += PB) {
                const int np = min(PB, npq - pq0);
                __syncthreads();
                // roots: lane (p, r) for p < np, r < N
                if (active && tl < np * N) {
                    const int p = tl / N, r = tl - p * N;
                    double aij, raij, akl, rakl, PA[3], PQ[3], QC[3], r_aa, theta, rr, fac;
                    prim(pq0 + p, aij, raij, akl, rakl, PA, PQ, QC, r_aa, theta, rr, fac);
                    double rt, wt;
                    if (omega2 > 0.) {
                        // long-range attenuation (see eri_lane)
                        const double aw = omega2 / (omega2 + theta);
                        rys_root_one<N>(a.rys, theta * rr * aw, r, rt, wt);
                        rt *= aw; wt *= sqrt(aw);
                    } else {
                        rys_root_one<N>(a.rys, theta * rr, r, rt, wt);
                    }
                    double *rw = my + p * S::G_DBL;
                    rw[2*r] = rt; rw[2*r+1] = wt;
                }
                __syncthreads();
                // 2D tables: lane (p, r, d) for p < np
                if (active && tl < np * 3 * N) {
                    const int p = tl / (3 * N), rd = tl - p * (3 * N);
                    const int r = rd / 3, d = rd - (rd / 3) * 3;
                    double aij, raij, akl, rakl, PA[3], PQ[3], QC[3], r_aa, theta, rr, fac;
                    prim(pq0 + p, aij, raij, akl, rakl, PA, PQ, QC, r_aa, theta, rr, fac);
                    const double *rw = my + p * S::G_DBL;
                    double rt_aa = rw[2*r] * r_aa;
                    double rt_aij = rt_aa * akl, rt_akl = rt_aa * aij;
                    double b10 = .5 * raij * (1. - rt_aij);
                    double b01 = .5 * rakl * (1. - rt_akl);
                    double b00 = .5 * rt_aa;
                    double g[GS];
                    double pa = d == 0 ? PA[0] : (d == 1 ? PA[1] : PA[2]);
                    double pq = d == 0 ? PQ[0] : (d == 1 ? PQ[1] : PQ[2]);
                    double qc = d == 0 ? QC[0] : (d == 1 ? QC[1] : QC[2]);
                    double rd_ = d == 0 ? sp.rij[0] : (d == 1 ? sp.rij[1] : sp.rij[2]);
                    double kd = d == 0 ? rkl[0] : (d == 1 ? rkl[1] : rkl[2]);
                    double g00 = d == 2 ? rw[2*r+1] * fac : 1.;
                    build_g<LI, LJ, LK, LL, double>(g, g00, pa - rt_aij * pq,
                                            qc + rt_akl * pq,
                                            b10, b01, b00, rd_, kd);
                    double *dst = my + p * S::G_DBL + 2 * N + (r * 3 + d) * GS;
#pragma unroll
                    for (int n = 0; n < GS; ++n) dst[n] = g[n];
                }
                __syncthreads();
                if (active && has_row) {
                    int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
                    // Cartesian powers of this row's i and j (runtime lookup)
#pragma unroll
                    for (int i = 0; i < NFI; ++i)
                        if (i == irow) { ax = ci.x[i]; ay = ci.y[i]; az = ci.z[i]; }
#pragma unroll
                    for (int j = 0; j < NFJ; ++j)
                        if (j == jrow) { bx = cj.x[j]; by = cj.y[j]; bz = cj.z[j]; }
                    const int ox = ax * SI + bx * SK, oy = ay * SI + by * SK, oz = az * SI + bz * SK;
                    for (int p = 0; p < np; ++p) {
                        const double *gb = my + p * S::G_DBL + 2 * N;
                        for (int r = 0; r < N; ++r) {
                            const double *gx = gb + (r * 3 + 0) * GS + ox;
                            const double *gy = gb + (r * 3 + 1) * GS + oy;
                            const double *gz = gb + (r * 3 + 2) * GS + oz;
                            double sx[SK], sy[SK], sz[SK];
#pragma unroll
                            for (int n = 0; n < SK; ++n) { sx[n] = gx[n]; sy[n] = gy[n]; sz[n] = gz[n]; }
#pragma unroll
                            for (int k = 0; k < NFK; ++k)
#pragma unroll
                            for (int l = 0; l < NFL; ++l) {
                                int kx = ck.x[k]*SJ + cl.x[l];
                                int ky = ck.y[k]*SJ + cl.y[l];
                                int kz = ck.z[k]*SJ + cl.z[l];
                                acc[k*NFL + l] = fma(sx[kx] * sy[ky], sz[kz], acc[k*NFL + l]);
                            }
                        }
                    }
                }
            }
            // ---- digestion of the team's quartet
            const int i0 = sp.i0, j0 = sp.j0;
            const int k0 = a.ao_loc[ksh], l0 = a.ao_loc[lsh];
            double *stage = my;       // reuse the team's LDS
            constexpr int DROW = S::DROW;
            const double *dkl = dsh + S::D_KL, *drow = dsh + S::D_R;
            for (int id = 0; id < a.n_dm; ++id) {
                const double *D = a.dm + id * nao2;
                // the quartet's density blocks into the team's LDS, one
                // cooperative gather
                __syncthreads();
                if (active) {
                    for (int n = tl; n < S::DSZ; n += TS) {
                        size_t idx;
                        if (n < S::D_R) {
                            idx = (size_t)(k0 + n / NFL) * nao + l0 + n % NFL;
                        } else if (n < S::D_IJ) {
                            const int r = (n - S::D_R) / DROW, c = (n - S::D_R) - r * DROW;
                            const int ao = r < NFI ? i0 + r : j0 + r - NFI;
                            idx = (size_t)ao * nao + (c < NFK ? k0 + c : l0 + c - NFK);
                        } else {
                            if (a.n_dm == 1) continue;
                            const int m = n - S::D_IJ;
                            idx = (size_t)(i0 + m / NFJ) * nao + j0 + m % NFJ;
                        }
                        dsh[n] = D[idx];
                    }
                }
                __syncthreads();
                const double *dij = a.n_dm == 1 ? dij_blk : dsh + S::D_IJ;
                if (a.vj && active && has_row) {
                    double s = 0.;
#pragma unroll
                    for (int k = 0; k < NFK; ++k)
#pragma unroll
                    for (int l = 0; l < NFL; ++l)
                        s = fma(acc[k*NFL + l], dkl[k*NFL + l], s);
                    if (a.n_dm == 1) vj_row += s;
                    else atomic_add(a.vj + id * nao2 + (i0+irow)*nao + j0+jrow, s);
                }
                for (int r0 = 0; r0 < NROWS; r0 += S::S_ROWS) {
                    const int r1 = min(r0 + S::S_ROWS, NROWS);
                    __syncthreads();
                    if (active && has_row && row >= r0 && row < r1) {
#pragma unroll
                        for (int n = 0; n < NFKL; ++n) stage[(row - r0) * NFKL + n] = acc[n];
                    }
                    __syncthreads();
                    if (!active) continue;
                    if (a.vj) {
                        double *J = a.vj + id * nao2;
                        for (int e = tl; e < NFKL; e += TS) {
                            int k = e / NFL, l = e - (e / NFL) * NFL;
                            double s = 0.;
                            for (int rr = r0; rr < r1; ++rr)
                                s = fma(stage[(rr - r0) * NFKL + e], dij[rr], s);
                            atomic_add(J + (k0+k)*nao + l0+l, s);
                        }
                    }
                    if (a.vk) {
                        double *K = a.vk + id * nao2;
                        const int ilo = r0 / NFJ, ihi = (r1 - 1) / NFJ;
                        const int ni = ihi - ilo + 1;
                        // K_ik (sum over j,l of D_jl) and K_il (sum over j,k of D_jk)
                        for (int e = tl; e < ni * (NFK + NFL); e += TS) {
                            int i = ilo + e / (NFK + NFL);
                            int m = e - (e / (NFK + NFL)) * (NFK + NFL);
                            double s = 0.;
                            for (int j = 0; j < NFJ; ++j) {
                                int rr = i * NFJ + j;
                                if (rr < r0 || rr >= r1) continue;
                                const double *srow = stage + (rr - r0) * NFKL;
                                const double *dj = drow + (NFI + j) * DROW;
                                if (m < NFK) {
                                    for (int l = 0; l < NFL; ++l)
                                        s = fma(srow[m*NFL + l], dj[NFK + l], s);
                                } else {
                                    int l = m - NFK;
                                    for (int k = 0; k < NFK; ++k)
                                        s = fma(srow[k*NFL + l], dj[k], s);
                                }
                            }
                            if (m < NFK) {
                                if (kr.kcols) lds_add(kr.p + i * kr.width + k0 + m - kr.k_ao0, s);
                                else atomic_add(K + (i0+i)*nao + k0+m, s);
                            } else if (kr.on) {
                                lds_add(kr.p + i * kr.width + kr.loff + l0 + m - NFK - kr.l_ao0, s);
                            } else {
                                atomic_add(K + (i0+i)*nao + l0+m-NFK, s);
                            }
                        }
                        // K_jk (sum over i,l of D_il) and K_jl (sum over i,k of D_ik)
                        for (int e = tl; e < NFJ * (NFK + NFL); e += TS) {
                            int j = e / (NFK + NFL);
                            int m = e - j * (NFK + NFL);
                            double s = 0.;
                            for (int i = ilo; i <= ihi; ++i) {
                                int rr = i * NFJ + j;
                                if (rr < r0 || rr >= r1) continue;
                                const double *srow = stage + (rr - r0) * NFKL;
                                const double *di = drow + i * DROW;
                                if (m < NFK) {
                                    for (int l = 0; l < NFL; ++l)
                                        s = fma(srow[m*NFL + l], di[NFK + l], s);
                                } else {
                                    int l = m - NFK;
                                    for (int k = 0; k < NFK; ++k)
                                        s = fma(srow[k*NFL + l], di[k], s);
                                }
                            }
                            if (m < NFK) {
                                if (kr.kcols) lds_add(kr.p + (NFI + j) * kr.width + k0 + m - kr.k_ao0, s);
                                else atomic_add(K + (j0+j)*nao + k0+m, s);
                            } else if (kr.on) {
                                lds_add(kr.p + (NFI + j) * kr.width + kr.loff + l0 + m - NFK - kr.l_ao0, s);
                            } else {
                                atomic_add(K + (j0+j)*nao + l0+m-NFK, s);
                            }
                        }
                    }
                }
            }
            __syncthreads();
            head += NQ;
        }
        // move the unprocessed remainder to the front of the queue
        int rem = qn - head;
        uint32_t v = 0;
        if (tid < rem) v = queue[head + tid];
        __syncthreads();
        if (tid < rem) queue[tid] = v;
        qn = rem;
        __syncthreads();
        if (flush) break;
    }
    if (a.vj && a.n_dm == 1 && has_row && vj_row != 0.)
        atomic_add(a.vj + (sp.i0 + irow)*nao + sp.j0 + jrow, vj_row);
    if (kr.on) {
        __syncthreads();
        krows_flush(kr, NFI, NFJ, sp.i0, sp.j0, a, tid, TB);
    }
    if (a.n_quartets && tid == 0 && nsurv) atomicAdd(a.n_quartets, nsurv);
    if (a.quartet_hash && hsum) atomicAdd(a.quartet_hash, hsum);
}

// ---------------------------------------------------------------- dispatch
// LDS bytes a workgroup may use (static + dynamic): the device limit, capped
// at 64 KB.  gfx950 reports 160 KB per workgroup, but every launch of this
// library stays within the 64 KB that has been validated on the hardware.
#define JK_LDS_CAP (64 * 1024)
static size_t lds_limit()
{
    static int cached_dev = -1;
    static size_t cached = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return JK_LDS_CAP;
    if (dev != cached_dev) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess
            || v <= 0 || v > JK_LDS_CAP)
            v = JK_LDS_CAP;
        cached = (size_t)v;
        cached_dev = dev;
    }
    return cached;
}

// LDS K rows for a launch: 2 = K_ik/K_jk and K_il/K_jl, 1 = K_il/K_jl only,
// 0 = none (several densities, no K, or too little LDS); `bytes` receives
// the dynamic LDS size.  max_mode 1 suits kernels whose many short work
// items would spend more on zeroing and flushing the wider rows (and on the
// occupancy they cost) than the K_ik/K_jk atomics they save.
// k_nao / l_nao: the K-row columns of a work item (the whole k and l groups
// of the task, or with kl column windows the widest window, l_nao = 0 and
// k_nao = kl_win_width).
template <int LI, int LJ>
static int krows_mode(const JKBuildArgs &a, int k_nao, int l_nao, size_t lds_static,
                      int max_mode, bool drows, size_t &bytes)
{
    constexpr size_t NFIJ = ncart(LI) + ncart(LJ);
    bytes = 0;
    if (!a.vk || a.n_dm != 1 || l_nao < 0 || k_nao < 0 || k_nao + l_nao <= 0) return 0;
    const size_t both = NFIJ * (size_t)(k_nao + l_nao) * sizeof(double);
    const size_t lonly = NFIJ * (size_t)l_nao * sizeof(double);
    if (max_mode >= 2 && drows && 2 * both + lds_static <= lds_limit()) {
        bytes = 2 * both;
        return 2 | KR_DROWS;
    }
    if (max_mode >= 2 && both + lds_static <= lds_limit()) { bytes = both; return 2; }
    if (max_mode >= 1 && l_nao > 0 && lonly + lds_static <= lds_limit()) {
        bytes = lonly;
        return 1;
    }
    return 0;
}

// LDS K rows of a per-lane launch.  Candidates: the rows over the whole k and
// l groups and, when the task has kl column windows, over one window, each
// with and without the density rows.  Ranked by the rows kept in LDS (all
// four K blocks, then K_il/K_jl only, then none), then by the workgroups a
// CU can hold with that LDS (hipOccupancyMaxActiveBlocksPerMultiprocessor,
// which also counts the kernel's registers), then whole groups before
// windows, then with density rows before without.  kl_win_policy 1 takes the
// windows whenever their rows fit.
struct KrChoice { int use_lds; size_t bytes; };

static KrChoice lane_krows_choice(const JKBuildArgs &a, size_t lds_static, bool win_ok,
                                  const void *kernel, int nfij)
{
    KrChoice best{0, 0};
    long best_key = -1;
    const bool have_win = a.kl_win[0].win && win_ok && LANE_KROWS >= 2;
    for (int win = 0; win < (have_win ? 2 : 1); ++win) {
        const int kn = win ? a.kl_win[0].width : a.k_nao, ln = win ? 0 : a.l_nao;
        for (int dr = (LANE_DROWS ? 1 : 0); dr >= 0; --dr) {
            size_t bytes = 0;
            // the generic krows_mode with this candidate's columns
            int mode = 0;
            if (a.vk && a.n_dm == 1 && kn + ln > 0) {
                const size_t both = (size_t)nfij * (kn + ln) * sizeof(double);
                const size_t lonly = (size_t)nfij * ln * sizeof(double);
                if (LANE_KROWS >= 2 && dr && 2 * both + lds_static <= lds_limit()) {
                    mode = 2 | KR_DROWS; bytes = 2 * both;
                } else if (LANE_KROWS >= 2 && !dr && both + lds_static <= lds_limit()) {
                    mode = 2; bytes = both;
                } else if (!win && LANE_KROWS >= 1 && ln > 0 && lonly + lds_static <= lds_limit()) {
                    mode = 1; bytes = lonly;
                }
            }
            if (win && (mode & KR_MODE_MASK) != 2) continue;
            int nb = 1;
#if LANE_KROWS_OCC
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, LANE_WAVES * WAVE,
                                                             bytes) != hipSuccess || nb < 1)
                nb = 1;
#endif
            long key = ((long)(mode & KR_MODE_MASK) << 24) | ((long)nb << 8) |
                       ((win ? 0 : 1) << 1) | (dr ? 1 : 0);
            if (a.kl_win_policy == 1 && win) key |= 1L << 30;
            if (key > best_key) {
                best_key = key;
                best.use_lds = mode | (win ? KR_WIN : 0);
                best.bytes = bytes;
            }
        }
    }
    return best;
}

template <int LI, int LJ, int LK, int LL>
struct ClassTraits {
    static constexpr int NF = ncart(LI) * ncart(LJ) * ncart(LK) * ncart(LL);
    static constexpr int GS = (LI+1)*(LJ+1)*(LK+1)*(LL+1);
    // per-lane path when the block and three 2D tables fit in LANE_MAX doubles,
    // or when chunks of at most LANE_MAX_CHUNKS row blocks of the block do
    static constexpr bool LANE = (NF + 3 * GS <= LANE_MAX ||
                                  (LaneChunk<LI, LJ, LK, LL>::FIT > 0 &&
                                   (LaneChunked<LI, LJ, LK, LL>::ON ||
                                    LaneChunk<LI, LJ, LK, LL>::NCH <= LANE_MAX_CHUNKS))) &&
                                 NF != LANE_SKIP_NF;
};

// one per-lane launch of precision mode PM (whole kl chunks or windows, as
// lane_krows_choice ranks them for that kernel)
template <int LI, int LJ, int LK, int LL, int PM>
static hipError_t launch_lane(const JKBuildArgs &a, hipStream_t stream, long nblocks,
                              int kl_chunk)
{
    const size_t lds_static = sizeof(uint32_t) * LANE_WAVES * 2 * 2 * WAVE;
    const long nblocks_w = a.kl_win[0].win ? (long)a.npairs_ij * a.kl_win[0].n_win : 0;
    const KrChoice c = lane_krows_choice(a, lds_static, nblocks_w <= 0x7fffffffL,
                                         (const void *)jk_lane_kernel<LI, LJ, LK, LL, PM>,
                                         ncart(LI) + ncart(LJ));
    if (c.use_lds & KR_WIN) {
        // K rows over one kl column window per work item
        JKBuildArgs aw = a;
        aw.pairs_kl = a.kl_win[0].pairs;
        aw.q_kl = a.kl_win[0].q;
        aw.pdata_kl = a.kl_win[0].pdata;
        hipLaunchKernelGGL((jk_lane_kernel<LI, LJ, LK, LL, PM>), dim3(nblocks_w),
                           dim3(LANE_WAVES * WAVE), c.bytes, stream, aw, 0, c.use_lds);
    } else {
        hipLaunchKernelGGL((jk_lane_kernel<LI, LJ, LK, LL, PM>), dim3(nblocks),
                           dim3(LANE_WAVES * WAVE), c.bytes, stream, a, kl_chunk, c.use_lds);
    }
    return hipGetLastError();
}

template <int LI, int LJ, int LK, int LL>
static hipError_t launch_class(const JKBuildArgs &a, hipStream_t stream)
{
    if constexpr (ClassTraits<LI, LJ, LK, LL>::LANE) {
        // whole kl lists per work item unless that leaves the GPU idle
        int nchunks = (LANE_ITEMS + a.npairs_ij - 1) / a.npairs_ij;
        int maxchunks = (a.npairs_kl + LANE_MIN_CHUNK - 1) / LANE_MIN_CHUNK;
        if (nchunks > maxchunks) nchunks = maxchunks;
        if (nchunks < 1) nchunks = 1;
        const int kl_chunk = (a.npairs_kl + nchunks - 1) / nchunks;
        nchunks = (a.npairs_kl + kl_chunk - 1) / kl_chunk;
        long nblocks = (long)a.npairs_ij * nchunks;
        if (nblocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
        // mixed precision: one kernel with both queues; FP64 only: the kernel
        // compiled without the FP32 arithmetic
        if (a.fp32_cut > -1e29f)
            return launch_lane<LI, LJ, LK, LL, 0>(a, stream, nblocks, kl_chunk);
        return launch_lane<LI, LJ, LK, LL, 1>(a, stream, nblocks, kl_chunk);
    } else {
        using S = TeamShape<LI, LJ, LK, LL>;
        const int kl_chunk = TEAM_KL_CHUNK;
        int nchunks = (a.npairs_kl + kl_chunk - 1) / kl_chunk;
        long nblocks = (long)a.npairs_ij * nchunks;
        if (nblocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
        // static LDS of the kernel: queue, wave counts, stop flag, team buffers
        size_t lds_static = sizeof(uint32_t) * 2 * S::TB + sizeof(int) * (S::TB / WAVE + 2) +
                            sizeof(double) * (S::NQ * (S::PER_QB + S::DSZ) + S::NFI * S::NFJ) +
                            256;
        const JKWindows &tw = a.kl_win[1];
        if (tw.win && a.vk && a.n_dm == 1 && (TEAM_KWIN || a.kl_win_policy == 1)) {
            // K rows over one kl column window per work item, when that
            // leaves enough work items to fill the GPU
            const size_t bytes = sizeof(double) * (S::NFI + S::NFJ) * (size_t)tw.width;
            const long nblocks_w = (long)a.npairs_ij * tw.n_win;
            if (bytes + lds_static <= lds_limit() && nblocks_w <= 0x7fffffffL &&
                (a.kl_win_policy == 1 || nblocks_w >= TEAM_KWIN_MIN_ITEMS)) {
                JKBuildArgs aw = a;
                aw.kl_win[0] = tw;
                aw.pairs_kl = tw.pairs;
                aw.q_kl = tw.q;
                aw.pdata_kl = tw.pdata;
                hipLaunchKernelGGL((jk_team_kernel<LI, LJ, LK, LL>), dim3(nblocks_w), dim3(S::TB),
                                   bytes, stream, aw, 0, 2 | KR_WIN);
                return hipGetLastError();
            }
        }
        size_t lds = 0;
        const int use_lds = krows_mode<LI, LJ>(a, a.k_nao, a.l_nao, lds_static, TEAM_KROWS,
                                               false, lds);
        hipLaunchKernelGGL((jk_team_kernel<LI, LJ, LK, LL>), dim3(nblocks), dim3(S::TB),
                           lds, stream, a, kl_chunk, use_lds);
    }
    return hipGetLastError();
}
