// Generic Rys kernels: the angular momenta are runtime values, so one kernel
// covers every class up to l = GEN_LMAX (g and h shells) and the first
// derivatives of the ERIs.  They serve the classes the templated per-lane and
// team kernels do not instantiate (any shell with l >= 4) and the per-atom
// energy gradient (no counterpart among the templated kernels).
//
// One wavefront per work item (one ij pair, a chunk of its kl list).  The
// wave scans 64 kl candidates at a time, applies the survival predicate and
// compacts the survivors into an LDS queue; each queued quartet is then
// evaluated by the whole wave:
//   * lanes r < N evaluate root r of every primitive quartet (Chebyshev
//     tables, long-range attenuation applied as for the templated kernels);
//   * lanes (r, d) < 3 N build the 2D table of root r and direction d in LDS,
//     in place: VRR in i and TRR in k at g[a + c SK], then the bra HRR
//     (stride SJ) and the ket HRR (stride SL), so that a table needs exactly
//     (LI+1)(LJ+1)(LK+1)(LL+1) doubles;
//   * every lane owns GEN_CH elements of the Cartesian block per pass and
//     accumulates their triple products (JK) or their derivative integrals
//     contracted with the two-particle density (gradient); a block larger
//     than 64 GEN_CH elements takes several passes over the primitives.
// Digestion (JK): every element adds into the six J/K blocks with FP64
// atomics (the symmetry factors of rys_contract_jk.cu:156-158,214-220).
// Gradient: the 3 x 3 derivative sums of centres i, j, k are reduced across
// the wave and added to the atoms of i, j, k; centre l gets minus their sum
// (translational invariance).
#include "jk_class_kernels.h"

#define GEN_LMAX        5
#define GEN_CH          8
#define GEN_KL_CHUNK    256
// LDS for the 2D tables of one chunk of roots (3 tables per root)
#define GEN_TAB_BYTES   (52 * 1024)
#define GEN_TAB_DBL     (GEN_TAB_BYTES / 8)

enum { GEN_JK = 0, GEN_IP1 = 1 };

struct GenExtra {
    double *ejk;        // natm x 3 (gradient mode)
    double jf, kf;      // factors of the J and K parts of the two-particle density
};

// Cartesian powers of component c of a shell of angular momentum l
// (order lx = l..0, then ly = l - lx..0).
__device__ __forceinline__ void cart_powers(int l, int c, int &x, int &y, int &z)
{
    int lx = l, start = 0;
    while (start + (l - lx + 1) <= c) { start += l - lx + 1; --lx; }
    x = lx;
    y = (l - lx) - (c - start);
    z = l - x - y;
}

// Root i of the N-point Rys rule at x (Chebyshev tables or the large-x
// limit); with omega2 > 0 the rule of the attenuated operator erf(omega r)/r:
// roots and weights of x' = x a, a = omega^2 / (omega^2 + theta), scaled by
// a and sqrt(a).
__device__ void gen_root(const JKRysTable &tab, int N, double x, double theta, double omega2,
                         int i, double &r, double &w)
{
    double sa = 1., aw = 1.;
    if (omega2 > 0.) {
        aw = omega2 / (omega2 + theta);
        sa = sqrt(aw);
        x *= aw;
    }
    if (x >= tab.xmax[N-1]) {
        const double ix = 1. / x;
        r = tab.coef[tab.large_r + N * (N - 1) / 2 + i] * ix;
        w = tab.coef[tab.large_w + N * (N - 1) / 2 + i] * sqrt(ix);
    } else {
        const int D1 = tab.degree + 1;
        const int k = (int)x;
        const double u = 2. * (x - k) - 1., u2 = 2. * u;
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
    r *= aw;
    w *= sa;
}

// 2D integrals of one root and direction, in place in g (LDS):
// g[i + j SJ + k SK + l SL] with SJ = LI+1, SK = SJ (LJ+1), SL = SK (LK+1).
__device__ void gen_build_g(double *g, int LI, int LJ, int LK, int LL, int SJ, int SK, int SL,
                            double g00, double c0, double cp, double b10, double b01,
                            double b00, double rij, double rkl)
{
    const int LIJ = LI + LJ, LKL = LK + LL;
    // VRR in a (c = 0), then TRR in c: G(a, c) at g[a + c SK]
    g[0] = g00;
    if (LIJ > 0) g[1] = c0 * g00;
    for (int a = 1; a < LIJ; ++a) g[a + 1] = c0 * g[a] + (a * b10) * g[a - 1];
    for (int c = 0; c < LKL; ++c) {
        const double *gc = g + c * SK;
        double *gn = g + (c + 1) * SK;
        for (int a = 0; a <= LIJ; ++a) {
            double v = cp * gc[a];
            if (c > 0) v += (c * b01) * gc[a - SK];
            if (a > 0) v += (a * b00) * gc[a - 1];
            gn[a] = v;
        }
    }
    // bra HRR: row j at h[j SJ + a], a <= LIJ - j; descending a keeps every
    // value of row j readable until row j+1 is complete
    for (int c = 0; c <= LKL; ++c) {
        double *h = g + c * SK;
        for (int j = 0; j < LJ; ++j)
            for (int a = LIJ - j - 1; a >= 0; --a)
                h[(j + 1) * SJ + a] = h[j * SJ + a + 1] - rij * h[j * SJ + a];
    }
    // ket HRR: the same in units of SK (c) and SL (l)
    for (int ij = 0; ij < SK; ++ij) {
        double *h = g + ij;
        for (int l = 0; l < LL; ++l)
            for (int c = LKL - l - 1; c >= 0; --c)
                h[(l + 1) * SL + c * SK] = h[l * SL + (c + 1) * SK] - rkl * h[l * SL + c * SK];
    }
}

// Survival predicate of the energy-gradient contraction
// (gpu4pyscf/lib/gvhf-rys/create_tasks.cu:858-877), float32 in the same order.
__device__ __forceinline__ bool survives_ejk(float q_kl, float kl_cutoff, uint32_t bas_ij,
                                             uint32_t bas_kl, int ish, int jsh, int nbas,
                                             const float *dc, bool do_j, bool do_k)
{
    if (!(q_kl >= kl_cutoff && bas_ij >= bas_kl)) return false;
    const int ksh = bas_kl / nbas, lsh = bas_kl - (bas_kl / nbas) * nbas;
    const float d_cutoff = kl_cutoff - q_kl;
    return (do_k && (dc[ish*nbas+lsh] + dc[jsh*nbas+ksh] > d_cutoff ||
                     dc[ish*nbas+ksh] + dc[jsh*nbas+lsh] > d_cutoff)) ||
           (do_j && dc[bas_ij] + dc[bas_kl] > d_cutoff);
}

template <int MODE>
__global__ __launch_bounds__(WAVE)
void jk_generic_kernel(JKBuildArgs a, int kl_chunk, GenExtra ex)
{
    __shared__ uint32_t queue[2 * WAVE];
    __shared__ double rw[2 * JK_RYS_NMAX];
    __shared__ double tab[GEN_TAB_DBL];
    const int lane = threadIdx.x;
    const int nbas = a.nbas, nao = a.nao;
    const size_t nao2 = (size_t)nao * nao;
    select_replica(a);
    const WorkItem wi = work_item(a.npairs_kl, kl_chunk);
    const uint32_t bas_ij = a.pairs_ij[wi.pij];
    const float q_ij = a.q_ij[wi.pij];
    const float kl_cutoff = a.cutoff - q_ij;
    const float penalty = MODE == GEN_JK ? a.dm_penalty : 0.f;
    if (a.q_kl[wi.kl0] + penalty + Q_COND_MARGIN < kl_cutoff) return;
    ShellPair sp;
    load_shell_pair(sp, bas_ij, nbas, a.bas, a.env, a.ao_loc);
    const int li = a.li, lj = a.lj, lk = a.lk, ll = a.ll;
    const int nfi = ncart(li), nfj = ncart(lj), nfk = ncart(lk), nfl = ncart(ll);
    const int nf = nfi * nfj * nfk * nfl;
    // table shape: one more unit of angular momentum on i, j, k for derivatives
    const int dd = MODE == GEN_IP1 ? 1 : 0;
    const int LI = li + dd, LJ = lj + dd, LK = lk + dd, LL = ll;
    const int SJ = LI + 1, SK = SJ * (LJ + 1), SL = SK * (LK + 1);
    const int GS = SL * (LL + 1);
    const int N = (li + lj + lk + ll + dd) / 2 + 1;
    int RC = GEN_TAB_DBL / (3 * GS);
    RC = RC < N ? RC : N;
    const double omega2 = a.omega * a.omega;
    const bool do_j = MODE == GEN_IP1 ? ex.jf != 0. : a.vj != nullptr;
    const bool do_k = MODE == GEN_IP1 ? ex.kf != 0. : a.vk != nullptr;
    const int nij = a.iprim * a.jprim, nkl = a.kprim * a.lprim;
    const double *pdij = a.pdata_ij + (size_t)wi.pij * nij * PREC;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int qn = 0;
    unsigned long long nsurv = 0, hsum = 0;
    bool stop = false;
    for (int base = wi.kl0; base < wi.kl1 || qn > 0; base += WAVE) {
        if (base < wi.kl1 && !stop) {
            const int pkl = base + lane;
            bool keep = false, fail_margin = false;
            if (pkl < wi.kl1) {
                const float q_kl = a.q_kl[pkl];
                const uint32_t bkl = a.pairs_kl[pkl];
                fail_margin = q_kl + penalty + Q_COND_MARGIN < kl_cutoff;
                if (MODE == GEN_JK)
                    keep = survives(q_kl, kl_cutoff, a.dm_penalty, bas_ij, bkl, sp.ish, sp.jsh,
                                    nbas, a.dm_cond, a.screen, a.swap_ijkl != 0,
                                    a.q_ij[wi.pij], a.cutoff);
                else
                    keep = survives_ejk(q_kl, kl_cutoff, bas_ij, bkl, sp.ish, sp.jsh, nbas,
                                        a.dm_cond, do_j, do_k);
                if (a.quartet_hash && keep) hsum += task_mix(a, bas_ij, bkl);
            }
            const uint64_t m = __ballot(keep);
            if (keep) queue[qn + __popcll(m & lt_mask)] = pkl;
            qn += __popcll(m);
            nsurv += __popcll(m);
            stop = __ballot(fail_margin) != 0;
            wave_sync();
        }
        const bool flush = base + WAVE >= wi.kl1 || stop;
        if (!(qn >= WAVE || (flush && qn > 0))) {
            if (flush) break;
            continue;
        }
        // evaluate every queued quartet, one at a time, with the whole wave
        for (int q = 0; q < qn; ++q) {
            const int pkl = queue[q];
            const uint32_t bas_kl = a.pairs_kl[pkl];
            const int ksh = bas_kl / nbas, lsh = bas_kl - (bas_kl / nbas) * nbas;
            double fac_sym = a.scale;
            if (sp.ish == sp.jsh) fac_sym *= .5;
            if (ksh == lsh) fac_sym *= .5;
            if (bas_ij == bas_kl) fac_sym *= .5;
            const double *rk = a.env + a.bas[ksh * 8 + 7], *rl = a.env + a.bas[lsh * 8 + 7];
            const double rkl[3] = {rl[0] - rk[0], rl[1] - rk[1], rl[2] - rk[2]};
            const int i0 = sp.i0, j0 = sp.j0, k0 = a.ao_loc[ksh], l0 = a.ao_loc[lsh];
            const double *pdkl = a.pdata_kl + (size_t)pkl * nkl * PREC;
            const double *expi = a.env + sp.pexpi, *expj = a.env + sp.pexpj;
            const double *expk = a.env + a.bas[ksh * 8 + 5];
            double gsum[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) gsum[t] = 0.;
            for (int e0 = 0; e0 < nf; e0 += WAVE * GEN_CH) {
                // this lane's elements of the pass: offsets into the tables,
                // Cartesian powers (gradient) and accumulators
                int ox[GEN_CH], oy[GEN_CH], oz[GEN_CH], pw[GEN_CH];
                double acc[GEN_CH];
#pragma unroll
                for (int c = 0; c < GEN_CH; ++c) {
                    const int n = e0 + c * WAVE + lane;
                    acc[c] = 0.;
                    ox[c] = oy[c] = oz[c] = pw[c] = 0;
                    if (n >= nf) continue;
                    const int l = n % nfl, k = (n / nfl) % nfk;
                    const int j = (n / (nfl * nfk)) % nfj, i = n / (nfl * nfk * nfj);
                    int ix, iy, iz, jx, jy, jz, kx, ky, kz, lx, ly, lz;
                    cart_powers(li, i, ix, iy, iz);
                    cart_powers(lj, j, jx, jy, jz);
                    cart_powers(lk, k, kx, ky, kz);
                    cart_powers(ll, l, lx, ly, lz);
                    ox[c] = ix + jx * SJ + kx * SK + lx * SL;
                    oy[c] = iy + jy * SJ + ky * SK + ly * SL;
                    oz[c] = iz + jz * SJ + kz * SK + lz * SL;
                    // powers packed 4 bits each: ix iy iz jx jy jz kx ky kz
                    pw[c] = ix | iy << 4 | iz << 8 | jx << 12 | jy << 16 | jz << 20 |
                            kx << 24 | ky << 28;
                    if (MODE == GEN_IP1) {
                        // two-particle density of the element (rys_contract_jk_ip1.cu:523-577)
                        const size_t I = i0 + i, J = j0 + j, K = k0 + k, L = l0 + l;
                        const double *D = a.dm;
                        double g2 = 0.;
                        for (int id = 0; id < a.n_dm; ++id, D += nao2) {
                            if (do_k) g2 += ex.kf * (D[J*nao+K] * D[L*nao+I] + D[J*nao+L] * D[K*nao+I]);
                        }
                        if (do_j) {
                            double dji = 0., dlk = 0.;
                            D = a.dm;
                            for (int id = 0; id < a.n_dm; ++id, D += nao2) {
                                dji += D[J*nao+I]; dlk += D[L*nao+K];
                            }
                            g2 += ex.jf * dji * dlk;
                        }
                        acc[c] = g2 * fac_sym;      // the element's weight
                        // kz is not packed: recovered from lk - kx - ky
                    }
                }
                for (int kp = 0; kp < nkl; ++kp) {
                    const double *qr = pdkl + kp * PREC;
                    const double akl = qr[0], rakl = qr[1];
                    const double QC[3] = {qr[2] - rk[0], qr[3] - rk[1], qr[4] - rk[2]};
                    const double Kkl = qr[5] * (MODE == GEN_JK ? fac_sym : 1.);
                    const double ak2 = 2. * expk[kp / a.lprim];
                    for (int ip = 0; ip < nij; ++ip) {
                        const double *pr = pdij + ip * PREC;
                        const double aij = pr[0], raij = pr[1];
                        const double ai2 = 2. * expi[ip / a.jprim], aj2 = 2. * expj[ip % a.jprim];
                        double PA[3], PQ[3];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            PA[d] = pr[2+d] - sp.ri[d];
                            PQ[d] = pr[2+d] - qr[2+d];
                        }
                        const double rr = PQ[0]*PQ[0] + PQ[1]*PQ[1] + PQ[2]*PQ[2];
                        const double r_aa = 1. / (aij + akl);
                        const double theta = aij * akl * r_aa;
                        const double fac = pr[5] * Kkl * sqrt(r_aa);
                        for (int r0 = 0; r0 < N; r0 += RC) {
                            const int nr = N - r0 < RC ? N - r0 : RC;
                            wave_sync();
                            if (lane < nr) {
                                double rt, wt;
                                gen_root(a.rys, N, theta * rr, theta, omega2, r0 + lane, rt, wt);
                                rw[2 * lane] = rt; rw[2 * lane + 1] = wt;
                            }
                            wave_sync();
                            if (lane < 3 * nr) {
                                const int r = lane / 3, d = lane - (lane / 3) * 3;
                                const double rt_aa = rw[2 * r] * r_aa;
                                const double rt_aij = rt_aa * akl, rt_akl = rt_aa * aij;
                                const double b10 = .5 * raij * (1. - rt_aij);
                                const double b01 = .5 * rakl * (1. - rt_akl);
                                const double b00 = .5 * rt_aa;
                                const double g00 = d == 2 ? rw[2 * r + 1] * fac : 1.;
                                gen_build_g(tab + (r * 3 + d) * GS, LI, LJ, LK, LL, SJ, SK, SL, g00,
                                            PA[d] - rt_aij * PQ[d], QC[d] + rt_akl * PQ[d],
                                            b10, b01, b00, sp.rij[d], rkl[d]);
                            }
                            wave_sync();
                            for (int r = 0; r < nr; ++r) {
                                const double *gx = tab + (r * 3) * GS;
                                const double *gy = gx + GS, *gz = gy + GS;
#pragma unroll
                                for (int c = 0; c < GEN_CH; ++c) {
                                    if (e0 + c * WAVE + lane >= nf) continue;
                                    const double x = gx[ox[c]], y = gy[oy[c]], z = gz[oz[c]];
                                    if (MODE == GEN_JK) {
                                        acc[c] = fma(x * y, z, acc[c]);
                                    } else {
                                        const int p = pw[c];
                                        const int ix = p & 15, iy = p >> 4 & 15, iz = p >> 8 & 15;
                                        const int jx = p >> 12 & 15, jy = p >> 16 & 15, jz = p >> 20 & 15;
                                        const int kx = p >> 24 & 15, ky = p >> 28 & 15;
                                        const int kz = lk - kx - ky;
                                        const double w = acc[c];
                                        // d/dA: 2 ai g(i+1) - i g(i-1), likewise B (j), C (k)
                                        const double dxi = ai2 * gx[ox[c] + 1] - (ix ? ix * gx[ox[c] - 1] : 0.);
                                        const double dyi = ai2 * gy[oy[c] + 1] - (iy ? iy * gy[oy[c] - 1] : 0.);
                                        const double dzi = ai2 * gz[oz[c] + 1] - (iz ? iz * gz[oz[c] - 1] : 0.);
                                        const double dxj = aj2 * gx[ox[c] + SJ] - (jx ? jx * gx[ox[c] - SJ] : 0.);
                                        const double dyj = aj2 * gy[oy[c] + SJ] - (jy ? jy * gy[oy[c] - SJ] : 0.);
                                        const double dzj = aj2 * gz[oz[c] + SJ] - (jz ? jz * gz[oz[c] - SJ] : 0.);
                                        const double dxk = ak2 * gx[ox[c] + SK] - (kx ? kx * gx[ox[c] - SK] : 0.);
                                        const double dyk = ak2 * gy[oy[c] + SK] - (ky ? ky * gy[oy[c] - SK] : 0.);
                                        const double dzk = ak2 * gz[oz[c] + SK] - (kz ? kz * gz[oz[c] - SK] : 0.);
                                        const double yz = y * z, xz = x * z, xy = x * y;
                                        gsum[0] = fma(w * dxi, yz, gsum[0]);
                                        gsum[1] = fma(w * dyi, xz, gsum[1]);
                                        gsum[2] = fma(w * dzi, xy, gsum[2]);
                                        gsum[3] = fma(w * dxj, yz, gsum[3]);
                                        gsum[4] = fma(w * dyj, xz, gsum[4]);
                                        gsum[5] = fma(w * dzj, xy, gsum[5]);
                                        gsum[6] = fma(w * dxk, yz, gsum[6]);
                                        gsum[7] = fma(w * dyk, xz, gsum[7]);
                                        gsum[8] = fma(w * dzk, xy, gsum[8]);
                                    }
                                }
                            }
                        }
                    }
                }
                if (MODE == GEN_JK) {
                    // digestion of this pass's elements
#pragma unroll
                    for (int c = 0; c < GEN_CH; ++c) {
                        const int n = e0 + c * WAVE + lane;
                        if (n >= nf) continue;
                        const int l = n % nfl, k = (n / nfl) % nfk;
                        const int j = (n / (nfl * nfk)) % nfj, i = n / (nfl * nfk * nfj);
                        const size_t I = i0 + i, J = j0 + j, K = k0 + k, L = l0 + l;
                        const double v = acc[c];
                        for (int id = 0; id < a.n_dm; ++id) {
                            const double *D = a.dm + id * nao2;
                            if (a.vj) {
                                double *Jm = a.vj + id * nao2;
                                atomic_add(Jm + K * nao + L, v * D[I * nao + J]);
                                atomic_add(Jm + I * nao + J, v * D[K * nao + L]);
                            }
                            if (a.vk) {
                                double *Km = a.vk + id * nao2;
                                atomic_add(Km + I * nao + K, v * D[J * nao + L]);
                                atomic_add(Km + I * nao + L, v * D[J * nao + K]);
                                atomic_add(Km + J * nao + K, v * D[I * nao + L]);
                                atomic_add(Km + J * nao + L, v * D[I * nao + K]);
                            }
                        }
                    }
                }
            }
            if (MODE == GEN_IP1) {
#pragma unroll
                for (int t = 0; t < 9; ++t) gsum[t] = wave_sum(gsum[t]);
                if (lane < 3) {
                    const int ia = a.bas[sp.ish * 8], ja = a.bas[sp.jsh * 8];
                    const int ka = a.bas[ksh * 8], la = a.bas[lsh * 8];
                    const double gi = lane == 0 ? gsum[0] : (lane == 1 ? gsum[1] : gsum[2]);
                    const double gj = lane == 0 ? gsum[3] : (lane == 1 ? gsum[4] : gsum[5]);
                    const double gk = lane == 0 ? gsum[6] : (lane == 1 ? gsum[7] : gsum[8]);
                    atomic_add(ex.ejk + ia * 3 + lane, gi);
                    atomic_add(ex.ejk + ja * 3 + lane, gj);
                    atomic_add(ex.ejk + ka * 3 + lane, gk);
                    atomic_add(ex.ejk + la * 3 + lane, -(gi + gj + gk));
                }
            }
        }
        wave_sync();
        qn = 0;
        if (flush) break;
    }
    if (lane == 0 && a.n_quartets && nsurv) atomicAdd(a.n_quartets, nsurv);
    if (a.quartet_hash && hsum) atomicAdd(a.quartet_hash, hsum);
}

static hipError_t launch_generic(const JKBuildArgs &a, int mode, const GenExtra &ex,
                                 hipStream_t stream)
{
    const int kl_chunk = GEN_KL_CHUNK;
    const int nchunks = (a.npairs_kl + kl_chunk - 1) / kl_chunk;
    const long nblocks = (long)a.npairs_ij * nchunks;
    if (nblocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    if (mode == GEN_JK)
        hipLaunchKernelGGL(jk_generic_kernel<GEN_JK>, dim3(nblocks), dim3(WAVE), 0, stream,
                           a, kl_chunk, ex);
    else
        hipLaunchKernelGGL(jk_generic_kernel<GEN_IP1>, dim3(nblocks), dim3(WAVE), 0, stream,
                           a, kl_chunk, ex);
    return hipGetLastError();
}

// Whether one root's three tables fit the LDS table buffer.
static bool generic_fits(int li, int lj, int lk, int ll, int dd)
{
    const long gs = (long)(li + dd + 1) * (lj + dd + 1) * (lk + dd + 1) * (ll + 1);
    return 3 * gs <= GEN_TAB_DBL;
}

hipError_t jk_dispatch_generic(const JKBuildArgs &a, hipStream_t stream)
{
    if (a.li > GEN_LMAX || a.lj > GEN_LMAX || a.lk > GEN_LMAX || a.ll > GEN_LMAX ||
        !generic_fits(a.li, a.lj, a.lk, a.ll, 0))
        return hipErrorNotSupported;
    GenExtra ex = {nullptr, 0., 0.};
    return launch_generic(a, GEN_JK, ex, stream);
}

hipError_t jk_dispatch_generic_ip1(const JKBuildArgs &a, double *ejk, double jf, double kf,
                                   hipStream_t stream)
{
    if (a.li > GEN_LMAX || a.lj > GEN_LMAX || a.lk > GEN_LMAX || a.ll > GEN_LMAX ||
        !generic_fits(a.li, a.lj, a.lk, a.ll, 1))
        return hipErrorNotSupported;
    GenExtra ex = {ejk, jf, kf};
    return launch_generic(a, GEN_IP1, ex, stream);
}
