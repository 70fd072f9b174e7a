// Setup kernels (overlap mask, primitive-pair records, q_cond, dm_cond) and
// the C-ABI of libjkhip.so.  The class kernels are in jk_class_kernels.h,
// instantiated by jk_classes.hip.
#include <stdarg.h>
#include <mutex>
#include "jk_class_kernels.h"

static thread_local char g_err[512];

static void set_err(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void jk_set_error(const char *msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
}

// class dispatch, one function per part of the class list (jk_classes.hip)
hipError_t jk_dispatch_part0(int key, const JKBuildArgs &a, hipStream_t stream);
hipError_t jk_dispatch_part1(int key, const JKBuildArgs &a, hipStream_t stream);
hipError_t jk_dispatch_part2(int key, const JKBuildArgs &a, hipStream_t stream);
hipError_t jk_dispatch_part3(int key, const JKBuildArgs &a, hipStream_t stream);
hipError_t jk_dispatch_part4(int key, const JKBuildArgs &a, hipStream_t stream);

// classes with a shell of l >= 4: the runtime-l kernels of jk_generic.hip
hipError_t jk_dispatch_generic(const JKBuildArgs &a, hipStream_t stream);
hipError_t jk_dispatch_generic_ip1(const JKBuildArgs &a, double *ejk, double jf, double kf,
                                   hipStream_t stream);
#define JK_TEMPLATED_LMAX 3

static hipError_t dispatch(const JKBuildArgs &a, hipStream_t stream)
{
    if (a.li > JK_TEMPLATED_LMAX || a.lj > JK_TEMPLATED_LMAX || a.lk > JK_TEMPLATED_LMAX ||
        a.ll > JK_TEMPLATED_LMAX)
        return jk_dispatch_generic(a, stream);
    const int key = a.li * 125 + a.lj * 25 + a.lk * 5 + a.ll;
    hipError_t (*const parts[])(int, const JKBuildArgs &, hipStream_t) = {
        jk_dispatch_part0, jk_dispatch_part1, jk_dispatch_part2, jk_dispatch_part3,
        jk_dispatch_part4};
    for (auto f : parts) {
        hipError_t e = f(key, a, stream);
        if (e != hipErrorNotSupported) return e;
    }
    return hipErrorNotSupported;
}

// ---------------------------------------------------------------- setup kernels
__global__ void ovlp_mask_kernel(int8_t *mask, const float *exps, const float *log_coeff,
                                 const double *coords, const int32_t *ls, int nbas,
                                 float log_cutoff)
{
    int jsh = blockIdx.x * blockDim.x + threadIdx.x;
    int ish = blockIdx.y * blockDim.y + threadIdx.y;
    if (ish >= nbas || jsh >= nbas) return;
    int8_t keep = 0;
    if (ish >= jsh) {
        // lib/pbc/overlap.cu:2163-2223, float32 in the same order
        float xi = coords[ish*3+0], yi = coords[ish*3+1], zi = coords[ish*3+2];
        float xj = coords[jsh*3+0], yj = coords[jsh*3+1], zj = coords[jsh*3+2];
        float ai = exps[ish], aj = exps[jsh];
        float aij = ai + aj;
        float fi = ai / aij, fj = aj / aij;
        float theta = ai * fj;
        float dx = xj - xi, dy = yj - yi, dz = zj - zi;
        float rr = dx * dx + dy * dy + dz * dz;
        float li = ls[ish], lj = ls[jsh];
        if (!(theta * rr > 50.f)) {
            float fac_norm = log_coeff[ish] + log_coeff[jsh] + 1.717f - 1.5f * logf(aij);
            float dr = sqrtf(rr);
            float dri = fj * dr, drj = fi * dr;
            float dri_fac = .5f * li * logf(.5f * li / aij + dri * dri + 1e-9f);
            float drj_fac = .5f * lj * logf(.5f * lj / aij + drj * drj + 1e-9f);
            float log_ovlp = fac_norm - theta * rr + dri_fac + drj_fac;
            keep = log_ovlp > log_cutoff;
        }
    }
    mask[(size_t)ish * nbas + jsh] = keep;
}

__global__ void pair_data_kernel(double *out, const uint32_t *pairs, int npairs, int iprim,
                                 int jprim, const int32_t *bas, const double *env, int nbas)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int np = iprim * jprim;
    if (t >= npairs * np) return;
    int p = t / np, ipjp = t - p * np;
    int ip = ipjp / jprim, jp = ipjp - ip * jprim;
    uint32_t bij = pairs[p];
    int ish = bij / nbas, jsh = bij - (bij / nbas) * nbas;
    const int32_t *bi = bas + ish * 8, *bj = bas + jsh * 8;
    const double *ri = env + bi[7], *rj = env + bj[7];
    double ai = env[bi[5] + ip], aj = env[bj[5] + jp];
    double a = ai + aj, ra = 1. / a;
    double rr = 0.;
    double *o = out + (size_t)t * 6;
    for (int d = 0; d < 3; ++d) {
        double dr = rj[d] - ri[d];
        rr += dr * dr;
        o[2 + d] = ri[d] + dr * aj * ra;
    }
    o[0] = a;
    o[1] = ra;
    o[5] = sqrt(PI_FAC) * env[bi[6] + ip] * env[bj[6] + jp] * exp(-ai * aj * ra * rr) * ra;
}

// max over the Cartesian diagonal (ab|ab) of a shell pair; by the
// Cauchy-Schwarz inequality this is max |(ij|ij)| over the whole block.
// Operator: sr / r + (lr - sr) erf(omega r)/r (omega = 0: lr / r); the
// long-range part uses the attenuated Rys rule of gen_root.  Shells up to
// l = 5.
#define QC_LMAX 5
__global__ void pair_qcond_kernel(float *q, const uint32_t *pairs, int npairs,
                                  const int32_t *bas, const double *env, int nbas,
                                  JKRysTable rys, double omega, double lr, double sr)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    uint32_t bij = pairs[p];
    int ish = bij / nbas, jsh = bij - (bij / nbas) * nbas;
    const int32_t *bi = bas + ish * 8, *bj = bas + jsh * 8;
    int li = bi[1], lj = bj[1];
    int lij = li + lj, n = lij + 1;
    int nfi = ncart(li), nfj = ncart(lj);
    const double *ri = env + bi[7], *rj = env + bj[7];
    double rij[3] = {rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]};
    double rr_ij = rij[0]*rij[0] + rij[1]*rij[1] + rij[2]*rij[2];
    constexpr int NC = ncart(QC_LMAX), L2 = 2 * QC_LMAX + 1;
    double out[NC * NC];
    for (int t = 0; t < nfi * nfj; ++t) out[t] = 0.;
    int px[NC], py[NC], pz[NC], qx[NC], qy[NC], qz[NC];
    {
        int c = 0;
        for (int x = li; x >= 0; --x) for (int y = li - x; y >= 0; --y) { px[c] = x; py[c] = y; pz[c] = li-x-y; ++c; }
        c = 0;
        for (int x = lj; x >= 0; --x) for (int y = lj - x; y >= 0; --y) { qx[c] = x; qy[c] = y; qz[c] = lj-x-y; ++c; }
    }
    // (operator omega, weight) of the one or two parts
    const double w_full = omega > 0. ? sr : lr;
    const double w_lr = omega > 0. ? lr - sr : 0.;
    for (int ip = 0; ip < bi[2]; ++ip)
    for (int jp = 0; jp < bj[2]; ++jp)
    for (int kp = 0; kp < bi[2]; ++kp)
    for (int lp = 0; lp < bj[2]; ++lp) {
        double ai = env[bi[5]+ip], aj = env[bj[5]+jp];
        double ak = env[bi[5]+kp], al = env[bj[5]+lp];
        double aij = ai + aj, akl = ak + al;
        double cicj = env[bi[6]+ip] * env[bj[6]+jp] * exp(-ai * aj / aij * rr_ij);
        double ckcl = env[bi[6]+kp] * env[bj[6]+lp] * exp(-ak * al / akl * rr_ij);
        double PQ[3], PA[3], QC[3];
        for (int d = 0; d < 3; ++d) {
            PA[d] = rij[d] * aj / aij;
            QC[d] = rij[d] * al / akl;
            PQ[d] = PA[d] - QC[d];
        }
        double rr = PQ[0]*PQ[0] + PQ[1]*PQ[1] + PQ[2]*PQ[2];
        double aijkl = aij + akl;
        double theta = aij * akl / aijkl;
        double fac = PI_FAC * cicj * ckcl / (aij * akl * sqrt(aijkl));
        for (int part = 0; part < 2; ++part) {
            const double wpart = part == 0 ? w_full : w_lr;
            if (wpart == 0.) continue;
            double rt[JK_RYS_NMAX], wt[JK_RYS_NMAX];
            if (part == 0) {
                rys_roots_rt(rys, n, theta * rr, rt, wt);
            } else {
                const double aw = omega * omega / (omega * omega + theta);
                rys_roots_rt(rys, n, theta * rr * aw, rt, wt);
                for (int ir = 0; ir < n; ++ir) { rt[ir] *= aw; wt[ir] *= sqrt(aw); }
            }
            for (int ir = 0; ir < n; ++ir) {
                double rt_aa = rt[ir] / aijkl;
                double rt_aij = rt_aa * akl, rt_akl = rt_aa * aij;
                double b10 = .5 / aij * (1. - rt_aij);
                double b01 = .5 / akl * (1. - rt_akl);
                double b00 = .5 * rt_aa;
                double gd[3][QC_LMAX + 1][QC_LMAX + 1];
                for (int d = 0; d < 3; ++d) {
                    double G[L2][L2];
                    double c0 = PA[d] - rt_aij * PQ[d];
                    double cp = QC[d] + rt_akl * PQ[d];
                    G[0][0] = d == 2 ? wt[ir] * fac * wpart : 1.;
                    if (lij > 0) G[1][0] = c0 * G[0][0];
                    for (int a2 = 1; a2 < lij; ++a2) G[a2+1][0] = c0 * G[a2][0] + a2 * b10 * G[a2-1][0];
                    for (int a2 = 0; a2 <= lij; ++a2) {
                        if (lij > 0) {
                            double v = cp * G[a2][0];
                            if (a2 > 0) v += a2 * b00 * G[a2-1][0];
                            G[a2][1] = v;
                        }
                        for (int c = 1; c < lij; ++c) {
                            double v = cp * G[a2][c] + c * b01 * G[a2][c-1];
                            if (a2 > 0) v += a2 * b00 * G[a2-1][c];
                            G[a2][c+1] = v;
                        }
                    }
                    // bra HRR for all (a, b, c); then ket HRR only for (c, e) = (a, b)
                    double H[L2][QC_LMAX + 1][L2];
                    for (int c = 0; c <= lij; ++c) {
                        for (int a2 = 0; a2 <= lij; ++a2) H[a2][0][c] = G[a2][c];
                        for (int b = 0; b < lj; ++b)
                            for (int a2 = 0; a2 < lij - b; ++a2)
                                H[a2][b+1][c] = H[a2+1][b][c] - rij[d] * H[a2][b][c];
                    }
                    for (int a2 = 0; a2 <= li; ++a2)
                    for (int b = 0; b <= lj; ++b) {
                        double T[L2][QC_LMAX + 1];
                        for (int c = 0; c <= lij; ++c) T[c][0] = H[a2][b][c];
                        for (int e = 0; e < lj; ++e)
                            for (int c = 0; c < lij - e; ++c)
                                T[c][e+1] = T[c+1][e] - rij[d] * T[c][e];
                        gd[d][a2][b] = T[a2][b];
                    }
                }
                for (int i = 0; i < nfi; ++i)
                for (int j = 0; j < nfj; ++j)
                    out[i*nfj+j] += gd[0][px[i]][qx[j]] * gd[1][py[i]][qy[j]] * gd[2][pz[i]][qz[j]];
            }
        }
    }
    double m = 0.;
    for (int t = 0; t < nfi * nfj; ++t) m = fmax(m, fabs(out[t]));
    q[p] = m > 0 ? (float)(.5 * log(m)) : -700.f;
}

__global__ void dm_block_max_kernel(double *work, const double *dm, int n_dm, int nao,
                                    const int32_t *ao_loc, int nbas)
{
    int jsh = blockIdx.x * blockDim.x + threadIdx.x;
    int ish = blockIdx.y;
    if (jsh >= nbas) return;
    int i0 = ao_loc[ish], i1 = ao_loc[ish+1], j0 = ao_loc[jsh], j1 = ao_loc[jsh+1];
    double m = 0.;
    for (int id = 0; id < n_dm; ++id) {
        const double *D = dm + (size_t)id * nao * nao;
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j)
                m = fmax(m, fabs(D[(size_t)i * nao + j]));
    }
    work[(size_t)ish * nbas + jsh] = m;
}

__global__ void dm_cond_log_kernel(float *cond, const double *work, int nbas, int hermi)
{
    int jsh = blockIdx.x * blockDim.x + threadIdx.x;
    int ish = blockIdx.y;
    if (jsh >= nbas) return;
    double v = work[(size_t)ish * nbas + jsh];
    if (hermi == 0) v = v + work[(size_t)jsh * nbas + ish];
    cond[(size_t)ish * nbas + jsh] = (float)log(v + 1e-300);
}

// Scratch of jk_dm_cond_noalloc: the block maxima of up to
// DM_COND_SCRATCH_NBAS^2 shell pairs (32 MiB), held by every GPU the library
// runs on until the process ends.  Host code never passes `scratch` itself
// to a kernel (in host code it names a host-side stand-in); it uses the
// device address from hipGetSymbolAddress, one per device.
#define DM_COND_SCRATCH_NBAS 2048
static __device__ __align__(256) double scratch[DM_COND_SCRATCH_NBAS * DM_COND_SCRATCH_NBAS];
#define MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES];

// device address of `scratch` on the current device (looked up once per
// device); nullptr with the error set on failure
static double *scratch_ptr(const char *who)
{
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        set_err("%s: no current device (%s)", who, hipGetErrorString(err));
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr) {
            set_err("%s: hipGetSymbolAddress(scratch) on device %d: %s", who, dev,
                    hipGetErrorString(err));
            return nullptr;
        }
        g_scratch_ptr[dev] = (double *)ptr;
    }
    return g_scratch_ptr[dev];
}

// ---------------------------------------------------------------- C-ABI
static bool table_ok(const JKRysTable &t)
{
    return t.coef != nullptr && t.degree > 0 && t.degree <= RYS_DEG_MAX;
}

extern "C" {

const char *jk_last_error(void) { return g_err; }

int jk_abi_version(void) { return JK_ABI_VERSION; }

int jk_build(const JKBuildArgs *args, void *stream);

int jk_build_tasks(const JKBuildArgs *args, const JKTask *tasks, int ntasks,
                   void *const *streams, int nstreams)
{
    g_err[0] = 0;
    if (!args || (ntasks > 0 && !tasks) || ntasks < 0 || nstreams < 1 || !streams) {
        set_err("jk_build_tasks: null pointer or bad count (%d tasks)", ntasks);
        return 1;
    }
    JKBuildArgs a = *args;
    for (int m = 0; m < ntasks; ++m) {
        const int t = ntasks - 1 - m;
        const JKTask &k = tasks[t];
        a.li = k.li; a.lj = k.lj; a.lk = k.lk; a.ll = k.ll;
        a.iprim = k.iprim; a.jprim = k.jprim; a.kprim = k.kprim; a.lprim = k.lprim;
        a.k_ao0 = k.k_ao0; a.k_nao = k.k_nao; a.l_ao0 = k.l_ao0; a.l_nao = k.l_nao;
        a.npairs_ij = k.npairs_ij; a.npairs_kl = k.npairs_kl;
        a.pairs_ij = k.pairs_ij; a.pairs_kl = k.pairs_kl;
        a.q_ij = k.q_ij; a.q_kl = k.q_kl;
        a.pdata_ij = k.pdata_ij; a.pdata_kl = k.pdata_kl;
        a.kl_win[0] = k.kl_win[0];
        a.kl_win[1] = k.kl_win[1];
        a.swap_ijkl = k.swap_ijkl;
        a.n_quartets = args->n_quartets ? args->n_quartets + t : nullptr;
        a.n_quartets_fp32 = args->n_quartets_fp32 ? args->n_quartets_fp32 + t : nullptr;
        a.quartet_hash = args->quartet_hash ? args->quartet_hash + t : nullptr;
        const int st = jk_build(&a, streams[m % nstreams]);
        if (st != 0) return st;
    }
    return 0;
}

// Host-side validation shared by jk_build and jk_energy_ip1; fills a copy
// of the arguments with scale 0 taken as 1.  Returns 0 or a status code
// (message set); -1: nothing to do (an empty pair list).
static int check_args(const JKBuildArgs *args, JKBuildArgs &a, const char *who)
{
    if (!args) { set_err("%s: null args", who); return 1; }
    a = *args;
    if (a.scale == 0.) a.scale = 1.;
    if (a.abi_version != JK_ABI_VERSION) {
        set_err("%s: ABI version mismatch (%d != JK_ABI_VERSION %d)", who, a.abi_version,
                JK_ABI_VERSION);
        return 1;
    }
    if (a.li < 0 || a.li > JK_LMAX || a.lj < 0 || a.lj > a.li || a.lk < 0 ||
        a.lk > a.li || a.ll < 0 || a.ll > a.lk) {
        set_err("%s: unsupported angular class (li=%d, lj=%d, ...)", who, a.li, a.lj);
        return 2;
    }
    if (a.npairs_ij <= 0 || a.npairs_kl <= 0) return -1;
    if (!a.bas || !a.env || !a.ao_loc || !a.pairs_ij || !a.pairs_kl || !a.q_ij ||
        !a.q_kl || !a.pdata_ij || !a.pdata_kl || !a.dm_cond || !a.dm || a.n_dm <= 0 ||
        a.nbas <= 0 || a.nao <= 0) {
        set_err("%s: null pointer or empty dimension", who);
        return 3;
    }
    if (!table_ok(a.rys)) { set_err("%s: Rys table not set", who); return 3; }
    if (!(a.omega >= 0.)) { set_err("%s: omega must be >= 0 (erf attenuation)", who); return 3; }
    for (int t = 0; t < 2; ++t) {
        const JKWindows &w = a.kl_win[t];
        if (w.win && (w.n_win <= 0 || w.width <= 0 || !w.pairs || !w.q || !w.pdata)) {
            set_err("%s: kl windows without their pair lists or sizes", who);
            return 3;
        }
    }
    return 0;
}

int jk_build(const JKBuildArgs *args, void *stream)
{
    g_err[0] = 0;
    JKBuildArgs a;
    {
        const int st = check_args(args, a, "jk_build");
        if (st == -1) return 0;
        if (st) return st;
    }
    if (!a.vj && !a.vk) { set_err("jk_build: neither vj nor vk given"); return 3; }
    if (a.screen < JK_SCREEN_JK || a.screen > JK_SCREEN_K ||
        (a.screen == JK_SCREEN_J && a.vk) || (a.screen == JK_SCREEN_K && a.vj)) {
        set_err("jk_build: screen %d does not match the requested matrices (J-only screen "
                "takes no vk, K-only no vj)", a.screen);
        return 3;
    }
    if (a.n_repl > 1 && a.repl_stride < (int64_t)a.n_dm * a.nao * a.nao) {
        set_err("jk_build: repl_stride smaller than %d x nao^2", a.n_dm);
        return 3;
    }
    int nroots = (a.li + a.lj + a.lk + a.ll) / 2 + 1;
    if (nroots > JK_RYS_NMAX || a.rys.xmax[nroots-1] <= 0) {
        set_err("jk_build: Rys table lacks %d roots", nroots);
        return 3;
    }
    hipError_t err = dispatch(a, (hipStream_t)stream);
    if (err != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "jk_build (%d%d|%d%d): %s", a.li, a.lj, a.lk, a.ll,
                 hipGetErrorString(err));
        return 4;
    }
    return 0;
}

int jk_energy_ip1(const JKBuildArgs *args, double *ejk, double j_factor, double k_factor,
                  void *stream)
{
    g_err[0] = 0;
    JKBuildArgs a;
    {
        const int st = check_args(args, a, "jk_energy_ip1");
        if (st == -1) return 0;
        if (st) return st;
    }
    if (a.swap_ijkl) {
        set_err("jk_energy_ip1: swap_ijkl is for J/K builds only");
        return 3;
    }
    if (!ejk || a.n_dm > 2) {
        set_err("jk_energy_ip1: ejk missing or n_dm = %d (1 or 2 densities)", a.n_dm);
        return 3;
    }
    const int nroots = (a.li + a.lj + a.lk + a.ll + 1) / 2 + 1;
    if (nroots > JK_RYS_NMAX || a.rys.xmax[nroots-1] <= 0) {
        set_err("jk_energy_ip1: Rys table lacks %d roots", nroots);
        return 3;
    }
    // factors of the per-element two-particle density, as the reference's
    // host code sets them (rys_contract_jk_ip1.cu:1789-1797): 8-fold
    // symmetry times the 1/2 of the Coulomb energy for J, and the symmetric
    // pair of exchange products for K, halved for a closed-shell density
    if (a.n_dm == 1) k_factor *= .5;
    const double jf = 4. * j_factor, kf = -2. * k_factor;
    hipError_t err = jk_dispatch_generic_ip1(a, ejk, jf, kf, (hipStream_t)stream);
    if (err != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "jk_energy_ip1 (%d%d|%d%d): %s", a.li, a.lj, a.lk, a.ll,
                 hipGetErrorString(err));
        return 4;
    }
    return 0;
}

int jk_pair_data(double *out, const uint32_t *pairs, int npairs, int iprim,
                 int jprim, const int32_t *bas, const double *env, int nbas,
                 void *stream)
{
    g_err[0] = 0;
    if (npairs <= 0) return 0;
    if (!out || !pairs || !bas || !env || iprim <= 0 || jprim <= 0 || nbas <= 0) {
        set_err("jk_pair_data: bad arguments"); return 3;
    }
    long n = (long)npairs * iprim * jprim;
    hipLaunchKernelGGL(pair_data_kernel, dim3((n + 127) / 128), dim3(128), 0,
                       (hipStream_t)stream, out, pairs, npairs, iprim, jprim, bas, env, nbas);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) { set_err("jk_pair_data: %s", hipGetErrorString(err)); return 4; }
    return 0;
}

int jk_ovlp_mask(int8_t *mask, const float *exps, const float *log_coeff,
                 const double *coords, const int32_t *ls, int nbas,
                 float log_cutoff, void *stream)
{
    g_err[0] = 0;
    if (!mask || !exps || !log_coeff || !coords || !ls || nbas <= 0) {
        set_err("jk_ovlp_mask: bad arguments"); return 3;
    }
    dim3 t(16, 16), b((nbas + 15) / 16, (nbas + 15) / 16);
    hipLaunchKernelGGL(ovlp_mask_kernel, b, t, 0, (hipStream_t)stream,
                       mask, exps, log_coeff, coords, ls, nbas, log_cutoff);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) { set_err("jk_ovlp_mask: %s", hipGetErrorString(err)); return 4; }
    return 0;
}

int jk_pair_qcond(float *q, const uint32_t *pairs, int npairs,
                  const int32_t *bas, const double *env, int nbas,
                  JKRysTable rys, double omega, double lr_factor, double sr_factor,
                  void *stream)
{
    g_err[0] = 0;
    if (npairs <= 0) return 0;
    if (!q || !pairs || !bas || !env || nbas <= 0 || !table_ok(rys) || !(omega >= 0.)) {
        set_err("jk_pair_qcond: bad arguments"); return 3;
    }
    hipLaunchKernelGGL(pair_qcond_kernel, dim3((npairs + 63) / 64), dim3(64), 0,
                       (hipStream_t)stream, q, pairs, npairs, bas, env, nbas, rys, omega,
                       lr_factor, sr_factor);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) { set_err("jk_pair_qcond: %s", hipGetErrorString(err)); return 4; }
    return 0;
}

int jk_dm_cond(float *dm_cond, double *work, const double *dm, int n_dm,
               int nao, const int32_t *ao_loc, int nbas, int hermi, void *stream)
{
    g_err[0] = 0;
    if (!dm_cond || !work || !dm || !ao_loc || n_dm <= 0 || nao <= 0 || nbas <= 0) {
        set_err("jk_dm_cond: bad arguments"); return 3;
    }
    dim3 b((nbas + 63) / 64, nbas);
    hipLaunchKernelGGL(dm_block_max_kernel, b, dim3(64), 0, (hipStream_t)stream,
                       work, dm, n_dm, nao, ao_loc, nbas);
    hipLaunchKernelGGL(dm_cond_log_kernel, b, dim3(64), 0, (hipStream_t)stream,
                       dm_cond, work, nbas, hermi);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) { set_err("jk_dm_cond: %s", hipGetErrorString(err)); return 4; }
    return 0;
}

int jk_dm_cond_noalloc(float *dm_cond, const double *dm, int n_dm, int nao,
                       const int32_t *ao_loc, int nbas, int hermi, void *stream)
{
    g_err[0] = 0;
    if (!dm_cond || !dm || !ao_loc || n_dm <= 0 || nao <= 0 || nbas <= 0) {
        set_err("jk_dm_cond_noalloc: bad arguments"); return 3;
    }
    if (nbas > DM_COND_SCRATCH_NBAS) {
        set_err("jk_dm_cond_noalloc: nbas = %d exceeds the static scratch (at most %d shells)",
                nbas, DM_COND_SCRATCH_NBAS);
        return 3;
    }
    double *work = scratch_ptr("jk_dm_cond_noalloc");
    if (!work) return 4;
    // dm_block_max_kernel writes every one of the nbas^2 entries it reads
    // back, so the scratch needs no clearing
    dim3 b((nbas + 63) / 64, nbas);
    hipLaunchKernelGGL(dm_block_max_kernel, b, dim3(64), 0, (hipStream_t)stream,
                       work, dm, n_dm, nao, ao_loc, nbas);
    hipLaunchKernelGGL(dm_cond_log_kernel, b, dim3(64), 0, (hipStream_t)stream,
                       dm_cond, work, nbas, hermi);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) { set_err("jk_dm_cond_noalloc: %s", hipGetErrorString(err)); return 4; }
    return 0;
}

}  // extern "C"
