/*
 * C-ABI of the MI355X direct-SCF J/K engine (libjkhip.so).
 *
 * Plain C: device buffers are passed as raw pointers (e.g. torch
 * tensor.data_ptr()), HIP streams as void*.  Every entry point returns 0 on
 * success and a non-zero status otherwise; jk_last_error() gives the message.
 * The callee never allocates device memory: the caller owns and zeroes
 * vj/vk, exactly as gpu4pyscf/scf/jk.py:538-539 does for the reference.
 *
 * Layout conventions (identical to gpu4pyscf/lib/gvhf-rys/vhf.cuh:5-24):
 *   bas: nbas x 8 int32 (ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF,
 *        PTR_EXP, PTR_COEFF, PTR_BAS_COORD); env: float64 with s/p
 *        coefficients already scaled by sqrt((2l+1)/4pi)
 *        (gpu4pyscf/gto/mole.py:1120-1132); ao_loc: nbas+1 int32.
 *   dm, vj, vk: n_dm x nao x nao float64, row major, sorted Cartesian AOs.
 *   pair ids: ish*nbas + jsh (uint32) with ish >= jsh.
 */
#ifndef JKHIP_H
#define JKHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JK_ABI_VERSION  13
#define JK_RYS_NMAX     11      /* max Rys roots: (hh|hh) */
#define JK_LMAX         5       /* max angular momentum on the GPU path */

/* Chebyshev tables of Rys roots/weights (device pointer + layout).
 * coef[off[n-1] + (k*2n + f)*(degree+1) + d] on unit intervals [k, k+1),
 * k < xmax[n-1]; above, u_i = large[n*(n-1)/2 + i] / x and
 * w_i = large[large_w + n*(n-1)/2 + i] / sqrt(x).
 * Boys grid (one to three roots): coef[boys + g*JK_BOYS_M + m] = F_m(g*boys_step),
 * boys_rows grid points covering x < xmax[2]. */
#define JK_BOYS_M       13
typedef struct {
    const double *coef;
    const float *coef32;        /* coef rounded to float (FP32 quartets) */
    int64_t off[JK_RYS_NMAX];
    double xmax[JK_RYS_NMAX];
    int64_t large_r;            /* offset of large-x nodes in coef */
    int64_t large_w;            /* offset of large-x weights in coef */
    int degree;
    int boys_rows;
    int64_t boys;               /* offset of the Boys grid in coef */
    double boys_step, boys_inv_step;
} JKRysTable;

/* Column windows of a kl pair list: pairs, q and pdata hold the npairs_kl kl
 * pairs ordered window by window, each window's pairs in the descending q
 * order of the task's list; window w is win[6w .. 6w+5] = {first pair, end
 * pair, k_ao0, k_nao, l_ao0, l_nao}: its pairs have their k AOs in
 * [k_ao0, k_ao0+k_nao) and their l AOs in [l_ao0, l_ao0+l_nao); width = max
 * over the windows of k_nao + l_nao.  A work item is then (ij pair, window)
 * and keeps the K rows of its i and j AOs over the window's columns in LDS.
 * No counterpart in the reference, whose K blocks are accumulated with
 * global atomics. */
typedef struct {
    const uint32_t *pairs;
    const float *q;
    const double *pdata;
    const int32_t *win;             /* device */
    int n_win;
    int width;
} JKWindows;

/* One angular-class task of the J/K build.
 * Replaces RYS_build_jk (gpu4pyscf/lib/gvhf-rys/rys_contract_jk.cu:344-350),
 * which the reference binds at gpu4pyscf/scf/jk.py:43 and calls per task at
 * gpu4pyscf/scf/jk.py:554-585. */
typedef struct {
    int abi_version;
    int nbas, nao;
    const int32_t *bas;             /* device */
    const double *env;              /* device */
    const int32_t *ao_loc;          /* device */
    /* class: all shells of a group share l and nprim */
    int li, lj, lk, ll;
    int iprim, jprim, kprim, lprim;
    /* AO columns [k_ao0, k_ao0 + k_nao) of the k group, same for l */
    int k_ao0, k_nao, l_ao0, l_nao;
    /* screened pair lists of the two group pairs, in descending-q bins */
    int npairs_ij, npairs_kl;
    const uint32_t *pairs_ij, *pairs_kl;    /* device */
    const float *q_ij, *q_kl;               /* device, log sqrt max|(ij|ij)| */
    const double *pdata_ij, *pdata_kl;      /* device, jk_pair_data() records */
    const float *dm_cond;                   /* device, nbas x nbas */
    float cutoff;                           /* log(direct_scf_tol) */
    float dm_penalty;                       /* max(dm_cond) */
    int n_dm;
    const double *dm;                       /* device */
    double *vj;                             /* device or NULL */
    double *vk;                             /* device or NULL */
    JKRysTable rys;
    unsigned long long *n_quartets;         /* device counter or NULL */
    /* Mixed precision: a surviving quartet with q_ij + q_kl + max of its six
     * dm_cond values below fp32_cut is evaluated in FP32 (roots and 2D
     * recursions; digestion and J/K stay FP64).  -1e30 disables it.  No
     * counterpart in the reference. */
    float fp32_cut;
    unsigned long long *n_quartets_fp32;    /* device counter or NULL */
    /* Set checksum of the surviving quartets, for set-equality tests: adds,
     * mod 2^64, jk_quartet_mix(bas_ij << 32 | bas_kl) of every surviving
     * quartet, where jk_quartet_mix is the SplitMix64 finaliser
     *   z = (z ^ z >> 30) * 0xbf58476d1ce4e5b9;
     *   z = (z ^ z >> 27) * 0x94d049bb133111eb;  return z ^ z >> 31.
     * Device counter or NULL.  No counterpart in the reference, whose
     * survivor queue (create_tasks.cu:114-180) is internal to the kernel. */
    unsigned long long *quartet_hash;
    /* Replicated accumulation: with n_repl > 1, vj and vk each point to
     * n_repl zeroed copies of the n_dm x nao x nao block, repl_stride
     * doubles apart, and workgroup b adds into copy b % n_repl; the caller
     * sums the copies afterwards.  Workgroups are dispatched round-robin
     * over the 8 XCDs of an MI355X, so n_repl = 8 gives each XCD its own
     * copy, which keeps the XCDs' atomic adds off each other's cache
     * lines.  0 or 1: plain accumulation into vj/vk.  No counterpart in the
     * reference. */
    int n_repl;
    int64_t repl_stride;
    /* Operator: scale * erf(omega r)/r for omega > 0, scale / r for
     * omega == 0 (scale multiplies every integral of the launch).  The
     * range-separated operators of gpu4pyscf/scf/jk.py:1122-1168,
     * lr erf(w r)/r + sr erfc(w r)/r = sr / r + (lr - sr) erf(w r)/r, are
     * built from two launches of each class task; the reference instead
     * doubles the Rys roots (rys_roots_for_k.cu:19-68). */
    double omega;
    double scale;           /* 0 is taken as 1 */
    /* Survival screen: which density blocks keep a quartet whose bound
     * q_ij + q_kl + dm_cond exceeds the cutoff.  JK_SCREEN_JK: any of the
     * six blocks D_ij, D_kl, D_ik, D_jk, D_il, D_jl (RYS_build_jk,
     * create_tasks.cu:155-163); JK_SCREEN_J: D_ij or D_kl, vk must be NULL
     * (RYS_build_j, create_tasks.cu:222-225); JK_SCREEN_K: D_ik, D_jk, D_il
     * or D_jl, vj must be NULL (RYS_build_k, create_tasks.cu:86-93). */
    int screen;
    /* Column windows of the kl list (JKWindows), for the K rows of a work
     * item accumulated in LDS: kl_win[0] for the per-lane kernels, kl_win[1]
     * (narrower) for the team kernels; a NULL win means none.
     * kl_win_policy 0: the launcher picks, per launch, whole groups or
     * windows: first the choice that keeps all four K blocks in LDS, then
     * the one whose LDS leaves the most workgroups per CU (the team kernels
     * take kl_win[1] only under policy 1); 1: windows whenever their rows
     * fit in LDS (tests).  Results do not depend on the choice beyond the
     * order of the floating-point additions. */
    JKWindows kl_win[2];
    int kl_win_policy;
    /* 1: the task (i, j, k, l) runs with bra and ket exchanged -- li..ll,
     * the prims, k_ao0..l_nao and the pair lists are those of (k, l, i, j)
     * -- so that the per-lane J atomics fall on the smaller block; the
     * canonical test becomes bas_ij <= bas_kl and the quartet checksum keeps
     * the task's order.  Only for tasks whose two group pairs differ.  No
     * counterpart in the reference. */
    int swap_ijkl;
} JKBuildArgs;

#define JK_SCREEN_JK    0
#define JK_SCREEN_J     1
#define JK_SCREEN_K     2

/* Accumulate one class task into vj/vk (before the hermi post-processing of
 * gpu4pyscf/scf/jk.py:594-616).  stream: hipStream_t or NULL. */
int jk_build(const JKBuildArgs *args, void *stream);

/* First derivatives of the two-electron energy per atom for one class task:
 * ejk[atom*3 + x] += d/dR_atom,x of
 *   sum over the task's surviving quartets of (ij|kl) Gamma_ijkl,
 * Gamma = j_factor J-part - k_factor K-part of the two-particle density of
 * the n_dm (1: closed shell, 2: alpha and beta) densities in args->dm, with
 * the factors and the screening of RYS_per_atom_jk_ip1
 * (gpu4pyscf/lib/gvhf-rys/rys_contract_jk_ip1.cu:1747-1800,
 * create_tasks.cu:829-889); args->dm_cond is the block log max of the
 * densities and args->dm_penalty is ignored.  ejk: natm x 3 FP64 on the
 * device, zeroed by the caller; atoms are read from bas[ATOM_OF].
 * Replaces RYS_per_atom_jk_ip1, bound by gpu4pyscf/grad/rhf.py:106-154. */
int jk_energy_ip1(const JKBuildArgs *args, double *ejk, double j_factor, double k_factor,
                  void *stream);

/* The per-task fields of JKBuildArgs, for a whole build in one call. */
typedef struct {
    int li, lj, lk, ll;
    int iprim, jprim, kprim, lprim;
    int k_ao0, k_nao, l_ao0, l_nao;
    int npairs_ij, npairs_kl;
    const uint32_t *pairs_ij, *pairs_kl;
    const float *q_ij, *q_kl;
    const double *pdata_ij, *pdata_kl;
    JKWindows kl_win[2];            /* as in JKBuildArgs */
    int swap_ijkl;                  /* as in JKBuildArgs */
} JKTask;

/* All class tasks of one build: the task loop of gpu4pyscf/scf/jk.py:554-585
 * in one call.  *args supplies the fields shared by the tasks; tasks[t] the
 * per-task ones.  Tasks are launched from the last to the first, the m-th
 * launch on streams[m % nstreams] (hipStream_t; NULL entries are the null
 * stream).  When args->n_quartets, n_quartets_fp32 or quartet_hash is set it
 * points to ntasks counters and task t adds into element t.  Stops at and
 * returns the status of the first task that fails.  No counterpart in the
 * reference, which calls RYS_build_jk once per task from Python. */
int jk_build_tasks(const JKBuildArgs *args, const JKTask *tasks, int ntasks,
                   void *const *streams, int nstreams);

/* ---- McMurchie-Davidson J-engine (basis of single-primitive shells,
 * SortedGTO(decontract=True)).  Replaces the reference's MD J path
 * (gpu4pyscf/scf/j_engine.py:50-236; dm_to_Rt / MD_build_j / Rt_to_dm of
 * gpu4pyscf/lib/gvhf-md, bound at j_engine.py:44-47). */

/* Hermite densities of one group pair's screened pairs (ish >= jsh, shells
 * of angular momenta li >= lj, li <= 4, li + lj <= 6):
 * dh[p*NH + h] = K_p (-1)^(t+u+v) sum_{c in ish, d in jsh} D~_cd E^{cd}_{tuv}
 * with NH = (L+1)(L+2)(L+3)/6, L = li + lj, (t, u, v) of Hermite index h in
 * the order total ascending, t descending, u descending; K_p the pair
 * prefactor of jk_pair_data (pdata, one primitive pair per pair) and
 * D~_cd = D_cd + D_dc for ish != jsh.  dm: nao x nao sorted Cartesian.
 * Replaces dm_to_Rt (gpu4pyscf/lib/gvhf-md/contract_Et.cu, called at
 * gpu4pyscf/scf/j_engine.py:380-394). */
int jk_md_dm_to_hermite(double *dh, const double *dm, int nao, const uint32_t *pairs,
                        int npairs, const double *pdata, int li, int lj, const int32_t *bas,
                        const double *env, const int32_t *ao_loc, int nbas, void *stream);

/* vj[u, v] += K_p sum_h E^{uv}_h jh[p*NH + h] for u in ish, v in jsh of every
 * pair p (and vj[v, u] for ish != jsh); plain adds (each pair owns its
 * block).  Replaces Rt_to_dm (gpu4pyscf/scf/j_engine.py:398-430). */
int jk_md_hermite_to_j(double *vj, const double *jh, int nao, const uint32_t *pairs,
                       int npairs, const double *pdata, int li, int lj, const int32_t *bas,
                       const double *env, const int32_t *ao_loc, int nbas, void *stream);

/* One bra group pair against one ket group pair:
 * jh_ij[p*NH_ij + h] += sum_{q surviving}
 *   (a_p + a_q)^-1/2 sum_g R_{h+g}(a_p a_q / (a_p + a_q), P_p - Q_q) dh_kl[q*NH_kl + g],
 * where (p, q) survives the J-only predicate of RYS_build_j in its float32
 * order (create_tasks.cu:213-225): kl_cutoff = cutoff - q_ij,
 * q_kl + dm_penalty >= kl_cutoff and max(dmax_ij, dmax_kl) > kl_cutoff - q_kl.
 * Both pair lists sorted by descending q; lij + lkl <= MD_LMAX (4).  boys: the
 * Boys grid of gen_md_boys.py (boys_rows x boys_m, F_m at g * boys_step).
 * Replaces MD_build_j (gpu4pyscf/lib/gvhf-md/md_j_driver.cu, called per task
 * at gpu4pyscf/scf/j_engine.py:199-222), which symmetrises over (ij) <-> (kl);
 * here every bra pair sums over every ket pair. */
typedef struct {
    int abi_version;
    int lij, lkl;
    int npairs_ij, npairs_kl;
    const double *pdata_ij, *pdata_kl;     /* device, jk_pair_data records */
    const float *q_ij, *q_kl;              /* device, log sqrt max|(ij|ij)| */
    const float *dmax_ij, *dmax_kl;        /* device, log max|D| of each pair block */
    float cutoff;                          /* log(direct_scf_tol) */
    float dm_penalty;                      /* max of the density's dm_cond */
    const double *dh_kl;                   /* device, jk_md_dm_to_hermite output */
    double *jh_ij;                         /* device, accumulated */
    const double *boys;                    /* device */
    int boys_rows, boys_m;
    double boys_step, boys_inv_step;
    /* survivor count and set checksum (sum mod 2^64 of the SplitMix64
     * finaliser of bas_ij << 32 | bas_kl, as JKBuildArgs.quartet_hash) over
     * the surviving (ij, kl) with bas_ij >= bas_kl: the canonical quartets
     * of the J-only set; device counters or NULL.  pairs_ij / pairs_kl (the
     * pair ids) are read only for them. */
    unsigned long long *n_quartets;
    unsigned long long *quartet_hash;
    const uint32_t *pairs_ij, *pairs_kl;
    /* operator: 1/r (0), erf(omega r)/r (omega > 0), erfc(|omega| r)/r
     * (omega < 0), the molecule's omega as in the reference's J-engine */
    double omega;
} MDBuildArgs;

int jk_md_build_j(const MDBuildArgs *args, void *stream);

/* Primitive-pair records of a pair list, geometry-only data precomputed once
 * per molecule: out[(p*iprim*jprim + ip*jprim + jp)*6 + (0..5)] =
 * (a, 1/a, Px, Py, Pz, K) with a = ai + aj, P = (ai Ri + aj Rj)/a and
 * K = sqrt(2 pi^2.5) ci cj exp(-ai aj/a |Ri - Rj|^2) / a, so that the ERI
 * prefactor of a primitive quartet is K_ij K_kl / sqrt(a_ij + a_kl).
 * No counterpart in the reference (it recomputes these in every kernel,
 * gpu4pyscf/lib/gvhf-rys/rys_contract_jk.cu:146-160). */
int jk_pair_data(double *out, const uint32_t *pairs, int npairs, int iprim,
                 int jprim, const int32_t *bas, const double *env, int nbas,
                 void *stream);

/* float32 shell-pair overlap mask (lower triangle), replaces
 * PBCovlp_mask_estimation (gpu4pyscf/lib/pbc/overlap.cu:2378-2387) as used by
 * _shell_overlap_mask (gpu4pyscf/pbc/gto/int1e.py:403-428).
 * exps, log_coeff: per-shell most diffuse primitive (float32, device);
 * coords: nbas x 3 float64 (device); ls: nbas int32 (device);
 * mask: nbas x nbas int8 (device, written fully). */
int jk_ovlp_mask(int8_t *mask, const float *exps, const float *log_coeff,
                 const double *coords, const int32_t *ls, int nbas,
                 float log_cutoff, void *stream);

/* q[p] = float32(log(max |(ij|ij)|) / 2), -700 if zero, for pairs p, of
 * the operator sr_factor / r + (lr_factor - sr_factor) erf(omega r)/r
 * (omega = 0: lr_factor / r); shells up to l = 5; evaluated in FP64.
 * Replaces int2e_qcond_estimator (gpu4pyscf/lib/gvhf-rys/
 * nr_sr_estimator.cu:416-430). */
int jk_pair_qcond(float *q, const uint32_t *pairs, int npairs,
                  const int32_t *bas, const double *env, int nbas,
                  JKRysTable rys, double omega, double lr_factor, double sr_factor,
                  void *stream);

/* dm_cond[ish,jsh] = float32(log(max_{dm, u in ish, v in jsh} |D| + 1e-300)),
 * with the block maxima symmetrised (A + A^T) first when hermi == 0.
 * Replaces condense('absmax') + log (gpu4pyscf/scf/jk.py:510-515,
 * gpu4pyscf/lib/cupy_helper.py:1068-1184).  work: nbas*nbas float64. */
int jk_dm_cond(float *dm_cond, double *work, const double *dm, int n_dm,
               int nao, const int32_t *ao_loc, int nbas, int hermi,
               void *stream);

/* jk_dm_cond without the work argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The nbas x nbas block maxima live in a
 * static device array of the library (32 MiB per GPU, held until the
 * process ends), so nbas must not exceed 2048 (status 3 otherwise); calls
 * no allocating routine.  Calls that overlap on one GPU share that array
 * and race: serialise them (one stream, or events). */
int jk_dm_cond_noalloc(float *dm_cond, const double *dm, int n_dm, int nao,
                       const int32_t *ao_loc, int nbas, int hermi, void *stream);

/* AO transform between a caller basis and the sorted Cartesian basis:
 * out[b, i, j] = sum_{p in [ptr[i], ptr[i+1])} sum_{q in [ptr[j], ptr[j+1])}
 *                coef[p] coef[q] in[b, idx[p], idx[q]]
 * for b < n_batch, i, j < n_out; in is (n_batch, n_in, n_in) and out
 * (n_batch, n_out, n_out), row major FP64; (ptr, idx, coef) are the
 * compressed rows of C (out = C in C^T) or of C^T (out = C^T in C).
 * Replaces the bra/ket sorted<->spherical/Cartesian passes
 * (gpu4pyscf/lib/gvhf-rys/mole_helper.cu:1734-1900) behind
 * SortedGTO.apply_C_mat_CT / apply_CT_mat_C (gpu4pyscf/gto/mole.py:906-938).
 * n_out <= 262140 and n_batch <= 65535 (status 3 otherwise). */
int jk_ao_transform(double *out, const double *in, int n_batch, int n_out, int n_in,
                    const int32_t *ptr, const int32_t *idx, const double *coef,
                    void *stream);

/* One-electron integrals on the sorted Cartesian basis (nao x nao, written
 * fully): overlap S, kinetic T and nuclear attraction
 * V = -sum_A Z_A <i|1/|r-R_A||j> of point charges.  coords: natm x 3 Bohr,
 * charges: natm (device).  The reference computes these for its SCF driver
 * (get_hcore / get_ovlp, gpu4pyscf/scf/hf.py:138-155, 880-883). */
int jk_int1e(double *S, double *T, double *V, const int32_t *bas, const double *env,
             const int32_t *ao_loc, int nbas, int nao, const double *coords,
             const double *charges, int natm, JKRysTable rys, void *stream);

/* ---- learned initial guess: BF16 MFMA network kernels (no counterpart in
 * the reference; they replace its minao/atom guess, gpu4pyscf/scf/hf.py:
 * 414-607).  bf16 buffers are uint16_t bit patterns; accumulation is FP32. */

/* C(m,n) = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn];
 * mode 0: C = acc + bias; mode 1: C = acc + bias, act = bf16(silu(C));
 * mode 2: C += acc.  bias may be NULL. */
int mlp_gemm_bf16(float *C, uint16_t *act, const uint16_t *A, const uint16_t *B,
                  const float *bias, int M, int N, int K, long sam, long sak, long sbk,
                  long sbn, int ldc, int mode, void *stream);

/* dZ = dA * silu'(Z) as bf16, db[n] += sum_m dZ(m,n); M x N row major. */
int mlp_silu_backward(uint16_t *dz, float *db, const float *dA, const float *Z, int M, int N,
                      void *stream);

/* db[n] += sum_m G(m,n). */
int mlp_colsum(float *db, const float *G, int M, int N, void *stream);

/* loss += scale sum mask (Y-T)^2; dy = 2 scale mask (Y-T) (FP32 and bf16). */
int mlp_mse_grad(float *dy, uint16_t *dy_bf16, float *loss, const float *Y, const float *T,
                 const float *mask, long n, float scale, void *stream);

/* AdamW step on FP32 master weights (bias-corrected, decoupled weight
 * decay), refreshing the bf16 copy. */
int mlp_adamw(float *w, uint16_t *wb, float *m, float *v, const float *g, long n, float lr,
              float beta1, float beta2, float eps, float wd, int step, void *stream);

/* bf16(in), round to nearest even. */
int mlp_to_bf16(uint16_t *out, const float *in, long n, void *stream);

/* Last error message of this thread ("" if none). */
const char *jk_last_error(void);

/* ABI version compiled into the library (JK_ABI_VERSION). */
int jk_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
