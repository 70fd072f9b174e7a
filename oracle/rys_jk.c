/*
 * CPU oracle for the direct-SCF J/K path.  TEST INFRASTRUCTURE ONLY: it is
 * the checker that the HIP path is compared against, never part of the
 * product.  Plain C + OpenMP, FP64.
 *
 * What it restates (gpu4pyscf v1.8.1):
 *   - Rys quadrature ERIs: 2D integrals by VRR in i, TRR in k, HRR for j and
 *     l, weight folded into the z factor, prefactor 2 pi^2.5 / (p q sqrt(p+q))
 *     -- lib/gvhf-rys/build_rys_gxyz.cuh:18-123, rys_contract_jk.cu:198-288.
 *   - survival predicate of a shell quartet -- lib/gvhf-rys/create_tasks.cu:
 *     114-180 (float32 compares in the reference's order).
 *   - digestion into J and K with the 8-fold symmetry factors --
 *     lib/gvhf-rys/rys_contract_jk.cu:156-158, 214-220, 291-330.
 *   - task enumeration over angular groups i>=j, k<=i, l<=k --
 *     gpu4pyscf/scf/jk.py:520-524.
 *
 * Rys roots and weights are computed here by an independent method (not the
 * Chebyshev tables of the product): the Rys measure exp(-x t^2) dt on [0,1]
 * (in u = t^2) is discretised with a 128-point Gauss-Legendre rule on
 * [0, min(1, 13/sqrt(x))], then the Lanczos process (Stieltjes procedure,
 * Gautschi, "Orthogonal Polynomials", 2004, sec. 2.2.3) with full
 * reorthogonalisation gives the Jacobi matrix, whose eigenvalues are the
 * nodes u_i and whose first eigenvector components give the weights
 * (Golub & Welsch, Math. Comp. 23, 221 (1969)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ANG_OF          1
#define NPRIM_OF        2
#define PTR_EXP         5
#define PTR_COEFF       6
#define PTR_BAS_COORD   7
#define BAS_SLOTS       8
#define PI_FAC          34.98683665524972497   /* 2 pi^2.5 */
#define MAXL            6
#define MAXROOTS        14
#define GL_M            128

static double gl_t[GL_M], gl_w[GL_M];
static int gl_ready = 0;

/* Gauss-Legendre nodes/weights on [0,1] by Newton iteration on P_M. */
static void gl_init(void)
{
    if (gl_ready) return;
    int M = GL_M;
    for (int i = 0; i < (M + 1) / 2; i++) {
        double z = cos(M_PI * (i + 0.75) / (M + 0.5)), z1, pp = 0;
        int it = 0;
        do {
            double p1 = 1., p2 = 0.;
            for (int j = 1; j <= M; j++) {
                double p3 = p2; p2 = p1;
                p1 = ((2. * j - 1.) * z * p2 - (j - 1.) * p3) / j;
            }
            pp = M * (z * p1 - p2) / (z * z - 1.);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > 1e-15 && ++it < 100);
        double w = 2. / ((1. - z * z) * pp * pp);
        gl_t[i] = .5 * (1. - z);       gl_w[i] = .5 * w;
        gl_t[M-1-i] = .5 * (1. + z);   gl_w[M-1-i] = .5 * w;
    }
    gl_ready = 1;
}

/* Eigen-decomposition of a symmetric tridiagonal matrix (implicit QL,
 * Numerical Recipes "tqli" scheme).  d: diagonal (in: n, out: eigenvalues);
 * e: off-diagonal e[1..n-1] (destroyed); z: n x n, in: identity, out: eigvecs
 * as columns (only row 0 is used by the caller). */
static void tqli(double *d, double *e, int n, double *z)
{
    for (int i = 1; i < n; i++) e[i-1] = e[i];
    e[n-1] = 0.;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m+1]);
                volatile double sum = fabs(e[m]) + dd;
                if (sum == dd) break;
            }
            if (m != l) {
                if (iter++ == 60) break;
                double g = (d[l+1] - d[l]) / (2. * e[l]);
                double r = hypot(g, 1.);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
                double s = 1., c = 1., p = 0.;
                int i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i], b = c * e[i];
                    e[i+1] = (r = hypot(f, g));
                    if (r == 0.) { d[i+1] -= p; e[m] = 0.; break; }
                    s = f / r; c = g / r;
                    g = d[i+1] - p;
                    r = (d[i] - g) * s + 2. * c * b;
                    d[i+1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; k++) {
                        f = z[k*n+i+1];
                        z[k*n+i+1] = s * z[k*n+i] + c * f;
                        z[k*n+i] = c * z[k*n+i] - s * f;
                    }
                }
                if (r == 0. && i >= l) continue;
                d[l] -= p; e[l] = g; e[m] = 0.;
            }
        } while (m != l);
    }
}

/* Rys nodes u_i = t_i^2 and weights w_i with sum_i w_i u_i^m = F_m(x). */
void oracle_rys_roots(int n, double x, double *u, double *w)
{
    gl_init();
    double tmax = x > 169. ? 13. / sqrt(x) : 1.;
    double pu[GL_M], v[MAXROOTS+1][GL_M], wk[GL_M];
    double alpha[MAXROOTS], beta[MAXROOTS+1];
    double mass = 0.;
    for (int m = 0; m < GL_M; m++) {
        double t = gl_t[m] * tmax;
        pu[m] = t * t;
        double pm = gl_w[m] * tmax * exp(-x * t * t);
        mass += pm;
        v[0][m] = sqrt(pm);
    }
    double nrm = sqrt(mass);
    for (int m = 0; m < GL_M; m++) v[0][m] /= nrm;
    beta[0] = 0.;
    for (int k = 0; k < n; k++) {
        double a = 0.;
        for (int m = 0; m < GL_M; m++) { wk[m] = pu[m] * v[k][m]; a += v[k][m] * wk[m]; }
        alpha[k] = a;
        if (k == n - 1) break;
        for (int m = 0; m < GL_M; m++) {
            wk[m] -= a * v[k][m];
            if (k > 0) wk[m] -= beta[k] * v[k-1][m];
        }
        for (int pass = 0; pass < 2; pass++) {
            for (int q = 0; q <= k; q++) {
                double s = 0.;
                for (int m = 0; m < GL_M; m++) s += wk[m] * v[q][m];
                for (int m = 0; m < GL_M; m++) wk[m] -= s * v[q][m];
            }
        }
        double b = 0.;
        for (int m = 0; m < GL_M; m++) b += wk[m] * wk[m];
        b = sqrt(b);
        beta[k+1] = b;
        for (int m = 0; m < GL_M; m++) v[k+1][m] = wk[m] / b;
    }
    double d[MAXROOTS], e[MAXROOTS], z[MAXROOTS*MAXROOTS];
    for (int i = 0; i < n; i++) {
        d[i] = alpha[i];
        e[i] = beta[i];
        for (int j = 0; j < n; j++) z[i*n+j] = (i == j);
    }
    tqli(d, e, n, z);
    /* sort ascending */
    int idx[MAXROOTS];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (d[idx[j]] < d[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < n; i++) {
        u[i] = d[idx[i]];
        double z0 = z[idx[i]];
        w[i] = mass * z0 * z0;
    }
}

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* lexical Cartesian powers of component c of shell l */
static void cart_pow(int l, int *px, int *py, int *pz)
{
    int c = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            px[c] = lx; py[c] = ly; pz[c] = l - lx - ly; c++;
        }
}

/* Operator of the ERIs: w_full / r + w_lr erf(omega r)/r.  The long-range
 * part uses the Rys rule of the attenuated operator: with
 * a = omega^2 / (omega^2 + theta) the roots and weights of x a, scaled by a
 * and sqrt(a) (the substitution t = t_max tau, t_max^2 = a, of the Rys
 * integral over t in [0, t_max]).  Set once before the calls, not thread
 * local. */
static double g_omega = 0., g_w_full = 1., g_w_lr = 0.;

/* lr erf(omega r)/r + sr erfc(omega r)/r = sr / r + (lr - sr) erf(omega r)/r
 * (omega > 0); omega == 0: lr / r. */
void oracle_set_operator(double omega, double lr, double sr)
{
    g_omega = omega > 0 ? omega : 0.;
    g_w_full = omega > 0 ? sr : lr;
    g_w_lr = omega > 0 ? lr - sr : 0.;
}

/* one contracted Cartesian shell */
typedef struct { int l, nprim; const double *exps, *coefs, *R; } Shell;

static Shell shell_of(const int *bas, const double *env, int sh)
{
    const int *b = bas + sh*BAS_SLOTS;
    Shell s = {b[ANG_OF], b[NPRIM_OF], env + b[PTR_EXP], env + b[PTR_COEFF], env + b[PTR_BAS_COORD]};
    return s;
}

/* Cartesian (ij|kl) block of four shells, layout
 * out[((i*nfj + j)*nfk + k)*nfl + l]. */
static void eri_gen(const Shell *si, const Shell *sj, const Shell *sk, const Shell *sl,
                    double *out)
{
    int li = si->l, lj = sj->l, lk = sk->l, ll = sl->l;
    int nfi = ncart(li), nfj = ncart(lj), nfk = ncart(lk), nfl = ncart(ll);
    int lij = li + lj, lkl = lk + ll;
    int nroots = (lij + lkl) / 2 + 1;
    const double *ri = si->R, *rj = sj->R, *rk = sk->R, *rl = sl->R;
    double rij[3], rkl[3];
    for (int d = 0; d < 3; d++) { rij[d] = rj[d] - ri[d]; rkl[d] = rl[d] - rk[d]; }
    double rr_ij = rij[0]*rij[0] + rij[1]*rij[1] + rij[2]*rij[2];
    double rr_kl = rkl[0]*rkl[0] + rkl[1]*rkl[1] + rkl[2]*rkl[2];
    int ip[3][36], jp[3][36], kp[3][36], lp[3][36];
    cart_pow(li, ip[0], ip[1], ip[2]);
    cart_pow(lj, jp[0], jp[1], jp[2]);
    cart_pow(lk, kp[0], kp[1], kp[2]);
    cart_pow(ll, lp[0], lp[1], lp[2]);
    int nout = nfi * nfj * nfk * nfl;
    memset(out, 0, sizeof(double) * nout);

    /* G[d][a][c] for a<=lij, c<=lkl; g[d][i][j][k][l] after HRR */
    enum { A1 = 2*MAXL+3 };
    static __thread double G[3][A1][A1];
    static __thread double H1[3][A1][MAXL+2][A1];        /* [i][j][c] */
    static __thread double g[3][MAXL+2][MAXL+2][MAXL+2][MAXL+2];
    double u[MAXROOTS], w[MAXROOTS];

    for (int pi = 0; pi < si->nprim; pi++)
    for (int pj = 0; pj < sj->nprim; pj++) {
        double ai = si->exps[pi], aj = sj->exps[pj];
        double aij = ai + aj;
        double Kab = exp(-ai * aj / aij * rr_ij);
        double cicj = si->coefs[pi] * sj->coefs[pj] * Kab;
        double P[3];
        for (int d = 0; d < 3; d++) P[d] = ri[d] + rij[d] * aj / aij;
        for (int pk = 0; pk < sk->nprim; pk++)
        for (int pl = 0; pl < sl->nprim; pl++) {
            double ak = sk->exps[pk], al = sl->exps[pl];
            double akl = ak + al;
            double Kcd = exp(-ak * al / akl * rr_kl);
            double ckcl = sk->coefs[pk] * sl->coefs[pl] * Kcd;
            double Q[3], PQ[3];
            for (int d = 0; d < 3; d++) { Q[d] = rk[d] + rkl[d] * al / akl; PQ[d] = P[d] - Q[d]; }
            double rr = PQ[0]*PQ[0] + PQ[1]*PQ[1] + PQ[2]*PQ[2];
            double theta = aij * akl / (aij + akl);
            double fac0 = PI_FAC * cicj * ckcl / (aij * akl * sqrt(aij + akl));
            for (int part = 0; part < 2; part++) {
            double wpart = part == 0 ? g_w_full : g_w_lr;
            if (wpart == 0.) continue;
            double fac = fac0 * wpart;
            if (part == 0) {
                oracle_rys_roots(nroots, theta * rr, u, w);
            } else {
                double aw = g_omega * g_omega / (g_omega * g_omega + theta);
                oracle_rys_roots(nroots, theta * rr * aw, u, w);
                for (int ir = 0; ir < nroots; ir++) { u[ir] *= aw; w[ir] *= sqrt(aw); }
            }
            for (int ir = 0; ir < nroots; ir++) {
                double rt = u[ir];
                double rt_aa = rt / (aij + akl);
                double rt_aij = rt_aa * akl, rt_akl = rt_aa * aij;
                double b10 = .5 / aij * (1 - rt_aij);
                double b01 = .5 / akl * (1 - rt_akl);
                double b00 = .5 * rt_aa;
                for (int d = 0; d < 3; d++) {
                    double c0 = rij[d] * aj / aij - rt_aij * PQ[d];
                    double cp = rkl[d] * al / akl + rt_akl * PQ[d];
                    G[d][0][0] = d == 2 ? w[ir] * fac : 1.;
                    if (lij > 0) G[d][1][0] = c0 * G[d][0][0];
                    for (int a = 1; a < lij; a++)
                        G[d][a+1][0] = c0 * G[d][a][0] + a * b10 * G[d][a-1][0];
                    for (int a = 0; a <= lij; a++) {
                        if (lkl > 0) {
                            G[d][a][1] = cp * G[d][a][0];
                            if (a > 0) G[d][a][1] += a * b00 * G[d][a-1][0];
                        }
                        for (int c = 1; c < lkl; c++) {
                            G[d][a][c+1] = cp * G[d][a][c] + c * b01 * G[d][a][c-1];
                            if (a > 0) G[d][a][c+1] += a * b00 * G[d][a-1][c];
                        }
                    }
                    /* HRR on the bra: (a, b+1) = (a+1, b) - (Rj - Ri)(a, b) */
                    for (int c = 0; c <= lkl; c++) {
                        for (int a = 0; a <= lij; a++) H1[d][a][0][c] = G[d][a][c];
                        for (int b = 0; b < lj; b++)
                            for (int a = 0; a <= lij - b - 1; a++)
                                H1[d][a][b+1][c] = H1[d][a+1][b][c] - rij[d] * H1[d][a][b][c];
                    }
                    /* HRR on the ket */
                    for (int a = 0; a <= li; a++)
                    for (int b = 0; b <= lj; b++) {
                        double tmp[A1][MAXL+2];
                        for (int c = 0; c <= lkl; c++) tmp[c][0] = H1[d][a][b][c];
                        for (int e = 0; e < ll; e++)
                            for (int c = 0; c <= lkl - e - 1; c++)
                                tmp[c][e+1] = tmp[c+1][e] - rkl[d] * tmp[c][e];
                        for (int c = 0; c <= lk; c++)
                            for (int e = 0; e <= ll; e++)
                                g[d][a][b][c][e] = tmp[c][e];
                    }
                }
                int n = 0;
                for (int i = 0; i < nfi; i++)
                for (int j = 0; j < nfj; j++)
                for (int k = 0; k < nfk; k++)
                for (int l = 0; l < nfl; l++, n++) {
                    out[n] += g[0][ip[0][i]][jp[0][j]][kp[0][k]][lp[0][l]]
                            * g[1][ip[1][i]][jp[1][j]][kp[1][k]][lp[1][l]]
                            * g[2][ip[2][i]][jp[2][j]][kp[2][k]][lp[2][l]];
                }
            }
            }
        }
    }
}

/* Cartesian (ij|kl) block of four sorted shells (one contraction each),
 * layout out[((i*nfj + j)*nfk + k)*nfl + l]. */
void oracle_eri_shell(const int *bas, const double *env,
                      int ish, int jsh, int ksh, int lsh, double *out)
{
    Shell a = shell_of(bas, env, ish), b = shell_of(bas, env, jsh);
    Shell c = shell_of(bas, env, ksh), d = shell_of(bas, env, lsh);
    eri_gen(&a, &b, &c, &d, out);
}

/* index of the Cartesian component (x, y, z) of a shell of l = x + y + z */
static inline int cart_index(int x, int y, int z)
{
    int l = x + y + z;
    return (l - x) * (l - x + 1) / 2 + (l - x - y);
}

/*
 * First derivatives of the two-electron energy per atom,
 *   ejk[A] = d/dR_A sum over unique quartets (ij|kl) Gamma_ijkl
 * with the two-particle density and factors of RYS_per_atom_jk_ip1
 * (gpu4pyscf/lib/gvhf-rys/rys_contract_jk_ip1.cu:523-577, 1789-1797):
 * Gamma = fac_sym (4 j_factor D_ji D_lk - 2 k_factor' (D_jk D_li + D_jl D_ki)),
 * summed over the n_dm (1 or 2) densities as there (k_factor' = k_factor/2
 * for one density), fac_sym = 1/2 for ish == jsh, ksh == lsh, ij == kl.
 * Derivatives: d/dA_x phi_a = 2 alpha phi_{a+1x} - a_x phi_{a-1x} at the
 * primitive level, evaluated as the ERI blocks of a shell of l+1 with
 * coefficients 2 alpha c and of l-1 with c; centre l by translational
 * invariance.  Every quartet of the sorted basis, no screening.  atom_of:
 * atom of each shell; ejk: natm x 3, accumulated.
 */
void oracle_ejk_ip1(int n_dm, int nao, const double *dms, const int *bas, int nbas,
                    const double *env, const int *ao_loc, const int *atom_of, int natm,
                    double j_factor, double k_factor, double *ejk)
{
    size_t nao2 = (size_t)nao * nao;
    if (n_dm == 1) k_factor *= .5;
    const double jf = 4. * j_factor, kf = -2. * k_factor;
    int npair = nbas * (nbas + 1) / 2;
#pragma omp parallel
    {
        double *acc = calloc((size_t)natm * 3, sizeof(double));
        double *bp = malloc(sizeof(double) * 36*36*36*36);
        double *bm = malloc(sizeof(double) * 36*36*36*36);
        double *c2a = malloc(sizeof(double) * 64);
#pragma omp for schedule(dynamic, 1)
        for (int pij = 0; pij < npair; pij++) {
            int ish = (int)((sqrt(8. * pij + 1.) - 1.) / 2.);
            while (ish * (ish + 1) / 2 > pij) ish--;
            while ((ish + 1) * (ish + 2) / 2 <= pij) ish++;
            int jsh = pij - ish * (ish + 1) / 2;
            for (int ksh = 0; ksh <= ish; ksh++)
            for (int lsh = 0; lsh <= ksh; lsh++) {
                int pkl = ksh * (ksh + 1) / 2 + lsh;
                if (pkl > pij) continue;
                Shell sh[4] = {shell_of(bas, env, ish), shell_of(bas, env, jsh),
                               shell_of(bas, env, ksh), shell_of(bas, env, lsh)};
                double fac = 1.;
                if (ish == jsh) fac *= .5;
                if (ksh == lsh) fac *= .5;
                if (pij == pkl) fac *= .5;
                int off[4] = {ao_loc[ish], ao_loc[jsh], ao_loc[ksh], ao_loc[lsh]};
                int nf[4], pw[4][3][36];
                for (int c = 0; c < 4; c++) {
                    nf[c] = ncart(sh[c].l);
                    cart_pow(sh[c].l, pw[c][0], pw[c][1], pw[c][2]);
                }
                int atoms[4] = {atom_of[ish], atom_of[jsh], atom_of[ksh], atom_of[lsh]};
                double grad[3][3] = {{0}};          /* centre (i, j, k) x direction */
                for (int cen = 0; cen < 3; cen++) {
                    /* blocks with the centre's shell raised and lowered */
                    Shell up = sh[cen], dn = sh[cen];
                    up.l += 1;
                    for (int p = 0; p < up.nprim; p++) c2a[p] = 2. * up.exps[p] * up.coefs[p];
                    up.coefs = c2a;
                    Shell q[4] = {sh[0], sh[1], sh[2], sh[3]};
                    q[cen] = up;
                    eri_gen(&q[0], &q[1], &q[2], &q[3], bp);
                    int has_dn = sh[cen].l > 0;
                    if (has_dn) {
                        dn.l -= 1;
                        q[cen] = dn;
                        eri_gen(&q[0], &q[1], &q[2], &q[3], bm);
                    }
                    int nfu[4] = {nf[0], nf[1], nf[2], nf[3]}, nfd[4] = {nf[0], nf[1], nf[2], nf[3]};
                    nfu[cen] = ncart(sh[cen].l + 1);
                    nfd[cen] = has_dn ? ncart(sh[cen].l - 1) : 1;
                    for (int i = 0; i < nf[0]; i++)
                    for (int j = 0; j < nf[1]; j++)
                    for (int k = 0; k < nf[2]; k++)
                    for (int l = 0; l < nf[3]; l++) {
                        size_t I = off[0] + i, J = off[1] + j, K = off[2] + k, L = off[3] + l;
                        double gam = 0.;
                        for (int id = 0; id < n_dm; id++) {
                            const double *D = dms + id * nao2;
                            gam += kf * (D[J*nao+K] * D[L*nao+I] + D[J*nao+L] * D[K*nao+I]);
                        }
                        double dji = 0., dlk = 0.;
                        for (int id = 0; id < n_dm; id++) {
                            dji += dms[id * nao2 + J*nao+I]; dlk += dms[id * nao2 + L*nao+K];
                        }
                        gam += jf * dji * dlk;
                        gam *= fac;
                        int idx[4] = {i, j, k, l};
                        for (int x = 0; x < 3; x++) {
                            int p0 = pw[cen][0][idx[cen]], p1 = pw[cen][1][idx[cen]],
                                p2 = pw[cen][2][idx[cen]];
                            int pp[3] = {p0, p1, p2};
                            pp[x] += 1;
                            int iu[4] = {i, j, k, l}, id_[4] = {i, j, k, l};
                            iu[cen] = cart_index(pp[0], pp[1], pp[2]);
                            double v = bp[((iu[0]*nfu[1] + iu[1])*nfu[2] + iu[2])*nfu[3] + iu[3]];
                            pp[x] -= 2;
                            if (pp[x] >= 0) {
                                id_[cen] = cart_index(pp[0], pp[1], pp[2]);
                                v -= (pp[x] + 1) *
                                     bm[((id_[0]*nfd[1] + id_[1])*nfd[2] + id_[2])*nfd[3] + id_[3]];
                            }
                            grad[cen][x] += gam * v;
                        }
                    }
                }
                for (int x = 0; x < 3; x++) {
                    acc[atoms[0]*3+x] += grad[0][x];
                    acc[atoms[1]*3+x] += grad[1][x];
                    acc[atoms[2]*3+x] += grad[2][x];
                    acc[atoms[3]*3+x] -= grad[0][x] + grad[1][x] + grad[2][x];
                }
            }
        }
#pragma omp critical
        for (int n = 0; n < natm * 3; n++) ejk[n] += acc[n];
        free(acc); free(bp); free(bm); free(c2a);
    }
}

/* q[p] = max |(ij|ij)| over the Cartesian components of pair p = ish*nbas+jsh */
void oracle_pair_diag_max(const int *bas, const double *env, int nbas,
                          const uint32_t *pairs, int npairs, double *qmax)
{
#pragma omp parallel
    {
        double *buf = malloc(sizeof(double) * 36*36*36*36);
#pragma omp for schedule(dynamic, 4)
        for (int p = 0; p < npairs; p++) {
            int ish = pairs[p] / nbas, jsh = pairs[p] % nbas;
            oracle_eri_shell(bas, env, ish, jsh, ish, jsh, buf);
            int nfi = ncart(bas[ish*BAS_SLOTS+ANG_OF]);
            int nfj = ncart(bas[jsh*BAS_SLOTS+ANG_OF]);
            int nf = nfi * nfj;
            double m = 0.;
            for (int n = 0; n < nf*nf; n++) m = fmax(m, fabs(buf[n]));
            qmax[p] = m;
        }
        free(buf);
    }
}

/* Reference predicate (create_tasks.cu:155-170), float32 arithmetic. */
/* Which density blocks decide survival (oracle_set_screen): 0 the six
 * blocks of a J/K build (create_tasks.cu:155-163, _fill_vjk_tasks), 1 the
 * two Coulomb blocks D_ij, D_kl of a J-only build (create_tasks.cu:222-225,
 * _fill_vj_tasks), 2 the four exchange blocks D_ik, D_jk, D_il, D_jl of a
 * K-only build (create_tasks.cu:86-93, _fill_vk_tasks). */
static int g_screen = 0;

void oracle_set_screen(int screen) { g_screen = screen; }

static inline int survives(float q_ij, float q_kl, float cutoff, float dm_penalty,
                           uint32_t bas_ij, uint32_t bas_kl,
                           int ish, int jsh, int ksh, int lsh, int nbas,
                           const float *dm_cond)
{
    float kl_cutoff = cutoff - q_ij;
    int keep = q_kl + dm_penalty >= kl_cutoff && bas_ij >= bas_kl;
    if (!keep) return 0;
    float d_cutoff = kl_cutoff - q_kl;
    int coulomb = dm_cond[bas_ij] > d_cutoff || dm_cond[bas_kl] > d_cutoff;
    int exchange = dm_cond[ish*nbas+ksh] > d_cutoff || dm_cond[jsh*nbas+ksh] > d_cutoff ||
                   dm_cond[ish*nbas+lsh] > d_cutoff || dm_cond[jsh*nbas+lsh] > d_cutoff;
    return g_screen == 1 ? coulomb : (g_screen == 2 ? exchange : (coulomb || exchange));
}

/*
 * J/K on the sorted Cartesian basis before the hermi post-processing, i.e.
 * what the reference's kernels accumulate into vj/vk (rys_contract_jk.cu:
 * 291-330).  Pair lists: group pair (gi, gj), gj <= gi, has its pairs at
 * pairs[pair_off[gi*(gi+1)/2+gj] : pair_off[..+1]] with q_cond alongside.
 * Returns the number of surviving quartets; if quartets != NULL the first
 * max_quartets survivors (bas_ij, bas_kl) are written there in task order.
 * ij_stride > 1 evaluates only every ij_stride-th ij pair of each task (a
 * bounded sample used to time the CPU baseline).
 */
long oracle_jk(int n_dm, int nao, const double *dms, double *vj, double *vk,
               const int *bas, int nbas, const double *env, const int *ao_loc,
               int ngroups, const int *pair_off, const uint32_t *pairs,
               const float *q_cond, const float *dm_cond, float cutoff,
               float dm_penalty, int with_j, int with_k,
               uint32_t *quartets, long max_quartets, int ij_stride)
{
    size_t nao2 = (size_t)nao * nao;
    long total = 0;
    /* enumerate tasks in the reference order */
    int ntask = 0;
    int (*tasks)[4] = malloc(sizeof(int[4]) * ngroups*ngroups*ngroups*ngroups);
    for (int i = 0; i < ngroups; i++)
    for (int j = 0; j <= i; j++)
    for (int k = 0; k <= i; k++)
    for (int l = 0; l <= k; l++) {
        tasks[ntask][0] = i; tasks[ntask][1] = j;
        tasks[ntask][2] = k; tasks[ntask][3] = l; ntask++;
    }
    long nq = 0;
    if (quartets) {
        /* serial enumeration of the survivor list (order is deterministic) */
        for (int t = 0; t < ntask; t++) {
            int gi = tasks[t][0], gj = tasks[t][1], gk = tasks[t][2], gl = tasks[t][3];
            int pij0 = pair_off[gi*(gi+1)/2+gj], pij1 = pair_off[gi*(gi+1)/2+gj+1];
            int pkl0 = pair_off[gk*(gk+1)/2+gl], pkl1 = pair_off[gk*(gk+1)/2+gl+1];
            for (int pij = pij0; pij < pij1; pij++) {
                uint32_t bij = pairs[pij];
                int ish = bij / nbas, jsh = bij % nbas;
                for (int pkl = pkl0; pkl < pkl1; pkl++) {
                    uint32_t bkl = pairs[pkl];
                    int ksh = bkl / nbas, lsh = bkl % nbas;
                    if (survives(q_cond[pij], q_cond[pkl], cutoff, dm_penalty,
                                 bij, bkl, ish, jsh, ksh, lsh, nbas, dm_cond)) {
                        if (nq < max_quartets) {
                            quartets[2*nq] = bij; quartets[2*nq+1] = bkl;
                        }
                        nq++;
                    }
                }
            }
        }
    }
    if (!vj && !vk) {
        free(tasks);
        return nq;
    }

#pragma omp parallel reduction(+:total)
    {
        double *pvj = calloc(n_dm * nao2, sizeof(double));
        double *pvk = calloc(n_dm * nao2, sizeof(double));
        double *buf = malloc(sizeof(double) * 36*36*36*36);
        for (int t = 0; t < ntask; t++) {
            int gi = tasks[t][0], gj = tasks[t][1], gk = tasks[t][2], gl = tasks[t][3];
            int pij0 = pair_off[gi*(gi+1)/2+gj], pij1 = pair_off[gi*(gi+1)/2+gj+1];
            int pkl0 = pair_off[gk*(gk+1)/2+gl], pkl1 = pair_off[gk*(gk+1)/2+gl+1];
#pragma omp for schedule(dynamic, 1) nowait
            for (int pij = pij0; pij < pij1; pij += ij_stride) {
                uint32_t bij = pairs[pij];
                int ish = bij / nbas, jsh = bij % nbas;
                for (int pkl = pkl0; pkl < pkl1; pkl++) {
                    uint32_t bkl = pairs[pkl];
                    int ksh = bkl / nbas, lsh = bkl % nbas;
                    if (!survives(q_cond[pij], q_cond[pkl], cutoff, dm_penalty,
                                  bij, bkl, ish, jsh, ksh, lsh, nbas, dm_cond))
                        continue;
                    total++;
                    oracle_eri_shell(bas, env, ish, jsh, ksh, lsh, buf);
                    double fac = 1.;
                    if (ish == jsh) fac *= .5;
                    if (ksh == lsh) fac *= .5;
                    if (bij == bkl) fac *= .5;
                    int i0 = ao_loc[ish], j0 = ao_loc[jsh], k0 = ao_loc[ksh], l0 = ao_loc[lsh];
                    int nfi = ao_loc[ish+1]-i0, nfj = ao_loc[jsh+1]-j0;
                    int nfk = ao_loc[ksh+1]-k0, nfl = ao_loc[lsh+1]-l0;
                    for (int id = 0; id < n_dm; id++) {
                        const double *D = dms + id * nao2;
                        double *J = pvj + id * nao2, *K = pvk + id * nao2;
                        int n = 0;
                        for (int i = i0; i < i0+nfi; i++)
                        for (int j = j0; j < j0+nfj; j++)
                        for (int k = k0; k < k0+nfk; k++)
                        for (int l = l0; l < l0+nfl; l++, n++) {
                            double v = buf[n] * fac;
                            if (with_k) {
                                K[i*nao+l] += v * D[j*nao+k];
                                K[i*nao+k] += v * D[j*nao+l];
                                K[j*nao+l] += v * D[i*nao+k];
                                K[j*nao+k] += v * D[i*nao+l];
                            }
                            if (with_j) {
                                J[k*nao+l] += v * D[i*nao+j];
                                J[i*nao+j] += v * D[k*nao+l];
                            }
                        }
                    }
                }
            }
        }
#pragma omp critical
        {
            for (size_t n = 0; n < n_dm * nao2; n++) {
                if (vj) vj[n] += pvj[n];
                if (vk) vk[n] += pvk[n];
            }
        }
        free(pvj); free(pvk); free(buf);
    }
    free(tasks);
    return total;
}

/* SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014; constants of
 * Vigna's splitmix64.c), the per-quartet term of a survivor set checksum. */
static inline uint64_t quartet_mix(uint32_t bij, uint32_t bkl)
{
    uint64_t z = ((uint64_t)bij << 32) | bkl;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

uint64_t oracle_splitmix64_mix(uint64_t z)
{
    return quartet_mix((uint32_t)(z >> 32), (uint32_t)z);
}

/*
 * Surviving quartets of every task (same enumeration and predicate as
 * oracle_jk, no integrals): counts[t] = number of survivors and
 * hashes[t] = sum mod 2^64 of quartet_mix(bas_ij, bas_kl) over them, for
 * t over all (i>=j, k<=i, l<=k) group quadruples in oracle_jk's order.  Every
 * kl candidate is tested (no early exit on the sorted bins).
 */
void oracle_quartet_sets(int nbas, int ngroups, const int *pair_off,
                         const uint32_t *pairs, const float *q_cond,
                         const float *dm_cond, float cutoff, float dm_penalty,
                         int64_t *counts, uint64_t *hashes)
{
    int t = 0;
    for (int i = 0; i < ngroups; i++)
    for (int j = 0; j <= i; j++)
    for (int k = 0; k <= i; k++)
    for (int l = 0; l <= k; l++, t++) {
        int pij0 = pair_off[i*(i+1)/2+j], pij1 = pair_off[i*(i+1)/2+j+1];
        int pkl0 = pair_off[k*(k+1)/2+l], pkl1 = pair_off[k*(k+1)/2+l+1];
        int64_t n = 0;
        uint64_t h = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+:n, h)
        for (int pij = pij0; pij < pij1; pij++) {
            uint32_t bij = pairs[pij];
            int ish = bij / nbas, jsh = bij % nbas;
            for (int pkl = pkl0; pkl < pkl1; pkl++) {
                uint32_t bkl = pairs[pkl];
                int ksh = bkl / nbas, lsh = bkl % nbas;
                if (survives(q_cond[pij], q_cond[pkl], cutoff, dm_penalty,
                             bij, bkl, ish, jsh, ksh, lsh, nbas, dm_cond)) {
                    n++;
                    h += quartet_mix(bij, bkl);
                }
            }
        }
        counts[t] = n;
        hashes[t] = h;
    }
}
