/*
 * ep_oracle.c -- CPU restatement of ElastoPlasm.jl v0.4.2's explicit MPM step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Restated kernel by kernel, in the
 * reference's loop and summation order; every function cites the Julia lines
 * it follows.  Compile with -ffp-contract=off so no multiply-add is fused,
 * as in Julia without @fastmath.  With or_set_threads(1) every scatter runs
 * in particle order; with more threads the scatters are OpenMP atomics, the
 * way KernelAbstractions' CPU backend runs `@atom` over Julia threads.
 *
 * Deliberate deviations, each a rounding-level or dead-code matter:
 *  - sym_eig: cyclic Jacobi on 0.5*(b+b') instead of LAPACK syevr/geev.
 *  - det / `\`: right-looking LU with partial pivoting instead of OpenBLAS
 *    getrf (det of a triangular matrix is the diagonal product, as Julia's
 *    generic det does).
 *  - Del*v: sequential row sums instead of OpenBLAS gemv.
 *  - J2 (src/.../retmap/J2.jl:21-75) is restated from its intent: s.cᵣ/s.c₀/
 *    s.τᵢ field paths, ϵpII[1,p] (not ϵpII[p,1]), two-argument kernel.
 */
#include "ep_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

int or_version(void) { return 1; }

void or_set_threads(int n) {
    g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}

#define PAR_FOR _Pragma("omp parallel for schedule(static) if(g_threads > 1)")

static inline void atom_add(double *a, double v) {
    if (g_threads > 1) {
#pragma omp atomic
        *a += v;
    } else {
        *a += v;
    }
}
/* Julia's fld for floats: floor of the exact quotient (computed through the
 * exact remainder), not floor of the rounded quotient. */
static inline double jfld(double a, double b) {
    double r = fmod(a, b);
    double d = round((a - r) / b);
    if (r != 0.0 && ((r < 0.0) != (b < 0.0))) d -= 1.0;
    return d;
}
static inline double jsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

/* ------------------------------------------------------------------------- */
/* small dense linear algebra                                                */
/* ------------------------------------------------------------------------- */

/* Cyclic Jacobi eigen-decomposition of a symmetric d x d matrix (column-major
 * A, only read), eigenvalues lam[d], eigenvectors V[:,k] column-major.
 * Stands in for LinearAlgebra.eigen (elast.jl:51, DP.jl:65, J2.jl:44). */
void or_sym_eig(int d, const double *A, double *lam, double *V) {
    double a[3][3];
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++)
            a[i][j] = 0.5 * (A[i + d * j] + A[j + d * i]);
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; sweep++) {
        int rotated = 0;
        for (int p = 0; p < d - 1; p++) {
            for (int q = p + 1; q < d; q++) {
                double apq = a[p][q];
                if (apq == 0.0) continue;
                if (fabs(apq) <= 1e-18 * (fabs(a[p][p]) + fabs(a[q][q]))) {
                    a[p][q] = a[q][p] = 0.0;
                    continue;
                }
                rotated = 1;
                double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                double t;
                if (fabs(theta) > 1e150) t = 0.5 / theta;
                else t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < d; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < d; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                a[p][q] = a[q][p] = 0.0;
                for (int k = 0; k < d; k++) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
        }
        if (!rotated) break;
    }
    for (int k = 0; k < d; k++) {
        lam[k] = a[k][k];
        for (int i = 0; i < d; i++) V[i + d * k] = v[i][k];
    }
}

/* M = V diag(f) V' (column-major), the n*diagm(f)*n' of elast.jl:52, DP.jl:66 */
static void vdvt(int d, const double *V, const double *f, double *M) {
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            double s = 0.0;
            for (int k = 0; k < d; k++) s += (V[i + d * k] * f[k]) * V[j + d * k];
            M[i + d * j] = s;
        }
}

/* LU with partial pivoting, in place; returns swap count, -1 if singular */
static int lu_fact(int n, double *A, int *piv) {
    int swaps = 0, singular = 0;
    for (int k = 0; k < n; k++) {
        int pk = k;
        double mx = fabs(A[k + n * k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i + n * k]) > mx) { mx = fabs(A[i + n * k]); pk = i; }
        piv[k] = pk;
        if (pk != k) {
            swaps++;
            for (int j = 0; j < n; j++) {
                double t = A[k + n * j]; A[k + n * j] = A[pk + n * j]; A[pk + n * j] = t;
            }
        }
        double akk = A[k + n * k];
        if (akk == 0.0) { singular = 1; continue; }
        double r = 1.0 / akk;
        for (int i = k + 1; i < n; i++) A[i + n * k] *= r;
        for (int j = k + 1; j < n; j++)
            for (int i = k + 1; i < n; i++) A[i + n * j] -= A[i + n * k] * A[k + n * j];
    }
    return singular ? -1 : swaps;
}

/* det (deform.jl:17-18) */
double or_det(int d, const double *A) {
    int triu = 1, tril = 1;
    for (int j = 0; j < d; j++)
        for (int i = 0; i < d; i++) {
            if (i > j && A[i + d * j] != 0.0) triu = 0;
            if (i < j && A[i + d * j] != 0.0) tril = 0;
        }
    if (triu || tril) {
        double P = 1.0;
        for (int i = 0; i < d; i++) P *= A[i + d * i];
        return P;
    }
    double M[9];
    int piv[3];
    memcpy(M, A, sizeof(double) * d * d);
    int s = lu_fact(d, M, piv);
    if (s < 0) return 0.0;
    double P = 1.0;
    for (int i = 0; i < d; i++) P *= M[i + d * i];
    return (s & 1) ? -P : P;
}

/* x = A \ b (DP.jl:64, J2.jl:43) */
void or_lu_solve(int n, const double *A, const double *b, double *x) {
    double M[36];
    int piv[6];
    memcpy(M, A, sizeof(double) * n * n);
    lu_fact(n, M, piv);
    double y[6];
    for (int i = 0; i < n; i++) y[i] = b[i];
    for (int k = 0; k < n; k++)
        if (piv[k] != k) { double t = y[k]; y[k] = y[piv[k]]; y[piv[k]] = t; }
    for (int i = 0; i < n; i++) {
        double s = y[i];
        for (int k = 0; k < i; k++) s -= M[i + n * k] * y[k];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int k = i + 1; k < n; k++) s -= M[i + n * k] * x[k];
        x[i] = s / M[i + n * i];
    }
}

/* inv (Bij via p2n.jl:249, 273): columns of the identity solved through the
 * same LU with partial pivoting */
void or_inv(int d, const double *A, double *Ai) {
    for (int j = 0; j < d; j++) {
        double e[3] = {0.0, 0.0, 0.0}, x[3];
        e[j] = 1.0;
        or_lu_solve(d, A, e, x);
        for (int i = 0; i < d; i++) Ai[i + d * j] = x[i];
    }
}

/* y = Del * x, Del [n,n] column-major */
static void matvec(int n, const double *A, const double *x, double *y) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += A[i + n * j] * x[j];
        y[i] = s;
    }
}

/* C = A*B for d x d column-major (Julia matmul2x2!/matmul3x3! term order) */
static void matmul(int d, const double *A, const double *B, double *C) {
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            double s = A[i] * B[d * j];
            for (int k = 1; k < d; k++) s += A[i + d * k] * B[k + d * j];
            C[i + d * j] = s;
        }
}
/* C = A*B' */
static void matmul_bt(int d, const double *A, const double *B, double *C) {
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            double s = A[i] * B[j];
            for (int k = 1; k < d; k++) s += A[i + d * k] * B[j + d * k];
            C[i + d * j] = s;
        }
}

/* mutate(ϵ,Χ,:voigt) (elast.jl:24-29) */
static void to_voigt(int d, const double *e, double X, double *o) {
    if (d == 2) { o[0] = e[0]; o[1] = e[3]; o[2] = X * e[2]; }
    else { o[0] = e[0]; o[1] = e[4]; o[2] = e[8]; o[3] = X * e[7]; o[4] = X * e[6]; o[5] = X * e[3]; }
}
/* mutate(ϵ,Χ,:tensor) (elast.jl:15-23) */
static void to_tensor(int d, const double *e, double X, double *o) {
    if (d == 2) { o[0] = e[0]; o[1] = X * e[2]; o[2] = X * e[2]; o[3] = e[1]; }
    else {
        o[0] = e[0];     o[3] = X * e[5]; o[6] = X * e[4];
        o[1] = X * e[5]; o[4] = e[1];     o[7] = X * e[3];
        o[2] = X * e[4]; o[5] = X * e[3]; o[8] = e[2];
    }
}

/* ------------------------------------------------------------------------- */
/* shape functions                                                           */
/* ------------------------------------------------------------------------- */

/* t1..t4_ϕ∂ϕ (bsmpm.jl:1-62); type 1..4 */
void or_bsmpm_1d(double xi, int type, double *phi, double *dphi) {
    double f = 0.0, df = 0.0;
    double x2 = xi * xi, x3 = x2 * xi;
    if (type == 1) {
        if (-2.0 <= xi && xi <= -1.0) { f = (1.0 / 6.0) * x3 + x2 + 2.0 * xi + 4.0 / 3.0; df = (3.0 / 6.0) * x2 + 2.0 * xi + 2.0; }
        else if (-1.0 <= xi && xi <= 0.0) { f = (-1.0 / 6.0) * x3 + xi + 1.0; df = (-3.0 / 6.0) * x2 + 1.0; }
        else if (0.0 <= xi && xi <= 1.0) { f = (1.0 / 6.0) * x3 - xi + 1.0; df = (3.0 / 6.0) * x2 - 1.0; }
        else if (1.0 <= xi && xi <= 2.0) { f = (-1.0 / 6.0) * x3 + x2 - 2.0 * xi + 4.0 / 3.0; df = (-3.0 / 6.0) * x2 + 2.0 * xi - 2.0; }
    } else if (type == 2) {
        if (-1.0 <= xi && xi <= 0.0) { f = (-1.0 / 3.0) * x3 - x2 + 2.0 / 3.0; df = (-3.0 / 3.0) * x2 - 2.0 * xi; }
        else if (0.0 <= xi && xi <= 1.0) { f = (1.0 / 2.0) * x3 - x2 + 2.0 / 3.0; df = (3.0 / 2.0) * x2 - 2.0 * xi; }
        else if (1.0 <= xi && xi <= 2.0) { f = (-1.0 / 6.0) * x3 + x2 - 2.0 * xi + 4.0 / 3.0; df = (-3.0 / 6.0) * x2 + 2.0 * xi - 2.0; }
    } else if (type == 3) {
        if (-2.0 <= xi && xi <= -1.0) { f = (1.0 / 6.0) * x3 + x2 + 2.0 * xi + 4.0 / 3.0; df = (3.0 / 6.0) * x2 + 2.0 * xi + 2.0; }
        else if (-1.0 <= xi && xi <= 0.0) { f = (-1.0 / 2.0) * x3 - x2 + 2.0 / 3.0; df = (-3.0 / 2.0) * x2 - 2.0 * xi; }
        else if (0.0 <= xi && xi <= 1.0) { f = (1.0 / 2.0) * x3 - x2 + 2.0 / 3.0; df = (3.0 / 2.0) * x2 - 2.0 * xi; }
        else if (1.0 <= xi && xi <= 2.0) { f = (-1.0 / 6.0) * x3 + x2 - 2.0 * xi + 4.0 / 3.0; df = (-3.0 / 6.0) * x2 + 2.0 * xi - 2.0; }
    } else if (type == 4) {
        if (-2.0 <= xi && xi <= -1.0) { f = (1.0 / 6.0) * x3 + x2 + 2.0 * xi + 4.0 / 3.0; df = (3.0 / 6.0) * x2 + 2.0 * xi + 2.0; }
        else if (-1.0 <= xi && xi <= 0.0) { f = (-1.0 / 2.0) * x3 - x2 + 2.0 / 3.0; df = (-3.0 / 2.0) * x2 - 2.0 * xi; }
        else if (0.0 <= xi && xi <= 1.0) { f = (1.0 / 3.0) * x3 - x2 + 2.0 / 3.0; df = (3.0 / 3.0) * x2 - 2.0 * xi; }
    }
    *phi = f;
    *dphi = df;
}

/* ϕ∂ϕ boundary classification by floating-point equality (bsmpm.jl:63-78);
 * returns 0 for "Invalid position". */
static int bsmpm_type(double xn, const double *xB, int dim, int i, double dx) {
    double lo = xB[i], hi = xB[i + dim];
    if (xn == lo) return 1;
    if (xn == lo + dx) return 2;
    if (lo + dx < xn && xn < hi - dx) return 3;
    if (xn == hi - dx) return 4;
    if (xn == hi) return 1;
    return 0;
}

/* S∂S (gimpm.jl:1-14) */
void or_gimp_1d(double dx, double h, double lp, double *S, double *dS) {
    double s = 0.0, ds = 0.0, ax = fabs(dx);
    if (ax < lp) {
        s = 1.0 - ((4.0 * (dx * dx) + (2.0 * lp) * (2.0 * lp)) / (8.0 * h * lp));
        ds = -((8.0 * dx) / (8.0 * h * lp));
    } else if ((ax >= lp) && (ax <= (h - lp))) {
        s = 1.0 - (ax / h);
        ds = jsign(dx) * (-1.0 / h);
    } else if ((ax >= (h - lp)) && (ax < (h + lp))) {
        double r = h + lp - ax;
        s = (r * r) / (4.0 * h * lp);
        ds = -jsign(dx) * r / (2.0 * h * lp);
    }
    *S = s;
    *dS = ds;
}

/* N∂N (smpm.jl:1-10) */
void or_smpm_1d(double dx, double h, double *N, double *dN) {
    if (-h < dx && dx <= 0.0) { *N = 1.0 + dx / h; *dN = 1.0 / h; }
    else if (0.0 < dx && dx <= h) { *N = 1.0 - dx / h; *dN = -1.0 / h; }
    else { *N = 0.0; *dN = 0.0; }
}

#define PHI(k, p, c) mp->phi[(k) + nn * ((p) + nmp * (int64_t)(c))]

/* p2e2n_{2,3}d (tplgy.jl:43-96) */
static void p2e2n(or_point *mp, const or_mesh *me) {
    const int dim = me->dim, nn = me->nn;
    const int64_t nmp = mp->nmp;
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++) {
        double el_f[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < dim; i++) {
            double e = jfld(mp->x[i + dim * p] - me->x0[i], me->h[i]);
            double hi = (double)(me->nel[i] - 1);
            el_f[i] = e < 0.0 ? 0.0 : (e > hi ? hi : e);
        }
        double elr;
        if (dim == 2) elr = 1.0 + el_f[1] + (double)me->nel[1] * el_f[0];
        else elr = 1.0 + el_f[2] + (double)me->nel[2] * el_f[0] + (double)me->nel[2] * (double)me->nel[0] * el_f[1];
        int64_t el = (int64_t)llround(elr);
        for (int k = 0; k < nn; k++) mp->p2n[k + nn * p] = me->e2n[k + nn * (el - 1)];
        mp->p2e[p] = el;
    }
}

/* bsmpm/gimpm/smpm_{2,3}d (bsmpm.jl:95-135, gimpm.jl:31-71, smpm.jl:27-66) */
static int basis(or_point *mp, const or_mesh *me, const or_instr *in) {
    const int dim = me->dim, nn = me->nn;
    const int64_t nmp = mp->nmp;
    int err = 0;
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++) {
        for (int k = 0; k < nn; k++) {
            int64_t no = mp->p2n[k + nn * p];
            if (no == 0) continue;
            no -= 1;
            double f[3] = {0.0, 0.0, 0.0}, df[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < dim; i++) {
                double xp = mp->x[i + dim * p], xn = me->x[i + dim * no], h = me->h[i];
                if (in->basis == OR_BSMPM) {
                    double xi = (xp - xn) / h;
                    int t = bsmpm_type(xn, me->xB, dim, i, h);
                    if (t == 0) {
#pragma omp atomic write
                        err = 1;
                        f[i] = df[i] = 0.0;
                        continue;
                    }
                    double a, b;
                    or_bsmpm_1d(xi, t, &a, &b);
                    f[i] = a;
                    df[i] = b / h;
                } else if (in->basis == OR_GIMPM) {
                    or_gimp_1d(xp - xn, h, mp->l[i + dim * p], &f[i], &df[i]);
                } else {
                    or_smpm_1d(xp - xn, h, &f[i], &df[i]);
                }
            }
            if (dim == 2) {
                PHI(k, p, 0) = f[0] * f[1];
                PHI(k, p, 1) = df[0] * f[1];
                PHI(k, p, 2) = f[0] * df[1];
            } else {
                PHI(k, p, 0) = f[0] * f[1] * f[2];
                PHI(k, p, 1) = df[0] * f[1] * f[2];
                PHI(k, p, 2) = f[0] * df[1] * f[2];
                PHI(k, p, 3) = f[0] * f[1] * df[2];
            }
        }
    }
    return err ? OR_E_DOMAIN : OR_OK;
}

/* Δnp_nd (misc.jl:1-14) */
static void delta_np(or_point *mp, const or_mesh *me) {
    const int dim = me->dim, nn = me->nn;
    const int64_t nmp = mp->nmp;
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++)
        for (int k = 0; k < nn; k++) {
            int64_t no = mp->p2n[k + nn * p];
            if (no == 0) continue;
            no -= 1;
            for (int i = 0; i < dim; i++)
                mp->dnp[k + nn * (i + dim * p)] = me->x[i + dim * no] - mp->x[i + dim * p];
        }
}

/* Dij_nd (misc.jl:15-27): Δnp as Δnp_nd, and Dᵢⱼ *assigned* (not summed)
 * at every existing node, so the last existing node of the stencil wins. */
static void dij_nd(or_point *mp, const or_mesh *me) {
    const int dim = me->dim, nn = me->nn, dd = dim * dim;
    const int64_t nmp = mp->nmp;
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++)
        for (int k = 0; k < nn; k++) {
            int64_t no = mp->p2n[k + nn * p];
            if (no == 0) continue;
            no -= 1;
            double dl[3];
            for (int i = 0; i < dim; i++) {
                dl[i] = me->x[i + dim * no] - mp->x[i + dim * p];
                mp->dnp[k + nn * (i + dim * p)] = dl[i];
            }
            const double N = PHI(k, p, 0);
            for (int j = 0; j < dim; j++)
                for (int i = 0; i < dim; i++) mp->Dij[i + dim * j + dd * p] = N * (dl[i] * dl[j]);
        }
}

/* shpfun (shpfun.jl:74-88) */
int or_shpfun(or_point *mp, const or_mesh *me, const or_instr *in) {
    p2e2n(mp, me);
    memset(mp->phi, 0, sizeof(double) * (size_t)me->nn * mp->nmp * (me->dim + 1));
    int r = basis(mp, me, in);
    if (in->trsfr == OR_TPIC) delta_np(mp, me);
    else if (in->trsfr == OR_APIC) dij_nd(mp, me);
    return r;
}

/* ------------------------------------------------------------------------- */
/* mapsto                                                                    */
/* ------------------------------------------------------------------------- */

/* APIC's B·D⁻¹·Δnp for one node (p2n.jl:243-249, 267-273): D⁻¹ = inv(D)
 * when |det D| > 1e-12, else the identity; Julia evaluates B*D⁻¹*x as
 * B*(D⁻¹*x). */
static void apic_dinv(int dim, const double *D, double *Di, int *ident) {
    *ident = !(fabs(or_det(dim, D)) > 1.0e-12);
    if (!*ident) or_inv(dim, D, Di);
}
static void apic_bw(int dim, const double *B, const double *Di, int ident, const double *x, double *o) {
    double w[3];
    for (int i = 0; i < dim; i++) {
        if (ident) { w[i] = x[i]; continue; }
        double s = Di[i] * x[0];
        for (int j = 1; j < dim; j++) s += Di[i + dim * j] * x[j];
        w[i] = s;
    }
    for (int i = 0; i < dim; i++) {
        double s = B[i] * w[0];
        for (int j = 1; j < dim; j++) s += B[i + dim * j] * w[j];
        o[i] = s;
    }
}

/* transform + std/tpic/apic_{2,3}d_p2n + p2n wrapper (p2n.jl:12-21, 58-102,
 * 140-191, 231-280, 284-296) */
void or_p2n(or_point *mp, or_mesh *me, const double *g, const or_instr *in) {
    const int dim = me->dim, nn = me->nn, nstr = dim == 2 ? 3 : 6;
    const int64_t nmp = mp->nmp, nno = me->nno[dim];
    if (in->deform == OR_FINITE) {
        PAR_FOR
        for (int64_t p = 0; p < nmp; p++) {
            double J = mp->J[p];
            for (int c = 0; c < nstr; c++) mp->sig[c + nstr * p] = mp->tau[c + nstr * p] / J;
        }
    }
    memset(me->m, 0, sizeof(double) * nno);
    memset(me->p, 0, sizeof(double) * dim * nno);
    memset(me->oobf, 0, sizeof(double) * dim * nno);
    const int tpic = in->trsfr == OR_TPIC, apic = in->trsfr == OR_APIC;
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++) {
        double ms = mp->rho[p] * mp->Om[p], Om = mp->Om[p];
        const double *s = mp->sig + nstr * p;
        const double *gv = mp->gradv + dim * dim * p;
        double Di[9];
        int ident = 1;
        if (apic) apic_dinv(dim, mp->Dij + dim * dim * p, Di, &ident);
        if (dim == 2) {
            double vx = mp->v[0 + 2 * p], vy = mp->v[1 + 2 * p];
            double px = ms * vx, py = ms * vy;
            double sxx = s[0], syy = s[1], sxy = s[2];
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                double N = PHI(k, p, 0), dNx = PHI(k, p, 1), dNy = PHI(k, p, 2);
                if (no == 0) continue;
                no -= 1;
                atom_add(&me->m[no], N * ms);
                if (tpic) {
                    double dx = mp->dnp[k + nn * (0 + 2 * p)], dy = mp->dnp[k + nn * (1 + 2 * p)];
                    atom_add(&me->p[0 + 2 * no], N * ms * (vx + gv[0] * dx + gv[2] * dy));
                    atom_add(&me->p[1 + 2 * no], N * ms * (vy + gv[1] * dx + gv[3] * dy));
                } else if (apic) {
                    double x[2] = {mp->dnp[k + nn * (0 + 2 * p)], mp->dnp[k + nn * (1 + 2 * p)]}, bw[2];
                    apic_bw(2, mp->Bij + 4 * p, Di, ident, x, bw);
                    atom_add(&me->p[0 + 2 * no], N * ms * (vx + bw[0]));
                    atom_add(&me->p[1 + 2 * no], N * ms * (vy + bw[1]));
                } else {
                    atom_add(&me->p[0 + 2 * no], N * px);
                    atom_add(&me->p[1 + 2 * no], N * py);
                }
                atom_add(&me->oobf[0 + 2 * no], -(Om * (dNx * sxx + dNy * sxy)));
                atom_add(&me->oobf[1 + 2 * no], -(Om * (dNx * sxy + dNy * syy) - N * (ms * g[1])));
            }
        } else {
            double vx = mp->v[0 + 3 * p], vy = mp->v[1 + 3 * p], vz = mp->v[2 + 3 * p];
            double px = ms * vx, py = ms * vy, pz = ms * vz;
            double sxx = s[0], syy = s[1], szz = s[2], syx = s[5], szy = s[3], szx = s[4];
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                double N = PHI(k, p, 0), dNx = PHI(k, p, 1), dNy = PHI(k, p, 2), dNz = PHI(k, p, 3);
                if (no == 0) continue;
                no -= 1;
                atom_add(&me->m[no], N * ms);
                if (tpic) {
                    double dx = mp->dnp[k + nn * (0 + 3 * p)], dy = mp->dnp[k + nn * (1 + 3 * p)],
                           dz = mp->dnp[k + nn * (2 + 3 * p)];
                    /* ∇vᵢⱼ[i,j] column-major: gv[i + 3*j] */
                    atom_add(&me->p[0 + 3 * no], N * ms * (vx + gv[0] * dx + gv[3] * dy + gv[6] * dz));
                    atom_add(&me->p[1 + 3 * no], N * ms * (vy + gv[1] * dx + gv[4] * dy + gv[7] * dz));
                    atom_add(&me->p[2 + 3 * no], N * ms * (vz + gv[2] * dx + gv[5] * dy + gv[8] * dz));
                } else if (apic) {
                    double x[3] = {mp->dnp[k + nn * (0 + 3 * p)], mp->dnp[k + nn * (1 + 3 * p)],
                                   mp->dnp[k + nn * (2 + 3 * p)]}, bw[3];
                    apic_bw(3, mp->Bij + 9 * p, Di, ident, x, bw);
                    atom_add(&me->p[0 + 3 * no], N * ms * (vx + bw[0]));
                    atom_add(&me->p[1 + 3 * no], N * ms * (vy + bw[1]));
                    atom_add(&me->p[2 + 3 * no], N * ms * (vz + bw[2]));
                } else {
                    atom_add(&me->p[0 + 3 * no], N * px);
                    atom_add(&me->p[1 + 3 * no], N * py);
                    atom_add(&me->p[2 + 3 * no], N * pz);
                }
                atom_add(&me->oobf[0 + 3 * no], -(Om * (dNx * sxx + dNy * syx + dNz * szx)));
                atom_add(&me->oobf[1 + 3 * no], -(Om * (dNx * syx + dNy * syy + dNz * szy)));
                atom_add(&me->oobf[2 + 3 * no], -(Om * (dNx * szx + dNy * szy + dNz * szz) - N * (ms * g[2])));
            }
        }
    }
}

/* euler + solve (solve.jl:14-63) */
void or_solve(or_mesh *me, double dt, const or_instr *in) {
    const int dim = me->dim;
    const int64_t nno = me->nno[dim];
    const double eta = in->damping;
    memset(me->a, 0, sizeof(double) * dim * nno);
    memset(me->v, 0, sizeof(double) * dim * nno);
    PAR_FOR
    for (int64_t no = 0; no < nno; no++) {
        if (me->m[no] == 0.0) continue;
        for (int d = 0; d < dim; d++) {
            if (me->bcs[d + dim * no]) {
                me->a[d + dim * no] = 0.0;
                me->v[d + dim * no] = 0.0;
            } else {
                double mi = 1.0 / me->m[no];
                double *o = me->oobf + dim * no;
                double nrm = o[0] * o[0];
                for (int k = 1; k < dim; k++) nrm += o[k] * o[k];
                nrm = sqrt(nrm);
                double D = eta * nrm * jsign(me->p[d + dim * no] * mi);
                if (fabs(me->p[d + dim * no] * mi) >= 1.0e-3) o[d] = o[d] - D;
                me->a[d + dim * no] = o[d] * mi;
                me->v[d + dim * no] = (me->p[d + dim * no] + dt * o[d]) * mi;
            }
        }
    }
}

/* picflip_n2p + augm_momentum + augm_velocity + n2p (n2p.jl:14-76,
 * augm.jl:13-53) */
void or_n2p(or_point *mp, or_mesh *me, double dt, const or_instr *in) {
    const int dim = me->dim, nn = me->nn;
    const int64_t nmp = mp->nmp, nno = me->nno[dim];
    const double C = in->C_pf;
    double vm0 = mp->vmax[0], vm1 = mp->vmax[1], vm2 = mp->vmax[2];
#pragma omp parallel for schedule(static) reduction(max : vm0, vm1, vm2) if (g_threads > 1)
    for (int64_t p = 0; p < nmp; p++) {
        for (int d = 0; d < dim; d++) {
            double vpic = 0.0;
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                if (no == 0) continue;
                vpic += PHI(k, p, 0) * me->v[d + dim * (no - 1)];
            }
            double aflip = 0.0, vflip = 0.0;
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                if (no == 0) continue;
                aflip += PHI(k, p, 0) * me->a[d + dim * (no - 1)];
                vflip += PHI(k, p, 0) * me->v[d + dim * (no - 1)];
            }
            (void)vflip;
            double *vp = &mp->v[d + dim * p];
            *vp = C * (*vp + dt * aflip) + (1.0 - C) * vpic;
            mp->x[d + dim * p] += dt * vpic;
            double av = fabs(*vp);
            if (d == 0) vm0 = vm0 > av ? vm0 : av;
            else if (d == 1) vm1 = vm1 > av ? vm1 : av;
            else vm2 = vm2 > av ? vm2 : av;
        }
    }
    mp->vmax[0] = vm0;
    mp->vmax[1] = vm1;
    if (dim == 3) mp->vmax[2] = vm2;
    /* Bij (Bij.jl:14-26, n2p.jl:59-62): Bᵢⱼ = Σ ϕ (v_n Δnpᵀ) from the grid
     * velocity of the Euler solve; without mUSL n2p.jl:72-74 computes the
     * same values a second time. */
    if (in->trsfr == OR_APIC) {
        const int dd = dim * dim;
        PAR_FOR
        for (int64_t p = 0; p < nmp; p++) {
            double B[9] = {0.0};
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                if (no == 0) continue;
                no -= 1;
                const double N = PHI(k, p, 0);
                for (int j = 0; j < dim; j++)
                    for (int i = 0; i < dim; i++)
                        B[i + dim * j] += N * (me->v[i + dim * no] * mp->dnp[k + nn * (j + dim * p)]);
            }
            memcpy(mp->Bij + dd * p, B, sizeof(double) * dd);
        }
    }
    if (in->musl) {
        memset(me->p, 0, sizeof(double) * dim * nno);
        memset(me->v, 0, sizeof(double) * dim * nno);
        PAR_FOR
        for (int64_t p = 0; p < nmp; p++)
            for (int d = 0; d < dim; d++)
                for (int k = 0; k < nn; k++) {
                    int64_t no = mp->p2n[k + nn * p];
                    if (no == 0) continue;
                    atom_add(&me->p[d + dim * (no - 1)], PHI(k, p, 0) * ((mp->rho[p] * mp->Om[p]) * mp->v[d + dim * p]));
                }
        PAR_FOR
        for (int64_t no = 0; no < nno; no++) {
            if (me->m[no] == 0.0) continue;
            for (int d = 0; d < dim; d++) {
                if (me->bcs[d + dim * no]) me->v[d + dim * no] = 0.0;
                else me->v[d + dim * no] = me->p[d + dim * no] * (1.0 / me->m[no]);
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* elastoplast                                                               */
/* ------------------------------------------------------------------------- */

/* GIMP domain update (domain.jl:1-52), chosen by basis.how per framework
 * (update.jl:7-29), after the deformation update and before F-bar
 * (update.jl:43-49).  Restated from intent: domain.jl reads mpts.Fᵢⱼ and
 * mpts.ΔFᵢⱼ, which are mpts.s fields (types.jl:38-63), here s.F and s.dF.
 * U = n diag(√λ) nᵀ from eigen(FᵀF). */
static void domain_point(or_point *mp, int dim, int64_t p, int how) {
    const int dd = dim * dim;
    double *l = mp->l + dim * p;
    const double *l0 = mp->l0 + dim * p, *F = mp->F + dd * p, *dF = mp->dF + dd * p;
    if (how == OR_DOM_UNDEFORMED) {
        for (int i = 0; i < dim; i++) l[i] = l0[i];
    } else if (how == OR_DOM_DETF || how == OR_DOM_DETDF) {
        const double d = or_det(dim, how == OR_DOM_DETF ? F : dF);
        for (int i = 0; i < dim; i++) l[i] = (how == OR_DOM_DETF ? l0[i] : l[i]) * d;
    } else if (how == OR_DOM_FII || how == OR_DOM_DFII) {
        const double *A = how == OR_DOM_FII ? F : dF;
        for (int i = 0; i < dim; i++) l[i] = (how == OR_DOM_FII ? l0[i] : l[i]) * A[i + dim * i];
    } else {
        const double *A = how == OR_DOM_UII ? F : dF;
        double C[9], lam[3], V[9];
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) {
                double s = A[0 + dim * i] * A[0 + dim * j];
                for (int k = 1; k < dim; k++) s += A[k + dim * i] * A[k + dim * j];
                C[i + dim * j] = s;
            }
        or_sym_eig(dim, C, lam, V);
        for (int i = 0; i < dim; i++) {
            double u = 0.0;
            for (int k = 0; k < dim; k++) u += V[i + dim * k] * sqrt(lam[k]) * V[i + dim * k];
            l[i] = u * (how == OR_DOM_UII ? l0[i] : l[i]);
        }
    }
}

/* finite/infinitesimal_deform (deform.jl:1-45), undeformed (domain.jl:1-6),
 * ΔJn/ΔJs/ΔJp (volumetric.jl:1-35), update (update.jl:43-64) */
void or_update(or_point *mp, or_mesh *me, double dt, const or_instr *in) {
    const int dim = me->dim, nn = me->nn, dd = dim * dim;
    const int64_t nmp = mp->nmp, nno = me->nno[dim];
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++) {
        double *gv = mp->gradv + dd * p;
        for (int c = 0; c < dd; c++) gv[c] = 0.0;
        for (int k = 0; k < nn; k++) {
            int64_t no = mp->p2n[k + nn * p];
            if (no == 0) continue;
            no -= 1;
            for (int i = 0; i < dim; i++)
                for (int j = 0; j < dim; j++)
                    gv[i + dim * j] += PHI(k, p, j + 1) * me->v[i + dim * no];
        }
        double *dF = mp->dF + dd * p, *F = mp->F + dd * p, Fo[9];
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) dF[i + dim * j] = (i == j ? 1.0 : 0.0) + dt * gv[i + dim * j];
        memcpy(Fo, F, sizeof(double) * dd);
        matmul(dim, dF, Fo, F);
        mp->dJ[p] = or_det(dim, dF);
        mp->J[p] = or_det(dim, F);
        mp->Om[p] = mp->J[p] * mp->Om0[p];
        if (in->deform == OR_FINITE) {
            mp->rho[p] = mp->rho0[p] / mp->J[p];
            mp->n[p] = 1.0 - 1.0 / mp->J[p] * (1.0 - mp->n0[p]);
        }
    }
    if (in->basis == OR_GIMPM) {
        PAR_FOR
        for (int64_t p = 0; p < nmp; p++) domain_point(mp, dim, p, in->domain);
    }
    if (in->locking) {
        memset(me->dJ, 0, sizeof(double) * dim * nno);
        double dimr = 1.0 / (double)dim;
        PAR_FOR
        for (int64_t p = 0; p < nmp; p++)
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                if (no == 0) continue;
                atom_add(&me->dJ[no - 1], PHI(k, p, 0) * ((mp->rho[p] * mp->Om[p]) * mp->dJ[p]));
            }
        PAR_FOR
        for (int64_t no = 0; no < nno; no++) {
            if (me->m[no] == 0.0) me->dJ[no] = 0.0;
            else me->dJ[no] /= me->m[no];
        }
        PAR_FOR
        for (int64_t p = 0; p < nmp; p++) {
            double s = 0.0;
            for (int k = 0; k < nn; k++) {
                int64_t no = mp->p2n[k + nn * p];
                if (no == 0) continue;
                s += PHI(k, p, 0) * me->dJ[no - 1] / mp->dJ[p];
            }
            double f = pow(s, dimr);
            for (int c = 0; c < dd; c++) mp->dF[c + dd * p] *= f;
        }
    }
}

/* finite_elast / infinitesimal_elast (elast.jl:45-80) */
static void elast_point(or_point *mp, int dim, int64_t p, const or_instr *in) {
    const int dd = dim * dim, nstr = dim == 2 ? 3 : 6;
    double *dF = mp->dF + dd * p;
    if (in->deform == OR_FINITE) {
        double *b = mp->b + dd * p, t1[9], lam[3], V[9], fl[3], ev[6];
        matmul(dim, dF, b, t1);
        matmul_bt(dim, t1, dF, b);
        or_sym_eig(dim, b, lam, V);
        for (int k = 0; k < dim; k++) fl[k] = log(lam[k]);
        double *e = mp->eps + dd * p;
        vdvt(dim, V, fl, e);
        for (int c = 0; c < dd; c++) e[c] = 0.5 * e[c];
        to_voigt(dim, e, 2.0, ev);
        matvec(nstr, in->Del, ev, mp->tau + nstr * p);
    } else {
        double *e = mp->eps + dd * p, *w = mp->omg + dd * p, ev[6], ds[6];
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) {
                e[i + dim * j] = 0.5 * (dF[i + dim * j] + dF[j + dim * i]) - (i == j ? 1.0 : 0.0);
                w[i + dim * j] = 0.5 * (dF[i + dim * j] - dF[j + dim * i]);
            }
        /* σJ = σω' + σ'ω vanishes identically for symmetric σ (elast.jl:76-77) */
        to_voigt(dim, e, 2.0, ev);
        matvec(nstr, in->Del, ev, ds);
        for (int c = 0; c < nstr; c++) mp->sig[c + nstr * p] += ds[c];
    }
}

/* σTr, materialParam, σn and the closed-form DP return (DP.jl:1-62) on one
 * Voigt stress; epII1 accumulates, dlam is set. */
void or_dp_return(int nstr, double *s, double c0, double cr, double phi,
                  double Hp, double Kc, double Gc, double epII2,
                  double *epII1, double *dlam) {
    const double psi = 0.0 * M_PI / 180.0;
    *dlam = 0.0;
    double c = c0 + Hp * epII2;
    if (c < cr) c = cr;
    double P, t0[6], tII, eta, etaB, xi;
    if (nstr == 3) {
        P = (s[0] + s[1]) / 2.0;
        t0[0] = s[0] - P; t0[1] = s[1] - P; t0[2] = s[2] - 0.0;
        tII = sqrt(0.5 * (t0[0] * t0[0] + t0[1] * t0[1]) + t0[2] * t0[2]);
        eta = 3.0 * tan(phi) / (sqrt(9.0 + 12.0 * tan(phi) * tan(phi)));
        etaB = 3.0 * tan(psi) / (sqrt(9.0 + 12.0 * tan(psi) * tan(psi)));
        xi = 3.0 * c / (sqrt(9.0 + 12.0 * tan(phi) * tan(phi)));
    } else {
        P = (s[0] + s[1] + s[2]) / 3.0;
        t0[0] = s[0] - P; t0[1] = s[1] - P; t0[2] = s[2] - P;
        t0[3] = s[3] - 0.0; t0[4] = s[4] - 0.0; t0[5] = s[5] - 0.0;
        tII = sqrt(0.5 * (t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]) + t0[3] * t0[3] + t0[4] * t0[4] + t0[5] * t0[5]);
        eta = 6.0 * sin(phi) / (sqrt(3.0) * (3.0 + sin(phi)));
        etaB = 6.0 * sin(psi) / (sqrt(3.0) * (3.0 + sin(psi)));
        xi = 6.0 * c * cos(phi) / (sqrt(3.0) * (3.0 + sin(phi)));
    }
    double sm = xi / eta, tP = xi - eta * (xi / eta);
    double fs = tII + eta * P - xi;
    double h = tII - tP - (sqrt(1.0 + eta * eta)) * (P - sm);
    int nd = nstr == 3 ? 2 : 3;
    if ((fs > 0.0 && P < sm) || (h > 0.0 && P >= sm)) {
        double dl = fs / (Gc + Kc * eta * etaB);
        double Pn = P - Kc * etaB * dl, tn = xi - eta * (P - Kc * etaB * dl);
        *dlam = dl;
        double r = tn / tII;
        for (int k = 0; k < nstr; k++) s[k] = t0[k] * r + (k < nd ? Pn : 0.0);
        *epII1 += dl * sqrt(1.0 / 3.0 + (2.0 / 9.0) * (etaB * etaB));
    }
    if (h <= 0.0 && P >= sm) {
        double dl = (P - sm) / Kc;
        double Pn = sm - P;
        *dlam = dl;
        double r = 0.0 / tII;
        for (int k = 0; k < nstr; k++) s[k] = t0[k] * r + (k < nd ? Pn : 0.0);
        *epII1 += sqrt(2.0) * dl / 3.0;
    }
}

/* get_J2 / yield_J2 (J2.jl:1-20) */
static double yield_j2(int nstr, const double *s, double kappa, double *n) {
    double P, x[6], J2;
    if (nstr == 3) {
        P = (s[0] + s[1]) / 2.0;
        x[0] = s[0] - P; x[1] = s[1] - P; x[2] = s[2] - 0.0;
        J2 = 0.5 * (x[0] * x[0] + x[1] * x[1] + 2.0 * (x[2] * x[2]));
    } else {
        P = (s[0] + s[1] + s[2]) / 3.0;
        x[0] = s[0] - P; x[1] = s[1] - P; x[2] = s[2] - P;
        x[3] = s[3] - 0.0; x[4] = s[4] - 0.0; x[5] = s[5] - 0.0;
        J2 = 0.5 * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + 2.0 * (x[3] * x[3]) + 2.0 * (x[4] * x[4]) + 2.0 * (x[5] * x[5]));
    }
    double xn = sqrt(2.0 * J2);
    for (int k = 0; k < nstr; k++) n[k] = x[k] / xn;
    return xn - kappa;
}

/* finite_J2 / infinitesimal_J2 cutting-plane loop (J2.jl:21-75), restated
 * from its intent. */
void or_j2_return(int nstr, double *s, double c0, double cr, double Hp,
                  const double *Del, double epII2, double *epII1,
                  int *active, int *iters) {
    const double ftol = 1e-9;
    const int itmax = 20;
    double kappa = fmax(cr, c0 + Hp * epII2), n[6];
    double f = yield_j2(nstr, s, kappa, n);
    *active = 0;
    *iters = 0;
    if (f > 0.0) {
        double g0 = epII2;
        int it = 1;
        while (fabs(f) > ftol && it < itmax) {
            double Dn[6], nDn = 0.0;
            /* ∂f∂σ'*Del*∂f∂σ: row vector n'*Del, then dot with n */
            double nD[6];
            for (int j = 0; j < nstr; j++) {
                double t = 0.0;
                for (int i = 0; i < nstr; i++) t += n[i] * Del[i + nstr * j];
                nD[j] = t;
            }
            for (int j = 0; j < nstr; j++) nDn += nD[j] * n[j];
            double dl = f / nDn;
            for (int i = 0; i < nstr; i++) {
                double t = 0.0;
                for (int j = 0; j < nstr; j++) t += (dl * Del[i + nstr * j]) * n[j];
                Dn[i] = t;
            }
            for (int i = 0; i < nstr; i++) s[i] -= Dn[i];
            g0 += dl;
            kappa = fmax(cr, c0 + Hp * g0);
            f = yield_j2(nstr, s, kappa, n);
            it += 1;
        }
        *epII1 = g0;
        *active = 1;
        *iters = it;
    }
}

/* b = exp(2 * tensor(Del\τ, 1/2)) (DP.jl:63-66, J2.jl:42-45) */
static void b_from_tau(or_point *mp, int dim, int64_t p, const or_instr *in) {
    const int dd = dim * dim, nstr = dim == 2 ? 3 : 6;
    double y[6], lam[3], V[9], fe[3];
    or_lu_solve(nstr, in->Del, mp->tau + nstr * p, y);
    double *e = mp->eps + dd * p;
    to_tensor(dim, y, 0.5, e);
    or_sym_eig(dim, e, lam, V);
    for (int k = 0; k < dim; k++) fe[k] = exp(2.0 * lam[k]);
    vdvt(dim, V, fe, mp->b + dd * p);
}

void or_elast(or_point *mp, const or_instr *in) {
    PAR_FOR
    for (int64_t p = 0; p < mp->nmp; p++) elast_point(mp, mp->dim, p, in);
}

/* plast (plast.jl:49-65) with the local (nonloc off) branch */
/* Symmetric pair weights w[p,q] of nonlocal.jl: an open-addressing map from
 * the pair (min, max) to its value; absent pairs read 0. */
typedef struct { int64_t *key; double *val; int64_t cap; } pairmap;

static int64_t pm_slot(const pairmap *m, int64_t p, int64_t q) {
    const int64_t a = p < q ? p : q, b = p < q ? q : p;
    const int64_t k = a * 0x100000000LL + b;
    uint64_t h = (uint64_t)k * 0x9E3779B97F4A7C15ULL;
    int64_t i = (int64_t)(h >> 20) & (m->cap - 1);
    while (m->key[i] != -1 && m->key[i] != k) i = (i + 1) & (m->cap - 1);
    m->key[i] = k;
    return i;
}

/* nonlocal (nonlocal.jl:18-47) as plast calls it (plast.jl:49-60), in the
 * reference's single-threaded order: "tplgy" lists, for every element, the
 * particles whose e2e neighbourhood (e2e.jl:20-61: within ceil(ls/h)
 * elements on every axis) holds it; "p->q" gives every plastic particle p
 * (Δλ of the last return != 0), for each particle q of its element's list in
 * ascending order whose weight is still 0, w = w[q,p] = (d/ls) exp(-(d/ls)^2),
 * d from the first two coordinates, adds w to W[p] and W[q] and records q in
 * p's p2p column; "p<-q" sets ϵpII[2,p] = Σ (w[p,q]/W[p]) ϵpII[1,q] over
 * that column, or ϵpII[1,p] when W[p] <= 1e-16.  ϵpII[2] of every other
 * particle is 0.  Returns -1 if memory runs out. */
int or_nonlocal(or_point *mp, const or_mesh *me, const or_instr *in) {
    const int dim = me->dim;
    const int64_t nmp = mp->nmp;
    int64_t nel[3] = {1, 1, 1}, nnel[3] = {0, 0, 0};
    for (int i = 0; i < dim; i++) {
        nel[i] = me->nel[i];
        nnel[i] = (int64_t)ceil(in->ls / me->h[i]);
    }
    const int64_t ntot = me->nel[dim];
    /* element coordinates from the 1-based element id (tplgy.jl:49, 82) */
    int64_t *ec = malloc(sizeof(int64_t) * 3 * (size_t)nmp);
    int64_t *cnt = calloc((size_t)ntot + 1, sizeof(int64_t));
    double *W = calloc((size_t)nmp, sizeof(double));
    if (!ec || !cnt || !W) { free(ec); free(cnt); free(W); return -1; }
    const int zl = dim - 1;
    for (int64_t p = 0; p < nmp; p++) {
        int64_t e = mp->p2e[p] - 1;
        ec[3 * p + zl] = e % nel[zl];
        e /= nel[zl];
        ec[3 * p + 0] = e % nel[0];
        if (dim == 3) ec[3 * p + 1] = e / nel[0];
    }
#define EL_ID(x, y, z) (dim == 2 ? (z) + nel[1] * (x) : (z) + nel[2] * ((x) + nel[0] * (y)))
#define NBR_LOOP(p, BODY)                                                                          \
    for (int64_t yy = (dim == 3 ? ec[3 * (p) + 1] - nnel[1] : 0); yy <= (dim == 3 ? ec[3 * (p) + 1] + nnel[1] : 0); yy++) \
        for (int64_t xx = ec[3 * (p)] - nnel[0]; xx <= ec[3 * (p)] + nnel[0]; xx++)               \
            for (int64_t zz = ec[3 * (p) + zl] - nnel[zl]; zz <= ec[3 * (p) + zl] + nnel[zl]; zz++) { \
                if (xx < 0 || xx >= nel[0] || zz < 0 || zz >= nel[zl] || (dim == 3 && (yy < 0 || yy >= nel[1]))) continue; \
                const int64_t el = EL_ID(xx, yy, zz);                                              \
                BODY                                                                               \
            }
    /* "tplgy": e2p[p, el] = p for el in p's neighbourhood -> per-element lists, p ascending */
    for (int64_t p = 0; p < nmp; p++) { NBR_LOOP(p, cnt[el + 1]++;) }
    for (int64_t e = 0; e < ntot; e++) cnt[e + 1] += cnt[e];
    int64_t *lst = malloc(sizeof(int64_t) * (size_t)(cnt[ntot] ? cnt[ntot] : 1));
    int64_t *fill = malloc(sizeof(int64_t) * (size_t)ntot);
    if (!lst || !fill) { free(ec); free(cnt); free(W); free(lst); free(fill); return -1; }
    memcpy(fill, cnt, sizeof(int64_t) * (size_t)ntot);
    for (int64_t p = 0; p < nmp; p++) { NBR_LOOP(p, lst[fill[el]++] = p;) }
#undef NBR_LOOP
#undef EL_ID
    /* pair map and p2p columns sized by the plastic particles' candidates */
    int64_t npair = 0;
    for (int64_t p = 0; p < nmp; p++)
        if (mp->dlam[p] != 0.0) { const int64_t e = mp->p2e[p] - 1; npair += cnt[e + 1] - cnt[e]; }
    pairmap m;
    m.cap = 16;
    while (m.cap < 2 * npair + 16) m.cap *= 2;
    m.key = malloc(sizeof(int64_t) * (size_t)m.cap);
    m.val = malloc(sizeof(double) * (size_t)m.cap);
    int64_t *p2p = malloc(sizeof(int64_t) * (size_t)(npair ? npair : 1));
    int64_t *p2p0 = malloc(sizeof(int64_t) * (size_t)(nmp + 1));
    if (!m.key || !m.val || !p2p || !p2p0) {
        free(ec); free(cnt); free(W); free(lst); free(fill); free(m.key); free(m.val); free(p2p); free(p2p0);
        return -1;
    }
    for (int64_t i = 0; i < m.cap; i++) { m.key[i] = -1; m.val[i] = 0.0; }
    for (int64_t p = 0; p < nmp; p++) mp->epII[1 + 2 * p] = 0.0;
    /* "p->q" */
    int64_t np2p = 0;
    for (int64_t p = 0; p < nmp; p++) {
        p2p0[p] = np2p;
        if (mp->dlam[p] == 0.0) continue;
        const int64_t e = mp->p2e[p] - 1;
        for (int64_t k = cnt[e]; k < cnt[e + 1]; k++) {
            const int64_t q = lst[k];
            const double xi = mp->x[0 + dim * p] - mp->x[0 + dim * q];
            const double eta = mp->x[1 + dim * p] - mp->x[1 + dim * q];
            const double d = sqrt(xi * xi + eta * eta);
            const int64_t s = pm_slot(&m, p, q);
            if (m.val[s] == 0.0) {
                const double w0 = d / in->ls * exp(-((d / in->ls) * (d / in->ls)));
                m.val[s] = w0;
                W[p] += w0;
                W[q] += w0;
                p2p[np2p++] = q;
            }
        }
    }
    p2p0[nmp] = np2p;
    /* "p<-q" */
    for (int64_t p = 0; p < nmp; p++) {
        if (mp->dlam[p] == 0.0) continue;
        if (!(W[p] > 1e-16)) {
            mp->epII[1 + 2 * p] = mp->epII[0 + 2 * p];
        } else {
            for (int64_t k = p2p0[p]; k < p2p0[p + 1]; k++) {
                const int64_t q = p2p[k];
                mp->epII[1 + 2 * p] += (m.val[pm_slot(&m, p, q)] / W[p]) * mp->epII[0 + 2 * q];
            }
        }
    }
    free(ec); free(cnt); free(W); free(lst); free(fill); free(m.key); free(m.val); free(p2p); free(p2p0);
    return 0;
}

void or_plast(or_point *mp, const or_instr *in) {
    const int dim = mp->dim, nstr = dim == 2 ? 3 : 6;
    const int64_t nmp = mp->nmp;
    /* without nonlocal regularisation ϵpII[2,:] = ϵpII[1,:] (plast.jl:61-63);
     * with it, or_step has run or_nonlocal first */
    if (!in->nonloc)
        for (int64_t p = 0; p < nmp; p++) mp->epII[1 + 2 * p] = mp->epII[0 + 2 * p];
    PAR_FOR
    for (int64_t p = 0; p < nmp; p++) {
        double *s = in->deform == OR_FINITE ? mp->tau + nstr * p : mp->sig + nstr * p;
        if (in->plast == OR_DP) {
            or_dp_return(nstr, s, mp->c0[p], mp->cr[p], mp->phif[p], in->Hp, in->Kc, in->Gc,
                         mp->epII[1 + 2 * p], &mp->epII[0 + 2 * p], &mp->dlam[p]);
            if (in->deform == OR_FINITE) b_from_tau(mp, dim, p, in);
        } else if (in->plast == OR_J2) {
            int active, iters;
            or_j2_return(nstr, s, mp->c0[p], mp->cr[p], in->Hp, in->Del, mp->epII[1 + 2 * p],
                         &mp->epII[0 + 2 * p], &active, &iters);
            if (active && in->deform == OR_FINITE) b_from_tau(mp, dim, p, in);
        }
    }
}

int or_step(or_point *mp, or_mesh *me, const or_instr *in, const double *g, double dt, int with_plast) {
    mp->dim = me->dim;
    int r = or_shpfun(mp, me, in);
    if (r != OR_OK) return r;
    or_p2n(mp, me, g, in);
    or_solve(me, dt, in);
    or_n2p(mp, me, dt, in);
    or_update(mp, me, dt, in);
    or_elast(mp, in);
    if (with_plast && in->plast != OR_PLAST_NONE) {
        if (in->nonloc && or_nonlocal(mp, me, in) != 0) return -1;
        or_plast(mp, in);
    }
    return OR_OK;
}

/* get_dt (physics.jl:21-26) */
double or_get_dt(or_point *mp, const or_mesh *me, double c, double t, double T) {
    double cmax = -INFINITY;
    for (int i = 0; i < me->dim; i++) {
        double r = me->h[i] / (mp->vmax[i] + c);
        if (r > cmax) cmax = r;
        mp->vmax[i] = 0.0;
    }
    double a = 0.5 * cmax, b = T - t;
    return a < b ? a : b;
}
