// ep_math.h -- per-particle math of the MPM step (shape functions, small
// dense linear algebra, constitutive updates) for gfx950 device code.
//
// Every function restates the ElastoPlasm.jl v0.4.2 line it cites with the
// same operation order, so the only differences from a sequential CPU run
// are fused multiply-adds and the order of atomic accumulation.
#pragma once
#include "ep_runtime.h"
#include <stdint.h>

namespace ep {

// Floor of the exact quotient a/h, which is what Julia's fld returns for
// floats (tplgy.jl:47-48, 79-81); the fma residuals are exact in sign.
__device__ __forceinline__ double exact_fld(double a, double h, double ih) {
    double q = floor(a * ih);
    if (fma(-q, h, a) < 0.0) q -= 1.0;
    else if (fma(-(q + 1.0), h, a) >= 0.0) q += 1.0;
    return q;
}

// t1..t4_ϕ∂ϕ (bsmpm.jl:1-62): piecewise cubic B-spline and derivative in
// units of the cell (divide the derivative by h).
template <typename T>
__device__ __forceinline__ void bspline(T xi, int type, T &f, T &df) {
    const T x2 = xi * xi, x3 = x2 * xi;
    const T c16 = T(1.0 / 6.0), c13 = T(1.0 / 3.0), c43 = T(4.0 / 3.0), c23 = T(2.0 / 3.0);
    f = T(0);
    df = T(0);
    if (type == 3) {
        if (T(-2) <= xi && xi <= T(-1)) { f = c16 * x3 + x2 + T(2) * xi + c43; df = T(0.5) * x2 + T(2) * xi + T(2); }
        else if (T(-1) <= xi && xi <= T(0)) { f = T(-0.5) * x3 - x2 + c23; df = T(-1.5) * x2 - T(2) * xi; }
        else if (T(0) <= xi && xi <= T(1)) { f = T(0.5) * x3 - x2 + c23; df = T(1.5) * x2 - T(2) * xi; }
        else if (T(1) <= xi && xi <= T(2)) { f = -c16 * x3 + x2 - T(2) * xi + c43; df = T(-0.5) * x2 + T(2) * xi - T(2); }
    } else if (type == 1) {
        if (T(-2) <= xi && xi <= T(-1)) { f = c16 * x3 + x2 + T(2) * xi + c43; df = T(0.5) * x2 + T(2) * xi + T(2); }
        else if (T(-1) <= xi && xi <= T(0)) { f = -c16 * x3 + xi + T(1); df = T(-0.5) * x2 + T(1); }
        else if (T(0) <= xi && xi <= T(1)) { f = c16 * x3 - xi + T(1); df = T(0.5) * x2 - T(1); }
        else if (T(1) <= xi && xi <= T(2)) { f = -c16 * x3 + x2 - T(2) * xi + c43; df = T(-0.5) * x2 + T(2) * xi - T(2); }
    } else if (type == 2) {
        if (T(-1) <= xi && xi <= T(0)) { f = -c13 * x3 - x2 + c23; df = T(-1) * x2 - T(2) * xi; }
        else if (T(0) <= xi && xi <= T(1)) { f = T(0.5) * x3 - x2 + c23; df = T(1.5) * x2 - T(2) * xi; }
        else if (T(1) <= xi && xi <= T(2)) { f = -c16 * x3 + x2 - T(2) * xi + c43; df = T(-0.5) * x2 + T(2) * xi - T(2); }
    } else if (type == 4) {
        if (T(-2) <= xi && xi <= T(-1)) { f = c16 * x3 + x2 + T(2) * xi + c43; df = T(0.5) * x2 + T(2) * xi + T(2); }
        else if (T(-1) <= xi && xi <= T(0)) { f = T(-0.5) * x3 - x2 + c23; df = T(-1.5) * x2 - T(2) * xi; }
        else if (T(0) <= xi && xi <= T(1)) { f = c13 * x3 - x2 + c23; df = T(1) * x2 - T(2) * xi; }
    }
}

template <typename T> __device__ __forceinline__ T jsign(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : x); }

// S∂S (gimpm.jl:1-14)
template <typename T>
__device__ __forceinline__ void gimp(T dx, T h, T lp, T &S, T &dS) {
    const T ax = fabs(dx);
    S = T(0);
    dS = T(0);
    if (ax < lp) {
        S = T(1) - ((T(4) * (dx * dx) + (T(2) * lp) * (T(2) * lp)) / (T(8) * h * lp));
        dS = -((T(8) * dx) / (T(8) * h * lp));
    } else if ((ax >= lp) && (ax <= (h - lp))) {
        S = T(1) - (ax / h);
        dS = jsign(dx) * (T(-1) / h);
    } else if ((ax >= (h - lp)) && (ax < (h + lp))) {
        const T r = h + lp - ax;
        S = (r * r) / (T(4) * h * lp);
        dS = -jsign(dx) * r / (T(2) * h * lp);
    }
}

// N∂N (smpm.jl:1-10)
template <typename T>
__device__ __forceinline__ void linear(T dx, T h, T &N, T &dN) {
    if (-h < dx && dx <= T(0)) { N = T(1) + dx / h; dN = T(1) / h; }
    else if (T(0) < dx && dx <= h) { N = T(1) - dx / h; dN = T(-1) / h; }
    else { N = T(0); dN = T(0); }
}

// ---------------------------------------------------------------------------
// small dense linear algebra, column-major d x d
// ---------------------------------------------------------------------------

// Cyclic Jacobi eigen-decomposition of a symmetric matrix; a is overwritten
// by the rotations, v receives eigenvectors (columns).  Stands in for the
// symmetric `eigen` of elast.jl:51 / DP.jl:65.
// FAST (the FP64 step's decomposition of the trial b, whose entries are
// O(1)): the tangent from d = a_qq - a_pp with one division,
// t = sgn(d) 2 a_pq / (|d| + sqrt(d² + 4 a_pq²)), and no confirming sweep after
// a sweep whose rotations all had |t| < 1e-9 (the off-diagonal left is then
// of order t² of the diagonal, below `tiny`).
template <int D, typename T, bool FAST = false>
__device__ __forceinline__ void sym_eig(T (&a)[D][D], T (&v)[D][D]) {
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) v[i][j] = (i == j) ? T(1) : T(0);
    const T tiny = sizeof(T) == 8 ? T(1e-18) : T(1e-9);
    for (int sweep = 0; sweep < 32; sweep++) {
        bool rotated = false;
        T tmax = T(0);
#pragma unroll
        for (int p = 0; p < D - 1; p++) {
#pragma unroll
            for (int q = p + 1; q < D; q++) {
                const T apq = a[p][q];
                if (apq == T(0)) continue;
                if (fabs(apq) <= tiny * (fabs(a[p][p]) + fabs(a[q][q]))) {
                    a[p][q] = a[q][p] = T(0);
                    continue;
                }
                rotated = true;
                // t = tan of the rotation angle, the smaller root of
                // t² + 2θt - 1 = 0 with θ = (a_qq - a_pp)/(2 a_pq).  θ is a
                // ratio, so tiny entries (the FP32 path decomposes b - I) do
                // not underflow as β² + a_pq² would
                T t;
                if constexpr (FAST) {
                    const T d = a[q][q] - a[p][p];
                    const T r = sqrt(d * d + T(4) * apq * apq);
                    t = (T(2) * apq) / (d >= T(0) ? d + r : d - r);
                    tmax = fmax(tmax, fabs(t));
                } else {
                    const T theta = (a[q][q] - a[p][p]) / (T(2) * apq);
                    if (fabs(theta) > T(sizeof(T) == 8 ? 1e150 : 1e18)) t = T(0.5) / theta;
                    else t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
                }
                const T c = rsqrt(t * t + T(1)), s = t * c;
                // symmetric two-sided update: the pivot pair in closed form,
                // the other entries of rows/columns p and q once
                a[p][p] -= t * apq;
                a[q][q] += t * apq;
                a[p][q] = a[q][p] = T(0);
#pragma unroll
                for (int k = 0; k < D; k++) {
                    if (k == p || k == q) continue;
                    const T akp = a[k][p], akq = a[k][q];
                    const T np = c * akp - s * akq, nq = s * akp + c * akq;
                    a[k][p] = a[p][k] = np;
                    a[k][q] = a[q][k] = nq;
                }
#pragma unroll
                for (int k = 0; k < D; k++) {
                    const T vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
        }
        if (!rotated || (FAST && tmax < T(1e-9))) break;
    }
}

// det (deform.jl:17-18): diagonal product when triangular (Julia's generic
// det), otherwise LU with partial pivoting.
template <int D, typename T>
__device__ __forceinline__ T det(const T (&A)[D][D]) {
    bool triu = true, tril = true;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            if (i > j && A[i][j] != T(0)) triu = false;
            if (i < j && A[i][j] != T(0)) tril = false;
        }
    if (triu || tril) {
        T P = T(1);
#pragma unroll
        for (int i = 0; i < D; i++) P *= A[i][i];
        return P;
    }
    T M[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) M[i][j] = A[i][j];
    int swaps = 0;
    bool singular = false;
#pragma unroll
    for (int k = 0; k < D; k++) {
        int pk = k;
        T mx = fabs(M[k][k]);
#pragma unroll
        for (int i = k + 1; i < D; i++)
            if (fabs(M[i][k]) > mx) { mx = fabs(M[i][k]); pk = i; }
        if (pk != k) {
            swaps++;
#pragma unroll
            for (int j = 0; j < D; j++) {
                // register-resident swap without dynamic indexing
#pragma unroll
                for (int i = k + 1; i < D; i++)
                    if (i == pk) { T t = M[k][j]; M[k][j] = M[i][j]; M[i][j] = t; }
            }
        }
        const T akk = M[k][k];
        if (akk == T(0)) { singular = true; continue; }
        const T r = T(1) / akk;
#pragma unroll
        for (int i = k + 1; i < D; i++) M[i][k] *= r;
#pragma unroll
        for (int j = k + 1; j < D; j++)
#pragma unroll
            for (int i = k + 1; i < D; i++) M[i][j] -= M[i][k] * M[k][j];
    }
    if (singular) return T(0);
    T P = T(1);
#pragma unroll
    for (int i = 0; i < D; i++) P *= M[i][i];
    return (swaps & 1) ? -P : P;
}

// det by cofactors, for the deformation gradients (ΔF ≈ I, F with J > 0),
// where it equals the LU determinant above up to rounding at a fraction of
// its instructions (no pivot search, no division)
template <int D, typename T>
__device__ __forceinline__ T det_cof(const T (&A)[D][D]) {
    if (D == 2) return A[0][0] * A[1][1] - A[0][1] * A[1][0];
    return A[0][0] * (A[1][1] * A[D - 1][D - 1] - A[1][D - 1] * A[D - 1][1]) -
           A[0][1] * (A[1][0] * A[D - 1][D - 1] - A[1][D - 1] * A[D - 1][0]) +
           A[0][D - 1] * (A[1][0] * A[D - 1][1] - A[1][1] * A[D - 1][0]);
}

// inv (p2n.jl:243-249): LU with partial pivoting as det above, then the
// columns of the identity by forward and back substitution.
template <int D>
__device__ __forceinline__ void inv_lu(const double (&A)[D][D], double (&Ai)[D][D]) {
    double M[D][D];
    int piv[D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) M[i][j] = A[i][j];
#pragma unroll
    for (int k = 0; k < D; k++) {
        int pk = k;
        double mx = fabs(M[k][k]);
#pragma unroll
        for (int i = k + 1; i < D; i++)
            if (fabs(M[i][k]) > mx) { mx = fabs(M[i][k]); pk = i; }
        piv[k] = pk;
#pragma unroll
        for (int j = 0; j < D; j++)
#pragma unroll
            for (int i = k + 1; i < D; i++)
                if (i == pk) { const double t = M[k][j]; M[k][j] = M[i][j]; M[i][j] = t; }
        const double akk = M[k][k];
        if (akk == 0.0) continue;
        const double r = 1.0 / akk;
#pragma unroll
        for (int i = k + 1; i < D; i++) M[i][k] *= r;
#pragma unroll
        for (int j = k + 1; j < D; j++)
#pragma unroll
            for (int i = k + 1; i < D; i++) M[i][j] -= M[i][k] * M[k][j];
    }
#pragma unroll
    for (int c = 0; c < D; c++) {
        double y[D], x[D];
#pragma unroll
        for (int i = 0; i < D; i++) y[i] = i == c ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < D; k++)
#pragma unroll
            for (int i = k + 1; i < D; i++)
                if (i == piv[k]) { const double t = y[k]; y[k] = y[i]; y[i] = t; }
#pragma unroll
        for (int i = 0; i < D; i++) {
            double s = y[i];
#pragma unroll
            for (int k = 0; k < i; k++) s -= M[i][k] * y[k];
            y[i] = s;
        }
#pragma unroll
        for (int i = D - 1; i >= 0; i--) {
            double s = y[i];
#pragma unroll
            for (int k = i + 1; k < D; k++) s -= M[i][k] * x[k];
            x[i] = s / M[i][i];
        }
#pragma unroll
        for (int i = 0; i < D; i++) Ai[i][c] = x[i];
    }
}

// Pre-factored Del for `Del \ τ` (DP.jl:64): LU with partial pivoting,
// computed once on the host by the same algorithm.
struct DelLU {
    double LU[36];   // column-major
    int piv[6];
};

template <int N, typename T>
__device__ __forceinline__ void lu_solve(const DelLU &f, const T *b, T *x) {
    T y[N];
#pragma unroll
    for (int i = 0; i < N; i++) y[i] = b[i];
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int pk = f.piv[k];
#pragma unroll
        for (int i = k + 1; i < N; i++)
            if (i == pk) { T t = y[k]; y[k] = y[i]; y[i] = t; }
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        T s = y[i];
#pragma unroll
        for (int k = 0; k < i; k++) s -= T(f.LU[i + N * k]) * y[k];
        y[i] = s;
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        T s = y[i];
#pragma unroll
        for (int k = i + 1; k < N; k++) s -= T(f.LU[i + N * k]) * x[k];
        x[i] = s / T(f.LU[i + N * i]);
    }
}

// σTr / materialParam / σn and the closed-form Drucker-Prager return
// (DP.jl:1-62) on one Voigt stress.  Returns 0 elastic, 1 shear, 2 apex.
// The step passes ψ = 0 as finite_DP / infinitesimal_DP hard-code it
// (DP.jl:37, 73); the surrogate's training samples use materialParam's ψ
// argument (DP.jl:13-24) for ψ in [0, φ].
// The friction angle enters through sf = tan φ (2D) or sin φ (3D) and
// cf = cos φ (3D); φ is constant per particle, so the step computes them once
// per particle (friction_trig) and the return below reads them.
template <int NSTR, typename T>
__device__ __forceinline__ void friction_trig(T phi, T &sf, T &cf) {
    if (NSTR == 3) { sf = tan(phi); cf = T(0); }
    else { sf = sin(phi); cf = cos(phi); }
}

template <int NSTR, typename T>
__device__ __forceinline__ int dp_return_trig(T *s, T c0, T cr, T sf, T cf, T Hp, T Kc, T Gc, T epII2, T &epII1,
                                              T &dlam, T psi = T(0)) {
    constexpr int ND = NSTR == 3 ? 2 : 3;
    dlam = T(0);
    T c = c0 + Hp * epII2;
    if (c < cr) c = cr;
    T P, t0[NSTR], tII, eta, etaB, xi;
    if (NSTR == 3) {
        P = (s[0] + s[1]) / T(2);
        t0[0] = s[0] - P; t0[1] = s[1] - P; t0[2] = s[2] - T(0);
        tII = sqrt(T(0.5) * (t0[0] * t0[0] + t0[1] * t0[1]) + t0[2] * t0[2]);
        const T tp = sf, tq = tan(psi);
        eta = T(3) * tp / (sqrt(T(9) + T(12) * tp * tp));
        etaB = T(3) * tq / (sqrt(T(9) + T(12) * tq * tq));
        xi = T(3) * c / (sqrt(T(9) + T(12) * tp * tp));
    } else {
        P = (s[0] + s[1] + s[2]) / T(3);
#pragma unroll
        for (int k = 0; k < 3; k++) t0[k] = s[k] - P;
#pragma unroll
        for (int k = 3; k < NSTR; k++) t0[k] = s[k] - T(0);
        tII = sqrt(T(0.5) * (t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]) + t0[3] * t0[3] + t0[4] * t0[4] + t0[5] * t0[5]);
        const T sp = sf, sq = sin(psi);
        eta = T(6) * sp / (sqrt(T(3)) * (T(3) + sp));
        etaB = T(6) * sq / (sqrt(T(3)) * (T(3) + sq));
        xi = T(6) * c * cf / (sqrt(T(3)) * (T(3) + sp));
    }
    const T sm = xi / eta, tP = xi - eta * (xi / eta);
    const T fs = tII + eta * P - xi;
    const T h = tII - tP - (sqrt(T(1) + eta * eta)) * (P - sm);
    int branch = 0;
    if ((fs > T(0) && P < sm) || (h > T(0) && P >= sm)) {
        const T dl = fs / (Gc + Kc * eta * etaB);
        const T Pn = P - Kc * etaB * dl, tn = xi - eta * (P - Kc * etaB * dl);
        dlam = dl;
        const T r = tn / tII;
#pragma unroll
        for (int k = 0; k < NSTR; k++) s[k] = t0[k] * r + (k < ND ? Pn : T(0));
        epII1 += dl * sqrt(T(1.0 / 3.0) + T(2.0 / 9.0) * (etaB * etaB));
        branch = 1;
    }
    if (h <= T(0) && P >= sm) {
        const T dl = (P - sm) / Kc;
        const T Pn = sm - P;
        dlam = dl;
        const T r = T(0) / tII;
#pragma unroll
        for (int k = 0; k < NSTR; k++) s[k] = t0[k] * r + (k < ND ? Pn : T(0));
        epII1 += sqrt(T(2)) * dl / T(3);
        branch = 2;
    }
    return branch;
}

template <int NSTR, typename T>
__device__ __forceinline__ int dp_return(T *s, T c0, T cr, T phi, T Hp, T Kc, T Gc, T epII2, T &epII1, T &dlam,
                                         T psi = T(0)) {
    T sf, cf;
    friction_trig<NSTR, T>(phi, sf, cf);
    return dp_return_trig<NSTR, T>(s, c0, cr, sf, cf, Hp, Kc, Gc, epII2, epII1, dlam, psi);
}

// The branch dp_return would take on s (0 elastic, 1 shear, 2 apex), with the
// same arithmetic and ψ = 0, without changing anything.
template <int NSTR, typename T>
__device__ __forceinline__ int dp_trial_trig(const T *s, T c0, T cr, T sf, T cf, T Hp, T epII2) {
    T c = c0 + Hp * epII2;
    if (c < cr) c = cr;
    T P, t0[NSTR], tII, eta, xi;
    if (NSTR == 3) {
        P = (s[0] + s[1]) / T(2);
        t0[0] = s[0] - P; t0[1] = s[1] - P; t0[2] = s[2] - T(0);
        tII = sqrt(T(0.5) * (t0[0] * t0[0] + t0[1] * t0[1]) + t0[2] * t0[2]);
        const T tp = sf;
        eta = T(3) * tp / (sqrt(T(9) + T(12) * tp * tp));
        xi = T(3) * c / (sqrt(T(9) + T(12) * tp * tp));
    } else {
        P = (s[0] + s[1] + s[2]) / T(3);
#pragma unroll
        for (int k = 0; k < 3; k++) t0[k] = s[k] - P;
#pragma unroll
        for (int k = 3; k < NSTR; k++) t0[k] = s[k] - T(0);
        tII = sqrt(T(0.5) * (t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]) + t0[3] * t0[3] + t0[4] * t0[4] + t0[5] * t0[5]);
        const T sp = sf;
        eta = T(6) * sp / (sqrt(T(3)) * (T(3) + sp));
        xi = T(6) * c * cf / (sqrt(T(3)) * (T(3) + sp));
    }
    const T sm = xi / eta, tP = xi - eta * (xi / eta);
    const T fs = tII + eta * P - xi;
    const T h = tII - tP - (sqrt(T(1) + eta * eta)) * (P - sm);
    if ((fs > T(0) && P < sm) || (h > T(0) && P >= sm)) return 1;
    if (h <= T(0) && P >= sm) return 2;
    return 0;
}

template <int NSTR, typename T>
__device__ __forceinline__ int dp_trial(const T *s, T c0, T cr, T phi, T Hp, T epII2) {
    T sf, cf;
    friction_trig<NSTR, T>(phi, sf, cf);
    return dp_trial_trig<NSTR, T>(s, c0, cr, sf, cf, Hp, epII2);
}

// get_J2 / yield_J2 (J2.jl:1-20)
template <int NSTR, typename T>
__device__ __forceinline__ T yield_j2(const T *s, T kappa, T *n) {
    T P, x[NSTR], J2;
    if (NSTR == 3) {
        P = (s[0] + s[1]) / T(2);
        x[0] = s[0] - P; x[1] = s[1] - P; x[2] = s[2] - T(0);
        J2 = T(0.5) * (x[0] * x[0] + x[1] * x[1] + T(2) * (x[2] * x[2]));
    } else {
        P = (s[0] + s[1] + s[2]) / T(3);
#pragma unroll
        for (int k = 0; k < 3; k++) x[k] = s[k] - P;
#pragma unroll
        for (int k = 3; k < NSTR; k++) x[k] = s[k] - T(0);
        J2 = T(0.5) * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + T(2) * (x[3] * x[3]) + T(2) * (x[4] * x[4]) + T(2) * (x[5] * x[5]));
    }
    const T xn = sqrt(T(2) * J2);
#pragma unroll
    for (int k = 0; k < NSTR; k++) n[k] = x[k] / xn;
    return xn - kappa;
}

// Cutting-plane J2 return (J2.jl:21-75, restated from its intent: ϵpII[1,p]
// receives γ, fields under mpts.s, ftol 1e-9, at most 19 corrections).
// Returns the number of corrections (0 = elastic), -1 if the cap was hit.
template <int NSTR, typename T>
__device__ __forceinline__ int j2_return(T *s, T c0, T cr, T Hp, const double *Del, T epII2, T &epII1) {
    const T ftol = T(1e-9);
    T kappa = fmax(cr, c0 + Hp * epII2), n[NSTR];
    T f = yield_j2<NSTR>(s, kappa, n);
    if (!(f > T(0))) return 0;
    T g0 = epII2;
    int it = 1;
    while (fabs(f) > ftol && it < 20) {
        T nD[NSTR];
#pragma unroll
        for (int j = 0; j < NSTR; j++) {
            T t = T(0);
#pragma unroll
            for (int i = 0; i < NSTR; i++) t += n[i] * T(Del[i + NSTR * j]);
            nD[j] = t;
        }
        T nDn = T(0);
#pragma unroll
        for (int j = 0; j < NSTR; j++) nDn += nD[j] * n[j];
        const T dl = f / nDn;
        T Dn[NSTR];
#pragma unroll
        for (int i = 0; i < NSTR; i++) {
            T t = T(0);
#pragma unroll
            for (int j = 0; j < NSTR; j++) t += (dl * T(Del[i + NSTR * j])) * n[j];
            Dn[i] = t;
        }
#pragma unroll
        for (int i = 0; i < NSTR; i++) s[i] -= Dn[i];
        g0 += dl;
        kappa = fmax(cr, c0 + Hp * g0);
        f = yield_j2<NSTR>(s, kappa, n);
        it += 1;
    }
    epII1 = g0;
    return (fabs(f) > ftol) ? -1 : it;
}

}  // namespace ep
