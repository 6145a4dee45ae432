/*
 * ep_oracle.h -- CPU restatement of ElastoPlasm.jl v0.4.2's explicit MPM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline
 * leg).  No product code may link, load or call it.
 *
 * Layout: every array mirrors the reference's Julia array in column-major
 * order with 0-based C offsets, e.g. x[d,nmp] -> x[i + dim*p],
 * phi_dphi[nn,nmp,dim+1] -> phi[k + nn*(p + nmp*c)], F[d,d,nmp] ->
 * F[i + dim*(j + dim*p)].  Node ids in e2n/p2n are the reference's 1-based
 * ids with 0 meaning "absent".
 *
 * Parity pins (tests/test_oracle_*.py): partition of unity and linear
 * completeness of the three bases, finite-difference derivatives, mass /
 * momentum / gravity-force conservation of P2G, rigid translation through
 * G2P + mUSL, affine velocity fields through the deformation update, matrix
 * log / exp against scipy.linalg, Drucker-Prager and J2 closed forms and
 * yield-surface residuals, the elastic-compaction closed form of
 * test/testset/test_collapse.jl, and an independent numpy restatement on tiny
 * inputs.  LAPACK/OpenBLAS rounding of eigen/det/\ and the thread order of
 * the reference's atomics are parity unpinned (no Julia in this image).
 */
#ifndef EP_ORACLE_H
#define EP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BSMPM = 0, OR_GIMPM = 1, OR_SMPM = 2 };
enum { OR_STD = 0, OR_TPIC = 1, OR_APIC = 2 };
enum { OR_DOM_UNDEFORMED = 0, OR_DOM_DETF = 1, OR_DOM_FII = 2, OR_DOM_UII = 3, OR_DOM_DETDF = 4, OR_DOM_DFII = 5,
       OR_DOM_DUII = 6 };
enum { OR_FINITE = 0, OR_INFINITESIMAL = 1 };
enum { OR_PLAST_NONE = 0, OR_DP = 1, OR_J2 = 2 };
enum { OR_OK = 0, OR_E_DOMAIN = -3 };

/* Mesh{T1,T2} (src/boot/needs/types.jl:141-167), used fields only. */
typedef struct or_mesh {
    int32_t dim, nn;
    int64_t nel[4];          /* nel[1..dim], total at [dim] */
    int64_t nno[4];          /* nno[1..dim], total at [dim] */
    double h[3], x0[3];
    double xB[6];            /* xB[dim,2] column-major: xB[i + dim*k] */
    double *x;               /* [dim,nno] */
    double *m;               /* mᵢ [nno] */
    double *oobf, *a, *p, *v;/* [dim,nno] */
    double *dJ;              /* ΔJ [dim,nno]; first nno entries used */
    int64_t *e2n;            /* [nn,nel] */
    uint8_t *bcs;            /* bcs.status [dim,nno] */
} or_mesh;

/* Point{T1,T2} + Solid{T1,T2} (types.jl:38-105), used fields only. */
typedef struct or_point {
    int64_t nmp;
    int32_t dim, pad0;
    double vmax[3];
    double *phi;             /* ϕ∂ϕ [nn,nmp,dim+1] */
    double *dnp;             /* Δnp [nn,dim,nmp] */
    int64_t *p2e;            /* [nmp] */
    int64_t *p2n;            /* [nn,nmp] */
    double *x, *l0, *l;      /* [dim,nmp] */
    double *n0, *n, *Om0, *Om, *dJ, *J; /* [nmp] */
    double *v;               /* s.v [dim,nmp] */
    double *rho0, *rho, *c0, *cr, *phif, *dlam; /* [nmp] */
    double *epII;            /* ϵpII [2,nmp] */
    double *sig, *tau;       /* σᵢ, τᵢ [nstr,nmp] */
    double *gradv, *dF, *F, *b, *eps, *omg; /* [dim,dim,nmp] */
    double *Bij, *Dij;       /* APIC Bᵢⱼ, Dᵢⱼ [dim,dim,nmp] (types.jl:81-83) */
} or_point;

/* instr.fwrk / instr.basis / cmpr (src/home/api/require.jl:34-88,
 * src/home/program/utils/setup/setup_cmpr.jl:120-143). */
typedef struct or_instr {
    int32_t basis, trsfr, musl, deform, plast, locking;
    int32_t nthreads;        /* 1 = sequential reference order */
    int32_t domain;          /* GIMP domain update, OR_DOM_* */
    double C_pf, damping;
    double Kc, Gc, Hp, c;
    double Del[36];          /* [nstr,nstr] column-major */
    int32_t nonloc, pad1;    /* instr.nonloc.status */
    double ls;               /* instr.nonloc.ls */
} or_instr;

int  or_version(void);
void or_set_threads(int n);

/* get_dt (src/home/program/utils/physics.jl:21-26) */
double or_det(int d, const double *A);
void or_inv(int d, const double *A, double *Ai);
double or_get_dt(or_point *mp, const or_mesh *me, double c, double t, double T);

/* shpfun (src/home/program/workflow/shpfun/shpfun.jl:74-88) */
int  or_shpfun(or_point *mp, const or_mesh *me, const or_instr *in);
/* mapsto (src/home/program/workflow/mapsto/mapsto.jl:70-78) */
void or_p2n(or_point *mp, or_mesh *me, const double *g, const or_instr *in);
void or_solve(or_mesh *me, double dt, const or_instr *in);
void or_n2p(or_point *mp, or_mesh *me, double dt, const or_instr *in);
/* elastoplast (src/home/program/workflow/elastoplast/elastoplast.jl:1-16) */
void or_update(or_point *mp, or_mesh *me, double dt, const or_instr *in);
void or_elast(or_point *mp, const or_instr *in);
void or_plast(or_point *mp, const or_instr *in);
int  or_nonlocal(or_point *mp, const or_mesh *me, const or_instr *in);

/* One step of elastoplastic!/elastodynamic! (src/home/program/workflow/
 * worflow.jl:30-38, 76-84) after dt and g were chosen. with_plast=0 is
 * `elasto`, 1 is `elastoplast`. */
int  or_step(or_point *mp, or_mesh *me, const or_instr *in,
             const double *g, double dt, int with_plast);

/* Point-wise helpers exposed for unit tests of the oracle itself. */
void or_bsmpm_1d(double xi, int type, double *phi, double *dphi);
void or_gimp_1d(double dx, double h, double lp, double *S, double *dS);
void or_smpm_1d(double dx, double h, double *N, double *dN);
void or_sym_eig(int d, const double *A, double *lam, double *V);
double or_det(int d, const double *A);
void or_lu_solve(int n, const double *A, const double *b, double *x);
void or_dp_return(int nstr, double *tau, double c0, double cr, double phi,
                  double Hp, double Kc, double Gc, double epII2,
                  double *epII1, double *dlam);
void or_j2_return(int nstr, double *tau, double c0, double cr, double Hp,
                  const double *Del, double epII2, double *epII1,
                  int *active, int *iters);

#ifdef __cplusplus
}
#endif
#endif
