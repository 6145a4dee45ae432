/*
 * elastoplasm.h -- C ABI of libelastoplasm_hip.so, the MI355X (gfx950)
 * implementation of ElastoPlasm.jl's explicit MPM step.
 *
 * The reference exposes its step as an operator table `instr.cairn` built by
 * kwargser (src/home/api/kwargser.jl:49-59) and called once per step by the
 * three wrappers shpfun / mapsto / elastoplast (src/home/program/workflow/
 * worflow.jl:34-36, 80-82).  These entry points replace the bodies of those
 * wrappers and of the time loop around them; Julia keeps its Point / Mesh
 * structs (src/boot/needs/types.jl:38-167) and passes raw pointers into them.
 *
 * Conventions
 *  - Every array argument points at a Julia column-major array: x[d,nmp] is
 *    x[i + dim*p], F[d,d,nmp] is F[i + dim*(j + dim*p)], σᵢ[nstr,nmp] is
 *    sig[c + nstr*p].  Values are Float64.  mesh.bcs.status is Matrix{Bool}
 *    (one byte per entry).
 *  - Every call returns EP_OK (0) or a negative EP_E_* code; ep_last_error()
 *    gives the message of the last failed call on the calling thread.
 *  - One context per host thread; a context owns one HIP stream and its
 *    device buffers.  Host arrays are only read/written inside the call that
 *    receives them; no pointer is retained.
 */
#ifndef ELASTOPLASM_H
#define ELASTOPLASM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EP_ABI_VERSION 3   /* 2: ep_particles_view.Bij, ep_config.domain; 3: ep_stats.resorts, adaptive re-sort */

enum ep_status {
    EP_OK = 0,
    EP_E_ARG = -1,        /* null pointer, size out of range */
    EP_E_CONFIG = -2,     /* unsupported basis/transfer/... (shpfun.jl:50, mapsto.jl:44, plast.jl:31) */
    EP_E_DOMAIN = -3,     /* bsmpm "Invalid position" (bsmpm.jl:75) */
    EP_E_NONFINITE = -4,  /* NaN/Inf in particle state after a step */
    EP_E_HIP = -5,        /* HIP runtime failure */
    EP_E_STATE = -6       /* call out of order (e.g. step before upload) */
};

enum ep_basis { EP_BSMPM = 0, EP_GIMPM = 1, EP_SMPM = 2 };         /* instr.basis.which */
enum ep_transfer { EP_STD = 0, EP_TPIC = 1, EP_APIC = 2 };         /* instr.fwrk.trsfr */
enum ep_deform { EP_FINITE = 0, EP_INFINITESIMAL = 1 };            /* instr.fwrk.deform */
enum ep_plast { EP_PLAST_NONE = 0, EP_DP = 1, EP_J2 = 2 };         /* instr.plast.constitutive */
enum ep_precision { EP_FP64 = 0, EP_FP32 = 1, EP_FP32_SURROGATE = 2 };
/* instr.basis.how, the GIMP domain update (domain.jl:1-52, update.jl:7-29):
 * finite: undeformed, detFij, Fii, Uii; infinitesimal: undeformed, detΔFij,
 * ΔFii, ΔUii.  The reference's kernels other than `undeformed` read
 * mpts.Fᵢⱼ / mpts.ΔFᵢⱼ, which live on mpts.s (types.jl:38-63), so they do
 * not run there; they are restated from their intent. */
enum ep_domain { EP_UNDEFORMED = 0, EP_DETF = 1, EP_FII = 2, EP_UII = 3, EP_DETDF = 4, EP_DFII = 5, EP_DUII = 6 };

/* Setup-time description of one problem: instr + mesh scalars + cmpr.
 * Replaces kwargser's cairn construction (kwargser.jl:49-59) and the
 * scalar parts of setup_mesh / setup_cmpr (setup_mesh.jl:26-99,
 * setup_cmpr.jl:120-143). */
typedef struct ep_config {
    int32_t dim;          /* 2 or 3 */
    int32_t basis;        /* enum ep_basis */
    int32_t transfer;     /* enum ep_transfer */
    int32_t musl;         /* instr.fwrk.musl */
    int32_t deform;       /* enum ep_deform */
    int32_t plast;        /* enum ep_plast */
    int32_t locking;      /* instr.fwrk.locking (F-bar) */
    int32_t precision;    /* enum ep_precision */
    int64_t nmp;          /* mpts.nmp */
    int64_t nel[3];       /* mesh.nel[1:dim] */
    int64_t nno[3];       /* mesh.nno[1:dim] */
    double h[3];          /* mesh.h */
    double x0[3];         /* mesh.x₀ */
    double xB[3][2];      /* mesh.xB[axis, lo/hi] */
    double C_pf;          /* instr.fwrk.C_pf */
    double damping;       /* instr.fwrk.damping */
    double Kc, Gc, Hp;    /* cmpr.Kc, cmpr.Gc, cmpr.Hp */
    double c;             /* cmpr.c, elastic wave speed for get_dt */
    double Del[36];       /* cmpr.Del, nstr x nstr column-major; with finite deformation it must be
                           * the isotropic form setup_cmpr builds (ep_create: EP_E_CONFIG otherwise) */
    double surrogate_tol; /* EP_FP32_SURROGATE: yield-residual tolerance */
    int32_t domain;       /* GIMP domain update, an ep_domain value */
    int32_t nonlocal;     /* instr.nonloc.status: non-local plastic strain (nonlocal.jl:18-47) */
    int32_t reserved[6];
    double nonlocal_ls;   /* instr.nonloc.ls */
} ep_config;

/* Pointers into mesh (types.jl:141-167). */
typedef struct ep_grid_view {
    const double *x;      /* mesh.x [dim, nno] */
    const uint8_t *bcs;   /* mesh.bcs.status [dim, nno] */
} ep_grid_view;

/* Pointers into mpts / mpts.s (types.jl:38-105).  On upload the fields the
 * configuration needs must be non-NULL; on download every non-NULL field is
 * written.  nstr = 3 (2D) or 6 (3D). */
typedef struct ep_particles_view {
    double *x, *v;        /* [dim, nmp]  mpts.x, mpts.s.v */
    double *l0;           /* [dim, nmp]  mpts.ℓ₀ (GIMP) */
    double *l;            /* [dim, nmp]  mpts.ℓ (download) */
    double *F, *b;        /* [dim,dim,nmp]  mpts.s.Fᵢⱼ, mpts.s.bᵢⱼ */
    double *gradv;        /* [dim,dim,nmp]  mpts.s.∇vᵢⱼ (TPIC state) */
    double *eps;          /* [dim,dim,nmp]  mpts.s.ϵᵢⱼ (download) */
    double *sig, *tau;    /* [nstr, nmp]   mpts.s.σᵢ, mpts.s.τᵢ */
    double *epII;         /* [2, nmp]      mpts.s.ϵpII */
    double *rho0, *rho;   /* [nmp] mpts.s.ρ₀, mpts.s.ρ */
    double *Om0, *Om;     /* [nmp] mpts.Ω₀, mpts.Ω */
    double *n0, *n;       /* [nmp] mpts.n₀, mpts.n */
    double *J, *dJ;       /* [nmp] mpts.J; mpts.ΔJ is ignored: ΔJ_p = det ΔF lives only inside a step */
    double *c0, *cr;      /* [nmp] mpts.s.c₀, mpts.s.cᵣ */
    double *phi;          /* [nmp] mpts.s.ϕ */
    double *dlam;         /* [nmp] mpts.s.Δλ */
    double *vmax;         /* [dim] mpts.vmax */
    double *Bij;          /* [dim,dim,nmp] mpts.Bᵢⱼ (APIC state, types.jl:81-83; Bij.jl:14-26);
                           * mpts.Dᵢⱼ and Δnp are recomputed each step (misc.jl:15-27) */
} ep_particles_view;

/* Nodal fields for parity checks (download only), [dim,nno] / [nno]. */
typedef struct ep_grid_fields {
    double *m, *p, *oobf, *a, *v;
} ep_grid_fields;

/* One call of ep_step advances up to `nsteps` steps of elastodynamic! or
 * elastoplastic! (worflow.jl:23-93) inside the window t < T.  Steps past
 * T are skipped on the device (no host round trip per step). */
typedef struct ep_step_params {
    double T;             /* checkpoint time of the window (ΔT in get_dt) */
    double G;             /* gravity magnitude, 9.81 in the reference */
    double tg;            /* gravity ramp end time (time.tg) */
    int32_t ramp;         /* 1: get_spacetime ramp (elastodynamic!), 0: get_g */
    int32_t with_plast;   /* 1: elastoplast, 0: elasto */
    int32_t dt_mode;      /* 0: get_dt CFL rule (physics.jl:21-26), 1: dt_fixed */
    int32_t record_dt;    /* 1: keep each step's dt for ep_get_dts */
    double dt_fixed;
} ep_step_params;

typedef struct ep_step_result {
    double t;             /* time.t[1] after the call */
    double last_dt;
    int64_t steps;        /* steps executed by this call */
    int64_t it;           /* total steps since upload */
} ep_step_result;

typedef struct ep_stats {
    int64_t launches;
    int64_t steps;
    int64_t fallbacks;    /* surrogate predictions replaced by the FP64 return */
    int64_t predictions;  /* surrogate predictions made */
    int64_t j2_maxit;     /* J2 particles that hit the iteration cap */
    double max_residual;  /* largest surrogate yield residual seen */
    int64_t resorts;      /* device re-sorts of the particles (ep_set_resort) */
} ep_stats;

typedef struct ep_ctx ep_ctx;

int ep_version(void);
const char *ep_last_error(void);

/* Replaces kwargser/init_* (kwargser.jl:49-59, shpfun.jl:14-60,
 * mapsto.jl:13-54, update.jl:1-42, elast.jl:93-105, plast.jl:12-34). */
int ep_create(const ep_config *cfg, ep_ctx **out);
void ep_destroy(ep_ctx *ctx);

/* Grid geometry and Dirichlet masks (setup_mesh.jl:115-188, get_bc.jl). */
int ep_upload_grid(ep_ctx *ctx, const ep_grid_view *grid);
/* Particle state (setup_mpts.jl:26-163 output or any later state). */
int ep_upload_particles(ep_ctx *ctx, const ep_particles_view *mp);
int ep_download_particles(ep_ctx *ctx, ep_particles_view *mp);
int ep_download_grid(ep_ctx *ctx, ep_grid_fields *g);

/* Keep the particles in brick-major cell order: the device re-sorts the
 * particles by the cell of their current position, so each block's node box
 * stays compact as the particles move.  every > 0: every `every` steps of
 * ep_step; every < 0 (the default, -5): adaptive, every -every steps the
 * host reads two device counters of the P2G blocks, the summed volume of
 * their node boxes and the number whose box overflowed LDS, and re-sorts
 * when the box volume per block has grown by 10 % over its value at the
 * first check after the previous sort, or when the overflow rate exceeds
 * both 0.2 % of the blocks and twice its rate at that check (the floor a
 * re-sort cannot lower); 0: never.  No counterpart in the
 * reference (its particles keep their setup order); results change only in
 * the summation order of the scatters.
 * Contexts in an ensemble are not re-sorted. */
int ep_set_resort(ep_ctx *ctx, int32_t every);
/* Launch shape (no counterpart in the reference): threads per particle /
 * node block (64, 128, 256 or 512) and cells per particle-sort brick
 * (brick[dim]); the particles are re-sorted into the new order at once.
 * Default 256 threads, bricks of 4x2x4 (3D) / 8x8 (2D) cells. */
int ep_set_layout(ep_ctx *ctx, int32_t block, const int32_t *brick);
/* Time state: time.t[1] (and the step counter). */
int ep_set_time(ep_ctx *ctx, double t);

/* shpfun + mapsto + elasto/elastoplast for up to nsteps steps, with get_dt /
 * get_spacetime on the device (worflow.jl:28-39, 74-85). Blocks until done. */
int ep_step(ep_ctx *ctx, int32_t nsteps, const ep_step_params *par, ep_step_result *res);
/* dt of each step of the last ep_step call with record_dt=1. */
int ep_get_dts(ep_ctx *ctx, double *dts, int64_t cap, int64_t *n);

/* Per-kernel HIP-event timing on the context's stream (no counterpart in the
 * reference).  ep_profile times every kernel launch; ep_profile_kernel times
 * only the launches of the kernel of that name ("p2g", "fbar", ...; the names
 * ep_kernel_times reports), so that the events of the other kernels stay out
 * of a timed run.  Both reset the accumulated times. */
int ep_profile(ep_ctx *ctx, int32_t enable);
int ep_profile_kernel(ep_ctx *ctx, const char *kernel);
int ep_kernel_times(ep_ctx *ctx, int32_t max_kernels, char *names, int32_t name_len,
                    double *total_ms, int64_t *launches, int32_t *count);
int ep_get_stats(ep_ctx *ctx, ep_stats *st);

/* The particle lattice of mpts_populate (mpts_populate.jl:26-56) and the
 * slump filter of get_slump (get_slump.jl:18-95) evaluated on the device.
 * axes[i] holds the n_axis[i] particle coordinates along axis i (the Julia
 * ranges xB_lo + h/(2 ni) : h/ni : xB_hi); the lattice runs along the last
 * axis fastest, then x, then y.  geometry 0 (get_slump): par = {zcut, wl, a,
 * X, Z}, a point is kept when z <= zcut and ((x - X) a - (z - Z) > 0 or
 * z < wl), (X, Z) being the last point of the filter line; geometry 1 (a
 * rectangular column): par = {xmax, zmax, -, -, -}, kept when x <= xmax and
 * z <= zmax.  The kept points go to x_out [dim, cap] in lattice order and
 * their count to *nmp (also when it exceeds cap, which is EP_E_ARG). */
int ep_slump_points(int32_t dim, const double *const *axes, const int64_t *n_axis, int32_t geometry,
                    const double *par, int64_t cap, double *x_out, int64_t *nmp);

/* Gaussian random field at n points by spectral randomisation, the method of
 * GRFS_gauss (src/home/program/utils/setup/mpts/GRF.jl:66-143):
 * g[p] = Σ_h (ab[2h] sin(V_h·x_p) + ab[2h+1] cos(V_h·x_p)) / sqrt(nh), with
 * x [dim, n] column-major and wave vectors waves [nh][dim].  The caller draws
 * the harmonics (seeded), so the field is reproducible. */
int ep_grf_gauss(int32_t dim, int64_t n, const double *x, int32_t nh, const double *waves, const double *ab,
                 double *g);
/* ep_grf_gauss for callers that cannot allocate device memory (code under
 * HIP graph capture, tight loops, harnesses that forbid hipMalloc); it has no
 * counterpart in the reference.  x [dim, n] and g [n] are DEVICE pointers
 * (checked); waves and ab are host arrays, staged into a static device table
 * of at most 16384 harmonics (nh above that is EP_E_ARG).  Calls no
 * hipMalloc/hipFree.  The table is per GPU and shared by every call on that
 * GPU, so calls that overlap on one GPU race.  Synchronous on the null
 * stream. */
int ep_grf_gauss_noalloc(int32_t dim, int64_t n, const double *x, int32_t nh, const double *waves,
                         const double *ab, double *g);

/* ---- ensemble (no counterpart in the reference) ------------------------------
 * Many independent contexts (e.g. slumps of different sizes and materials,
 * each created and uploaded as usual) advanced together: one step of every
 * instance is one replay of a captured HIP graph whose kernels cover the
 * blocks of all instances.  Each instance keeps its own dt (get_dt rule),
 * time window and state; ep_download_particles on an instance reads it back.
 * All instances share dim, the B-spline basis, transfer scheme, deformation
 * framework, plasticity and precision (EP_FP64 or EP_FP32).  The ensemble
 * holds its contexts' device buffers: destroy it before any of its contexts,
 * and do not re-upload a member meanwhile (ep_ensemble_tune does its own
 * re-uploads); members are not re-sorted (ep_set_resort) while they belong
 * to an ensemble. */
typedef struct ep_ensemble ep_ensemble;
typedef struct ep_tune_result {
    int32_t block;            /* threads per particle / node block chosen */
    int32_t brick[3];         /* cells per particle-sort brick chosen (2D: x, z) */
    double ms_per_step;       /* of the chosen layout */
    int32_t ntried;
    int32_t tried_block[16];
    int32_t tried_brick[16][3];
    double tried_ms[16];
} ep_tune_result;
int ep_ensemble_create(ep_ctx *const *ctxs, int32_t n, ep_ensemble **out);
void ep_ensemble_destroy(ep_ensemble *e);
/* res: n entries (may be NULL); record_dt must be 0. */
int ep_ensemble_step(ep_ensemble *e, int32_t nsteps, const ep_step_params *par, ep_step_result *res);
/* Autotuner: times `steps` graph replays for every block size (64, 128,
 * 256) and sort-brick shape (2D 4x4, 8x8, 16x16; 3D 2x2x2, 4x2x4, 8x4x8)
 * from the current state, keeps the fastest and restores the state. */
int ep_ensemble_tune(ep_ensemble *e, int32_t steps, const ep_step_params *par, ep_tune_result *out);
int ep_ensemble_get_stats(ep_ensemble *e, ep_stats *st);

/* ---- learned return mapping (no counterpart in the reference) -------------
 * A network 8-128-128-128-4 (SiLU) that predicts the returned principal
 * Kirchhoff stresses and Δλ of the 3D Drucker-Prager return (DP.jl:1-62)
 * from the trial principal log strains and E, ν, c, φ, ψ; bf16 MFMA with
 * FP32 accumulation, trained with Adam on the device.  Used by contexts
 * created with precision EP_FP32_SURROGATE: every plastic trial state is
 * predicted, checked against the flow rule and yield condition (scaled
 * residual <= tol) and replaced by the FP64 closed-form return otherwise. */
#define EP_SURR_NPARAM 34692
typedef struct ep_surrogate ep_surrogate;
int ep_surrogate_create(uint64_t seed, ep_surrogate **out);
void ep_surrogate_destroy(ep_surrogate *s);
/* FP32 master parameters, [W1 (128x8), b1, W2 (128x128), b2, W3, b3,
 * W4 (4x128), b4], W row-major [out][in]; set resets the Adam moments. */
int ep_surrogate_get_params(ep_surrogate *s, float *params);
int ep_surrogate_set_params(ep_surrogate *s, const float *params);
/* Device training set of n samples: inputs x [n][8], scaled targets y [n][4],
 * raw [n][6] = FP64 (τ'1, τ'2, τ'3, Δλ, fs, S): the returned principal
 * stresses, Δλ, the trial overstress and the residual scale. */
int ep_surrogate_sample(ep_surrogate *s, int64_t n, uint64_t seed);
/* Training samples from a step's own trial states: the plastic trial states
 * that the last ep_step of `ctx` (3D, EP_FP32_SURROGATE) queued for the
 * learned return, as the network's inputs (sorted principal log strains, E,
 * ν, the particle's cohesion and φ, ψ = 0), with the FP64 closed-form return
 * as targets; states that do not end on the cone by a shear return are left
 * out.  append = 0 replaces the training set, 1 appends to it; *count = the
 * samples added.  For fine-tuning a network on the states a simulation
 * actually produces. */
int ep_surrogate_harvest(ep_surrogate *s, ep_ctx *ctx, int32_t append, int64_t *count);
int ep_surrogate_get_samples(ep_surrogate *s, int64_t first, int64_t n, float *x, float *y, double *raw);
/* `steps` Adam steps over consecutive batches of the training set, learning
 * rate cosine from lr0 to lr1; loss[steps] (may be NULL) = batch MSE. */
int ep_surrogate_train(ep_surrogate *s, int64_t steps, int64_t batch, double lr0, double lr1, float *loss);
/* Gradient of the MSE over n samples (n a multiple of 32), no update. */
int ep_surrogate_grad(ep_surrogate *s, const float *x, const float *y, int64_t n, float *grad, float *loss);
int ep_surrogate_predict(ep_surrogate *s, const float *x, int64_t n, float *y);
/* Network used by an EP_FP32_SURROGATE context; tol <= 0 takes
 * cfg.surrogate_tol, or 1e-3 when that is 0.  The network must outlive the
 * context's steps. */
int ep_attach_surrogate(ep_ctx *ctx, ep_surrogate *s, double tol);

#ifdef __cplusplus
}
#endif
#endif
