// ep_api.hip -- C ABI of libelastoplasm_hip.so (include/elastoplasm.h).
//
// The context owns one HIP stream, a control block in device memory and the
// particle / grid buffers.  Upload sorts particles by background cell in
// brick-major order (brick_key) so that a block of neighbouring threads
// touches a compact box of nodes; download undoes the permutation.
#include "../../include/elastoplasm.h"
#include "ep_mlp.h"
#include "ep_batch.h"
#ifndef EP_CPU_CHECK
#include <hipcub/hipcub.hpp>
#endif

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <string>
#include <vector>

using namespace ep;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(EP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

enum KernelId { K_BEGIN, K_ZERO, K_P2G, K_EULER, K_G2P, K_BIJ, K_AUGV, K_DEFORM, K_FBAR, K_SURR, K_END, K_SORT,
                K_NONLOCAL, K_COUNT };
const char *kKernelNames[K_COUNT] = {"begin", "zero_grid", "p2g", "euler", "g2p", "bij", "augv", "deform", "fbar",
                                     "surrogate", "end", "resort", "nonlocal"};

// Default re-sort policy of a context: adaptive, checked every 5 steps;
// re-sort when the p2g node boxes have grown by kBoxGrowth over their volume
// after the last sort, or when more than kSpillFrac of the p2g blocks of the
// window fell back to global atomics.
constexpr int kResortAdaptive = -5;
constexpr double kSpillFrac = 0.002;
// re-sort once the p2g node boxes have grown by this factor over their size
// just after the previous sort
constexpr double kBoxGrowth = 1.1;

// LU with partial pivoting of Del (same algorithm the device solve expects).
void factor_del(int n, const double *A, DelLU &f) {
    std::memcpy(f.LU, A, sizeof(double) * n * n);
    for (int k = 0; k < n; k++) {
        int pk = k;
        double mx = std::fabs(f.LU[k + n * k]);
        for (int i = k + 1; i < n; i++)
            if (std::fabs(f.LU[i + n * k]) > mx) { mx = std::fabs(f.LU[i + n * k]); pk = i; }
        f.piv[k] = pk;
        if (pk != k)
            for (int j = 0; j < n; j++) std::swap(f.LU[k + n * j], f.LU[pk + n * j]);
        const double akk = f.LU[k + n * k];
        if (akk == 0.0) continue;
        const double r = 1.0 / akk;
        for (int i = k + 1; i < n; i++) f.LU[i + n * k] *= r;
        for (int j = k + 1; j < n; j++)
            for (int i = k + 1; i < n; i++) f.LU[i + n * j] -= f.LU[i + n * k] * f.LU[k + n * j];
    }
}

double host_exact_fld(double a, double h) {
    double q = std::floor(a / h);
    if (std::fma(-q, h, a) < 0.0) q -= 1.0;
    else if (std::fma(-(q + 1.0), h, a) >= 0.0) q += 1.0;
    return q;
}

}  // namespace

struct ep_ctx {
    ep_config cfg{};
    int dim = 0, nstr = 0;
    hipStream_t stream = nullptr;
    // grid
    Geom geom{};
    GState gs{};
    std::vector<void *> grid_allocs;   // one allocation per nodal field
    double *axis_block = nullptr;
    int8_t *type_block = nullptr;
    uint8_t *bc_dev = nullptr;
    bool grid_ready = false;
    // particles
    PState<double> ps{};               // FP64 particle state (EP_FP64)
    PState<float> psf{};               // FP32 particle state (EP_FP32, EP_FP32_SURROGATE)
    bool fp32 = false;
    std::vector<void *> part_allocs;   // one allocation per particle field
    long nmp = 0;                      // particles
    long nslot = 0;                    // device slots: particles + padding
    std::vector<long> perm;            // device slot -> host particle index, -1 for padding
    std::vector<long> donor;           // device slot -> a particle of the same cell (for padding)
    std::vector<double> n0_h, rho0_h, Om0_h;
    bool part_ready = false;
    // periodic re-sort into brick-major cell order (k_sort_keys): slot list
    // on the device, spare field buffers of each element type, sort scratch
    // resort_every > 0: every that many steps; < 0: adaptive, the p2g box
    // counters are read every -resort_every steps; 0: never
    int resort_every = kResortAdaptive;
    long since_sort = 0;
    int spill_seen = 0;                // ctrl->spill at the last check
    double spill_floor = 0.0;          // overflow rate (per block and step) just after the last sort
    unsigned long long box_seen = 0;   // ctrl->boxsum at the last check
    double box_floor = 0.0;            // box nodes per block and step just after the last sort
    bool sorted_fresh = true;          // the next check measures that floor
    long resorts = 0;
    bool in_ensemble = false;          // an ensemble holds this context's field pointers
    int *perm_d = nullptr;             // device copy of perm, permuted with the fields
    bool perm_dirty = false;
    double *spare_d[kPermK] = {};
    float *spare_f[kPermK] = {};
    int *spare_i = nullptr;
    std::vector<void *> sort_allocs;
    unsigned *sort_key[2] = {nullptr, nullptr};
    int *sort_val = nullptr, *sort_idx = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    // launch shape: threads per particle/node block and cells per sort brick
    int tb = EP_BLOCK;
    long brick[3] = {4, 2, 4};
    // learned return (EP_FP32_SURROGATE): queue of plastic trial states
    std::vector<void *> q_allocs;
    // non-local regularisation: element bins of the particles, scan scratch
    std::vector<void *> nl_allocs;
    void *nl_tmp = nullptr;
    size_t nl_tmp_bytes = 0;
    ep_surrogate *surr = nullptr;
    float surr_tol = 0.f;
    // control
    Ctrl *ctrl = nullptr;
    double *dt_log = nullptr;
    long dt_cap = 0;
    unsigned long long *vmax_part = nullptr;
    Mat mat{};
    // profiling
    bool profile = false;
    int prof_kid = -1;                  // >= 0: only this kernel is timed
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kernel;
    size_t ev_used = 0;
    double k_ms[K_COUNT] = {};
    long k_launches[K_COUNT] = {};
    ep_stats stats{};
    double t_host = 0.0;
};

static int sort_buffers(ep_ctx *c);

extern "C" int ep_version(void) { return EP_ABI_VERSION; }
extern "C" const char *ep_last_error(void) { return g_err.c_str(); }

extern "C" int ep_create(const ep_config *cfg, ep_ctx **out) {
    if (!cfg || !out) return fail(EP_E_ARG, "ep_create: null argument");
    *out = nullptr;
    if (cfg->dim != 2 && cfg->dim != 3) return fail(EP_E_CONFIG, "dim must be 2 or 3");
    if (cfg->basis < EP_BSMPM || cfg->basis > EP_SMPM)
        return fail(EP_E_CONFIG, "unsupported shape function basis");
    if (cfg->transfer != EP_STD && cfg->transfer != EP_TPIC && cfg->transfer != EP_APIC)
        return fail(EP_E_CONFIG, "unsupported transfer scheme");
    // update.jl:7-29: the domain kernels of each framework; GIMP only
    {
        const int dm = cfg->domain;
        const bool ok = dm == EP_UNDEFORMED ||
                        (cfg->basis == EP_GIMPM && cfg->deform == EP_FINITE && dm >= EP_DETF && dm <= EP_UII) ||
                        (cfg->basis == EP_GIMPM && cfg->deform == EP_INFINITESIMAL && dm >= EP_DETDF && dm <= EP_DUII);
        if (!ok) return fail(EP_E_CONFIG, "domain update does not match the basis / deformation framework");
    }
    if (cfg->deform != EP_FINITE && cfg->deform != EP_INFINITESIMAL)
        return fail(EP_E_CONFIG, "unsupported deformation framework");
    if (cfg->plast < EP_PLAST_NONE || cfg->plast > EP_J2) return fail(EP_E_CONFIG, "InvalidReturnMapping");
    if (cfg->precision != EP_FP64 && cfg->precision != EP_FP32 && cfg->precision != EP_FP32_SURROGATE)
        return fail(EP_E_CONFIG, "unknown precision mode");
    if (cfg->nonlocal && !(cfg->nonlocal_ls > 0.0)) return fail(EP_E_CONFIG, "non-local regularisation needs ls > 0");
    if (cfg->nonlocal && cfg->precision == EP_FP32_SURROGATE)
        return fail(EP_E_CONFIG, "the learned return runs without non-local regularisation");
    if (cfg->precision == EP_FP32_SURROGATE &&
        (cfg->dim != 3 || cfg->deform != EP_FINITE || cfg->plast != EP_DP))
        return fail(EP_E_CONFIG, "the learned return needs 3D finite deformation with Drucker-Prager");
    if (cfg->nmp <= 0) return fail(EP_E_ARG, "nmp must be positive");
    if (cfg->nmp > 0x7fffffffL) return fail(EP_E_ARG, "nmp exceeds 2^31 - 1");
    for (int i = 0; i < cfg->dim; i++) {
        if (cfg->nel[i] < 1 || cfg->nno[i] < 2 || cfg->nno[i] > (1 << 24) || !(cfg->h[i] > 0.0))
            return fail(EP_E_ARG, "bad mesh extent");
    }
    if (cfg->deform == EP_FINITE) {
        // the finite-strain update works in the principal frame of the trial
        // b, which needs an isotropic Del (the reference's, from E and ν):
        // equal normal diagonal a, equal normal off-diagonal b, shear (a-b)/2
        // on the diagonal, zeros elsewhere
        const int ns = cfg->dim == 2 ? 3 : 6, nd = cfg->dim;
        const double *Dl = cfg->Del;
        const double a = Dl[0], b = Dl[1], tol = 1e-12 * std::fabs(a);
        bool iso = a > 0.0;
        for (int i = 0; i < ns && iso; i++)
            for (int j = 0; j < ns; j++) {
                const double v = Dl[i + ns * j];
                double want = 0.0;
                if (i < nd && j < nd) want = i == j ? a : b;
                else if (i == j) want = 0.5 * (a - b);
                const bool zero = !(i < nd && j < nd) && i != j;
                if (zero ? v != 0.0 : !(std::fabs(v - want) <= tol)) { iso = false; break; }
            }
        if (!iso) return fail(EP_E_CONFIG, "finite deformation needs an isotropic Del (the principal-frame update)");
    }
    ep_ctx *c = new ep_ctx();
    c->cfg = *cfg;
    c->fp32 = cfg->precision != EP_FP64;
    c->dim = cfg->dim;
    c->nstr = cfg->dim == 2 ? 3 : 6;
    if (c->dim == 2) { c->brick[0] = 8; c->brick[1] = 8; c->brick[2] = 1; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(EP_E_HIP, "hipStreamCreate failed");
    }
    Mat &M = c->mat;
    M.Kc = cfg->Kc; M.Gc = cfg->Gc; M.Hp = cfg->Hp; M.c = cfg->c;
    M.E = 9.0 * cfg->Kc * cfg->Gc / (3.0 * cfg->Kc + cfg->Gc);
    M.nu = (3.0 * cfg->Kc - 2.0 * cfg->Gc) / (2.0 * (3.0 * cfg->Kc + cfg->Gc));
    M.C_pf = cfg->C_pf; M.damping = cfg->damping;
    std::memcpy(M.Del, cfg->Del, sizeof(double) * c->nstr * c->nstr);
    factor_del(c->nstr, cfg->Del, M.lu);
    M.finite = cfg->deform == EP_FINITE;
    M.tpic = cfg->transfer == EP_TPIC;
    M.apic = cfg->transfer == EP_APIC;
    M.domain = cfg->basis == EP_GIMPM ? cfg->domain : EP_UNDEFORMED;
    M.musl = cfg->musl != 0;
    M.locking = cfg->locking != 0;
    M.plast = cfg->plast;
    M.nonlocal = cfg->nonlocal != 0 && cfg->plast != EP_PLAST_NONE;
    M.nl.ls = cfg->nonlocal_ls;
    for (int i = 0; i < 3; i++)
        M.nl.nnel[i] = i < cfg->dim ? int(std::ceil(cfg->nonlocal_ls / cfg->h[i])) : 0;
    Geom &g = c->geom;
    long nno_tot = 1;
    for (int i = 0; i < 3; i++) {
        g.nno[i] = i < c->dim ? int(cfg->nno[i]) : 1;
        g.nel[i] = i < c->dim ? int(cfg->nel[i]) : 1;
        g.h[i] = i < c->dim ? cfg->h[i] : 1.0;
        g.ih[i] = 1.0 / g.h[i];
        g.x0[i] = i < c->dim ? cfg->x0[i] : 0.0;
        nno_tot *= g.nno[i];
    }
    g.nno_tot = nno_tot;
    if (hipMalloc(&c->ctrl, sizeof(Ctrl)) != hipSuccess) {
        ep_destroy(c);
        return fail(EP_E_HIP, "hipMalloc(ctrl) failed");
    }
    Ctrl h{};
    if (hipMemcpy(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess) {
        ep_destroy(c);
        return fail(EP_E_HIP, "hipMemcpy(ctrl) failed");
    }
    *out = c;
    return EP_OK;
}

extern "C" void ep_destroy(ep_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void *a : c->grid_allocs) hipFree(a);
    hipFree(c->axis_block);
    hipFree(c->type_block);
    hipFree(c->bc_dev);
    for (void *a : c->part_allocs) hipFree(a);
    for (void *a : c->q_allocs) hipFree(a);
    for (void *a : c->sort_allocs) hipFree(a);
    for (void *a : c->nl_allocs) hipFree(a);
    hipFree(c->ctrl);
    hipFree(c->dt_log);
    hipFree(c->vmax_part);
    for (auto e : c->ev_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// B-spline node type by floating-point equality with xB (bsmpm.jl:63-76).
static int8_t bsmpm_type(double xn, double lo, double hi, double dx) {
    if (xn == lo) return 1;
    if (xn == lo + dx) return 2;
    if (lo + dx < xn && xn < hi - dx) return 3;
    if (xn == hi - dx) return 4;
    if (xn == hi) return 1;
    return 0;
}

extern "C" int ep_upload_grid(ep_ctx *c, const ep_grid_view *gv) {
    if (!c || !gv || !gv->x || !gv->bcs) return fail(EP_E_ARG, "ep_upload_grid: null argument");
    const int D = c->dim;
    Geom &g = c->geom;
    const long nno = g.nno_tot;
    // per-axis coordinates of a tensor-product grid (get_coords.jl:36-52)
    std::vector<double> ax;
    std::vector<int8_t> ty;
    long off[3] = {0, 0, 0};
    for (int i = 0; i < D; i++) {
        off[i] = long(ax.size());
        for (int k = 0; k < g.nno[i]; k++) {
            long no;
            if (D == 2) no = (i == 0) ? long(k) * g.nno[1] : long(k);
            else no = (i == 0) ? long(k) * g.nno[2] : (i == 1 ? long(k) * g.nno[2] * g.nno[0] : long(k));
            ax.push_back(gv->x[i + D * no]);
            ty.push_back(bsmpm_type(gv->x[i + D * no], c->cfg.xB[i][0], c->cfg.xB[i][1], g.h[i]));
        }
    }
    // verify the tensor-product structure on every node
    for (long no = 0; no < nno; no++) {
        long idx[3];
        if (D == 2) { idx[1] = no % g.nno[1]; idx[0] = no / g.nno[1]; }
        else { idx[2] = no % g.nno[2]; long r = no / g.nno[2]; idx[0] = r % g.nno[0]; idx[1] = r / g.nno[0]; }
        for (int i = 0; i < D; i++)
            if (gv->x[i + D * no] != ax[off[i] + idx[i]])
                return fail(EP_E_ARG, "mesh.x is not the reference's tensor-product node layout");
    }
    if (c->cfg.basis == EP_BSMPM) {
        // particles never reach invalid-type nodes only if every node is typed
        for (size_t k = 0; k < ty.size(); k++)
            if (ty[k] == 0) {
                // kept: the kernel raises EP_E_DOMAIN if a particle touches it
            }
    }
    std::vector<uint8_t> bc(nno, 0);
    for (long no = 0; no < nno; no++)
        for (int i = 0; i < D; i++)
            if (gv->bcs[i + D * no]) bc[no] |= uint8_t(1u << i);
    hipFree(c->axis_block);
    hipFree(c->type_block);
    hipFree(c->bc_dev);
    for (void *a : c->grid_allocs) hipFree(a);
    c->grid_allocs.clear();
    c->axis_block = nullptr; c->type_block = nullptr; c->bc_dev = nullptr;
    HIPCHK(hipMalloc(&c->axis_block, sizeof(double) * ax.size()));
    HIPCHK(hipMalloc(&c->type_block, ty.size()));
    HIPCHK(hipMalloc(&c->bc_dev, nno));
    HIPCHK(hipMemcpy(c->axis_block, ax.data(), sizeof(double) * ax.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->type_block, ty.data(), ty.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->bc_dev, bc.data(), nno, hipMemcpyHostToDevice));
    for (int i = 0; i < 3; i++) {
        g.X[i] = i < D ? c->axis_block + off[i] : c->axis_block;
        g.T[i] = i < D ? c->type_block + off[i] : c->type_block;
        // the run of interior-type (3) nodes around the middle of the axis
        g.t3lo[i] = 1;
        g.t3hi[i] = 0;
        if (i < D && c->cfg.basis == EP_BSMPM) {
            const int8_t *t = ty.data() + off[i];
            int lo = g.nno[i] / 2, hi = lo;
            if (t[lo] == 3) {
                while (lo > 0 && t[lo - 1] == 3) lo--;
                while (hi + 1 < g.nno[i] && t[hi + 1] == 3) hi++;
                g.t3lo[i] = lo;
                g.t3hi[i] = hi;
            }
        }
    }
    // m, p[D], f[D], a[D], v[D], pa[D], dJ: one allocation each
    GState &G = c->gs;
    G = GState{};
    auto galloc = [&](double **dst) -> int {
        void *a = nullptr;
        HIPCHK(hipMalloc(&a, sizeof(double) * nno));
        HIPCHK(hipMemset(a, 0, sizeof(double) * nno));
        c->grid_allocs.push_back(a);
        *dst = static_cast<double *>(a);
        return EP_OK;
    };
    int r;
    if ((r = galloc(&G.m)) != EP_OK) return r;
    for (int d = 0; d < D; d++) {
        if ((r = galloc(&G.p[d])) != EP_OK || (r = galloc(&G.f[d])) != EP_OK || (r = galloc(&G.a[d])) != EP_OK ||
            (r = galloc(&G.v[d])) != EP_OK || (r = galloc(&G.pa[d])) != EP_OK)
            return r;
    }
    if ((r = galloc(&G.dJ)) != EP_OK) return r;
    G.bc = c->bc_dev;
    G.nno = nno;
    c->grid_ready = true;
    return EP_OK;
}

// Sort key of a particle: its background cell in brick-major order.  Bricks
// of 4x2x4 cells (3D, x-y-z) or 8x8 cells (2D) hold ~256 particles at the
// reference's 2^d particles per cell, so a 256-thread block touches a
// compact node box; inside a brick cells run z fastest, as the grid.
static long brick_key(const ep_ctx *c, const double *x, long p) {
    const int D = c->dim;
    const Geom &g = c->geom;
    const long *b = c->brick;   // cells per brick along each axis
    long e[3];
    for (int i = 0; i < D; i++) {
        double q = host_exact_fld(x[i + D * p] - g.x0[i], g.h[i]);
        if (!(q >= 0.0)) q = 0.0;
        if (q > g.nel[i] - 1) q = g.nel[i] - 1;
        e[i] = long(q);
    }
    return cell_key(D, e, b[0], b[1], b[2], g.nel);
}

// Particle fields in device order, one allocation per field, stored as T.
template <typename T>
static int upload_fields(ep_ctx *c, const ep_particles_view *v, PState<T> &P) {
    const int D = c->dim, NS = c->nstr, DD = D * D;
    const long n = c->nslot;
    const Mat &M = c->mat;
    // one device allocation per particle field
    for (void *a : c->part_allocs) hipFree(a);
    c->part_allocs.clear();
    P = PState<T>{};
    bool oom = false;
    auto take = [&]() -> T * {
        void *a = nullptr;
        if (hipMalloc(&a, sizeof(T) * n) != hipSuccess) { oom = true; return nullptr; }
        c->part_allocs.push_back(a);
        return static_cast<T *>(a);
    };
    auto take_d = [&]() -> double * {
        void *a = nullptr;
        if (hipMalloc(&a, sizeof(double) * n) != hipSuccess) { oom = true; return nullptr; }
        c->part_allocs.push_back(a);
        return static_cast<double *>(a);
    };
    for (int k = 0; k < 2; k++) for (int i = 0; i < D; i++) P.x[k][i] = take_d();
    for (int i = 0; i < D; i++) P.v[i] = take();
    for (int i = 0; i < DD; i++) P.F[i] = take();
    for (int i = 0; i < (D == 2 ? 3 : 6); i++) P.b[i] = take();
    for (int i = 0; i < NS; i++) P.s[i] = take();
    P.epII = take(); P.dlam = take(); P.J = take(); P.rho = take();
    P.Om0 = take(); P.c0 = take(); P.cr = take(); P.phi = take();
    P.sphi = take(); P.cphi = take();
    if (c->cfg.basis == EP_GIMPM) for (int i = 0; i < D; i++) P.l[i] = take();
    if (M.domain != EP_UNDEFORMED) for (int i = 0; i < D; i++) P.l0[i] = take();
    if (M.tpic) for (int i = 0; i < DD; i++) P.gv[i] = take();
    if (M.apic) for (int i = 0; i < DD; i++) P.Bm[i] = take();
    if (M.nonlocal) P.epII2 = take();
    // re-sort: the device slot list and one spare buffer per element type
    auto take_i = [&]() -> int * {
        void *a = nullptr;
        if (hipMalloc(&a, sizeof(int) * n) != hipSuccess) { oom = true; return nullptr; }
        c->part_allocs.push_back(a);
        return static_cast<int *>(a);
    };
    for (int k = 0; k < kPermK; k++) {
        c->spare_d[k] = take_d();
        c->spare_f[k] = nullptr;
        if constexpr (std::is_same<T, float>::value) c->spare_f[k] = take();
    }
    c->perm_d = take_i();
    c->spare_i = take_i();
    if (oom) return fail(EP_E_HIP, "hipMalloc of particle fields failed");
    if (M.nonlocal) {
        for (void *a : c->nl_allocs) hipFree(a);
        c->nl_allocs.clear();
        NLocal &nl = c->mat.nl;
        nl.ncell = 1;
        for (int i = 0; i < D; i++) nl.ncell *= c->geom.nel[i];
        auto ti = [&](size_t k) -> int * {
            void *a = nullptr;
            if (hipMalloc(&a, sizeof(int) * k) != hipSuccess) { oom = true; return nullptr; }
            c->nl_allocs.push_back(a);
            return static_cast<int *>(a);
        };
        nl.count = ti(size_t(nl.ncell) + 1);
        nl.start = ti(size_t(nl.ncell) + 1);
        nl.list = ti(size_t(n));
        nl.cell = ti(size_t(n));
        nl.orig = c->perm_d;
#ifndef EP_CPU_CHECK
        if (!oom) {
            size_t bytes = 0;
            if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, nl.count, nl.start, int(nl.ncell + 1), c->stream) !=
                hipSuccess)
                oom = true;
            void *a = nullptr;
            if (!oom && hipMalloc(&a, bytes ? bytes : 1) == hipSuccess) {
                c->nl_allocs.push_back(a);
                c->nl_tmp = a;
                c->nl_tmp_bytes = bytes;
            } else {
                oom = true;
            }
        }
#endif
        if (oom) return fail(EP_E_HIP, "hipMalloc of the non-local bins failed");
    }
    P.n = n;
    // host staging in device order.  A padding slot copies a particle of its
    // cell (position, material constants) unless the field has a value for
    // padding (dead < 1e300); its volume Ω₀ = 0 marks it for the kernels.
    auto put = [&](auto *dst, auto getter, double dead) -> int {
        using E = std::remove_pointer_t<decltype(dst)>;
        std::vector<E> st(n);
        for (long q = 0; q < n; q++) {
            const long p = c->perm[q];
            st[q] = p >= 0 ? E(getter(p)) : (dead < 1e300 ? E(dead) : E(getter(c->donor[q])));
        }
        hipError_t e = hipMemcpy(dst, st.data(), sizeof(E) * n, hipMemcpyHostToDevice);
        return e == hipSuccess ? EP_OK : fail(EP_E_HIP, "hipMemcpy particle upload failed");
    };
    int r;
    const double COPY = 1e301;
#define PUTD(dst, expr, dead) do { if ((r = put(dst, [&](long p) { return (expr); }, dead)) != EP_OK) return r; } while (0)
#define PUT(dst, expr) PUTD(dst, expr, COPY)
    for (int i = 0; i < D; i++) { PUT(P.x[0][i], v->x[i + D * p]); PUT(P.x[1][i], v->x[i + D * p]); }
    for (int i = 0; i < D; i++) PUTD(P.v[i], v->v[i + D * p], 0.0);
    for (int i = 0; i < DD; i++) PUTD(P.F[i], v->F[i + DD * p], (i % (D + 1) == 0) ? 1.0 : 0.0);
    if (M.finite) {
        // symmetric part of bᵢⱼ: 2D xx,yy,xy; 3D xx,yy,zz,yz,xz,xy
        static const int i2[3][2] = {{0, 0}, {1, 1}, {0, 1}};
        static const int i3[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
        for (int k = 0; k < (D == 2 ? 3 : 6); k++) {
            const int a = D == 2 ? i2[k][0] : i3[k][0], bb = D == 2 ? i2[k][1] : i3[k][1];
            // the FP32 path stores B = b - I
            const double off = (c->fp32 && a == bb) ? 1.0 : 0.0;
            PUTD(P.b[k], 0.5 * (v->b[a + D * bb + DD * p] + v->b[bb + D * a + DD * p]) - off,
                 (a == bb ? 1.0 : 0.0) - off);
        }
        for (int k = 0; k < NS; k++) PUTD(P.s[k], v->tau[k + NS * p], 0.0);
    } else {
        for (int k = 0; k < (D == 2 ? 3 : 6); k++) PUT(P.b[k], (k < D && !c->fp32) ? 1.0 : 0.0);
        for (int k = 0; k < NS; k++) PUTD(P.s[k], v->sig[k + NS * p], 0.0);
    }
    PUTD(P.epII, v->epII[0 + 2 * p], 0.0);
    PUTD(P.dlam, v->dlam ? v->dlam[p] : 0.0, 0.0);
    PUTD(P.J, v->J[p], 1.0);
    PUT(P.rho, M.finite ? v->rho0[p] : v->rho[p]);
    PUTD(P.Om0, v->Om0[p], 0.0);
    PUT(P.c0, v->c0[p]);
    PUT(P.cr, v->cr[p]);
    PUT(P.phi, v->phi[p]);
    // `undeformed`: ℓ = ℓ₀ at every step, so ℓ comes from ℓ₀
    if (c->cfg.basis == EP_GIMPM)
        for (int i = 0; i < D; i++) PUT(P.l[i], M.domain != EP_UNDEFORMED ? v->l[i + D * p] : v->l0[i + D * p]);
    if (M.domain != EP_UNDEFORMED) for (int i = 0; i < D; i++) PUT(P.l0[i], v->l0[i + D * p]);
    if (M.tpic) for (int i = 0; i < DD; i++) PUTD(P.gv[i], v->gradv[i + DD * p], 0.0);
    if (M.apic) for (int i = 0; i < DD; i++) PUTD(P.Bm[i], v->Bij[i + DD * p], 0.0);
    {
        std::vector<int> pi(n);
        for (long q = 0; q < n; q++) pi[q] = int(c->perm[q]);
        if (hipMemcpy(c->perm_d, pi.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess)
            return fail(EP_E_HIP, "hipMemcpy particle upload failed");
    }
    {
        const dim3 gf((n + 255) / 256);
        if (D == 2) EP_LAUNCH((k_friction<3, T>), gf, 256, c->stream, P.phi, P.sphi, P.cphi, n);
        else EP_LAUNCH((k_friction<6, T>), gf, 256, c->stream, P.phi, P.sphi, P.cphi, n);
        HIPCHK(hipGetLastError());
    }
    for (void *a : c->sort_allocs) hipFree(a);
    c->sort_allocs.clear();
    c->sort_key[0] = c->sort_key[1] = nullptr;
    c->sort_val = c->sort_idx = nullptr;
    c->sort_tmp = nullptr;
    c->sort_tmp_bytes = 0;
    c->since_sort = 0;
    c->sorted_fresh = true;
    c->perm_dirty = false;
#undef PUT
#undef PUTD
    // a context that re-sorts (the default) gets its re-sort buffers now
    return c->resort_every != 0 ? sort_buffers(c) : EP_OK;
}

extern "C" int ep_upload_particles(ep_ctx *c, const ep_particles_view *v) {
    if (!c || !v) return fail(EP_E_ARG, "ep_upload_particles: null argument");
    if (!c->grid_ready) return fail(EP_E_STATE, "upload the grid first");
    const int D = c->dim, NS = c->nstr, DD = D * D;
    const long n = c->cfg.nmp;
    const Mat &M = c->mat;
    if (!v->x || !v->v || !v->F || !v->epII || !v->rho0 || !v->Om0 || !v->J || !v->c0 || !v->cr || !v->phi)
        return fail(EP_E_ARG, "missing particle field (x, v, F, epII, rho0, Om0, J, c0, cr, phi)");
    if (M.finite && (!v->b || !v->tau)) return fail(EP_E_ARG, "finite deformation needs b and tau");
    if (!M.finite && (!v->sig || !v->rho)) return fail(EP_E_ARG, "infinitesimal deformation needs sig and rho");
    if (c->cfg.basis == EP_GIMPM && !v->l0) return fail(EP_E_ARG, "GIMP needs l0");
    if (M.domain != EP_UNDEFORMED && !v->l) return fail(EP_E_ARG, "a GIMP domain update needs l");
    if (M.tpic && !v->gradv) return fail(EP_E_ARG, "TPIC needs gradv");
    if (M.apic && !v->Bij) return fail(EP_E_ARG, "APIC needs Bij");
    for (long p = 0; p < n; p++)
        if (!(v->Om0[p] > 0.0)) return fail(EP_E_ARG, "particle volumes Om0 must be positive");
    // cell-sorted order.  Cell runs could be padded to a multiple of G slots
    // (padding slots have Ω₀ = 0 and every kernel skips them); no kernel
    // needs aligned runs, so G = 1.
    std::vector<long> key(n), order(n);
    for (long p = 0; p < n; p++) key[p] = brick_key(c, v->x, p);
    std::iota(order.begin(), order.end(), 0L);
    std::stable_sort(order.begin(), order.end(), [&](long a, long b) { return key[a] < key[b]; });
    c->perm.clear();
    c->donor.clear();
    const long G = 1;
    for (long a = 0; a < n;) {
        long b = a;
        while (b < n && key[order[b]] == key[order[a]]) b++;
        for (long k = a; k < b; k++) { c->perm.push_back(order[k]); c->donor.push_back(order[k]); }
        while (long(c->perm.size()) % G) { c->perm.push_back(-1); c->donor.push_back(order[a]); }
        a = b;
    }
    c->nmp = n;
    c->nslot = long(c->perm.size());
    if (c->nslot > 0x7fffffffL) return fail(EP_E_ARG, "too many particle slots");
    {
        const int r = c->fp32 ? upload_fields<float>(c, v, c->psf) : upload_fields<double>(c, v, c->ps);
        if (r != EP_OK) return r;
    }
    if (c->cfg.precision == EP_FP32_SURROGATE) {
        for (void *a : c->q_allocs) hipFree(a);
        c->q_allocs.clear();
        auto qa = [&](size_t bytes) -> void * {
            void *a = nullptr;
            if (hipMalloc(&a, bytes) != hipSuccess) return nullptr;
            c->q_allocs.push_back(a);
            return a;
        };
        SurrQ &q = c->mat.q;
        const long ns = c->nslot;
        q.idx = static_cast<int *>(qa(sizeof(int) * ns));
        bool ok = q.idx != nullptr;
        for (int k = 0; k < 3; k++) ok = ok && (q.e[k] = static_cast<float *>(qa(sizeof(float) * ns))) != nullptr;
        for (int k = 0; k < 9; k++) ok = ok && (q.V[k] = static_cast<float *>(qa(sizeof(float) * ns))) != nullptr;
        if (!ok) return fail(EP_E_HIP, "hipMalloc of the learned-return queue failed");
    }
    c->n0_h.assign(n, 0.1);
    if (v->n0) for (long p = 0; p < n; p++) c->n0_h[p] = v->n0[p];
    c->rho0_h.assign(v->rho0, v->rho0 + n);
    c->Om0_h.assign(v->Om0, v->Om0 + n);
    // control block: vmax from mpts.vmax, parity 0
    Ctrl h{};
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    for (int i = 0; i < 3; i++) {
        double vm = (v->vmax && i < D) ? v->vmax[i] : 0.0;
        unsigned long long bits;
        std::memcpy(&bits, &vm, 8);
        h.vmax[i] = bits;
    }
    h.parity = 0;
    h.err = 0;
    hipFree(c->vmax_part);
    c->vmax_part = nullptr;
    h.nblk = int((c->nslot + c->tb - 1) / c->tb);
    HIPCHK(hipMalloc(&c->vmax_part, sizeof(unsigned long long) * 3 * size_t(h.nblk)));
    HIPCHK(hipMemset(c->vmax_part, 0, sizeof(unsigned long long) * 3 * size_t(h.nblk)));
    h.vmax_part = c->vmax_part;
    HIPCHK(hipMemcpy(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice));
    c->part_ready = true;
    return EP_OK;
}

extern "C" int ep_set_resort(ep_ctx *c, int32_t every) {
    if (!c) return fail(EP_E_ARG, "null context");
    c->resort_every = every;
    c->since_sort = 0;
    return EP_OK;
}

extern "C" int ep_set_time(ep_ctx *c, double t) {
    if (!c) return fail(EP_E_ARG, "null context");
    Ctrl h{};
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    h.t = t;
    HIPCHK(hipMemcpy(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice));
    return EP_OK;
}

template <typename T>
static int download_fields(ep_ctx *c, ep_particles_view *v, const PState<T> &P, int par) {
    const int D = c->dim, NS = c->nstr, DD = D * D;
    const long n = c->nslot;
    const Mat &M = c->mat;
    std::vector<double> st(n);
    std::vector<T> stt(n);
    auto get = [&](const auto *src) -> int {
        using E = std::remove_cv_t<std::remove_pointer_t<decltype(src)>>;
        hipError_t e;
        if constexpr (std::is_same<E, double>::value) {
            e = hipMemcpy(st.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost);
        } else {
            e = hipMemcpy(stt.data(), src, sizeof(E) * n, hipMemcpyDeviceToHost);
            for (long q = 0; q < n; q++) st[q] = double(stt[q]);
        }
        return e == hipSuccess ? EP_OK : fail(EP_E_HIP, "hipMemcpy particle download failed");
    };
    int r;
#define GET(src, stmt) do { if ((r = get(src)) != EP_OK) return r; for (long q = 0; q < n; q++) { const long p = c->perm[q]; if (p < 0) continue; const double val = st[q]; stmt; } } while (0)
    if (v->x) for (int i = 0; i < D; i++) GET(P.x[par][i], v->x[i + D * p] = val);
    if (v->v) for (int i = 0; i < D; i++) GET(P.v[i], v->v[i + D * p] = val);
    if (v->F) for (int i = 0; i < DD; i++) GET(P.F[i], v->F[i + DD * p] = val);
    const long np_ = c->nmp;
    std::vector<double> J(np_), tau;
    GET(P.J, J[p] = val);
    if (v->J) for (long p = 0; p < np_; p++) v->J[p] = J[p];
    if (v->Om) for (long p = 0; p < np_; p++) v->Om[p] = J[p] * c->Om0_h[p];
    if (M.finite) {
        if (v->rho) for (long p = 0; p < np_; p++) v->rho[p] = c->rho0_h[p] / J[p];
        if (v->n) for (long p = 0; p < np_; p++) v->n[p] = 1.0 - 1.0 / J[p] * (1.0 - c->n0_h[p]);
    }
    if (M.finite && v->b) {
        static const int i2[3][2] = {{0, 0}, {1, 1}, {0, 1}};
        static const int i3[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
        for (int k = 0; k < (D == 2 ? 3 : 6); k++) {
            const int a = D == 2 ? i2[k][0] : i3[k][0], bb = D == 2 ? i2[k][1] : i3[k][1];
            const double off = (c->fp32 && a == bb) ? 1.0 : 0.0;
            GET(P.b[k], (v->b[a + D * bb + DD * p] = val + off, v->b[bb + D * a + DD * p] = val + off));
        }
    }
    tau.assign(size_t(NS) * np_, 0.0);
    for (int k = 0; k < NS; k++) GET(P.s[k], tau[k + NS * p] = val);
    if (M.finite) {
        if (v->tau) std::memcpy(v->tau, tau.data(), sizeof(double) * NS * np_);
        if (v->sig) for (long p = 0; p < np_; p++) for (int k = 0; k < NS; k++) v->sig[k + NS * p] = tau[k + NS * p] / J[p];
        if (v->eps) {
            // ϵᵢⱼ = tensor(Del \ τ, ½) (DP.jl:64)
            for (long p = 0; p < np_; p++) {
                double y[6] = {0, 0, 0, 0, 0, 0}, yy[6];
                for (int k = 0; k < NS; k++) y[k] = tau[k + NS * p];
                // forward/back substitution with the host LU of Del
                for (int k = 0; k < NS; k++) yy[k] = y[k];
                for (int k = 0; k < NS; k++) { int pk = M.lu.piv[k]; if (pk != k) std::swap(yy[k], yy[pk]); }
                for (int i = 0; i < NS; i++) { double s = yy[i]; for (int k = 0; k < i; k++) s -= M.lu.LU[i + NS * k] * yy[k]; yy[i] = s; }
                double x[6];
                for (int i = NS - 1; i >= 0; i--) { double s = yy[i]; for (int k = i + 1; k < NS; k++) s -= M.lu.LU[i + NS * k] * x[k]; x[i] = s / M.lu.LU[i + NS * i]; }
                double *e = v->eps + DD * p;
                if (D == 2) { e[0] = x[0]; e[3] = x[1]; e[1] = e[2] = 0.5 * x[2]; }
                else { e[0] = x[0]; e[4] = x[1]; e[8] = x[2]; e[5] = e[7] = 0.5 * x[3]; e[2] = e[6] = 0.5 * x[4]; e[1] = e[3] = 0.5 * x[5]; }
            }
        }
    } else {
        if (v->sig) std::memcpy(v->sig, tau.data(), sizeof(double) * NS * np_);
    }
    if (v->epII) GET(P.epII, (v->epII[0 + 2 * p] = val, v->epII[1 + 2 * p] = val));
    // ϵpII[2]: the non-local value of the last step (plast.jl:49-60)
    if (v->epII && M.nonlocal) GET(P.epII2, v->epII[1 + 2 * p] = val);
    if (v->dlam) GET(P.dlam, v->dlam[p] = val);
    if (v->gradv && M.tpic) for (int i = 0; i < DD; i++) GET(P.gv[i], v->gradv[i + DD * p] = val);
    if (v->Bij && M.apic) for (int i = 0; i < DD; i++) GET(P.Bm[i], v->Bij[i + DD * p] = val);
    if (v->l && c->cfg.basis == EP_GIMPM) for (int i = 0; i < D; i++) GET(P.l[i], v->l[i + D * p] = val);
    if (v->l0 && c->cfg.basis == EP_GIMPM)
        for (int i = 0; i < D; i++) GET(M.domain != EP_UNDEFORMED ? P.l0[i] : P.l[i], v->l0[i + D * p] = val);
    if (v->c0) GET(P.c0, v->c0[p] = val);
    if (v->cr) GET(P.cr, v->cr[p] = val);
    if (v->phi) GET(P.phi, v->phi[p] = val);
    if (v->rho0) for (long p = 0; p < np_; p++) v->rho0[p] = c->rho0_h[p];
    if (v->Om0) for (long p = 0; p < np_; p++) v->Om0[p] = c->Om0_h[p];
    if (v->n0) for (long p = 0; p < np_; p++) v->n0[p] = c->n0_h[p];
    if (!M.finite && v->rho) GET(P.rho, v->rho[p] = val);
#undef GET
    return EP_OK;
}

extern "C" int ep_download_particles(ep_ctx *c, ep_particles_view *v) {
    if (!c || !v) return fail(EP_E_ARG, "ep_download_particles: null argument");
    if (!c->part_ready) return fail(EP_E_STATE, "no particles uploaded");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int D = c->dim, NS = c->nstr, DD = D * D;
    const long n = c->nmp;
    const Mat &M = c->mat;
    Ctrl h{};
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    if (c->perm_dirty) {
        std::vector<int> pi(size_t(c->nslot));
        HIPCHK(hipMemcpy(pi.data(), c->perm_d, sizeof(int) * size_t(c->nslot), hipMemcpyDeviceToHost));
        for (long q = 0; q < c->nslot; q++) c->perm[q] = pi[q];
        c->perm_dirty = false;
    }
    {
        const int r = c->fp32 ? download_fields<float>(c, v, c->psf, h.parity)
                              : download_fields<double>(c, v, c->ps, h.parity);
        if (r != EP_OK) return r;
    }
    if (v->vmax)
        for (int i = 0; i < D; i++) std::memcpy(&v->vmax[i], &h.vmax[i], 8);
    return EP_OK;
}

extern "C" int ep_download_grid(ep_ctx *c, ep_grid_fields *gf) {
    if (!c || !gf) return fail(EP_E_ARG, "ep_download_grid: null argument");
    if (!c->grid_ready) return fail(EP_E_STATE, "no grid uploaded");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int D = c->dim;
    const long nno = c->gs.nno;
    std::vector<double> st(nno);
    if (gf->m) HIPCHK(hipMemcpy(gf->m, c->gs.m, sizeof(double) * nno, hipMemcpyDeviceToHost));
    double *const *srcs[4] = {c->gs.p, c->gs.f, c->gs.a, c->gs.v};
    double *dsts[4] = {gf->p, gf->oobf, gf->a, gf->v};
    for (int k = 0; k < 4; k++) {
        if (!dsts[k]) continue;
        for (int d = 0; d < D; d++) {
            HIPCHK(hipMemcpy(st.data(), srcs[k][d], sizeof(double) * nno, hipMemcpyDeviceToHost));
            for (long no = 0; no < nno; no++) dsts[k][d + D * no] = st[no];
        }
    }
    return EP_OK;
}

// ---------------------------------------------------------------------------
// step
// ---------------------------------------------------------------------------
template <typename L>
static void timed(ep_ctx *c, int kid, L &&launch) {
    if (c->profile && (c->prof_kid < 0 || c->prof_kid == kid)) {
        if (c->ev_used + 2 > c->ev_pool.size()) {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                hipEventCreate(&e);
                c->ev_pool.push_back(e);
            }
            c->ev_kernel.resize(c->ev_pool.size() / 2 + 64);
        }
        hipEventRecord(c->ev_pool[c->ev_used], c->stream);
        launch();
        hipEventRecord(c->ev_pool[c->ev_used + 1], c->stream);
        c->ev_kernel[c->ev_used / 2] = kid;
        c->ev_used += 2;
    } else {
        launch();
    }
    c->k_launches[kid]++;
    c->stats.launches++;
}

static mlp::SurrNet surr_net(ep_ctx *c);

template <int D, int B, typename T, bool SURR>
static void launch_step(ep_ctx *c, const PState<T> &P) {
    const int TB = c->tb;
    const long n = c->nslot, nno = c->gs.nno;
    const dim3 gp((n + TB - 1) / TB), gg((nno + TB - 1) / TB);
    hipStream_t s = c->stream;
    Ctrl *ct = c->ctrl;
    const Geom g = c->geom;
    const GState G = c->gs;
    const Mat M = c->mat;
    timed(c, K_BEGIN, [&] { EP_LAUNCH(k_begin<D>, 1, 64, s, ct, g, M); });
    timed(c, K_ZERO, [&] { EP_LAUNCH(k_zero_grid<D>, gg, TB, s, ct, G); });
    if (M.tpic || M.apic) timed(c, K_P2G, [&] { EP_LAUNCH((k_p2g<D, B, T, true>), gp, TB, s, ct, g, P, G, M); });
    else timed(c, K_P2G, [&] { EP_LAUNCH((k_p2g<D, B, T, false>), gp, TB, s, ct, g, P, G, M); });
    timed(c, K_EULER, [&] { EP_LAUNCH(k_euler<D>, gg, TB, s, ct, G, M); });
    timed(c, K_G2P, [&] { EP_LAUNCH((k_g2p<D, B, T>), gp, TB, s, ct, g, P, G, M); });
    if (M.apic) timed(c, K_BIJ, [&] { EP_LAUNCH((k_bij<D, B, T>), gp, TB, s, ct, g, P, G); });
    if (M.nonlocal && M.plast != EP_PLAST_NONE) {
        timed(c, K_NONLOCAL, [&] {
            const NLocal nl = M.nl;
            const dim3 gc((nl.ncell + 1 + 255) / 256), gs((n + 255) / 256);
            EP_LAUNCH(k_nl_zero, gc, 256, s, ct, nl.count, nl.ncell + 1);
            EP_LAUNCH((k_nl_count<D, T>), gs, 256, s, ct, g, P, nl);
#ifdef EP_CPU_CHECK
            {
                // the checking build runs kernels synchronously on host threads
                int acc = 0;
                for (long e = 0; e <= nl.ncell; e++) { nl.start[e] = acc; acc += nl.count[e]; }
            }
#else
            hipcub::DeviceScan::ExclusiveSum(c->nl_tmp, c->nl_tmp_bytes, nl.count, nl.start, int(nl.ncell + 1), s);
#endif
            EP_LAUNCH(k_nl_zero, gc, 256, s, ct, nl.count, nl.ncell + 1);
            EP_LAUNCH((k_nl_fill<T>), gs, 256, s, ct, n, nl);
            EP_LAUNCH((k_nonlocal<D, T>), gs, 256, s, ct, g, P, nl);
        });
    }
    if (M.musl) timed(c, K_AUGV, [&] { EP_LAUNCH(k_augv<D>, gg, TB, s, ct, G); });
    if (M.locking) timed(c, K_DEFORM, [&] { EP_LAUNCH((k_deform<D, B, T, true>), gp, TB, s, ct, g, P, G, M); });
    else timed(c, K_DEFORM, [&] { EP_LAUNCH((k_deform<D, B, T, false, SURR>), gp, TB, s, ct, g, P, G, M); });
    if (M.locking) timed(c, K_FBAR, [&] { EP_LAUNCH((k_fbar<D, B, T, SURR>), gp, TB, s, ct, g, P, G, M); });
    if constexpr (SURR) {
        const mlp::SurrNet net = surr_net(c);
        // a fixed grid that strides over the queue; its length is read on the device
        const int nb = EP_BLOCK == 256 ? 512 : 4;
        timed(c, K_SURR, [&] { EP_LAUNCH(mlp::k_surr, nb, EP_BLOCK, s, ct, P, M, net); });
    }
    timed(c, K_END, [&] { EP_LAUNCH(k_end, 1, 4 * EP_BLOCK, s, ct); });
}

template <typename T>
static void dispatch_step_t(ep_ctx *c, const PState<T> &P) {
    const int D = c->dim, B = c->cfg.basis;
    if (D == 2) {
        if (B == EP_BSMPM) launch_step<2, 0, T, false>(c, P);
        else if (B == EP_GIMPM) launch_step<2, 1, T, false>(c, P);
        else launch_step<2, 2, T, false>(c, P);
    } else {
        if (B == EP_BSMPM) launch_step<3, 0, T, false>(c, P);
        else if (B == EP_GIMPM) launch_step<3, 1, T, false>(c, P);
        else launch_step<3, 2, T, false>(c, P);
    }
}

// Re-sort (k_sort_keys, k_permute): a stable radix sort of the slots by
// their brick-major cell key, then every particle field gathered through the
// sorted slot list into a spare buffer that takes the field's place.  The
// slot count stays; padding slots move to the end.
// Keys, values, slot index and hipCUB's temporary storage of the re-sort,
// allocated once per particle upload so that no re-sort inside a run
// allocates.
static int sort_buffers(ep_ctx *c) {
    const long n = c->nslot;
    auto alloc = [&](size_t bytes) -> void * {
        void *a = nullptr;
        if (hipMalloc(&a, bytes) != hipSuccess) return nullptr;
        c->sort_allocs.push_back(a);
        return a;
    };
    c->sort_key[0] = static_cast<unsigned *>(alloc(sizeof(unsigned) * n));
    c->sort_key[1] = static_cast<unsigned *>(alloc(sizeof(unsigned) * n));
    c->sort_val = static_cast<int *>(alloc(sizeof(int) * n));
    c->sort_idx = static_cast<int *>(alloc(sizeof(int) * n));
    if (!c->sort_key[0] || !c->sort_key[1] || !c->sort_val || !c->sort_idx) {
        c->sort_idx = nullptr;
        return fail(EP_E_HIP, "hipMalloc of the re-sort buffers failed");
    }
#ifndef EP_CPU_CHECK
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, c->sort_key[0], c->sort_key[1], c->sort_val,
                                              c->sort_idx, int(n), 0, 32, c->stream));
    c->sort_tmp = alloc(bytes);
    c->sort_tmp_bytes = bytes;
    if (!c->sort_tmp) {
        c->sort_idx = nullptr;
        return fail(EP_E_HIP, "hipMalloc of the re-sort buffers failed");
    }
#endif
    return EP_OK;
}

template <int D, typename T>
static int resort_t(ep_ctx *c, PState<T> &P) {
    const long n = c->nslot;
    hipStream_t s = c->stream;
    const int TB = 256;
    const dim3 gp((n + TB - 1) / TB);
    if (!c->sort_idx) {
        const int r = sort_buffers(c);
        if (r != EP_OK) return r;
    }
    EP_LAUNCH((k_sort_keys<D, T>), gp, TB, s, c->ctrl, c->geom, P, n, int(c->brick[0]), int(c->brick[1]),
              int(c->brick[2]), c->sort_key[0], c->sort_val);
#ifdef EP_CPU_CHECK
    {
        // the checking build runs kernels on host threads, synchronously
        std::vector<int> ord(static_cast<size_t>(n));
        std::iota(ord.begin(), ord.end(), 0);
        const unsigned *key = c->sort_key[0];
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
        std::memcpy(c->sort_idx, ord.data(), sizeof(int) * size_t(n));
    }
#else
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, c->sort_tmp_bytes, c->sort_key[0], c->sort_key[1],
                                              c->sort_val, c->sort_idx, int(n), 0, 32, s));
#endif
    // every field into sorted order, kPermK of one element type per launch
    // (the slot index read once per group); each spare then takes its
    // field's place
    std::vector<double **> fd;
    std::vector<float **> ff;
    std::vector<int **> fi;
    auto add = [&](auto *&f) {
        if (!f) return;
        using E = std::remove_pointer_t<std::remove_reference_t<decltype(f)>>;
        if constexpr (std::is_same<E, double>::value) fd.push_back(&f);
        else if constexpr (std::is_same<E, float>::value) ff.push_back(&f);
        else fi.push_back(&f);
    };
    constexpr int DD = D * D, NB = D == 2 ? 3 : 6;
    for (int k = 0; k < 2; k++)
        for (int i = 0; i < D; i++) add(P.x[k][i]);
    for (int i = 0; i < D; i++) add(P.v[i]);
    for (int i = 0; i < DD; i++) add(P.F[i]);
    for (int i = 0; i < NB; i++) add(P.b[i]);
    for (int i = 0; i < NB; i++) add(P.s[i]);
    add(P.epII);
    add(P.dlam);
    add(P.J);
    add(P.rho);
    add(P.Om0);
    add(P.c0);
    add(P.cr);
    add(P.phi);
    add(P.sphi);
    add(P.cphi);
    for (int i = 0; i < D; i++) { add(P.l[i]); add(P.l0[i]); }
    for (int i = 0; i < DD; i++) { add(P.gv[i]); add(P.Bm[i]); }
    add(P.epII2);
    add(c->perm_d);
    auto permute = [&](auto &fields, auto *const *spares) {
        using E = std::remove_pointer_t<std::remove_pointer_t<typename std::remove_reference_t<decltype(fields)>::value_type>>;
        for (size_t g0 = 0; g0 < fields.size(); g0 += kPermK) {
            PermArgs<E> a{};
            a.cnt = int(std::min<size_t>(kPermK, fields.size() - g0));
            for (int k = 0; k < a.cnt; k++) {
                a.src[k] = *fields[g0 + k];
                a.dst[k] = spares[k];
            }
            EP_LAUNCH((k_permute<E>), gp, TB, s, a, c->sort_idx, n);
            for (int k = 0; k < a.cnt; k++) std::swap(*fields[g0 + k], const_cast<E *&>(spares[k]));
        }
    };
    permute(fd, c->spare_d);
    permute(ff, c->spare_f);
    permute(fi, &c->spare_i);
    // the non-local pass reads each particle's original index through
    // nl.orig, which must follow perm_d into its new buffer
    c->mat.nl.orig = c->perm_d;
    c->perm_dirty = true;
    return EP_OK;
}

static int resort(ep_ctx *c) {
    int r = EP_OK;
    timed(c, K_SORT, [&] {
        if (c->dim == 2) r = c->fp32 ? resort_t<2, float>(c, c->psf) : resort_t<2, double>(c, c->ps);
        else r = c->fp32 ? resort_t<3, float>(c, c->psf) : resort_t<3, double>(c, c->ps);
    });
    return r;
}

static void dispatch_step(ep_ctx *c) {
    if (c->cfg.precision == EP_FP32_SURROGATE) {
        const int B = c->cfg.basis;
        if (B == EP_BSMPM) launch_step<3, 0, float, true>(c, c->psf);
        else if (B == EP_GIMPM) launch_step<3, 1, float, true>(c, c->psf);
        else launch_step<3, 2, float, true>(c, c->psf);
    } else if (c->fp32) {
        dispatch_step_t<float>(c, c->psf);
    } else {
        dispatch_step_t<double>(c, c->ps);
    }
}

extern "C" int ep_step(ep_ctx *c, int32_t nsteps, const ep_step_params *par, ep_step_result *res) {
    if (!c || !par) return fail(EP_E_ARG, "ep_step: null argument");
    if (!c->part_ready || !c->grid_ready) return fail(EP_E_STATE, "upload grid and particles first");
    if (nsteps < 0) return fail(EP_E_ARG, "nsteps must be >= 0");
    if (c->cfg.precision == EP_FP32_SURROGATE && !c->surr)
        return fail(EP_E_STATE, "precision EP_FP32_SURROGATE needs ep_attach_surrogate first");
    Ctrl h{};
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    const long it0 = h.it;
    h.T = par->T;
    h.G = par->G;
    h.tg = par->tg;
    h.ramp = par->ramp;
    h.with_plast = par->with_plast;
    h.dt_mode = par->dt_mode;
    h.dt_fixed = par->dt_fixed;
    h.record = par->record_dt;
    h.steps = 0;
    h.spill = 0;
    h.boxsum = 0;
    h.err = 0;
    h.j2cap = 0;
    h.s_acc = 0;
    h.s_fb = 0;
    h.s_maxr = 0;
    if (par->record_dt) {
        if (c->dt_cap < nsteps) {
            hipFree(c->dt_log);
            c->dt_log = nullptr;
            HIPCHK(hipMalloc(&c->dt_log, sizeof(double) * std::max<long>(nsteps, 1)));
            c->dt_cap = nsteps;
        }
        h.dt_log = c->dt_log;
        h.dt_cap = c->dt_cap;
    }
    HIPCHK(hipMemcpy(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice));
    c->ev_used = 0;
    c->spill_seen = 0;
    c->box_seen = 0;
    for (int k = 0; k < nsteps; k++) {
        const int every = c->resort_every;
        if (every != 0 && !c->in_ensemble && c->since_sort >= (every > 0 ? every : -every)) {
            bool go = every > 0;
            if (!go) {
                // adaptive: re-sort once the p2g node boxes have grown by
                // kBoxGrowth since the last sort (the particles left their
                // blocks' bricks), or once more than kSpillFrac of the p2g
                // blocks of the window fell back to global atomics
                int spill = 0;
                unsigned long long box = 0;
                HIPCHK(hipMemcpyAsync(&spill, &c->ctrl->spill, sizeof(int), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipMemcpyAsync(&box, &c->ctrl->boxsum, sizeof(box), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
                const long nblk = (c->nslot + c->tb - 1) / c->tb;
                const double win = double(nblk) * double(c->since_sort);
                const double rate = double(spill - c->spill_seen) / win;
                const double brate = double(box - c->box_seen) / win;
                c->spill_seen = spill;
                c->box_seen = box;
                if (c->sorted_fresh) {
                    // what a fresh sort leaves is the floor a re-sort cannot lower
                    c->spill_floor = rate;
                    c->box_floor = brate;
                    c->sorted_fresh = false;
                } else {
                    go = (rate > kSpillFrac && rate > 2.0 * c->spill_floor) || brate > kBoxGrowth * c->box_floor;
                }
            }
            if (go) {
                const int r = resort(c);
                if (r != EP_OK) return r;
                c->resorts++;
                c->sorted_fresh = true;
            }
            c->since_sort = 0;
        }
        dispatch_step(c);
        c->since_sort++;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->profile) {
        for (size_t e = 0; e < c->ev_used; e += 2) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->ev_pool[e], c->ev_pool[e + 1]);
            c->k_ms[c->ev_kernel[e / 2]] += ms;
        }
    }
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    c->stats.steps += h.steps;
    c->stats.j2_maxit += h.j2cap;
    c->stats.predictions += int64_t(h.s_acc + h.s_fb);
    c->stats.fallbacks += int64_t(h.s_fb);
    {
        float mr;
        std::memcpy(&mr, &h.s_maxr, 4);
        if (double(mr) > c->stats.max_residual) c->stats.max_residual = double(mr);
    }
    if (res) {
        res->t = h.t;
        res->last_dt = h.dt;
        res->steps = h.steps;
        res->it = h.it;
    }
    (void)it0;
    if (h.err & ERR_DOMAIN) return fail(EP_E_DOMAIN, "Invalid position: a particle reached a node outside xB");
    if (h.err & ERR_NONFINITE) return fail(EP_E_NONFINITE, "non-finite particle state");
    return EP_OK;
}

extern "C" int ep_get_dts(ep_ctx *c, double *dts, int64_t cap, int64_t *n) {
    if (!c || !dts || !n) return fail(EP_E_ARG, "ep_get_dts: null argument");
    Ctrl h{};
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    const long k = std::min<long>(std::min<long>(h.steps, cap), c->dt_cap);
    if (k > 0) HIPCHK(hipMemcpy(dts, c->dt_log, sizeof(double) * k, hipMemcpyDeviceToHost));
    *n = k;
    return EP_OK;
}

extern "C" int ep_profile(ep_ctx *c, int32_t enable) {
    if (!c) return fail(EP_E_ARG, "null context");
    c->profile = enable != 0;
    c->prof_kid = -1;
    for (int k = 0; k < K_COUNT; k++) { c->k_ms[k] = 0.0; c->k_launches[k] = 0; }
    return EP_OK;
}

extern "C" int ep_profile_kernel(ep_ctx *c, const char *name) {
    if (!c || !name) return fail(EP_E_ARG, "null argument");
    int kid = -1;
    for (int k = 0; k < K_COUNT; k++)
        if (std::strcmp(name, kKernelNames[k]) == 0) kid = k;
    if (kid < 0) return fail(EP_E_ARG, std::string("ep_profile_kernel: no kernel named ") + name);
    c->profile = true;
    c->prof_kid = kid;
    for (int k = 0; k < K_COUNT; k++) { c->k_ms[k] = 0.0; c->k_launches[k] = 0; }
    return EP_OK;
}

extern "C" int ep_kernel_times(ep_ctx *c, int32_t max_kernels, char *names, int32_t name_len,
                               double *total_ms, int64_t *launches, int32_t *count) {
    if (!c || !count) return fail(EP_E_ARG, "null argument");
    int k = 0;
    for (int i = 0; i < K_COUNT && k < max_kernels; i++) {
        if (c->k_launches[i] == 0 || (c->prof_kid >= 0 && i != c->prof_kid)) continue;
        if (names && name_len > 0) {
            std::strncpy(names + size_t(k) * name_len, kKernelNames[i], name_len - 1);
            names[size_t(k) * name_len + name_len - 1] = 0;
        }
        if (total_ms) total_ms[k] = c->k_ms[i];
        if (launches) launches[k] = c->k_launches[i];
        k++;
    }
    *count = k;
    return EP_OK;
}

extern "C" int ep_get_stats(ep_ctx *c, ep_stats *st) {
    if (!c || !st) return fail(EP_E_ARG, "null argument");
    *st = c->stats;
    st->resorts = c->resorts;
    return EP_OK;
}

// ---------------------------------------------------------------------------
// learned return mapping: network, training samples, training
// ---------------------------------------------------------------------------
struct ep_surrogate {
    hipStream_t stream = nullptr;
    float *prm = nullptr, *am = nullptr, *av = nullptr, *grad = nullptr;
    unsigned short *img_f = nullptr, *img_b = nullptr;
    long adam_t = 0;
    long ns = 0;                                   // device training set
    float *X = nullptr, *Y = nullptr;
    double *raw = nullptr;
    long bcap = 0;                                 // training scratch for batches up to bcap
    int nsplit = 0;
    unsigned short *act = nullptr, *del = nullptr;
    float *lossp = nullptr, *part = nullptr;
    float *loss = nullptr;
    long loss_cap = 0;
};

static mlp::SurrNet surr_net(ep_ctx *c) { return mlp::SurrNet{c->surr->img_f, c->surr->prm, c->surr_tol}; }

namespace {
constexpr int kMlpBlock = EP_BLOCK;   // kernels with wave intrinsics: one wave per block in the CPU check

void surr_free_train(ep_surrogate *s) {
    hipFree(s->act); hipFree(s->del); hipFree(s->lossp); hipFree(s->part);
    s->act = s->del = nullptr;
    s->lossp = s->part = nullptr;
    s->bcap = 0;
}

// number of batch splits of the weight-gradient product: a power of two up to
// 64 that leaves every split a multiple of 32 samples
int surr_splits(long batch) {
    int S = 64;
    while (S > 1 && (batch % (32L * S) != 0 || batch / S < 256)) S >>= 1;
    return S;
}

int surr_ensure(ep_surrogate *s, long batch) {
    const int S = surr_splits(batch);
    if (batch <= s->bcap && S <= s->nsplit) return EP_OK;
    surr_free_train(s);
    HIPCHK(hipMalloc(&s->act, sizeof(unsigned short) * mlp::ACT_ROWS * batch));
    HIPCHK(hipMalloc(&s->del, sizeof(unsigned short) * mlp::DEL_ROWS * batch));
    HIPCHK(hipMalloc(&s->lossp, sizeof(float) * (batch / 32)));
    HIPCHK(hipMalloc(&s->part, sizeof(float) * size_t(S) * mlp::NPARAM));
    s->bcap = batch;
    s->nsplit = S;
    return EP_OK;
}

// forward + backward + weight gradient of one batch (and an Adam step when
// a.update); the batch loss goes to *loss_dev
void surr_batch(ep_surrogate *s, const float *X, const float *Y, long first, long batch, mlp::AdamArgs a,
                float *loss_dev) {
    const int S = surr_splits(batch);
    mlp::TrainBuf tb{X, Y, first, batch, s->act, s->del, s->lossp};
    const long ntile = batch / 32;
    const long per = kMlpBlock / 64;
    const int gb = int(std::min<long>((ntile + per - 1) / per, 1024));
    EP_LAUNCH(mlp::k_mlp_fb, gb, kMlpBlock, s->stream, s->img_f, s->img_b, s->prm, tb);
    EP_LAUNCH(mlp::k_mlp_dw, mlp::DW_TILES * S, 64, s->stream, s->act, s->del, batch, S, s->part);
    EP_LAUNCH(mlp::k_adam, (mlp::NPARAM + 255) / 256, 256, s->stream, s->part, S, s->grad, s->prm, s->am, s->av,
              s->img_f, s->img_b, a);
    EP_LAUNCH(mlp::k_loss_sum, 1, 256, s->stream, s->lossp, ntile, batch, loss_dev);
}
}  // namespace

extern "C" int ep_surrogate_create(uint64_t seed, ep_surrogate **out) {
    if (!out) return fail(EP_E_ARG, "ep_surrogate_create: null argument");
    *out = nullptr;
    ep_surrogate *s = new ep_surrogate();
    auto bad = [&](const char *m) { ep_surrogate_destroy(s); return fail(EP_E_HIP, m); };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bad("hipStreamCreate failed");
    const size_t pb = sizeof(float) * mlp::NPARAM;
    if (hipMalloc(&s->prm, pb) != hipSuccess || hipMalloc(&s->am, pb) != hipSuccess ||
        hipMalloc(&s->av, pb) != hipSuccess || hipMalloc(&s->grad, pb) != hipSuccess ||
        hipMalloc(&s->img_f, sizeof(unsigned short) * mlp::IF_N) != hipSuccess ||
        hipMalloc(&s->img_b, sizeof(unsigned short) * mlp::IB_N) != hipSuccess ||
        hipMalloc(&s->loss, sizeof(float)) != hipSuccess)
        return bad("hipMalloc of the network failed");
    s->loss_cap = 1;
    // the padding of the swizzled images stays zero
    if (hipMemset(s->img_f, 0, sizeof(unsigned short) * mlp::IF_N) != hipSuccess ||
        hipMemset(s->img_b, 0, sizeof(unsigned short) * mlp::IB_N) != hipSuccess)
        return bad("hipMemset failed");
    // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases of each layer
    std::vector<float> w(mlp::NPARAM);
    for (int k = 0; k < mlp::NPARAM; k++) {
        const int fan = k < mlp::OW2 ? mlp::NI : mlp::NH;
        const double u = mlp::u01(seed, (unsigned long long)k, 0);
        w[k] = float((2.0 * u - 1.0) / std::sqrt(double(fan)));
    }
    *out = s;
    return ep_surrogate_set_params(s, w.data());
}

extern "C" void ep_surrogate_destroy(ep_surrogate *s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    surr_free_train(s);
    hipFree(s->prm); hipFree(s->am); hipFree(s->av); hipFree(s->grad);
    hipFree(s->img_f); hipFree(s->img_b);
    hipFree(s->X); hipFree(s->Y); hipFree(s->raw); hipFree(s->loss);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

extern "C" int ep_surrogate_set_params(ep_surrogate *s, const float *params) {
    if (!s || !params) return fail(EP_E_ARG, "ep_surrogate_set_params: null argument");
    for (int k = 0; k < mlp::NPARAM; k++)
        if (!std::isfinite(params[k])) return fail(EP_E_ARG, "non-finite network parameter");
    const size_t pb = sizeof(float) * mlp::NPARAM;
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(s->prm, params, pb, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->am, 0, pb));
    HIPCHK(hipMemset(s->av, 0, pb));
    s->adam_t = 0;
    EP_LAUNCH(mlp::k_mlp_pack, (mlp::NPARAM + 255) / 256, 256, s->stream, s->prm, s->img_f, s->img_b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return EP_OK;
}

extern "C" int ep_surrogate_get_params(ep_surrogate *s, float *params) {
    if (!s || !params) return fail(EP_E_ARG, "ep_surrogate_get_params: null argument");
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(params, s->prm, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost));
    return EP_OK;
}

extern "C" int ep_surrogate_sample(ep_surrogate *s, int64_t n, uint64_t seed) {
    if (!s) return fail(EP_E_ARG, "null surrogate");
    if (n <= 0) return fail(EP_E_ARG, "n must be positive");
    HIPCHK(hipStreamSynchronize(s->stream));
    hipFree(s->X); hipFree(s->Y); hipFree(s->raw);
    s->X = s->Y = nullptr;
    s->raw = nullptr;
    s->ns = 0;
    HIPCHK(hipMalloc(&s->X, sizeof(float) * 8 * n));
    HIPCHK(hipMalloc(&s->Y, sizeof(float) * 4 * n));
    HIPCHK(hipMalloc(&s->raw, sizeof(double) * 6 * n));
    s->ns = n;
    EP_LAUNCH(mlp::k_mlp_samples, (n + 255) / 256, 256, s->stream, long(n), (unsigned long long)seed, s->X, s->Y,
              s->raw);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return EP_OK;
}

extern "C" int ep_surrogate_get_samples(ep_surrogate *s, int64_t first, int64_t n, float *x, float *y, double *raw) {
    if (!s) return fail(EP_E_ARG, "null surrogate");
    if (first < 0 || n < 0 || first + n > s->ns) return fail(EP_E_ARG, "sample range outside the training set");
    HIPCHK(hipStreamSynchronize(s->stream));
    if (x) HIPCHK(hipMemcpy(x, s->X + 8 * first, sizeof(float) * 8 * n, hipMemcpyDeviceToHost));
    if (y) HIPCHK(hipMemcpy(y, s->Y + 4 * first, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
    if (raw) HIPCHK(hipMemcpy(raw, s->raw + 6 * first, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
    return EP_OK;
}

extern "C" int ep_surrogate_train(ep_surrogate *s, int64_t steps, int64_t batch, double lr0, double lr1,
                                  float *loss) {
    if (!s) return fail(EP_E_ARG, "null surrogate");
    if (steps < 0 || batch <= 0 || batch % 32 != 0) return fail(EP_E_ARG, "batch must be a positive multiple of 32");
    if (s->ns < batch) return fail(EP_E_STATE, "ep_surrogate_sample a training set of at least one batch first");
    int r;
    if ((r = surr_ensure(s, batch)) != EP_OK) return r;
    if (s->loss_cap < steps) {
        hipFree(s->loss);
        s->loss = nullptr;
        HIPCHK(hipMalloc(&s->loss, sizeof(float) * std::max<int64_t>(steps, 1)));
        s->loss_cap = std::max<int64_t>(steps, 1);
    }
    const long nb = s->ns / batch;
    const double b1 = 0.9, b2 = 0.999;
    for (int64_t i = 0; i < steps; i++) {
        s->adam_t++;
        const double lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(3.14159265358979323846 * double(i) / double(steps)));
        mlp::AdamArgs a{float(lr), float(b1), float(b2), 1e-8f, float(1.0 / (1.0 - std::pow(b1, double(s->adam_t)))),
                        float(1.0 / (1.0 - std::pow(b2, double(s->adam_t)))), 1};
        surr_batch(s, s->X, s->Y, (i % nb) * batch, batch, a, s->loss + i);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    if (loss && steps > 0) HIPCHK(hipMemcpy(loss, s->loss, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return EP_OK;
}

extern "C" int ep_surrogate_grad(ep_surrogate *s, const float *x, const float *y, int64_t n, float *grad,
                                 float *loss) {
    if (!s || !x || !y || !grad) return fail(EP_E_ARG, "ep_surrogate_grad: null argument");
    if (n <= 0 || n % 32 != 0) return fail(EP_E_ARG, "n must be a positive multiple of 32");
    int r;
    if ((r = surr_ensure(s, n)) != EP_OK) return r;
    float *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc(&dx, sizeof(float) * 8 * n));
    if (hipMalloc(&dy, sizeof(float) * 4 * n) != hipSuccess) { hipFree(dx); return fail(EP_E_HIP, "hipMalloc failed"); }
    int rc = EP_OK;
    if (hipMemcpy(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy, y, sizeof(float) * 4 * n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(EP_E_HIP, "hipMemcpy failed");
    } else {
        mlp::AdamArgs a{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0};
        surr_batch(s, dx, dy, 0, n, a, s->loss);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess ||
            hipMemcpy(grad, s->grad, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost) != hipSuccess ||
            (loss && hipMemcpy(loss, s->loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
            rc = fail(EP_E_HIP, "gradient evaluation failed");
    }
    hipFree(dx);
    hipFree(dy);
    return rc;
}

extern "C" int ep_surrogate_predict(ep_surrogate *s, const float *x, int64_t n, float *y) {
    if (!s || !x || !y) return fail(EP_E_ARG, "ep_surrogate_predict: null argument");
    if (n <= 0) return fail(EP_E_ARG, "n must be positive");
    const long np = (n + 31) / 32 * 32;
    float *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc(&dx, sizeof(float) * 8 * np));
    if (hipMalloc(&dy, sizeof(float) * 4 * np) != hipSuccess) { hipFree(dx); return fail(EP_E_HIP, "hipMalloc failed"); }
    int rc = EP_OK;
    if (hipMemset(dx, 0, sizeof(float) * 8 * np) != hipSuccess ||
        hipMemcpy(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(EP_E_HIP, "hipMemcpy failed");
    } else {
        const long per = kMlpBlock / 64;
        const int gb = int(std::min<long>((np / 32 + per - 1) / per, 1024));
        EP_LAUNCH(mlp::k_mlp_predict, gb, kMlpBlock, s->stream, s->img_f, s->prm, dx, np, dy);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess ||
            hipMemcpy(y, dy, sizeof(float) * 4 * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(EP_E_HIP, "prediction failed");
    }
    hipFree(dx);
    hipFree(dy);
    return rc;
}

extern "C" int ep_surrogate_harvest(ep_surrogate *s, ep_ctx *c, int32_t append, int64_t *count) {
    if (!s || !c) return fail(EP_E_ARG, "ep_surrogate_harvest: null argument");
    if (count) *count = 0;
    if (c->cfg.precision != EP_FP32_SURROGATE || c->dim != 3 || !c->mat.q.idx)
        return fail(EP_E_CONFIG, "ep_surrogate_harvest: needs a 3D EP_FP32_SURROGATE context");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    int q = 0;
    HIPCHK(hipMemcpy(&q, &c->ctrl->qcount, sizeof(int), hipMemcpyDeviceToHost));
    const long keep = append ? s->ns : 0;
    const long cap = keep + std::max(q, 0);
    float *X = nullptr, *Y = nullptr;
    double *raw = nullptr;
    int *cnt = nullptr;
    auto release = [&] { hipFree(X); hipFree(Y); hipFree(raw); hipFree(cnt); };
    if (hipMalloc(&X, sizeof(float) * 8 * std::max(cap, 1L)) != hipSuccess ||
        hipMalloc(&Y, sizeof(float) * 4 * std::max(cap, 1L)) != hipSuccess ||
        hipMalloc(&raw, sizeof(double) * 6 * std::max(cap, 1L)) != hipSuccess ||
        hipMalloc(&cnt, sizeof(int)) != hipSuccess) {
        release();
        return fail(EP_E_HIP, "hipMalloc failed");
    }
    int n = 0;
    bool ok = hipMemset(cnt, 0, sizeof(int)) == hipSuccess;
    if (ok && keep > 0)
        ok = hipMemcpy(X, s->X, sizeof(float) * 8 * keep, hipMemcpyDeviceToDevice) == hipSuccess &&
             hipMemcpy(Y, s->Y, sizeof(float) * 4 * keep, hipMemcpyDeviceToDevice) == hipSuccess &&
             hipMemcpy(raw, s->raw, sizeof(double) * 6 * keep, hipMemcpyDeviceToDevice) == hipSuccess;
    if (ok && q > 0) {
        EP_LAUNCH(mlp::k_mlp_harvest, (q + 255) / 256, 256, c->stream, c->ctrl, c->psf, c->mat, keep, cap, X, Y, raw,
                  cnt);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(c->stream) == hipSuccess &&
             hipMemcpy(&n, cnt, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (!ok) {
        release();
        return fail(EP_E_HIP, "ep_surrogate_harvest: device copy or kernel failed");
    }
    hipFree(cnt);
    hipFree(s->X); hipFree(s->Y); hipFree(s->raw);
    s->X = X;
    s->Y = Y;
    s->raw = raw;
    s->ns = keep + std::min<long>(n, cap - keep);
    if (count) *count = n;
    return EP_OK;
}

extern "C" int ep_attach_surrogate(ep_ctx *c, ep_surrogate *s, double tol) {
    if (!c) return fail(EP_E_ARG, "null context");
    if (c->cfg.precision != EP_FP32_SURROGATE) return fail(EP_E_CONFIG, "the context's precision is not EP_FP32_SURROGATE");
    if (s) HIPCHK(hipStreamSynchronize(s->stream));
    c->surr = s;
    const double t = tol > 0.0 ? tol : (c->cfg.surrogate_tol > 0.0 ? c->cfg.surrogate_tol : 1e-3);
    c->surr_tol = float(t);
    return EP_OK;
}

// ---------------------------------------------------------------------------
// ensemble: many contexts advanced by one captured graph per step
// ---------------------------------------------------------------------------
struct ep_ensemble {
    std::vector<ep_ctx *> ctx;
    int n = 0, dim = 0, basis = 0;
    bool fp32 = false;
    hipStream_t stream = nullptr;
    void *inst = nullptr;                  // Inst<T>[n]
    int *pm_inst = nullptr, *pm_first = nullptr, *gm_inst = nullptr, *gm_first = nullptr;
    long npblk = 0, ngblk = 0;
    double *collect = nullptr;             // [n][5]
    hipGraphExec_t exec = nullptr;
    int tb = EP_BLOCK;
    ep_stats stats{};
};

namespace {

void ens_free_tables(ep_ensemble *e) {
    if (e->exec) hipGraphExecDestroy(e->exec);
    e->exec = nullptr;
    hipFree(e->inst); hipFree(e->pm_inst); hipFree(e->pm_first); hipFree(e->gm_inst); hipFree(e->gm_first);
    e->inst = nullptr;
    e->pm_inst = e->pm_first = e->gm_inst = e->gm_first = nullptr;
}

// Give a context tb threads per block: its per-block vmax partials and the
// block count in its control block follow.
int ctx_set_block(ep_ctx *c, int tb) {
    if (c->tb == tb && c->vmax_part) return EP_OK;
    c->tb = tb;
    Ctrl h{};
    HIPCHK(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    hipFree(c->vmax_part);
    c->vmax_part = nullptr;
    h.nblk = int((c->nslot + tb - 1) / tb);
    HIPCHK(hipMalloc(&c->vmax_part, sizeof(unsigned long long) * 3 * size_t(h.nblk)));
    HIPCHK(hipMemset(c->vmax_part, 0, sizeof(unsigned long long) * 3 * size_t(h.nblk)));
    h.vmax_part = c->vmax_part;
    HIPCHK(hipMemcpy(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice));
    return EP_OK;
}

template <typename T>
int ens_tables(ep_ensemble *e) {
    ens_free_tables(e);
    const int n = e->n, tb = e->tb;
    std::vector<Inst<T>> I(n);
    std::vector<int> pf(n + 1), gf(n + 1), pi, gi;
    long pb = 0, gb = 0;
    for (int i = 0; i < n; i++) {
        ep_ctx *c = e->ctx[i];
        int r;
        if ((r = ctx_set_block(c, tb)) != EP_OK) return r;
        if constexpr (std::is_same<T, float>::value) I[i] = Inst<T>{c->ctrl, c->geom, c->psf, c->gs, c->mat};
        else I[i] = Inst<T>{c->ctrl, c->geom, c->ps, c->gs, c->mat};
        const long np = (c->nslot + tb - 1) / tb, ng = (c->gs.nno + tb - 1) / tb;
        pf[i] = int(pb);
        gf[i] = int(gb);
        for (long k = 0; k < np; k++) pi.push_back(i);
        for (long k = 0; k < ng; k++) gi.push_back(i);
        pb += np;
        gb += ng;
        if (pb > 0x7fffffffL || gb > 0x7fffffffL) return fail(EP_E_ARG, "ensemble exceeds 2^31 blocks");
    }
    pf[n] = int(pb);
    gf[n] = int(gb);
    e->npblk = pb;
    e->ngblk = gb;
    HIPCHK(hipMalloc(&e->inst, sizeof(Inst<T>) * n));
    HIPCHK(hipMalloc(&e->pm_inst, sizeof(int) * std::max<long>(pb, 1)));
    HIPCHK(hipMalloc(&e->gm_inst, sizeof(int) * std::max<long>(gb, 1)));
    HIPCHK(hipMalloc(&e->pm_first, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&e->gm_first, sizeof(int) * (n + 1)));
    HIPCHK(hipMemcpy(e->inst, I.data(), sizeof(Inst<T>) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->pm_inst, pi.data(), sizeof(int) * pb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->gm_inst, gi.data(), sizeof(int) * gb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->pm_first, pf.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->gm_first, gf.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    return EP_OK;
}

// One step of every instance, captured once into a graph.  No memset or
// copy nodes: the grid is cleared by kb_zero_grid.
template <int D, int B, typename T>
int ens_capture(ep_ensemble *e) {
    int r;
    if ((r = ens_tables<T>(e)) != EP_OK) return r;
    const Inst<T> *I = static_cast<const Inst<T> *>(e->inst);
    const BlockMap pm{e->pm_inst, e->pm_first}, gm{e->gm_inst, e->gm_first};
    const Mat &M = e->ctx[0]->mat;
    const int n = e->n, tb = e->tb;
    hipStream_t s = e->stream;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    EP_LAUNCH((kb_begin<D, T>), (n + 63) / 64, 64, s, I, n);
    EP_LAUNCH((kb_zero_grid<D, T>), e->ngblk, tb, s, I, gm);
    if (M.tpic || M.apic) EP_LAUNCH((kb_p2g<D, B, T, true>), e->npblk, tb, s, I, pm);
    else EP_LAUNCH((kb_p2g<D, B, T, false>), e->npblk, tb, s, I, pm);
    EP_LAUNCH((kb_euler<D, T>), e->ngblk, tb, s, I, gm);
    EP_LAUNCH((kb_g2p<D, B, T>), e->npblk, tb, s, I, pm);
    if (M.apic) EP_LAUNCH((kb_bij<D, B, T>), e->npblk, tb, s, I, pm);
    if (M.musl) EP_LAUNCH((kb_augv<D, T>), e->ngblk, tb, s, I, gm);
    if (M.locking) {
        EP_LAUNCH((kb_deform<D, B, T, true>), e->npblk, tb, s, I, pm);
        EP_LAUNCH((kb_fbar<D, B, T>), e->npblk, tb, s, I, pm);
    } else {
        EP_LAUNCH((kb_deform<D, B, T, false>), e->npblk, tb, s, I, pm);
    }
    EP_LAUNCH(kb_end<T>, n, 4 * EP_BLOCK, s, I);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(s, &g));
    const hipError_t ie = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) { e->exec = nullptr; return fail(EP_E_HIP, "hipGraphInstantiate failed"); }
    return EP_OK;
}

int ens_build(ep_ensemble *e) {
    const bool f = e->fp32;
    if (e->dim == 2) return f ? ens_capture<2, 0, float>(e) : ens_capture<2, 0, double>(e);
    return f ? ens_capture<3, 0, float>(e) : ens_capture<3, 0, double>(e);
}

template <typename T>
void ens_params(ep_ensemble *e, const StepPar &q) {
    EP_LAUNCH(kb_params<T>, (e->n + 63) / 64, 64, e->stream, static_cast<const Inst<T> *>(e->inst), e->n, q);
}
template <typename T>
void ens_collect(ep_ensemble *e) {
    EP_LAUNCH(kb_collect<T>, (e->n + 63) / 64, 64, e->stream, static_cast<const Inst<T> *>(e->inst), e->n,
              e->collect);
}

// nsteps graph replays; per-instance results into res (may be NULL)
int ens_run(ep_ensemble *e, int32_t nsteps, const ep_step_params *par, ep_step_result *res, float *ms) {
    const StepPar q{par->T, par->G, par->tg, par->dt_fixed, par->ramp, par->with_plast, par->dt_mode};
    if (e->fp32) ens_params<float>(e, q);
    else ens_params<double>(e, q);
    hipEvent_t a = nullptr, b = nullptr;
    if (ms) {
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, e->stream));
    }
    for (int k = 0; k < nsteps; k++) HIPCHK(hipGraphLaunch(e->exec, e->stream));
    if (ms) {
        HIPCHK(hipEventRecord(b, e->stream));
        HIPCHK(hipEventSynchronize(b));
        HIPCHK(hipEventElapsedTime(ms, a, b));
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
    if (e->fp32) ens_collect<float>(e);
    else ens_collect<double>(e);
    HIPCHK(hipGetLastError());
    std::vector<double> out(size_t(5) * e->n);
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(out.data(), e->collect, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
    int code = EP_OK;
    for (int i = 0; i < e->n; i++) {
        const double *o = &out[size_t(5) * i];
        const long st = long(o[2]);
        e->ctx[i]->stats.steps += st;
        e->stats.steps += st;
        e->stats.j2_maxit += long(o[4]);
        if (res) {
            res[i].t = o[0];
            res[i].last_dt = o[1];
            res[i].steps = st;
            res[i].it = 0;
        }
        const unsigned err = unsigned(o[3]);
        if (code == EP_OK && (err & ERR_DOMAIN))
            code = fail(EP_E_DOMAIN, "Invalid position in ensemble instance " + std::to_string(i));
        else if (code == EP_OK && (err & ERR_NONFINITE))
            code = fail(EP_E_NONFINITE, "non-finite state in ensemble instance " + std::to_string(i));
    }
    e->stats.launches += long(nsteps);
    return code;
}

struct HostState {
    std::vector<double> f[25];
    ep_particles_view v{};
    double t = 0.0;
};

// download an instance into host buffers laid out as the reference's arrays
int ens_save(ep_ctx *c, HostState &h) {
    const long n = c->nmp;
    const int D = c->dim, DD = D * D, NS = c->nstr;
    double **slot[] = {&h.v.x, &h.v.v, &h.v.l0, &h.v.l, &h.v.F, &h.v.b, &h.v.gradv, &h.v.eps, &h.v.sig, &h.v.tau,
                       &h.v.epII, &h.v.rho0, &h.v.rho, &h.v.Om0, &h.v.Om, &h.v.n0, &h.v.n, &h.v.J, &h.v.dJ,
                       &h.v.c0, &h.v.cr, &h.v.phi, &h.v.dlam, &h.v.vmax, &h.v.Bij};
    const long len[] = {D * n, D * n, D * n, D * n, DD * n, DD * n, DD * n, DD * n, NS * n, NS * n, 2 * n,
                        n, n, n, n, n, n, n, n, n, n, n, n, 3, DD * n};
    for (int k = 0; k < 25; k++) {
        h.f[k].assign(size_t(len[k]), 0.0);
        *slot[k] = h.f[k].data();
    }
    int r;
    if ((r = ep_download_particles(c, &h.v)) != EP_OK) return r;
    Ctrl hc{};
    HIPCHK(hipMemcpy(&hc, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    h.t = hc.t;
    return EP_OK;
}

int ens_restore(ep_ctx *c, HostState &h) {
    int r;
    if ((r = ep_upload_particles(c, &h.v)) != EP_OK) return r;
    return ep_set_time(c, h.t);
}

}  // namespace

// Launch shape of one context: threads per particle / node block and cells
// per particle-sort brick; the particles are re-sorted into the new brick
// order at once.
extern "C" int ep_set_layout(ep_ctx *c, int32_t block, const int32_t *brick) {
    if (!c || !brick) return fail(EP_E_ARG, "ep_set_layout: null argument");
    if (c->in_ensemble) return fail(EP_E_STATE, "ep_set_layout: the context belongs to an ensemble");
    if (block != 64 && block != 128 && block != 256 && block != 512)
        return fail(EP_E_ARG, "ep_set_layout: block must be 64, 128, 256 or 512");
    for (int i = 0; i < c->dim; i++)
        if (brick[i] < 1 || brick[i] > 64) return fail(EP_E_ARG, "ep_set_layout: brick sides must be 1..64");
    HIPCHK(hipStreamSynchronize(c->stream));
    const int r = ctx_set_block(c, block);
    if (r != EP_OK) return r;
    for (int i = 0; i < 3; i++) c->brick[i] = i < c->dim ? brick[i] : 1;
    if (c->part_ready) {
        const int q = resort(c);
        if (q != EP_OK) return q;
        HIPCHK(hipStreamSynchronize(c->stream));
        c->since_sort = 0;
        c->sorted_fresh = true;
    }
    return EP_OK;
}

extern "C" int ep_ensemble_create(ep_ctx *const *ctxs, int32_t n, ep_ensemble **out) {
    if (!ctxs || !out || n <= 0) return fail(EP_E_ARG, "ep_ensemble_create: bad argument");
    *out = nullptr;
    const ep_config &c0 = ctxs[0]->cfg;
    for (int i = 0; i < n; i++) {
        const ep_ctx *c = ctxs[i];
        if (!c) return fail(EP_E_ARG, "null context in the ensemble");
        if (!c->grid_ready || !c->part_ready) return fail(EP_E_STATE, "upload every instance before ep_ensemble_create");
        const ep_config &k = c->cfg;
        if (k.dim != c0.dim || k.basis != c0.basis || k.transfer != c0.transfer || k.musl != c0.musl ||
            k.deform != c0.deform || k.plast != c0.plast || k.locking != c0.locking || k.precision != c0.precision)
            return fail(EP_E_CONFIG, "ensemble instances must share dim, basis, scheme, plasticity and precision");
    }
    if (c0.basis != EP_BSMPM) return fail(EP_E_CONFIG, "ensembles run the cubic B-spline basis");
    if (c0.precision == EP_FP32_SURROGATE) return fail(EP_E_CONFIG, "ensembles run EP_FP64 or EP_FP32");
    for (int k = 0; k < n; k++)
        if (ctxs[k]->mat.nonlocal) return fail(EP_E_CONFIG, "ensembles run without non-local regularisation");
    ep_ensemble *e = new ep_ensemble();
    e->ctx.assign(ctxs, ctxs + n);
    // the ensemble's instance table holds the contexts' field pointers
    for (ep_ctx *c : e->ctx) c->in_ensemble = true;
    e->n = n;
    e->dim = c0.dim;
    e->basis = c0.basis;
    e->fp32 = c0.precision == EP_FP32;
    e->tb = ctxs[0]->tb;
    int r = EP_OK;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) r = fail(EP_E_HIP, "hipStreamCreate failed");
    else if (hipMalloc(&e->collect, sizeof(double) * 5 * n) != hipSuccess) r = fail(EP_E_HIP, "hipMalloc failed");
    else r = ens_build(e);
    if (r != EP_OK) { ep_ensemble_destroy(e); return r; }
    *out = e;
    return EP_OK;
}

extern "C" void ep_ensemble_destroy(ep_ensemble *e) {
    if (!e) return;
    for (ep_ctx *c : e->ctx) c->in_ensemble = false;
    if (e->stream) hipStreamSynchronize(e->stream);
    ens_free_tables(e);
    hipFree(e->collect);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

extern "C" int ep_ensemble_step(ep_ensemble *e, int32_t nsteps, const ep_step_params *par, ep_step_result *res) {
    if (!e || !par) return fail(EP_E_ARG, "ep_ensemble_step: null argument");
    if (nsteps < 0) return fail(EP_E_ARG, "nsteps must be >= 0");
    if (par->record_dt) return fail(EP_E_ARG, "ensembles do not record dt");
    for (ep_ctx *c : e->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    return ens_run(e, nsteps, par, res, nullptr);
}

extern "C" int ep_ensemble_tune(ep_ensemble *e, int32_t steps, const ep_step_params *par, ep_tune_result *out) {
    if (!e || !par || !out) return fail(EP_E_ARG, "ep_ensemble_tune: null argument");
    if (steps < 1) return fail(EP_E_ARG, "steps must be >= 1");
    static const int blocks[] = {64, 128, 256};
    static const long b2[][3] = {{4, 4, 1}, {8, 8, 1}, {16, 16, 1}};
    static const long b3[][3] = {{2, 2, 2}, {4, 2, 4}, {8, 4, 8}};
    std::vector<HostState> saved(size_t(e->n));
    int r;
    for (int i = 0; i < e->n; i++)
        if ((r = ens_save(e->ctx[i], saved[i])) != EP_OK) return r;
    *out = ep_tune_result{};
    double best = 1e300;
    int bi = -1;
    int k = 0;
    for (int ib = 0; ib < 3; ib++) {
        if (blocks[ib] > EP_BLOCK) continue;
        for (int jb = 0; jb < 3; jb++) {
            const long *br = e->dim == 2 ? b2[jb] : b3[jb];
            e->tb = blocks[ib];
            for (int i = 0; i < e->n; i++) {
                ep_ctx *c = e->ctx[i];
                for (int a = 0; a < 3; a++) c->brick[a] = br[a];
                c->tb = e->tb;
                if ((r = ens_restore(c, saved[i])) != EP_OK) return r;
            }
            if ((r = ens_build(e)) != EP_OK) return r;
            float ms = 0.f;
            if ((r = ens_run(e, 3, par, nullptr, nullptr)) != EP_OK) return r;   // warm-up, graph upload
            if ((r = ens_run(e, steps, par, nullptr, &ms)) != EP_OK) return r;
            const double per = double(ms) / steps;
            if (k < 16) {
                out->tried_block[k] = e->tb;
                for (int a = 0; a < 3; a++) out->tried_brick[k][a] = int32_t(br[a]);
                out->tried_ms[k] = per;
            }
            if (per < best) { best = per; bi = k; }
            k++;
        }
    }
    out->ntried = k;
    out->block = out->tried_block[bi];
    for (int a = 0; a < 3; a++) out->brick[a] = out->tried_brick[bi][a];
    out->ms_per_step = best;
    // restore the saved state in the chosen layout
    e->tb = out->block;
    for (int i = 0; i < e->n; i++) {
        ep_ctx *c = e->ctx[i];
        for (int a = 0; a < 3; a++) c->brick[a] = out->brick[a];
        c->tb = e->tb;
        if ((r = ens_restore(c, saved[i])) != EP_OK) return r;
    }
    return ens_build(e);
}

extern "C" int ep_ensemble_get_stats(ep_ensemble *e, ep_stats *st) {
    if (!e || !st) return fail(EP_E_ARG, "null argument");
    *st = e->stats;
    return EP_OK;
}

// ---------------------------------------------------------------------------
// Particle lattice and slump geometry on the device: mpts_populate's lattice
// (mpts_populate.jl:26-56) filtered as get_slump does (get_slump.jl:18-95)
// or to a rectangular column, kept in lattice order.
// ---------------------------------------------------------------------------
namespace {
struct LatticeSpec {
    int dim, geometry;       // geometry: 0 wedge (get_slump), 1 column
    long n[3];               // points per axis (x, y, z in 3D; x, z in 2D)
    const double *ax[3];     // per-axis coordinates
    double par[5];           // wedge: zcut, wl, a, X, Z; column: xmax, zmax
};

// The lattice runs along the last axis fastest, then x, then y (the vec
// order of mpts_populate.jl:34-43).
__device__ __host__ inline void lattice_point(const LatticeSpec &L, long i, double *x) {
    const long nz = L.n[L.dim - 1];
    const long iz = i % nz, r = i / nz;
    if (L.dim == 2) {
        x[0] = L.ax[0][r];
        x[1] = L.ax[1][iz];
    } else {
        x[0] = L.ax[0][r % L.n[0]];
        x[1] = L.ax[1][r / L.n[0]];
        x[2] = L.ax[2][iz];
    }
}

__device__ __host__ inline bool lattice_keep(const LatticeSpec &L, const double *x) {
#pragma clang fp contract(off)
    const double xp = x[0], zp = x[L.dim - 1];
    if (L.geometry == 1) return xp <= L.par[0] && zp <= L.par[1];
    if (!(zp <= L.par[0])) return false;
    // Δx nx + Δz nz of the last line point, nx = a, nz = -1 (get_slump.jl:30-41, 59-71)
    const double dx = xp - L.par[3], dz = zp - L.par[4];
    const double s = dx * L.par[2] + dz * -1.0;
    return s > 0.0 || zp < L.par[1];
}

__global__ void k_lattice_flags(LatticeSpec L, long n, unsigned char *flag) {
    const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    double x[3];
    lattice_point(L, i, x);
    flag[i] = lattice_keep(L, x) ? 1 : 0;
}

__global__ void k_lattice_gather(LatticeSpec L, const int *sel, long m, double *out) {
    const long k = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (k >= m) return;
    double x[3];
    lattice_point(L, sel[k], x);
    for (int d = 0; d < L.dim; d++) out[d + L.dim * k] = x[d];
}
}  // namespace

extern "C" int ep_slump_points(int32_t dim, const double *const *axes, const int64_t *n_axis, int32_t geometry,
                               const double *par, int64_t cap, double *x_out, int64_t *nmp) {
    if (!axes || !n_axis || !par || !nmp || (dim != 2 && dim != 3) || (geometry != 0 && geometry != 1) || cap < 0)
        return fail(EP_E_ARG, "ep_slump_points: bad argument");
    LatticeSpec L{};
    L.dim = dim;
    L.geometry = geometry;
    long n = 1;
    for (int i = 0; i < dim; i++) {
        if (!axes[i] || n_axis[i] < 1) return fail(EP_E_ARG, "ep_slump_points: empty axis");
        L.n[i] = n_axis[i];
        n *= n_axis[i];
    }
    if (n > 0x7fffffffL) return fail(EP_E_ARG, "ep_slump_points: lattice exceeds 2^31 - 1 points");
    for (int k = 0; k < 5; k++) L.par[k] = par[k];
    std::vector<void *> mem;
    auto done = [&](int r) {
        for (void *a : mem) hipFree(a);
        return r;
    };
    auto alloc = [&](size_t bytes) -> void * {
        void *a = nullptr;
        if (hipMalloc(&a, bytes ? bytes : 1) != hipSuccess) return nullptr;
        mem.push_back(a);
        return a;
    };
    double *dax[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < dim; i++) {
        dax[i] = static_cast<double *>(alloc(sizeof(double) * L.n[i]));
        if (!dax[i] || hipMemcpy(dax[i], axes[i], sizeof(double) * L.n[i], hipMemcpyHostToDevice) != hipSuccess)
            return done(fail(EP_E_HIP, "ep_slump_points: axis upload failed"));
        L.ax[i] = dax[i];
    }
    auto *flag = static_cast<unsigned char *>(alloc(size_t(n)));
    auto *sel = static_cast<int *>(alloc(sizeof(int) * size_t(n)));
    auto *cnt = static_cast<int *>(alloc(sizeof(int)));
    if (!flag || !sel || !cnt) return done(fail(EP_E_HIP, "ep_slump_points: hipMalloc failed"));
    EP_LAUNCH(k_lattice_flags, (n + 255) / 256, 256, 0, L, n, flag);
    int m = 0;
#ifdef EP_CPU_CHECK
    for (long i = 0; i < n; i++)
        if (flag[i]) sel[m++] = int(i);
#else
    {
        size_t bytes = 0;
        hipcub::CountingInputIterator<int> it(0);
        if (hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flag, sel, cnt, int(n)) != hipSuccess)
            return done(fail(EP_E_HIP, "ep_slump_points: selection failed"));
        void *tmp = alloc(bytes);
        if (!tmp || hipcub::DeviceSelect::Flagged(tmp, bytes, it, flag, sel, cnt, int(n)) != hipSuccess ||
            hipMemcpy(&m, cnt, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            return done(fail(EP_E_HIP, "ep_slump_points: selection failed"));
    }
#endif
    *nmp = m;
    if (m > cap) return done(fail(EP_E_ARG, "ep_slump_points: cap " + std::to_string(cap) + " < " + std::to_string(m) +
                                                " kept points (*nmp holds the count)"));
    if (m == 0) return done(EP_OK);
    if (!x_out) return done(fail(EP_E_ARG, "ep_slump_points: null output"));
    auto *dout = static_cast<double *>(alloc(sizeof(double) * dim * size_t(m)));
    if (!dout) return done(fail(EP_E_HIP, "ep_slump_points: hipMalloc failed"));
    EP_LAUNCH(k_lattice_gather, (m + 255) / 256, 256, 0, L, sel, long(m), dout);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(x_out, dout, sizeof(double) * dim * size_t(m), hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(EP_E_HIP, "ep_slump_points: lattice evaluation failed"));
    return done(EP_OK);
}

// ---------------------------------------------------------------------------
// Gaussian random field (spectral randomisation, the method of GRFS_gauss,
// GRF.jl:66-143): g(x) = Σ_h (a_h sin(V_h·x) + b_h cos(V_h·x)) / sqrt(nh)
// at every particle, one thread per particle, harmonics staged in LDS.
// ---------------------------------------------------------------------------
namespace {
constexpr int GRF_TILE = 256;   // harmonics per LDS stage
__global__ void k_grf(int dim, long n, const double *x, int nh, const double *waves, const double *ab, double *g) {
    __shared__ double sv[GRF_TILE * 3], sab[GRF_TILE * 2];
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    double xp[3] = {0.0, 0.0, 0.0};
    if (p < n)
        for (int i = 0; i < dim; i++) xp[i] = x[i + dim * p];
    double acc = 0.0;
    for (int h0 = 0; h0 < nh; h0 += GRF_TILE) {
        const int m = nh - h0 < GRF_TILE ? nh - h0 : GRF_TILE;
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            for (int i = 0; i < 3; i++) sv[3 * k + i] = i < dim ? waves[dim * (h0 + k) + i] : 0.0;
            sab[2 * k] = ab[2 * (h0 + k)];
            sab[2 * k + 1] = ab[2 * (h0 + k) + 1];
        }
        __syncthreads();
        for (int k = 0; k < m; k++) {
            const double ph = sv[3 * k] * xp[0] + sv[3 * k + 1] * xp[1] + sv[3 * k + 2] * xp[2];
            double s, c;
            sincos(ph, &s, &c);
            acc += sab[2 * k] * s + sab[2 * k + 1] * c;
        }
    }
    if (p < n) g[p] = acc / sqrt(double(nh));
}
}  // namespace

extern "C" int ep_grf_gauss(int32_t dim, int64_t n, const double *x, int32_t nh, const double *waves,
                            const double *ab, double *g) {
    if (!x || !waves || !ab || !g || n < 0 || nh <= 0 || dim < 1 || dim > 3)
        return fail(EP_E_ARG, "ep_grf_gauss: bad argument");
    if (n == 0) return EP_OK;
    double *dx = nullptr, *dw = nullptr, *dab = nullptr, *dg = nullptr;
    auto done = [&](int r) { hipFree(dx); hipFree(dw); hipFree(dab); hipFree(dg); return r; };
    if (hipMalloc(&dx, sizeof(double) * dim * n) != hipSuccess || hipMalloc(&dw, sizeof(double) * dim * nh) != hipSuccess ||
        hipMalloc(&dab, sizeof(double) * 2 * nh) != hipSuccess || hipMalloc(&dg, sizeof(double) * n) != hipSuccess)
        return done(fail(EP_E_HIP, "hipMalloc failed"));
    if (hipMemcpy(dx, x, sizeof(double) * dim * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dw, waves, sizeof(double) * dim * nh, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dab, ab, sizeof(double) * 2 * nh, hipMemcpyHostToDevice) != hipSuccess)
        return done(fail(EP_E_HIP, "hipMemcpy failed"));
    EP_LAUNCH(k_grf, (n + EP_BLOCK - 1) / EP_BLOCK, EP_BLOCK, 0, int(dim), long(n), dx, int(nh), dw, dab, dg);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(g, dg, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(EP_E_HIP, "random field evaluation failed"));
    return done(EP_OK);
}

// ---------------------------------------------------------------------------
// ep_grf_gauss_noalloc: the random field with device-resident points and
// values, for callers that cannot allocate.  The harmonic table (the staging
// buffers ep_grf_gauss allocates) lives in `scratch`, a static device array
// of GRF_NH_MAX * 5 doubles (640 KiB) that each GPU the library runs on holds
// until the process ends.  Host code never uses the array's name as a device
// address: its address on each device comes from hipGetSymbolAddress, taken
// once per device with that device current.  Every call on a GPU overwrites
// that GPU's table, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int GRF_NH_MAX = 16384;
static __device__ __align__(256) double scratch[GRF_NH_MAX * 5];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

int scratch_on_current_device(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(EP_E_HIP, "ep_grf_gauss_noalloc: no current device");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(EP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return EP_OK;
}

bool device_resident(const void *p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}
}  // namespace

extern "C" int ep_grf_gauss_noalloc(int32_t dim, int64_t n, const double *x, int32_t nh, const double *waves,
                                    const double *ab, double *g) {
    if (!x || !waves || !ab || !g || n < 0 || nh <= 0 || dim < 1 || dim > 3)
        return fail(EP_E_ARG, "ep_grf_gauss_noalloc: bad argument");
    if (nh > GRF_NH_MAX)
        return fail(EP_E_ARG, "ep_grf_gauss_noalloc: nh = " + std::to_string(nh) + " exceeds the static harmonic table (" +
                                  std::to_string(GRF_NH_MAX) + ")");
    if (n == 0) return EP_OK;
    if (!device_resident(x) || !device_resident(g))
        return fail(EP_E_ARG, "ep_grf_gauss_noalloc: x and g must be device memory");
    double *scr = nullptr;
    const int r = scratch_on_current_device(&scr);
    if (r != EP_OK) return r;
    double *dw = scr, *dab = scr + size_t(dim) * nh;
    if (hipMemcpy(dw, waves, sizeof(double) * dim * nh, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dab, ab, sizeof(double) * 2 * nh, hipMemcpyHostToDevice) != hipSuccess)
        return fail(EP_E_HIP, "ep_grf_gauss_noalloc: staging the harmonics failed");
    EP_LAUNCH(k_grf, (n + EP_BLOCK - 1) / EP_BLOCK, EP_BLOCK, 0, int(dim), long(n), x, int(nh), dw, dab, g);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(0) != hipSuccess)
        return fail(EP_E_HIP, "ep_grf_gauss_noalloc: random field evaluation failed");
    return EP_OK;
}
