// ep_kernels.h -- gfx950 kernels of one explicit MPM step.
//
// Data layout in HBM (all structure-of-arrays, particles in cell-sorted
// order, see ep_api.hip):
//   particle field f of particle p  -> base_f[p]
//   grid field f of node n           -> base_f[n], node n = i_last + nno_last*(...)
//   the node numbering is the reference's (get_coords.jl:36-52): 2D n = iz +
//   nno_z*ix, 3D n = iz + nno_z*(ix + nno_x*iy).
// Shape functions are never stored: every particle kernel re-evaluates the
// 4^d stencil weights from the start-of-step position (two position buffers
// alternate between steps, selected by ctrl->parity on the device).
#pragma once
#include "ep_math.h"
#include <type_traits>

namespace ep {

constexpr int MAXD = 3;
// largest block of the particle kernels (threads); ep_set_layout chooses 64..512
#define EP_PBLK_MAX 512

struct Geom {
    int nno[3];
    int nel[3];
    double h[3];
    double ih[3];              // 1/h
    double x0[3];
    const double *X[3];        // per-axis node coordinates
    const int8_t *T[3];        // per-axis B-spline node type (1..4, 0 invalid)
    int t3lo[3], t3hi[3];      // per axis, the run of type-3 nodes around the middle
    long nno_tot;
};

template <typename T>
struct PState {
    double *x[2][3];           // positions (FP64 in every path), two buffers, ctrl->parity picks
    T *v[3];
    T *F[9];                   // F[i + d*j]
    T *b[6];                   // symmetric b (FP64 path) or B = b - I (FP32 path):
                               // 2D xx,yy,xy; 3D xx,yy,zz,yz,xz,xy
    T *s[6];                   // τ (finite) or σ (infinitesimal), Voigt
    T *epII, *dlam, *J;
    T *rho;                    // ρ₀ (finite: ρ = ρ₀/J) or the constant ρ (infinitesimal)
    T *Om0, *c0, *cr, *phi;
    T *sphi, *cphi;            // tan φ (2D) or sin φ, cos φ (3D) of the constant φ, set at upload
    T *l[3];                   // GIMP half lengths ℓ the basis uses
    T *l0[3];                  // ℓ₀, kept for domain updates other than `undeformed`
    T *gv[9];                  // ∇v (TPIC state)
    T *Bm[9];                  // APIC Bᵢⱼ [i + D*j] (Bij.jl:14-26)
    T *epII2;                  // ϵpII[2]: non-local plastic strain of the step (non-local contexts)
    long n;
};

struct GState {
    double *m, *p[3], *f[3], *a[3], *v[3], *pa[3], *dJ;
    const uint8_t *bc;         // bit d set: Dirichlet in axis d
    long nno;
};

struct Ctrl {
    double t, dt, T, G, tg, dt_fixed;
    double g[3];
    unsigned long long vmax[3];  // bit patterns of non-negative doubles
    long it, steps;
    int active, parity, ramp, dt_mode, with_plast, record;
    unsigned int err;
    int j2cap;
    double *dt_log;
    long dt_cap;
    unsigned long long *vmax_part;  // [3][nblk] per-block maxima of |v_p| written by G2P
    int nblk;
    int spill;                      // particle blocks of p2g whose node box did not fit LDS (cumulative)
    unsigned long long boxsum;      // node-box volumes of the p2g blocks (cumulative): how far the order has decayed
    // learned return mapping (FP32 + surrogate path)
    int qcount;                     // plastic trial states queued this step
    unsigned long long s_acc, s_fb; // predictions accepted / replaced by the FP64 return
    unsigned s_maxr;                // bit pattern of the largest scaled residual
};

enum : unsigned { ERR_DOMAIN = 1u, ERR_NONFINITE = 2u };

// Queue of plastic trial states for the learned return (capacity nmp):
// particle slot, principal log strains sorted descending and the matching
// eigenvectors, V[i + 3k] = component i of eigenvector k.
struct SurrQ {
    int *idx;
    float *e[3];
    float *V[9];
};

// Non-local regularisation: particles binned by the element of their
// start-of-step position (p2e) each step.
struct NLocal {
    int *count, *start, *list, *cell;  // per element: count / first list entry; per slot: element
    const int *orig;                   // slot -> the caller's particle index (the reference's order)
    long ncell;
    int nnel[3];                       // ceil(ls / h) per axis
    double ls;
};

struct Mat {
    double Kc, Gc, Hp, c, C_pf, damping;
    double E, nu;
    double Del[36];
    DelLU lu;
    int finite, tpic, musl, locking, plast, apic, domain, nonlocal;
    SurrQ q;
    NLocal nl;
};

// -------------------------------------------------------------------------
// stencil weights
// -------------------------------------------------------------------------
template <int D, typename T>
struct W {
    T w[D][4], dw[D][4];
    int base[D];
    unsigned ok;    // bit (4*axis + k): node exists
    // A lane past the last particle: no node, zero weights, and a base
    // inside the block's node box so that box walks stay in bounds.
    __device__ __forceinline__ void retire(const int *lo) {
        ok = 0u;
#pragma unroll
        for (int i = 0; i < D; i++) {
            base[i] = lo[i];
#pragma unroll
            for (int k = 0; k < 4; k++) { w[i][k] = T(0); dw[i][k] = T(0); }
        }
    }
};

// shpfun: p2e2n + basis (tplgy.jl:43-96, bsmpm.jl:95-135, gimpm.jl:31-71,
// smpm.jl:27-66).  Returns false on a bsmpm "Invalid position".
template <int D, int B, typename T>
__device__ __forceinline__ bool weights(const Geom &g, const double *xp, const T *lp, W<D, T> &w) {
    bool good = true;
    w.ok = 0u;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const double a = xp[i] - g.x0[i];
        double e = exact_fld(a, g.h[i], g.ih[i]);
        const double hi = double(g.nel[i] - 1);
        e = e < 0.0 ? 0.0 : (e > hi ? hi : e);
        const int base = int(e) - 1;
        w.base[i] = base;
        const T h = T(g.h[i]), ih = T(g.ih[i]);
        if (B == 0 && base >= g.t3lo[i] && base + 3 <= g.t3hi[i]) {
            // Four interior nodes: when every ξ lies where bsmpm's first
            // matching branch is the expected piece (k = 0: (1, 2], 1: (0, 1],
            // 2: (-1, 0], 3: [-2, -1]), those pieces are the uniform cubic
            // B-spline's, written in u = ξ₁ and t = -ξ₂ (= 1 - u up to
            // rounding): (2-ξ₀)³/6 = t³/6, 2/3 - u² + u³/2, 2/3 - t² + t³/2,
            // (2+ξ₃)³/6 = u³/6, and their derivatives -t²/2, 1.5u² - 2u,
            // 2t - 1.5t², u²/2 (bsmpm.jl:1-62 expanded; equal up to rounding)
            T xi[4];
#pragma unroll
            for (int k = 0; k < 4; k++) xi[k] = T((xp[i] - g.X[i][base + k]) * g.ih[i]);
            if (xi[0] > T(1) && xi[0] <= T(2) && xi[1] > T(0) && xi[1] <= T(1) && xi[2] > T(-1) && xi[2] <= T(0) &&
                xi[3] >= T(-2) && xi[3] <= T(-1)) {
                const T c16 = T(1.0 / 6.0), c23 = T(2.0 / 3.0);
                const T u = xi[1], t = -xi[2];
                const T u2 = u * u, u3 = u2 * u, t2 = t * t, t3 = t2 * t;
                w.w[i][0] = c16 * t3;
                w.dw[i][0] = T(-0.5) * t2 * ih;
                w.w[i][1] = T(0.5) * u3 - u2 + c23;
                w.dw[i][1] = (T(1.5) * u2 - T(2) * u) * ih;
                w.w[i][2] = T(0.5) * t3 - t2 + c23;
                w.dw[i][2] = (T(2) * t - T(1.5) * t2) * ih;
                w.w[i][3] = c16 * u3;
                w.dw[i][3] = T(0.5) * u2 * ih;
                w.ok |= 0xFu << (4 * i);
                continue;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int n = base + k;
            T f = T(0), df = T(0);
            if (n >= 0 && n < g.nno[i]) {
                w.ok |= 1u << (4 * i + k);
                // x_p - x_n is formed in FP64 before the weight arithmetic
                const double dxd = xp[i] - g.X[i][n];
                if (B == 0) {
                    const int ty = g.T[i][n];
                    if (ty == 0) good = false;
                    T ff, dd;
                    bspline<T>(T(dxd * g.ih[i]), ty, ff, dd);
                    f = ff;
                    df = dd * ih;
                } else if (B == 1) {
                    gimp<T>(T(dxd), h, lp[i], f, df);
                } else {
                    linear<T>(T(dxd), h, f, df);
                }
            }
            w.w[i][k] = f;
            w.dw[i][k] = df;
        }
    }
    return good;
}

// Iterate the stencil in the reference's e2n order (e2n.jl:37-75):
// 2D x outer, z inner; 3D y outer, x, z inner.  fn(node, N, dN[D]).
template <int D, typename T, typename Fn>
__device__ __forceinline__ void for_stencil(const Geom &g, const W<D, T> &w, Fn &&fn) {
    if (D == 2) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (!(w.ok >> i & 1u)) continue;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (!(w.ok >> (4 + j) & 1u)) continue;
                const long no = long(w.base[1] + j) + long(g.nno[1]) * long(w.base[0] + i);
                T dN[D];
                const int ix[2] = {w.base[0] + i, w.base[1] + j};
                const T N = w.w[0][i] * w.w[1][j];
                dN[0] = w.dw[0][i] * w.w[1][j];
                dN[1] = w.w[0][i] * w.dw[1][j];
                fn(no, N, dN, ix);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!(w.ok >> (4 + k) & 1u)) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (!(w.ok >> i & 1u)) continue;
                const T wxy = w.w[0][i] * w.w[1][k];
                const T dxy = w.dw[0][i] * w.w[1][k];
                const T xdy = w.w[0][i] * w.dw[1][k];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (!(w.ok >> (8 + j) & 1u)) continue;
                    const long no = long(w.base[2] + j) +
                                    long(g.nno[2]) * (long(w.base[0] + i) + long(g.nno[0]) * long(w.base[1] + k));
                    T dN[D];
                    const int ix[3] = {w.base[0] + i, w.base[1] + k, w.base[2] + j};
                    const T N = wxy * w.w[2][j];
                    dN[0] = dxy * w.w[2][j];
                    dN[1] = xdy * w.w[2][j];
                    dN[2] = wxy * w.dw[2][j];
                    fn(no, N, dN, ix);
                }
            }
        }
    }
}

__device__ __forceinline__ void atomic_add(double *a, double v) { unsafeAtomicAdd(a, v); }

__device__ __forceinline__ void atomic_max_pos(unsigned long long *a, double v) {
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ __forceinline__ T mass_of(const PState<T> &P, const Mat &M, long p, T J) {
    // ρ*Ω with ρ = ρ₀/J (deform.jl:21, finite) and Ω = J*Ω₀ (deform.jl:19)
    const T Om = J * P.Om0[p];
    return M.finite ? (P.rho[p] / J) * Om : P.rho[p] * Om;
}


// -------------------------------------------------------------------------
// Block node box.  A block of 256 brick-sorted particles touches a small box
// of nodes; the block reduces the box of its stencils and
//  - for gathers, stages the nodal fields of the box in LDS with coalesced
//    loads, so each particle reads its 4^d nodes from LDS;
//  - for scatters, accumulates contributions in LDS (ds_add_f64) and flushes
//    the box with one global atomic per node value.
// When the box does not fit the LDS budget (a block spanning distant cells)
// the kernels fall back to global loads / global atomics.
// -------------------------------------------------------------------------
constexpr int ACC_LDS = 4096;   // doubles of LDS per block (32 KiB)
// P2G's box (7 fields in 3D): 39.9 KiB, so that 4 blocks still share a CU's
// 160 KiB as with ACC_LDS (occupancy is LDS-bound either way)
constexpr int ACC_LDS_P2G = 5104;
// G2P's box (6 fields staged, then 3 scattered): 24 KiB, so that 6 blocks
// share a CU (its 80 VGPRs allow 6 waves per SIMD)
constexpr int ACC_LDS_G2P = 3072;
// deform (3 fields staged, 1 scattered): 16 KiB; F-bar (4 fields): 24 KiB.
// Blocks per CU are then bounded by registers, not LDS (FP32 deform 64
// VGPRs: 8 waves per SIMD; FP32 F-bar 99: 5)
constexpr int ACC_LDS_DEFORM = 2048;
constexpr int ACC_LDS_FBAR = 3072;
// Box layout.  The LDS serves a 64-bit access as four quarter-waves of 16
// lanes, each through 32 four-byte banks, i.e. 16 bank pairs per double; the
// measured bank-conflict counts of the walks match that model exactly.
// A box node at local (z, x, y) keeps field v of an NV-field box at
//   16*NV*G + 16*v + (z & 3) + 4*(x & 3),   G = z/4 + nz4*(x/4 + nx4*y),
// so its bank pair is (z & 3) + 4*(x & 3) whatever the field.  Scatter walks
// start each lane at the node with z = lane (mod 4), x = lane/4 (mod 4)
// (for_stencil_box): the 16 lanes of a quarter-wave then hit 16 different
// bank pairs, hence 16 different nodes, at every step of the walk for any
// spread of particles over cells.  Fields of one node sit 16 doubles apart,
// immediate offsets of one address.
template <int D>
struct BlockBox {
    int lo[D], ext[D];
    int nloc;
    long vol;                  // nodes of the block's box whether or not it fits
    int nz4, nx4, ey;          // box extent in groups of 4 along z and x; y extent (1 in 2D)
    int cap;                   // doubles of LDS behind buf
    bool any;                  // the block has a live particle
    EP_LDS double *buf;

    __device__ __forceinline__ void begin(EP_LDS double *smem, int *sint, bool live, const int *base,
                                          int capacity = ACC_LDS) {
        buf = smem;
        cap = capacity;
        if (threadIdx.x < D) { sint[threadIdx.x] = 0x7fffffff; sint[D + threadIdx.x] = -0x7fffffff; }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int i = 0; i < D; i++) { atomicMin(&sint[i], base[i]); atomicMax(&sint[D + i], base[i]); }
        }
        __syncthreads();
        long n = 1;
        // the box is block-uniform: keep it in scalar registers
        int mn[D], mx[D];
#pragma unroll
        for (int i = 0; i < D; i++) { mn[i] = uniform(sint[i]); mx[i] = uniform(sint[D + i]); }
        bool ok = mn[0] != 0x7fffffff;
        vol = 1;
#pragma unroll
        for (int i = 0; i < D; i++) vol *= ok ? long(mx[i]) - long(mn[i]) + 4 : 0;
#pragma unroll
        for (int i = 0; i < D; i++) {
            lo[i] = mn[i];
            const long e = long(mx[i]) - long(mn[i]) + 4;
            ok = ok && e >= 4 && e <= cap;
            ext[i] = ok ? int(e) : 4;
            n *= ext[i];
        }
        any = mn[0] != 0x7fffffff;
        nloc = (ok && n <= cap) ? int(n) : 0;
        nz4 = (ext[D - 1] + 3) >> 2;
        nx4 = (ext[0] + 3) >> 2;
        ey = D == 3 ? ext[1] : 1;
        __syncthreads();   // everyone read sint before a later begin() rewrites it
    }
    // doubles one field group of the layout spans (NV fields: NV times this)
    __device__ __forceinline__ int span() const { return 16 * nz4 * nx4 * ey; }
    __device__ __forceinline__ bool fits(int nv) const { return nloc > 0 && long(span()) * nv <= cap; }
    // address of field 0 of local node (z, x, y) in an NV-field box
    template <int NV>
    __device__ __forceinline__ int addr(int z, int x, int y) const {
        return 16 * NV * ((z >> 2) + nz4 * ((x >> 2) + nx4 * y)) + (z & 3) + 4 * (x & 3);
    }
    // box entries run as the grid: last axis (z) fastest, then x, then y;
    // local coordinates of entry k and its global node, or -1 outside the grid
    __device__ __forceinline__ long node(const Geom &g, int k, int &z, int &x, int &y) const {
        z = k % ext[D - 1];
        const int r = k / ext[D - 1];
        x = D == 2 ? r : r % ext[0];
        y = D == 2 ? 0 : r / ext[0];
        if (D == 2) {
            const int ix0 = lo[0] + x, ix1 = lo[1] + z;
            if (ix0 < 0 || ix1 < 0 || ix0 >= g.nno[0] || ix1 >= g.nno[1]) return -1;
            return long(ix1) + long(g.nno[1]) * ix0;
        }
        const int ix0 = lo[0] + x, ix1 = lo[1] + y, ix2 = lo[2] + z;
        if (ix0 < 0 || ix1 < 0 || ix2 < 0 || ix0 >= g.nno[0] || ix1 >= g.nno[1] || ix2 >= g.nno[2]) return -1;
        return long(ix2) + long(g.nno[2]) * (long(ix0) + long(g.nno[0]) * ix1);
    }
    // the box memory as elements of type E (a gather of the FP32 path stages
    // FP32 node values: the gathers convert every node value to the particle
    // precision anyway, so staging them converted gives the same results with
    // half the LDS reads)
    template <typename E>
    __device__ __forceinline__ EP_LDS E *as() const { return (EP_LDS E *)buf; }
    // LDS field v of every box node = E(src[v][node]) (0 outside the grid); caller checks fits(NV)
    template <int NV, typename E = double>
    __device__ __forceinline__ void stage(const Geom &g, const double *const *src) {
        EP_LDS E *eb = as<E>();
        for (int k = threadIdx.x; k < nloc; k += blockDim.x) {
            int z, x, y;
            const long no = node(g, k, z, x, y);
            const int a = addr<NV>(z, x, y);
#pragma unroll
            for (int v = 0; v < NV; v++) eb[a + 16 * v] = no >= 0 ? E(src[v][no]) : E(0);
        }
        __syncthreads();
    }
    __device__ __forceinline__ void zero(int nv) {
        __syncthreads();   // finished reading staged values
        const int n = span() * nv;
        for (int k = threadIdx.x; k < n; k += blockDim.x) buf[k] = 0.0;
        __syncthreads();
    }
    template <int NV>
    __device__ __forceinline__ void flush(const Geom &g, double *const *dst) {
        __syncthreads();
        for (int k = threadIdx.x; k < nloc; k += blockDim.x) {
            int z, x, y;
            const long no = node(g, k, z, x, y);
            if (no < 0) continue;
            const int a = addr<NV>(z, x, y);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                const double val = buf[a + 16 * v];
                if (val != 0.0) unsafeAtomicAdd(&dst[v][no], val);
            }
        }
    }
};

// Scatter targets: the block's LDS box or the global nodal arrays.  Each
// scatter loop is instantiated once per sink, so LDS traffic stays ds_*.
template <int D, int NV>
struct LdsSink {
    BlockBox<D> *b;
    __device__ __forceinline__ void operator()(int v, long, int li, double val) const {
        lds_atomic_add(&b->buf[li + 16 * v], val);
    }
};
struct GlobSink {
    double *const *dst;
    __device__ __forceinline__ void operator()(int v, long no, int, double val) const {
        unsafeAtomicAdd(&dst[v][no], val);
    }
};
// Gather sources: the staged LDS box (elements of type E) or the global nodal arrays.
template <int D, int NV, typename E = double>
struct LdsSrc {
    const BlockBox<D> *b;
    __device__ __forceinline__ double operator()(int v, long, int li) const { return double(b->template as<E>()[li + 16 * v]); }
};
struct GlobSrc {
    const double *const *src;
    __device__ __forceinline__ double operator()(int v, long no, int) const { return src[v][no]; }
};

__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// a[i & 3] from registers.  Written as two bit tests so that it lowers to
// v_cndmask selects: the compare chain i == 0 / 1 / 2 was turned into a
// switch and lowered as divergent branches (about 20 instructions, 12 of them
// scalar, per select).
template <typename T>
__device__ __forceinline__ T sel4(const T *a, int i) {
    const bool b0 = (i & 1) != 0, b1 = (i & 2) != 0;
    const T lo = b0 ? a[1] : a[0];
    const T hi = b0 ? a[3] : a[2];
    return b1 ? hi : lo;
}

// Scatter iteration over the 4^d stencil with a lane-dependent rotation of
// the x and z offsets: the particles of one cell sit in consecutive lanes,
// and the rotation sends them to different nodes at the same instant, so
// their LDS atomics do not collide.  Scatter order does not matter (the
// reference accumulates with atomics too); gathers keep for_stencil's
// reference order.
template <int D, typename T, typename Fn>
__device__ __forceinline__ void for_stencil_rot(const Geom &g, const W<D, T> &w, Fn &&fn) {
    const int lane = threadIdx.x & 63;
    if (D == 2) {
        const int rz = lane & 3;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (!(w.ok >> i & 1u)) continue;
#pragma unroll
            for (int j0 = 0; j0 < 4; j0++) {
                const int j = (j0 + rz) & 3;
                if (!(w.ok >> (4 + j) & 1u)) continue;
                const int ix[2] = {w.base[0] + i, w.base[1] + j};
                const long no = long(ix[1]) + long(g.nno[1]) * long(ix[0]);
                const T wz = sel4(w.w[1], j), dwz = sel4(w.dw[1], j);
                T dN[D];
                const T N = w.w[0][i] * wz;
                dN[0] = w.dw[0][i] * wz;
                dN[1] = w.w[0][i] * dwz;
                fn(no, N, dN, ix);
            }
        }
    } else {
        const int rz = lane & 3, rx = (lane >> 2) & 3;
        // the fallback of blocks whose box does not fit: rolled, so that it
        // does not raise the register count of the kernel's box path
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            if (!(w.ok >> (4 + k) & 1u)) continue;
#pragma unroll 1
            for (int i0 = 0; i0 < 4; i0++) {
                const int i = (i0 + rx) & 3;
                if (!(w.ok >> i & 1u)) continue;
                const T wx = sel4(w.w[0], i), dwx = sel4(w.dw[0], i);
                const T wy = sel4(w.w[1], k), dwy = sel4(w.dw[1], k);
                const T wxy = wx * wy;
                const T dxy = dwx * wy;
                const T xdy = wx * dwy;
#pragma unroll
                for (int j0 = 0; j0 < 4; j0++) {
                    const int j = (j0 + rz) & 3;
                    if (!(w.ok >> (8 + j) & 1u)) continue;
                    const int ix[3] = {w.base[0] + i, w.base[1] + k, w.base[2] + j};
                    const long no = long(ix[2]) + long(g.nno[2]) * (long(ix[0]) + long(g.nno[0]) * long(ix[1]));
                    const T wz = sel4(w.w[2], j), dwz = sel4(w.dw[2], j);
                    T dN[D];
                    const T N = wxy * wz;
                    dN[0] = dxy * wz;
                    dN[1] = xdy * wz;
                    dN[2] = wxy * dwz;
                    fn(no, N, dN, ix);
                }
            }
        }
    }
}

// Stencil walk over a block's LDS node box: no existence tests (nodes
// outside the grid have zero weights, and their box slots are never
// flushed), the box address of each visited node from per-axis parts
// computed once per particle.  Scatters (ROT) start lane l's walk at the node
// whose local z is l (mod 4) and local x is l/4 (mod 4), and y at offset
// (l/16) mod 4: with the box layout above, the 16 lanes of a quarter-wave
// touch 16 different LDS bank pairs at every step, whatever cells their
// particles lie in.  Without ROT the order is for_stencil's reference order,
// and the skipped nodes of for_stencil add exact zeros here.  fn(node, box
// address of field 0, N, dN, axis indices); field v is 16*v further.
template <int D, typename T, bool ROT, int NV, typename Fn>
__device__ __forceinline__ void for_stencil_box(const BlockBox<D> &b, const W<D, T> &w, Fn &&fn) {
    const int lane = threadIdx.x & 63;
    const int zb = w.base[D - 1] - b.lo[D - 1], xb = w.base[0] - b.lo[0];
    const int yb = D == 3 ? w.base[1] - b.lo[1] : 0;
    const int rz = ROT ? ((lane - zb) & 3) : 0;
    const int rx = ROT ? (((lane >> 2) - xb) & 3) : 0;
    const int ry = ROT ? ((lane >> 4) & 3) : 0;
    T wz[4], dwz[4], wx[4], dwx[4];
    int oz[4], ox[4], az[4], ax[4];
#pragma unroll
    for (int j0 = 0; j0 < 4; j0++) {
        const int j = (j0 + rz) & 3;
        oz[j0] = j;
        wz[j0] = ROT ? sel4(w.w[D - 1], j) : w.w[D - 1][j0];
        dwz[j0] = ROT ? sel4(w.dw[D - 1], j) : w.dw[D - 1][j0];
        const int z = zb + j;
        az[j0] = 16 * NV * (z >> 2) + (z & 3);
        const int i = (j0 + rx) & 3;
        ox[j0] = i;
        wx[j0] = ROT ? sel4(w.w[0], i) : w.w[0][j0];
        dwx[j0] = ROT ? sel4(w.dw[0], i) : w.dw[0][j0];
        const int x = xb + i;
        ax[j0] = 16 * NV * b.nz4 * (x >> 2) + 4 * (x & 3);
    }
    if constexpr (D == 2) {
#pragma unroll
        for (int i0 = 0; i0 < 4; i0++) {
#pragma unroll
            for (int j0 = 0; j0 < 4; j0++) {
                T dN[D];
                const T N = wx[i0] * wz[j0];
                dN[0] = dwx[i0] * wz[j0];
                dN[1] = wx[i0] * dwz[j0];
                const int ix[2] = {w.base[0] + ox[i0], w.base[1] + oz[j0]};
                fn(0L, ax[i0] + az[j0], N, dN, ix);
            }
        }
    } else {
        const int sy = 16 * NV * b.nz4 * b.nx4;
        if constexpr (ROT && NV >= 5) {
            // many-field scatters (P2G): one (y, x) row of 4 nodes per
            // iteration, the row's weights picked with selects; rolling the 16
            // rows keeps the register count down (FP64 P2G 225 -> 148 VGPRs,
            // 3 waves per SIMD instead of 2: 1603 -> 1210 us)
#pragma unroll 1
            for (int m = 0; m < 16; m++) {
                const int k = ((m >> 2) + ry) & 3, i = ((m & 3) + rx) & 3;
                const T wy = sel4(w.w[1], k), dwy = sel4(w.dw[1], k);
                const T wxv = sel4(w.w[0], i), dwxv = sel4(w.dw[0], i);
                const int x = xb + i;
                const int row = sy * (yb + k) + 16 * NV * b.nz4 * (x >> 2) + 4 * (x & 3);
                const T wxy = wxv * wy;
                const T dxy = dwxv * wy;
                const T xdy = wxv * dwy;
#pragma unroll
                for (int j0 = 0; j0 < 4; j0++) {
                    T dN[D];
                    const T N = wxy * wz[j0];
                    dN[0] = dxy * wz[j0];
                    dN[1] = xdy * wz[j0];
                    dN[2] = wxy * dwz[j0];
                    const int ix[3] = {w.base[0] + i, w.base[1] + k, w.base[2] + oz[j0]};
                    fn(0L, row + az[j0], N, dN, ix);
                }
            }
        } else {
#pragma unroll
            for (int k0 = 0; k0 < 4; k0++) {
                const int k = ROT ? ((k0 + ry) & 3) : k0;
                const T wy = ROT ? sel4(w.w[1], k) : w.w[1][k0];
                const T dwy = ROT ? sel4(w.dw[1], k) : w.dw[1][k0];
                const int ay = sy * (yb + k);
#pragma unroll
                for (int i0 = 0; i0 < 4; i0++) {
                    const T wxy = wx[i0] * wy;
                    const T dxy = dwx[i0] * wy;
                    const T xdy = wx[i0] * dwy;
                    const int row = ay + ax[i0];
#pragma unroll
                    for (int j0 = 0; j0 < 4; j0++) {
                        T dN[D];
                        const T N = wxy * wz[j0];
                        dN[0] = dxy * wz[j0];
                        dN[1] = xdy * wz[j0];
                        dN[2] = wxy * dwz[j0];
                        const int ix[3] = {w.base[0] + ox[i0], w.base[1] + k, w.base[2] + oz[j0]};
                        fn(0L, row + az[j0], N, dN, ix);
                    }
                }
            }
        }
    }
}

// The 3D box scatter by rows: fn(row, wxy, dxy, xdy) once per (y, x) row of
// the stencil, in for_stencil_box's bank-aware order (ROT); the row's 4 z
// nodes sit at row + z.a[j0] with weights z.w[j0], derivatives z.dw[j0].
// A scatter whose nodal values are linear in (N, ∂N/∂z) for fixed row
// factors (P2G's m, p and f) then spends two products per value and node.
template <typename T>
struct ZNodes {
    T w[4], dw[4];
    int a[4];
};
template <typename T, int NV, typename Fn>
__device__ __forceinline__ void box_rows_3d(const BlockBox<3> &b, const W<3, T> &w, ZNodes<T> &z, Fn &&fn) {
    const int lane = threadIdx.x & 63;
    const int zb = w.base[2] - b.lo[2], xb = w.base[0] - b.lo[0], yb = w.base[1] - b.lo[1];
    const int rz = (lane - zb) & 3, rx = ((lane >> 2) - xb) & 3, ry = (lane >> 4) & 3;
#pragma unroll
    for (int j0 = 0; j0 < 4; j0++) {
        const int j = (j0 + rz) & 3;
        z.w[j0] = sel4(w.w[2], j);
        z.dw[j0] = sel4(w.dw[2], j);
        const int zz = zb + j;
        z.a[j0] = 16 * NV * (zz >> 2) + (zz & 3);
    }
    const int sy = 16 * NV * b.nz4 * b.nx4;
#pragma unroll 1
    for (int m = 0; m < 16; m++) {
        const int k = ((m >> 2) + ry) & 3, i = ((m & 3) + rx) & 3;
        const T wy = sel4(w.w[1], k), dwy = sel4(w.dw[1], k);
        const T wxv = sel4(w.w[0], i), dwxv = sel4(w.dw[0], i);
        const int x = xb + i;
        fn(sy * (yb + k) + 16 * NV * b.nz4 * (x >> 2) + 4 * (x & 3), wxv * wy, dwxv * wy, wxv * dwy);
    }
}

// Walks for the two sinks / sources; fn(node, box index, N, dN, axis
// indices).  On the box walks the axis indices may lie outside the grid
// (zero weight): clamp them before reading node coordinates.
template <int D, typename T, int NV, typename Fn>
__device__ __forceinline__ void scatter_walk(const Geom &, const W<D, T> &w, const LdsSink<D, NV> &s, Fn &&fn) {
    for_stencil_box<D, T, true, NV>(*s.b, w, fn);
}
template <int D, typename T, typename Fn>
__device__ __forceinline__ void scatter_walk(const Geom &g, const W<D, T> &w, const GlobSink &, Fn &&fn) {
    for_stencil_rot<D, T>(g, w, [&](long no, T N, const T *dN, const int *ix) { fn(no, 0, N, dN, ix); });
}
template <int D, typename T, int NV, typename E, typename Fn>
__device__ __forceinline__ void gather_walk(const Geom &, const W<D, T> &w, const LdsSrc<D, NV, E> &s, Fn &&fn) {
    for_stencil_box<D, T, false, NV>(*s.b, w, fn);
}
template <int D, typename T, typename Fn>
__device__ __forceinline__ void gather_walk(const Geom &g, const W<D, T> &w, const GlobSrc &, Fn &&fn) {
    for_stencil<D, T>(g, w, [&](long no, T N, const T *dN, const int *ix) { fn(no, 0, N, dN, ix); });
}

// Block max of non-negative doubles (as ordered bit patterns): wavefront
// butterfly, then across the block's waves through LDS; thread 0 returns it.
__device__ __forceinline__ unsigned long long block_max_bits(unsigned long long b, unsigned long long *sred) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off);
        b = o > b ? o : b;
    }
    const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) sred[wv] = b;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < nw; k++) b = sred[k] > b ? sred[k] : b;
    __syncthreads();
    return b;
}

// N block maxima with one pair of barriers (sred holds N x 16 entries, so
// blocks of at most 1024 threads); thread 0 returns them in b
template <int N>
__device__ __forceinline__ void block_max_bits_n(unsigned long long (&b)[N], unsigned long long *sred) {
#pragma unroll
    for (int d = 0; d < N; d++)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const unsigned long long o = __shfl_xor(b[d], off);
            b[d] = o > b[d] ? o : b[d];
        }
    const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int d = 0; d < N; d++) sred[16 * d + wv] = b[d];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < N; d++)
            for (int k = 1; k < nw; k++) b[d] = sred[16 * d + k] > b[d] ? sred[16 * d + k] : b[d];
    __syncthreads();
}

template <int D>
__device__ __forceinline__ void stencil_base(const Geom &g, const double *xp, int *base) {
#pragma unroll
    for (int i = 0; i < D; i++) {
        double e = exact_fld(xp[i] - g.x0[i], g.h[i], g.ih[i]);
        const double hi = double(g.nel[i] - 1);
        e = e < 0.0 ? 0.0 : (e > hi ? hi : e);
        base[i] = int(e) - 1;
    }
}

// -------------------------------------------------------------------------
// step kernels
// -------------------------------------------------------------------------

// get_dt / get_spacetime (physics.jl:21-26, 80-90) + loop guard
// `while T > time.t[1]` (worflow.jl:28, 74).
template <int D>
__device__ __forceinline__ void b_begin(Ctrl *c, const Geom &g, const Mat &M) {
    const double t = c->t;
    c->qcount = 0;
    c->active = (c->T > t) ? 1 : 0;
    if (!c->active) return;
    double dt;
    if (c->dt_mode == 0) {
        double cmax = -HUGE_VAL;
        for (int i = 0; i < D; i++) {
            const double vm = __longlong_as_double((long long)c->vmax[i]);
            const double r = g.h[i] / (vm + M.c);
            cmax = r > cmax ? r : cmax;
        }
        const double a = 0.5 * cmax, b = c->T - t;
        dt = a < b ? a : b;
    } else {
        dt = c->dt_fixed;
    }
    double G = c->G;
    if (c->ramp) G = (t <= c->tg) ? (c->G * t / c->tg) : c->G;
    for (int i = 0; i < 3; i++) c->g[i] = 0.0;
    c->g[D - 1] = -G;
    c->dt = dt;
    if (c->record && c->steps < c->dt_cap) c->dt_log[c->steps] = dt;
}

template <int D>
__global__ void k_begin(Ctrl *c, Geom g, Mat M) {
    if (threadIdx.x == 0) b_begin<D>(c, g, M);
}

// End of step: vmax = max over G2P's per-block maxima (the atomic max of
// n2p.jl:37), then time.t[1] += dt (worflow.jl:38, 84).
__device__ __forceinline__ void b_end(Ctrl *c) {
    __shared__ unsigned long long sred[48];   // 3 x 16 waves of a 1024-thread block
    if (!c->active) return;
    const int nblk = c->nblk;
    const unsigned long long *part = c->vmax_part;
    unsigned long long b[3] = {0ull, 0ull, 0ull};
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const unsigned long long o = part[d * nblk + k];
            b[d] = o > b[d] ? o : b[d];
        }
    }
    block_max_bits_n<3>(b, sred);
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int d = 0; d < 3; d++) c->vmax[d] = b[d];
    c->t = c->t + c->dt;
    c->it += 1;
    c->steps += 1;
    c->parity ^= 1;
}

__global__ void k_end(Ctrl *c) { b_end(c); }

// zero mᵢ, p, oobf (p2n.jl:290-292), the mUSL momentum and ΔJ (update.jl:53)
template <int D>
__device__ __forceinline__ void b_zero_grid(const Ctrl *c, const GState &G, long blk) {
    const long n = blk * long(blockDim.x) + threadIdx.x;
    if (n >= G.nno || !c->active) return;
    G.m[n] = 0.0;
    G.dJ[n] = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) { G.p[d][n] = 0.0; G.f[d][n] = 0.0; G.pa[d][n] = 0.0; }
}

// kernel of one context: block b of the launch is block b of the particles / nodes
template <int D>
__global__ void k_zero_grid(const Ctrl *c, GState G) { b_zero_grid<D>(c, G, blockIdx.x); }

// APIC's affine matrix B·D⁻¹ (p2n.jl:243-249, 267-273).  Dij_nd
// (misc.jl:15-27) assigns Dᵢⱼ = N δδᵀ at every existing stencil node instead
// of summing, so D is that of the last existing node in e2n order (y, x, z
// nesting, e2n.jl): the largest existing index on each axis.  D⁻¹ = inv(D)
// when |det D| > 1e-12, else the identity; a rank-one D gives the identity.
template <int D, typename T>
__device__ __forceinline__ void apic_affine(const Geom &g, const W<D, T> &w, const double *xp, const PState<T> &P,
                                            long q, T (&A)[D][D]) {
    double Bm[D][D], Dm[D][D], dl[D], N = 1.0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) Bm[i][j] = double(P.Bm[i + D * j][q]);
#pragma unroll
    for (int i = 0; i < D; i++) {
        const unsigned bits = (w.ok >> (4 * i)) & 0xFu;
        const int k = bits ? 31 - __builtin_clz(bits) : 0;
        const double wk = k == 0 ? double(w.w[i][0]) : (k == 1 ? double(w.w[i][1]) : (k == 2 ? double(w.w[i][2]) : double(w.w[i][3])));
        N = i == 0 ? wk : N * wk;
        dl[i] = g.X[i][clampi(w.base[i] + k, g.nno[i])] - xp[i];
    }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) Dm[i][j] = N * (dl[i] * dl[j]);
    if (fabs(det<D, double>(Dm)) > 1.0e-12) {
        double Di[D][D];
        inv_lu<D>(Dm, Di);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) {
                double s = Bm[i][0] * Di[0][j];
#pragma unroll
                for (int k = 1; k < D; k++) s += Bm[i][k] * Di[k][j];
                A[i][j] = T(s);
            }
    } else {
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) A[i][j] = T(Bm[i][j]);
    }
}

// transform + std/tpic/apic_{2,3}d_p2n (p2n.jl:12-21, 58-102, 140-191, 231-280)
// AFF: the affine momentum term of TPIC / APIC, its own instantiation so that
// the standard transfer carries no ∇v registers through its scatter
template <int D, int B, typename T, bool AFF>
__device__ __forceinline__ void b_p2g(const Ctrl *c, const Geom &g, const PState<T> &P, const GState &G, const Mat &M, long blk) {
    constexpr int NS = D == 2 ? 3 : 6;
    constexpr int NV = 1 + 2 * D;
    __shared__ double smem[ACC_LDS_P2G];
    __shared__ int sint[8];
    if (!c->active) return;
    const long p = blk * long(blockDim.x) + threadIdx.x;
    const long q = p < P.n ? p : 0;
    const bool live = p < P.n && P.Om0[q] != T(0);   // Ω₀ = 0: a padding slot
    const int par = c->parity;
    double xp[D];
    T lp[D];
#pragma unroll
    for (int i = 0; i < D; i++) { xp[i] = P.x[par][i][q]; lp[i] = B == 1 ? P.l[i][q] : T(0); }
    W<D, T> w;
    if (!weights<D, B, T>(g, xp, lp, w) && live) atomicOr(const_cast<unsigned *>(&c->err), ERR_DOMAIN);
    BlockBox<D> box;
    box.begin((EP_LDS double *)smem, sint, live, w.base, ACC_LDS_P2G);
    const bool lds = box.fits(NV);
    if (threadIdx.x == 0 && box.any) {
        if (!lds) atomicAdd(const_cast<int *>(&c->spill), 1);
        atomicAdd(const_cast<unsigned long long *>(&c->boxsum), (unsigned long long)box.vol);
    }
    double *dst[NV];
    dst[0] = G.m;
#pragma unroll
    for (int d = 0; d < D; d++) { dst[1 + d] = G.p[d]; dst[1 + D + d] = G.f[d]; }
    const T J = P.J[q];
    const T Om = J * P.Om0[q];
    const T ms = mass_of(P, M, q, J);
    T s[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) s[k] = M.finite ? P.s[k][q] / J : P.s[k][q];
    T v[D], pm[D], gv[D][D];
#pragma unroll
    for (int i = 0; i < D; i++) { v[i] = P.v[i][q]; pm[i] = ms * v[i]; }
    if constexpr (AFF) {
        if (M.tpic) {
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) gv[i][j] = P.gv[i + D * j][q];
        }
        // APIC scatters TPIC's form with B·D⁻¹ in place of ∇v
        if (M.apic) apic_affine<D, T>(g, w, xp, P, q, gv);
    }
    const T gz = T(c->g[D - 1]);
    constexpr bool tpic = AFF;
    const T mgz = ms * gz;
    // 3D standard transfer into the LDS box, by stencil rows: with
    // so = -Ω σ and N = wxy wz, ∂N = (dxy wz, xdy wz, wxy dwz),
    //   m N = (m wxy) wz,  p_i N = (p_i wxy) wz,
    //   f_x = (so_xx dxy + so_xy xdy) wz + (so_xz wxy) dwz, likewise y,
    //   f_z = (so_xz dxy + so_yz xdy + m g_z wxy) wz + (so_zz wxy) dwz,
    // the same sums as the per-node form below up to rounding.  FP64 p2g
    // 1138 -> 1105 us; FP32 p2g 799 -> 754 us (once the box's min/max no
    // longer cost a scalar loop; before, the extra VGPRs of the row factors
    // had made FP32 slower, 949 -> 1116 us).
    constexpr bool rows = D == 3 && !AFF;
    auto scatter_rows = [&](const LdsSink<D, NV> &sink) {
        if constexpr (rows) {
            T so[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) so[k] = -(Om * s[k]);
            ZNodes<T> z;
            box_rows_3d<T, NV>(*sink.b, w, z, [&](int row, T wxy, T dxy, T xdy) {
                const T mw = ms * wxy, p0 = pm[0] * wxy, p1 = pm[1] * wxy, p2 = pm[2] * wxy;
                const T ax = so[0] * dxy + so[5] * xdy, bx = so[4] * wxy;
                const T ay = so[5] * dxy + so[1] * xdy, by = so[3] * wxy;
                const T az = so[4] * dxy + so[3] * xdy + mgz * wxy, bz = so[2] * wxy;
#pragma unroll
                for (int j0 = 0; j0 < 4; j0++) {
                    const int li = row + z.a[j0];
                    const T wz = z.w[j0], dwz = z.dw[j0];
                    sink(0, 0, li, double(mw * wz));
                    sink(1, 0, li, double(p0 * wz));
                    sink(2, 0, li, double(p1 * wz));
                    sink(3, 0, li, double(p2 * wz));
                    sink(4, 0, li, double(ax * wz + bx * dwz));
                    sink(5, 0, li, double(ay * wz + by * dwz));
                    sink(6, 0, li, double(az * wz + bz * dwz));
                }
            });
        }
    };
    auto scatter = [&](const auto &sink) {
    scatter_walk<D, T>(g, w, sink, [&](long no, int li, T N, const T *dN, const int *ix) {
        sink(0, no, li, double(N * ms));
        if (D == 2) {
            if constexpr (tpic) {
                // Δnp = x_n - x_p (misc.jl:1-14)
                const T dx = T(g.X[0][clampi(ix[0], g.nno[0])] - xp[0]);
                const T dy = T(g.X[1][clampi(ix[1], g.nno[1])] - xp[1]);
                sink(1, no, li, double(N * ms * (v[0] + gv[0][0] * dx + gv[0][1] * dy)));
                sink(2, no, li, double(N * ms * (v[1] + gv[1][0] * dx + gv[1][1] * dy)));
            } else {
                sink(1, no, li, double(N * pm[0]));
                sink(2, no, li, double(N * pm[1]));
            }
            sink(3, no, li, double(-(Om * (dN[0] * s[0] + dN[1] * s[2]))));
            sink(4, no, li, double(-(Om * (dN[0] * s[2] + dN[1] * s[1]) - N * mgz)));
        } else {
            if constexpr (tpic) {
                const T dx = T(g.X[0][clampi(ix[0], g.nno[0])] - xp[0]);
                const T dy = T(g.X[1][clampi(ix[1], g.nno[1])] - xp[1]);
                const T dz = T(g.X[2][clampi(ix[2], g.nno[2])] - xp[2]);
                sink(1, no, li, double(N * ms * (v[0] + gv[0][0] * dx + gv[0][1] * dy + gv[0][2] * dz)));
                sink(2, no, li, double(N * ms * (v[1] + gv[1][0] * dx + gv[1][1] * dy + gv[1][2] * dz)));
                sink(3, no, li, double(N * ms * (v[2] + gv[2][0] * dx + gv[2][1] * dy + gv[2][2] * dz)));
            } else {
                sink(1, no, li, double(N * pm[0]));
                sink(2, no, li, double(N * pm[1]));
                sink(3, no, li, double(N * pm[2]));
            }
            // Voigt [xx,yy,zz,yz,xz,xy]: σyx = s[5], σzy = s[3], σzx = s[4]
            sink(4, no, li, double(-(Om * (dN[0] * s[0] + dN[1] * s[5] + dN[2] * s[4]))));
            sink(5, no, li, double(-(Om * (dN[0] * s[5] + dN[1] * s[1] + dN[2] * s[3]))));
            sink(6, no, li, double(-(Om * (dN[0] * s[4] + dN[1] * s[3] + dN[2] * s[2]) - N * mgz)));
        }
    });
    };
    // When the block's particles are spread too far for one box, one pass
    // per wave, each with the box of that wave's particles; a wave whose box
    // still does not fit scatters with global atomics.
    const int wave = threadIdx.x >> 6;
    const int npass = lds ? 1 : (blockDim.x + 63) >> 6;
    for (int pass = 0; pass < npass; pass++) {
        const bool mine = lds || wave == pass;
        if (!lds) box.begin((EP_LDS double *)smem, sint, live && mine, w.base, ACC_LDS_P2G);
        const bool sl = box.fits(NV);
        if (sl) box.zero(NV);
        if (mine) {
            if (!live) w.retire(box.lo);
            if (rows && sl) scatter_rows(LdsSink<D, NV>{&box});
            else if (sl) scatter(LdsSink<D, NV>{&box});
            else scatter(GlobSink{dst});
        }
        if (sl) box.template flush<NV>(g, dst);
    }
}

// waves per SIMD the register allocation of k_p2g aims at: 2 is what
// 512-thread blocks imply; FP32 fits 4 (at most 128 VGPRs) without spilling
// and runs 8 % faster for it (FP64 would spill)
template <typename T> constexpr int kP2gWaves = std::is_same<T, float>::value ? 4 : 2;

// kernel of one context: block b of the launch is block b of the particles / nodes
template <int D, int B, typename T, bool AFF>
__global__ void __launch_bounds__(EP_PBLK_MAX) EP_WAVES_PER_EU(kP2gWaves<T>) k_p2g(const Ctrl *c, Geom g, PState<T> P, GState G, Mat M) { b_p2g<D, B, T, AFF>(c, g, P, G, M, blockIdx.x); }

// euler (solve.jl:14-40); a and v of massless nodes are the zeros that
// solve() writes first (solve.jl:58-59).
template <int D>
__device__ __forceinline__ void b_euler(const Ctrl *c, const GState &G, const Mat &M, long blk) {
    const long no = blk * long(blockDim.x) + threadIdx.x;
    if (no >= G.nno || !c->active) return;
    const double dt = c->dt, eta = M.damping;
    const double m = G.m[no];
    double o[D], pp[D];
#pragma unroll
    for (int d = 0; d < D; d++) { o[d] = G.f[d][no]; pp[d] = G.p[d][no]; }
    double a[D], v[D];
#pragma unroll
    for (int d = 0; d < D; d++) { a[d] = 0.0; v[d] = 0.0; }
    if (m != 0.0) {
        const unsigned bc = G.bc[no];
#pragma unroll
        for (int d = 0; d < D; d++) {
            if (bc >> d & 1u) {
                a[d] = 0.0;
                v[d] = 0.0;
            } else {
                const double mi = 1.0 / m;
                double nrm = o[0] * o[0];
#pragma unroll
                for (int k = 1; k < D; k++) nrm += o[k] * o[k];
                nrm = sqrt(nrm);
                const double Dm = eta * nrm * jsign(pp[d] * mi);
                if (fabs(pp[d] * mi) >= 1.0e-3) o[d] = o[d] - Dm;
                a[d] = o[d] * mi;
                v[d] = (pp[d] + dt * o[d]) * mi;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < D; d++) { G.f[d][no] = o[d]; G.a[d][no] = a[d]; G.v[d][no] = v[d]; }
}

// kernel of one context: block b of the launch is block b of the particles / nodes
template <int D>
__global__ void k_euler(const Ctrl *c, GState G, Mat M) { b_euler<D>(c, G, M, blockIdx.x); }

// picflip_n2p + augm_momentum (n2p.jl:14-40, augm.jl:13-25)
template <int D, int B, typename T>
__device__ __forceinline__ void b_g2p(Ctrl *c, const Geom &g, const PState<T> &P, const GState &G, const Mat &M, long blk) {
    __shared__ double smem[ACC_LDS_G2P];
    __shared__ int sint[8];
    __shared__ unsigned long long sred[16 * D];
    if (!c->active) return;
    const long p = blk * long(blockDim.x) + threadIdx.x;
    const long q = p < P.n ? p : 0;
    const bool live = p < P.n && P.Om0[q] != T(0);   // Ω₀ = 0: a padding slot
    const int par = c->parity;
    double xp[D];
    T lp[D];
#pragma unroll
    for (int i = 0; i < D; i++) { xp[i] = P.x[par][i][q]; lp[i] = B == 1 ? P.l[i][q] : T(0); }
    W<D, T> w;
    weights<D, B, T>(g, xp, lp, w);
    BlockBox<D> box;
    box.begin((EP_LDS double *)smem, sint, live, w.base, ACC_LDS_G2P);
    if (!live) w.retire(box.lo);
    const bool staged = box.fits(2 * D);
    if (staged) {
        const double *src[2 * D];
#pragma unroll
        for (int d = 0; d < D; d++) { src[d] = G.v[d]; src[D + d] = G.a[d]; }
        box.template stage<2 * D, T>(g, src);
    }
    T vpic[D], aflip[D];
#pragma unroll
    for (int d = 0; d < D; d++) { vpic[d] = T(0); aflip[d] = T(0); }
    const double *gsrc[2 * D];
#pragma unroll
    for (int d = 0; d < D; d++) { gsrc[d] = G.v[d]; gsrc[D + d] = G.a[d]; }
    auto gather = [&](const auto &src) {
        gather_walk<D, T>(g, w, src, [&](long no, int li, T N, const T *, const int *) {
#pragma unroll
            for (int d = 0; d < D; d++) {
                vpic[d] += N * T(src(d, no, li));
                aflip[d] += N * T(src(D + d, no, li));
            }
        });
    };
    if (staged) gather(LdsSrc<D, 2 * D, T>{&box});
    else gather(GlobSrc{gsrc});
    const T dt = T(c->dt), C = T(M.C_pf);
    T vn[D];
    unsigned long long bm[D];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < D; d++) {
        vn[d] = C * (P.v[d][q] + dt * aflip[d]) + (T(1) - C) * vpic[d];
        const double xn = xp[d] + double(dt * vpic[d]);
        if (live) {
            P.v[d][p] = vn[d];
            P.x[par ^ 1][d][p] = xn;
            if (!isfinite(vn[d]) || !isfinite(xn)) bad = true;
        }
        bm[d] = live ? (unsigned long long)__double_as_longlong(double(fabs(vn[d]))) : 0ull;
    }
    block_max_bits_n<D>(bm, sred);
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < D; d++) c->vmax_part[d * c->nblk + blk] = bm[d];
    if (bad) atomicOr(&c->err, ERR_NONFINITE);
    if (M.musl) {
        const bool lds = box.fits(D);
        if (lds) box.zero(D);
        double *dst[D];
#pragma unroll
        for (int d = 0; d < D; d++) dst[d] = G.pa[d];
        const T ms = mass_of(P, M, q, P.J[q]);
        auto scatter = [&](const auto &sink) {
            scatter_walk<D, T>(g, w, sink, [&](long no, int li, T N, const T *, const int *) {
#pragma unroll
                for (int d = 0; d < D; d++) sink(d, no, li, double(N * (ms * vn[d])));
            });
        };
        if (lds) scatter(LdsSink<D, D>{&box});
        else scatter(GlobSink{dst});
        if (lds) box.template flush<D>(g, dst);
    }
}

// kernel of one context: block b of the launch is block b of the particles / nodes
template <int D, int B, typename T>
__global__ void __launch_bounds__(EP_PBLK_MAX) k_g2p(Ctrl *c, Geom g, PState<T> P, GState G, Mat M) { b_g2p<D, B, T>(c, g, P, G, M, blockIdx.x); }

// Bij (Bij.jl:14-26, n2p.jl:59-62): APIC's Bᵢⱼ = Σ N v_n Δnpᵀ from the
// Euler grid velocity at the position before the PIC/FLIP update (the other
// position buffer), between G2P and the mUSL velocity; without mUSL
// n2p.jl:72-74 computes the same values a second time.
template <int D, int B, typename T>
__device__ __forceinline__ void b_bij(const Ctrl *c, const Geom &g, const PState<T> &P, const GState &G, long blk) {
    __shared__ double smem[ACC_LDS];
    __shared__ int sint[8];
    if (!c->active) return;
    const long p = blk * long(blockDim.x) + threadIdx.x;
    const long q = p < P.n ? p : 0;
    const bool live = p < P.n && P.Om0[q] != T(0);
    const int par = c->parity;   // G2P wrote x into par ^ 1
    double xp[D];
    T lp[D];
#pragma unroll
    for (int i = 0; i < D; i++) { xp[i] = P.x[par][i][q]; lp[i] = B == 1 ? P.l[i][q] : T(0); }
    W<D, T> w;
    weights<D, B, T>(g, xp, lp, w);
    BlockBox<D> box;
    box.begin((EP_LDS double *)smem, sint, live, w.base);
    if (!live) w.retire(box.lo);
    const bool staged = box.fits(D);
    const double *src[D];
#pragma unroll
    for (int d = 0; d < D; d++) src[d] = G.v[d];
    if (staged) box.template stage<D, T>(g, src);
    T Bacc[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) Bacc[i][j] = T(0);
    auto gather = [&](const auto &s) {
        gather_walk<D, T>(g, w, s, [&](long no, int li, T N, const T *, const int *ix) {
            T dl[D];
#pragma unroll
            for (int j = 0; j < D; j++) dl[j] = T(g.X[j][clampi(ix[j], g.nno[j])] - xp[j]);
#pragma unroll
            for (int i = 0; i < D; i++) {
                const T vi = T(s(i, no, li));
#pragma unroll
                for (int j = 0; j < D; j++) Bacc[i][j] += N * (vi * dl[j]);
            }
        });
    };
    if (staged) gather(LdsSrc<D, D, T>{&box});
    else gather(GlobSrc{src});
    if (live) {
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) P.Bm[i + D * j][p] = Bacc[i][j];
    }
}

template <int D, int B, typename T>
__global__ void __launch_bounds__(EP_PBLK_MAX) k_bij(const Ctrl *c, Geom g, PState<T> P, GState G) { b_bij<D, B, T>(c, g, P, G, blockIdx.x); }

// Sort key of a cell (e0, e1[, e2]) in brick-major order: bricks of b0 x b1
// (x b2) cells; inside a brick cells run z fastest, as the grid.  Bricks
// follow a snake: along z within a column of bricks, the direction reversing
// from one column to the next, the columns along x reversing from one y row
// to the next, so consecutive bricks always share a face and a block of
// particles that straddles two bricks still touches a compact node box.
__host__ __device__ inline long cell_key(int D, const long *e, long b0, long b1, long b2, const int *nel) {
    if (D == 2) {
        const long nbz = (nel[1] + b1 - 1) / b1;
        const long bx = e[0] / b0;
        long bz = e[1] / b1;
        if (bx & 1) bz = nbz - 1 - bz;
        const long brick = bz + nbz * bx;
        return brick * (b0 * b1) + (e[1] % b1) + b1 * (e[0] % b0);
    }
    const long nbz = (nel[2] + b2 - 1) / b2, nbx = (nel[0] + b0 - 1) / b0;
    const long by = e[1] / b1;
    long bx = e[0] / b0, bz = e[2] / b2;
    if (by & 1) bx = nbx - 1 - bx;
    const long col = bx + nbx * by;
    if (col & 1) bz = nbz - 1 - bz;
    const long brick = bz + nbz * col;
    return brick * (b0 * b1 * b2) + (e[2] % b2) + b2 * ((e[0] % b0) + b0 * (e[1] % b1));
}

// Re-sort.  Particles drift out of the brick-major cell order the upload
// gives them, and each block's node box grows with the spread of its
// particles.  A re-sort recomputes every slot's key from its current
// position (the upload's order: bricks of b0 x b1 x b2 cells, cells inside a
// brick), sorts the slots stably by key (padding last) and gathers every
// particle field through the sorted slot list.
template <int D, typename T>
__global__ void __launch_bounds__(256) k_sort_keys(const Ctrl *c, Geom g, PState<T> P, long nslot, int b0, int b1, int b2,
                                                   unsigned *key, int *val) {
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (p >= nslot) return;
    unsigned k = 0xFFFFFFFFu;
    if (P.Om0[p] != T(0)) {
        long e[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < D; i++) {
            double q = exact_fld(P.x[c->parity][i][p] - g.x0[i], g.h[i], g.ih[i]);
            if (!(q >= 0.0)) q = 0.0;
            if (q > double(g.nel[i] - 1)) q = double(g.nel[i] - 1);
            e[i] = long(q);
        }
        const long kk = cell_key(D, e, b0, b1, b2, g.nel);
        k = unsigned(kk);
    }
    key[p] = k;
    val[p] = int(p);
}

// the friction angle's trigonometric values the Drucker-Prager return reads,
// once per upload (φ does not change during a run)
template <int NSTR, typename T>
__global__ void __launch_bounds__(256) k_friction(const T *phi, T *sphi, T *cphi, long n) {
    const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    T sf, cf;
    friction_trig<NSTR, T>(phi[i], sf, cf);
    sphi[i] = sf;
    cphi[i] = cf;
}

// up to kPermK particle fields of one element type gathered into sorted
// order by one launch: dst[k][i] = src[k][idx[i]]
constexpr int kPermK = 6;
template <typename E>
struct PermArgs {
    const E *src[kPermK];
    E *dst[kPermK];
    int cnt;
};

template <typename E>
__global__ void __launch_bounds__(256) k_permute(PermArgs<E> a, const int *idx, long n) {
    const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const long j = idx[i];
#pragma unroll
    for (int k = 0; k < kPermK; k++)
        if (k < a.cnt) a.dst[k][i] = a.src[k][j];
}

// augm_velocity (augm.jl:37-53) after mesh.v .= 0 (n2p.jl:67)
template <int D>
__device__ __forceinline__ void b_augv(const Ctrl *c, const GState &G, long blk) {
    const long no = blk * long(blockDim.x) + threadIdx.x;
    if (no >= G.nno || !c->active) return;
    const double m = G.m[no];
    const unsigned bc = G.bc[no];
#pragma unroll
    for (int d = 0; d < D; d++) {
        double v = 0.0;
        if (m != 0.0 && !(bc >> d & 1u)) v = G.pa[d][no] * (1.0 / m);
        G.v[d][no] = v;
    }
}

// kernel of one context: block b of the launch is block b of the particles / nodes
template <int D>
__global__ void k_augv(const Ctrl *c, GState G) { b_augv<D>(c, G, blockIdx.x); }

// The FP32 path carries the kinematics as deviations from the identity so
// that small strains keep their relative precision: K = ΔF - I instead of
// ΔF, δJ = ΔJ - 1, B = b - I, ε = ½ log1p(eig B), b_new - I = expm1(2ε).
template <typename T> constexpr bool kDev = std::is_same<T, float>::value;

// det(I + A) - 1 expanded so no 1 + small is rounded first
template <int D, typename T>
__device__ __forceinline__ T det1m(const T (&A)[D][D]) {
    if (D == 2) return A[0][0] + A[1][1] + (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
    const T tr = A[0][0] + A[1][1] + A[D - 1][D - 1];
    const T m2 = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) + (A[0][0] * A[D - 1][D - 1] - A[0][D - 1] * A[D - 1][0]) +
                 (A[1][1] * A[D - 1][D - 1] - A[1][D - 1] * A[D - 1][1]);
    const T d3 = A[0][0] * (A[1][1] * A[D - 1][D - 1] - A[1][D - 1] * A[D - 1][1]) -
                 A[0][1] * (A[1][0] * A[D - 1][D - 1] - A[1][D - 1] * A[D - 1][0]) +
                 A[0][D - 1] * (A[1][0] * A[D - 1][1] - A[1][1] * A[D - 1][0]);
    return tr + m2 + d3;
}

// b <- ΔF b ΔF' (elast.jl:49) on the stored symmetric b
template <int D, typename T>
__device__ __forceinline__ void push_b(const PState<T> &P, long p, const T (&dF)[D][D], T (&bn)[D][D]) {
    T b[D][D], t1[D][D];
    if (D == 2) {
        b[0][0] = P.b[0][p]; b[1][1] = P.b[1][p]; b[0][1] = b[1][0] = P.b[2][p];
    } else {
        b[0][0] = P.b[0][p]; b[1][1] = P.b[1][p]; b[D - 1][D - 1] = P.b[D - 1][p];
        b[1][D - 1] = b[D - 1][1] = P.b[3][p]; b[0][D - 1] = b[D - 1][0] = P.b[4][p]; b[0][1] = b[1][0] = P.b[5][p];
    }
    if constexpr (kDev<T>) {
        // B' = K + K' + B + K B + B K' + K K' + K B K'   (K = ΔF - I)
        T KB[D][D];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) {
                T a = dF[i][0] * b[0][j];
#pragma unroll
                for (int k = 1; k < D; k++) a += dF[i][k] * b[k][j];
                KB[i][j] = a + b[i][j] + dF[i][j];      // K B + B + K
            }
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = i; j < D; j++) {
                T a = KB[i][0] * dF[j][0];
#pragma unroll
                for (int k = 1; k < D; k++) a += KB[i][k] * dF[j][k];   // (K B + B + K) K'
                const T v = KB[i][j] + dF[j][i] + a;
                bn[i][j] = v;
                bn[j][i] = v;
            }
        return;
    }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            T a = dF[i][0] * b[0][j];
#pragma unroll
            for (int k = 1; k < D; k++) a += dF[i][k] * b[k][j];
            t1[i][j] = a;
        }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = i; j < D; j++) {
            T a = t1[i][0] * dF[j][0];
#pragma unroll
            for (int k = 1; k < D; k++) a += t1[i][k] * dF[j][k];
            bn[i][j] = a;
            bn[j][i] = a;
        }
}

template <int D, typename T>
__device__ __forceinline__ void store_sym(T *const *dst, long p, const T (&m)[D][D]) {
    if (D == 2) { dst[0][p] = m[0][0]; dst[1][p] = m[1][1]; dst[2][p] = m[0][1]; }
    else {
        dst[0][p] = m[0][0]; dst[1][p] = m[1][1]; dst[2][p] = m[D - 1][D - 1];
        dst[3][p] = m[1][D - 1]; dst[4][p] = m[0][D - 1]; dst[5][p] = m[0][1];
    }
}

// M = V diag(f) V'
template <int D, typename T>
__device__ __forceinline__ void vdvt(const T (&V)[D][D], const T *f, T (&m)[D][D]) {
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = i; j < D; j++) {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < D; k++) acc += (V[i][k] * f[k]) * V[j][k];
            m[i][j] = acc;
            m[j][i] = acc;
        }
}

template <int D, typename T>
__device__ __forceinline__ void sym_to_voigt(const T (&m)[D][D], T *s) {
    if (D == 2) { s[0] = m[0][0]; s[1] = m[1][1]; s[2] = m[0][1]; }
    else { s[0] = m[0][0]; s[1] = m[1][1]; s[2] = m[D - 1][D - 1]; s[3] = m[1][D - 1]; s[4] = m[0][D - 1]; s[5] = m[0][1]; }
}

// Elastic predictor + plastic corrector on one particle (elast.jl:45-80,
// plast.jl:49-65, DP.jl:33-99, J2.jl:21-75); dF is the (F-bar scaled) ΔF,
// or ΔF - I on the FP32 path.
//
// Finite strain with Drucker-Prager (or no plasticity) runs in the principal
// frame of the trial b: ε = ½ log b, τ = Del·ε and the return map
// (isotropic Del, ψ = 0) all keep b's eigenvectors, so one eigen-
// decomposition gives ε and τ (elast.jl:51-54), the DP return acts on the
// principal stresses (its invariants are frame-independent, DP.jl:1-12) and
// the recomputed b = exp(2 tensor(Del\τ, ½)) (DP.jl:63-66) is
// V diag(exp 2ε_new) V' with the same V.  The reference's second eigen-
// decomposition of ε_new yields the same matrix up to rounding.  The J2
// cutting plane weights shear with Del's G against the 2·s² of J₂, which is
// not coaxial, so J2 keeps the reference's Voigt-space iteration.
template <int D, typename T, bool SURR = false>
__device__ __forceinline__ void constitutive(Ctrl *c, const PState<T> &P, const Mat &M, long p, const T (&dF)[D][D]) {
    constexpr int NS = D == 2 ? 3 : 6;
    T s[NS];
    bool bad = false;
    if (M.finite) {
        T bn[D][D], a[D][D], V[D][D];
        push_b<D, T>(P, p, dF, bn);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) a[i][j] = bn[i][j];
        sym_eig<D, T, !kDev<T>>(a, V);
        if constexpr (std::is_same<T, float>::value) {
            // FP64 island of the FP32 path: near-repeated eigenvalues of
            // B = b - I are decomposed again in FP64.
            T mx = T(0), gap = T(1e30f);
#pragma unroll
            for (int i = 0; i < D; i++) {
                mx = fmax(mx, fabs(a[i][i]));
#pragma unroll
                for (int j = i + 1; j < D; j++) gap = fmin(gap, fabs(a[i][i] - a[j][j]));
            }
            if (gap < T(1e-4f) * mx) {
                double ad[D][D], Vd[D][D];
#pragma unroll
                for (int i = 0; i < D; i++)
#pragma unroll
                    for (int j = 0; j < D; j++) ad[i][j] = double(bn[i][j]);
                sym_eig<D, double>(ad, Vd);
#pragma unroll
                for (int i = 0; i < D; i++)
#pragma unroll
                    for (int j = 0; j < D; j++) { a[i][j] = T(ad[i][j]); V[i][j] = T(Vd[i][j]); }
            }
        }
        T ep[D], sp[NS];
#pragma unroll
        for (int k = 0; k < D; k++) ep[k] = kDev<T> ? T(0.5) * log1p(a[k][k]) : T(0.5) * log(a[k][k]);
        // principal Kirchhoff stress: Del on [ε1, ε2(, ε3), 0...]; the shear
        // rows of the isotropic Del (ep_create checks it) give exact zeros
#pragma unroll
        for (int i = 0; i < NS; i++) {
            T acc = T(0);
            if (i < D) {
#pragma unroll
                for (int j = 0; j < D; j++) acc += T(M.Del[i + NS * j]) * ep[j];
            }
            sp[i] = acc;
        }
        const bool plast = c->with_plast && M.plast != 0;
        if constexpr (SURR) {
            // Plastic trial states go to the learned return (k_surr), which
            // writes τ, b, ϵpII and Δλ; the queue slots are reserved with one
            // atomic per wave.
            const bool pl = plast && M.plast == 1 &&
                            dp_trial_trig<NS, T>(sp, P.c0[p], P.cr[p], P.sphi[p], P.cphi[p], T(M.Hp), P.epII[p]) != 0;
            const unsigned long long mask = __ballot(pl);
            if (mask) {
                const int lane = int(__lane_id());
                const int leader = __ffsll(mask) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&c->qcount, __popcll(mask));
                base = __shfl(base, leader);
                if (pl) {
                    const int slot = base + __popcll(mask & ((1ull << lane) - 1ull));
                    int o[D];
#pragma unroll
                    for (int k = 0; k < D; k++) o[k] = k;
                    // principal values sorted descending, eigenvectors alike
#pragma unroll
                    for (int a0 = 0; a0 < D; a0++)
#pragma unroll
                        for (int b0 = 0; b0 + 1 < D - a0; b0++)
                            if (ep[o[b0]] < ep[o[b0 + 1]]) { const int t = o[b0]; o[b0] = o[b0 + 1]; o[b0 + 1] = t; }
                    M.q.idx[slot] = int(p);
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        M.q.e[k][slot] = float(ep[o[k]]);
#pragma unroll
                        for (int i = 0; i < D; i++) M.q.V[i + 3 * k][slot] = float(V[i][o[k]]);
                    }
                    return;
                }
            }
        }
        if (plast && M.plast == 1) {
            T e1 = P.epII[p], e2 = M.nonlocal ? P.epII2[p] : e1, dl;
            const int br = dp_return_trig<NS, T>(sp, P.c0[p], P.cr[p], P.sphi[p], P.cphi[p], T(M.Hp), T(M.Kc),
                                                 T(M.Gc), e2, e1, dl);
            P.dlam[p] = dl;
            P.epII[p] = e1;
            // b from the returned stress, exp(2 Del⁻¹τ) (DP.jl:63-66).  An
            // elastic trial keeps the trial b, which that formula gives back
            // up to rounding (Del⁻¹ Del ε = ε, exp 2ε = the eigenvalues of b)
            // at the cost of a 6x6 solve and d exponentials per particle
            if (br != 0) {
                T y[NS];
                lu_solve<NS, T>(M.lu, sp, y);
                T fe[D];
#pragma unroll
                for (int k = 0; k < D; k++) fe[k] = kDev<T> ? expm1(T(2) * y[k]) : exp(T(2) * y[k]);
                vdvt<D, T>(V, fe, bn);
            }
        }
        T tp[D];
#pragma unroll
        for (int k = 0; k < D; k++) tp[k] = sp[k];
        T tm[D][D];
        vdvt<D, T>(V, tp, tm);
        sym_to_voigt<D, T>(tm, s);
        if (plast && M.plast == 2) {
            T e1 = P.epII[p], e2 = M.nonlocal ? P.epII2[p] : e1;
            const int it = j2_return<NS, T>(s, P.c0[p], P.cr[p], T(M.Hp), M.Del, e2, e1);
            if (it != 0) {
                P.epII[p] = e1;
                if (it < 0) atomicAdd(&c->j2cap, 1);
                // ε = tensor(Del \ τ, ½); b = exp(2ε) (J2.jl:42-45)
                T y[NS];
                lu_solve<NS, T>(M.lu, s, y);
                T e[D][D], W2[D][D];
                if (D == 2) { e[0][0] = y[0]; e[1][1] = y[1]; e[0][1] = e[1][0] = T(0.5) * y[2]; }
                else {
                    e[0][0] = y[0]; e[1][1] = y[1]; e[D - 1][D - 1] = y[D - 1];
                    e[1][D - 1] = e[D - 1][1] = T(0.5) * y[3]; e[0][D - 1] = e[D - 1][0] = T(0.5) * y[4];
                    e[0][1] = e[1][0] = T(0.5) * y[5];
                }
                sym_eig<D, T>(e, W2);
                T fe[D];
#pragma unroll
                for (int k = 0; k < D; k++) fe[k] = kDev<T> ? expm1(T(2) * e[k][k]) : exp(T(2) * e[k][k]);
                vdvt<D, T>(W2, fe, bn);
            }
        }
        store_sym<D, T>(P.b, p, bn);
    } else {
        // infinitesimal_elast (elast.jl:69-80); the Jaumann term σω'+σ'ω is
        // identically zero for symmetric σ
        T ev[NS];
        T e[D][D];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++)
                e[i][j] = kDev<T> ? T(0.5) * (dF[i][j] + dF[j][i])
                                  : T(0.5) * (dF[i][j] + dF[j][i]) - (i == j ? T(1) : T(0));
        if (D == 2) { ev[0] = e[0][0]; ev[1] = e[1][1]; ev[2] = T(2) * e[0][1]; }
        else { ev[0] = e[0][0]; ev[1] = e[1][1]; ev[2] = e[D - 1][D - 1]; ev[3] = T(2) * e[1][D - 1]; ev[4] = T(2) * e[0][D - 1]; ev[5] = T(2) * e[0][1]; }
#pragma unroll
        for (int i = 0; i < NS; i++) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < NS; j++) acc += T(M.Del[i + NS * j]) * ev[j];
            s[i] = P.s[i][p] + acc;
        }
        if (c->with_plast && M.plast == 1) {
            T e1 = P.epII[p], e2 = M.nonlocal ? P.epII2[p] : e1, dl;
            dp_return_trig<NS, T>(s, P.c0[p], P.cr[p], P.sphi[p], P.cphi[p], T(M.Hp), T(M.Kc), T(M.Gc), e2, e1, dl);
            P.dlam[p] = dl;
            P.epII[p] = e1;
        } else if (c->with_plast && M.plast == 2) {
            T e1 = P.epII[p], e2 = M.nonlocal ? P.epII2[p] : e1;
            const int it = j2_return<NS, T>(s, P.c0[p], P.cr[p], T(M.Hp), M.Del, e2, e1);
            if (it != 0) {
                P.epII[p] = e1;
                if (it < 0) atomicAdd(&c->j2cap, 1);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NS; k++) {
        P.s[k][p] = s[k];
        if (!isfinite(s[k])) bad = true;
    }
    if (bad) atomicOr(&c->err, ERR_NONFINITE);
}

// ∇v gather (deform.jl:5-12) from a gather source, in the reference order
template <int D, typename T, typename Src>
__device__ __forceinline__ void grad_v(const Geom &g, const W<D, T> &w, const Src &src, int v0, T (&gv)[D][D]) {
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) gv[i][j] = T(0);
    gather_walk<D, T>(g, w, src, [&](long no, int li, T, const T *dN, const int *) {
        T vn[D];
#pragma unroll
        for (int i = 0; i < D; i++) vn[i] = T(src(v0 + i, no, li));
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) gv[i][j] += dN[j] * vn[i];
    });
}

// ∇v gather from a staged 3D box by stencil rows (y outer, x, z inner, the
// reference's node order): with N = wxy wz,
//   Σ v_n ∂N/∂x = Σ_rows dxy Σ_z wz v_n,  ∂/∂y: xdy Σ_z wz v_n,
//   ∂/∂z: wxy Σ_z dwz v_n,
// six products per node instead of twelve; the sums of grad_v up to
// rounding.  With Q, also Σ N q_n of box field 3 (F-bar's ΔJ_n / m_n).
template <typename T, int NV, bool Q>
__device__ __forceinline__ T grad_v_rows(const BlockBox<3> &b, const W<3, T> &w, T (&gv)[3][3]) {
    const int zb = w.base[2] - b.lo[2], xb = w.base[0] - b.lo[0], yb = w.base[1] - b.lo[1];
    int az[4];
#pragma unroll
    for (int j = 0; j < 4; j++) az[j] = 16 * NV * ((zb + j) >> 2) + ((zb + j) & 3);
    const int sy = 16 * NV * b.nz4 * b.nx4;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) gv[i][j] = T(0);
    T q = T(0);
    const EP_LDS T *bb = b.template as<T>();   // staged in the particle precision
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const T wxy = w.w[0][i] * w.w[1][k], dxy = w.dw[0][i] * w.w[1][k], xdy = w.w[0][i] * w.dw[1][k];
            const int x = xb + i;
            const int row = sy * (yb + k) + 16 * NV * b.nz4 * (x >> 2) + 4 * (x & 3);
            T s[3] = {T(0), T(0), T(0)}, t[3] = {T(0), T(0), T(0)}, sq = T(0);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int li = row + az[j];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T vn = bb[li + 16 * c];
                    s[c] += w.w[2][j] * vn;
                    t[c] += w.dw[2][j] * vn;
                }
                if (Q) sq += w.w[2][j] * bb[li + 48];
            }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                gv[c][0] += dxy * s[c];
                gv[c][1] += xdy * s[c];
                gv[c][2] += wxy * t[c];
            }
            if (Q) q += wxy * sq;
        }
    }
    return q;
}

// ΔF = I + dt ∇v, F_new = ΔF F (deform.jl:14-15)
template <int D, typename T>
__device__ __forceinline__ void delta_f(T dt, const T (&gv)[D][D], T (&dF)[D][D]) {
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++)
            // FP32 path: dF holds K = dt ∇v = ΔF - I
            dF[i][j] = kDev<T> ? dt * gv[i][j] : (i == j ? T(1) : T(0)) + dt * gv[i][j];
}

// Fn = ΔF F (FP32: F + K F)
template <int D, typename T>
__device__ __forceinline__ void push_f(const PState<T> &P, long q, const T (&dF)[D][D], T (&Fn)[D][D]) {
    T F[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) F[i][j] = P.F[i + D * j][q];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            T a = dF[i][0] * F[0][j];
#pragma unroll
            for (int k = 1; k < D; k++) a += dF[i][k] * F[k][j];
            Fn[i][j] = kDev<T> ? F[i][j] + a : a;
        }
}

// GIMP domain update (domain.jl:1-52) chosen by basis.how per framework
// (update.jl:7-29), after the deformation update and before F-bar
// (update.jl:43-49).  Restated from intent: domain.jl reads mpts.Fᵢⱼ and
// mpts.ΔFᵢⱼ, which are fields of mpts.s.  dF is ΔF (FP32 path: ΔF - I), Fn
// the updated F and J = det Fn; U = n diag(√λ) nᵀ from eigen(AᵀA).
enum { DOM_UNDEFORMED = 0, DOM_DETF = 1, DOM_FII = 2, DOM_UII = 3, DOM_DETDF = 4, DOM_DFII = 5, DOM_DUII = 6 };
template <int D, typename T>
__device__ __forceinline__ void domain_update(const PState<T> &P, const Mat &M, long p, const T (&dF)[D][D],
                                              const T (&Fn)[D][D], T J) {
    const int how = M.domain;
    if (how == DOM_UNDEFORMED) return;       // ℓ stays ℓ₀
    const bool inc = how >= DOM_DETDF;       // infinitesimal: scales the current ℓ
    T A[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++)
            A[i][j] = inc ? ((kDev<T> && i == j) ? T(1) + dF[i][j] : dF[i][j]) : Fn[i][j];
    T f[D];
    if (how == DOM_DETF || how == DOM_DETDF) {
        const T d = how == DOM_DETF ? J : (kDev<T> ? T(1) + det1m<D, T>(dF) : det<D, T>(dF));
#pragma unroll
        for (int i = 0; i < D; i++) f[i] = d;
    } else if (how == DOM_FII || how == DOM_DFII) {
#pragma unroll
        for (int i = 0; i < D; i++) f[i] = A[i][i];
    } else {
        T Cm[D][D], V[D][D];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) {
                T s = A[0][i] * A[0][j];
#pragma unroll
                for (int k = 1; k < D; k++) s += A[k][i] * A[k][j];
                Cm[i][j] = s;
            }
        sym_eig<D, T>(Cm, V);
#pragma unroll
        for (int i = 0; i < D; i++) {
            T u = T(0);
#pragma unroll
            for (int k = 0; k < D; k++) u += V[i][k] * sqrt(Cm[k][k]) * V[i][k];
            f[i] = u;
        }
    }
#pragma unroll
    for (int i = 0; i < D; i++) P.l[i][p] = f[i] * (inc ? P.l[i][p] : P.l0[i][p]);
}

// finite/infinitesimal_deform (deform.jl:1-45).
// LOCK = false: the whole particle update (deform + elast + plast) in one
//   pass.
// LOCK = true: only the ΔJn scatter of F-bar (volumetric.jl:1-11), which
//   needs ΔJ = det ΔF and ρΩ = ρ₀/J·JΩ₀ of the updated J; nothing is
//   written back, k_fbar recomputes ΔF from the same grid velocity.
template <int D, int B, typename T, bool LOCK, bool SURR = false>
__device__ __forceinline__ void b_deform(Ctrl *c, const Geom &g, const PState<T> &P, const GState &G, const Mat &M, long blk) {
    __shared__ double smem[ACC_LDS_DEFORM];
    __shared__ int sint[8];
    if (!c->active) return;
    const long p = blk * long(blockDim.x) + threadIdx.x;
    const long q = p < P.n ? p : 0;
    const bool live = p < P.n && P.Om0[q] != T(0);   // Ω₀ = 0: a padding slot
    const int par = c->parity;
    double xp[D];
    T lp[D];
#pragma unroll
    for (int i = 0; i < D; i++) { xp[i] = P.x[par][i][q]; lp[i] = B == 1 ? P.l[i][q] : T(0); }
    W<D, T> w;
    weights<D, B, T>(g, xp, lp, w);
    BlockBox<D> box;
    box.begin((EP_LDS double *)smem, sint, live, w.base, ACC_LDS_DEFORM);
    if (!live) w.retire(box.lo);
    const bool staged = box.fits(D);
    if (staged) box.template stage<D, T>(g, G.v);
    T gv[D][D];
    if constexpr (D == 3) {
        if (staged) grad_v_rows<T, D, false>(box, w, gv);
        else grad_v<D, T>(g, w, GlobSrc{G.v}, 0, gv);
    } else {
        if (staged) grad_v<D, T>(g, w, LdsSrc<D, D, T>{&box}, 0, gv);
        else grad_v<D, T>(g, w, GlobSrc{G.v}, 0, gv);
    }
    T dF[D][D];
    delta_f<D, T>(T(c->dt), gv, dF);
    if (LOCK) {
        const T dJ = kDev<T> ? det1m<D, T>(dF) : det_cof<D, T>(dF);   // FP32 path: δJ = ΔJ - 1
        // J of the updated F enters here only through the particle mass
        // (ρ₀/J · JΩ₀ finite, ρ · JΩ₀ infinitesimal; deform.jl:19-21, 41-43).
        // It is taken as J_old ΔJ, which equals det(ΔF F) up to rounding, so
        // F is not read here (k_fbar updates it): 9 loads fewer, and no
        // registers for F and ΔF F
        const T Jo = P.J[q];
        const T J = kDev<T> ? Jo + Jo * dJ : Jo * dJ;
        const bool lds = box.fits(1);
        if (lds) box.zero(1);
        double *dst[1] = {G.dJ};
        // FP64: m_p ΔJ_p with m_p = ρΩ of the updated J (volumetric.jl:1-11).
        // FP32: the node then holds ΔJ_n - 1 = Σ N (m_p ΔJ_p - m_p,old) / m_n,
        // m_n = Σ N m_p,old from P2G; m_p = m_p,old for finite deformation
        // (ρ₀/J·JΩ₀), m_p = m_p,old (1 + δJ) for infinitesimal (ρ·JΩ₀), so
        // the scattered quantity is m δJ or m_old δJ (2 + δJ).
        T qv;
        if (!kDev<T> || M.finite) qv = mass_of(P, M, q, J) * dJ;
        else qv = mass_of(P, M, q, P.J[q]) * dJ * (T(2) + dJ);
        auto scatter = [&](const auto &sink) {
            scatter_walk<D, T>(g, w, sink, [&](long no, int li, T N, const T *, const int *) {
                sink(0, no, li, double(N * qv));
            });
        };
        if (lds) scatter(LdsSink<D, 1>{&box});
        else scatter(GlobSink{dst});
        if (lds) box.template flush<1>(g, dst);
    } else if (live) {
        T Fn[D][D];
        push_f<D, T>(P, q, dF, Fn);
        const T J = det_cof<D, T>(Fn);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) {
                P.F[i + D * j][p] = Fn[i][j];
                if (M.tpic) P.gv[i + D * j][p] = gv[i][j];
            }
        P.J[p] = J;
        if (B == 1) domain_update<D, T>(P, M, p, dF, Fn, J);
        constitutive<D, T, SURR>(c, P, M, p, dF);
    }
}

// kernel of one context: block b of the launch is block b of the particles / nodes
template <int D, int B, typename T, bool LOCK, bool SURR = false>
__global__ void __launch_bounds__(EP_PBLK_MAX) k_deform(Ctrl *c, Geom g, PState<T> P, GState G, Mat M) { b_deform<D, B, T, LOCK, SURR>(c, g, P, G, M, blockIdx.x); }

// F-bar (volumetric.jl:12-35) then elast/plast.  The box stages the grid
// velocity and ΔJ_n/m_n exactly as ΔJs writes it (0 where m_n = 0); ΔF,
// ΔJ and F_new are recomputed from the staged velocity with the same
// arithmetic as k_deform, F and J are written here, and ΔF is scaled by
// (Σ N ΔJ_n/ΔJ_p)^(1/d) before the constitutive update.
template <int D, int B, typename T, bool SURR = false>
__device__ __forceinline__ void b_fbar(Ctrl *c, const Geom &g, const PState<T> &P, const GState &G, const Mat &M, long blk) {
    __shared__ double smem[ACC_LDS_FBAR];
    __shared__ int sint[8];
    if (!c->active) return;
    const long p = blk * long(blockDim.x) + threadIdx.x;
    const long q = p < P.n ? p : 0;
    const bool live = p < P.n && P.Om0[q] != T(0);   // Ω₀ = 0: a padding slot
    const int par = c->parity;
    double xp[D];
    T lp[D];
#pragma unroll
    for (int i = 0; i < D; i++) { xp[i] = P.x[par][i][q]; lp[i] = B == 1 ? P.l[i][q] : T(0); }
    W<D, T> w;
    weights<D, B, T>(g, xp, lp, w);
    BlockBox<D> box;
    box.begin((EP_LDS double *)smem, sint, live, w.base, ACC_LDS_FBAR);
    if (!live) w.retire(box.lo);
    const bool staged = box.fits(D + 1);
    if (staged) {
        for (int k = threadIdx.x; k < box.nloc; k += blockDim.x) {
            int z, x, y;
            const long no = box.node(g, k, z, x, y);
            const int a = box.template addr<D + 1>(z, x, y);
            double qv = 0.0;
            if (no >= 0) {
                const double m = G.m[no];
                qv = m == 0.0 ? 0.0 : G.dJ[no] / m;
            }
            EP_LDS T *eb = box.template as<T>();   // staged in the particle precision
#pragma unroll
            for (int d = 0; d < D; d++) eb[a + 16 * d] = no >= 0 ? T(G.v[d][no]) : T(0);
            eb[a + 16 * D] = T(qv);
        }
        __syncthreads();
    }
    if (!live) return;
    T gv[D][D];
    T qsum = T(0);   // Σ N ΔJ_n/m_n from the staged box (3D)
    if constexpr (D == 3) {
        if (staged) qsum = grad_v_rows<T, D + 1, true>(box, w, gv);
        else grad_v<D, T>(g, w, GlobSrc{G.v}, 0, gv);
    } else {
        if (staged) grad_v<D, T>(g, w, LdsSrc<D, D + 1, T>{&box}, 0, gv);
        else grad_v<D, T>(g, w, GlobSrc{G.v}, 0, gv);
    }
    T dF[D][D], Fn[D][D];
    delta_f<D, T>(T(c->dt), gv, dF);
    const T dJp = kDev<T> ? det1m<D, T>(dF) : det_cof<D, T>(dF);
    T acc = T(0);
    // FP64: Σ N ΔJ_n/ΔJ_p as volumetric.jl:27-32 (3D: (Σ N ΔJ_n)/ΔJ_p, one
    // division instead of 64).  FP32: the nodal values are
    // δJ_n = ΔJ_n - 1 and Σ N δJ_n is accumulated, so that with Σ N = 1
    // ΔJ̄/ΔJ_p - 1 = (Σ N δJ_n - δJ_p)/(1 + δJ_p) keeps its relative precision.
    // This is the weights' last use: done before F is read, the weights are
    // not live beside F and ΔF F (FP64: 144 -> 130 VGPRs)
    const T rdiv = kDev<T> ? T(1) : dJp;
    if (D == 3 && staged) {
        acc = kDev<T> ? qsum : qsum / rdiv;
    } else if (staged) {
        gather_walk<D, T>(g, w, LdsSrc<D, D + 1, T>{&box}, [&](long, int li, T N, const T *, const int *) {
            acc += N * box.template as<T>()[li + 16 * D] / rdiv;
        });
    } else {
        for_stencil<D, T>(g, w, [&](long no, T N, const T *, const int *) {
            const double m = G.m[no];
            acc += N * (m == 0.0 ? T(0) : T(G.dJ[no] / m)) / rdiv;
        });
    }
    push_f<D, T>(P, p, dF, Fn);
    const T J = det_cof<D, T>(Fn);
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            P.F[i + D * j][p] = Fn[i][j];
            if (M.tpic) P.gv[i + D * j][p] = gv[i][j];
        }
    P.J[p] = J;
    if (B == 1) domain_update<D, T>(P, M, p, dF, Fn, J);
    if constexpr (kDev<T>) {
        const T rm1 = (acc - dJp) / (T(1) + dJp);
        const T fm1 = expm1(log1p(rm1) / T(D));   // (ΔJ̄/ΔJ_p)^(1/d) - 1
        // K̄ = f (I + K) - I = K + (f - 1)(I + K)
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) dF[i][j] = dF[i][j] + fm1 * ((i == j ? T(1) : T(0)) + dF[i][j]);
    } else {
        // (ΔJ̄/ΔJ_p)^(1/d): cbrt / sqrt instead of pow (within an ulp of
        // pow(acc, 1/d), at a fraction of its cost); a negative ratio is NaN
        // as for pow, and ends the step as non-finite
        const T f = !(acc >= T(0)) ? T(NAN) : (D == 3 ? cbrt(acc) : sqrt(acc));
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) dF[i][j] *= f;
    }
    constitutive<D, T, SURR>(c, P, M, p, dF);
}

// kernel of one context: block b of the launch is block b of the particles / nodes
// k_fbar's register budget: FP32 at 5 waves per SIMD (at most 96 VGPRs, a
// few spilled) runs 10 % faster than at the 4 its 104 VGPRs allow (6 waves
// spill too much); FP64 at 4 waves (128 VGPRs, 2 spilled) 9 % faster than at
// the 3 its 130 allow
template <typename T> constexpr int kFbarWaves = kDev<T> ? 5 : 4;

template <int D, int B, typename T, bool SURR = false>
__global__ void __launch_bounds__(EP_PBLK_MAX) EP_WAVES_PER_EU(kFbarWaves<T>) k_fbar(Ctrl *c, Geom g, PState<T> P, GState G, Mat M) { b_fbar<D, B, T, SURR>(c, g, P, G, M, blockIdx.x); }


// ---------------------------------------------------------------------------
// Non-local regularisation (nonlocal.jl:18-47 as plast.jl:49-60 runs it, the
// neighbourhood of e2e.jl:20-61).  Per step, after G2P: the particles are
// binned by the element of their start-of-step position (the p2e of shpfun,
// tplgy.jl:43-96), then every plastic particle p (Δλ of the last return != 0)
// sums over the particles q of the elements within ceil(ls/h) on every axis,
// with ω = (d/ls) exp(-(d/ls)²) and d from the first two coordinates of the
// updated positions:
//   W_p = Σ_q ω,   ϵpII[2,p] = Σ_q' ω ϵpII[1,q'] / W_p  (ϵpII[1,p] if W_p <= 1e-16)
// where q' leaves out the plastic particles that come before p in the
// caller's order (the reference's sequential "p->q" pass has already given
// their pair a weight and skips it).  Every other particle gets ϵpII[2] = 0.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ long nl_element(const Geom &g, const double *xp, int *e) {
#pragma unroll
    for (int i = 0; i < D; i++) {
        double q = exact_fld(xp[i] - g.x0[i], g.h[i], g.ih[i]);
        const double hi = double(g.nel[i] - 1);
        q = q < 0.0 ? 0.0 : (q > hi ? hi : q);
        e[i] = int(q);
    }
    return D == 2 ? long(e[1]) + long(g.nel[1]) * e[0] : long(e[2]) + long(g.nel[2]) * (long(e[0]) + long(g.nel[0]) * e[1]);
}

__global__ void __launch_bounds__(256) k_nl_zero(const Ctrl *c, int *a, long n) {
    const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (i < n && c->active && c->with_plast) a[i] = 0;
}

template <int D, typename T>
__global__ void __launch_bounds__(256) k_nl_count(const Ctrl *c, Geom g, PState<T> P, NLocal nl) {
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (p >= P.n || !c->active || !c->with_plast) return;
    if (P.Om0[p] == T(0)) { nl.cell[p] = -1; return; }
    double xp[D];
#pragma unroll
    for (int i = 0; i < D; i++) xp[i] = P.x[c->parity][i][p];
    int e[3];
    const long el = nl_element<D>(g, xp, e);
    nl.cell[p] = int(el);
    atomicAdd(&nl.count[el], 1);
}

template <typename T>
__global__ void __launch_bounds__(256) k_nl_fill(const Ctrl *c, long n, NLocal nl) {
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (p >= n || !c->active || !c->with_plast) return;
    const int el = nl.cell[p];
    if (el < 0) return;
    nl.list[nl.start[el] + atomicAdd(&nl.count[el], 1)] = int(p);
}

template <int D, typename T>
__global__ void __launch_bounds__(256) k_nonlocal(const Ctrl *c, Geom g, PState<T> P, NLocal nl) {
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (p >= P.n || !c->active || !c->with_plast) return;
    if (P.Om0[p] == T(0) || P.dlam[p] == T(0)) { P.epII2[p] = T(0); return; }
    const int np = c->parity ^ 1;   // G2P wrote the updated positions here
    const double x0 = P.x[np][0][p], x1 = P.x[np][1][p];
    double xs[D];
#pragma unroll
    for (int i = 0; i < D; i++) xs[i] = P.x[c->parity][i][p];
    int e[3];
    nl_element<D>(g, xs, e);
    const int op = nl.orig[p];
    const double ils = 1.0 / nl.ls;
    double W = 0.0, S = 0.0;
    int lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < D; i++) {
        lo[i] = e[i] - nl.nnel[i] < 0 ? 0 : e[i] - nl.nnel[i];
        hi[i] = e[i] + nl.nnel[i] > g.nel[i] - 1 ? g.nel[i] - 1 : e[i] + nl.nnel[i];
    }
    const int ylo = D == 3 ? lo[1] : 0, yhi = D == 3 ? hi[1] : 0;
    for (int ey = ylo; ey <= yhi; ey++)
        for (int ex = lo[0]; ex <= hi[0]; ex++)
            for (int ez = lo[D - 1]; ez <= hi[D - 1]; ez++) {
                const long el = D == 2 ? long(ez) + long(g.nel[1]) * ex
                                       : long(ez) + long(g.nel[2]) * (long(ex) + long(g.nel[0]) * ey);
                const int a = nl.start[el], b = nl.start[el + 1];
                for (int k = a; k < b; k++) {
                    const int q = nl.list[k];
                    const double xi = x0 - P.x[np][0][q], eta = x1 - P.x[np][1][q];
                    const double r = sqrt(xi * xi + eta * eta) * ils;
                    const double w = r * exp(-(r * r));
                    W += w;
                    if (!(P.dlam[q] != T(0) && nl.orig[q] < op)) S += w * double(P.epII[q]);
                }
            }
    P.epII2[p] = W > 1e-16 ? T(S / W) : P.epII[p];
}

}  // namespace ep
