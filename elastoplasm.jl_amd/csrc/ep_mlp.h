// ep_mlp.h -- learned return mapping for the FP32 Drucker-Prager step.
//
// A multilayer perceptron 8-128-128-128-4 with SiLU maps the trial principal
// log strains and the material parameters (E, ν, c, φ, ψ) to the returned
// principal Kirchhoff stresses and the plastic multiplier of the closed-form
// Drucker-Prager return (DP.jl:1-62, 3D formulas).  Weights and activations
// enter gfx950's v_mfma_f32_32x32x16_bf16 (the 8-input layer and the
// 4-output transpose, K = 8, use v_mfma_f32_32x32x8_bf16) as bf16, every sum
// accumulates in FP32, the FP32 master weights are trained with Adam by the
// kernels below.
//
// One wave evaluates 32 samples: the samples are the N = 32 columns of every
// MFMA, the features are rows, so a layer's 32x32 output tiles (FP32, in the
// D layout) become the next layer's B operands without a trip through LDS:
// D element v of a lane holds row 8(v/4) + 4(lane/32) + v%4, which is exactly
// the k index that B slot v%4 of k-chunk (v/4) expects.  Weight matrices are
// A operands, pre-arranged ("swizzled") so that each lane loads its four bf16
// of one MFMA as one 8-byte LDS read with no bank conflicts.
#pragma once
#include "ep_kernels.h"

namespace ep {
namespace mlp {

constexpr int NI = 8, NH = 128, NO = 4;
// flat FP32 parameter vector: W[out][in] row-major, then the bias
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NO * NH, NPARAM = OB4 + NO;
// the biases as the kernels keep them in LDS: b1, b2, b3, b4 side by side
constexpr int BIAS_B1 = 0, BIAS_B2 = NH, BIAS_B3 = 2 * NH, BIAS_B4 = 3 * NH, BIAS_N = 3 * NH + NO;

// Position of A[i][k] in a swizzled operand image with KC chunks of 8 k
// (32x32x8 operands, 4 bf16 per lane): tile (i/32, k/8), lane
// (i%32) + 32((k%8)/4), slot k%4.
__host__ __device__ constexpr int swz(int i, int k, int KC) {
    return ((((i >> 5) * KC + (k >> 3)) * 64 + (i & 31) + 32 * ((k & 7) >> 2)) << 2) + (k & 3);
}
// The same for 32x32x16 operands (8 bf16 per lane, KC chunks of 16 k): a
// 16-k chunk is the two 8-k chunks 2c, 2c+1 side by side in each lane, slots
// 0..3 from the first and 4..7 from the second: tile (i/32, k/16), lane
// (i%32) + 32((k%8)/4), slot 4((k%16)/8) + k%4.  A layer's D tiles then feed
// the next layer as they are: B chunk c of a lane is D elements 8c..8c+7,
// rows 16c + 4(lane/32) + {0..3} and 16c + 8 + 4(lane/32) + {0..3}, which is
// this k order, so the 16-k MFMA sums the same 16 products as the two 8-k
// MFMAs it replaces.
__host__ __device__ constexpr int swz16(int i, int k, int KC) {
    return ((((i >> 5) * KC + (k >> 4)) * 64 + (i & 31) + 32 * ((k & 7) >> 2)) << 3) + 4 * ((k & 15) >> 3) + (k & 3);
}
// forward images: W1 (4 row tiles x 1 chunk of 8, swz), W2, W3 (4 x 8
// chunks of 16, swz16), W4 (rows 0..3 of 1 x 8 chunks of 16, swz16)
constexpr int IF_W1 = 0, IF_W2 = IF_W1 + 4 * 1 * 256, IF_W3 = IF_W2 + 4 * 8 * 512, IF_W4 = IF_W3 + 4 * 8 * 512,
              IF_N = IF_W4 + 1 * 8 * 512;
// backward images (transposes): W4' (4 x 1 chunk of 8, k 0..3, swz), W3',
// W2' (4 x 8 chunks of 16, swz16)
constexpr int IB_W4T = 0, IB_W3T = IB_W4T + 4 * 1 * 256, IB_W2T = IB_W3T + 4 * 8 * 512, IB_N = IB_W2T + 4 * 8 * 512;

// ---- input / output scaling ------------------------------------------------
// x = [ε1/0.2, ε2/0.2, ε3/0.2 (sorted, descending), (log10 E - 6.5)/1.5,
//      (ν - 0.325)/0.125, log10 c - 4, (φ - 30°)/15°, 2ψ/φ - 1]
// y = [(τ'_k - τtr_k)/fs (k = 1..3), G Δλ/fs],  fs = τII + η P - ξ of the
// trial state (DP.jl:47), > 0 for every state the network stands in for.
// Scaled by the overstress, the plastic correction of a shear return is
// O(1) whatever the overstress, so the network's relative error becomes an
// error of that size in the correction only: small for the small
// overstresses of an explicit step.  The residual check is scaled by
// S = |τtr|_2 + ξ.
constexpr double EPS_SCALE = 0.2;
constexpr int SAMPLE_TRIES = 128;

template <typename T>
__host__ __device__ inline void features(const T *e, T E, T nu, T c, T phi, T psi, T *x) {
    const T d2r = T(3.14159265358979323846 / 180.0);
    x[0] = e[0] / T(EPS_SCALE);
    x[1] = e[1] / T(EPS_SCALE);
    x[2] = e[2] / T(EPS_SCALE);
    x[3] = (log10(E) - T(6.5)) / T(1.5);
    x[4] = (nu - T(0.325)) / T(0.125);
    x[5] = log10(c) - T(4);
    x[6] = (phi - T(30) * d2r) / (T(15) * d2r);
    x[7] = phi > T(0) ? T(2) * psi / phi - T(1) : T(-1);
}

// 3D materialParam (DP.jl:18-22)
template <typename T>
__host__ __device__ inline void dp_params(T phi, T psi, T c, T &eta, T &etaB, T &xi) {
    const T sp = sin(phi), sq = sin(psi);
    eta = T(6) * sp / (sqrt(T(3)) * (T(3) + sp));
    etaB = T(6) * sq / (sqrt(T(3)) * (T(3) + sq));
    xi = T(6) * c * cos(phi) / (sqrt(T(3)) * (T(3) + sp));
}

// ---- bf16 ------------------------------------------------------------------
__host__ __device__ inline unsigned short bf16_rn(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf16_f(unsigned short h) {
    const unsigned u = unsigned(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// four floats as bf16 (round to nearest even, the same rounding as bf16_rn)
__device__ inline v4s pack4(float a, float b, float c, float d) {
    v4s r;
    r[0] = short(ep_bf16_bits(a));
    r[1] = short(ep_bf16_bits(b));
    r[2] = short(ep_bf16_bits(c));
    r[3] = short(ep_bf16_bits(d));
    return r;
}
// logistic σ(z) = 1 / (1 + e^-z) with the hardware exp2 and reciprocal (about
// 1 ulp, far inside the bf16 rounding that follows); SiLU = z σ(z),
// SiLU' = σ (1 + z (1 - σ))
__device__ inline float sigmoid(float z) { return ep_rcp(1.f + ep_exp2(-1.4426950408889634f * z)); }
// row of D element v in row tile mt
__device__ inline int drow(int mt, int v, int lane) { return 32 * mt + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

template <int MT>
__device__ inline void zero_acc(v16f (&acc)[MT]) {
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[m][v] = 0.f;
}

// acc[m] = Σ_kc A(m, kc) B(kc) over KC chunks of 8 k (32x32x8, swz
// images); chunks outermost so the MT accumulation chains interleave.
template <int MT, int KC>
__device__ inline void gemm_tile(const EP_LDS unsigned short *img, const v4s (&b)[KC], v16f (&acc)[MT], int lane) {
    zero_acc<MT>(acc);
#pragma unroll
    for (int kc = 0; kc < KC; kc++)
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const v4s a = *(const EP_LDS v4s *)(img + (((m * KC + kc) * 64 + lane) << 2));
            acc[m] = EP_MFMA_32X32X8_BF16(a, b[kc], acc[m]);
        }
}
// The same over 2*KC16 chunks of 8 k with 32x32x16 MFMAs (swz16 images): B
// chunks 2c and 2c+1 form the 16-k operand of chunk c, A is one 16-byte LDS
// read per lane.
template <int MT, int KC16>
__device__ inline void gemm_tile16(const EP_LDS unsigned short *img, const v4s (&b)[2 * KC16], v16f (&acc)[MT],
                                   int lane) {
    zero_acc<MT>(acc);
#pragma unroll
    for (int kc = 0; kc < KC16; kc++) {
        v8s bb;
#pragma unroll
        for (int q = 0; q < 4; q++) { bb[q] = b[2 * kc][q]; bb[4 + q] = b[2 * kc + 1][q]; }
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const v8s a = *(const EP_LDS v8s *)(img + (((m * KC16 + kc) * 64 + lane) << 3));
            acc[m] = EP_MFMA_32X32X16_BF16(a, bb, acc[m]);
        }
    }
}

// bias + SiLU of a 128-row layer, packed as the next layer's 16 B chunks;
// optionally the bf16 SiLU' (for backward) and the activations themselves.
__device__ inline void activate(const v16f (&acc)[4], const EP_LDS float *bias, int lane, v4s (&h)[16], v4s *dh) {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float z[4];
#pragma unroll
            for (int q = 0; q < 4; q++) z[q] = acc[m][4 * r + q] + bias[drow(m, 4 * r + q, lane)];
            float sg[4];
#pragma unroll
            for (int q = 0; q < 4; q++) sg[q] = sigmoid(z[q]);
            h[4 * m + r] = pack4(z[0] * sg[0], z[1] * sg[1], z[2] * sg[2], z[3] * sg[3]);
            if (dh)
                dh[4 * m + r] = pack4(sg[0] * (1.f + z[0] * (1.f - sg[0])), sg[1] * (1.f + z[1] * (1.f - sg[1])),
                                      sg[2] * (1.f + z[2] * (1.f - sg[2])), sg[3] * (1.f + z[3] * (1.f - sg[3])));
        }
}

// Forward pass of one 32-sample tile.  xb: this lane's B operand of layer 1
// (features 4(lane/32)..+3 of sample lane%32).  On return lanes 0..31 hold
// the four outputs of their sample in y[0..3].  With TRAIN the bf16
// activations and SiLU' are kept for the backward pass.
template <bool TRAIN>
__device__ inline void forward(const EP_LDS unsigned short *wf, const EP_LDS float *bias, v4s xb, int lane, float (&y)[4],
                               v4s (&h1)[16], v4s (&h2)[16], v4s (&h3)[16], v4s *d1, v4s *d2, v4s *d3) {
    v16f acc[4];
    {
        const v4s b1[1] = {xb};
        gemm_tile<4, 1>(wf + IF_W1, b1, acc, lane);
    }
    activate(acc, bias + BIAS_B1, lane, h1, TRAIN ? d1 : nullptr);
    gemm_tile16<4, 8>(wf + IF_W2, h1, acc, lane);
    activate(acc, bias + BIAS_B2, lane, h2, TRAIN ? d2 : nullptr);
    gemm_tile16<4, 8>(wf + IF_W3, h2, acc, lane);
    activate(acc, bias + BIAS_B3, lane, h3, TRAIN ? d3 : nullptr);
    v16f o[1];
    gemm_tile16<1, 8>(wf + IF_W4, h3, o, lane);
#pragma unroll
    for (int q = 0; q < 4; q++) y[q] = o[0][q] + bias[BIAS_B4 + q];   // rows 0..3 live in lanes 0..31, v = 0..3
}

// copy a swizzled image (n elements, multiple of 8) from HBM to LDS
__device__ inline void load_image(EP_LDS unsigned short *dst, const unsigned short *src, int n) {
    for (int i = threadIdx.x * 8; i < n; i += blockDim.x * 8) {
        const v4u v = *(const v4u *)(src + i);
        *(EP_LDS v4u *)(dst + i) = v;
    }
}
// the biases b1, b2, b3, b4 side by side in LDS (BIAS_*): each activation
// reads 64 of them per lane, which from global memory put an L1 round trip
// in every layer of every tile
__device__ inline void load_bias(EP_LDS float *dst, const float *prm) {
    for (int i = threadIdx.x; i < BIAS_N; i += blockDim.x)
        dst[i] = i < BIAS_B4 ? prm[(i / NH == 0 ? OB1 : i / NH == 1 ? OB2 : OB3) + i % NH] : prm[OB4 + i - BIAS_B4];
}

// ---- training --------------------------------------------------------------
struct TrainBuf {
    const float *X, *Y;         // samples [n][8], targets [n][4]
    long first;                 // first sample of this batch
    long batch;
    unsigned short *act;        // x̂ [8][batch], ĥ1..ĥ3 [128][batch] (bf16, batch contiguous)
    unsigned short *del;        // ĝ4 [4][batch], ĝ3..ĝ1 [128][batch]
    float *lossp;               // [batch/32] per-tile Σ (y - t)²
};
constexpr long ACT_H0 = 0, ACT_H1 = 8, ACT_H2 = 8 + 128, ACT_H3 = 8 + 256, ACT_ROWS = 8 + 384;   // row offsets
constexpr long DEL_G4 = 0, DEL_G3 = 4, DEL_G2 = 4 + 128, DEL_G1 = 4 + 256, DEL_ROWS = 4 + 384;

__device__ inline void store_rows(unsigned short *dst, long batch, long col, const v4s (&h)[16], int lane) {
#pragma unroll
    for (int kc = 0; kc < 16; kc++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = 8 * kc + 4 * (lane >> 5) + q;   // = drow(kc/4, 4(kc%4)+q, lane)
            dst[row * batch + col] = (unsigned short)h[kc][q];
        }
}

// Forward + backward of one 32-sample tile per wave; writes the bf16
// activations and output-side gradients that the weight gradient needs.
// Loss: mean over the batch and the 4 outputs of (y - t)².
__global__ void __launch_bounds__(256) k_mlp_fb(const unsigned short *img_f, const unsigned short *img_b,
                                                const float *prm, TrainBuf tb) {
    __shared__ __attribute__((aligned(16))) unsigned short sf[IF_N];
    __shared__ __attribute__((aligned(16))) unsigned short sb[IB_N];
    __shared__ float sbias[BIAS_N];
    load_image((EP_LDS unsigned short *)sf, img_f, IF_N);
    load_image((EP_LDS unsigned short *)sb, img_b, IB_N);
    load_bias((EP_LDS float *)sbias, prm);
    __syncthreads();
    const int lane = __lane_id(), nw = blockDim.x >> 6;
    const long ntile = tb.batch / 32;
    const float gscale = 2.f / float(4 * tb.batch);
    for (long t = long(blockIdx.x) * nw + (threadIdx.x >> 6); t < ntile; t += long(gridDim.x) * nw) {
        const long col = t * 32 + (lane & 31);
        const long smp = tb.first + col;
        const int hh = lane >> 5;
        float xv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) xv[q] = tb.X[smp * 8 + 4 * hh + q];
        const v4s xb = pack4(xv[0], xv[1], xv[2], xv[3]);
#pragma unroll
        for (int q = 0; q < 4; q++) tb.act[(ACT_H0 + 4 * hh + q) * tb.batch + col] = (unsigned short)xb[q];
        v4s h1[16], h2[16], h3[16], d1[16], d2[16], d3[16];
        float y[4];
        forward<true>((const EP_LDS unsigned short *)sf, (const EP_LDS float *)sbias, xb, lane, y, h1, h2, h3, d1, d2,
                      d3);
        store_rows(tb.act + ACT_H1 * tb.batch, tb.batch, col, h1, lane);
        store_rows(tb.act + ACT_H2 * tb.batch, tb.batch, col, h2, lane);
        store_rows(tb.act + ACT_H3 * tb.batch, tb.batch, col, h3, lane);
        // output gradient: lanes 0..31 own the sample's 4 outputs
        float g4[4], l2 = 0.f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float r = y[q] - tb.Y[smp * 4 + q];
            l2 += r * r;
            g4[q] = hh == 0 ? gscale * r : 0.f;
        }
        const v4s gb4 = pack4(g4[0], g4[1], g4[2], g4[3]);
        if (hh == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) tb.del[(DEL_G4 + q) * tb.batch + col] = (unsigned short)gb4[q];
        }
        // Σ (y - t)² of the tile, lanes 0..31
        float ls = hh == 0 ? l2 : 0.f;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            unsigned long long u = 0;
            __builtin_memcpy(&u, &ls, 4);
            const unsigned long long o = __shfl_xor(u, off);
            float of;
            __builtin_memcpy(&of, &o, 4);
            ls += of;
        }
        if (lane == 0) tb.lossp[t] = ls;
        // backward through layer 4 (K = 4 outputs, padded to 8), then 3, 2
        v16f acc[4];
        v4s g[16];
        {
            const v4s b4[1] = {gb4};
            gemm_tile<4, 1>((const EP_LDS unsigned short *)sb + IB_W4T, b4, acc, lane);
        }
        auto grad_act = [&](const v4s (&dz)[16]) {
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const v4s s = dz[4 * m + r];
                    g[4 * m + r] = pack4(acc[m][4 * r + 0] * bf16_f((unsigned short)s[0]),
                                         acc[m][4 * r + 1] * bf16_f((unsigned short)s[1]),
                                         acc[m][4 * r + 2] * bf16_f((unsigned short)s[2]),
                                         acc[m][4 * r + 3] * bf16_f((unsigned short)s[3]));
                }
        };
        grad_act(d3);
        store_rows(tb.del + DEL_G3 * tb.batch, tb.batch, col, g, lane);
        gemm_tile16<4, 8>((const EP_LDS unsigned short *)sb + IB_W3T, g, acc, lane);
        grad_act(d2);
        store_rows(tb.del + DEL_G2 * tb.batch, tb.batch, col, g, lane);
        gemm_tile16<4, 8>((const EP_LDS unsigned short *)sb + IB_W2T, g, acc, lane);
        grad_act(d1);
        store_rows(tb.del + DEL_G1 * tb.batch, tb.batch, col, g, lane);
    }
}

// Weight gradients dW_l = ĝ_l ĥ_{l-1}' and bias gradients Σ ĝ_l as a
// split-K MFMA product over the batch.  Task = (layer tile, split); split s
// writes its partial sums to part[s][param] (every parameter exactly once
// per split), k_adam adds the splits in a fixed order.
struct DwTask {
    int layer, mt, nt;
};
__device__ inline DwTask dw_task(int t) {
    // layer 1: 4 row tiles x 1 column tile; layers 2, 3: 4 x 4; layer 4: 1 x 4
    if (t < 4) return {1, t, 0};
    t -= 4;
    if (t < 16) return {2, t >> 2, t & 3};
    t -= 16;
    if (t < 16) return {3, t >> 2, t & 3};
    t -= 16;
    return {4, 0, t};
}
constexpr int DW_TILES = 4 + 16 + 16 + 4;

__global__ void __launch_bounds__(64) k_mlp_dw(const unsigned short *act, const unsigned short *del, long batch,
                                               int nsplit, float *part) {
    const int lane = __lane_id();
    const int task = blockIdx.x % DW_TILES, split = blockIdx.x / DW_TILES;
    const DwTask T = dw_task(task);
    // A = ĝ_l rows (outputs), B = ĥ_{l-1} rows (inputs); M, N: valid rows / columns
    long arow0, brow0;
    int M, N, ow, ob, nin;
    if (T.layer == 1) { arow0 = DEL_G1; brow0 = ACT_H0; M = 128; N = 8; ow = OW1; ob = OB1; nin = NI; }
    else if (T.layer == 2) { arow0 = DEL_G2; brow0 = ACT_H1; M = 128; N = 128; ow = OW2; ob = OB2; nin = NH; }
    else if (T.layer == 3) { arow0 = DEL_G3; brow0 = ACT_H2; M = 128; N = 128; ow = OW3; ob = OB3; nin = NH; }
    else { arow0 = DEL_G4; brow0 = ACT_H3; M = 4; N = 128; ow = OW4; ob = OB4; nin = NH; }
    const int i = 32 * T.mt + (lane & 31);   // A row of this lane
    const int j = 32 * T.nt + (lane & 31);   // B column of this lane
    const bool arow_ok = i < M, bcol_ok = j < N;
    // the batch is the K dimension: 32x32x16 products, lane l covering
    // batch entries 16j + 8(l/32) + {0..7} of each chunk j (16 bytes)
    const long kspan = batch / nsplit, k0 = long(split) * kspan;
    const unsigned short *ap = del + (arow0 + (arow_ok ? i : 0)) * batch + k0 + 8 * (lane >> 5);
    const unsigned short *bp = act + (brow0 + (bcol_ok ? j : 0)) * batch + k0 + 8 * (lane >> 5);
    v16f acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = 0.f;
    float bsum = 0.f;
    v8s zero;
#pragma unroll
    for (int q = 0; q < 8; q++) zero[q] = 0;
    for (long k = 0; k < kspan; k += 16) {
        v8s a = *(const v8s *)(ap + k);
        v8s b = *(const v8s *)(bp + k);
        if (!arow_ok) a = zero;
        if (!bcol_ok) b = zero;
        if (T.nt == 0) {
            float s4 = 0.f;
#pragma unroll
            for (int q = 0; q < 8; q++) s4 += bf16_f((unsigned short)a[q]);
            bsum += s4;
        }
        acc = EP_MFMA_32X32X16_BF16(a, b, acc);
    }
    float *out = part + long(split) * NPARAM;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int r = 32 * T.mt + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int cc = 32 * T.nt + (lane & 31);
        if (r < M && cc < N) out[ow + r * nin + cc] = acc[v];
    }
    if (T.nt == 0) {
        unsigned long long u = 0;
        __builtin_memcpy(&u, &bsum, 4);
        const unsigned long long o = __shfl_xor(u, 32);
        float of;
        __builtin_memcpy(&of, &o, 4);
        if (lane < 32 && i < M) out[ob + i] = bsum + of;
    }
}

// swizzled bf16 images of parameter k (no-op for biases)
__device__ inline void pack_param(int k, float w, unsigned short *img_f, unsigned short *img_b) {
    const unsigned short h = bf16_rn(w);
    if (k < OB1) {
        const int o = (k - OW1) / NI, i = (k - OW1) % NI;
        img_f[IF_W1 + swz(o, i, 1)] = h;
    } else if (k >= OW2 && k < OB2) {
        const int o = (k - OW2) / NH, i = (k - OW2) % NH;
        img_f[IF_W2 + swz16(o, i, 8)] = h;
        img_b[IB_W2T + swz16(i, o, 8)] = h;
    } else if (k >= OW3 && k < OB3) {
        const int o = (k - OW3) / NH, i = (k - OW3) % NH;
        img_f[IF_W3 + swz16(o, i, 8)] = h;
        img_b[IB_W3T + swz16(i, o, 8)] = h;
    } else if (k >= OW4 && k < OB4) {
        const int o = (k - OW4) / NH, i = (k - OW4) % NH;
        img_f[IF_W4 + swz16(o, i, 8)] = h;
        img_b[IB_W4T + swz(i, o, 1)] = h;
    }
}

__global__ void k_mlp_pack(const float *prm, unsigned short *img_f, unsigned short *img_b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NPARAM) pack_param(k, prm[k], img_f, img_b);
}

struct AdamArgs {
    float lr, b1, b2, eps, c1, c2;   // c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
    int update;                      // 0: only sum the splits into grad
};
__global__ void k_adam(const float *part, int nsplit, float *grad, float *prm, float *m, float *v,
                       unsigned short *img_f, unsigned short *img_b, AdamArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPARAM) return;
    float g = 0.f;
    for (int s = 0; s < nsplit; s++) g += part[long(s) * NPARAM + k];
    grad[k] = g;
    if (!a.update) return;
    const float mk = a.b1 * m[k] + (1.f - a.b1) * g;
    const float vk = a.b2 * v[k] + (1.f - a.b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    const float w = prm[k] - a.lr * (mk * a.c1) / (sqrtf(vk * a.c2) + a.eps);
    prm[k] = w;
    pack_param(k, w, img_f, img_b);
}

__global__ void k_loss_sum(const float *lossp, long ntile, long batch, float *out) {
    __shared__ float red[256];
    float s = 0.f;
    for (long t = threadIdx.x; t < ntile; t += blockDim.x) s += lossp[t];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] / float(4 * batch);
}

// ---- samples ---------------------------------------------------------------
// counter-based uniform draw in [0, 1): splitmix64 of (seed, sample, k)
__host__ __device__ inline double u01(unsigned long long seed, unsigned long long i, int k) {
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + (i * 8ull + unsigned(k) + 1ull) * 0xD1B54A32D192ED03ull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

// Sample i: trial principal log strains U[-0.2, 0.2]^3 sorted descending,
// E log-uniform [1e5, 1e8], ν U[0.2, 0.45], c log-uniform [1e3, 1e5],
// φ U[15°, 45°], ψ U[0, φ]; target = the FP64 return (3D DP.jl formulas with
// materialParam's ψ) of the isotropic trial stress Del ε.  Draws that do
// not end on the cone by a shear return (elastic, apex, or τn = ξ - η Pn
// < 0, where DP.jl's σn flips the deviator) are redrawn with the next seed
// of the sequence seed + a·0x632BE59BD9B4E019, a = 0, 1, ...: the network
// only ever stands in for regular shear returns.
// raw[i] = [τ'1, τ'2, τ'3, Δλ, fs, S] in FP64.
// The FP64 return of the isotropic trial stress Del ε (principal, sorted
// strains e) with cohesion c (no softening within the return) and whether it
// is a regular shear return onto the cone: trial tr, returned s, Δλ, the
// trial overstress fs and the residual scale S.
__device__ inline bool dp_sample(const double (&e)[3], double E, double nu, double c, double phi, double psi,
                                 double (&tr)[3], double (&s)[6], double &G, double &dl, double &fs, double &S) {
    const double K = E / (3.0 * (1.0 - 2.0 * nu));
    G = E / (2.0 * (1.0 + nu));
    const double D11 = K + 4.0 / 3.0 * G, D12 = K - 2.0 / 3.0 * G;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double a0 = k == 0 ? D11 : D12, a1 = k == 1 ? D11 : D12, a2 = k == 2 ? D11 : D12;
        tr[k] = a0 * e[0] + a1 * e[1] + a2 * e[2];
        s[k] = tr[k];
    }
    s[3] = s[4] = s[5] = 0.0;
    double e1 = 0.0;
    dl = 0.0;
    const int br = dp_return<6, double>(s, c, c, phi, 0.0, K, G, 0.0, e1, dl, psi);
    double eta, etaB, xi;
    dp_params<double>(phi, psi, c, eta, etaB, xi);
    const double P = (tr[0] + tr[1] + tr[2]) / 3.0;
    const double d0 = tr[0] - P, d1 = tr[1] - P, d2 = tr[2] - P;
    fs = sqrt(0.5 * (d0 * d0 + d1 * d1 + d2 * d2)) + eta * P - xi;
    S = sqrt(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2]) + xi;
    return br == 1 && fs > 0.0 && xi - eta * ((s[0] + s[1] + s[2]) / 3.0) >= 0.0;
}

// sample i of a training set: features, scaled targets, raw FP64 values
__device__ inline void put_sample(long i, const double (&e)[3], double E, double nu, double c, double phi, double psi,
                                  bool good, const double (&tr)[3], const double (&s)[6], double G, double dl,
                                  double fs, double S, float *X, float *Y, double *raw) {
    double x[8];
    features<double>(e, E, nu, c, phi, psi, x);
#pragma unroll
    for (int k = 0; k < 8; k++) X[i * 8 + k] = float(x[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) Y[i * 4 + k] = good ? float((s[k] - tr[k]) / fs) : 0.f;
    Y[i * 4 + 3] = good ? float(G * dl / fs) : 0.f;
    if (raw) {
        double *r = raw + i * 6;
        r[0] = s[0]; r[1] = s[1]; r[2] = s[2]; r[3] = dl; r[4] = good ? fs : 0.0; r[5] = S;
    }
}

__global__ void k_mlp_samples(long n, unsigned long long seed, float *X, float *Y, double *raw) {
    const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double d2r = 3.14159265358979323846 / 180.0;
    double e[3], E, nu, c, phi, psi, G, tr[3], s[6], dl, fs, S;
    bool good = false;
    for (int a = 0; a < SAMPLE_TRIES && !good; a++) {
        const unsigned long long sd = seed + (unsigned long long)a * 0x632BE59BD9B4E019ull;
#pragma unroll
        for (int k = 0; k < 3; k++) e[k] = 0.4 * u01(sd, i, k) - 0.2;
        if (e[0] < e[1]) { const double t = e[0]; e[0] = e[1]; e[1] = t; }
        if (e[1] < e[2]) { const double t = e[1]; e[1] = e[2]; e[2] = t; }
        if (e[0] < e[1]) { const double t = e[0]; e[0] = e[1]; e[1] = t; }
        E = pow(10.0, 5.0 + 3.0 * u01(sd, i, 3));
        nu = 0.2 + 0.25 * u01(sd, i, 4);
        c = pow(10.0, 3.0 + 2.0 * u01(sd, i, 5));
        phi = (15.0 + 30.0 * u01(sd, i, 6)) * d2r;
        psi = phi * u01(sd, i, 7);
        good = dp_sample(e, E, nu, c, phi, psi, tr, s, G, dl, fs, S);
    }
    put_sample(i, e, E, nu, c, phi, psi, good, tr, s, G, dl, fs, S, X, Y, raw);
}

// Training samples from the step's own trial states: the plastic trial
// states k_fbar queued in the last step of an EP_FP32_SURROGATE context,
// with the inputs k_surr gives the network (sorted principal log strains,
// E, ν, the particle's cohesion and φ, ψ = 0) and the FP64 return of
// k_mlp_samples as targets.  The cohesion is the one of the trial state,
// c0 + Hp ϵpII before the step's return, max cr (ϵpII less Δλ/√3, the
// increment of a shear return with ψ = 0).  States that do not end on the
// cone by a shear return are left out; the others are appended at
// first + atomicAdd(count) up to cap.
__global__ void k_mlp_harvest(const Ctrl *c, PState<float> P, Mat M, long first, long cap, float *X, float *Y,
                              double *raw, int *count) {
    const long slot = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (slot >= long(c->qcount)) return;
    const long p = M.q.idx[slot];
    double e[3];
#pragma unroll
    for (int k = 0; k < 3; k++) e[k] = double(M.q.e[k][slot]);
    const double ep0 = double(P.epII[p]) - double(P.dlam[p]) * sqrt(1.0 / 3.0);
    double cc = double(P.c0[p]) + M.Hp * ep0;
    if (cc < double(P.cr[p])) cc = double(P.cr[p]);
    // the network sees the FP32 cohesion k_surr forms
    cc = double(float(cc));
    const double E = double(float(M.E)), nu = double(float(M.nu)), phi = double(P.phi[p]);
    double tr[3], s[6], G, dl, fs, S;
    if (!dp_sample(e, E, nu, cc, phi, 0.0, tr, s, G, dl, fs, S)) return;
    const long i = first + atomicAdd(count, 1);
    if (i < cap) put_sample(i, e, E, nu, cc, phi, 0.0, true, tr, s, G, dl, fs, S, X, Y, raw);
}

// plain inference on [n][8] inputs (n a multiple of 32) -> [n][4]
__global__ void __launch_bounds__(256) k_mlp_predict(const unsigned short *img_f, const float *prm, const float *X,
                                                     long n, float *Y) {
    __shared__ __attribute__((aligned(16))) unsigned short sf[IF_N];
    __shared__ float sbias[BIAS_N];
    load_image((EP_LDS unsigned short *)sf, img_f, IF_N);
    load_bias((EP_LDS float *)sbias, prm);
    __syncthreads();
    const int lane = __lane_id(), nw = blockDim.x >> 6;
    for (long t = long(blockIdx.x) * nw + (threadIdx.x >> 6); t * 32 < n; t += long(gridDim.x) * nw) {
        const long smp = t * 32 + (lane & 31);
        const int hh = lane >> 5;
        const v4s xb = pack4(X[smp * 8 + 4 * hh], X[smp * 8 + 4 * hh + 1], X[smp * 8 + 4 * hh + 2],
                             X[smp * 8 + 4 * hh + 3]);
        v4s h1[16], h2[16], h3[16];
        float y[4];
        forward<false>((const EP_LDS unsigned short *)sf, (const EP_LDS float *)sbias, xb, lane, y, h1, h2, h3, nullptr,
                       nullptr, nullptr);
        if (hh == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) Y[smp * 4 + q] = y[q];
        }
    }
}

// ---- learned return inside the step ---------------------------------------
struct SurrNet {
    const unsigned short *img_f;
    const float *prm;
    float tol;
};

// Residual-checked learned return for the plastic trial states that k_fbar
// (or k_deform) queued.  Lanes 0..31 of a wave own one queued particle each:
// the prediction (τ' = τtr + fs y[0..2], Δλ = fs y[3]/G) is accepted when
// the trial state takes the shear branch and both the flow rule of
// DP.jl:50-53 (τ' = τtr - G Δλ dev/τII - K ηB Δλ) and the yield condition
// f(τ') = τII' + η P' - ξ = 0 hold to tol·S; otherwise the FP64
// closed-form return replaces it.  The apex
// branch (Pn = σm - P, DP.jl:56-62) satisfies neither and always falls back.
// Then b = V diag(exp 2 Del⁻¹τ') V' and τ = V diag(τ') V' as on the plain
// FP32 path.
__global__ void __launch_bounds__(256) k_surr(Ctrl *c, PState<float> P, Mat M, SurrNet net) {
    __shared__ __attribute__((aligned(16))) unsigned short sf[IF_N];
    __shared__ float sbias[BIAS_N];
    if (!c->active) return;
    load_image((EP_LDS unsigned short *)sf, net.img_f, IF_N);
    load_bias((EP_LDS float *)sbias, net.prm);
    __syncthreads();
    const int count = c->qcount;
    const int lane = __lane_id(), nw = blockDim.x >> 6, hh = lane >> 5;
    const float E = float(M.E), nu = float(M.nu), Kf = float(M.Kc), Gf = float(M.Gc), Hp = float(M.Hp);
    // statistics of the wave's tiles, added to the control block once at the
    // end (atomics on the same three addresses after every tile serialise:
    // 2.5 ms per step when most of the column yields)
    unsigned long long n_acc = 0, n_fb = 0, rr_max = 0, any_bad = 0;
    for (long t = long(blockIdx.x) * nw + (threadIdx.x >> 6); t * 32 < count; t += long(gridDim.x) * nw) {
        const long slot = t * 32 + (lane & 31);
        const bool valid = slot < count;
        const long sq = valid ? slot : t * 32;
        const long p = M.q.idx[sq];
        float e[3];
#pragma unroll
        for (int k = 0; k < 3; k++) e[k] = M.q.e[k][sq];
        const float epII2 = P.epII[p];
        float cc = P.c0[p] + Hp * epII2;
        if (cc < P.cr[p]) cc = P.cr[p];
        const float phi = P.phi[p];
        float x[8];
        features<float>(e, E, nu, cc, phi, 0.f, x);
        const v4s xb = pack4(x[4 * hh], x[4 * hh + 1], x[4 * hh + 2], x[4 * hh + 3]);
        v4s h1[16], h2[16], h3[16];
        float y[4];
        forward<false>((const EP_LDS unsigned short *)sf, (const EP_LDS float *)sbias, xb, lane, y, h1, h2, h3, nullptr,
                       nullptr, nullptr);
        bool ok = false, fb = false, bad = false;
        float res = 0.f;
        if (hh == 0 && valid) {
            float tr[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                float a = 0.f;
#pragma unroll
                for (int j = 0; j < 3; j++) a += float(M.Del[i + 6 * j]) * e[j];
                tr[i] = a;
            }
            const float Pt = (tr[0] + tr[1] + tr[2]) / 3.f;
            float d[3];
#pragma unroll
            for (int k = 0; k < 3; k++) d[k] = tr[k] - Pt;
            const float tII = sqrtf(0.5f * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
            float eta, etaB, xi;
            dp_params<float>(phi, 0.f, cc, eta, etaB, xi);
            const float S = sqrtf(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2]) + xi;
            const float sm = xi / eta, tP = xi - eta * (xi / eta);
            const float fs = tII + eta * Pt - xi;
            const float hy = tII - tP - sqrtf(1.f + eta * eta) * (Pt - sm);
            const bool shear = (fs > 0.f && Pt < sm) || (hy > 0.f && Pt >= sm);
            float tp[3];
#pragma unroll
            for (int k = 0; k < 3; k++) tp[k] = tr[k] + fs * y[k];
            float dl = fs * y[3] / Gf;
            float r = 0.f;
#pragma unroll
            for (int k = 0; k < 3; k++)
                r = fmaxf(r, fabsf(tp[k] - (tr[k] - Gf * dl * d[k] / tII - Kf * etaB * dl)));
            const float Pp = (tp[0] + tp[1] + tp[2]) / 3.f;
            const float q0 = tp[0] - Pp, q1 = tp[1] - Pp, q2 = tp[2] - Pp;
            const float tIIp = sqrtf(0.5f * (q0 * q0 + q1 * q1 + q2 * q2));
            r = fmaxf(r, fabsf(tIIp + eta * Pp - xi));
            res = r / S;
            if (!(res <= 3.0e38f)) res = 3.0e38f;
            ok = shear && fs > 0.f && tII > 0.f && res <= net.tol;
            float epn;
            if (ok) {
                epn = epII2 + dl * sqrtf(1.f / 3.f + 2.f / 9.f * etaB * etaB);
            } else {
                fb = true;
                double s64[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    double a = 0.0;
#pragma unroll
                    for (int j = 0; j < 3; j++) a += M.Del[i + 6 * j] * double(e[j]);
                    s64[i] = a;
                }
                double e1 = double(epII2), dld = 0.0;
                dp_return<6, double>(s64, double(P.c0[p]), double(P.cr[p]), double(phi), M.Hp, M.Kc, M.Gc,
                                     double(epII2), e1, dld);
#pragma unroll
                for (int k = 0; k < 3; k++) tp[k] = float(s64[k]);
                dl = float(dld);
                epn = float(e1);
            }
            // b - I = V diag(expm1(2 Del⁻¹τ')) V', τ = V diag(τ') V'
            float sv[6] = {tp[0], tp[1], tp[2], 0.f, 0.f, 0.f}, yv[6];
            lu_solve<6, float>(M.lu, sv, yv);
            float V[3][3], fe[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                fe[k] = expm1f(2.f * yv[k]);
#pragma unroll
                for (int i = 0; i < 3; i++) V[i][k] = M.q.V[i + 3 * k][sq];
            }
            float bn[3][3], tm[3][3];
            vdvt<3, float>(V, fe, bn);
            vdvt<3, float>(V, tp, tm);
            store_sym<3, float>(P.b, p, bn);
            float sv2[6];
            sym_to_voigt<3, float>(tm, sv2);
#pragma unroll
            for (int k = 0; k < 6; k++) {
                P.s[k][p] = sv2[k];
                if (!isfinite(sv2[k])) bad = true;
            }
            P.epII[p] = epn;
            P.dlam[p] = dl;
        }
        n_acc += __popcll(__ballot(ok));
        n_fb += __popcll(__ballot(fb));
        any_bad |= __ballot(bad);
        unsigned rb;
        __builtin_memcpy(&rb, &res, 4);
        rr_max = rb > rr_max ? rb : rr_max;   // non-negative floats order as their bits
    }
    unsigned long long rr = rr_max;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(rr, off);
        rr = o > rr ? o : rr;
    }
    if (lane == 0) {
        if (n_acc) atomicAdd(&c->s_acc, n_acc);
        if (n_fb) atomicAdd(&c->s_fb, n_fb);
        if (rr) atomicMax(&c->s_maxr, unsigned(rr));
        if (any_bad) atomicOr(&c->err, ERR_NONFINITE);
    }
}

}  // namespace mlp
}  // namespace ep
