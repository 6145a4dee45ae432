// ep_batch.h -- an ensemble of independent domains (slumps of different
// sizes and materials) advanced by one launch per stage of the step.
//
// Every instance keeps its own control block (t, dt, vmax, window), grid and
// particle arrays.  A launch covers the blocks of all instances back to
// back; a block finds its instance in a block -> instance table and runs the
// same block body as the single-domain kernels with that instance's
// descriptors, which it reads from one table in HBM (uniform per block, so
// scalar loads).  Instances are padded to whole blocks, so a block never
// spans two instances.
#pragma once
#include "ep_kernels.h"

namespace ep {

template <typename T>
struct Inst {
    Ctrl *c;
    Geom g;
    PState<T> P;
    GState G;
    Mat M;
};

// blocks [first[i], first[i+1]) of a launch belong to instance inst[b]
struct BlockMap {
    const int *inst;
    const int *first;
};

template <int D, typename T>
__global__ void kb_begin(const Inst<T> *I, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b_begin<D>(I[i].c, I[i].g, I[i].M);
}

template <typename T>
__global__ void __launch_bounds__(1024) kb_end(const Inst<T> *I) { b_end(I[blockIdx.x].c); }

template <int D, typename T>
__global__ void kb_zero_grid(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    b_zero_grid<D>(I[i].c, I[i].G, long(blockIdx.x) - m.first[i]);
}

template <int D, typename T>
__global__ void kb_euler(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    b_euler<D>(I[i].c, I[i].G, I[i].M, long(blockIdx.x) - m.first[i]);
}

template <int D, typename T>
__global__ void kb_augv(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    b_augv<D>(I[i].c, I[i].G, long(blockIdx.x) - m.first[i]);
}

template <int D, int B, typename T, bool AFF>
__global__ void __launch_bounds__(EP_PBLK_MAX) kb_p2g(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    const Inst<T> &in = I[i];
    b_p2g<D, B, T, AFF>(in.c, in.g, in.P, in.G, in.M, long(blockIdx.x) - m.first[i]);
}

template <int D, int B, typename T>
__global__ void __launch_bounds__(EP_PBLK_MAX) kb_g2p(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    const Inst<T> &in = I[i];
    b_g2p<D, B, T>(in.c, in.g, in.P, in.G, in.M, long(blockIdx.x) - m.first[i]);
}

template <int D, int B, typename T>
__global__ void __launch_bounds__(EP_PBLK_MAX) kb_bij(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    const Inst<T> &in = I[i];
    b_bij<D, B, T>(in.c, in.g, in.P, in.G, long(blockIdx.x) - m.first[i]);
}

template <int D, int B, typename T, bool LOCK>
__global__ void __launch_bounds__(EP_PBLK_MAX) kb_deform(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    const Inst<T> &in = I[i];
    b_deform<D, B, T, LOCK, false>(in.c, in.g, in.P, in.G, in.M, long(blockIdx.x) - m.first[i]);
}

template <int D, int B, typename T>
__global__ void __launch_bounds__(EP_PBLK_MAX) kb_fbar(const Inst<T> *I, BlockMap m) {
    const int i = m.inst[blockIdx.x];
    const Inst<T> &in = I[i];
    b_fbar<D, B, T, false>(in.c, in.g, in.P, in.G, in.M, long(blockIdx.x) - m.first[i]);
}

// step parameters of one ep_ensemble_step call, for every instance
struct StepPar {
    double T, G, tg, dt_fixed;
    int ramp, with_plast, dt_mode;
};
template <typename T>
__global__ void kb_params(const Inst<T> *I, int n, StepPar q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ctrl *c = I[i].c;
    c->T = q.T;
    c->G = q.G;
    c->tg = q.tg;
    c->dt_fixed = q.dt_fixed;
    c->ramp = q.ramp;
    c->with_plast = q.with_plast;
    c->dt_mode = q.dt_mode;
    c->record = 0;
    c->steps = 0;
    c->err = 0;
    c->j2cap = 0;
}
// per instance: t, dt of the last step, steps done, error bits, J2 cap hits
template <typename T>
__global__ void kb_collect(const Inst<T> *I, int n, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Ctrl *c = I[i].c;
    out[5 * i + 0] = c->t;
    out[5 * i + 1] = c->dt;
    out[5 * i + 2] = double(c->steps);
    out[5 * i + 3] = double(c->err);
    out[5 * i + 4] = double(c->j2cap);
}

}  // namespace ep
