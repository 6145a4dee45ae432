// ep_cpu_check.h -- CPU checking build of the HIP sources (tests only).
//
// Compiled with g++ -DEP_CPU_CHECK (and AddressSanitizer) instead of hipcc,
// the kernels run on host threads: the blocks of a launch one after the
// other, the threads of a block as std::threads that meet at a std::barrier
// for __syncthreads, device memory on the host heap, atomics through
// std::atomic_ref.  It exists to find out-of-bounds accesses, races on
// block-local memory and wrong launch shapes on the CPU before a kernel
// reaches the GPU, and to run the parity tests without a GPU.  It is never
// part of libelastoplasm_hip.so.
#pragma once
#include <atomic>
#include <algorithm>
#include <barrier>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using std::isfinite;
using std::isnan;

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#ifdef EP_CPU_FIBERS
// Fiber variant (-DEP_CPU_FIBERS, no AddressSanitizer): the threads of a block
// are fibers of one host thread that switch at __syncthreads, and blocks run
// in parallel on a pool of host threads, so block-local memory is per host
// thread.  Fast enough for whole trajectories of long runs on the CPU.
#define __shared__ static thread_local
#else
#define __shared__ static
#endif

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct EmuIdx {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local EmuIdx threadIdx, blockIdx, blockDim, gridDim;
struct EmuBlock {
    std::barrier<> *bar = nullptr;
    unsigned long long *shfl = nullptr;
    unsigned char *alive = nullptr;
    unsigned long long *mf_a = nullptr, *mf_b = nullptr;
};
inline thread_local EmuBlock emu_block;

// A thread that has returned from the kernel drops out of later barriers, as
// an exited wave does on the GPU.
#ifdef EP_CPU_FIBERS
inline void emu_fiber_yield();
inline void __syncthreads() { emu_fiber_yield(); }
#else
inline void __syncthreads() { emu_block.bar->arrive_and_wait(); }
#endif
// device math the host library lacks
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
inline float rsqrt(float x) { return 1.0f / std::sqrt(x); }

inline void unsafeAtomicAdd(double *a, double v) { std::atomic_ref<double>(*a).fetch_add(v); }
inline void unsafeAtomicAdd(float *a, float v) { std::atomic_ref<float>(*a).fetch_add(v); }
inline double atomicAdd(double *a, double v) { return std::atomic_ref<double>(*a).fetch_add(v); }
inline float atomicAdd(float *a, float v) { return std::atomic_ref<float>(*a).fetch_add(v); }
inline int atomicAdd(int *a, int v) { return std::atomic_ref<int>(*a).fetch_add(v); }
inline unsigned atomicAdd(unsigned *a, unsigned v) { return std::atomic_ref<unsigned>(*a).fetch_add(v); }
inline unsigned long long atomicAdd(unsigned long long *a, unsigned long long v) {
    return std::atomic_ref<unsigned long long>(*a).fetch_add(v);
}
template <class T>
inline T emu_atomic_min(T *a, T v) {
    std::atomic_ref<T> r(*a);
    T o = r.load();
    while (v < o && !r.compare_exchange_weak(o, v)) {
    }
    return o;
}
template <class T>
inline T emu_atomic_max(T *a, T v) {
    std::atomic_ref<T> r(*a);
    T o = r.load();
    while (v > o && !r.compare_exchange_weak(o, v)) {
    }
    return o;
}
inline int atomicMin(int *a, int v) { return emu_atomic_min(a, v); }
inline int atomicMax(int *a, int v) { return emu_atomic_max(a, v); }
inline unsigned long long atomicMax(unsigned long long *a, unsigned long long v) { return emu_atomic_max(a, v); }
inline unsigned atomicOr(unsigned *a, unsigned v) { return std::atomic_ref<unsigned>(*a).fetch_or(v); }
inline unsigned atomicMax(unsigned *a, unsigned v) { return emu_atomic_max(a, v); }
typedef unsigned v4u __attribute__((vector_size(16)));
inline long long __double_as_longlong(double d) {
    long long r;
    std::memcpy(&r, &d, 8);
    return r;
}
inline double __longlong_as_double(long long l) {
    double r;
    std::memcpy(&r, &l, 8);
    return r;
}
// Lane exchange inside a 64-wide wavefront; every running thread of the block
// calls it (the kernels that use the wave intrinsics below run one wave per
// block in this build).
inline unsigned long long __shfl_xor(unsigned long long v, int mask) {
    const unsigned t = threadIdx.x;
    emu_block.shfl[t] = v;
    __syncthreads();
    const unsigned src = (t & ~63u) | ((t & 63u) ^ unsigned(mask));
    const unsigned long long r = src < blockDim.x ? emu_block.shfl[src] : v;
    __syncthreads();
    return r;
}
inline unsigned __lane_id() { return threadIdx.x & 63u; }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __ffsll(unsigned long long v) { return __builtin_ffsll((long long)v); }
// Lanes that have exited contribute 0, as on the GPU.
inline unsigned long long __ballot(int pred) {
    const unsigned t = threadIdx.x;
    emu_block.shfl[t] = pred ? 1ull : 0ull;
    __syncthreads();
    const unsigned base = t & ~63u;
    unsigned long long m = 0;
    for (unsigned w = 0; w < 64 && base + w < blockDim.x; w++)
        if (emu_block.alive[base + w] && emu_block.shfl[base + w]) m |= 1ull << w;
    __syncthreads();
    return m;
}
template <class T>
inline T __shfl(T v, int src) {
    static_assert(sizeof(T) <= 8, "shfl of more than 8 bytes");
    const unsigned t = threadIdx.x;
    unsigned long long u = 0;
    std::memcpy(&u, &v, sizeof(T));
    emu_block.shfl[t] = u;
    __syncthreads();
    const unsigned long long r = emu_block.shfl[(t & ~63u) + (unsigned(src) & 63u)];
    __syncthreads();
    T out;
    std::memcpy(&out, &r, sizeof(T));
    return out;
}

// v_mfma_f32_32x32x8_bf16 on one wave: lane l holds A[i = l%32][k = 4(l/32)+v]
// and B[k = 4(l/32)+v][j = l%32] (v = 0..3); D[i][j] is returned in lane
// j + 32*((i/4)%2), element 4*(i/8) + i%4.
typedef short v4s __attribute__((vector_size(8)));
typedef float v16f __attribute__((vector_size(64)));
inline float emu_bf16(short h) {
    const unsigned u = unsigned((unsigned short)h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline v16f emu_mfma_32x32x8bf16(v4s a, v4s b, v16f c) {
    const unsigned t = threadIdx.x;
    std::memcpy(&emu_block.mf_a[t], &a, 8);
    std::memcpy(&emu_block.mf_b[t], &b, 8);
    __syncthreads();
    const unsigned base = t & ~63u, lane = t & 63u, j = lane & 31u, hh = lane >> 5;
    v16f d;
    for (int v = 0; v < 16; v++) {
        const unsigned i = 8u * unsigned(v >> 2) + 4u * hh + unsigned(v & 3);
        float acc = 0.f;
        for (unsigned k = 0; k < 8; k++) {
            short as[4], bs[4];
            std::memcpy(as, &emu_block.mf_a[base + i + 32u * (k >> 2)], 8);
            std::memcpy(bs, &emu_block.mf_b[base + j + 32u * (k >> 2)], 8);
            acc += emu_bf16(as[k & 3]) * emu_bf16(bs[k & 3]);
        }
        d[v] = c[v] + acc;
    }
    __syncthreads();
    return d;
}
// v_mfma_f32_32x32x16_bf16: lane l holds A[i = l%32][k = 8(l/32)+v] and
// B[k = 8(l/32)+v][j = l%32] (v = 0..7); D as for the 32x32x8 form.  The
// same sum as two 32x32x8 products over k = 8(l/32)+{0..3} and +{4..7}.
typedef short v8s __attribute__((vector_size(16)));
inline v16f emu_mfma_32x32x16bf16(v8s a, v8s b, v16f c) {
    v4s alo, ahi, blo, bhi;
    std::memcpy(&alo, &a, 8);
    std::memcpy(&ahi, reinterpret_cast<const char *>(&a) + 8, 8);
    std::memcpy(&blo, &b, 8);
    std::memcpy(&bhi, reinterpret_cast<const char *>(&b) + 8, 8);
    return emu_mfma_32x32x8bf16(ahi, bhi, emu_mfma_32x32x8bf16(alo, blo, c));
}

// ---- runtime ----
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
typedef void *hipStream_t;
struct EmuEvent {
    double t = 0.0;
};
typedef EmuEvent *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
#define hipStreamNonBlocking 1u

template <class T>
inline hipError_t hipMalloc(T **p, size_t n) {
    *p = static_cast<T *>(std::malloc(n ? n : 1));
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) {
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = reinterpret_cast<hipStream_t>(1);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline double emu_now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline hipError_t hipEventCreate(hipEvent_t *e) {
    *e = new EmuEvent();
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    e->t = emu_now_ms();
    return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    *ms = float(b->t - a->t);
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
// static device arrays are host arrays here; their "device address" is their address
#define __align__(n) alignas(n)
#define HIP_SYMBOL(X) X
template <class T>
inline hipError_t hipGetSymbolAddress(void **p, const T &sym) {
    *p = const_cast<void *>(static_cast<const void *>(&sym));
    return hipSuccess;
}
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
enum hipMemoryType { hipMemoryTypeUnregistered = 0, hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2,
                     hipMemoryTypeManaged = 3 };
struct hipPointerAttribute_t {
    hipMemoryType type;
};
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *) {
    a->type = hipMemoryTypeDevice;   // every buffer of this build is host memory standing in for device memory
    return hipSuccess;
}
inline const char *hipGetErrorString(hipError_t) { return "cpu check build"; }

// One persistent team of host threads per block size; the launching thread
// joins the team's start/done barriers, the team's own barrier serves
// __syncthreads.
struct EmuTeam {
    unsigned n;
    std::barrier<> start, done;
    std::unique_ptr<std::barrier<>> bar;   // fresh per block: exited threads drop out of it
    std::vector<unsigned long long> shfl, mf_a, mf_b;
    std::vector<unsigned char> alive;
    std::vector<std::thread> th;
    std::function<void(unsigned)> job;
    bool quit = false;
    explicit EmuTeam(unsigned n_)
        : n(n_), start(n_ + 1), done(n_ + 1), shfl(n_), mf_a(n_), mf_b(n_), alive(n_) {
        for (unsigned t = 0; t < n; t++)
            th.emplace_back([this, t] {
                for (;;) {
                    start.arrive_and_wait();
                    if (quit) return;
                    emu_block = {bar.get(), shfl.data(), alive.data(), mf_a.data(), mf_b.data()};
                    job(t);
                    alive[t] = 0;
                    bar->arrive_and_drop();
                    done.arrive_and_wait();
                }
            });
    }
    ~EmuTeam() {
        quit = true;
        start.arrive_and_wait();
        for (auto &x : th) x.join();
    }
    void run(const std::function<void(unsigned)> &f) {
        job = f;
        bar.reset(new std::barrier<>(n));
        std::fill(alive.begin(), alive.end(), 1);
        start.arrive_and_wait();
        done.arrive_and_wait();
    }
};

inline EmuTeam &emu_team(unsigned n) {
    static std::mutex mu;
    static std::map<unsigned, std::unique_ptr<EmuTeam>> teams;
    std::lock_guard<std::mutex> lk(mu);
    auto &t = teams[n];
    if (!t) t.reset(new EmuTeam(n));
    return *t;
}

// Graph capture: launches on a capturing thread are recorded and replayed
// by hipGraphLaunch.
struct EmuGraph {
    std::vector<std::function<void()>> nodes;
};
typedef EmuGraph *hipGraph_t;
typedef EmuGraph *hipGraphExec_t;
typedef void *hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline thread_local EmuGraph *emu_capture = nullptr;
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    emu_capture = new EmuGraph();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    *g = emu_capture;
    emu_capture = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    *e = new EmuGraph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto &f : e->nodes) f();
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }

template <class... P, class... A>
inline void emu_launch_now(void (*k)(P...), dim3 g, dim3 b, A... args);

template <class... P, class... A>
inline void emu_launch(void (*k)(P...), dim3 g, dim3 b, A... args) {
    if (emu_capture) {
        emu_capture->nodes.push_back([=] { emu_launch_now(k, g, b, args...); });
        return;
    }
    emu_launch_now(k, g, b, args...);
}

#ifdef EP_CPU_FIBERS
// ---- fibers: a block's threads on one host thread ----
// emu_ctx_swap saves the callee-saved registers on the current stack, stores
// the stack pointer in *save and resumes the stack `load` (x86-64 SysV).
extern "C" void emu_ctx_swap(void **save, void *load);
asm(R"(
    .text
    .globl emu_ctx_swap
    .type emu_ctx_swap, @function
emu_ctx_swap:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_ctx_swap, .-emu_ctx_swap
)");

struct EmuFiberBlock {
    static constexpr size_t kStack = 512 * 1024;
    unsigned n = 0;
    std::vector<std::unique_ptr<char[]>> stack;
    std::vector<void *> sp;
    void *sched = nullptr;
    unsigned cur = 0;
    const std::function<void(unsigned)> *job = nullptr;
    std::vector<unsigned long long> shfl, mf_a, mf_b;
    std::vector<unsigned char> alive;
    std::vector<unsigned> order;
    void size(unsigned n_) {
        if (n_ == n) return;
        n = n_;
        stack.clear();
        for (unsigned t = 0; t < n; t++) stack.emplace_back(new char[kStack]);
        sp.assign(n, nullptr);
        shfl.assign(n, 0);
        mf_a.assign(n, 0);
        mf_b.assign(n, 0);
        alive.assign(n, 0);
        order.resize(n);
    }
};
inline thread_local EmuFiberBlock *emu_fb = nullptr;

[[noreturn]] inline void emu_fiber_entry() {
    EmuFiberBlock *f = emu_fb;
    const unsigned t = f->cur;
    (*f->job)(t);
    f->alive[t] = 0;
    void *dead;
    emu_ctx_swap(&dead, f->sched);
    std::abort();   // never resumed
}

inline void emu_fiber_yield() {
    EmuFiberBlock *f = emu_fb;
    emu_ctx_swap(&f->sp[f->cur], f->sched);
}

// Run one block: every live fiber runs to its next __syncthreads (or its
// end) in each pass, in the order `order`, so a pass is one barrier phase.
inline void emu_fiber_run(EmuFiberBlock &f, const std::function<void(unsigned)> &job, unsigned long long seed) {
    f.job = &job;
    for (unsigned t = 0; t < f.n; t++) {
        // initial frame: six zero registers, then emu_fiber_entry as the return address
        char *top = f.stack[t].get() + EmuFiberBlock::kStack;
        top = reinterpret_cast<char *>(reinterpret_cast<uintptr_t>(top) & ~uintptr_t(15));
        void **s = reinterpret_cast<void **>(top);
        *--s = nullptr;
        *--s = reinterpret_cast<void *>(&emu_fiber_entry);
        for (int r = 0; r < 6; r++) *--s = nullptr;
        f.sp[t] = s;
        f.alive[t] = 1;
        f.order[t] = t;
    }
    if (seed) {
        // a seeded permutation of the fibers: another order of block-local atomics
        unsigned long long z = seed;
        for (unsigned t = f.n - 1; t > 0; t--) {
            z += 0x9E3779B97F4A7C15ull;
            unsigned long long y = z;
            y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
            y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
            y ^= y >> 31;
            std::swap(f.order[t], f.order[unsigned(y % (t + 1))]);
        }
    }
    emu_fb = &f;
    bool any = true;
    while (any) {
        any = false;
        for (unsigned k = 0; k < f.n; k++) {
            const unsigned t = f.order[k];
            if (!f.alive[t]) continue;
            f.cur = t;
            threadIdx = {t, 0, 0};
            emu_block = {nullptr, f.shfl.data(), f.alive.data(), f.mf_a.data(), f.mf_b.data()};
            emu_ctx_swap(&f.sched, f.sp[t]);
            if (f.alive[t]) any = true;
        }
    }
}

// Host-thread pool: the blocks of a launch are shared out dynamically.
struct EmuPool {
    unsigned nw;
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    unsigned long long gen = 0;
    unsigned busy = 0;
    bool quit = false;
    std::function<void(unsigned)> work;   // argument: worker id
    explicit EmuPool(unsigned n) : nw(n) {
        for (unsigned w = 1; w < nw; w++)
            th.emplace_back([this, w] {
                unsigned long long seen = 0;
                for (;;) {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return quit || gen != seen; });
                    if (quit) return;
                    seen = gen;
                    lk.unlock();
                    work(w);
                    lk.lock();
                    if (--busy == 0) cv_done.notify_all();
                }
            });
    }
    ~EmuPool() {
        {
            std::lock_guard<std::mutex> lk(mu);
            quit = true;
        }
        cv.notify_all();
        for (auto &x : th) x.join();
    }
    void run(std::function<void(unsigned)> f) {
        {
            std::lock_guard<std::mutex> lk(mu);
            work = std::move(f);
            busy = nw - 1;
            gen++;
        }
        cv.notify_all();
        work(0);
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return busy == 0; });
    }
};
inline EmuPool &emu_pool() {
    static EmuPool p([] {
        const char *e = std::getenv("EP_EMU_THREADS");
        const int n = e ? std::atoi(e) : int(std::thread::hardware_concurrency());
        return unsigned(n < 1 ? 1 : n);
    }());
    return p;
}
inline std::atomic<unsigned long long> emu_launch_count{0};

template <class... P, class... A>
inline void emu_launch_now(void (*k)(P...), dim3 g, dim3 b, A... args) {
    static thread_local EmuFiberBlock fb;
    // EP_EMU_SEED: non-zero permutes each block's fiber order (a different
    // order of block-local atomics per block and launch)
    static const unsigned long long seed0 = [] {
        const char *e = std::getenv("EP_EMU_SEED");
        return e ? std::strtoull(e, nullptr, 10) : 0ull;
    }();
    const unsigned long long launch = emu_launch_count.fetch_add(1);
    std::atomic<unsigned> next{0};
    emu_pool().run([&](unsigned) {
        fb.size(b.x);
        for (;;) {
            const unsigned bx = next.fetch_add(1);
            if (bx >= g.x) break;
            blockIdx = {bx, 0, 0};
            blockDim = {b.x, 1, 1};
            gridDim = {g.x, 1, 1};
            const unsigned long long s = seed0 ? (seed0 * 0x100000001B3ull) ^ (launch << 24) ^ bx : 0ull;
            emu_fiber_run(fb, [&](unsigned) { k(args...); }, s);
        }
    });
}
#else
template <class... P, class... A>
inline void emu_launch_now(void (*k)(P...), dim3 g, dim3 b, A... args) {
    EmuTeam &team = emu_team(b.x);
    for (unsigned bx = 0; bx < g.x; bx++) {
        team.run([&, bx](unsigned t) {
            threadIdx = {t, 0, 0};
            blockIdx = {bx, 0, 0};
            blockDim = {b.x, 1, 1};
            gridDim = {g.x, 1, 1};
            k(args...);
        });
    }
}
#endif
#define EP_LAUNCH(kern, grid, block, stream, ...) emu_launch(kern, dim3(grid), dim3(block), __VA_ARGS__)
