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
#include <barrier>
#include <chrono>
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
#define __shared__ static

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct EmuIdx {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local EmuIdx threadIdx, blockIdx, blockDim;
struct EmuBlock {
    std::barrier<> *bar = nullptr;
    unsigned long long *shfl = nullptr;
};
inline thread_local EmuBlock emu_block;

inline void __syncthreads() { emu_block.bar->arrive_and_wait(); }

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
// Lane exchange inside a 64-wide wavefront; every thread of the block calls it.
inline unsigned long long __shfl_xor(unsigned long long v, int mask) {
    const unsigned t = threadIdx.x;
    emu_block.shfl[t] = v;
    __syncthreads();
    const unsigned src = (t & ~63u) | ((t & 63u) ^ unsigned(mask));
    const unsigned long long r = src < blockDim.x ? emu_block.shfl[src] : v;
    __syncthreads();
    return r;
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
inline const char *hipGetErrorString(hipError_t) { return "cpu check build"; }

// One persistent team of host threads per block size; the launching thread
// joins the team's start/done barriers, the team's own barrier serves
// __syncthreads.
struct EmuTeam {
    unsigned n;
    std::barrier<> start, done, bar;
    std::vector<unsigned long long> shfl;
    std::vector<std::thread> th;
    std::function<void(unsigned)> job;
    bool quit = false;
    explicit EmuTeam(unsigned n_) : n(n_), start(n_ + 1), done(n_ + 1), bar(n_), shfl(n_) {
        for (unsigned t = 0; t < n; t++)
            th.emplace_back([this, t] {
                for (;;) {
                    start.arrive_and_wait();
                    if (quit) return;
                    emu_block = {&bar, shfl.data()};
                    job(t);
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

template <class... P, class... A>
inline void emu_launch(void (*k)(P...), dim3 g, dim3 b, A... args) {
    EmuTeam &team = emu_team(b.x);
    for (unsigned bx = 0; bx < g.x; bx++) {
        team.run([&, bx](unsigned t) {
            threadIdx = {t, 0, 0};
            blockIdx = {bx, 0, 0};
            blockDim = {b.x, 1, 1};
            k(args...);
        });
    }
}
#define EP_LAUNCH(kern, grid, block, stream, ...) emu_launch(kern, dim3(grid), dim3(block), __VA_ARGS__)
