// copy_probe.hip -- a plain device-to-device copy of a known size, used to
// calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters against exact byte
// counts before they are applied to the step kernels.
//   copy_probe <MiB> <repeats>
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void copy_f64(const double *__restrict__ a, double *__restrict__ b, long n) {
    for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) b[i] = a[i];
}

int main(int argc, char **argv) {
    const long mib = argc > 1 ? atol(argv[1]) : 512;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const long n = mib * 1024 * 1024 / 8;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) return 1;
    if (hipMemset(a, 0, n * 8) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
        if (hipEventRecord(e0, 0) != hipSuccess) return 1;
        hipLaunchKernelGGL(copy_f64, dim3(8192), dim3(256), 0, 0, a, b, n);
        if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        best = ms < best ? ms : best;
    }
    printf("copy_probe: %ld bytes read + %ld written per launch, best %.3f ms, %.1f GB/s\n", n * 8, n * 8, best,
           2.0 * n * 8 / (best * 1e-3) / 1e9);
    return hipFree(a) == hipSuccess && hipFree(b) == hipSuccess ? 0 : 1;
}
