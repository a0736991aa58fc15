// Byte-counter calibration for rocprofv3 --pmc: one device-to-device copy of
// a known size (reads and writes 256 MiB once per launch, three launches).
// The FETCH_SIZE / WRITE_SIZE a profiler reports for copy_f4 divided by the
// bytes printed here is the scale factor of those counters on this GPU.
//
// With the argument "bw" it instead measures copy bandwidth: copy_f4 over
// 16 MiB .. 2 GiB (20 back-to-back launches per size, hipEvents), i.e. the
// ceiling a streaming kernel reaches while its working set fits the last-level
// cache and beyond it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

__global__ __launch_bounds__(256) void copy_f4(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

static int bandwidth()
{
    const long long maxb = 2048LL << 20;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, maxb) != hipSuccess || hipMalloc(&b, maxb) != hipSuccess) {
        std::printf("allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, maxb) != hipSuccess || hipMemset(b, 0, maxb) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    std::printf("| copied [MiB] | us per launch | GB/s (read + written) |\n|---:|---:|---:|\n");
    for (long long mb = 16; mb <= 2048; mb *= 2) {
        const long long n = (mb << 20) / 16;
        const int reps = 20;
        for (int k = 0; k < 3; k++) copy_f4<<<1024, 256>>>(a, b, n);
        if (hipEventRecord(e0) != hipSuccess) return 1;
        for (int k = 0; k < reps; k++) copy_f4<<<1024, 256>>>(a, b, n);
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        const double us = ms * 1e3 / reps;
        std::printf("| %lld | %.1f | %.0f |\n", mb, us, 2.0 * (double)(mb << 20) / us / 1e3);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && std::strcmp(argv[1], "bw") == 0) return bandwidth();
    const long long bytes = 256LL << 20, n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::printf("allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
    for (int k = 0; k < 3; k++) copy_f4<<<4096, 256>>>(a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::printf("copy failed\n");
        return 1;
    }
    std::printf("copy_f4: %lld bytes read and %lld bytes written per launch, 3 launches\n", bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
