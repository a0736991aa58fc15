// Byte-counter calibration for rocprofv3 --pmc: one device-to-device copy of
// a known size (reads and writes 256 MiB once per launch, three launches).
// The FETCH_SIZE / WRITE_SIZE a profiler reports for copy_f4 divided by the
// bytes printed here is the scale factor of those counters on this GPU.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ __launch_bounds__(256) void copy_f4(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

int main()
{
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
    hipFree(a);
    hipFree(b);
    return 0;
}
