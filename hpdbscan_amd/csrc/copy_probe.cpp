// A plain device-to-device copy of a known size: the yardstick for reading
// the HBM traffic counters.  FETCH_SIZE and WRITE_SIZE of one dispatch of
// k_copy are compared with the bytes it must move (`bytes` each way), which
// gives the factor that turns the counters into bytes for the clustering
// kernels.  Also prints the copy bandwidth it reaches.
//
//     copy_probe [bytes = 1 GiB] [launches = 5]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src,
                                              uint4* __restrict__ dst, int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

int main(int argc, char** argv) {
    const int64_t bytes = argc > 1 ? std::atoll(argv[1]) : (int64_t(1) << 30);
    const int launches = argc > 2 ? std::atoi(argv[2]) : 5;
    if (bytes < 16 || bytes % 16 != 0 || launches < 1) {
        std::fprintf(stderr, "usage: copy_probe [bytes, a multiple of 16] [launches >= 1]\n");
        return 2;
    }
    const int64_t n = bytes / 16;
    uint4 *src = nullptr, *dst = nullptr;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMemset(src, 1, bytes));
    CHECK(hipMemset(dst, 0, bytes));
    const int64_t want = (n + 255) / 256;
    const unsigned grid = static_cast<unsigned>(want < 8192 ? want : 8192);
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, src, dst, n);  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(a, nullptr));
    for (int r = 0; r < launches; ++r)
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, src, dst, n);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b, nullptr));
    CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    const double per = ms / launches;
    std::printf("{\"kernel\": \"k_copy\", \"bytes_read\": %lld, \"bytes_written\": %lld, "
                "\"dispatches\": %d, \"ms_per_launch\": %.4f, \"GB_per_s\": %.1f}\n",
                static_cast<long long>(bytes), static_cast<long long>(bytes), launches + 1, per,
                2.0 * static_cast<double>(bytes) / (per * 1e-3) / 1e9);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    return 0;
}
