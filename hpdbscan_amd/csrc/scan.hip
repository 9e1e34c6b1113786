// Exclusive prefix sum over uint32 (reduce / scan-of-sums / downsweep).
// Used for the radix-sort digit offsets, the occupied-cell numbering and the
// CSR offsets of the neighbour ranges.
#include "common.hpp"
#include "internal.hpp"

namespace hpd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (l >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kScanThreads / 64];
    const int w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (__lane_id() == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        const uint32_t s = wsum[i];
        if (i < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              int64_t n,
                                                              uint32_t* __restrict__ sums) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t idx = base + static_cast<int64_t>(i) * kScanThreads + threadIdx.x;
        if (idx < n) acc += in[idx];
    }
    uint32_t tot;
    (void)block_exclusive_scan(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Scans one tile per block: thread t owns elements [t*16, t*16+16) of the
// tile, adds the tile offset (offs[blockIdx.x], or 0 when offs is null).
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* in, uint32_t* out,
                                                            int64_t n,
                                                            const uint32_t* __restrict__ offs) {
    __shared__ uint32_t tile[kScanTile];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int idx = i * kScanThreads + threadIdx.x;
        const int64_t g = base + idx;
        tile[idx] = g < n ? in[g] : 0u;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = tile[threadIdx.x * kScanItems + i];
        acc += v[i];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(acc, &tot) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        tile[threadIdx.x * kScanItems + i] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int idx = i * kScanThreads + threadIdx.x;
        const int64_t g = base + idx;
        if (g < n) out[g] = tile[idx];
    }
}

// Totals need the last input value, which an in-place scan overwrites; this
// kernel runs before the downsweep and stashes it.
__global__ void k_scan_stash_last(const uint32_t* __restrict__ in, int64_t n, uint32_t* stash) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *stash = n > 0 ? in[n - 1] : 0u;
}

__global__ void k_scan_finish_total(const uint32_t* __restrict__ out, int64_t n,
                                    const uint32_t* stash, uint32_t* total_dev) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total_dev = (n > 0 ? out[n - 1] : 0u) + *stash;
}

int64_t ntiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

}  // namespace

size_t scan_tmp_words(int64_t n) {
    size_t words = 4;  // stash + padding
    int64_t m = n;
    while (m > kScanTile) {
        m = ntiles(m);
        words += static_cast<size_t>(align_up(static_cast<size_t>(m), 64));
    }
    return words;
}

static void scan_rec(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, hipStream_t s) {
    if (n <= 0) return;
    if (n <= kScanTile) {
        hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                           static_cast<const uint32_t*>(nullptr));
        return;
    }
    const int64_t nb = ntiles(n);
    uint32_t* sums = tmp;
    uint32_t* rest = tmp + align_up(static_cast<size_t>(nb), 64);
    hipLaunchKernelGGL(k_scan_reduce, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in,
                       n, sums);
    scan_rec(sums, sums, nb, rest, s);
    hipLaunchKernelGGL(k_scan_down, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in,
                       out, n, static_cast<const uint32_t*>(sums));
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                        uint32_t* total_dev, hipStream_t s) {
    uint32_t* stash = tmp;
    if (total_dev) hipLaunchKernelGGL(k_scan_stash_last, dim3(1), dim3(64), 0, s, in, n, stash);
    scan_rec(in, out, n, tmp + 4, s);
    if (total_dev)
        hipLaunchKernelGGL(k_scan_finish_total, dim3(1), dim3(64), 0, s, out, n,
                           static_cast<const uint32_t*>(stash), total_dev);
    HPD_HIP(hipGetLastError());
}

}  // namespace hpd
