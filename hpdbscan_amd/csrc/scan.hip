// Exclusive prefix sum over uint32: a reduce of 4096-element tiles, then a
// downsweep whose blocks add their tile offsets (two launches up to 16.7M
// elements; beyond, the tile sums are scanned recursively in between).
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

// the length of a scan: n, or the device-written length capped at n
__device__ __forceinline__ int64_t scan_len(int64_t n, const uint32_t* n_dev) {
    if (!n_dev) return n;
    const int64_t v = static_cast<int64_t>(*n_dev);
    return v < n ? v : n;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              int64_t n,
                                                              uint32_t* __restrict__ sums,
                                                              const uint32_t* __restrict__ n_dev) {
    n = scan_len(n, n_dev);
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
// tile.  The tile's offset is offs[blockIdx.x] (offs = scanned tile sums),
// or, with sums instead, the block's own sum of sums[0, blockIdx.x) (the
// raw tile sums; used up to kScanTile tiles, where that is at most 16 loads
// per thread and saves the separate scan of the sums), or 0 when both are
// null.  The thread holding element n - 1 writes the total to total_dev
// (input and output may alias: every block reads its tile before writing).
// With total_at_len the total goes to out[n] instead (n: the scan's length,
// which may be device-written; block 0 writes 0 there for an empty scan).
__device__ __forceinline__ void scan_tile(const uint32_t* in, uint32_t* out, int64_t n,
                                          const uint32_t* __restrict__ offs,
                                          const uint32_t* __restrict__ sums,
                                          uint32_t* __restrict__ total_dev, uint32_t blk,
                                          bool total_at_len = false) {
    __shared__ uint32_t tile[kScanTile];
    if (total_at_len) {
        total_dev = out + n;
        if (n == 0) {
            if (blk == 0 && threadIdx.x == 0) *total_dev = 0u;
            return;
        }
    }
    const int64_t base = static_cast<int64_t>(blk) * kScanTile;
    uint32_t off = offs ? offs[blk] : 0u;
    if (sums) {
        uint32_t acc = 0;
        for (uint32_t j = threadIdx.x; j < blk; j += kScanThreads) acc += sums[j];
        (void)block_exclusive_scan(acc, &off);
    }
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
    uint32_t run = block_exclusive_scan(acc, &tot) + off;
    const int64_t last = n - 1 - base;  // tile index of element n - 1
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int idx = static_cast<int>(threadIdx.x) * kScanItems + i;
        tile[idx] = run;
        if (total_dev && idx == last) *total_dev = run + v[i];
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

// Scans one tile per block: thread t owns elements [t*16, t*16+16) of the
// tile.  The tile's offset is offs[blockIdx.x] (offs = scanned tile sums),
// or, with sums instead, the block's own sum of sums[0, blockIdx.x) (the
// raw tile sums; used up to kScanTile tiles, where that is at most 16 loads
// per thread and saves the separate scan of the sums), or 0 when both are
// null.  The thread holding element n - 1 writes the total to total_dev
// (input and output may alias: every block reads its tile before writing).
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* in, uint32_t* out,
                                                            int64_t n,
                                                            const uint32_t* __restrict__ offs,
                                                            const uint32_t* __restrict__ sums,
                                                            uint32_t* __restrict__ total_dev,
                                                            const uint32_t* __restrict__ n_dev,
                                                            bool total_at_len) {
    scan_tile(in, out, scan_len(n, n_dev), offs, sums, total_dev, blockIdx.x, total_at_len);
}

// In-place scans of several arrays of one length n (array blockIdx.y), the
// total of each to its entry n.
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce_set(ScanSet set, int64_t n,
                                                                  uint32_t* __restrict__ sums,
                                                                  int64_t stride,
                                                                  const uint32_t* __restrict__ n_dev) {
    n = scan_len(n, n_dev);
    const uint32_t* in = set.a[blockIdx.y];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t idx = base + static_cast<int64_t>(i) * kScanThreads + threadIdx.x;
        if (idx < n) acc += in[idx];
    }
    uint32_t tot;
    (void)block_exclusive_scan(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.y * stride + blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down_set(ScanSet set, int64_t n,
                                                                const uint32_t* __restrict__ sums,
                                                                int64_t stride,
                                                                const uint32_t* __restrict__ n_dev) {
    uint32_t* a = set.a[blockIdx.y];
    scan_tile(a, a, scan_len(n, n_dev), nullptr, sums ? sums + blockIdx.y * stride : nullptr,
              nullptr, blockIdx.x, true);
}

__global__ void k_scan_zero_total(uint32_t* total_dev) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total_dev = 0u;
}

__global__ void k_copy_at_count(const uint32_t* __restrict__ a, const uint32_t* __restrict__ idx,
                                uint32_t* __restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = a[*idx];
}

int64_t ntiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

}  // namespace

size_t scan_tmp_words(int64_t n) {
    size_t words = 4;  // padding
    int64_t m = n;
    while (m > kScanTile) {
        m = ntiles(m);
        words += static_cast<size_t>(align_up(static_cast<size_t>(m), 64));
    }
    return words;
}

static void scan_rec(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                     uint32_t* total_dev, hipStream_t s, const uint32_t* n_dev = nullptr,
                     bool total_at_len = false) {
    const uint32_t* none = nullptr;
    if (n <= 0) {
        uint32_t* t = total_at_len ? out : total_dev;
        if (t) hipLaunchKernelGGL(k_scan_zero_total, dim3(1), dim3(64), 0, s, t);
        return;
    }
    if (n <= kScanTile) {
        hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(kScanThreads), 0, s, in, out, n, none, none,
                           total_dev, n_dev, total_at_len);
        return;
    }
    const int64_t nb = ntiles(n);
    uint32_t* sums = tmp;
    uint32_t* rest = tmp + align_up(static_cast<size_t>(nb), 64);
    hipLaunchKernelGGL(k_scan_reduce, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in,
                       n, sums, n_dev);
    if (nb <= kScanTile) {
        // each block sums its predecessors' tile sums itself
        hipLaunchKernelGGL(k_scan_down, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s,
                           in, out, n, none, static_cast<const uint32_t*>(sums), total_dev, n_dev,
                           total_at_len);
        return;
    }
    // tiles past a device-written length have zero sums
    scan_rec(sums, sums, nb, rest, nullptr, s);
    hipLaunchKernelGGL(k_scan_down, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in,
                       out, n, static_cast<const uint32_t*>(sums), none, total_dev, n_dev,
                       total_at_len);
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                        uint32_t* total_dev, hipStream_t s) {
    scan_rec(in, out, n, tmp + 4, total_dev, s);
    HPD_HIP(hipGetLastError());
}

void exclusive_scan_u32_dev(const uint32_t* in, uint32_t* out, Count n, uint32_t* tmp,
                            hipStream_t s) {
    scan_rec(in, out, n.host, tmp + 4, nullptr, s, n.dev, true);
    HPD_HIP(hipGetLastError());
}

namespace {
__global__ void k_copy_words(const WordCopies c) {
    if (static_cast<int>(threadIdx.x) < c.k) *c.dst[threadIdx.x] = *c.src[threadIdx.x];
}
}  // namespace

void copy_words(const WordCopies& c, hipStream_t s) {
    if (c.k <= 0) return;
    hipLaunchKernelGGL(k_copy_words, dim3(1), dim3(64), 0, s, c);
    HPD_HIP(hipGetLastError());
}

void copy_at_count(const uint32_t* a, const uint32_t* idx_dev, uint32_t* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_at_count, dim3(1), dim3(64), 0, s, a, idx_dev, dst);
    HPD_HIP(hipGetLastError());
}

void exclusive_scan_set_u32(const ScanSet& set, int k, int64_t n, uint32_t* tmp, size_t tmp_words,
                            hipStream_t s, const uint32_t* n_dev) {
    const int64_t nb = ntiles(n);
    if (n <= 0 || k <= 0 || k > kMaxScanSet || nb > kScanTile ||
        static_cast<size_t>(k) * align_up(static_cast<size_t>(nb), 64) + 4 > tmp_words) {
        for (int i = 0; i < k; ++i) {
            if (n_dev)
                scan_rec(set.a[i], set.a[i], n, tmp + 4, nullptr, s, n_dev, true);
            else
                exclusive_scan_u32(set.a[i], set.a[i], n, tmp, set.a[i] + n, s);
        }
        HPD_HIP(hipGetLastError());
        return;
    }
    const int64_t stride = static_cast<int64_t>(align_up(static_cast<size_t>(nb), 64));
    uint32_t* sums = tmp + 4;
    if (nb > 1)
        hipLaunchKernelGGL(k_scan_reduce_set, dim3(static_cast<unsigned>(nb), k), dim3(kScanThreads),
                           0, s, set, n, sums, stride, n_dev);
    hipLaunchKernelGGL(k_scan_down_set, dim3(static_cast<unsigned>(nb), k), dim3(kScanThreads), 0, s,
                       set, n, static_cast<const uint32_t*>(nb > 1 ? sums : nullptr), stride, n_dev);
    HPD_HIP(hipGetLastError());
}

}  // namespace hpd
