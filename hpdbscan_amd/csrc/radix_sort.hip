// Stable LSD radix sort of (cell id, row index) pairs, 8 bits per pass.
//
// Replaces the reference's cell histogram + serial counting-sort scatter
// (spatial_index.h:144-216).  Per pass:
//   1. k_radix_hist    per-tile digit histogram, stored digit-major
//                      hist[digit * ntiles + tile]
//   2. exclusive scan  of hist -> global offset of (digit, tile)
//   3. k_radix_scatter stable rank of every key inside its tile (wave
//                      ballots + per-wave LDS counters), the tile staged
//                      in LDS in digit order, then written out so that
//                      consecutive threads write consecutive places
// A tile is 4096 keys: 4 waves x 16 iterations x 64 lanes; the stable order
// inside a tile is (wave, iteration, lane), i.e. input order.
#include "common.hpp"
#include "internal.hpp"

namespace hpd {

namespace {

constexpr int kThreads = 256;
constexpr int kIters = 16;
constexpr int kTile = kThreads * kIters;  // 4096
constexpr int kBuckets = 256;

__host__ __device__ inline int64_t tiles_of(int64_t n) { return (n + kTile - 1) / kTile; }

template <typename K>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const K* __restrict__ keys, int64_t n,
                                                         int shift, uint32_t* __restrict__ hist,
                                                         int64_t ntiles) {
    __shared__ uint32_t cnt[kBuckets];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + w * (kTile / 4);
#pragma unroll 4
    for (int it = 0; it < kIters; ++it) {
        const int64_t idx = base + it * 64 + l;
        if (idx < n) {
            const uint32_t dgt = static_cast<uint32_t>((keys[idx] >> shift) & 0xff);
            atomicAdd(&cnt[dgt], 1u);
        }
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

template <typename K>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const K* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            K* __restrict__ kout,
                                                            uint32_t* __restrict__ vout, int64_t n,
                                                            int shift,
                                                            const uint32_t* __restrict__ offs,
                                                            int64_t ntiles) {
    __shared__ uint32_t cnt[4][kBuckets];
    __shared__ uint32_t gbase[kBuckets];
    __shared__ uint32_t tbase[kBuckets];  // digit's first place in the staged tile
    __shared__ uint32_t wsum[kThreads / 64];
    __shared__ K skey[kTile];
    __shared__ uint32_t sval[kTile];
    const int w = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 4; ++i) cnt[i][threadIdx.x] = 0;
    gbase[threadIdx.x] = offs[static_cast<int64_t>(threadIdx.x) * ntiles + blockIdx.x];
    __syncthreads();

    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + w * (kTile / 4);
    const uint64_t lt = lanemask_lt();
    K key[kIters];
    uint32_t val[kIters];
    uint32_t rank[kIters];
    uint32_t dgt[kIters];
#pragma unroll
    for (int it = 0; it < kIters; ++it) {
        const int64_t idx = base + it * 64 + l;
        const bool valid = idx < n;
        const K k = valid ? kin[idx] : K(0);
        const uint32_t d = static_cast<uint32_t>((k >> shift) & 0xff);
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t c = 0;
        if (valid) c = cnt[w][d];
        if (valid && (peers & lt) == 0) cnt[w][d] = c + static_cast<uint32_t>(__popcll(peers));
        rank[it] = c + static_cast<uint32_t>(__popcll(peers & lt));
        dgt[it] = valid ? d : 0xffffffffu;
        key[it] = k;
        val[it] = valid ? vin[idx] : 0u;
    }
    __syncthreads();
    // exclusive prefix of each digit over the four waves, and of the digits'
    // tile totals over the digits (thread t owns digit t)
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t c = cnt[i][threadIdx.x];
        cnt[i][threadIdx.x] = run;
        run += c;
    }
    uint32_t inc = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (l >= o) inc += t;
    }
    if (l == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t woff = 0;
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i)
        if (i < w) woff += wsum[i];
    tbase[threadIdx.x] = woff + inc - run;
    __syncthreads();
    // stage the tile in digit order (stable: waves, iterations and lanes are
    // in input order) ...
#pragma unroll
    for (int it = 0; it < kIters; ++it) {
        if (dgt[it] != 0xffffffffu) {
            const uint32_t at = tbase[dgt[it]] + cnt[w][dgt[it]] + rank[it];
            skey[at] = key[it];
            sval[at] = val[it];
        }
    }
    __syncthreads();
    // ... and write it out: consecutive threads write consecutive places of
    // one digit's run
    const int64_t left = n - static_cast<int64_t>(blockIdx.x) * kTile;
    const int valid_in_tile = left < kTile ? static_cast<int>(left) : kTile;
#pragma unroll
    for (int it = 0; it < kIters; ++it) {
        const int at = it * kThreads + threadIdx.x;
        if (at < valid_in_tile) {
            const K k = skey[at];
            const uint32_t d = static_cast<uint32_t>((k >> shift) & 0xff);
            const uint32_t dst = gbase[d] + (static_cast<uint32_t>(at) - tbase[d]);
            kout[dst] = k;
            vout[dst] = sval[at];
        }
    }
}

}  // namespace

size_t radix_hist_words(int64_t n) {
    return static_cast<size_t>(kBuckets) * static_cast<size_t>(tiles_of(n > 0 ? n : 1));
}

template <typename K>
bool radix_sort_pairs(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                      int bits, uint32_t* hist, uint32_t* tmp, hipStream_t s) {
    if (n <= 1 || bits <= 0) return false;
    const int64_t nt = tiles_of(n);
    const int64_t nh = static_cast<int64_t>(kBuckets) * nt;
    K* kin = keys_a;
    uint32_t* vin = vals_a;
    K* kout = keys_b;
    uint32_t* vout = vals_b;
    bool in_b = false;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist<K>, dim3(static_cast<unsigned>(nt)), dim3(kThreads), 0, s,
                           kin, n, shift, hist, nt);
        exclusive_scan_u32(hist, hist, nh, tmp, nullptr, s);
        hipLaunchKernelGGL(k_radix_scatter<K>, dim3(static_cast<unsigned>(nt)), dim3(kThreads), 0,
                           s, kin, vin, kout, vout, n, shift, hist, nt);
        HPD_HIP(hipGetLastError());
        K* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
        in_b = !in_b;
    }
    return in_b;
}

template bool radix_sort_pairs<uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, int64_t, int,
                                         uint32_t*, uint32_t*, hipStream_t);
template bool radix_sort_pairs<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t, int,
                                         uint32_t*, uint32_t*, hipStream_t);

}  // namespace hpd
