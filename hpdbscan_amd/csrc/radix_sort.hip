// Stable LSD radix sort of (cell id, row index) pairs, 8 or 9 bits per pass.
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
// inside a tile is (wave, iteration, lane), i.e. input order.  9-bit digits
// (512 buckets, two per thread) are used where they save a pass.
#include "common.hpp"
#include "hpdbscan_capi.h"
#include "internal.hpp"

namespace hpd {

namespace {

constexpr int kThreads = 256;
constexpr int kIters = 16;
constexpr int kTile = kThreads * kIters;  // 4096
// passes that carry rows stage them in LDS too: half the tile keeps two
// more blocks resident per CU
constexpr int kItersRows = 8;
constexpr int kTileMin = kThreads * kItersRows;
constexpr int kMaxBuckets = 512;          // 9-bit digits

__host__ __device__ inline int64_t tiles_of(int64_t n, int64_t tile = kTile) {
    return (n + tile - 1) / tile;
}

template <typename K, int kBits, int kIt = kIters>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const K* __restrict__ keys, int64_t n,
                                                         int shift, uint32_t* __restrict__ hist,
                                                         int64_t ntiles,
                                                         const int* __restrict__ bits_dev) {
    if (bits_dev && shift >= *bits_dev) return;  // a pass beyond the key's bits
    constexpr int B = 1 << kBits;
    constexpr uint32_t kMask = B - 1;
    __shared__ uint32_t cnt[B];
    for (int dg = threadIdx.x; dg < B; dg += kThreads) cnt[dg] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    constexpr int tile = kThreads * kIt;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * tile + w * (tile / 4);
#pragma unroll 4
    for (int it = 0; it < kIt; ++it) {
        const int64_t idx = base + it * 64 + l;
        if (idx < n) {
            const uint32_t dgt = static_cast<uint32_t>((keys[idx] >> shift) & kMask);
            atomicAdd(&cnt[dgt], 1u);
        }
    }
    __syncthreads();
    for (int dg = threadIdx.x; dg < B; dg += kThreads)
        hist[static_cast<int64_t>(dg) * ntiles + blockIdx.x] = cnt[dg];
}

// Rows that a pass carries along with the pairs: row i of `in` (one word W
// of D values of T) goes where pair i goes, into `out` (words), or on the
// last pass into D planes of stride ld (the coordinates in sorted order).
// The row is read with its key, staged in LDS with it and written with it.
struct NoRows {
    static constexpr bool kOn = false;
    using W = uint32_t;
    __device__ W load(int64_t) const { return 0; }
    __device__ void store(int64_t, W) const {}
};
template <typename T, int D>
struct MoveRows {
    static constexpr bool kOn = true;
    using W = std::conditional_t<sizeof(T) * D == 8, uint64_t, uint32_t>;
    static_assert(sizeof(T) * D == sizeof(W), "a row is one 4- or 8-byte word");
    const W* in;
    W* out;      // ld == 0
    T* planes;   // ld > 0
    int64_t ld;
    __device__ __forceinline__ W load(int64_t i) const { return in[i]; }
    __device__ __forceinline__ void store(int64_t to, W v) const {
        if (ld > 0) {
            T x[D];
            __builtin_memcpy(x, &v, sizeof(W));
#pragma unroll
            for (int j = 0; j < D; ++j) planes[static_cast<int64_t>(j) * ld + to] = x[j];
        } else {
            out[to] = v;
        }
    }
};

template <typename K, int kBits, typename Rows = NoRows, int kIt = kIters>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const K* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            K* __restrict__ kout,
                                                            uint32_t* __restrict__ vout, int64_t n,
                                                            int shift,
                                                            const uint32_t* __restrict__ offs,
                                                            int64_t ntiles,
                                                            const int* __restrict__ bits_dev,
                                                            Rows rows = Rows{}) {
    if (bits_dev && shift >= *bits_dev) return;  // a pass beyond the key's bits
    constexpr int B = 1 << kBits;
    constexpr uint32_t kMask = B - 1;
    constexpr int kOwn = B / kThreads;  // consecutive digits per thread
    __shared__ uint32_t cnt[4][B];
    __shared__ uint32_t gbase[B];
    __shared__ uint32_t tbase[B];  // digit's first place in the staged tile
    __shared__ uint32_t wsum[kThreads / 64];
    constexpr int tile = kThreads * kIt;
    __shared__ K skey[tile];
    __shared__ uint32_t sval[tile];
    __shared__ typename Rows::W srow[Rows::kOn ? tile : 1];
    const int w = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    for (int dg = threadIdx.x; dg < B; dg += kThreads) {
#pragma unroll
        for (int i = 0; i < 4; ++i) cnt[i][dg] = 0;
        gbase[dg] = offs[static_cast<int64_t>(dg) * ntiles + blockIdx.x];
    }
    __syncthreads();

    const int64_t base = static_cast<int64_t>(blockIdx.x) * tile + w * (tile / 4);
    const uint64_t lt = lanemask_lt();
    K key[kIt];
    uint32_t val[kIt];
    uint32_t rank[kIt];
    uint32_t dgt[kIt];
    typename Rows::W row[Rows::kOn ? kIt : 1];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
        const int64_t idx = base + it * 64 + l;
        const bool valid = idx < n;
        const K k = valid ? kin[idx] : K(0);
        const uint32_t d = static_cast<uint32_t>((k >> shift) & kMask);
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < kBits; ++b) {
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
        if constexpr (Rows::kOn) row[it] = valid ? rows.load(idx) : 0;
    }
    __syncthreads();
    // exclusive prefix of each digit over the four waves, and of the digits'
    // tile totals over the digits (thread t owns digits [t * kOwn, ..))
    uint32_t tot[kOwn];
    uint32_t run = 0;
#pragma unroll
    for (int o = 0; o < kOwn; ++o) {
        const int dg = threadIdx.x * kOwn + o;
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t c = cnt[i][dg];
            cnt[i][dg] = acc;
            acc += c;
        }
        tot[o] = acc;
        run += acc;
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
    uint32_t first = woff + inc - run;
#pragma unroll
    for (int o = 0; o < kOwn; ++o) {
        tbase[threadIdx.x * kOwn + o] = first;
        first += tot[o];
    }
    __syncthreads();
    // stage the tile in digit order (stable: waves, iterations and lanes are
    // in input order) ...
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
        if (dgt[it] != 0xffffffffu) {
            const uint32_t at = tbase[dgt[it]] + cnt[w][dgt[it]] + rank[it];
            skey[at] = key[it];
            sval[at] = val[it];
            if constexpr (Rows::kOn) srow[at] = row[it];
        }
    }
    __syncthreads();
    // ... and write it out: consecutive threads write consecutive places of
    // one digit's run
    const int64_t left = n - static_cast<int64_t>(blockIdx.x) * tile;
    const int valid_in_tile = left < tile ? static_cast<int>(left) : tile;
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
        const int at = it * kThreads + threadIdx.x;
        if (at < valid_in_tile) {
            const K k = skey[at];
            const uint32_t d = static_cast<uint32_t>((k >> shift) & kMask);
            const uint32_t dst = gbase[d] + (static_cast<uint32_t>(at) - tbase[d]);
            kout[dst] = k;
            vout[dst] = sval[at];
            if constexpr (Rows::kOn) rows.store(dst, srow[at]);
        }
    }
}

// After passes beyond a device-written bit count did nothing, the result of
// the `launched` passes lies in the b buffers when an odd number of them ran:
// copy it back to the a buffers.
template <typename K>
__global__ __launch_bounds__(kThreads) void k_radix_settle(K* __restrict__ keys_a,
                                                           uint32_t* __restrict__ vals_a,
                                                           const K* __restrict__ keys_b,
                                                           const uint32_t* __restrict__ vals_b,
                                                           int64_t n, int digit_bits,
                                                           const int* __restrict__ bits_dev) {
    const int ran = (*bits_dev + digit_bits - 1) / digit_bits;
    if ((ran & 1) == 0) return;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kThreads) {
        keys_a[i] = keys_b[i];
        vals_a[i] = vals_b[i];
    }
}

template <typename K, int kBits>
bool sort_passes(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n, int bits,
                 uint32_t* hist, uint32_t* tmp, hipStream_t s, const int* bits_dev = nullptr) {
    const int64_t nt = tiles_of(n);
    const int64_t nh = static_cast<int64_t>(1 << kBits) * nt;
    K* kin = keys_a;
    uint32_t* vin = vals_a;
    K* kout = keys_b;
    uint32_t* vout = vals_b;
    bool in_b = false;
    for (int shift = 0; shift < bits; shift += kBits) {
        hipLaunchKernelGGL((k_radix_hist<K, kBits>), dim3(static_cast<unsigned>(nt)),
                           dim3(kThreads), 0, s, kin, n, shift, hist, nt, bits_dev);
        exclusive_scan_u32(hist, hist, nh, tmp, nullptr, s);
        hipLaunchKernelGGL((k_radix_scatter<K, kBits>), dim3(static_cast<unsigned>(nt)),
                           dim3(kThreads), 0, s, kin, vin, kout, vout, n, shift, hist, nt,
                           bits_dev);
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

// sort_passes carrying rows: pass 0 reads them from `pts`, passes in
// between alternate between rows_a and rows_b, the last writes the planes
template <typename K, int kBits, typename T, int D>
bool sort_passes_rows(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                      int bits, uint32_t* hist, uint32_t* tmp, const T* pts, T* rows_a,
                      T* rows_b, T* soa, int64_t ld, hipStream_t s) {
    const int64_t nt = tiles_of(n, kTileMin);
    const int64_t nh = static_cast<int64_t>(1 << kBits) * nt;
    K* kin = keys_a;
    uint32_t* vin = vals_a;
    K* kout = keys_b;
    uint32_t* vout = vals_b;
    bool in_b = false;
    const T* rin = pts;
    T* rnext = rows_a;
    for (int shift = 0; shift < bits; shift += kBits) {
        const bool last = shift + kBits >= bits;
        using W = typename MoveRows<T, D>::W;
        MoveRows<T, D> mv{reinterpret_cast<const W*>(rin), reinterpret_cast<W*>(rnext), soa,
                          last ? ld : 0};
        hipLaunchKernelGGL((k_radix_hist<K, kBits, kItersRows>), dim3(static_cast<unsigned>(nt)),
                           dim3(kThreads), 0, s, kin, n, shift, hist, nt, nullptr);
        exclusive_scan_u32(hist, hist, nh, tmp, nullptr, s);
        hipLaunchKernelGGL((k_radix_scatter<K, kBits, MoveRows<T, D>, kItersRows>),
                           dim3(static_cast<unsigned>(nt)), dim3(kThreads), 0, s, kin, vin, kout,
                           vout, n, shift, hist, nt, nullptr, mv);
        HPD_HIP(hipGetLastError());
        K* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
        in_b = !in_b;
        rin = rnext;
        rnext = rnext == rows_a ? rows_b : rows_a;
    }
    return in_b;
}

}  // namespace

int radix_pass_count(int bits) {
    if (bits <= 0) return 0;
    const int digit = (bits + 8) / 9 < (bits + 7) / 8 ? 9 : 8;
    return (bits + digit - 1) / digit;
}

size_t radix_hist_words(int64_t n) {
    return static_cast<size_t>(kMaxBuckets) * static_cast<size_t>(tiles_of(n > 0 ? n : 1, kTileMin));
}

// 9-bit digits where they save a pass (e.g. 25-27 key bits: 3 passes, not 4)
template <typename K>
bool radix_sort_pairs(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                      int bits, uint32_t* hist, uint32_t* tmp, hipStream_t s) {
    if (n <= 1 || bits <= 0) return false;
    if ((bits + 8) / 9 < (bits + 7) / 8)
        return sort_passes<K, 9>(keys_a, vals_a, keys_b, vals_b, n, bits, hist, tmp, s);
    return sort_passes<K, 8>(keys_a, vals_a, keys_b, vals_b, n, bits, hist, tmp, s);
}

template <typename K, typename T, int D>
bool radix_sort_pairs_rows(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                           int bits, uint32_t* hist, uint32_t* tmp, const T* pts, T* rows_a,
                           T* rows_b, T* soa, int64_t ld, hipStream_t s) {
    if (n <= 1 || bits <= 0)
        throw Error(HPD_EINVAL, "radix_sort_pairs_rows needs n > 1 and bits > 0");
    if ((bits + 8) / 9 < (bits + 7) / 8)
        return sort_passes_rows<K, 9, T, D>(keys_a, vals_a, keys_b, vals_b, n, bits, hist, tmp,
                                            pts, rows_a, rows_b, soa, ld, s);
    return sort_passes_rows<K, 8, T, D>(keys_a, vals_a, keys_b, vals_b, n, bits, hist, tmp, pts,
                                        rows_a, rows_b, soa, ld, s);
}

#define HPD_INST_SORT_ROWS(K, T, D)                                                              \
    template bool radix_sort_pairs_rows<K, T, D>(K*, uint32_t*, K*, uint32_t*, int64_t, int,     \
                                                 uint32_t*, uint32_t*, const T*, T*, T*, T*,     \
                                                 int64_t, hipStream_t);
#define HPD_INST_SORT_ROWS_K(T, D) HPD_INST_SORT_ROWS(uint32_t, T, D) HPD_INST_SORT_ROWS(uint64_t, T, D)
HPD_INST_SORT_ROWS_K(float, 1)
HPD_INST_SORT_ROWS_K(float, 2)
HPD_INST_SORT_ROWS_K(double, 1)

template <typename K>
void radix_sort_pairs_dev(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                          int max_bits, const int* bits_dev, uint32_t* hist, uint32_t* tmp,
                          hipStream_t s) {
    if (n <= 1 || max_bits <= 0) return;
    sort_passes<K, 8>(keys_a, vals_a, keys_b, vals_b, n, max_bits, hist, tmp, s, bits_dev);
    const int64_t want = (n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_radix_settle<K>, dim3(static_cast<unsigned>(want < 4096 ? want : 4096)),
                       dim3(kThreads), 0, s, keys_a, vals_a, static_cast<const K*>(keys_b),
                       static_cast<const uint32_t*>(vals_b), n, 8, bits_dev);
    HPD_HIP(hipGetLastError());
}

template void radix_sort_pairs_dev<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t,
                                             int, const int*, uint32_t*, uint32_t*, hipStream_t);

template bool radix_sort_pairs<uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, int64_t, int,
                                         uint32_t*, uint32_t*, hipStream_t);
template bool radix_sort_pairs<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t, int,
                                         uint32_t*, uint32_t*, hipStream_t);

}  // namespace hpd
