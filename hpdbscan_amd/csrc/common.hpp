// Shared device-side definitions: element traits, the exact squared-distance
// test, and small wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace hpd {

constexpr int kWave = 64;

// --------------------------------------------------------------------------
// Squared distance in the reference's operation order.
//
// spatial_index.h:751-754 computes `offset += (p[d]-q[d])^2` in T.  The
// reference build (CMakeLists.txt:12, -Ofast, SSE2) vectorises the loop:
//   float : four partial sums a_j over d = j (mod 4), reduced as
//           (a0 + a2) + (a1 + a3), then the next two terms as one pair
//           (t + t'), then a last single term.  D < 4 adds left to right.
//   double: two partial sums over d = j (mod 2), reduced a0 + a1, then a
//           last single term.
//   integer types: every step wraps modulo 2^bits(T).
// Kernels are built with -ffp-contract=off, so no step is fused into an FMA.
// --------------------------------------------------------------------------
template <typename T>
struct Coord {
    using acc = T;
};

template <int D>
__device__ __forceinline__ float dist2(const float* p, const float* q) {
    float t[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const float u = p[j] - q[j];
        t[j] = u * u;
    }
    float s;
    int base;
    if constexpr (D >= 4) {
        float a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = t[j];
#pragma unroll
        for (int k = 1; k < D / 4; ++k) {
#pragma unroll
            for (int j = 0; j < 4; ++j) a[j] = a[j] + t[4 * k + j];
        }
        s = (a[0] + a[2]) + (a[1] + a[3]);
        base = (D / 4) * 4;
        if constexpr (D - (D / 4) * 4 >= 2) {
            s = s + (t[base] + t[base + 1]);
            base += 2;
        }
        if constexpr (D % 2 == 1) s = s + t[base];
    } else if constexpr (D == 3) {
        s = (t[0] + t[1]) + t[2];
        (void)base;
    } else if constexpr (D == 2) {
        s = t[0] + t[1];
        (void)base;
    } else {
        s = t[0];
        (void)base;
    }
    return s;
}

template <int D>
__device__ __forceinline__ double dist2(const double* p, const double* q) {
    double t[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double u = p[j] - q[j];
        t[j] = u * u;
    }
    if constexpr (D == 1) {
        return t[0];
    } else {
        double a0 = t[0], a1 = t[1];
#pragma unroll
        for (int k = 1; k < D / 2; ++k) {
            a0 = a0 + t[2 * k];
            a1 = a1 + t[2 * k + 1];
        }
        double s = a0 + a1;
        if constexpr (D % 2 == 1) s = s + t[D - 1];
        return s;
    }
}

template <typename T>
struct IntTraits {
    using U = std::make_unsigned_t<T>;
};

// Integer data: T distance = p - q; offset += distance * distance, in T.
template <int D, typename T>
__device__ __forceinline__ std::enable_if_t<std::is_integral<T>::value, T> dist2(const T* p,
                                                                                const T* q) {
    using U = typename IntTraits<T>::U;
    U s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const U t = static_cast<U>(static_cast<U>(p[j]) - static_cast<U>(q[j]));
        s = static_cast<U>(s + static_cast<U>(static_cast<uint64_t>(t) * static_cast<uint64_t>(t)));
    }
    return static_cast<T>(s);
}

// `offset <= EPS2` (spatial_index.h:756)
template <int D, typename T>
__device__ __forceinline__ bool within(const T* p, const T* q, T eps2) {
    return dist2<D>(p, q) <= eps2;
}

// --------------------------------------------------------------------------
// wave helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = __lane_id();
    return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}

// Adds v of every thread of the block to *out with one atomic per block
// (wave sums meet in LDS).  Every thread of the block must call it.  One
// atomic per wave on one address serialises a kernel of thousands of waves.
template <int kBlockThreads>
__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long* out) {
    constexpr int kWaves = kBlockThreads / 64;
    __shared__ unsigned long long part[kWaves];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane_id() == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += part[w];
        if (t) atomicAdd(out, t);
    }
}

// Two sums at once, to *out0 and *out1 (same rules as block_add).
template <int kBlockThreads>
__device__ __forceinline__ void block_add2(unsigned long long v0, unsigned long long v1,
                                           unsigned long long* out0, unsigned long long* out1) {
    constexpr int kWaves = kBlockThreads / 64;
    __shared__ unsigned long long part2[2][kWaves];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v0 += __shfl_xor(v0, o, 64);
        v1 += __shfl_xor(v1, o, 64);
    }
    if (lane_id() == 0) {
        part2[0][threadIdx.x >> 6] = v0;
        part2[1][threadIdx.x >> 6] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t0 = 0, t1 = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            t0 += part2[0][w];
            t1 += part2[1][w];
        }
        if (t0) atomicAdd(out0, t0);
        if (t1) atomicAdd(out1, t1);
    }
}

// --------------------------------------------------------------------------
// Lock-free union-find over sorted positions.  parent[x] <= x always; a root
// is hooked only under a smaller root, so every tree's root is its smallest
// member, which is exactly the reference's cluster id - 1.
// --------------------------------------------------------------------------
// Loads of the parent array bypass the per-CU cache (agent-scope relaxed
// atomics) so that a spinning thread always sees other CUs' hooks.
__device__ __forceinline__ int ld_agent(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_agent(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uf_find(int* parent, int x) {
    int cur = ld_agent(&parent[x]);
    if (cur != x) {
        int prev = x;
        int next;
        while (cur > (next = ld_agent(&parent[cur]))) {
            st_agent(&parent[prev], next);  // path halving: still an ancestor of prev
            prev = cur;
            cur = next;
        }
    }
    return cur;
}

__device__ __forceinline__ void uf_unite(int* parent, int a, int b) {
    int ra = uf_find(parent, a);
    int rb = uf_find(parent, b);
    while (ra != rb) {
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        // hook the larger root ra under rb
        const int old = atomicCAS(&parent[ra], ra, rb);
        if (old == ra) return;
        ra = uf_find(parent, old);
        rb = uf_find(parent, rb);
    }
}

}  // namespace hpd
