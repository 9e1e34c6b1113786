// eps-grid construction on the device:
//   k_minmax_partial / k_minmax_final  per-dimension min and max
//                                      (spatial_index.h:99-122, with the
//                                      reference's initial values :531-532)
//   k_cell_ids                         linearised cell id of every row
//                                      (spatial_index.h:144-165)
//   k_gather_soa                       coordinates in sorted order, one
//                                      array per dimension
#include <cfloat>
#include <limits>

#include "common.hpp"
#include "grid.hpp"
#include "internal.hpp"

namespace hpd {

namespace {

constexpr int kMMThreads = 256;

// The reference starts max at numeric_limits<T>::min(): FLT_MIN / DBL_MIN for
// floating types (the smallest positive normal), the lowest value for
// integers.
template <typename T>
__device__ __forceinline__ T max_init() {
    return std::numeric_limits<T>::min();
}
template <typename T>
__device__ __forceinline__ T min_init() {
    return std::numeric_limits<T>::max();
}

// std::min(m, c) == (c < m) ? c : m ; std::max(m, c) == (m < c) ? c : m
template <typename T>
__device__ __forceinline__ T ref_min(T m, T c) {
    return (c < m) ? c : m;
}
template <typename T>
__device__ __forceinline__ T ref_max(T m, T c) {
    return (m < c) ? c : m;
}

template <typename T, int D>
__global__ __launch_bounds__(kMMThreads) void k_minmax_partial(const T* __restrict__ pts, int64_t n,
                                                               T* __restrict__ part,
                                                               int* __restrict__ nonfinite) {
    T mn[D], mx[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        mn[j] = min_init<T>();
        mx[j] = max_init<T>();
    }
    bool bad = false;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kMMThreads;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kMMThreads + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const T c = pts[i * D + j];
            mn[j] = ref_min(mn[j], c);
            mx[j] = ref_max(mx[j], c);
            if constexpr (std::is_floating_point<T>::value) bad |= !isfinite(c);
        }
    }
    // NaN never wins the comparisons above, so non-finite input is flagged
    // separately (the host rejects it)
    if (__any(bad) && __lane_id() == 0) atomicOr(nonfinite, 1);
    __shared__ T smn[D][kMMThreads];
    __shared__ T smx[D][kMMThreads];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        smn[j][threadIdx.x] = mn[j];
        smx[j][threadIdx.x] = mx[j];
    }
    __syncthreads();
    for (int o = kMMThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
#pragma unroll
            for (int j = 0; j < D; ++j) {
                smn[j][threadIdx.x] = ref_min(smn[j][threadIdx.x], smn[j][threadIdx.x + o]);
                smx[j][threadIdx.x] = ref_max(smx[j][threadIdx.x], smx[j][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            part[static_cast<int64_t>(blockIdx.x) * 2 * D + j] = smn[j][0];
            part[static_cast<int64_t>(blockIdx.x) * 2 * D + D + j] = smx[j][0];
        }
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kMMThreads) void k_minmax_final(const T* __restrict__ part, int nblocks,
                                                             T* __restrict__ out) {
    __shared__ T smn[D][kMMThreads];
    __shared__ T smx[D][kMMThreads];
    T mn[D], mx[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        mn[j] = min_init<T>();
        mx[j] = max_init<T>();
    }
    for (int b = threadIdx.x; b < nblocks; b += kMMThreads) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            mn[j] = ref_min(mn[j], part[static_cast<int64_t>(b) * 2 * D + j]);
            mx[j] = ref_max(mx[j], part[static_cast<int64_t>(b) * 2 * D + D + j]);
        }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
        smn[j][threadIdx.x] = mn[j];
        smx[j][threadIdx.x] = mx[j];
    }
    __syncthreads();
    for (int o = kMMThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
#pragma unroll
            for (int j = 0; j < D; ++j) {
                smn[j][threadIdx.x] = ref_min(smn[j][threadIdx.x], smn[j][threadIdx.x + o]);
                smx[j][threadIdx.x] = ref_max(smx[j][threadIdx.x], smx[j][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            out[j] = smn[j][0];
            out[D + j] = smx[j][0];
        }
    }
}

// floor((x - min) / eps) as the reference build computes it
// (spatial_index.h:156): (x - min) * rcp with rcp = 1/eps rounded once, in
// float for float and integer data, in double for double data.
template <typename T>
__device__ __forceinline__ uint64_t cell_index(T x, T mn, float rcp_f, double rcp_d);

template <>
__device__ __forceinline__ uint64_t cell_index<float>(float x, float mn, float rcp_f, double) {
    const float v = (x - mn) * rcp_f;
    return static_cast<uint64_t>(floorf(v));
}
template <>
__device__ __forceinline__ uint64_t cell_index<double>(double x, double mn, float, double rcp_d) {
    const double v = (x - mn) * rcp_d;
    return static_cast<uint64_t>(floor(v));
}

// (x - min) is evaluated in T's promoted type and converted to float
template <typename T>
__device__ __forceinline__ float int_diff_f32(T x, T mn) {
    using P = decltype(x - mn);
    if constexpr (sizeof(T) < sizeof(int)) {
        return static_cast<float>(static_cast<int>(x) - static_cast<int>(mn));
    } else {
        using U = std::make_unsigned_t<P>;
        return static_cast<float>(static_cast<P>(static_cast<U>(x) - static_cast<U>(mn)));
    }
}

#define HPD_INT_CELL_INDEX(T)                                                             \
    template <>                                                                           \
    __device__ __forceinline__ uint64_t cell_index<T>(T x, T mn, float rcp_f, double) {   \
        const float v = int_diff_f32<T>(x, mn) * rcp_f;                                   \
        return static_cast<uint64_t>(floorf(v));                                          \
    }
HPD_INT_CELL_INDEX(int8_t)
HPD_INT_CELL_INDEX(int16_t)
HPD_INT_CELL_INDEX(int32_t)
HPD_INT_CELL_INDEX(int64_t)
HPD_INT_CELL_INDEX(uint8_t)
HPD_INT_CELL_INDEX(uint16_t)
HPD_INT_CELL_INDEX(uint32_t)
HPD_INT_CELL_INDEX(uint64_t)
#undef HPD_INT_CELL_INDEX

template <typename T, int D, typename K>
__global__ __launch_bounds__(256) void k_cell_ids(const T* __restrict__ pts, int64_t n,
                                                  const T* __restrict__ minv, CellParams cp,
                                                  float rcp_f, double rcp_d, K* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    T row[D];
#pragma unroll
    for (int j = 0; j < D; ++j) row[j] = pts[i * D + j];
    uint64_t cell = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int j = cp.swapped[k];
        T xj = row[0];
        T mj = minv[0];
#pragma unroll
        for (int t = 1; t < D; ++t)
            if (t == j) {
                xj = row[t];
                mj = minv[t];
            }
        cell += cell_index<T>(xj, mj, rcp_f, rcp_d) * cp.stride[k];
    }
    keys[i] = static_cast<K>(cell);
    vals[i] = static_cast<uint32_t>(i);
}

template <typename T, int D>
__global__ __launch_bounds__(256) void k_gather_soa(const T* __restrict__ pts, int64_t n,
                                                    const uint32_t* __restrict__ order,
                                                    T* __restrict__ soa, int64_t ld) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t o = order[i];
#pragma unroll
    for (int j = 0; j < D; ++j) soa[static_cast<int64_t>(j) * ld + i] = pts[o * D + j];
}

inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

int minmax_partial_blocks(int64_t n) {
    const int64_t b = (n + kMMThreads * 8 - 1) / (kMMThreads * 8);
    return static_cast<int>(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

template <typename T, int D>
void launch_minmax(const T* pts, int64_t n, T* part, T* out, int* nonfinite, hipStream_t s) {
    const int nb = minmax_partial_blocks(n);
    hipLaunchKernelGGL((k_minmax_partial<T, D>), dim3(nb), dim3(kMMThreads), 0, s, pts, n, part,
                       nonfinite);
    hipLaunchKernelGGL((k_minmax_final<T, D>), dim3(1), dim3(kMMThreads), 0, s, part, nb, out);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D, typename K>
void launch_cell_ids(const T* pts, int64_t n, const T* minv, const CellParams& cp, float rcp_f,
                     double rcp_d, K* keys, uint32_t* vals, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_cell_ids<T, D, K>), dim3(blocks_for(n, 256)), dim3(256), 0, s, pts, n,
                       minv, cp, rcp_f, rcp_d, keys, vals);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_gather_soa(const T* pts, int64_t n, const uint32_t* order, T* soa, int64_t ld,
                       hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_gather_soa<T, D>), dim3(blocks_for(n, 256)), dim3(256), 0, s, pts, n,
                       order, soa, ld);
    HPD_HIP(hipGetLastError());
}

#define HPD_INST_GRID(T, D)                                                                     \
    template void launch_minmax<T, D>(const T*, int64_t, T*, T*, int*, hipStream_t);                         \
    template void launch_cell_ids<T, D, uint32_t>(const T*, int64_t, const T*, const CellParams&, \
                                                  float, double, uint32_t*, uint32_t*,          \
                                                  hipStream_t);                                 \
    template void launch_cell_ids<T, D, uint64_t>(const T*, int64_t, const T*, const CellParams&, \
                                                  float, double, uint64_t*, uint32_t*,          \
                                                  hipStream_t);                                 \
    template void launch_gather_soa<T, D>(const T*, int64_t, const uint32_t*, T*, int64_t,      \
                                          hipStream_t);

#define HPD_INST_GRID_ALL_D(T) \
    HPD_INST_GRID(T, 1)        \
    HPD_INST_GRID(T, 2)        \
    HPD_INST_GRID(T, 3)        \
    HPD_INST_GRID(T, 4)        \
    HPD_INST_GRID(T, 5)        \
    HPD_INST_GRID(T, 6)        \
    HPD_INST_GRID(T, 7)        \
    HPD_INST_GRID(T, 8)

HPD_FOR_EACH_DTYPE(HPD_INST_GRID_ALL_D)

}  // namespace hpd
