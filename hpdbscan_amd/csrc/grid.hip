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
#include <type_traits>

#include "cellkey.hpp"
#include "common.hpp"
#include "grid.hpp"
#include "internal.hpp"

namespace hpd {

namespace {

constexpr int kMMThreads = 256;
constexpr int kMMRows = 8 * kMMThreads;  // rows per partial block

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

// first job whose partial-block range contains block b (bfirst ascending)
__device__ __forceinline__ int32_t job_of_block(const int32_t* __restrict__ bfirst, int32_t njobs,
                                                int32_t b) {
    int32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (bfirst[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One block reduces up to kMMRows rows of one job; bfirst[j] is the first
// block of job j.
template <typename T, int D>
__global__ __launch_bounds__(kMMThreads) void k_minmax_partial(const T* __restrict__ pts,
                                                               const JobParams<T>* __restrict__ jobs,
                                                               int32_t njobs,
                                                               const int32_t* __restrict__ bfirst,
                                                               T* __restrict__ part,
                                                               int* __restrict__ nonfinite) {
    const int32_t b = static_cast<int32_t>(blockIdx.x);
    const int32_t j = job_of_block(bfirst, njobs, b);
    const int64_t r0 = jobs[j].row_begin + static_cast<int64_t>(b - bfirst[j]) * kMMRows;
    const int64_t r1 = min(r0 + kMMRows, jobs[j].row_end);
    T mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        mn[k] = min_init<T>();
        mx[k] = max_init<T>();
    }
    bool bad = false;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kMMThreads) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T c = pts[i * D + k];
            mn[k] = ref_min(mn[k], c);
            mx[k] = ref_max(mx[k], c);
            if constexpr (std::is_floating_point<T>::value) bad |= !isfinite(c);
        }
    }
    // NaN never wins the comparisons above, so non-finite input is flagged
    // separately (the host rejects it)
    if (__any(bad) && __lane_id() == 0) atomicOr(nonfinite, 1);
    __shared__ T smn[D][kMMThreads];
    __shared__ T smx[D][kMMThreads];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        smn[k][threadIdx.x] = mn[k];
        smx[k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = kMMThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                smn[k][threadIdx.x] = ref_min(smn[k][threadIdx.x], smn[k][threadIdx.x + o]);
                smx[k][threadIdx.x] = ref_max(smx[k][threadIdx.x], smx[k][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            part[static_cast<int64_t>(b) * 2 * D + k] = smn[k][0];
            part[static_cast<int64_t>(b) * 2 * D + D + k] = smx[k][0];
        }
    }
}

// One block per job: min/max over the job's partial blocks.
template <typename T, int D>
__global__ __launch_bounds__(kMMThreads) void k_minmax_final(const T* __restrict__ part,
                                                             const int32_t* __restrict__ bfirst,
                                                             T* __restrict__ out) {
    const int32_t j = static_cast<int32_t>(blockIdx.x);
    __shared__ T smn[D][kMMThreads];
    __shared__ T smx[D][kMMThreads];
    T mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        mn[k] = min_init<T>();
        mx[k] = max_init<T>();
    }
    for (int32_t b = bfirst[j] + threadIdx.x; b < bfirst[j + 1]; b += kMMThreads) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            mn[k] = ref_min(mn[k], part[static_cast<int64_t>(b) * 2 * D + k]);
            mx[k] = ref_max(mx[k], part[static_cast<int64_t>(b) * 2 * D + D + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        smn[k][threadIdx.x] = mn[k];
        smx[k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = kMMThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                smn[k][threadIdx.x] = ref_min(smn[k][threadIdx.x], smn[k][threadIdx.x + o]);
                smx[k][threadIdx.x] = ref_max(smx[k][threadIdx.x], smx[k][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            out[static_cast<int64_t>(j) * 2 * D + k] = smn[k][0];
            out[static_cast<int64_t>(j) * 2 * D + D + k] = smx[k][0];
        }
    }
}

template <typename T, int D, typename K>
__global__ __launch_bounds__(256) void k_cell_ids(const T* __restrict__ pts, int64_t n,
                                                  const JobParams<T>* __restrict__ jobs,
                                                  int32_t njobs, int sub_k, int sub_bits,
                                                  K* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const JobParams<T>& jp = jobs[job_of_row_wave(jobs, njobs, i)];
    keys[i] = static_cast<K>(row_key<T, D>(pts + i * D, jp, sub_k, sub_bits));
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

// float / double -> integer as x86-64 cvttss2si / cvttsd2si produce it
// (out of range or NaN: the "integer indefinite" value)
__device__ __forceinline__ int64_t d2i64_x86(double v) {
    if (!(v >= -9223372036854775808.0 && v < 9223372036854775808.0)) return INT64_MIN;
    return static_cast<int64_t>(v);
}
__device__ __forceinline__ uint64_t d2u64_x86(double v) {
    if (v < 9223372036854775808.0) return static_cast<uint64_t>(d2i64_x86(v));
    return static_cast<uint64_t>(d2i64_x86(v - 9223372036854775808.0)) ^ 0x8000000000000000ull;
}

// cells of one dimension, ceil((max - min) / eps) + 1 (spatial_index.h:124-131)
// with the reciprocal form and the flushing of the reference's build (device
// code flushes denormal inputs and results itself)
template <typename T>
__device__ __forceinline__ uint64_t cells_of_dev(T mx, T mn, float rcp_f, double rcp_d) {
    if constexpr (std::is_same<T, float>::value) {
        const float v = (mx - mn) * rcp_f;
        return d2u64_x86(static_cast<double>(ceilf(v))) + 1;
    } else if constexpr (std::is_same<T, double>::value) {
        const double v = (mx - mn) * rcp_d;
        return d2u64_x86(ceil(v)) + 1;
    } else {
        const float v = int_diff_f32<T>(mx, mn) * rcp_f;
        return d2u64_x86(static_cast<double>(ceilf(v))) + 1;
    }
}

// The grid of every job from its min / max (mm: 2D values per job), as the
// host computes it after reading them back: cells per dimension, dimension
// order (an insertion sort ascending by cells, which is what std::sort does
// for so few keys: ties keep their order), strides; then key bases in job
// order.  flags[0] |= 1: a grid of more than 2^64 cells; *sort_bits = bits
// of the largest key + sub_bits; *total = all jobs' cells.  One block.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_geometry(JobParams<T>* __restrict__ jobs, int32_t njobs,
                                                  const T* __restrict__ mm, int sub_bits,
                                                  uint64_t* __restrict__ cells,
                                                  int* __restrict__ flags,
                                                  int* __restrict__ sort_bits,
                                                  unsigned long long* __restrict__ total) {
    __shared__ unsigned long long part[256];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    // per job: geometry and number of cells (0 for a job without rows)
    for (int32_t j = threadIdx.x; j < njobs; j += 256) {
        JobParams<T>& jp = jobs[j];
        uint64_t tot = 0;
        if (jp.row_end > jp.row_begin) {
            const T* mn = &mm[static_cast<int64_t>(j) * 2 * D];
            uint64_t c[D];
            tot = 1;
            bool of = false;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                c[k] = cells_of_dev<T>(mn[D + k], mn[k], jp.rcp_f, jp.rcp_d);
                uint64_t t;
                if (__builtin_mul_overflow(tot, c[k], &t)) of = true;
                tot = t;
            }
            if (of) atomicOr(&bad, 1);
            int sw[D];
#pragma unroll
            for (int k = 0; k < D; ++k) sw[k] = k;
            for (int k = 1; k < D; ++k) {
                const int v = sw[k];
                int i = k - 1;
                while (i >= 0 && c[v] < c[sw[i]]) {
                    sw[i + 1] = sw[i];
                    --i;
                }
                sw[i + 1] = v;
            }
            uint64_t acc = 1;
            jp.cp.d = D;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                jp.minv[k] = mn[k];
                jp.cp.swapped[k] = sw[k];
                jp.cp.stride[k] = acc;
                jp.cp.ncell[k] = c[sw[k]];
                acc *= c[sw[k]];
            }
        }
        cells[j] = tot;
    }
    __syncthreads();
    // key bases: each thread sums a contiguous run of jobs, the runs are
    // scanned, then every thread assigns its run's bases
    const int32_t per = (njobs + 255) / 256;
    const int32_t j0 = static_cast<int32_t>(threadIdx.x) * per;
    const int32_t j1 = min(njobs, j0 + per);
    unsigned long long sum = 0;
    bool of = false;
    for (int32_t j = j0; j < j1; ++j) of |= __builtin_add_overflow(sum, cells[j], &sum);
    part[threadIdx.x] = sum;
    if (of) atomicOr(&bad, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long acc = 0;
        bool o2 = false;
        for (int t = 0; t < 256; ++t) {
            const unsigned long long v = part[t];
            part[t] = acc;
            o2 |= __builtin_add_overflow(acc, v, &acc);
        }
        if (o2) bad = 1;
        *total = acc;
        const uint64_t maxkey = acc > 0 ? acc - 1 : 0;
        const int kb = maxkey == 0 ? 0 : 64 - __clzll(static_cast<long long>(maxkey));
        *sort_bits = kb + sub_bits;
        if (kb + sub_bits > 64) bad = 1;
    }
    __syncthreads();
    unsigned long long base = part[threadIdx.x];
    for (int32_t j = j0; j < j1; ++j) {
        jobs[j].key_base = base;
        base += cells[j];
    }
    if (threadIdx.x == 0 && bad) flags[0] |= 1;
}

inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

template <typename T, int D>
void launch_geometry(JobParams<T>* jobs, int32_t njobs, const T* mm, int sub_bits, uint64_t* cells,
                     int* flags, int* sort_bits, unsigned long long* total, hipStream_t s) {
    hipLaunchKernelGGL((k_geometry<T, D>), dim3(1), dim3(256), 0, s, jobs, njobs, mm, sub_bits, cells,
                       flags, sort_bits, total);
    HPD_HIP(hipGetLastError());
}

int minmax_rows_per_block() { return kMMRows; }

template <typename T, int D>
void launch_minmax(const T* pts, const JobParams<T>* jobs, int32_t njobs, const int32_t* bfirst,
                   int32_t nblocks, T* part, T* out, int* nonfinite, hipStream_t s) {
    if (nblocks > 0)
        hipLaunchKernelGGL((k_minmax_partial<T, D>), dim3(nblocks), dim3(kMMThreads), 0, s, pts, jobs,
                           njobs, bfirst, part, nonfinite);
    hipLaunchKernelGGL((k_minmax_final<T, D>), dim3(njobs), dim3(kMMThreads), 0, s, part, bfirst, out);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D, typename K>
void launch_cell_ids(const T* pts, int64_t n, const JobParams<T>* jobs, int32_t njobs, int sub_k,
                     int sub_bits, K* keys, uint32_t* vals, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_cell_ids<T, D, K>), dim3(blocks_for(n, 256)), dim3(256), 0, s, pts, n,
                       jobs, njobs, sub_k, sub_bits, keys, vals);
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
    template void launch_geometry<T, D>(JobParams<T>*, int32_t, const T*, int, uint64_t*, int*,  \
                                        int*, unsigned long long*, hipStream_t);                  \
    template void launch_minmax<T, D>(const T*, const JobParams<T>*, int32_t, const int32_t*,    \
                                      int32_t, T*, T*, int*, hipStream_t);                       \
    template void launch_cell_ids<T, D, uint32_t>(const T*, int64_t, const JobParams<T>*,        \
                                                  int32_t, int, int, uint32_t*, uint32_t*,       \
                                                  hipStream_t);                                  \
    template void launch_cell_ids<T, D, uint64_t>(const T*, int64_t, const JobParams<T>*,        \
                                                  int32_t, int, int, uint64_t*, uint32_t*,       \
                                                  hipStream_t);                                  \
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
