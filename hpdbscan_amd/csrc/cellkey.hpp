// The cell key of a row: the reference's linearised cell id
// (spatial_index.h:144-165) offset by the job's key base, followed by the
// row's sub-box digit.  Shared by the kernel that keys the rows for the sort
// and the one that labels rows in input order, which must find the very key
// the sort used.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "internal.hpp"

namespace hpd {

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

// (x - min) * (1/eps) as the cell index computes it (float product for
// float data, double for double), returned exactly as a double
template <typename T>
__device__ __forceinline__ double scaled_value(T x, T mn, float rcp_f, double rcp_d);
template <>
__device__ __forceinline__ double scaled_value<float>(float x, float mn, float rcp_f, double) {
    return static_cast<double>((x - mn) * rcp_f);
}
template <>
__device__ __forceinline__ double scaled_value<double>(double x, double mn, float, double rcp_d) {
    return (x - mn) * rcp_d;
}
template <typename T>
__device__ __forceinline__ double scaled_value(T, T, float, double) {
    return 0.0;  // integer data has no sub-boxes
}

// key = (key_base + cell id) << sub_bits | sub-box digit, where the digit is
// the box of the cell's sub_k^D split along the same scaled coordinate
// v = (x - min) * (1/eps) that the cell index floors (v held exactly in
// double): sub_j = clamp(floor(v_j sub_k) - sub_k floor(v_j), 0, sub_k - 1)
template <typename T, int D>
__device__ __forceinline__ uint64_t row_key(const T* __restrict__ p, const JobParams<T>& jp,
                                            int sub_k, int sub_bits) {
    T row[D];
#pragma unroll
    for (int j = 0; j < D; ++j) row[j] = p[j];
    uint64_t cell = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int j = jp.cp.swapped[k];
        T xj = row[0];
        T mj = jp.minv[0];
#pragma unroll
        for (int t = 1; t < D; ++t)
            if (t == j) {
                xj = row[t];
                mj = jp.minv[t];
            }
        cell += cell_index<T>(xj, mj, jp.rcp_f, jp.rcp_d) * jp.cp.stride[k];
    }
    uint64_t sub = 0, mul = 1;
    if (sub_bits > 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const double v = scaled_value<T>(row[j], jp.minv[j], jp.rcp_f, jp.rcp_d);
            const double c = floor(v);
            long long sd = static_cast<long long>(floor(v * sub_k)) - static_cast<long long>(c) * sub_k;
            sd = sd < 0 ? 0 : (sd >= sub_k ? sub_k - 1 : sd);
            sub += static_cast<uint64_t>(sd) * mul;
            mul *= static_cast<uint64_t>(sub_k);
        }
    }
    return ((jp.key_base + cell) << sub_bits) | sub;
}

}  // namespace hpd
