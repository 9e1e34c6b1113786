// Launchers of the eps-grid kernels (grid.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

// X-macro over the element types the library instantiates (the same ten as
// the reference's src/hpdbscan.cpp:19-30).
#define HPD_FOR_EACH_DTYPE(X) \
    X(float)                  \
    X(double)                 \
    X(int8_t)                 \
    X(int16_t)                \
    X(int32_t)                \
    X(int64_t)                \
    X(uint8_t)                \
    X(uint16_t)               \
    X(uint32_t)               \
    X(uint64_t)

namespace hpd {

int minmax_partial_blocks(int64_t n);

// minmax: part holds 2*D*minmax_partial_blocks(n) values; out receives
// min[0..D) then max[0..D); *nonfinite (zeroed by the caller) becomes 1 if a
// coordinate is NaN or infinite.
template <typename T, int D>
void launch_minmax(const T* pts, int64_t n, T* part, T* out, int* nonfinite, hipStream_t s);

template <typename T, int D, typename K>
void launch_cell_ids(const T* pts, int64_t n, const T* minv, const CellParams& cp, float rcp_f,
                     double rcp_d, K* keys, uint32_t* vals, hipStream_t s);

// soa[j * ld + i] = pts[order[i] * D + j]
template <typename T, int D>
void launch_gather_soa(const T* pts, int64_t n, const uint32_t* order, T* soa, int64_t ld,
                       hipStream_t s);

}  // namespace hpd
