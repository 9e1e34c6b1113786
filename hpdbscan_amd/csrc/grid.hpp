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

// rows reduced by one block of the min/max kernel
int minmax_rows_per_block();

// The static pipeline's grid of every job, on the device: from the jobs'
// min / max (mm) it fills cp, minv and key_base of every job (host-written
// row ranges, rcp_f, rcp_d), cells[j] = job j's cells, *sort_bits, *total;
// flags[0] |= 1 when the grids cannot be keyed in 64 bits (the host path
// then takes over).  Same arithmetic as the host's (make_geometry).
template <typename T, int D>
void launch_geometry(JobParams<T>* jobs, int32_t njobs, const T* mm, int sub_bits, uint64_t* cells,
                     int* flags, int* sort_bits, unsigned long long* total, hipStream_t s);

// Per job j: out[j*2D .. j*2D+D) = min, out[j*2D+D .. (j+1)*2D) = max over
// the job's rows.  bfirst[j] = first partial block of job j (host-made,
// ceil(rows / minmax_rows_per_block()) blocks per job, bfirst[njobs] =
// nblocks); part holds 2*D*nblocks values; *nonfinite (zeroed by the caller)
// becomes 1 if a coordinate is NaN or infinite.
template <typename T, int D>
void launch_minmax(const T* pts, const JobParams<T>* jobs, int32_t njobs, const int32_t* bfirst,
                   int32_t nblocks, T* part, T* out, int* nonfinite, hipStream_t s);

// key = (job key_base + linearised cell id (spatial_index.h:144-165)) <<
// sub_bits | sub-box digit (sub_bits = 0: no digit; the cell's k^D split
// with k = sub_k), val = row
template <typename T, int D, typename K>
void launch_cell_ids(const T* pts, int64_t n, const JobParams<T>* jobs, int32_t njobs, int sub_k,
                     int sub_bits, K* keys, uint32_t* vals, hipStream_t s);

// soa[j * ld + i] = pts[order[i] * D + j]
template <typename T, int D>
void launch_gather_soa(const T* pts, int64_t n, const uint32_t* order, T* soa, int64_t ld,
                       hipStream_t s);

}  // namespace hpd
