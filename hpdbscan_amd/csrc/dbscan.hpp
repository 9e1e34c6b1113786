// Launchers of the clustering kernels (dbscan.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace hpd {

template <typename T>
struct DbscanArgs {
    const T* soa = nullptr;          // sorted coordinates, [D][ld]
    int64_t n = 0;
    int64_t ld = 0;                  // row stride of soa: n rounded up to 16
    const int2* items = nullptr;     // (cell, first row) per 64-row chunk
    int32_t nitems = 0;
    const int32_t* ustart = nullptr; // first row of each occupied cell, [m+1]
    const uint32_t* roff = nullptr;  // CSR offsets into ranges, [m+1]
    const int2* ranges = nullptr;    // [first, end) rows of neighbour rows
    T eps2{};                        // EPS2 converted to T
    uint32_t min_pts = 0;
};

template <typename T, int D>
void launch_core(const DbscanArgs<T>& a, uint8_t* core, unsigned long long* ncore, hipStream_t s);

template <typename T, int D>
void launch_union(const DbscanArgs<T>& a, const uint8_t* core, int32_t* parent, hipStream_t s);

// roots of the cores, then border / noise labels; labels are written in
// input order through `order` (sorted position -> input row).
// stats[0] += noise rows, stats[1] += ambiguous border rows
template <typename T, int D>
void launch_labels(const DbscanArgs<T>& a, const uint8_t* core, int32_t* parent,
                   const uint32_t* order, int64_t* labels, unsigned long long* stats,
                   hipStream_t s);

}  // namespace hpd
