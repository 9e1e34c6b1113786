// Test infrastructure only: a C-ABI wrapper around the compiled reference.
//
// This file is compiled against the reference headers where they lie
// (include/hpdbscan.h, include/spatial_index.h of Markus-Goetz/hpdbscan) with
// the reference's own flags (CMakeLists.txt:12, `-Ofast`, plus OpenMP), and
// the result lands in oracle/_ref/.  Nothing in the product links it: only the
// tests, the golden-fixture script and the CPU-baseline leg of bench.py load
// it, as the checker / the timed CPU reference.
//
// Entry points (all return 0 on success, 1 on std::invalid_argument,
// 2 on any other exception; the message is copied into `err`):
//   hpd_ref_cluster_<t>   HPDBSCAN(eps, minPts).cluster<T>(data, n, d, threads)
//                          (hpdbscan.h:288, :476)
//   hpd_ref_spatial_<t>   SpatialIndex<T>(dataset, eps) public state: the
//                          sorted cell id per sorted position (cell_of,
//                          spatial_index.h:678) and the sorted position of
//                          every input row (recover_initial_order applied to
//                          the identity, spatial_index.h:770)
//   hpd_ref_candidates_<t> sum over points of |get_neighbors(cell_of(p))|
//                          (spatial_index.h:682), the reference's candidate
//                          count Σcand
#include <unordered_set>  // hpdbscan.h:215 uses it under WITH_OUTPUT only
#include <cstring>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>

#include "hpdbscan.h"

namespace {

void copy_err(char* err, int errlen, const char* msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg, static_cast<size_t>(errlen) - 1);
        err[errlen - 1] = '\0';
    }
}

template <typename T>
int run_cluster(const T* data, int n, int d, float eps, uint64_t min_pts, int threads,
                int64_t* out, char* err, int errlen) {
    try {
        HPDBSCAN h(eps, static_cast<size_t>(min_pts));
        Clusters c = h.cluster<T>(const_cast<T*>(data), n, d, threads);
        for (size_t i = 0; i < c.size(); ++i) out[i] = static_cast<int64_t>(c[i]);
        return 0;
    } catch (const std::invalid_argument& e) {
        copy_err(err, errlen, e.what());
        return 1;
    } catch (const std::exception& e) {
        copy_err(err, errlen, e.what());
        return 2;
    }
}

template <typename T>
int run_spatial(const T* data, int n, int d, float eps, uint64_t* cells_sorted,
                int64_t* sorted_pos_of_input) {
    hsize_t chunk[2] = {static_cast<hsize_t>(n), static_cast<hsize_t>(d)};
    Dataset ds(const_cast<T*>(data), chunk, HDF5_Types<T>::map());
    SpatialIndex<T> index(ds, eps);
    for (int i = 0; i < n; ++i) cells_sorted[i] = static_cast<uint64_t>(index.cell_of(i));
    Clusters c(static_cast<size_t>(n));
    std::iota(c.begin(), c.end(), 0);
    index.recover_initial_order(c);
    for (int i = 0; i < n; ++i) sorted_pos_of_input[i] = static_cast<int64_t>(c[i]);
    return 0;
}

template <typename T>
int run_candidates(const T* data, int n, int d, float eps, uint64_t* total) {
    hsize_t chunk[2] = {static_cast<hsize_t>(n), static_cast<hsize_t>(d)};
    Dataset ds(const_cast<T*>(data), chunk, HDF5_Types<T>::map());
    SpatialIndex<T> index(ds, eps);
    uint64_t sum = 0;
    Cell prev = NOT_VISITED;
    uint64_t cur = 0;
    for (int i = 0; i < n; ++i) {
        const Cell c = index.cell_of(i);
        if (c != prev) {
            cur = index.get_neighbors(c).size();
            prev = c;
        }
        sum += cur;
    }
    *total = sum;
    return 0;
}

}  // namespace

#define HPD_REF_ENTRY(suffix, T)                                                              \
    extern "C" int hpd_ref_cluster_##suffix(const T* data, int n, int d, float eps,           \
                                            uint64_t min_pts, int threads, int64_t* out,      \
                                            char* err, int errlen) {                          \
        return run_cluster<T>(data, n, d, eps, min_pts, threads, out, err, errlen);           \
    }                                                                                         \
    extern "C" int hpd_ref_spatial_##suffix(const T* data, int n, int d, float eps,           \
                                            uint64_t* cells_sorted, int64_t* pos) {           \
        return run_spatial<T>(data, n, d, eps, cells_sorted, pos);                            \
    }                                                                                         \
    extern "C" int hpd_ref_candidates_##suffix(const T* data, int n, int d, float eps,        \
                                               uint64_t* total) {                             \
        return run_candidates<T>(data, n, d, eps, total);                                     \
    }

HPD_REF_ENTRY(f32, float)
HPD_REF_ENTRY(f64, double)
HPD_REF_ENTRY(i8, int8_t)
HPD_REF_ENTRY(i16, int16_t)
HPD_REF_ENTRY(i32, int32_t)
HPD_REF_ENTRY(i64, int64_t)
HPD_REF_ENTRY(u8, uint8_t)
HPD_REF_ENTRY(u16, uint16_t)
HPD_REF_ENTRY(u32, uint32_t)
HPD_REF_ENTRY(u64, uint64_t)
