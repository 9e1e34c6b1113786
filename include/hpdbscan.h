/*
 * Drop-in C++ class for the reference's HPDBSCAN (include/hpdbscan.h of
 * Markus-Goetz/hpdbscan), running on the MI355X through the C-ABI of
 * hpdbscan_capi.h.  Link with -lhpdbscan_amd.
 *
 *   HPDBSCAN(float epsilon, size_t min_points)        hpdbscan.h:288-298
 *   Clusters cluster(T* data, int dim0, int dim1[, int threads])
 *                                                      hpdbscan.h:475-486
 *   Clusters cluster(path, dataset[, threads])         hpdbscan.h:300-304
 *   Clusters cluster<T>(Dataset&[, int threads])       hpdbscan.h:358-359
 *                                     (Dataset: include hpdbscan_dataset.h)
 *
 * Same types (Cluster = ptrdiff_t, Clusters = std::vector<Cluster>,
 * constants.h:24-26), same label convention (0 noise, -k core, +k border of
 * cluster k, README.md:54-56), same exceptions (std::invalid_argument
 * "epsilon needs to be positive" for epsilon <= 0).  `threads` is accepted
 * for source compatibility; the GPU path does not use host threads.
 * The input is read, never modified, as in the reference (dataset.h:40-42).
 *
 * The HDF5 overloads read the 2-D dataset in its stored element type
 * (hpdbscan.h:306-355) through libhpdbscan_amd_h5.so (hpdbscan_h5.h): link
 * with -lhpdbscan_amd_h5 when calling them.  IO::write_hdf5 (io.h:113-194)
 * is hpdbscan_amd::write_hdf5 below.
 *
 * Differences: a NaN epsilon or non-finite coordinates throw
 * std::invalid_argument instead of producing undefined grid sizes.
 * One HPDBSCAN object owns one device context; copies share it and must not
 * be used from two threads at once.
 */
#ifndef HPDBSCAN_AMD_HPDBSCAN_H
#define HPDBSCAN_AMD_HPDBSCAN_H

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hpdbscan_capi.h"
#include "hpdbscan_h5.h"

using Cluster = ptrdiff_t;
using Clusters = std::vector<Cluster>;

// The reference's in-memory dataset (dataset.h:26-62); defined, with the
// member template HPDBSCAN::cluster<T>(Dataset&, int), by hpdbscan_dataset.h,
// which needs hdf5.h as the reference's does.
struct Dataset;

namespace hpdbscan_amd {

template <typename T>
struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = HPD_DTYPE_F32; };
template <> struct dtype_of<double> { static constexpr int value = HPD_DTYPE_F64; };
template <> struct dtype_of<int8_t> { static constexpr int value = HPD_DTYPE_I8; };
template <> struct dtype_of<int16_t> { static constexpr int value = HPD_DTYPE_I16; };
template <> struct dtype_of<int32_t> { static constexpr int value = HPD_DTYPE_I32; };
template <> struct dtype_of<int64_t> { static constexpr int value = HPD_DTYPE_I64; };
template <> struct dtype_of<uint8_t> { static constexpr int value = HPD_DTYPE_U8; };
template <> struct dtype_of<uint16_t> { static constexpr int value = HPD_DTYPE_U16; };
template <> struct dtype_of<uint32_t> { static constexpr int value = HPD_DTYPE_U32; };
template <> struct dtype_of<uint64_t> { static constexpr int value = HPD_DTYPE_U64; };

inline void throw_status(int status) {
    if (status == HPD_OK) return;
    const std::string msg = hpd_last_error();
    if (status == HPD_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Status of an hpd_h5_* / hpd_cluster_h5 call as the reference's exceptions.
inline void throw_h5_status(int status) {
    if (status == HPD_OK) return;
    const std::string msg = hpd_h5_last_error();
    if (status == HPD_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// IO::write_hdf5 (io.h:113-194): the labels as a 1-D native-long dataset of
// `path`, created if missing; an existing dataset of that name is written in
// place.
inline void write_hdf5(const std::string& path, const std::string& dataset,
                       const Clusters& clusters) {
    static_assert(sizeof(Cluster) == sizeof(int64_t), "ptrdiff_t must be 64-bit");
    throw_h5_status(hpd_h5_write_clusters(path.c_str(), dataset.c_str(),
                                          reinterpret_cast<const int64_t*>(clusters.data()),
                                          static_cast<int64_t>(clusters.size())));
}

}  // namespace hpdbscan_amd

class HPDBSCAN {
    float m_epsilon;
    size_t m_min_points;
    std::shared_ptr<hpd_context> m_ctx;

    hpd_context* context() {
        if (!m_ctx) {
            hpd_context* c = nullptr;
            hpdbscan_amd::throw_status(hpd_context_create(0, &c));
            m_ctx.reset(c, [](hpd_context* p) { hpd_context_destroy(p); });
        }
        return m_ctx.get();
    }

public:
    HPDBSCAN(float epsilon, size_t min_points) : m_epsilon(epsilon), m_min_points(min_points) {
        if (epsilon <= 0.0) {
            throw std::invalid_argument("epsilon needs to be positive");
        }
    }

    Clusters cluster(const std::string& path, const std::string& dataset) {
        return cluster(path, dataset, 0);
    }

    Clusters cluster(const std::string& path, const std::string& dataset, int threads) {
        (void)threads;
        int64_t rows = 0;
        int32_t cols = 0, dtype = 0;
        hpdbscan_amd::throw_h5_status(
            hpd_h5_matrix_info(path.c_str(), dataset.c_str(), &rows, &cols, &dtype));
        Clusters out(static_cast<size_t>(rows));
        static_assert(sizeof(Cluster) == sizeof(int64_t), "ptrdiff_t must be 64-bit");
        hpdbscan_amd::throw_h5_status(hpd_cluster_h5(context(), path.c_str(), dataset.c_str(),
                                                     m_epsilon, static_cast<int64_t>(m_min_points),
                                                     reinterpret_cast<int64_t*>(out.data()), rows));
        return out;
    }

    template <typename T>
    Clusters cluster(T* data, int dim0, int dim1, int threads) {
        (void)threads;
        if (dim0 < 0 || dim1 < 0) throw std::invalid_argument("negative data shape");
        Clusters out(static_cast<size_t>(dim0));
        static_assert(sizeof(Cluster) == sizeof(int64_t), "ptrdiff_t must be 64-bit");
        hpdbscan_amd::throw_status(hpd_cluster(context(), hpdbscan_amd::dtype_of<T>::value, data,
                                               dim0, dim1, m_epsilon, m_min_points,
                                               reinterpret_cast<int64_t*>(out.data())));
        return out;
    }

    template <typename T>
    Clusters cluster(T* data, int dim0, int dim1) {
        return cluster(data, dim0, dim1, 0);
    }

    // cluster<T>(Dataset&, int threads) (hpdbscan.h:358-359): the rows of
    // dataset.m_p read as T; defined in hpdbscan_dataset.h
    template <typename T>
    Clusters cluster(Dataset& dataset, int threads);

    template <typename T>
    Clusters cluster(Dataset& dataset) {
        return cluster<T>(dataset, 0);
    }
};

#endif  // HPDBSCAN_AMD_HPDBSCAN_H
