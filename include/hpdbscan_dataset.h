/*
 * The reference's in-memory Dataset (include/dataset.h:26-62 of
 * Markus-Goetz/hpdbscan) and HPDBSCAN::cluster<T>(Dataset&, int threads)
 * (include/hpdbscan.h:358-359, instantiated for ten element types in
 * src/hpdbscan.cpp:19-30), for C++ callers that build a Dataset themselves.
 * Needs hdf5.h (hsize_t, hid_t, H5Tcopy) like the reference's header; link
 * with -lhdf5 besides -lhpdbscan_amd.
 *
 * Same members with the same meaning, since caller code reads and writes
 * them: m_shape (rows, dims), m_chunk (the rows of this process: equal to
 * m_shape on one node), m_offset, m_type (an HDF5 type id, copied) and m_p
 * (the rows, allocated with room for m_shape[0] * m_shape[1] elements of
 * m_type's precision).  The constructor from a pointer copies the rows, as
 * the reference's does.  Unlike the reference, a Dataset cannot be copied
 * (the reference's implicit copy would free m_p twice).
 *
 * cluster<T>(dataset, threads) clusters the m_chunk[0] x m_chunk[1] rows of
 * m_p read as T on the GPU (threads is accepted and unused) and leaves the
 * dataset as it was: the reference sorts m_p in place and restores the
 * input order before returning (spatial_index.h:184-216, 852-872); this
 * build never writes it.
 */
#ifndef HPDBSCAN_AMD_DATASET_H
#define HPDBSCAN_AMD_DATASET_H

#include <hdf5.h>

#include <climits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>

#include "hpdbscan.h"

struct Dataset {
    hsize_t m_shape[2];
    hsize_t m_chunk[2];
    hsize_t m_offset[2] = {0, 0};
    hid_t m_type;
    void* m_p;

    Dataset(const hsize_t shape[2], hid_t type) {
        m_shape[0] = m_chunk[0] = shape[0];
        m_shape[1] = m_chunk[1] = shape[1];
        m_type = H5Tcopy(type);
        const size_t bytes = static_cast<size_t>(shape[0] * shape[1] * H5Tget_precision(type) / 8);
        m_p = std::malloc(bytes > 0 ? bytes : 1);
        if (!m_p) {
            H5Tclose(m_type);
            throw std::bad_alloc();
        }
    }

    template <typename T>
    Dataset(T* data, const hsize_t shape[2], hid_t type) : Dataset(shape, type) {
        std::memcpy(m_p, data, static_cast<size_t>(shape[0] * shape[1]) * sizeof(T));
    }

    Dataset(const Dataset&) = delete;
    Dataset& operator=(const Dataset&) = delete;

    ~Dataset() {
        H5Tclose(m_type);
        std::free(m_p);
    }
};

template <typename T>
Clusters HPDBSCAN::cluster(Dataset& dataset, int threads) {
    if (dataset.m_chunk[0] > static_cast<hsize_t>(INT_MAX) ||
        dataset.m_chunk[1] > static_cast<hsize_t>(INT_MAX))
        throw std::invalid_argument("dataset shape exceeds the int range of cluster(T*, int, int)");
    return cluster<T>(static_cast<T*>(dataset.m_p), static_cast<int>(dataset.m_chunk[0]),
                      static_cast<int>(dataset.m_chunk[1]), threads);
}

#endif  // HPDBSCAN_AMD_DATASET_H
