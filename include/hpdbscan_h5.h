/*
 * HDF5 file layer of the MI355X HPDBSCAN library (libhpdbscan_amd_h5.so,
 * linked against libhdf5 1.10 and libhpdbscan_amd.so).  Host code only: it
 * reads the rows into host memory and hands them to hpd_cluster, which does
 * all of the clustering on the GPU.
 *
 * Same file format as the reference:
 *   input   a 2-D dataset (rows x dims) of any of the ten element types
 *           of hpd_dtype, in any byte order (HDF5 converts on read)
 *   output  a 1-D dataset of n native longs (int64), written into the
 *           output file, which is created if it does not exist; an existing
 *           dataset of that name is written in place
 *
 * The reference interface each entry point replaces:
 *   hpd_h5_matrix_info    the checks of IO::read_hdf5 (include/io.h:35-95)
 *                         and the element-type dispatch of
 *                         HPDBSCAN::cluster(path, dataset, threads)
 *                         (include/hpdbscan.h:306-355)
 *   hpd_h5_read_matrix    IO::read_hdf5's read (include/io.h:96-110)
 *   hpd_cluster_h5        HPDBSCAN::cluster(path, dataset[, threads])
 *                         (include/hpdbscan.h:300-356)
 *   hpd_h5_write_clusters IO::write_hdf5 (include/io.h:113-194)
 *   hpd_h5_clusters_length,
 *   hpd_h5_read_clusters  no counterpart (read a labels dataset back)
 *   hpd_h5_write_matrix   no counterpart (writes an input dataset; the
 *                         reference's own scripts use h5py for that)
 *
 * Errors carry the reference's messages and exception kinds:
 *   HPD_EINVAL (std::invalid_argument): "Could not open <path>",
 *     "Could not open dataset <name>", "Data in dataset <name> should be a
 *     matrix", "Unsupported signed integer precision", "Unsupported unsigned
 *     integer precision", "Unsupported floating point precision",
 *     "Unsupported data set type"
 *   HPD_EIO (std::runtime_error): "Could not retrieve dataset type",
 *     "Could not retrieve dataspace extent, file seems to be corrupted",
 *     "Could not retrieve dataset extents, file seems to be corrupted",
 *     "Failed to read data", "Failed to write data, target data set is
 *     probably of wrong size or type"
 * hpd_h5_last_error() returns the message for the calling thread.  HDF5's
 * own error stack printing is off while these calls run, as in the
 * reference.
 */
#ifndef HPDBSCAN_H5_H
#define HPDBSCAN_H5_H

#include <stdint.h>

#include "hpdbscan_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed hpd_h5_* / hpd_cluster_h5 call on this thread. */
const char* hpd_h5_last_error(void);

/* Shape and element type (an hpd_dtype) of the 2-D dataset `dataset` of
 * `path`. */
int hpd_h5_matrix_info(const char* path, const char* dataset, int64_t* rows, int32_t* cols,
                       int32_t* dtype);

/* Reads the 2-D dataset into `out` (rows*cols values of `dtype`, row-major).
 * rows and cols must be the dataset's shape (HPD_EINVAL otherwise); dtype
 * may differ from the stored type, HDF5 then converts. */
int hpd_h5_read_matrix(const char* path, const char* dataset, int32_t dtype, void* out,
                       int64_t rows, int32_t cols);

/* Clusters the 2-D dataset of `path` in its stored element type:
 * labels_out gets `rows` int64 labels in file order, with n == rows of
 * hpd_h5_matrix_info (HPD_EINVAL otherwise).  Errors of hpd_cluster are
 * passed on with its status and message. */
int hpd_cluster_h5(hpd_context* ctx, const char* path, const char* dataset, float eps,
                   int64_t min_pts, int64_t* labels_out, int64_t n);

/* Writes n labels as the 1-D native-long dataset `dataset` of `path`. */
int hpd_h5_write_clusters(const char* path, const char* dataset, const int64_t* labels,
                          int64_t n);

/* Length of the 1-D dataset `dataset` of `path`. */
int hpd_h5_clusters_length(const char* path, const char* dataset, int64_t* n);

/* Reads a 1-D dataset of exactly n elements as int64. */
int hpd_h5_read_clusters(const char* path, const char* dataset, int64_t* labels, int64_t n);

/* Creates the 2-D dataset `dataset` (rows x cols of `dtype`, native byte
 * order) in `path` (created if missing); HPD_EINVAL if it exists. */
int hpd_h5_write_matrix(const char* path, const char* dataset, int32_t dtype, const void* data,
                        int64_t rows, int32_t cols);

#ifdef __cplusplus
}
#endif

#endif /* HPDBSCAN_H5_H */
