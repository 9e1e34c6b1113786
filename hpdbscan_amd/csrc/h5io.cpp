// HDF5 file layer (include/hpdbscan_h5.h): reads a 2-D dataset of any of
// the ten element types, clusters it with hpd_cluster, writes the labels as
// a 1-D native-long dataset.  The file format, the order of the HDF5 calls
// that decide what happens to existing files and datasets, and the error
// messages follow include/io.h:35-194 and include/hpdbscan.h:300-356.
#include <hdf5.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hpdbscan_capi.h"
#include "../../include/hpdbscan_h5.h"

namespace {

thread_local std::string g_error;

struct Failure {
    int status;
    std::string message;
};

[[noreturn]] void fail(int status, std::string message) { throw Failure{status, std::move(message)}; }

// Owns one HDF5 identifier and releases it with the matching close call.
class Handle {
public:
    using Closer = herr_t (*)(hid_t);
    Handle(hid_t id, Closer close) : id_(id), close_(close) {}
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    ~Handle() {
        if (id_ >= 0) close_(id_);
    }
    hid_t get() const { return id_; }
    bool ok() const { return id_ >= 0; }

private:
    hid_t id_;
    Closer close_;
};

// HDF5 prints its error stack to stderr by default; the reference switches
// that off around its I/O (io.h:43-47, :137-141).
class QuietErrors {
public:
    QuietErrors() {
        H5Eget_auto2(H5E_DEFAULT, &func_, &data_);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietErrors() { H5Eset_auto2(H5E_DEFAULT, func_, data_); }

private:
    H5E_auto2_t func_ = nullptr;
    void* data_ = nullptr;
};

hid_t native_type(int32_t dtype) {
    switch (dtype) {
        case HPD_DTYPE_F32: return H5T_NATIVE_FLOAT;
        case HPD_DTYPE_F64: return H5T_NATIVE_DOUBLE;
        case HPD_DTYPE_I8: return H5T_NATIVE_SCHAR;
        case HPD_DTYPE_I16: return H5T_NATIVE_SHORT;
        case HPD_DTYPE_I32: return H5T_NATIVE_INT;
        case HPD_DTYPE_I64: return H5T_NATIVE_LONG;
        case HPD_DTYPE_U8: return H5T_NATIVE_UCHAR;
        case HPD_DTYPE_U16: return H5T_NATIVE_USHORT;
        case HPD_DTYPE_U32: return H5T_NATIVE_UINT;
        case HPD_DTYPE_U64: return H5T_NATIVE_ULONG;
        default: fail(HPD_EINVAL, "unknown dtype code " + std::to_string(dtype));
    }
}

size_t dtype_size(int32_t dtype) {
    static const size_t sizes[] = {4, 8, 1, 2, 4, 8, 1, 2, 4, 8};
    return sizes[dtype];
}

// The element-type dispatch of HPDBSCAN::cluster(path, ...)
// (hpdbscan.h:306-355): class, sign and precision of the stored type.
int32_t dtype_of(hid_t type) {
    const H5T_class_t cls = H5Tget_class(type);
    const size_t precision = H5Tget_precision(type);
    if (cls == H5T_INTEGER) {
        const bool is_signed = H5Tget_sign(type) == H5T_SGN_2;
        switch (precision) {
            case 8: return is_signed ? HPD_DTYPE_I8 : HPD_DTYPE_U8;
            case 16: return is_signed ? HPD_DTYPE_I16 : HPD_DTYPE_U16;
            case 32: return is_signed ? HPD_DTYPE_I32 : HPD_DTYPE_U32;
            case 64: return is_signed ? HPD_DTYPE_I64 : HPD_DTYPE_U64;
            default:
                fail(HPD_EINVAL, is_signed ? "Unsupported signed integer precision"
                                           : "Unsupported unsigned integer precision");
        }
    }
    if (cls == H5T_FLOAT) {
        if (precision == 32) return HPD_DTYPE_F32;
        if (precision == 64) return HPD_DTYPE_F64;
        fail(HPD_EINVAL, "Unsupported floating point precision");
    }
    fail(HPD_EINVAL, "Unsupported data set type");
}

// An open 2-D input dataset with its shape and element type: the checks of
// IO::read_hdf5 (io.h:49-85) in the same order.
struct Matrix {
    Handle file;
    Handle dset;
    Handle type;
    Handle space;
    hsize_t extents[2] = {0, 0};

    Matrix(const char* path, const char* name)
        : file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose),
          dset(file.ok() ? H5Dopen2(file.get(), name, H5P_DEFAULT) : -1, H5Dclose),
          type(dset.ok() ? H5Dget_type(dset.get()) : -1, H5Tclose),
          space(type.ok() ? H5Dget_space(dset.get()) : -1, H5Sclose) {
        if (!file.ok()) fail(HPD_EINVAL, std::string("Could not open ") + path);
        if (!dset.ok()) fail(HPD_EINVAL, std::string("Could not open dataset ") + name);
        if (!type.ok()) fail(HPD_EIO, "Could not retrieve dataset type");
        if (!space.ok())
            fail(HPD_EIO, "Could not retrieve dataspace extent, file seems to be corrupted");
        if (H5Sget_simple_extent_ndims(space.get()) != 2)
            fail(HPD_EINVAL, std::string("Data in dataset ") + name + " should be a matrix");
        if (H5Sget_simple_extent_dims(space.get(), extents, nullptr) < 0)
            fail(HPD_EIO, "Could not retrieve dataset extents, file seems to be corrupted");
    }
};

// An open 1-D dataset (a labels vector) and its length.
struct Vector {
    Handle file;
    Handle dset;
    Handle space;
    hsize_t count = 0;

    Vector(const char* path, const char* name)
        : file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose),
          dset(file.ok() ? H5Dopen2(file.get(), name, H5P_DEFAULT) : -1, H5Dclose),
          space(dset.ok() ? H5Dget_space(dset.get()) : -1, H5Sclose) {
        if (!file.ok()) fail(HPD_EINVAL, std::string("Could not open ") + path);
        if (!dset.ok()) fail(HPD_EINVAL, std::string("Could not open dataset ") + name);
        if (!space.ok())
            fail(HPD_EIO, "Could not retrieve dataspace extent, file seems to be corrupted");
        if (H5Sget_simple_extent_ndims(space.get()) != 1 ||
            H5Sget_simple_extent_dims(space.get(), &count, nullptr) < 0)
            fail(HPD_EINVAL, std::string("Data in dataset ") + name + " should be a vector");
    }
};

void check_shape(const Matrix& m, int64_t rows, int32_t cols) {
    if (rows < 0 || cols < 0 || static_cast<hsize_t>(rows) != m.extents[0] ||
        static_cast<hsize_t>(cols) != m.extents[1])
        fail(HPD_EINVAL, "buffer shape (" + std::to_string(rows) + ", " + std::to_string(cols) +
                             ") does not match the dataset's (" + std::to_string(m.extents[0]) +
                             ", " + std::to_string(m.extents[1]) + ")");
}

void read_into(const Matrix& m, hid_t mem_type, void* out) {
    if (m.extents[0] == 0 || m.extents[1] == 0) return;
    if (H5Dread(m.dset.get(), mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        fail(HPD_EIO, "Failed to read data");
}

// Opens `path` for writing, creating it first if it does not exist
// (io.h:158-163: an exclusive create that may fail, then an open).
hid_t open_for_write(const char* path) {
    {
        Handle created(H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    }
    const hid_t file = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (file < 0) fail(HPD_EINVAL, std::string("Could not open ") + path);
    return file;
}

template <typename F>
int guarded(F&& f) {
    try {
        QuietErrors quiet;
        f();
        return HPD_OK;
    } catch (const Failure& e) {
        g_error = e.message;
        return e.status;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return HPD_ENOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return HPD_EIO;
    }
}

void need(const void* p, const char* what) {
    if (!p) fail(HPD_EINVAL, std::string(what) + " is null");
}

}  // namespace

extern "C" {

const char* hpd_h5_last_error(void) { return g_error.c_str(); }

int hpd_h5_matrix_info(const char* path, const char* dataset, int64_t* rows, int32_t* cols,
                       int32_t* dtype) {
    return guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        Matrix m(path, dataset);
        const int32_t t = dtype_of(m.type.get());
        if (rows) *rows = static_cast<int64_t>(m.extents[0]);
        if (cols) *cols = static_cast<int32_t>(m.extents[1]);
        if (dtype) *dtype = t;
    });
}

int hpd_h5_read_matrix(const char* path, const char* dataset, int32_t dtype, void* out,
                       int64_t rows, int32_t cols) {
    return guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        const hid_t mem = native_type(dtype);
        Matrix m(path, dataset);
        check_shape(m, rows, cols);
        if (rows > 0 && cols > 0) need(out, "out");
        read_into(m, mem, out);
    });
}

int hpd_cluster_h5(hpd_context* ctx, const char* path, const char* dataset, float eps,
                   int64_t min_pts, int64_t* labels_out, int64_t n) {
    int status = HPD_OK;
    std::vector<unsigned char> rows_buf;
    int64_t rows = 0;
    int32_t cols = 0, dtype = 0;
    status = guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        Matrix m(path, dataset);
        dtype = dtype_of(m.type.get());
        rows = static_cast<int64_t>(m.extents[0]);
        cols = static_cast<int32_t>(m.extents[1]);
        if (n != rows)
            fail(HPD_EINVAL, "labels buffer holds " + std::to_string(n) + " entries, the dataset " +
                                 std::to_string(rows) + " rows");
        rows_buf.resize(static_cast<size_t>(rows) * static_cast<size_t>(cols) * dtype_size(dtype));
        read_into(m, native_type(dtype), rows_buf.data());
    });
    if (status != HPD_OK) return status;
    status = hpd_cluster(ctx, dtype, rows_buf.data(), rows, cols, eps, min_pts, labels_out);
    if (status != HPD_OK) g_error = hpd_last_error();
    return status;
}

int hpd_h5_write_clusters(const char* path, const char* dataset, const int64_t* labels,
                          int64_t n) {
    return guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        if (n < 0) fail(HPD_EINVAL, "negative label count");
        if (n > 0) need(labels, "labels");
        Handle file(open_for_write(path), H5Fclose);
        hsize_t count = static_cast<hsize_t>(n);
        Handle mem_space(H5Screate_simple(1, &count, nullptr), H5Sclose);
        {
            // a dataset of that name may exist already: then the create
            // fails and the existing one is written (io.h:170-177)
            Handle created(H5Dcreate2(file.get(), dataset, H5T_NATIVE_LONG, mem_space.get(),
                                      H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                           H5Dclose);
        }
        Handle dset(H5Dopen2(file.get(), dataset, H5P_DEFAULT), H5Dclose);
        if (!dset.ok()) fail(HPD_EINVAL, std::string("Could not open dataset ") + dataset);
        Handle file_space(H5Dget_space(dset.get()), H5Sclose);
        if (!file_space.ok())
            fail(HPD_EIO, "Could not retrieve dataspace extent, file seems to be corrupted");
        const hsize_t offset = 0;
        const herr_t sel = H5Sselect_hyperslab(file_space.get(), H5S_SELECT_SET, &offset, nullptr,
                                               &count, nullptr);
        if (sel < 0 || H5Dwrite(dset.get(), H5T_NATIVE_LONG, mem_space.get(), file_space.get(),
                                H5P_DEFAULT, labels) < 0)
            fail(HPD_EIO, "Failed to write data, target data set is probably of wrong size or type");
    });
}

int hpd_h5_clusters_length(const char* path, const char* dataset, int64_t* n) {
    return guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        need(n, "n");
        Vector v(path, dataset);
        *n = static_cast<int64_t>(v.count);
    });
}

int hpd_h5_read_clusters(const char* path, const char* dataset, int64_t* labels, int64_t n) {
    return guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        Vector v(path, dataset);
        if (n < 0 || static_cast<hsize_t>(n) != v.count)
            fail(HPD_EINVAL, "dataset " + std::string(dataset) + " holds " +
                                 std::to_string(v.count) + " entries, not " + std::to_string(n));
        if (v.count == 0) return;
        need(labels, "labels");
        if (H5Dread(v.dset.get(), H5T_NATIVE_LONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, labels) < 0)
            fail(HPD_EIO, "Failed to read data");
    });
}

int hpd_h5_write_matrix(const char* path, const char* dataset, int32_t dtype, const void* data,
                        int64_t rows, int32_t cols) {
    return guarded([&] {
        need(path, "path");
        need(dataset, "dataset");
        const hid_t mem = native_type(dtype);
        if (rows < 0 || cols < 0) fail(HPD_EINVAL, "negative data shape");
        if (rows > 0 && cols > 0) need(data, "data");
        Handle file(open_for_write(path), H5Fclose);
        if (H5Lexists(file.get(), dataset, H5P_DEFAULT) > 0)
            fail(HPD_EINVAL, std::string("dataset ") + dataset + " exists already");
        const hsize_t dims[2] = {static_cast<hsize_t>(rows), static_cast<hsize_t>(cols)};
        Handle space(H5Screate_simple(2, dims, nullptr), H5Sclose);
        Handle dset(H5Dcreate2(file.get(), dataset, mem, space.get(), H5P_DEFAULT, H5P_DEFAULT,
                               H5P_DEFAULT),
                    H5Dclose);
        if (!dset.ok()) fail(HPD_EIO, std::string("Could not create dataset ") + dataset);
        if (rows > 0 && cols > 0 &&
            H5Dwrite(dset.get(), mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            fail(HPD_EIO, "Failed to write data");
    });
}

}  // extern "C"
