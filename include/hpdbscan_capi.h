/*
 * C-ABI of the MI355X HPDBSCAN library (libhpdbscan_amd.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an hpd_status;
 * on failure hpd_last_error() returns a message for the calling thread.
 *
 * The reference interface each entry point replaces:
 *   hpd_cluster            HPDBSCAN::cluster<T>(T* data, int dim0, int dim1,
 *                          int threads)          include/hpdbscan.h:475-486
 *                          (and HPDBSCAN::HPDBSCAN(float, size_t) argument
 *                          checks, include/hpdbscan.h:288-298)
 *   hpd_cluster_device     the same, with the rows already in device memory
 *                          and the labels left there
 *   hpd_context_*          no counterpart: the reference keeps no state
 *                          between calls; the context owns the device
 *                          workspace and the stream that replace the OpenMP
 *                          team of include/hpdbscan.h:364
 */
#ifndef HPDBSCAN_CAPI_H
#define HPDBSCAN_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hpd_status {
    HPD_OK = 0,
    HPD_EINVAL = 1,      /* maps to std::invalid_argument */
    HPD_ENOMEM = 2,      /* device or host allocation failed */
    HPD_EDEVICE = 3,     /* HIP runtime error or no usable device */
    HPD_EUNSUPPORTED = 4, /* valid input this build does not handle */
    HPD_ETIMEOUT = 5,     /* hpd_server_wait: the request is not finished yet */
    HPD_EIO = 6           /* file I/O failed (libhpdbscan_amd_h5.so); maps to
                           * std::runtime_error */
} hpd_status;

/* Element types of the input rows; the reference instantiates the same ten
 * (src/hpdbscan.cpp:19-30). */
typedef enum hpd_dtype {
    HPD_DTYPE_F32 = 0,
    HPD_DTYPE_F64 = 1,
    HPD_DTYPE_I8 = 2,
    HPD_DTYPE_I16 = 3,
    HPD_DTYPE_I32 = 4,
    HPD_DTYPE_I64 = 5,
    HPD_DTYPE_U8 = 6,
    HPD_DTYPE_U16 = 7,
    HPD_DTYPE_U32 = 8,
    HPD_DTYPE_U64 = 9
} hpd_dtype;

/* How a border point that lies within eps of cores of two or more clusters
 * picks its cluster.  The reference (hpdbscan.h:80-91 with rules.h) takes the
 * cluster of the adjacent core that was processed with the smallest label,
 * which depends on the OpenMP schedule when threads > 1. */
typedef enum hpd_border_rule {
    /* the adjacent cluster with the smallest id; one extra pass, no
     * dependence on scheduling */
    HPD_BORDER_MIN_CLUSTER = 0,
    /* the cluster the reference assigns when run with one thread */
    HPD_BORDER_ONE_THREAD = 1
} hpd_border_rule;

typedef struct hpd_context hpd_context;

/* Per-call figures of the last clustering run on a context. */
typedef struct hpd_run_stats {
    int64_t n;                   /* rows */
    int32_t d;                   /* dimensions */
    int32_t key_bits;            /* significant bits of the cell id */
    uint64_t total_cells;        /* product of cells per dimension */
    uint64_t occupied_cells;     /* cells holding at least one row */
    uint64_t core_points;
    uint64_t noise_points;
    uint64_t ambiguous_borders;  /* border rows touching >= 2 clusters */
    uint64_t sum_candidates;     /* Σ_p |candidate list of p| (the
                                    reference's distance tests) */
    /* device time per phase in milliseconds (HIP events on the context's
       stream; zero unless hpd_context_set_timing(ctx, 1)) */
    double ms_grid;              /* min/max, cell ids */
    double ms_sort;              /* radix sort by cell, gather */
    double ms_index;             /* occupied cells, neighbour ranges */
    double ms_core;              /* neighbour counts, core flags */
    double ms_union;             /* core-core union-find */
    double ms_border;            /* roots, border labels, scatter */
    double ms_total;             /* first to last event */
    double ms_hook_kernel;       /* box union: hook forest and its roots */
    double ms_union_kernel;      /* box union: union of the remaining pairs */
    double ms_border_kernel;     /* border / noise labels of non-core rows */
    uint64_t boxes;              /* sub-boxes of the occupied cells */
    int32_t box_union;           /* 1 if the union ran on sub-boxes */
    uint64_t neighbor_rows;      /* [first, end) slot ranges of all cells */
    uint64_t open_box_pairs;     /* box pairs the hook forest left open */
    double ms_core_kernel;       /* the core count kernel alone (part of ms_core) */
    uint64_t core_tests;         /* row-row distance tests the core count executed
                                    (box mode, timing level 2; bounding boxes
                                    decide the rest) */
} hpd_run_stats;

/* Thread-local message describing the last failure ("" if none). */
const char* hpd_last_error(void);

/* Library version string. */
const char* hpd_version(void);

/* Number of usable HIP devices (0 if none). */
int hpd_device_count(void);

/* Create a context bound to HIP device `device`.  The context owns a HIP
 * stream and a grow-only device workspace; it is not thread-safe (use one
 * per host thread). */
int hpd_context_create(int device, hpd_context** out);
int hpd_context_destroy(hpd_context* ctx);

/* Options. */
int hpd_context_set_border_rule(hpd_context* ctx, int rule);
/* Phase timing: 1 records HIP events between the phases (ms_* of the
 * statistics); 2 also runs an instrumented core count kernel that counts
 * the distance tests it executes (core_tests; slower, for measurement). */
int hpd_context_set_timing(hpd_context* ctx, int enabled);
int hpd_context_get_stats(const hpd_context* ctx, hpd_run_stats* out);

/* Graph mode (no counterpart in the reference): with enabled != 0,
 * floating-point runs of this context (plain and batched calls, smallest-
 * cluster border rule) take the static pipeline: every table is sized by its
 * worst case from the number of rows, jobs and dimensions and every count
 * stays in device memory, so no host read-back happens between the first
 * kernel and the last.  The sequence is captured once in a hipGraph and
 * replayed by later calls with the same shape (dtype, d, rows, job offsets,
 * data and label buffers, options); eps, minPts and the data itself may
 * change.  The labels are identical to the plain pipeline's.  It uses more
 * device memory (about 3^(d-1) * 16 + 260 bytes per row) and falls back to
 * the plain pipeline, transparently, for loose sub-boxes, grids beyond 64
 * key bits or too many open box pairs; non-finite input then fails as
 * usual.  Disabling it frees the graph. */
int hpd_context_set_graph(hpd_context* ctx, int enabled);
/* Graph launches, captures and fallbacks to the plain pipeline so far. */
int hpd_context_get_graph_stats(const hpd_context* ctx, uint64_t* launches, uint64_t* captures,
                                uint64_t* fallbacks);

/* Work split of the neighbour kernels on sub-boxes (floating-point data):
 * cells with at most sparse_max rows (0..64) are processed one GPU thread
 * per row in the border pass, and in the core count too when sparse_core is
 * nonzero; other cells one 64-lane wavefront per 64-row chunk.  Results do
 * not depend on it; hpdbscan_amd.autotune picks it for a workload.  No
 * counterpart in the reference (OpenMP schedule(dynamic, 32), hpdbscan.h:65). */
int hpd_context_set_work_split(hpd_context* ctx, int32_t sparse_max, int sparse_core);
int hpd_context_get_work_split(const hpd_context* ctx, int32_t* sparse_max, int* sparse_core);

/* Most box pairs the union keeps in its open-pair lists (0: the default,
 * 2n for n rows, which the lists' memory holds).  When a run has more, the
 * union walks every box's neighbourhood instead (decided on the device).
 * Results do not depend on it.  No counterpart in the reference. */
int hpd_context_set_open_pair_limit(hpd_context* ctx, uint64_t limit);

/* Sub-boxes per cell and dimension (the tile of the pair kernels): each
 * eps-cell is split into parts^d boxes, each box's bounding box deciding
 * whole groups of distance tests.  0 (the default) takes 2 up to 3-D and 3
 * above, which keeps every box's diagonal below eps; 1 to 8 set it.  Boxes
 * wider than eps fall back to the row-level union.  Results do not depend
 * on it; hpdbscan_amd.autotune picks it for a workload.  No counterpart in
 * the reference (its cells are the only tiles, spatial_index.h:144-165). */
int hpd_context_set_box_split(hpd_context* ctx, int32_t parts);
int hpd_context_get_box_split(const hpd_context* ctx, int32_t* parts);

/*
 * Cluster n rows of d values of type `dtype` (row-major, host memory).
 * labels_out (host, n entries) receives the reference's labels in input
 * order: 0 noise, -k core point of cluster k, +k border point of cluster k,
 * where k - 1 is the smallest sorted position of a core point of the
 * cluster (include/hpdbscan.h:80, spatial_index.h:743, README.md:54-56).
 *
 * Errors: HPD_EINVAL if eps is not > 0 (message "epsilon needs to be
 * positive", as hpdbscan.h:295-297), if n < 0, d < 1, or a pointer is NULL
 * while n > 0.  NaN eps is rejected too (the reference accepts it and then
 * computes garbage grid sizes).
 */
int hpd_cluster(hpd_context* ctx, int dtype, const void* data, int64_t n, int32_t d, float eps,
                uint64_t min_pts, int64_t* labels_out);

/* Same as hpd_cluster with `data` (n*d values) and `labels_out` (n int64) in
 * device memory of the context's device.  Work is ordered after all prior
 * work on `stream` (a hipStream_t, may be NULL for the context's own stream)
 * and labels_out is complete when the call returns. */
int hpd_cluster_device(hpd_context* ctx, int dtype, const void* data, int64_t n, int32_t d,
                       float eps, uint64_t min_pts, int64_t* labels_out, void* stream);

/* hpd_cluster_device without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, same labels.
 * Every workspace buffer comes from a static device array of 128 MiB that
 * the library holds on each device it runs on until the process ends; the
 * call allocates neither device nor host memory.  Inputs whose workspace
 * does not fit (about 400k rows of 3-D float32 data; the bound depends on n,
 * d, the element type and the border rule) fail with HPD_EUNSUPPORTED before
 * any work, or, when the neighbour tables of the data do not fit the rest,
 * before any label is written.  All calls on one device share the array:
 * calls that overlap in time on one device (from several threads, contexts
 * or streams) race and must be serialised by the caller. */
int hpd_cluster_device_noalloc(hpd_context* ctx, int dtype, const void* data, int64_t n,
                               int32_t d, float eps, uint64_t min_pts, int64_t* labels_out,
                               void* stream);

/*
 * Batched clustering: njobs independent point clouds of the same dtype and
 * dimension, concatenated row-wise; job j is rows [offsets[j], offsets[j+1])
 * (offsets has njobs + 1 entries, offsets[0] = 0) with its own eps[j] and
 * min_pts[j].  labels_out[offsets[j] ..] receives exactly the labels
 * hpd_cluster would return for job j alone (cluster ids count from the
 * job's own sorted positions).  All jobs run through one launch of every
 * kernel.  No counterpart in the reference, which clusters one dataset per
 * call (hpdbscan.h:475-486); offsets, eps and min_pts are host arrays.
 */
int hpd_cluster_batch(hpd_context* ctx, int dtype, const void* data, const int64_t* offsets,
                      int32_t njobs, int32_t d, const float* eps, const uint64_t* min_pts,
                      int64_t* labels_out);

/* Same with `data` and `labels_out` in device memory; ordered on `stream`
 * (NULL: the context's stream), complete on return. */
int hpd_cluster_batch_device(hpd_context* ctx, int dtype, const void* data,
                             const int64_t* offsets, int32_t njobs, int32_t d, const float* eps,
                             const uint64_t* min_pts, int64_t* labels_out, void* stream);

/*
 * Continuous-batching server (no counterpart in the reference).
 *
 * hpd_server_submit copies one point cloud and returns a ticket at once
 * (hpd_server_submit_device takes one already in device memory).  A
 * scheduler groups waiting requests of the same dtype and dimension into a
 * batch whenever one of `lanes` workers is idle; each worker owns a context
 * (HIP stream + workspace) and runs its batch with hpd_cluster_batch, so the
 * batches of different lanes overlap on the GPU.  Admission charges every
 * request rows * (bytes_per_row_estimate + row bytes) against
 * hbm_budget_bytes; a batch waits until it fits beside the running ones, a
 * request larger than the whole budget (or than max_batch_rows) is refused
 * with HPD_ENOMEM at submit.  Results are identical to hpd_cluster on the
 * same request alone.  Destroying the server finishes running batches and
 * drops queued requests; do not wait on tickets while destroying.
 */
typedef struct hpd_server hpd_server;

typedef struct hpd_server_config {
    int device;                     /* HIP device */
    int lanes;                      /* concurrent batches (contexts / streams) */
    size_t hbm_budget_bytes;        /* device memory all running batches may use */
    size_t bytes_per_row_estimate;  /* workspace charged per row (plus the row itself) */
    int32_t max_batch_jobs;         /* requests per batch */
    int64_t max_batch_rows;         /* rows per batch */
} hpd_server_config;

typedef struct hpd_request_info {
    int64_t rows;
    int32_t batch_jobs;  /* requests in the batch it ran in */
    double queue_ms;     /* submit -> batch start */
    double latency_ms;   /* submit -> result ready */
} hpd_request_info;

typedef struct hpd_server_stats {
    uint64_t completed, failed, batches;
    double mean_batch_jobs;
    int64_t rows;
    double p50_ms, p99_ms, max_ms;  /* submit -> result latency of completed requests */
    double wall_s;                  /* first submit -> last completion */
    double jobs_per_s, mrows_per_s;
    uint64_t queued;
    size_t charged_bytes;
} hpd_server_stats;

void hpd_server_default_config(hpd_server_config* cfg);
int hpd_server_create(const hpd_server_config* cfg, hpd_server** out);
int hpd_server_submit(hpd_server* srv, int dtype, const void* data, int64_t n, int32_t d,
                      float eps, uint64_t min_pts, uint64_t* ticket);
/* A request whose rows and labels stay in device memory (of the server's
 * device): its rows are copied device to device into its batch and its n
 * labels device to device into labels_out, so nothing crosses PCIe.  Both
 * buffers must stay valid, and the rows unchanged, until hpd_server_wait
 * returns for the ticket. */
int hpd_server_submit_device(hpd_server* srv, int dtype, const void* data, int64_t n, int32_t d,
                             float eps, uint64_t min_pts, int64_t* labels_out, uint64_t* ticket);
/* Wait for a ticket (timeout_ms < 0: forever) and copy its n labels (host
 * requests; a device request's labels are in its own buffer, and labels_out
 * and n are ignored).  HPD_ETIMEOUT leaves the ticket valid; any other
 * return consumes it. */
int hpd_server_wait(hpd_server* srv, uint64_t ticket, int64_t* labels_out, int64_t n,
                    double timeout_ms, hpd_request_info* info);
int hpd_server_get_stats(hpd_server* srv, hpd_server_stats* out);
int hpd_server_destroy(hpd_server* srv);
/* Thread-local message of the last failing hpd_server_* call. */
const char* hpd_server_last_error(void);

/* Size limit of one run's neighbour table (no counterpart in the reference,
 * which builds candidate lists per thread, spatial_index.h:682-737): a run
 * whose occupied cells have more than 2^31-1 non-empty neighbour rows in
 * total (rows of 3 adjacent cells; at most occupied_cells * 3^(d-1)) fails
 * with HPD_EUNSUPPORTED before any table is sized from the total.  This
 * applies the same check to a given total: HPD_OK if it fits. */
int hpd_check_neighbor_rows(uint64_t neighbor_rows);

/* Device-memory helpers for callers of hpd_cluster_device that have no HIP
 * runtime of their own (no counterpart in the reference).  Allocations are
 * on the context's device; copies are ordered on the context's stream and
 * complete before the call returns. */
int hpd_device_alloc(hpd_context* ctx, size_t bytes, void** out);
int hpd_device_free(hpd_context* ctx, void* ptr);
int hpd_memcpy_to_device(hpd_context* ctx, void* dst, const void* src, size_t bytes);
int hpd_memcpy_to_host(hpd_context* ctx, void* dst, const void* src, size_t bytes);
/* The context's hipStream_t. */
int hpd_context_stream(hpd_context* ctx, void** stream_out);
/* Wait for all work on the context's stream. */
int hpd_context_synchronize(hpd_context* ctx);

#ifdef __cplusplus
}
#endif

#endif /* HPDBSCAN_CAPI_H */
