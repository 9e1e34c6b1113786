// C-ABI of libhpdbscan_amd.so and the host orchestration of one clustering
// run.  The steps follow the reference's single-node path:
//
//   HPDBSCAN::HPDBSCAN / cluster<T>(T*, int, int)    hpdbscan.h:288-298, 475-486
//   SpatialIndex<T>::SpatialIndex                     spatial_index.h:528-628
//     min/max -> cells per dim -> dimension order -> cell ids -> sort
//   local_dbscan -> apply_rules -> recover_initial_order
//                                                     hpdbscan.h:372-448
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hpdbscan_capi.h"
#include "boxes.hpp"
#include "dbscan.hpp"
#include "grid.hpp"
#include "index.hpp"
#include "internal.hpp"

namespace hpd {

// --------------------------------------------------------------------------
// errors
// --------------------------------------------------------------------------
static thread_local std::string g_last_error;

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw Error(e == hipErrorOutOfMemory ? HPD_ENOMEM : HPD_EDEVICE,
                    std::string(what) + ": " + hipGetErrorString(e));
    }
}

// --------------------------------------------------------------------------
// arena
// --------------------------------------------------------------------------
Arena::~Arena() {
    if (base_ && owned_) (void)hipFree(base_);
}

void Arena::reserve(size_t bytes) {
    used_ = 0;
    if (bytes <= cap_) return;
    if (!owned_)
        throw Error(HPD_EUNSUPPORTED, "the run needs " + std::to_string(bytes) +
                                          " bytes of device workspace here; " +
                                          std::to_string(cap_) + " are available");
    if (base_) {
        HPD_HIP(hipFree(base_));
        base_ = nullptr;
        cap_ = 0;
    }
    const size_t want = align_up(bytes + bytes / 8, size_t(1) << 20);
    void* p = nullptr;
    HPD_HIP(hipMalloc(&p, want));
    base_ = static_cast<char*>(p);
    cap_ = want;
}

void* Arena::take(size_t bytes) {
    const size_t b = align_up(bytes > 0 ? bytes : 1, 256);
    if (used_ + b > cap_) throw Error(HPD_ENOMEM, "internal: device workspace exhausted");
    void* p = base_ + used_;
    used_ += b;
    return p;
}

// --------------------------------------------------------------------------
// host-side arithmetic of the grid (reference build semantics)
// --------------------------------------------------------------------------

// The reference's -Ofast build runs with flush-to-zero and
// denormals-are-zero (crtfastmath.o); these helpers apply the same flushing
// to the few scalar steps evaluated on the host.
static inline float flushf(float x) { return std::fabs(x) < FLT_MIN ? 0.0f * x : x; }
static inline double flushd(double x) { return std::fabs(x) < DBL_MIN ? 0.0 * x : x; }

// float/double -> integer as x86-64 cvttss2si / cvttsd2si produce it
static inline int64_t f2i64_x86(double v) {
    if (!(v >= -9223372036854775808.0 && v < 9223372036854775808.0)) return INT64_MIN;
    return static_cast<int64_t>(v);
}
static inline int32_t f2i32_x86(double v) {
    if (!(v >= -2147483648.0 && v < 2147483648.0)) return INT32_MIN;
    return static_cast<int32_t>(v);
}
static inline uint64_t f2u64_x86(double v) {
    if (v < 9223372036854775808.0) return static_cast<uint64_t>(f2i64_x86(v));
    return static_cast<uint64_t>(f2i64_x86(v - 9223372036854775808.0)) ^ 0x8000000000000000ull;
}

template <typename T>
static float int_diff_f32_host(T x, T mn) {
    using P = decltype(x - mn);
    if constexpr (sizeof(T) < sizeof(int)) {
        return static_cast<float>(static_cast<int>(x) - static_cast<int>(mn));
    } else {
        using U = std::make_unsigned_t<P>;
        return static_cast<float>(static_cast<P>(static_cast<U>(x) - static_cast<U>(mn)));
    }
}

// cells per dimension, ceil((max - min) / eps) + 1 (spatial_index.h:124-131)
template <typename T>
static uint64_t cells_of(T mx, T mn, float rcp_f, double rcp_d) {
    if constexpr (std::is_same<T, float>::value) {
        const float v = flushf(flushf(flushf(mx) - flushf(mn)) * rcp_f);
        return f2u64_x86(static_cast<double>(std::ceil(v))) + 1;
    } else if constexpr (std::is_same<T, double>::value) {
        const double v = flushd(flushd(flushd(mx) - flushd(mn)) * rcp_d);
        return f2u64_x86(std::ceil(v)) + 1;
    } else {
        const float v = flushf(int_diff_f32_host<T>(mx, mn) * rcp_f);
        return f2u64_x86(static_cast<double>(std::ceil(v))) + 1;
    }
}

// EPS2 = (float)(eps * eps) (hpdbscan.h:53), converted to T when passed to
// region_query (spatial_index.h:739)
template <typename T>
static T eps2_as(float eps) {
    const float f = flushf(flushf(eps) * flushf(eps));
    if constexpr (std::is_same<T, float>::value) {
        return f;
    } else if constexpr (std::is_same<T, double>::value) {
        return static_cast<double>(f);
    } else if constexpr (std::is_same<T, uint64_t>::value) {
        return f2u64_x86(static_cast<double>(f));
    } else if constexpr (std::is_same<T, int64_t>::value) {
        return f2i64_x86(static_cast<double>(f));
    } else if constexpr (std::is_same<T, uint32_t>::value) {
        return static_cast<uint32_t>(f2i64_x86(static_cast<double>(f)));
    } else {
        return static_cast<T>(f2i32_x86(static_cast<double>(f)));
    }
}

template <typename T>
static Geometry make_geometry(const T* mn, const T* mx, int d, float eps, float* rcp_f_out,
                              double* rcp_d_out) {
    Geometry g;
    g.d = d;
    const float rcp_f = flushf(1.0f / flushf(eps));
    const double rcp_d = flushd(1.0 / flushd(static_cast<double>(eps)));
    *rcp_f_out = rcp_f;
    *rcp_d_out = rcp_d;
    uint64_t total = 1;
    bool overflow = false;
    for (int j = 0; j < d; ++j) {
        g.cells[j] = cells_of<T>(mx[j], mn[j], rcp_f, rcp_d);
        uint64_t t;
        if (__builtin_mul_overflow(total, g.cells[j], &t)) overflow = true;
        total *= g.cells[j];
    }
    g.total = total;
    g.overflow = overflow;
    // dimension order: the reference's own call, std::sort ascending by cell
    // count (spatial_index.h:135-139)
    size_t sw[kMaxDims];
    std::iota(sw, sw + d, size_t(0));
    std::sort(sw, sw + d, [&](size_t a, size_t b) { return g.cells[a] < g.cells[b]; });
    uint64_t acc = 1;
    for (int k = 0; k < d; ++k) {
        g.swapped[k] = static_cast<int>(sw[k]);
        g.stride[k] = acc;
        g.ncell[k] = g.cells[g.swapped[k]];
        acc *= g.ncell[k];
    }
    const uint64_t maxkey = total - 1;
    g.key_bits = maxkey == 0 ? 0 : 64 - __builtin_clzll(maxkey);
    return g;
}

// Host array of n elements: inline storage up to N (a single job needs no
// heap allocation), the heap beyond.
template <typename U, size_t N>
class HostArray {
public:
    explicit HostArray(size_t n) : n_(n) {
        if (n > N) heap_.resize(n);
        p_ = n > N ? heap_.data() : inl_;
        for (size_t i = 0; i < n; ++i) p_[i] = U{};
    }
    HostArray(const HostArray&) = delete;
    HostArray& operator=(const HostArray&) = delete;
    U* data() { return p_; }
    U& operator[](size_t i) { return p_[i]; }
    size_t size() const { return n_; }

private:
    size_t n_;
    U inl_[N];
    std::vector<U> heap_;
    U* p_;
};

// --------------------------------------------------------------------------
// context
// --------------------------------------------------------------------------
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena main;
    Arena aux;
    Arena io;  // input rows and labels of the host-memory entry point
    // the arenas a run takes its workspace from: main / aux, or the static
    // scratch (hpd_cluster_device_noalloc)
    Arena* main_use = &main;
    Arena* aux_use = &aux;
    int border_rule = HPD_BORDER_MIN_CLUSTER;
    // box mode: cells with at most this many rows are handled one thread
    // per row in the border pass (and in the core count if sparse_core),
    // the others one wave per 64-row chunk (hpd_context_set_work_split;
    // defaults picked by hpdbscan_amd.autotune on the 10M x 3-D workload)
    int32_t sparse_max = 8;
    // sub-boxes per cell and dimension (hpd_context_set_box_split; 0: 2 up
    // to 3-D, 3 above, which keeps every sub-box's diagonal below eps)
    int32_t box_split = 0;
    uint64_t open_pair_limit = 0;  // 0: 2n
    bool sparse_core = false;
    bool timing = false;
    bool count_tests = false;  // hpd_context_set_timing level 2
    hpd_run_stats stats{};
    uint64_t* host_scratch = nullptr;  // pinned, 64 words
    std::vector<hipEvent_t> events;
    // hpd_context_set_graph: floating-point runs go through the static
    // pipeline, captured in a hipGraph and replayed while the shape stays
    bool graph = false;
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        int dtype = -1, d = 0, border_rule = -1, sparse_max = -1, box_split = -1;
        bool count_tests = false;
        bool sparse_core = false;
        int64_t n = 0;
        const void* data = nullptr;
        const void* labels = nullptr;
        const void* arena_main = nullptr;
        const void* arena_aux = nullptr;
        std::vector<int64_t> offsets;
    } graph_entry;
    uint64_t graph_launches = 0, graph_captures = 0, graph_fallbacks = 0;
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        HPD_HIP(hipGetDevice(&prev));
        if (prev != dev) HPD_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

// Named points of a run at which HIP events are recorded on the run's
// stream (when timing is on).
enum Mark : int {
    M_START,
    M_GRID,        // min/max and cell ids done
    M_SORT,        // both sorts and the gather done
    M_INDEX,       // cells, neighbour rows, boxes done
    M_CORE0,       // the core count kernel starts (after the full boxes' flags)
    M_CORE,        // core flags done
    M_HOOK0,       // box union: hook forest starts
    M_HOOK1,       // box union: hook forest and its roots done
    M_JOIN1,       // box union: union kernel done
    M_UNION,       // clusters of all cores known
    M_BORDER0,     // core labels written, border pass starts
    M_LABELS,      // all labels written
    M_COUNT
};

class PhaseTimer {
public:
    PhaseTimer(Context& c, hipStream_t s) : ctx_(c), s_(s) {}
    void mark(Mark m) {
        hipEvent_t e = event(m);
        if (e) HPD_HIP(hipEventRecord(e, s_));
    }
    // the event of mark m for the caller to record (null when timing is off)
    hipEvent_t event(Mark m) {
        if (!ctx_.timing) return nullptr;
        while (ctx_.events.size() < M_COUNT) {
            hipEvent_t e;
            HPD_HIP(hipEventCreate(&e));
            ctx_.events.push_back(e);
        }
        seen_[m] = true;
        return ctx_.events[m];
    }
    // milliseconds between two recorded marks (after a stream sync); 0 if
    // either was not recorded in this run
    double ms(Mark i, Mark j) const {
        if (!ctx_.timing || !seen_[i] || !seen_[j]) return 0.0;
        float v = 0.0f;
        HPD_HIP(hipEventElapsedTime(&v, ctx_.events[i], ctx_.events[j]));
        return static_cast<double>(v);
    }

private:
    Context& ctx_;
    hipStream_t s_;
    bool seen_[M_COUNT] = {};
};

// --------------------------------------------------------------------------
// one clustering run
// --------------------------------------------------------------------------
// Neighbour rows of one run are indexed with 32-bit offsets (signed in the
// kernels' loops).
constexpr uint64_t kMaxNeighborRows = 0x7fffffffull;

static void check_neighbor_rows(uint64_t total) {
    if (total > kMaxNeighborRows)
        throw Error(HPD_EUNSUPPORTED,
                    "the run has " + std::to_string(total) +
                        " neighbour rows (non-empty rows of 3 adjacent cells, summed over the "
                        "occupied cells); at most 2^31-1 are supported");
}

// sub-boxes per cell and dimension of a run of dimension d
static int split_of(const Context& c, int d) { return c.box_split > 0 ? c.box_split : box_split(d); }

static inline int bits_for(uint64_t max_value) {
    return max_value == 0 ? 0 : 64 - __builtin_clzll(max_value);
}

// worst-case device workspace of one run with n rows
template <typename T, int D>
static size_t workspace_bytes(int64_t n, int64_t njobs, bool one_thread) {
    const size_t nn = static_cast<size_t>(n);
    const size_t L = align_up(nn, 16);
    const size_t hist_words = radix_hist_words(n);
    const size_t tmp_words = scan_tmp_words(static_cast<int64_t>(std::max(hist_words, 2 * nn + 2)));
    size_t b = 0;
    b += 2 * 8 * nn + 2 * 4 * nn;              // keys / values of the sort
    b += 4 * hist_words + 4 * tmp_words;       // radix histogram, scan scratch
    b += sizeof(T) * L * D;                    // coordinates in slot order
    b += 4 * (2 * nn + 2);                     // run flags + scan
    b += 8 * nn + 4 * 4 * (nn + 1);            // ucell, ustart, nranges, nchunks, core chunks
    b += L + 3 * 4 * L + 3 * 4 * nn;           // core, parent, rsnap, rootslot, cmin/cmax/rcenter
    b += 8 * nn + 2 * 4 * nn + 3 * 4 * (nn + 1) + 5 * 4 * nn + nn;  // scratch64, positions, box tables
    b += 2 * D * 8 * nn;                       // box bounding boxes
    b += gather_part_bytes(n, D, static_cast<int>(sizeof(T)));  // their per-block shares
    if (sort_moves_rows<T, D>()) b += 2 * sizeof(T) * D * nn;   // rows between sort passes
    const size_t nj = static_cast<size_t>(njobs);
    const size_t mm_blocks = nn / static_cast<size_t>(minmax_rows_per_block()) + nj + 1;
    b += sizeof(T) * 2 * D * (mm_blocks + nj) + 4 * (mm_blocks + nj + 2) + 64;  // min/max
    b += sizeof(JobParams<T>) * nj + 4 * nn;   // job table, job of every cell
    if (one_thread) b += 8 * nn + 4 * nn + nn + 4;  // ambiguous rows, f, involved, flag
    b += 8 * (16 + kNoiseStripes);            // statistics, counters, noise counters
    b += 54 * 256 + 64 * 1024;                 // alignment
    return b;
}

template <typename T, int D, typename K>
static void run_keyed(Context& ctx, const T* pts, int64_t n, const JobParams<T>* jobs,
                      int32_t njobs, int key_bits, uint64_t key_total, bool boxes,
                      int64_t* labels, hipStream_t s, PhaseTimer& tm) {
    Arena& A = *ctx.main_use;
    Arena& X = *ctx.aux_use;
    const size_t nn = static_cast<size_t>(n);
    const int64_t ld = (n + 15) & ~int64_t(15);  // rows are read in aligned batches of <= 16
    const size_t L = static_cast<size_t>(ld);
    // each key buffer is followed by its value buffer, so that the pair the
    // sort leaves free is one block (the dense cell table's room)
    K* keys_a = A.take_n<K>(nn);
    uint32_t* vals_a = A.take_n<uint32_t>(nn);
    K* keys_b = A.take_n<K>(nn);
    uint32_t* vals_b = A.take_n<uint32_t>(nn);
    const size_t hist_words = radix_hist_words(n);
    uint32_t* hist = A.take_n<uint32_t>(hist_words);
    const size_t tmp_words = scan_tmp_words(static_cast<int64_t>(std::max(hist_words, 2 * nn + 2)));
    uint32_t* tmp = A.take_n<uint32_t>(tmp_words);
    T* soa = A.take_n<T>(L * D);
    uint32_t* flags = A.take_n<uint32_t>(2 * nn + 2);
    K* ucell = A.take_n<K>(nn);
    int32_t* ustart = A.take_n<int32_t>(nn + 1);
    uint32_t* nranges = A.take_n<uint32_t>(nn + 1);
    uint32_t* nchunks = A.take_n<uint32_t>(nn + 1);
    uint8_t* core = A.take_n<uint8_t>(L);
    int32_t* parent = A.take_n<int32_t>(L);
    int32_t* rsnap = A.take_n<int32_t>(L);
    int32_t* rootslot = A.take_n<int32_t>(L);
    int32_t* cmin = A.take_n<int32_t>(nn);
    int32_t* cmax = A.take_n<int32_t>(nn);
    int32_t* rcenter = A.take_n<int32_t>(nn);
    int32_t* job_of_cell = A.take_n<int32_t>(nn);
    unsigned long long* dstats = A.take_n<unsigned long long>(16 + kNoiseStripes);
    uint32_t* counters = reinterpret_cast<uint32_t*>(dstats + 8);  // m, ranges, items, boxes, box items, loose
    unsigned long long* noise_stripes = dstats + 16;
    HPD_HIP(hipMemsetAsync(dstats, 0, (16 + kNoiseStripes) * sizeof(unsigned long long), s));
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx.host_scratch);

    // cell ids (spatial_index.h:144-165), offset by the job's key base; in
    // box mode followed by the row's sub-box digit, so that one stable sort
    // orders the rows by cell, then sub-box, then input row
    const int sub_k = split_of(ctx, D);
    const int sub_bits = boxes ? box_sub_bits(D, sub_k) : 0;
    launch_cell_ids<T, D, K>(pts, n, jobs, njobs, sub_k, sub_bits, keys_a, vals_a, s);
    tm.mark(M_GRID);

    // stable sort (spatial_index.h:167-216): slot -> input row; jobs stay
    // contiguous and in order because their key ranges are.  Rows of up to
    // 16 bytes travel with their pairs into soa (a gather of random rows
    // afterwards reads a whole cache line per row).
    const int sort_bits = key_bits + sub_bits;
    const bool rows_sorted = sort_moves_rows<T, D>() && n > 1 && sort_bits > 0;
    bool in_b = false;
    if (rows_sorted) {
        const int passes = radix_pass_count(sort_bits);
        T* rows_a = passes > 1 ? A.take_n<T>(nn * D) : nullptr;
        T* rows_b = passes > 2 ? A.take_n<T>(nn * D) : nullptr;
        in_b = radix_sort_pairs_rows<K, T, D>(keys_a, vals_a, keys_b, vals_b, n, sort_bits, hist,
                                              tmp, pts, rows_a, rows_b, soa, ld, s);
    } else {
        in_b = radix_sort_pairs<K>(keys_a, vals_a, keys_b, vals_b, n, sort_bits, hist, tmp, s);
    }
    K* keys = in_b ? keys_b : keys_a;
    uint32_t* inrow_of_slot = in_b ? vals_b : vals_a;
    K* keys_free = in_b ? keys_a : keys_b;
    uint32_t* vals_free = in_b ? vals_a : vals_b;

    // occupied cells (spatial_index.h:167-182); their order and the order of
    // input rows inside them are the reference's sorted order (box mode:
    // from the boxes, below)
    if (!boxes) launch_cells<K>(keys, n, sub_bits, flags, tmp, ucell, ustart, &counters[0], s);

    // box mode: slots are rows sorted by cell, sub-box, input row; the
    // reference's sorted positions are only materialised where needed
    uint32_t* pos_of_slot = nullptr;
    int32_t *box_of_slot = nullptr, *bstart = nullptr, *bcell = nullptr, *cbox = nullptr;
    uint32_t* bcnt = nullptr;
    int32_t *bmincore = nullptr, *bparent = nullptr, *compmin = nullptr, *bcluster = nullptr;
    uint8_t* btight = nullptr;
    void* bbox = nullptr;
    uint64_t* scratch64 = nullptr;  // n words: open-pair lists, cluster keys, position sort
    if (boxes) {
        scratch64 = A.take_n<uint64_t>(nn);
        pos_of_slot = A.take_n<uint32_t>(nn);
        box_of_slot = A.take_n<int32_t>(nn);
        bstart = A.take_n<int32_t>(nn + 1);
        bcell = A.take_n<int32_t>(nn);
        cbox = A.take_n<int32_t>(nn + 1);
        bcnt = A.take_n<uint32_t>(nn + 1);
        bmincore = A.take_n<int32_t>(nn);
        bparent = A.take_n<int32_t>(nn);
        compmin = A.take_n<int32_t>(nn);
        bcluster = A.take_n<int32_t>(nn);
        btight = A.take_n<uint8_t>(nn);
        bbox = A.take(2 * D * sizeof(T) * nn);
        void* bpart = A.take(gather_part_bytes(n, D, static_cast<int>(sizeof(T))));
        // box flags and their scan in the sort's free buffers; then the
        // occupied cells from the boxes (per-box cell flags and their scan
        // in the run-flag buffer)
        launch_box_table<K>(keys, n, nullptr, nullptr, vals_free,
                            reinterpret_cast<uint32_t*>(keys_free), tmp, bstart, bcell, box_of_slot,
                            &counters[3], nullptr, 2 * D, static_cast<int>(sizeof(T)), s, sub_bits,
                            flags);
        launch_cells_of_boxes<K>(keys, n, sub_bits, bstart, &counters[3], flags, flags + nn + 1,
                                 tmp, bcell, ucell, ustart, &counters[0], s);
        launch_gather_boxes<T, D>(rows_sorted ? nullptr : pts, n, inrow_of_slot, soa, ld,
                                  box_of_slot, bstart, bbox, bpart, s);
    } else if (!rows_sorted) {
        launch_gather_soa<T, D>(pts, n, inrow_of_slot, soa, ld, s);
    }
    // one read-back for the occupied cells and the boxes
    HPD_HIP(hipMemcpyAsync(hc, counters, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    const int32_t m = static_cast<int32_t>(hc[0]);
    const int32_t nbox = boxes ? static_cast<int32_t>(hc[3]) : 0;
    // box mode: a dense table of every cell id in the sort's spare key
    // buffer, when the grid is small enough (the neighbour rows then read
    // the table instead of searching the occupied cells)
    int32_t* ctab = nullptr;
    const size_t free_pair = align_up(nn * sizeof(K), 256) + nn * sizeof(uint32_t);
    if (boxes && key_total > 0 && key_total <= free_pair / sizeof(int32_t)) {
        ctab = reinterpret_cast<int32_t*>(keys_free);  // keys_free and vals_free
        launch_cell_table<K>(ucell, host_count(m), ctab, key_total, s);
    }
    tm.mark(M_SORT);

    // neighbour ranges (spatial_index.h:682-737) and work items
    uint32_t* ccnt = boxes ? A.take_n<uint32_t>(nn + 1) : nullptr;
    // the count pass keeps the rows it finds in scratch64 (idle until the
    // union) when every cell's 3^(D-1) rows fit, and the fill copies them
    int32_t stash_rows = 1;
    for (int j = 1; j < D; ++j) stash_rows *= 3;
    int2* stash = scratch64 && static_cast<uint64_t>(m) * static_cast<uint64_t>(stash_rows) <= nn
                      ? reinterpret_cast<int2*>(scratch64)
                      : nullptr;
    // up to 27 rows per cell (D <= 4): the copy puts the nearest rows first
    const bool by_class = stash && D <= 4;
    launch_neighbor_count<K, T>(ucell, ustart, host_count(m), jobs, njobs, job_of_cell, nranges, ccnt,
                                &dstats[0], &dstats[4], stash, stash_rows, rcenter, by_class, s,
                                ctab);
    launch_chunk_counts(ustart, host_count(m), nchunks, s);
    // box mode: dense-cell chunks and sparse-cell rows as separate work
    // lists; their per-cell counts live in the (now unused) run flags
    uint32_t* dcnt = flags;
    uint32_t* scnt = flags + nn + 1;
    if (boxes) launch_split_counts(ustart, host_count(m), ctx.sparse_max, dcnt, scnt, s);
    // the per-cell counts, scanned together (totals at entry m)
    ScanSet counts{};
    int ncounts = 0;
    counts.a[ncounts++] = nranges;
    counts.a[ncounts++] = nchunks;
    if (boxes) {
        counts.a[ncounts++] = ccnt;
        counts.a[ncounts++] = dcnt;
        counts.a[ncounts++] = scnt;
    }
    exclusive_scan_set_u32(counts, ncounts, m, tmp, tmp_words, s);
    if (boxes) {
        // the box chunks (counted and scanned) and the tight flags (their
        // loose count decides the union) before the read-back below
        launch_box_items(bstart, host_count(nbox), bcnt, tmp, &counters[4], s);
        BoxArgs<T> bt;
        bt.nbox = nbox;
        bt.bstart = bstart;
        bt.bcell = bcell;
        bt.bbox = bbox;
        bt.jobs = jobs;
        bt.job_of_cell = job_of_cell;
        launch_box_tight<T, D>(bt, btight, &counters[5], s);
    }
    // the scans' totals into the counters (one launch), then one read-back
    // of the 64-bit row total (dstats[4]) and the counters (dstats + 8)
    WordCopies totals;
    totals.add(&nranges[m], &counters[1]);
    totals.add(&nchunks[m], &counters[2]);
    if (boxes) {
        totals.add(&ccnt[m], &counters[11]);
        totals.add(&dcnt[m], &counters[6]);
        totals.add(&scnt[m], &counters[7]);
    }
    copy_words(totals, s);
    unsigned long long* hs_rows = reinterpret_cast<unsigned long long*>(ctx.host_scratch) + 12;
    HPD_HIP(hipMemcpyAsync(hs_rows, &dstats[4], 10 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    hc = reinterpret_cast<uint32_t*>(hs_rows + 4);  // the counters as read back
    // the scan of nranges and every range index are 32-bit: check the
    // 64-bit total first
    check_neighbor_rows(*hs_rows);
    const uint32_t total_ranges = hc[1];
    const int32_t ncitems = boxes ? static_cast<int32_t>(hc[11]) : 0;
    const int32_t nitems = static_cast<int32_t>(hc[2]);
    const int32_t nditems = boxes ? static_cast<int32_t>(hc[6]) : 0;
    const int32_t nsitems = boxes ? static_cast<int32_t>(hc[7]) : 0;
    const int32_t nbitems = boxes ? static_cast<int32_t>(hc[4]) : 0;
    // every box of two or more rows is tight: the union runs on boxes
    const bool all_tight = boxes && hc[5] == 0;

    X.reserve(2 * align_up(static_cast<size_t>(total_ranges) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nitems) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nbitems) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nditems) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nsitems) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nbitems) * sizeof(int2), 256) +
                    2 * align_up(static_cast<size_t>(ncitems) * sizeof(int2), 256) + 2048);
    int2* ranges = X.take_n<int2>(total_ranges);
    int2* rcells = X.take_n<int2>(total_ranges);
    int2* items = X.take_n<int2>(static_cast<size_t>(nitems));
    int2* bitems = boxes ? X.take_n<int2>(static_cast<size_t>(nbitems)) : nullptr;
    int2* ditems = boxes ? X.take_n<int2>(static_cast<size_t>(nditems)) : nullptr;
    int2* sitems = boxes ? X.take_n<int2>(static_cast<size_t>(nsitems)) : nullptr;
    int2* uitems = boxes ? X.take_n<int2>(static_cast<size_t>(nbitems)) : nullptr;
    int2* citems = boxes ? X.take_n<int2>(static_cast<size_t>(ncitems)) : nullptr;
    int2* citems_need = boxes ? X.take_n<int2>(static_cast<size_t>(ncitems)) : nullptr;
    if (stash)
        launch_neighbor_copy(stash, stash_rows, ustart, m, nranges, ranges, rcells, rcenter,
                             by_class, s);
    else
        launch_neighbor_fill<K, T>(ucell, ustart, host_count(m), jobs, job_of_cell, nranges, ranges,
                                   rcells, rcenter, s, ctab);
    launch_work_items(ustart, host_count(m), nchunks, host_count(nitems), items, s);
    if (boxes) {
        launch_split_items(ustart, host_count(m), ctx.sparse_max, dcnt, scnt, ditems,
                           host_count(nditems), sitems, s);
        launch_expand_chunks(ccnt, host_count(m), ustart, host_count(ncitems), citems, s);
    }

    DbscanArgs<T> a;
    a.soa = soa;
    a.n = n;
    a.ld = ld;
    a.items = items;
    a.nitems = nitems;
    a.ustart = ustart;
    a.roff = nranges;
    a.ranges = ranges;
    a.rcells = rcells;
    a.rcenter = rcenter;
    a.m = m;
    a.jobs = jobs;
    a.njobs = njobs;
    a.job_of_cell = job_of_cell;
    a.pos_of_slot = nullptr;  // slots are positions unless boxes reorder them
    a.inrow = inrow_of_slot;

    BoxArgs<T> ba;
    bool box_union = false;
    if (boxes) {
        launch_cell_boxes(bcell, host_count(nbox), host_count(m), cbox, s);
        launch_box_items_fill(bstart, host_count(nbox), bcnt, host_count(nbitems), bitems, s);
        ba.soa = soa;
        ba.ld = ld;
        ba.nbox = nbox;
        ba.bstart = bstart;
        ba.bcell = bcell;
        ba.cbox = cbox;
        ba.bitems = bitems;
        ba.nbitems = nbitems;
        ba.bbox = bbox;
        ba.jobs = jobs;
        ba.job_of_cell = job_of_cell;
        ba.rcenter = rcenter;
        a.box_of_slot = box_of_slot;
        a.btight = btight;
        a.cbox = cbox;
        a.bstart = bstart;
        a.bcell = bcell;
        a.bbox = bbox;
        a.sparse_core = ctx.sparse_core;
        a.bmincore = bmincore;
        fill_i32(bmincore, INT_MAX, nbox, s);
        a.ditems = ditems;
        a.nditems = nditems;
        a.citems = citems;
        a.ncitems = ncitems;
        a.citems_need = citems_need;
        a.ncitems_need_dev = &counters[12];
        a.sitems = sitems;
        a.nsitems = nsitems;
    }
    tm.mark(M_INDEX);

    launch_core<T, D>(a, core, ctx.count_tests ? &dstats[7] : nullptr, s, tm.event(M_CORE0));
    tm.mark(M_CORE);

    // box lists of the union (box mode leaves the slot-level parent free)
    // (the lists' lengths stay on the device: boxes with cores, and the
    // roots of each join round, bounded by the number of boxes)
    int32_t* blist = parent;
    box_union = all_tight;
    if (box_union) {
        // compmin holds the forest's roots until the clusters are known
        tm.mark(M_HOOK0);
        launch_select_boxes(host_count(nbox), 0, bmincore, blist, &counters[8], s);
        // (the slot-level union's rsnap is free in box mode: the list of
        // the larger boxes)
        launch_select_boxes(host_count(nbox), 2, bmincore, rsnap, &counters[14], s, bstart);
        launch_box_hook<T, D>(ba, nranges, rcells, bmincore, core, blist,
                              Count{nbox, &counters[8], true}, bparent, rsnap, &counters[14], s);
        launch_box_roots(host_count(nbox), bmincore, bparent, compmin, s);
        for (int round = 0; round < kRootJoinRounds<D>; ++round) {
            HPD_HIP(hipMemsetAsync(&counters[9], 0, sizeof(uint32_t), s));
            launch_select_boxes(host_count(nbox), 1, compmin, blist, &counters[9], s);
            launch_box_root_join<T, D>(ba, nranges, rcells, compmin, core, blist,
                                       Count{nbox, &counters[9]}, bparent, s);
            launch_box_roots(host_count(nbox), bmincore, bparent, compmin, s);
        }
        tm.mark(M_HOOK1);
        // the open pairs of every box, as a list in scratch64; offsets in
        // bcnt (free after the box items)
        int32_t* list = reinterpret_cast<int32_t*>(scratch64);
        unsigned long long cap = 2 * static_cast<unsigned long long>(nn);
        if (ctx.open_pair_limit > 0 && ctx.open_pair_limit < cap) cap = ctx.open_pair_limit;
        launch_box_scan<T, D>(ba, nranges, rcells, compmin, bcnt, nullptr, &dstats[5], s);
        // offsets (meaningful when the total fits; the kernels that use them
        // check the 64-bit total on the device, and the full join runs
        // instead when it does not fit) and the chunks of the boxes that
        // have open pairs
        exclusive_scan_u32(bcnt, bcnt, nbox, tmp, &bcnt[nbox], s);
        // sparse data: the small boxes' pairs one thread per box
        const int32_t small_rows = small_union_rows(n, nbox);
        launch_select_items(host_count(nbitems), bitems, bcnt, uitems, &counters[10], s, bstart,
                            small_rows);
        launch_box_scan<T, D>(ba, nranges, rcells, compmin, bcnt, list, nullptr, s, &dstats[5], cap);
        launch_box_union_list<T, D>(ba, uitems, Count{nbitems, &counters[10]}, core, bcnt, list,
                                    bparent, s, &dstats[5], cap, small_rows);
        launch_box_join<T, D>(ba, nranges, rcells, bmincore, core, compmin, bparent, s, &dstats[5],
                              cap);
        tm.mark(M_JOIN1);
        launch_box_clusters(host_count(nbox), bmincore, bcell, cbox, bstart, inrow_of_slot, ustart,
                            bparent, reinterpret_cast<unsigned long long*>(scratch64), compmin,
                            bcluster, s);
    } else {
        if (boxes) {
            // a loose box: the slot-level union needs the sorted positions
            launch_positions(n, m, box_of_slot, bcell, inrow_of_slot,
                             reinterpret_cast<uint32_t*>(keys), vals_free,
                             reinterpret_cast<uint32_t*>(keys_free),
                             reinterpret_cast<uint32_t*>(scratch64), hist, tmp, pos_of_slot, s);
            a.pos_of_slot = pos_of_slot;
        }
        if (!compmin) compmin = A.take_n<int32_t>(nn);
        launch_union_points<T, D>(a, core, parent, rsnap, cmin, cmax, compmin, s);
    }
    tm.mark(M_UNION);
    const bool one_thread = ctx.border_rule == HPD_BORDER_ONE_THREAD;
    int2* amb = one_thread ? A.take_n<int2>(nn) : nullptr;
    // box mode reads rootslot only under the one-thread border rule; without
    // it the rows of full boxes are labelled in input order, through a table
    // indexed by cell key in the sort's spare key buffer when the keys fit
    // (launched only when every key fits the table; *full_ok stays 0 else)
    const bool full_rows = box_union && !one_thread &&
                           key_total <= (nn * sizeof(K) / sizeof(int32_t)) >> sub_bits;
    unsigned int* full_ok = full_rows ? &counters[13] : nullptr;
    if (full_rows)
        launch_label_full_rows<T, D, K>(pts, a, sub_k, sub_bits, keys, host_count(nbox),
                                        bcluster, reinterpret_cast<int32_t*>(keys_free),
                                        static_cast<int64_t>(nn * sizeof(K) / 4), labels,
                                        full_ok, s);
    launch_label_cores<T>(a, core, box_union ? bcluster : nullptr, parent, compmin,
                          box_union && !one_thread ? nullptr : rootslot, cmin, cmax, labels,
                          &dstats[1], s, full_ok);
    tm.mark(M_BORDER0);
    launch_border<T, D>(a, core, box_union ? bcluster : nullptr, rootslot, cmin, cmax, labels,
                        &dstats[2], amb, noise_stripes, s);

    unsigned long long* hs = reinterpret_cast<unsigned long long*>(ctx.host_scratch);
    HPD_HIP(hipMemcpyAsync(hs, dstats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    if (box_union) ctx.stats.open_box_pairs = hs[5];
    ctx.stats.core_tests = hs[7];
    if (one_thread && hs[3] > 0) {
        // the reference's choice at one thread for the border rows next to
        // two or more clusters (hpdbscan.h:74-92)
        const int32_t namb = static_cast<int32_t>(hs[3]);
        if (boxes && !a.pos_of_slot) {  // the relaxation compares sorted positions
            launch_positions(n, m, box_of_slot, bcell, inrow_of_slot,
                             reinterpret_cast<uint32_t*>(keys), vals_free,
                             reinterpret_cast<uint32_t*>(keys_free),
                             reinterpret_cast<uint32_t*>(scratch64), hist, tmp, pos_of_slot, s);
            a.pos_of_slot = pos_of_slot;
        }
        int32_t* f = A.take_n<int32_t>(nn);
        uint8_t* involved = A.take_n<uint8_t>(nn);
        unsigned int* changed = A.take_n<unsigned int>(1);
        HPD_HIP(hipMemsetAsync(involved, 0, nn, s));
        launch_one_thread_mark<T, D>(a, core, rootslot, amb, namb, involved, f, s);
        // sweep k leaves every core whose best decreasing path has at most
        // k steps final, so sweep n + 1 at the latest changes nothing.  A
        // sweep that changes nothing is a fixed point; the host reads the
        // flag of the last sweep of a batch, and the batches double (up to
        // 64 sweeps), so a long chain costs few read-backs
        int64_t sweeps = 0;
        for (int batch = 1;; batch = std::min(2 * batch, 64)) {
            if (sweeps > n + 1) throw Error(HPD_EDEVICE, "border relaxation did not converge");
            for (int i = 0; i < batch; ++i) {
                if (i == batch - 1) HPD_HIP(hipMemsetAsync(changed, 0, sizeof(unsigned int), s));
                launch_one_thread_relax<T, D>(a, core, rootslot, involved, f, changed, s);
            }
            sweeps += batch;
            HPD_HIP(hipMemcpyAsync(hc + 32, changed, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
            HPD_HIP(hipStreamSynchronize(s));
            if (hc[32] == 0) break;
        }
        launch_one_thread_assign<T, D>(a, core, rootslot, amb, namb, f, labels, s);
    }
    tm.mark(M_LABELS);
    // labels are complete on return; the timing events are read next
    if (ctx.timing || (one_thread && hs[3] > 0)) HPD_HIP(hipStreamSynchronize(s));
    ctx.stats.sum_candidates = hs[0];
    ctx.stats.core_points = hs[1];
    ctx.stats.noise_points = hs[2];
    ctx.stats.ambiguous_borders = hs[3];
    ctx.stats.occupied_cells = static_cast<uint64_t>(m);
    ctx.stats.neighbor_rows = total_ranges;
    ctx.stats.boxes = static_cast<uint64_t>(nbox);
    ctx.stats.box_union = box_union ? 1 : 0;
}

// --------------------------------------------------------------------------
// the static pipeline (hpd_context_set_graph)
// --------------------------------------------------------------------------
// The same kernels as run_keyed, with every table sized by its worst case
// from n, the number of jobs and D, and every count that run_keyed reads back
// to size the next step left in device memory (a Count with a device
// address, the kernels reading it; wave kernels walk their items from a
// bounded grid).  The grid of every job is computed on the device
// (launch_geometry) and the radix passes beyond its key bits do nothing, so
// no host read-back happens between the first kernel and the last: the
// sequence can be captured in a hipGraph and replayed.  Conditions on which
// run_keyed branches become flags read after the run; any of them hands the
// run to the plain pipeline: non-finite input, grids beyond 64 key bits,
// loose boxes, more open box pairs than the list holds.  Box mode only
// (floating-point data), smallest-cluster border rule.

// Device words the static pipeline keeps its counts in.
enum StaticSlot : int {
    SS_M = 0,          // occupied cells
    SS_RANGES = 1,     // neighbour rows
    SS_ITEMS = 2,      // 64-row chunks of every cell
    SS_NBOX = 3,       // boxes
    SS_BITEMS = 4,     // 64-slot chunks of every box
    SS_LOOSE = 5,      // boxes of two or more rows that are not tight
    SS_DITEMS = 6,     // chunks of dense cells
    SS_SITEMS = 7,     // rows of sparse cells
    SS_NLIST = 8,      // boxes with cores
    SS_JOIN0 = 9,      // roots of the first root-join round
    SS_UITEMS = 10,    // box chunks with open pairs
    SS_CITEMS = 11,    // chunks of cells with >= minPts candidates
    SS_NEED = 12,      // those holding a row outside the full tight boxes
    SS_JOIN1 = 13,     // roots of the second round
    SS_NONFINITE = 14, // non-finite coordinates
    SS_GRID = 15,      // grid beyond 64 key bits
    SS_BITS = 16,      // key bits + sub-box bits
    SS_ROWTAB = 17,    // the row-order label table holds every key
    SS_HOOKBIG = 18,   // boxes with cores of more than 32 rows
    SS_WORDS = 32
};

// worst-case sizes of the static pipeline's work lists
static inline int64_t item_bound(int64_t n) { return n + n / 64 + 16; }  // chunks: <= m + n / 64
template <int D>
static inline int64_t range_bound(int64_t n) {
    int64_t r = n;
    for (int j = 1; j < D; ++j) r *= 3;
    return r;
}

template <typename T, int D>
static size_t static_aux_bytes(int64_t n) {
    return 2 * align_up(static_cast<size_t>(range_bound<D>(n)) * sizeof(int2), 256) +
           7 * align_up(static_cast<size_t>(item_bound(n)) * sizeof(int2), 256) + 4096;
}

template <typename T, int D>
static size_t static_main_bytes(int64_t n, int32_t njobs) {
    return workspace_bytes<T, D>(n, njobs, false) + sizeof(uint64_t) * static_cast<size_t>(njobs) +
           SS_WORDS * 4 + 64 * 1024 + 8 * 256;
}

// The body of a static run on stream s: every launch of the run, nothing
// else (no read-back, no synchronisation, no memset), in the workspace the
// caller reserved and started taking from; jobs / bfirst_dev hold the
// host-made job table (row ranges, eps2, minPts, reciprocals) and min/max
// blocks.  sc (SS_WORDS words) and dstats (16 + kNoiseStripes words) end up
// holding the counts read afterwards.
template <typename T, int D>
static void static_body(Context& ctx, const T* pts, int64_t n, JobParams<T>* jobs,
                        const int32_t* bfirst_dev, int32_t njobs, int32_t nblocks, int64_t* labels,
                        uint32_t* sc, unsigned long long* dstats, hipStream_t s) {
    using K = uint64_t;
    Arena& A = ctx.main;
    Arena& X = ctx.aux;
    X.reset();
    const size_t nn = static_cast<size_t>(n);
    const int64_t ld = (n + 15) & ~int64_t(15);
    const size_t L = static_cast<size_t>(ld);
    const int64_t IB = item_bound(n);
    // zeroed by a kernel (no memset in a graph)
    fill_i32(reinterpret_cast<int32_t*>(sc), 0, SS_WORDS, s);
    fill_i32(reinterpret_cast<int32_t*>(dstats), 0, 2 * (16 + kNoiseStripes), s);
    unsigned long long* noise_stripes = dstats + 16;

    // min / max, then every job's grid on the device
    T* part = A.take_n<T>(static_cast<size_t>(nblocks) * 2 * D + 1);
    T* mm = A.take_n<T>(static_cast<size_t>(njobs) * 2 * D + 16);
    uint64_t* jcells = A.take_n<uint64_t>(static_cast<size_t>(njobs));
    unsigned long long* total_cells = dstats + 6;
    launch_minmax<T, D>(pts, jobs, njobs, bfirst_dev, nblocks, part, mm,
                        reinterpret_cast<int*>(&sc[SS_NONFINITE]), s);
    const int sub_k = split_of(ctx, D);
    const int sub_bits = box_sub_bits(D, sub_k);
    launch_geometry<T, D>(jobs, njobs, mm, sub_bits, jcells, reinterpret_cast<int*>(&sc[SS_GRID]),
                          reinterpret_cast<int*>(&sc[SS_BITS]), total_cells, s);

    K* keys_a = A.take_n<K>(nn);
    K* keys_b = A.take_n<K>(nn);
    uint32_t* vals_a = A.take_n<uint32_t>(nn);
    uint32_t* vals_b = A.take_n<uint32_t>(nn);
    const size_t hist_words = radix_hist_words(n);
    uint32_t* hist = A.take_n<uint32_t>(hist_words);
    const size_t tmp_words = scan_tmp_words(static_cast<int64_t>(std::max(hist_words, 2 * nn + 2)));
    uint32_t* tmp = A.take_n<uint32_t>(tmp_words);
    T* soa = A.take_n<T>(L * D);
    uint32_t* flags = A.take_n<uint32_t>(2 * nn + 2);
    K* ucell = A.take_n<K>(nn);
    int32_t* ustart = A.take_n<int32_t>(nn + 1);
    uint32_t* nranges = A.take_n<uint32_t>(nn + 1);
    uint32_t* nchunks = A.take_n<uint32_t>(nn + 1);
    uint8_t* core = A.take_n<uint8_t>(L);
    int32_t* parent = A.take_n<int32_t>(L);
    int32_t* hook_big = A.take_n<int32_t>(L);  // (rsnap of the slot-level union)
    int32_t* rootslot = A.take_n<int32_t>(L);
    int32_t* cmin = A.take_n<int32_t>(nn);
    int32_t* cmax = A.take_n<int32_t>(nn);
    int32_t* rcenter = A.take_n<int32_t>(nn);
    int32_t* job_of_cell = A.take_n<int32_t>(nn);
    uint64_t* scratch64 = A.take_n<uint64_t>(nn);
    (void)A.take_n<uint32_t>(nn);  // sorted positions (loose boxes only)
    int32_t* box_of_slot = A.take_n<int32_t>(nn);
    int32_t* bstart = A.take_n<int32_t>(nn + 1);
    int32_t* bcell = A.take_n<int32_t>(nn);
    int32_t* cbox = A.take_n<int32_t>(nn + 1);
    uint32_t* bcnt = A.take_n<uint32_t>(nn + 1);
    int32_t* bmincore = A.take_n<int32_t>(nn);
    int32_t* bparent = A.take_n<int32_t>(nn);
    int32_t* compmin = A.take_n<int32_t>(nn);
    int32_t* bcluster = A.take_n<int32_t>(nn);
    uint8_t* btight = A.take_n<uint8_t>(nn);
    void* bbox = A.take(2 * D * sizeof(T) * nn);
    void* bpart = A.take(gather_part_bytes(n, D, static_cast<int>(sizeof(T))));
    uint32_t* ccnt = A.take_n<uint32_t>(nn + 1);

    // cell ids and the sort (passes past the key's bits do nothing)
    launch_cell_ids<T, D, K>(pts, n, jobs, njobs, sub_k, sub_bits, keys_a, vals_a, s);
    radix_sort_pairs_dev<K>(keys_a, vals_a, keys_b, vals_b, n, 64,
                            reinterpret_cast<const int*>(&sc[SS_BITS]), hist, tmp, s);
    K* keys = keys_a;
    uint32_t* inrow_of_slot = vals_a;
    // the boxes, then the occupied cells from them
    launch_box_table<K>(keys, n, nullptr, nullptr, vals_b, reinterpret_cast<uint32_t*>(keys_b), tmp,
                        bstart, bcell, box_of_slot, &sc[SS_NBOX], nullptr, 2 * D,
                        static_cast<int>(sizeof(T)), s, sub_bits, flags);
    launch_cells_of_boxes<K>(keys, n, sub_bits, bstart, &sc[SS_NBOX], flags, flags + nn + 1, tmp,
                             bcell, ucell, ustart, &sc[SS_M], s);
    launch_gather_boxes<T, D>(pts, n, inrow_of_slot, soa, ld, box_of_slot, bstart, bbox, bpart, s);

    // neighbour rows and work lists, counted and laid out on the device
    const Count M{n, &sc[SS_M]};
    const Count NB{n, &sc[SS_NBOX]};
    uint32_t* dcnt = flags;
    uint32_t* scnt = flags + nn + 1;
    launch_neighbor_count<K, T>(ucell, ustart, M, jobs, njobs, job_of_cell, nranges, ccnt,
                                &dstats[0], &dstats[4], nullptr, 1, rcenter, false, s);
    launch_chunk_counts(ustart, M, nchunks, s);
    launch_split_counts(ustart, M, ctx.sparse_max, dcnt, scnt, s);
    ScanSet counts{};
    counts.a[0] = nranges;
    counts.a[1] = nchunks;
    counts.a[2] = ccnt;
    counts.a[3] = dcnt;
    counts.a[4] = scnt;
    exclusive_scan_set_u32(counts, 5, n, tmp, tmp_words, s, &sc[SS_M]);
    copy_at_count(nranges, &sc[SS_M], &sc[SS_RANGES], s);
    copy_at_count(nchunks, &sc[SS_M], &sc[SS_ITEMS], s);
    copy_at_count(ccnt, &sc[SS_M], &sc[SS_CITEMS], s);
    copy_at_count(dcnt, &sc[SS_M], &sc[SS_DITEMS], s);
    copy_at_count(scnt, &sc[SS_M], &sc[SS_SITEMS], s);
    launch_box_items(bstart, NB, bcnt, tmp, &sc[SS_BITEMS], s);

    const int64_t R = range_bound<D>(n);
    int2* ranges = X.take_n<int2>(static_cast<size_t>(R));
    int2* rcells = X.take_n<int2>(static_cast<size_t>(R));
    int2* items = X.take_n<int2>(static_cast<size_t>(IB));
    int2* bitems = X.take_n<int2>(static_cast<size_t>(IB));
    int2* ditems = X.take_n<int2>(static_cast<size_t>(IB));
    int2* sitems = X.take_n<int2>(static_cast<size_t>(IB));
    int2* uitems = X.take_n<int2>(static_cast<size_t>(IB));
    int2* citems = X.take_n<int2>(static_cast<size_t>(IB));
    int2* citems_need = X.take_n<int2>(static_cast<size_t>(IB));
    launch_neighbor_fill<K, T>(ucell, ustart, M, jobs, job_of_cell, nranges, ranges, rcells, rcenter,
                               s);
    launch_work_items(ustart, M, nchunks, Count{IB, &sc[SS_ITEMS]}, items, s);
    launch_split_items(ustart, M, ctx.sparse_max, dcnt, scnt, ditems, Count{IB, &sc[SS_DITEMS]},
                       sitems, s);
    launch_expand_chunks(ccnt, M, ustart, Count{IB, &sc[SS_CITEMS]}, citems, s);
    launch_cell_boxes(bcell, NB, M, cbox, s);
    launch_box_items_fill(bstart, NB, bcnt, Count{IB, &sc[SS_BITEMS]}, bitems, s);

    BoxArgs<T> ba;
    ba.soa = soa;
    ba.ld = ld;
    ba.nbox = static_cast<int32_t>(n);
    ba.nbox_dev = &sc[SS_NBOX];
    ba.bstart = bstart;
    ba.bcell = bcell;
    ba.cbox = cbox;
    ba.bitems = bitems;
    ba.nbitems = static_cast<int32_t>(IB);
    ba.bbox = bbox;
    ba.jobs = jobs;
    ba.job_of_cell = job_of_cell;
    ba.rcenter = rcenter;
    launch_box_tight<T, D>(ba, btight, &sc[SS_LOOSE], s);
    fill_i32(bmincore, INT_MAX, n, s);

    DbscanArgs<T> a;
    a.soa = soa;
    a.n = n;
    a.ld = ld;
    a.items = items;
    a.nitems = static_cast<int32_t>(IB);
    a.ustart = ustart;
    a.roff = nranges;
    a.ranges = ranges;
    a.rcells = rcells;
    a.rcenter = rcenter;
    a.m = static_cast<int32_t>(n);
    a.jobs = jobs;
    a.njobs = njobs;
    a.job_of_cell = job_of_cell;
    a.pos_of_slot = nullptr;
    a.inrow = inrow_of_slot;
    a.box_of_slot = box_of_slot;
    a.btight = btight;
    a.cbox = cbox;
    a.bstart = bstart;
    a.bcell = bcell;
    a.bbox = bbox;
    a.sparse_core = ctx.sparse_core;
    a.bmincore = bmincore;
    a.ditems = ditems;
    a.nditems = static_cast<int32_t>(IB);
    a.nditems_dev = &sc[SS_DITEMS];
    a.citems = citems;
    a.ncitems = static_cast<int32_t>(IB);
    a.ncitems_dev = &sc[SS_CITEMS];
    a.citems_need = citems_need;
    a.ncitems_need_dev = &sc[SS_NEED];
    a.sitems = sitems;
    a.nsitems = static_cast<int32_t>(n);
    a.nsitems_dev = &sc[SS_SITEMS];
    launch_core<T, D>(a, core, ctx.count_tests ? &dstats[7] : nullptr, s);

    // the box union (its result is discarded when a box is loose)
    int32_t* blist = parent;
    launch_select_boxes(NB, 0, bmincore, blist, &sc[SS_NLIST], s);
    launch_select_boxes(NB, 2, bmincore, hook_big, &sc[SS_HOOKBIG], s, bstart);
    launch_box_hook<T, D>(ba, nranges, rcells, bmincore, core, blist, Count{n, &sc[SS_NLIST]},
                          bparent, hook_big, &sc[SS_HOOKBIG], s);
    launch_box_roots(NB, bmincore, bparent, compmin, s);
    for (int round = 0; round < kRootJoinRounds<D>; ++round) {
        uint32_t* cnt = &sc[round == 0 ? SS_JOIN0 : SS_JOIN1];
        launch_select_boxes(NB, 1, compmin, blist, cnt, s);
        launch_box_root_join<T, D>(ba, nranges, rcells, compmin, core, blist, Count{n, cnt}, bparent,
                                   s);
        launch_box_roots(NB, bmincore, bparent, compmin, s);
    }
    int32_t* list = reinterpret_cast<int32_t*>(scratch64);
    const unsigned long long cap = 2 * static_cast<unsigned long long>(nn);
    launch_box_scan<T, D>(ba, nranges, rcells, compmin, bcnt, nullptr, &dstats[5], s);
    exclusive_scan_u32_dev(bcnt, bcnt, NB, tmp, s);
    // (the box count is on the device: the small boxes' pairs always one
    // thread per box, the static pipeline's batches of small clouds being
    // sparse)
    launch_select_items(Count{IB, &sc[SS_BITEMS]}, bitems, bcnt, uitems, &sc[SS_UITEMS], s, bstart,
                        kSmallUnionRows);
    launch_box_scan<T, D>(ba, nranges, rcells, compmin, bcnt, list, nullptr, s, &dstats[5], cap);
    launch_box_union_list<T, D>(ba, uitems, Count{IB, &sc[SS_UITEMS]}, core, bcnt, list, bparent, s,
                                &dstats[5], cap, kSmallUnionRows);
    launch_box_clusters(NB, bmincore, bcell, cbox, bstart, inrow_of_slot, ustart, bparent,
                        reinterpret_cast<unsigned long long*>(scratch64), compmin, bcluster, s);

    // the rows of full boxes labelled in input order when the keys fit the
    // table (in the sort's spare key buffer), the other cores in slot order
    launch_label_full_rows<T, D, K>(pts, a, sub_k, sub_bits, keys, NB, bcluster,
                                    reinterpret_cast<int32_t*>(keys_b),
                                    static_cast<int64_t>(nn * sizeof(K) / 4), labels,
                                    &sc[SS_ROWTAB], s);
    launch_label_cores<T>(a, core, bcluster, parent, compmin, nullptr, cmin, cmax, labels,
                          &dstats[1], s, &sc[SS_ROWTAB]);
    launch_border<T, D>(a, core, bcluster, rootslot, cmin, cmax, labels, &dstats[2], nullptr,
                        noise_stripes, s);
}

// Ends a stream capture if the body throws halfway.
struct CaptureGuard {
    hipStream_t s;
    bool open = true;
    ~CaptureGuard() {
        if (open) {
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        }
    }
};

// One batch through the static pipeline: replayed from the context's graph
// when the shape (dtype, D, rows, job offsets, buffers, options, workspace)
// matches the captured one, captured first otherwise.  Returns false when a
// flag hands the batch to the plain pipeline.
template <typename T, int D>
static bool run_static(Context& ctx, int dtype, const T* pts, int64_t n, const int64_t* offsets,
                       int32_t njobs, const float* eps, const uint64_t* min_pts, int64_t* labels,
                       hipStream_t s) {
    if (range_bound<D>(n) > 0x7fffffffll) return false;  // 32-bit range offsets
    ctx.main.reserve(static_main_bytes<T, D>(n, njobs));
    ctx.aux.reserve(static_aux_bytes<T, D>(n));
    Arena& A = ctx.main;
    A.reset();
    JobParams<T>* jobs_dev = A.take_n<JobParams<T>>(static_cast<size_t>(njobs));
    int32_t* bfirst_dev = A.take_n<int32_t>(static_cast<size_t>(njobs) + 1);
    uint32_t* sc = A.take_n<uint32_t>(SS_WORDS);
    unsigned long long* dstats = A.take_n<unsigned long long>(16 + kNoiseStripes);

    // host-made part of the job table: row ranges, eps2, minPts, reciprocals
    HostArray<JobParams<T>, 1> hj(static_cast<size_t>(njobs));
    HostArray<int32_t, 2> bfirst(static_cast<size_t>(njobs) + 1);
    const int rpb = minmax_rows_per_block();
    int32_t nblocks = 0;
    for (int32_t j = 0; j < njobs; ++j) {
        JobParams<T>& jp = hj[j];
        jp = JobParams<T>{};
        jp.row_begin = offsets[j];
        jp.row_end = offsets[j + 1];
        jp.min_pts = static_cast<uint32_t>(std::min<uint64_t>(min_pts[j], 0xffffffffull));
        jp.eps2 = eps2_as<T>(eps[j]);
        jp.gap2_max = static_cast<double>(jp.eps2) * (1.0 + 1e-5) + 1e-30;
        jp.rcp_f = flushf(1.0f / flushf(eps[j]));
        jp.rcp_d = flushd(1.0 / flushd(static_cast<double>(eps[j])));
        jp.cp.d = D;
        bfirst[j] = nblocks;
        nblocks += static_cast<int32_t>((offsets[j + 1] - offsets[j] + rpb - 1) / rpb);
    }
    bfirst[njobs] = nblocks;
    HPD_HIP(hipMemcpyAsync(jobs_dev, hj.data(), sizeof(JobParams<T>) * njobs, hipMemcpyHostToDevice,
                           s));
    HPD_HIP(hipMemcpyAsync(bfirst_dev, bfirst.data(), sizeof(int32_t) * (njobs + 1),
                           hipMemcpyHostToDevice, s));

    Context::GraphEntry& ge = ctx.graph_entry;
    const bool same = ge.exec && ge.dtype == dtype && ge.d == D && ge.n == n && ge.data == pts &&
                      ge.labels == labels && ge.arena_main == A.base() &&
                      ge.arena_aux == ctx.aux.base() && ge.border_rule == ctx.border_rule &&
                      ge.sparse_max == ctx.sparse_max && ge.sparse_core == ctx.sparse_core &&
                      ge.box_split == ctx.box_split && ge.count_tests == ctx.count_tests &&
                      ge.offsets.size() == static_cast<size_t>(njobs) + 1 &&
                      std::equal(ge.offsets.begin(), ge.offsets.end(), offsets);
    if (!same) {
        if (ge.exec) {
            (void)hipGraphExecDestroy(ge.exec);
            ge.exec = nullptr;
        }
        // captured on the context's stream, launched on s
        HPD_HIP(hipStreamSynchronize(ctx.stream));
        HPD_HIP(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
        CaptureGuard guard{ctx.stream};
        static_body<T, D>(ctx, pts, n, jobs_dev, bfirst_dev, njobs, nblocks, labels, sc, dstats,
                          ctx.stream);
        hipGraph_t g = nullptr;
        guard.open = false;
        HPD_HIP(hipStreamEndCapture(ctx.stream, &g));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HPD_HIP(e);
        ge.exec = exec;
        ge.dtype = dtype;
        ge.d = D;
        ge.n = n;
        ge.data = pts;
        ge.labels = labels;
        ge.arena_main = A.base();
        ge.arena_aux = ctx.aux.base();
        ge.border_rule = ctx.border_rule;
        ge.sparse_max = ctx.sparse_max;
        ge.box_split = ctx.box_split;
        ge.count_tests = ctx.count_tests;
        ge.sparse_core = ctx.sparse_core;
        ge.offsets.assign(offsets, offsets + njobs + 1);
        ++ctx.graph_captures;
    }
    HPD_HIP(hipGraphLaunch(ge.exec, s));
    ++ctx.graph_launches;
    // one read-back after the run: counts, flags and statistics
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx.host_scratch);
    unsigned long long* hs = reinterpret_cast<unsigned long long*>(ctx.host_scratch) + 16;
    HPD_HIP(hipMemcpyAsync(hc, sc, SS_WORDS * 4, hipMemcpyDeviceToHost, s));
    HPD_HIP(hipMemcpyAsync(hs, dstats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    const bool ok = hc[SS_NONFINITE] == 0 && hc[SS_GRID] == 0 && hc[SS_LOOSE] == 0 &&
                    hs[5] <= 2 * static_cast<unsigned long long>(n) &&
                    hs[4] <= static_cast<unsigned long long>(range_bound<D>(n));
    if (!ok) {
        ++ctx.graph_fallbacks;
        return false;
    }
    ctx.stats.sum_candidates = hs[0];
    ctx.stats.core_points = hs[1];
    ctx.stats.noise_points = hs[2];
    ctx.stats.ambiguous_borders = hs[3];
    ctx.stats.open_box_pairs = hs[5];
    ctx.stats.core_tests = hs[7];
    ctx.stats.total_cells = hs[6];
    ctx.stats.key_bits = static_cast<int32_t>(hc[SS_BITS]) - box_sub_bits(D, split_of(ctx, D));
    ctx.stats.occupied_cells = hc[SS_M];
    ctx.stats.neighbor_rows = hc[SS_RANGES];
    ctx.stats.boxes = hc[SS_NBOX];
    ctx.stats.box_union = 1;
    return true;
}

// A batch of jobs: rows [offsets[j], offsets[j+1]) of `pts` form job j with
// its own eps[j] and min_pts[j].  A plain call is a batch of one.
template <typename T, int D>
static void run_batch(Context& ctx, const T* pts, int64_t n, const int64_t* offsets,
                      int32_t njobs, const float* eps, const uint64_t* min_pts, int64_t* labels,
                      hipStream_t s) {
    if constexpr (std::is_floating_point<T>::value) {
        // the static pipeline, captured once per shape and replayed; it hands
        // the run back when one of its flags says so
        if (ctx.graph && ctx.border_rule == HPD_BORDER_MIN_CLUSTER && ctx.main_use == &ctx.main &&
            !ctx.timing &&
            run_static<T, D>(ctx, static_cast<int>(sizeof(T)), pts, n, offsets, njobs, eps, min_pts,
                             labels, s))
            return;
    }
    PhaseTimer tm(ctx, s);
    tm.mark(M_START);
    Arena& A = *ctx.main_use;
    A.reserve(workspace_bytes<T, D>(n, njobs, ctx.border_rule == HPD_BORDER_ONE_THREAD));

    // job table, first pass: row ranges and min/max blocks.  The host
    // vectors outlive every copy from them: the run ends with a stream sync.
    HostArray<JobParams<T>, 1> hj(static_cast<size_t>(njobs));
    HostArray<int32_t, 2> bfirst(static_cast<size_t>(njobs) + 1);
    const int rpb = minmax_rows_per_block();
    int32_t nblocks = 0;
    for (int32_t j = 0; j < njobs; ++j) {
        hj[j] = JobParams<T>{};
        hj[j].row_begin = offsets[j];
        hj[j].row_end = offsets[j + 1];
        bfirst[j] = nblocks;
        nblocks += static_cast<int32_t>((offsets[j + 1] - offsets[j] + rpb - 1) / rpb);
    }
    bfirst[njobs] = nblocks;
    JobParams<T>* jobs_dev = A.take_n<JobParams<T>>(static_cast<size_t>(njobs));
    int32_t* bfirst_dev = A.take_n<int32_t>(static_cast<size_t>(njobs) + 1);
    T* part = A.take_n<T>(static_cast<size_t>(nblocks) * 2 * D + 1);
    T* mm = A.take_n<T>(static_cast<size_t>(njobs) * 2 * D + 16);
    int* nonfinite = reinterpret_cast<int*>(A.take_n<int>(4));
    HPD_HIP(hipMemcpyAsync(jobs_dev, hj.data(), sizeof(JobParams<T>) * njobs, hipMemcpyHostToDevice, s));
    HPD_HIP(hipMemcpyAsync(bfirst_dev, bfirst.data(), sizeof(int32_t) * (njobs + 1),
                           hipMemcpyHostToDevice, s));
    HPD_HIP(hipMemsetAsync(nonfinite, 0, sizeof(int), s));

    // min / max per job (spatial_index.h:99-122)
    launch_minmax<T, D>(pts, jobs_dev, njobs, bfirst_dev, nblocks, part, mm, nonfinite, s);
    HostArray<T, 2 * D> hmm(static_cast<size_t>(njobs) * 2 * D);
    int hbad = 0;
    HPD_HIP(hipMemcpyAsync(hmm.data(), mm, sizeof(T) * njobs * 2 * D, hipMemcpyDeviceToHost, s));
    HPD_HIP(hipMemcpyAsync(ctx.host_scratch, nonfinite, sizeof(int), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    std::memcpy(&hbad, ctx.host_scratch, sizeof(int));
    if (hbad) throw Error(HPD_EINVAL, "input contains NaN or infinite coordinates");

    // per-job grid (spatial_index.h:124-142) and key ranges
    uint64_t key_total = 0;
    for (int32_t j = 0; j < njobs; ++j) {
        JobParams<T>& jp = hj[j];
        jp.key_base = key_total;
        jp.min_pts = static_cast<uint32_t>(std::min<uint64_t>(min_pts[j], 0xffffffffull));
        jp.eps2 = eps2_as<T>(eps[j]);
        jp.gap2_max = static_cast<double>(jp.eps2) * (1.0 + 1e-5) + 1e-30;
        jp.cp.d = D;
        if (jp.row_end == jp.row_begin) continue;
        float rcp_f;
        double rcp_d;
        const T* mn = &hmm[static_cast<size_t>(j) * 2 * D];
        const Geometry g = make_geometry<T>(mn, mn + D, D, eps[j], &rcp_f, &rcp_d);
        if (g.overflow)
            throw Error(HPD_EUNSUPPORTED,
                        "the eps-grid has more than 2^64 cells (the reference's cell ids wrap)");
        jp.rcp_f = rcp_f;
        jp.rcp_d = rcp_d;
        for (int k = 0; k < D; ++k) {
            jp.minv[k] = mn[k];
            jp.cp.swapped[k] = g.swapped[k];
            jp.cp.stride[k] = g.stride[k];
            jp.cp.ncell[k] = g.ncell[k];
        }
        uint64_t t;
        if (__builtin_add_overflow(key_total, g.total, &t))
            throw Error(HPD_EUNSUPPORTED, "the jobs' grids have more than 2^64 cells together");
        key_total = t;
        if (njobs == 1) {
            ctx.stats.total_cells = g.total;
        }
    }
    if (njobs != 1) ctx.stats.total_cells = key_total;
    const int key_bits = bits_for(key_total > 0 ? key_total - 1 : 0);
    ctx.stats.key_bits = key_bits;
    HPD_HIP(hipMemcpyAsync(jobs_dev, hj.data(), sizeof(JobParams<T>) * njobs, hipMemcpyHostToDevice, s));

    // floating-point data runs on sub-boxes (the sort key then carries the
    // sub-box digit below the cell id); integer data, or a grid so large
    // that the digit does not fit beside the cell id, runs on cells
    const int split_bits = box_sub_bits(D, split_of(ctx, D));
    const bool boxes = std::is_floating_point<T>::value && key_bits + split_bits <= 64;
    const int sort_bits = key_bits + (boxes ? split_bits : 0);
    if (sort_bits <= 32)
        run_keyed<T, D, uint32_t>(ctx, pts, n, jobs_dev, njobs, key_bits, key_total, boxes, labels,
                                  s, tm);
    else
        run_keyed<T, D, uint64_t>(ctx, pts, n, jobs_dev, njobs, key_bits, key_total, boxes, labels,
                                  s, tm);
    if (ctx.timing) {
        ctx.stats.ms_grid = tm.ms(M_START, M_GRID);
        ctx.stats.ms_sort = tm.ms(M_GRID, M_SORT);
        ctx.stats.ms_index = tm.ms(M_SORT, M_INDEX);
        ctx.stats.ms_core = tm.ms(M_INDEX, M_CORE);
        ctx.stats.ms_union = tm.ms(M_CORE, M_UNION);
        ctx.stats.ms_border = tm.ms(M_UNION, M_LABELS);
        ctx.stats.ms_total = tm.ms(M_START, M_LABELS);
        ctx.stats.ms_hook_kernel = tm.ms(M_HOOK0, M_HOOK1);
        ctx.stats.ms_union_kernel = tm.ms(M_HOOK1, M_JOIN1);
        ctx.stats.ms_border_kernel = tm.ms(M_BORDER0, M_LABELS);
        ctx.stats.ms_core_kernel = tm.ms(M_CORE0, M_CORE);
    }
}

template <typename T>
static void run_typed(Context& ctx, const T* pts, int64_t n, int d, const int64_t* offsets,
                      int32_t njobs, const float* eps, const uint64_t* min_pts, int64_t* labels,
                      hipStream_t s) {
    switch (d) {
        case 1: return run_batch<T, 1>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 2: return run_batch<T, 2>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 3: return run_batch<T, 3>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 4: return run_batch<T, 4>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 5: return run_batch<T, 5>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 6: return run_batch<T, 6>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 7: return run_batch<T, 7>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        case 8: return run_batch<T, 8>(ctx, pts, n, offsets, njobs, eps, min_pts, labels, s);
        default:
            throw Error(HPD_EUNSUPPORTED, "dimensions above " + std::to_string(kMaxDims) +
                                              " are not supported by this build");
    }
}

static size_t dtype_size(int dt) {
    switch (dt) {
        case HPD_DTYPE_F32: return 4;
        case HPD_DTYPE_F64: return 8;
        case HPD_DTYPE_I8: case HPD_DTYPE_U8: return 1;
        case HPD_DTYPE_I16: case HPD_DTYPE_U16: return 2;
        case HPD_DTYPE_I32: case HPD_DTYPE_U32: return 4;
        case HPD_DTYPE_I64: case HPD_DTYPE_U64: return 8;
        default: return 0;
    }
}

static void run_any(Context& ctx, int dt, const void* pts, int64_t n, int d,
                    const int64_t* offsets, int32_t njobs, const float* eps,
                    const uint64_t* min_pts, int64_t* labels, hipStream_t s) {
#define HPD_RUN_CASE(CODE, T)                                                                \
    case CODE:                                                                               \
        return run_typed(ctx, static_cast<const T*>(pts), n, d, offsets, njobs, eps, min_pts, \
                         labels, s);
    switch (dt) {
        HPD_RUN_CASE(HPD_DTYPE_F32, float)
        HPD_RUN_CASE(HPD_DTYPE_F64, double)
        HPD_RUN_CASE(HPD_DTYPE_I8, int8_t)
        HPD_RUN_CASE(HPD_DTYPE_I16, int16_t)
        HPD_RUN_CASE(HPD_DTYPE_I32, int32_t)
        HPD_RUN_CASE(HPD_DTYPE_I64, int64_t)
        HPD_RUN_CASE(HPD_DTYPE_U8, uint8_t)
        HPD_RUN_CASE(HPD_DTYPE_U16, uint16_t)
        HPD_RUN_CASE(HPD_DTYPE_U32, uint32_t)
        HPD_RUN_CASE(HPD_DTYPE_U64, uint64_t)
        default: throw Error(HPD_EINVAL, "unknown dtype code " + std::to_string(dt));
    }
#undef HPD_RUN_CASE
}

// workspace_bytes of a plain call (one job) of any dtype and dimension
template <typename T>
static size_t workspace_typed(int64_t n, int d, bool one_thread) {
    switch (d) {
        case 1: return workspace_bytes<T, 1>(n, 1, one_thread);
        case 2: return workspace_bytes<T, 2>(n, 1, one_thread);
        case 3: return workspace_bytes<T, 3>(n, 1, one_thread);
        case 4: return workspace_bytes<T, 4>(n, 1, one_thread);
        case 5: return workspace_bytes<T, 5>(n, 1, one_thread);
        case 6: return workspace_bytes<T, 6>(n, 1, one_thread);
        case 7: return workspace_bytes<T, 7>(n, 1, one_thread);
        case 8: return workspace_bytes<T, 8>(n, 1, one_thread);
        default:
            throw Error(HPD_EUNSUPPORTED, "dimensions above " + std::to_string(kMaxDims) +
                                              " are not supported by this build");
    }
}

static size_t workspace_any(int dt, int64_t n, int d, bool one_thread) {
    switch (dt) {
        case HPD_DTYPE_F32: return workspace_typed<float>(n, d, one_thread);
        case HPD_DTYPE_F64: return workspace_typed<double>(n, d, one_thread);
        case HPD_DTYPE_I8: return workspace_typed<int8_t>(n, d, one_thread);
        case HPD_DTYPE_I16: return workspace_typed<int16_t>(n, d, one_thread);
        case HPD_DTYPE_I32: return workspace_typed<int32_t>(n, d, one_thread);
        case HPD_DTYPE_I64: return workspace_typed<int64_t>(n, d, one_thread);
        case HPD_DTYPE_U8: return workspace_typed<uint8_t>(n, d, one_thread);
        case HPD_DTYPE_U16: return workspace_typed<uint16_t>(n, d, one_thread);
        case HPD_DTYPE_U32: return workspace_typed<uint32_t>(n, d, one_thread);
        case HPD_DTYPE_U64: return workspace_typed<uint64_t>(n, d, one_thread);
        default: throw Error(HPD_EINVAL, "unknown dtype code " + std::to_string(dt));
    }
}

// --------------------------------------------------------------------------
// static scratch of hpd_cluster_device_noalloc
// --------------------------------------------------------------------------
// Every workspace buffer of a run of hpd_cluster_device_noalloc lives in
// this array (one copy per device, held until the process ends).  Host code
// never passes its name to a kernel: the name is a host-side stand-in; the
// device address comes from hipGetSymbolAddress, looked up once per device
// when a context is created on it (with that device current).
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static void init_scratch_for_current_device(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw Error(HPD_EUNSUPPORTED, "device index above " + std::to_string(kMaxDevices - 1));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    HPD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(HPD_EDEVICE, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[device] = ptr;
}

static void* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Points a context's runs at fixed arenas and turns its timing off (event
// creation) for the duration of one call.
class FixedWorkspace {
public:
    FixedWorkspace(Context& c, void* base, size_t main_bytes, size_t total_bytes)
        : c_(c), timing_(c.timing) {
        main_.attach(base, main_bytes);
        aux_.attach(static_cast<char*>(base) + main_bytes, total_bytes - main_bytes);
        c.main_use = &main_;
        c.aux_use = &aux_;
        c.timing = false;
    }
    ~FixedWorkspace() {
        c_.main_use = &c_.main;
        c_.aux_use = &c_.aux;
        c_.timing = timing_;
    }
    FixedWorkspace(const FixedWorkspace&) = delete;
    FixedWorkspace& operator=(const FixedWorkspace&) = delete;

private:
    Context& c_;
    bool timing_;
    Arena main_, aux_;
};

static void validate(const Context* ctx, int dt, const void* data, int64_t n, int32_t d, float eps,
                     const void* labels) {
    if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
    // hpdbscan.h:295-297 (NaN is rejected as well)
    if (!(eps > 0.0f)) throw Error(HPD_EINVAL, "epsilon needs to be positive");
    if (dtype_size(dt) == 0) throw Error(HPD_EINVAL, "unknown dtype code " + std::to_string(dt));
    if (n < 0) throw Error(HPD_EINVAL, "the number of rows must not be negative");
    if (n > 0x7fffffffll) throw Error(HPD_EINVAL, "at most 2^31-1 rows (the reference takes an int)");
    if (d < 1) throw Error(HPD_EINVAL, "the number of dimensions must be positive");
    if (n > 0 && (!data || !labels)) throw Error(HPD_EINVAL, "data and labels must not be NULL");
}

}  // namespace hpd

using hpd::Context;
using hpd::Error;

struct hpd_context {
    Context impl;
};

template <typename F>
static int guarded(F&& f) {
    try {
        f();
        hpd::g_last_error.clear();
        return HPD_OK;
    } catch (const Error& e) {
        hpd::g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        hpd::g_last_error = "host allocation failed";
        return HPD_ENOMEM;
    } catch (const std::exception& e) {
        hpd::g_last_error = e.what();
        return HPD_EDEVICE;
    }
}

extern "C" {

const char* hpd_last_error(void) { return hpd::g_last_error.c_str(); }

const char* hpd_version(void) { return "hpdbscan-amd 0.1.0 (gfx950)"; }

int hpd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int hpd_context_create(int device, hpd_context** out) {
    return guarded([&] {
        if (!out) throw Error(HPD_EINVAL, "out is NULL");
        *out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            throw Error(HPD_EDEVICE, "no HIP device is available");
        if (device < 0 || device >= count)
            throw Error(HPD_EINVAL, "device index " + std::to_string(device) + " out of range");
        auto* c = new hpd_context();
        try {
            c->impl.device = device;
            hpd::DeviceGuard g(device);
            hpd::init_scratch_for_current_device(device);
            HPD_HIP(hipStreamCreateWithFlags(&c->impl.stream, hipStreamNonBlocking));
            void* hs = nullptr;
            HPD_HIP(hipHostMalloc(&hs, 64 * sizeof(uint64_t), hipHostMallocDefault));
            c->impl.host_scratch = static_cast<uint64_t*>(hs);
        } catch (...) {
            if (c->impl.stream) (void)hipStreamDestroy(c->impl.stream);
            delete c;
            throw;
        }
        *out = c;
    });
}

int hpd_context_destroy(hpd_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        {
            hpd::DeviceGuard g(ctx->impl.device);
            if (ctx->impl.stream) (void)hipStreamSynchronize(ctx->impl.stream);
            if (ctx->impl.graph_entry.exec) (void)hipGraphExecDestroy(ctx->impl.graph_entry.exec);
            for (auto e : ctx->impl.events) (void)hipEventDestroy(e);
            if (ctx->impl.host_scratch) (void)hipHostFree(ctx->impl.host_scratch);
            if (ctx->impl.stream) (void)hipStreamDestroy(ctx->impl.stream);
        }
        delete ctx;
    });
}

int hpd_context_set_border_rule(hpd_context* ctx, int rule) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        if (rule != HPD_BORDER_MIN_CLUSTER && rule != HPD_BORDER_ONE_THREAD)
            throw Error(HPD_EINVAL, "unknown border rule");
        ctx->impl.border_rule = rule;
    });
}

int hpd_context_set_timing(hpd_context* ctx, int enabled) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        ctx->impl.timing = enabled != 0;
        ctx->impl.count_tests = enabled >= 2;
    });
}

int hpd_context_set_graph(hpd_context* ctx, int enabled) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        Context& c = ctx->impl;
        c.graph = enabled != 0;
        if (!c.graph && c.graph_entry.exec) {
            hpd::DeviceGuard g(c.device);
            HPD_HIP(hipStreamSynchronize(c.stream));
            (void)hipGraphExecDestroy(c.graph_entry.exec);
            c.graph_entry = Context::GraphEntry{};
        }
    });
}

int hpd_context_get_graph_stats(const hpd_context* ctx, uint64_t* launches, uint64_t* captures,
                                uint64_t* fallbacks) {
    return guarded([&] {
        if (!ctx || !launches || !captures || !fallbacks) throw Error(HPD_EINVAL, "NULL argument");
        *launches = ctx->impl.graph_launches;
        *captures = ctx->impl.graph_captures;
        *fallbacks = ctx->impl.graph_fallbacks;
    });
}

int hpd_context_set_open_pair_limit(hpd_context* ctx, uint64_t limit) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        ctx->impl.open_pair_limit = limit;
    });
}

int hpd_context_set_work_split(hpd_context* ctx, int32_t sparse_max, int sparse_core) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        if (sparse_max < 0 || sparse_max > 64)
            throw Error(HPD_EINVAL, "sparse_max must be in [0, 64]");
        ctx->impl.sparse_max = sparse_max;
        ctx->impl.sparse_core = sparse_core != 0;
    });
}

int hpd_context_set_box_split(hpd_context* ctx, int32_t parts) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        if (parts < 0 || parts > 8) throw Error(HPD_EINVAL, "box split must be in [0, 8]");
        ctx->impl.box_split = parts;
    });
}

int hpd_context_get_box_split(const hpd_context* ctx, int32_t* parts) {
    return guarded([&] {
        if (!ctx || !parts) throw Error(HPD_EINVAL, "NULL argument");
        *parts = ctx->impl.box_split;
    });
}

int hpd_context_get_work_split(const hpd_context* ctx, int32_t* sparse_max, int* sparse_core) {
    return guarded([&] {
        if (!ctx || !sparse_max || !sparse_core) throw Error(HPD_EINVAL, "NULL argument");
        *sparse_max = ctx->impl.sparse_max;
        *sparse_core = ctx->impl.sparse_core ? 1 : 0;
    });
}

int hpd_context_get_stats(const hpd_context* ctx, hpd_run_stats* out) {
    return guarded([&] {
        if (!ctx || !out) throw Error(HPD_EINVAL, "NULL argument");
        *out = ctx->impl.stats;
    });
}

int hpd_cluster_device(hpd_context* ctx, int dtype, const void* data, int64_t n, int32_t d,
                       float eps, uint64_t min_pts, int64_t* labels_out, void* stream) {
    return guarded([&] {
        hpd::validate(ctx ? &ctx->impl : nullptr, dtype, data, n, d, eps, labels_out);
        Context& c = ctx->impl;
        c.stats = hpd_run_stats{};
        c.stats.n = n;
        c.stats.d = d;
        if (n == 0) return;
        hpd::DeviceGuard g(c.device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
        const int64_t offsets[2] = {0, n};
        hpd::run_any(c, dtype, data, n, d, offsets, 1, &eps, &min_pts, labels_out, s);
    });
}

int hpd_cluster_device_noalloc(hpd_context* ctx, int dtype, const void* data, int64_t n,
                               int32_t d, float eps, uint64_t min_pts, int64_t* labels_out,
                               void* stream) {
    return guarded([&] {
        hpd::validate(ctx ? &ctx->impl : nullptr, dtype, data, n, d, eps, labels_out);
        Context& c = ctx->impl;
        c.stats = hpd_run_stats{};
        c.stats.n = n;
        c.stats.d = d;
        if (n == 0) return;
        // the workspace of the main arena, checked before anything runs; the
        // neighbour tables and work lists take the rest and are checked when
        // their sizes are known (before any label is written)
        const size_t need = hpd::workspace_any(dtype, n, d, c.border_rule == HPD_BORDER_ONE_THREAD);
        if (need > hpd::kScratchBytes)
            throw Error(HPD_EUNSUPPORTED,
                        "hpd_cluster_device_noalloc: " + std::to_string(n) + " rows of " +
                            std::to_string(d) + " dimensions need " + std::to_string(need) +
                            " bytes of workspace; its static scratch holds " +
                            std::to_string(hpd::kScratchBytes) +
                            " bytes (use hpd_cluster_device for larger inputs)");
        hpd::DeviceGuard g(c.device);
        void* base = hpd::scratch_of(c.device);
        if (!base) throw Error(HPD_EDEVICE, "no scratch address for this device");
        hpd::FixedWorkspace ws(c, base, hpd::align_up(need, 256), hpd::kScratchBytes);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
        const int64_t offsets[2] = {0, n};
        hpd::run_any(c, dtype, data, n, d, offsets, 1, &eps, &min_pts, labels_out, s);
    });
}

int hpd_cluster(hpd_context* ctx, int dtype, const void* data, int64_t n, int32_t d, float eps,
                uint64_t min_pts, int64_t* labels_out) {
    return guarded([&] {
        hpd::validate(ctx ? &ctx->impl : nullptr, dtype, data, n, d, eps, labels_out);
        Context& c = ctx->impl;
        c.stats = hpd_run_stats{};
        c.stats.n = n;
        c.stats.d = d;
        if (n == 0) return;
        hpd::DeviceGuard g(c.device);
        const size_t row_bytes = hpd::dtype_size(dtype) * static_cast<size_t>(d);
        const size_t in_bytes = row_bytes * static_cast<size_t>(n);
        const size_t out_bytes = sizeof(int64_t) * static_cast<size_t>(n);
        // input rows and output labels live in their own grow-only buffer
        c.io.reserve(hpd::align_up(in_bytes, 256) + out_bytes + 512);
        void* din = c.io.take(in_bytes);
        int64_t* dout = c.io.take_n<int64_t>(static_cast<size_t>(n));
        HPD_HIP(hipMemcpyAsync(din, data, in_bytes, hipMemcpyHostToDevice, c.stream));
        const int64_t offsets[2] = {0, n};
        hpd::run_any(c, dtype, din, n, d, offsets, 1, &eps, &min_pts, dout, c.stream);
        HPD_HIP(hipMemcpyAsync(labels_out, dout, out_bytes, hipMemcpyDeviceToHost, c.stream));
        HPD_HIP(hipStreamSynchronize(c.stream));
    });
}

static void validate_batch(const hpd_context* ctx, int dtype, const void* data,
                           const int64_t* offsets, int32_t njobs, int32_t d, const float* eps,
                           const uint64_t* min_pts, const void* labels) {
    if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
    if (njobs < 0) throw Error(HPD_EINVAL, "the number of jobs must not be negative");
    if (njobs == 0) return;
    if (!offsets || !eps || !min_pts) throw Error(HPD_EINVAL, "offsets, eps and min_pts must not be NULL");
    if (offsets[0] != 0) throw Error(HPD_EINVAL, "offsets[0] must be 0");
    for (int32_t j = 0; j < njobs; ++j) {
        if (offsets[j + 1] < offsets[j]) throw Error(HPD_EINVAL, "offsets must not decrease");
        if (!(eps[j] > 0.0f))
            throw Error(HPD_EINVAL, "epsilon needs to be positive (job " + std::to_string(j) + ")");
    }
    hpd::validate(&ctx->impl, dtype, data, offsets[njobs], d, 1.0f, labels);
}

int hpd_cluster_batch_device(hpd_context* ctx, int dtype, const void* data,
                             const int64_t* offsets, int32_t njobs, int32_t d, const float* eps,
                             const uint64_t* min_pts, int64_t* labels_out, void* stream) {
    return guarded([&] {
        validate_batch(ctx, dtype, data, offsets, njobs, d, eps, min_pts, labels_out);
        Context& c = ctx->impl;
        c.stats = hpd_run_stats{};
        const int64_t n = njobs ? offsets[njobs] : 0;
        c.stats.n = n;
        c.stats.d = d;
        if (n == 0) return;
        hpd::DeviceGuard g(c.device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
        hpd::run_any(c, dtype, data, n, d, offsets, njobs, eps, min_pts, labels_out, s);
    });
}

int hpd_cluster_batch(hpd_context* ctx, int dtype, const void* data, const int64_t* offsets,
                      int32_t njobs, int32_t d, const float* eps, const uint64_t* min_pts,
                      int64_t* labels_out) {
    return guarded([&] {
        validate_batch(ctx, dtype, data, offsets, njobs, d, eps, min_pts, labels_out);
        Context& c = ctx->impl;
        c.stats = hpd_run_stats{};
        const int64_t n = njobs ? offsets[njobs] : 0;
        c.stats.n = n;
        c.stats.d = d;
        if (n == 0) return;
        hpd::DeviceGuard g(c.device);
        const size_t in_bytes = hpd::dtype_size(dtype) * static_cast<size_t>(d) * static_cast<size_t>(n);
        const size_t out_bytes = sizeof(int64_t) * static_cast<size_t>(n);
        c.io.reserve(hpd::align_up(in_bytes, 256) + out_bytes + 512);
        void* din = c.io.take(in_bytes);
        int64_t* dout = c.io.take_n<int64_t>(static_cast<size_t>(n));
        HPD_HIP(hipMemcpyAsync(din, data, in_bytes, hipMemcpyHostToDevice, c.stream));
        hpd::run_any(c, dtype, din, n, d, offsets, njobs, eps, min_pts, dout, c.stream);
        HPD_HIP(hipMemcpyAsync(labels_out, dout, out_bytes, hipMemcpyDeviceToHost, c.stream));
        HPD_HIP(hipStreamSynchronize(c.stream));
    });
}

int hpd_check_neighbor_rows(uint64_t neighbor_rows) {
    return guarded([&] { hpd::check_neighbor_rows(neighbor_rows); });
}

int hpd_device_alloc(hpd_context* ctx, size_t bytes, void** out) {
    return guarded([&] {
        if (!ctx || !out) throw Error(HPD_EINVAL, "NULL argument");
        *out = nullptr;
        hpd::DeviceGuard g(ctx->impl.device);
        HPD_HIP(hipMalloc(out, bytes > 0 ? bytes : 1));
    });
}

int hpd_device_free(hpd_context* ctx, void* ptr) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        if (!ptr) return;
        hpd::DeviceGuard g(ctx->impl.device);
        HPD_HIP(hipStreamSynchronize(ctx->impl.stream));
        HPD_HIP(hipFree(ptr));
    });
}

int hpd_memcpy_to_device(hpd_context* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!ctx || (bytes && (!dst || !src))) throw Error(HPD_EINVAL, "NULL argument");
        hpd::DeviceGuard g(ctx->impl.device);
        HPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->impl.stream));
        HPD_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

int hpd_memcpy_to_host(hpd_context* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!ctx || (bytes && (!dst || !src))) throw Error(HPD_EINVAL, "NULL argument");
        hpd::DeviceGuard g(ctx->impl.device);
        HPD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->impl.stream));
        HPD_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

int hpd_context_stream(hpd_context* ctx, void** stream_out) {
    return guarded([&] {
        if (!ctx || !stream_out) throw Error(HPD_EINVAL, "NULL argument");
        *stream_out = static_cast<void*>(ctx->impl.stream);
    });
}

int hpd_context_synchronize(hpd_context* ctx) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        hpd::DeviceGuard g(ctx->impl.device);
        HPD_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

}  // extern "C"
