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
#include <cstring>
#include <limits>
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
    if (base_) (void)hipFree(base_);
}

void Arena::reserve(size_t bytes) {
    used_ = 0;
    if (bytes <= cap_) return;
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
    std::vector<size_t> sw(static_cast<size_t>(d));
    std::iota(sw.begin(), sw.end(), 0);
    std::sort(sw.begin(), sw.end(), [&](size_t a, size_t b) { return g.cells[a] < g.cells[b]; });
    uint64_t acc = 1;
    for (int k = 0; k < d; ++k) {
        g.swapped[k] = static_cast<int>(sw[static_cast<size_t>(k)]);
        g.stride[k] = acc;
        g.ncell[k] = g.cells[g.swapped[k]];
        acc *= g.ncell[k];
    }
    const uint64_t maxkey = total - 1;
    g.key_bits = maxkey == 0 ? 0 : 64 - __builtin_clzll(maxkey);
    return g;
}

// --------------------------------------------------------------------------
// context
// --------------------------------------------------------------------------
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena main;
    Arena aux;
    Arena io;  // input rows and labels of the host-memory entry point
    int border_rule = HPD_BORDER_MIN_CLUSTER;
    bool timing = false;
    hpd_run_stats stats{};
    uint64_t* host_scratch = nullptr;  // pinned, 64 words
    std::vector<hipEvent_t> events;
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

class PhaseTimer {
public:
    PhaseTimer(Context& c, hipStream_t s) : ctx_(c), s_(s) {}
    void mark() {
        if (!ctx_.timing) return;
        if (idx_ >= ctx_.events.size()) {
            hipEvent_t e;
            HPD_HIP(hipEventCreate(&e));
            ctx_.events.push_back(e);
        }
        HPD_HIP(hipEventRecord(ctx_.events[idx_], s_));
        ++idx_;
    }
    // milliseconds between mark i and mark j (after a stream sync)
    double ms(size_t i, size_t j) const {
        if (!ctx_.timing || j >= idx_) return 0.0;
        float v = 0.0f;
        HPD_HIP(hipEventElapsedTime(&v, ctx_.events[i], ctx_.events[j]));
        return static_cast<double>(v);
    }
    size_t count() const { return idx_; }

private:
    Context& ctx_;
    hipStream_t s_;
    size_t idx_ = 0;
};

// --------------------------------------------------------------------------
// one clustering run
// --------------------------------------------------------------------------
static inline int bits_for(uint64_t max_value) {
    return max_value == 0 ? 0 : 64 - __builtin_clzll(max_value);
}

// worst-case device workspace of one run with n rows
template <typename T, int D>
static size_t workspace_bytes(int64_t n) {
    const size_t nn = static_cast<size_t>(n);
    const size_t L = align_up(nn, 16);
    const size_t hist_words = radix_hist_words(n);
    const size_t tmp_words = scan_tmp_words(static_cast<int64_t>(std::max(hist_words, 2 * nn + 2)));
    size_t b = 0;
    b += 4 * 8 * nn + 4 * 4 * nn;              // keys / values of both sorts
    b += 4 * hist_words + 4 * tmp_words;       // radix histogram, scan scratch
    b += sizeof(T) * L * D;                    // coordinates in slot order
    b += 4 * (2 * nn + 2);                     // run flags + scan
    b += 8 * nn + 3 * 4 * (nn + 1);            // ucell, ustart, nranges, nchunks
    b += L + 3 * 4 * L + 3 * 4 * nn;           // core, parent, rsnap, rootslot, cmin/cmax/rcenter
    b += 2 * 4 * nn + 3 * 4 * (nn + 1) + 5 * 4 * nn + nn;  // box tables
    b += 2 * D * 8 * nn;                       // box bounding boxes
    b += sizeof(T) * 2 * D * 1024 + 40 * 256 + 64 * 1024;  // min/max partials, alignment
    return b;
}

template <typename T, int D, typename K>
static void run_keyed(Context& ctx, const T* pts, int64_t n, float eps, uint64_t min_pts,
                      int64_t* labels, hipStream_t s, const Geometry& g, float rcp_f,
                      double rcp_d, const T* minv_dev, PhaseTimer& tm) {
    constexpr bool kFloat = std::is_floating_point<T>::value;
    Arena& A = ctx.main;
    const size_t nn = static_cast<size_t>(n);
    const int64_t ld = (n + 15) & ~int64_t(15);  // rows are read in aligned batches of <= 16
    const size_t L = static_cast<size_t>(ld);
    K* keys_a = A.take_n<K>(nn);
    K* keys_b = A.take_n<K>(nn);
    uint32_t* vals_a = A.take_n<uint32_t>(nn);
    uint32_t* vals_b = A.take_n<uint32_t>(nn);
    const size_t hist_words = radix_hist_words(n);
    uint32_t* hist = A.take_n<uint32_t>(hist_words);
    uint32_t* tmp = A.take_n<uint32_t>(scan_tmp_words(static_cast<int64_t>(std::max(hist_words, 2 * nn + 2))));
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
    unsigned long long* dstats = A.take_n<unsigned long long>(16);
    uint32_t* counters = reinterpret_cast<uint32_t*>(dstats + 8);  // m, ranges, items, boxes, box items, loose
    HPD_HIP(hipMemsetAsync(dstats, 0, 16 * sizeof(unsigned long long), s));
    uint32_t* hc = reinterpret_cast<uint32_t*>(ctx.host_scratch);

    CellParams cp{};
    cp.d = D;
    for (int k = 0; k < D; ++k) {
        cp.swapped[k] = g.swapped[k];
        cp.stride[k] = g.stride[k];
        cp.ncell[k] = g.ncell[k];
    }

    // cell ids (spatial_index.h:144-165)
    launch_cell_ids<T, D, K>(pts, n, minv_dev, cp, rcp_f, rcp_d, keys_a, vals_a, s);
    tm.mark();  // 1: grid done

    // stable sort by cell (spatial_index.h:167-216): position -> input row
    const bool in_b = radix_sort_pairs<K>(keys_a, vals_a, keys_b, vals_b, n, g.key_bits, hist, tmp, s);
    K* keys = in_b ? keys_b : keys_a;
    uint32_t* inrow_of_pos = in_b ? vals_b : vals_a;

    // occupied cells (spatial_index.h:167-182)
    launch_cells<K>(keys, n, flags, tmp, ucell, ustart, &counters[0], s);
    HPD_HIP(hipMemcpyAsync(hc, counters, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    const int32_t m = static_cast<int32_t>(hc[0]);

    // slot order: in box mode, rows regrouped by sub-box inside each cell
    const bool boxes = kFloat;
    const uint32_t* pos_of_slot = nullptr;
    uint32_t* inrow_of_slot = inrow_of_pos;
    int32_t *box_of_slot = nullptr, *bstart = nullptr, *bcell = nullptr, *cbox = nullptr;
    uint32_t* bcnt = nullptr;
    int32_t *bmincore = nullptr, *bparent = nullptr, *compmin = nullptr, *bcluster = nullptr;
    uint8_t* btight = nullptr;
    void* bbox = nullptr;
    if constexpr (kFloat) {
        const int sub_bits = box_sub_bits(D);
        const int bits2 = bits_for(static_cast<uint64_t>(m - 1)) + sub_bits;
        uint64_t* k2a = A.take_n<uint64_t>(nn);
        uint64_t* k2b = A.take_n<uint64_t>(nn);
        uint32_t* v2a = A.take_n<uint32_t>(nn);
        uint32_t* v2b = A.take_n<uint32_t>(nn);
        inrow_of_slot = A.take_n<uint32_t>(nn);
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
        const uint32_t* slot_pos;
        const void* sorted2;
        if (bits2 <= 32) {
            uint32_t* ka = reinterpret_cast<uint32_t*>(k2a);
            uint32_t* kb = reinterpret_cast<uint32_t*>(k2b);
            launch_box_keys<T, D, uint32_t>(pts, n, inrow_of_pos, flags, flags + n, minv_dev, rcp_f,
                                            rcp_d, ka, v2a, s);
            const bool ib = radix_sort_pairs<uint32_t>(ka, v2a, kb, v2b, n, bits2, hist, tmp, s);
            slot_pos = ib ? v2b : v2a;
            sorted2 = ib ? static_cast<const void*>(kb) : static_cast<const void*>(ka);
            launch_box_table<uint32_t>(static_cast<const uint32_t*>(sorted2), n, sub_bits, flags, tmp,
                                       bstart, bcell, box_of_slot, &counters[3], s);
        } else {
            launch_box_keys<T, D, uint64_t>(pts, n, inrow_of_pos, flags, flags + n, minv_dev, rcp_f,
                                            rcp_d, k2a, v2a, s);
            const bool ib = radix_sort_pairs<uint64_t>(k2a, v2a, k2b, v2b, n, bits2, hist, tmp, s);
            slot_pos = ib ? v2b : v2a;
            sorted2 = ib ? static_cast<const void*>(k2b) : static_cast<const void*>(k2a);
            launch_box_table<uint64_t>(static_cast<const uint64_t*>(sorted2), n, sub_bits, flags, tmp,
                                       bstart, bcell, box_of_slot, &counters[3], s);
        }
        pos_of_slot = slot_pos;
        launch_gather_boxes<T, D>(pts, n, pos_of_slot, inrow_of_pos, soa, ld, inrow_of_slot, s);
    } else {
        launch_gather_soa<T, D>(pts, n, inrow_of_pos, soa, ld, s);
    }
    tm.mark();  // 2: sort done

    // neighbour ranges (spatial_index.h:682-737) and work items
    launch_neighbor_count<K>(ucell, ustart, m, cp, nranges, &dstats[0], s);
    exclusive_scan_u32(nranges, nranges, m, tmp, &nranges[m], s);
    launch_chunk_counts(ustart, m, nchunks, s);
    exclusive_scan_u32(nchunks, nchunks, m, tmp, &nchunks[m], s);
    HPD_HIP(hipMemcpyAsync(&counters[1], &nranges[m], sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    HPD_HIP(hipMemcpyAsync(&counters[2], &nchunks[m], sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    HPD_HIP(hipMemcpyAsync(hc, counters, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    const uint32_t total_ranges = hc[1];
    const int32_t nitems = static_cast<int32_t>(hc[2]);
    const int32_t nbox = boxes ? static_cast<int32_t>(hc[3]) : 0;
    int32_t nbitems = 0;
    if (boxes) {
        launch_box_items(bstart, nbox, bcnt, tmp, &counters[4], s);
        HPD_HIP(hipMemcpyAsync(hc + 4, &counters[4], sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HPD_HIP(hipStreamSynchronize(s));
        nbitems = static_cast<int32_t>(hc[4]);
    }

    ctx.aux.reserve(2 * align_up(static_cast<size_t>(total_ranges) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nitems) * sizeof(int2), 256) +
                    align_up(static_cast<size_t>(nbitems) * sizeof(int2), 256) + 1024);
    int2* ranges = ctx.aux.take_n<int2>(total_ranges);
    int2* rcells = ctx.aux.take_n<int2>(total_ranges);
    int2* items = ctx.aux.take_n<int2>(static_cast<size_t>(nitems));
    int2* bitems = boxes ? ctx.aux.take_n<int2>(static_cast<size_t>(nbitems)) : nullptr;
    launch_neighbor_fill<K>(ucell, ustart, m, cp, nranges, ranges, rcells, rcenter, s);
    launch_work_items(ustart, m, nchunks, items, s);

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
    a.eps2 = eps2_as<T>(eps);
    a.min_pts = static_cast<uint32_t>(std::min<uint64_t>(min_pts, 0xffffffffull));
    a.pos_of_slot = pos_of_slot;
    a.inrow = inrow_of_slot;

    BoxArgs<T> ba;
    bool box_union = false;
    if constexpr (kFloat) {
        launch_cell_boxes(bcell, nbox, m, cbox, s);
        launch_box_items_fill(bstart, nbox, bcnt, bitems, s);
        ba.soa = soa;
        ba.ld = ld;
        ba.nbox = nbox;
        ba.bstart = bstart;
        ba.bcell = bcell;
        ba.cbox = cbox;
        ba.bitems = bitems;
        ba.nbitems = nbitems;
        ba.pos_of_slot = pos_of_slot;
        ba.bbox = bbox;
        launch_box_reduce<T, D>(ba, nullptr, nullptr, s);
        launch_box_tight<T, D>(ba, a.eps2, a.min_pts, btight, &counters[5], s);
        a.box_of_slot = box_of_slot;
        a.btight = btight;
    }
    tm.mark();  // 3: index done

    launch_core<T, D>(a, core, &dstats[1], s);
    tm.mark();  // 4: core done

    if constexpr (kFloat) {
        HPD_HIP(hipMemcpyAsync(hc + 5, &counters[5], sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HPD_HIP(hipStreamSynchronize(s));
        box_union = hc[5] == 0;  // every box of two or more rows is tight
    }
    if (box_union) {
        launch_box_reduce<T, D>(ba, core, bmincore, s);
        launch_box_union<T, D>(ba, nranges, rcells, bmincore, core, a.eps2, bparent, compmin,
                               bcluster, s);
    } else {
        if (!compmin) compmin = A.take_n<int32_t>(nn);
        launch_union_points<T, D>(a, core, parent, rsnap, cmin, cmax, compmin, s);
    }
    tm.mark();  // 5: union done
    launch_labels<T, D>(a, core, box_union ? bcluster : nullptr, parent, compmin, rootslot, cmin,
                        cmax, labels, &dstats[2], s);
    tm.mark();  // 6: labels done

    unsigned long long* hs = reinterpret_cast<unsigned long long*>(ctx.host_scratch);
    HPD_HIP(hipMemcpyAsync(hs, dstats, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    ctx.stats.sum_candidates = hs[0];
    ctx.stats.core_points = hs[1];
    ctx.stats.noise_points = hs[2];
    ctx.stats.ambiguous_borders = hs[3];
    ctx.stats.occupied_cells = static_cast<uint64_t>(m);
    ctx.stats.boxes = static_cast<uint64_t>(nbox);
    ctx.stats.box_union = box_union ? 1 : 0;
}

template <typename T, int D>
static void run_dims(Context& ctx, const T* pts, int64_t n, float eps, uint64_t min_pts,
                     int64_t* labels, hipStream_t s) {
    PhaseTimer tm(ctx, s);
    tm.mark();  // 0: start
    Arena& A = ctx.main;
    A.reserve(workspace_bytes<T, D>(n));

    // min / max (spatial_index.h:99-122)
    const int nb = minmax_partial_blocks(n);
    T* part = A.take_n<T>(static_cast<size_t>(nb) * 2 * D);
    T* mm = A.take_n<T>(2 * D + 8);
    int* nonfinite = reinterpret_cast<int*>(mm + 2 * D);
    HPD_HIP(hipMemsetAsync(nonfinite, 0, sizeof(int), s));
    launch_minmax<T, D>(pts, n, part, mm, nonfinite, s);
    T hmm[2 * D];
    int hbad = 0;
    HPD_HIP(hipMemcpyAsync(ctx.host_scratch, mm, sizeof(hmm) + sizeof(int), hipMemcpyDeviceToHost, s));
    HPD_HIP(hipStreamSynchronize(s));
    std::memcpy(hmm, ctx.host_scratch, sizeof(hmm));
    std::memcpy(&hbad, reinterpret_cast<const char*>(ctx.host_scratch) + sizeof(hmm), sizeof(int));
    if (hbad) throw Error(HPD_EINVAL, "input contains NaN or infinite coordinates");
    float rcp_f;
    double rcp_d;
    const Geometry g = make_geometry<T>(hmm, hmm + D, D, eps, &rcp_f, &rcp_d);
    if (g.overflow)
        throw Error(HPD_EUNSUPPORTED,
                    "the eps-grid has more than 2^64 cells (the reference's cell ids wrap)");
    ctx.stats.total_cells = g.total;
    ctx.stats.key_bits = g.key_bits;

    if (g.key_bits <= 32)
        run_keyed<T, D, uint32_t>(ctx, pts, n, eps, min_pts, labels, s, g, rcp_f, rcp_d, mm, tm);
    else
        run_keyed<T, D, uint64_t>(ctx, pts, n, eps, min_pts, labels, s, g, rcp_f, rcp_d, mm, tm);
    if (ctx.timing) {
        ctx.stats.ms_grid = tm.ms(0, 1);
        ctx.stats.ms_sort = tm.ms(1, 2);
        ctx.stats.ms_index = tm.ms(2, 3);
        ctx.stats.ms_core = tm.ms(3, 4);
        ctx.stats.ms_union = tm.ms(4, 5);
        ctx.stats.ms_border = tm.ms(5, 6);
        ctx.stats.ms_total = tm.ms(0, 6);
        ctx.stats.core_launches = 1;
        ctx.stats.ms_core_kernel = ctx.stats.ms_core;
    }
}

template <typename T>
static void run_typed(Context& ctx, const T* pts, int64_t n, int d, float eps, uint64_t min_pts,
                      int64_t* labels, hipStream_t s) {
    switch (d) {
        case 1: return run_dims<T, 1>(ctx, pts, n, eps, min_pts, labels, s);
        case 2: return run_dims<T, 2>(ctx, pts, n, eps, min_pts, labels, s);
        case 3: return run_dims<T, 3>(ctx, pts, n, eps, min_pts, labels, s);
        case 4: return run_dims<T, 4>(ctx, pts, n, eps, min_pts, labels, s);
        case 5: return run_dims<T, 5>(ctx, pts, n, eps, min_pts, labels, s);
        case 6: return run_dims<T, 6>(ctx, pts, n, eps, min_pts, labels, s);
        case 7: return run_dims<T, 7>(ctx, pts, n, eps, min_pts, labels, s);
        case 8: return run_dims<T, 8>(ctx, pts, n, eps, min_pts, labels, s);
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

static void run_any(Context& ctx, int dt, const void* pts, int64_t n, int d, float eps,
                    uint64_t min_pts, int64_t* labels, hipStream_t s) {
    switch (dt) {
        case HPD_DTYPE_F32: return run_typed(ctx, static_cast<const float*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_F64: return run_typed(ctx, static_cast<const double*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_I8: return run_typed(ctx, static_cast<const int8_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_I16: return run_typed(ctx, static_cast<const int16_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_I32: return run_typed(ctx, static_cast<const int32_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_I64: return run_typed(ctx, static_cast<const int64_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_U8: return run_typed(ctx, static_cast<const uint8_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_U16: return run_typed(ctx, static_cast<const uint16_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_U32: return run_typed(ctx, static_cast<const uint32_t*>(pts), n, d, eps, min_pts, labels, s);
        case HPD_DTYPE_U64: return run_typed(ctx, static_cast<const uint64_t*>(pts), n, d, eps, min_pts, labels, s);
        default: throw Error(HPD_EINVAL, "unknown dtype code " + std::to_string(dt));
    }
}

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
        if (rule == HPD_BORDER_ONE_THREAD)
            throw Error(HPD_EUNSUPPORTED, "the one-thread border rule is not available yet");
        ctx->impl.border_rule = rule;
    });
}

int hpd_context_set_timing(hpd_context* ctx, int enabled) {
    return guarded([&] {
        if (!ctx) throw Error(HPD_EINVAL, "context is NULL");
        ctx->impl.timing = enabled != 0;
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
        hpd::run_any(c, dtype, data, n, d, eps, min_pts, labels_out, s);
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
        hpd::run_any(c, dtype, din, n, d, eps, min_pts, dout, c.stream);
        HPD_HIP(hipMemcpyAsync(labels_out, dout, out_bytes, hipMemcpyDeviceToHost, c.stream));
        HPD_HIP(hipStreamSynchronize(c.stream));
    });
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
