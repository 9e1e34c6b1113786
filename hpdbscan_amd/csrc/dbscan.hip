// Neighbour-cell kernels of the clustering pass.
//
// Replaces local_dbscan (hpdbscan.h:51-101), region_query
// (spatial_index.h:739-768), the rules merge (rules.h, hpdbscan.h:198-211)
// and the scatter back to input order (spatial_index.h:852-872).
//
// The result the reference computes is order independent except for one
// case (border rows next to cores of two clusters):
//   core(p)  <=>  |{q in candidates(p) : dist2(p, q) <= EPS2}| >= minPts
//   cluster of a core = 1 + smallest sorted position of a core in its
//                       eps-connected component of cores
//   border row: within eps of a core; noise: neither.
// The kernels compute exactly that:
//   k_core    count within-eps candidates per row, stop at minPts
//   k_hook    every core points at the smallest adjacent core before it
//             (a forest whose roots are local minima; parent[x] <= x)
//   k_roots   root of every core (path-halving find), snapshot of the
//             roots, and per occupied cell the smallest / largest root of
//             its cores
//   k_union   union-find over the core pairs not yet known to be connected:
//             a neighbour row whose cells hold only cores of the chunk's
//             single root is skipped, a candidate whose snapshot root equals
//             the row's root is skipped, every other within-eps core pair
//             is united (larger root hooked under the smaller)
//   k_roots   final roots; core rows get -(root + 1)
//   k_border  smallest adjacent root for non-core rows (0 = noise); a
//             neighbourhood whose cells hold cores of one root only stops at
//             the first hit
// Each core pair is either tested or provably in one component already, so
// the components, and with them every label, are exact.
//
// Work decomposition: one wave per 64-row chunk of an occupied cell; lanes
// own the chunk's rows, the candidate rows stream through the wave in
// aligned batches fetched with wave-uniform addresses.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "boxes.hpp"
#include "cellkey.hpp"
#include "common.hpp"
#include "dbscan.hpp"
#include "grid.hpp"
#include "internal.hpp"
#include "visit.hpp"

namespace hpd {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;

// sorted position of a slot
__device__ __forceinline__ int32_t slot_pos(const uint32_t* __restrict__ pos_of_slot, int32_t s) {
    return pos_of_slot ? static_cast<int32_t>(pos_of_slot[s]) : s;
}

struct Chunk {
    int32_t u;       // occupied cell
    int32_t first;   // first row of the chunk
    int32_t end;     // end row of the cell
    int32_t r0, r1;  // its neighbour ranges (ascending, or nearest first)
};

__device__ __forceinline__ bool fetch_chunk(const int2* __restrict__ items, int32_t nitems,
                                            const int32_t* __restrict__ ustart,
                                            const uint32_t* __restrict__ roff, Chunk& c) {
    const int32_t wid = uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock +
                                static_cast<int32_t>(threadIdx.x >> 6));
    if (wid >= nitems) return false;
    const int2 it = items[wid];
    c.u = uniform(it.x);
    c.first = uniform(it.y);
    c.end = uniform(ustart[c.u + 1]);
    c.r0 = uniform(static_cast<int32_t>(roff[c.u]));
    c.r1 = uniform(static_cast<int32_t>(roff[c.u + 1]));
    return true;
}

// Chunk of work item wid (no bounds check).
__device__ __forceinline__ void chunk_at(const int2* __restrict__ items, int32_t wid,
                                         const int32_t* __restrict__ ustart,
                                         const uint32_t* __restrict__ roff, Chunk& c) {
    const int2 it = items[wid];
    c.u = uniform(it.x);
    c.first = uniform(it.y);
    c.end = uniform(ustart[c.u + 1]);
    c.r0 = uniform(static_cast<int32_t>(roff[c.u]));
    c.r1 = uniform(static_cast<int32_t>(roff[c.u + 1]));
}

// Number of work items: host-known, or written by an earlier kernel.
__device__ __forceinline__ int32_t item_count(int32_t host, const unsigned int* dev) {
    return dev ? uniform(static_cast<int32_t>(*dev)) : host;
}

// The work item of the calling wave: one wave per item where the grid covers
// the host-known count or its bound (waves past a device-written count
// exit); the waves of a smaller grid walk items wid, wid + wave_stride(), ...
__device__ __forceinline__ int32_t first_wave() {
    return uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock +
                   static_cast<int32_t>(threadIdx.x >> 6));
}
__device__ __forceinline__ int32_t wave_stride() {
    return static_cast<int32_t>(gridDim.x) * kWavesPerBlock;
}

// The boxes of chunk c's neighbour rows (for_each_row_box).
template <typename T, int D, typename Fn>
__device__ __forceinline__ void for_each_nbr_box(const DbscanArgs<T>& a, const Chunk& c,
                                                 const int32_t* __restrict__ tag, Fn&& fn,
                                                 bool own_cell_first = false) {
    for_each_row_box<T, D>(c.r0, c.r1, uniform(a.rcenter[c.u]), a.rcells, a.cbox, a.bstart,
                           static_cast<const OrdKey<T>*>(a.bbox), tag, fn, INT_MAX,
                           own_cell_first ? uniform(a.cbox[c.u]) : -1);
}

__device__ __forceinline__ int32_t wave_min(int32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Every occupied cell in [a, b) holds either no core or only cores of root R.
__device__ __forceinline__ bool cells_single_root(int2 cs, const int32_t* __restrict__ cmin,
                                                  const int32_t* __restrict__ cmax, int32_t R) {
    for (int32_t c = cs.x; c < cs.y; ++c) {
        const int32_t mn = cmin[c], mx = cmax[c];
        if (mn <= mx && (mn != R || mx != R)) return false;
    }
    return true;
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_core(const DbscanArgs<T> a, uint8_t* __restrict__ core,
                                                   unsigned long long* __restrict__ ncore) {
    Chunk c;
    if (!fetch_chunk(a.items, a.nitems, a.ustart, a.roff, c)) return;
    const JobParams<T>& jp = a.jobs[uniform(a.job_of_cell[c.u])];
    const T eps2 = jp.eps2;
    const uint32_t min_pts = jp.min_pts;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    T pc[D];
    load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
    uint32_t cnt = 0;
    // rows of a tight box with >= minPts rows are core without counting
    if (a.btight && valid && (a.btight[a.box_of_slot[p]] & 2)) cnt = min_pts;
    auto body = [&](int32_t, const T* qc, uint8_t, int32_t) {
        cnt += within<D>(pc, qc, eps2) ? 1u : 0u;
    };
    auto stop = [&] { return static_cast<bool>(__all(!valid || cnt >= min_pts)); };
    bool done = stop();
    // the cell's own row first: most neighbours are there
    const int32_t center = uniform(a.rcenter[c.u]);
    for (int32_t i = -1; i < c.r1 - c.r0 && !done; ++i) {
        const int32_t r = i < 0 ? center : c.r0 + i;
        if (i >= 0 && r == center) continue;
        const int2 rg = a.ranges[r];
        done = visit_range<T, D>(a.soa, a.ld, nullptr, nullptr, uniform(rg.x), uniform(rg.y), body,
                                 stop);
    }
    const bool is_core = valid && cnt >= min_pts;
    if (valid) core[p] = is_core ? 1 : 0;
    (void)ncore;
}

// Box-mode core counts.  For every neighbour box: lanes whose row is farther
// than eps from the box's bounding box skip it, lanes whose row is within
// eps of the box's farthest corner count all its rows, the other lanes
// test the box's rows one by one.  Exactness: with d the exact distance
// and gmin2 <= d^2 <= gmax2, the reference's float expression is within a
// few ulps of d^2, so gmin2 > EPS2 (1 + 1e-5) (+ 1e-30) rules every member
// out and gmax2 <= EPS2 (1 - 1e-5) rules every member in.
// kCount: also add the row-row distance tests executed to *ntests (an
// instrumented variant: the count is one more live value in a kernel that
// has no VGPR to spare at 8 waves per SIMD)
template <typename T, int D, bool kCount = false>
__device__ __forceinline__ void core_boxes_item(const DbscanArgs<T>& a, uint8_t* __restrict__ core,
                                                unsigned long long* __restrict__ ntests,
                                                int32_t wid) {
    Chunk c;
    chunk_at(a.ditems, wid, a.ustart, a.roff, c);
    uint32_t tests = 0;  // row-row distance tests executed (lanes x rows visited)
    // the first group's neighbour boxes are loaded ahead, beside the rows
    // and their boxes (the chunk was selected because a row needs counting)
    const int32_t center = uniform(a.rcenter[c.u]);
    const int32_t own_box = uniform(a.cbox[c.u]);
    RowBoxes pre;
    row_boxes(c.r0, c.r1, center, 0, a.rcells, a.cbox, INT_MAX, pre.rb0, pre.rcnt);
    const JobParams<T>& jp = a.jobs[uniform(a.job_of_cell[c.u])];
    const T eps2 = jp.eps2;
    const uint32_t min_pts = jp.min_pts;
    const double gap2_max = jp.gap2_max;
    const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    T pc[D];
    load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
    uint32_t cnt = 0;
    // rows of a tight box with >= minPts rows are core without counting
    const int32_t bp = valid ? a.box_of_slot[p] : -1 - lane_id();
    const bool full = valid && (a.btight[bp] & 2);
    if (full) cnt = min_pts;
    bool need = valid && cnt < min_pts;
    if (__any(need)) {
        // the group's boxes are first classified against the bounding box of
        // the rows still counting, one box per lane: a box farther than eps
        // from all of them is skipped, one within eps of all of them counts
        // whole for every row (both exactly as the per-row tests below would
        // decide, box_box bounding them monotonically); only the others are
        // walked box by box
        double clo[D], chi[D];
        rows_bbox<T, D>(pc, need, clo, chi);
        for_each_row_box_group<T, D>(
            c.r0, c.r1, center, a.rcells, a.cbox, a.bstart,
            static_cast<const OrdKey<T>*>(a.bbox), nullptr,
            [&](int32_t nk, int32_t bl, const double* lol, const double* hil, int32_t q0l,
                int32_t q1l, int32_t) {
                double cmin2, cmax2;
                box_box<D>(clo, chi, lol, hil, cmin2, cmax2);
                const bool mine = lane_id() < nk;
                const bool all_in = mine && cmax2 <= full2;
                const bool some = mine && !all_in && cmin2 <= gap2_max;
                uint32_t add = all_in ? static_cast<uint32_t>(q1l - q0l) : 0u;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o, 64);
                if (need) cnt += add;
                need = valid && cnt < min_pts;
                const uint64_t somem = ballot(some);
                const uint64_t needm = ballot(need);
                if (D <= 4 && __popcll(needm) < __popcll(somem)) {
                    // fewer rows than boxes to decide: row by row, every
                    // box on its lane at once (the same point-box test as
                    // box by box below); the boxes within eps of a row's
                    // farthest corner count whole, the others that are
                    // near it are marked for the row and tested row by row
                    uint64_t part_rows = 0;  // on a box's lane: rows it is partial for
                    for (uint64_t rows = needm; rows != 0; rows &= rows - 1) {
                        const int r = static_cast<int>(__builtin_ctzll(rows));
                        double v[D];
#pragma unroll
                        for (int e = 0; e < D; ++e) v[e] = bcast(static_cast<double>(pc[e]), r);
                        double gmin2, gmax2;
                        point_box_dd<D>(lol, hil, v, gmin2, gmax2);
                        const bool near = some && gmin2 <= gap2_max;
                        const bool whole = near && gmax2 <= full2;
                        uint32_t w = whole ? static_cast<uint32_t>(q1l - q0l) : 0u;
#pragma unroll
                        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
                        if (lane_id() == r) cnt += w;
                        if (near && !whole) part_rows |= uint64_t{1} << r;
                    }
                    need = valid && cnt < min_pts;
                    for (uint64_t todo = ballot(part_rows != 0); todo != 0 && __any(need);
                         todo &= todo - 1) {
                        const int j = static_cast<int>(__builtin_ctzll(todo));
                        const bool part = need && ((bcast64(part_rows, j) >> lane_id()) & 1);
                        if (__any(part)) {
                            const uint32_t lanes = static_cast<uint32_t>(__popcll(ballot(part)));
                            auto body = [&](int32_t, const T* qc, uint8_t, int32_t) {
                                cnt += (part && within<D>(pc, qc, eps2)) ? 1u : 0u;
                                if constexpr (kCount) tests += lanes;
                            };
                            auto stop = [&] {
                                return static_cast<bool>(__all(!part || cnt >= min_pts));
                            };
                            visit_range<T, D>(a.soa, a.ld, nullptr, nullptr, bcast(q0l, j),
                                              bcast(q1l, j), body, stop);
                        }
                        need = valid && cnt < min_pts;
                    }
                    return !__any(need);
                }
                for (uint64_t todo = somem; todo != 0 && __any(need); todo &= todo - 1) {
                    const int j = static_cast<int>(__builtin_ctzll(todo));
                    double lo[D], hi[D];
#pragma unroll
                    for (int e = 0; e < D; ++e) {
                        lo[e] = bcast(lol[e], j);
                        hi[e] = bcast(hil[e], j);
                    }
                    const int32_t q0 = bcast(q0l, j), q1 = bcast(q1l, j);
                    double gmin2, gmax2;
                    point_box_lh<T, D>(lo, hi, pc, gmin2, gmax2);
                    const bool near = need && gmin2 <= gap2_max;
                    if (!__any(near)) continue;
                    const bool whole = near && gmax2 <= full2;
                    if (whole) cnt += static_cast<uint32_t>(q1 - q0);
                    const bool part = near && !whole;
                    if (__any(part)) {
                        const uint32_t lanes = static_cast<uint32_t>(__popcll(ballot(part)));
                        auto body = [&](int32_t, const T* qc, uint8_t, int32_t) {
                            cnt += (part && within<D>(pc, qc, eps2)) ? 1u : 0u;
                            if constexpr (kCount) tests += lanes;
                        };
                        auto stop = [&] { return static_cast<bool>(__all(!part || cnt >= min_pts)); };
                        visit_range<T, D>(a.soa, a.ld, nullptr, nullptr, q0, q1, body, stop);
                    }
                    need = valid && cnt < min_pts;
                }
                return !__any(need);
            },
            INT_MAX, own_box, &pre);
    }
    const bool is_core = valid && cnt >= min_pts;
    if (valid) core[p] = is_core ? 1 : 0;
    if constexpr (kCount)
        if (tests != 0 && lane_id() == 0) atomicAdd(ntests, static_cast<unsigned long long>(tests));
    if (a.bmincore) {
        // smallest core input row per box.  A box holds its rows in input
        // order, and every row of a tight box with >= minPts rows is core:
        // its value is its first slot's input row, stored by that lane (no
        // atomics on one address from every chunk of a large box).  Other
        // boxes: one inside this chunk is stored by its first lane, one that
        // extends past it is merged atomically.
        const uint32_t row = valid ? a.inrow[p] : 0u;
        if (full && a.bstart[bp] == p) a.bmincore[bp] = static_cast<int32_t>(row);
        const int32_t m =
            run_suffix_min<int32_t>(is_core && !full ? static_cast<int32_t>(row) : INT_MAX, bp);
        const bool first = run_head(bp);  // every lane takes part in the shuffles
        if (valid && first && m != INT_MAX) {
            if (a.bstart[bp] >= c.first && a.bstart[bp + 1] <= c.first + kWave)
                a.bmincore[bp] = m;
            else
                atomicMin(&a.bmincore[bp], m);
        }
    }
}

// One wave per work item.
template <typename T, int D, bool kCount = false>
__global__ __launch_bounds__(kThreads, 8) void k_core_boxes(const DbscanArgs<T> a,
                                                         uint8_t* __restrict__ core,
                                                         unsigned long long* __restrict__ ntests) {
    const int32_t wid = first_wave();
    if (wid < item_count(a.nditems, a.nditems_dev))
        core_boxes_item<T, D, kCount>(a, core, ntests, wid);
}

// The waves walk items wid0, wid0 + 1, ...: the items of a device-written
// count beyond what a one-wave-per-item grid covered.  The walk keeps the
// arguments live across items, more than the 64 VGPRs of 8 waves per SIMD
// hold without spilling, so it runs at 4.
template <typename T, int D, bool kCount = false>
__global__ __launch_bounds__(kThreads, 4) void k_core_boxes_walk(
    const DbscanArgs<T> a, uint8_t* __restrict__ core, unsigned long long* __restrict__ ntests,
    int32_t wid0) {
    const int32_t nwork = item_count(a.nditems, a.nditems_dev);
    for (int32_t wid = wid0 + first_wave(); wid < nwork; wid += wave_stride())
        core_boxes_item<T, D, kCount>(a, core, ntests, wid);
}

// Thread-per-row core counts for the rows of sparse cells (a.sitems): the
// same box-by-box count as k_core_boxes, with every thread walking its own
// row's neighbourhood.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_core_sparse(const DbscanArgs<T> a,
                                                          uint8_t* __restrict__ core,
                                                          unsigned long long* __restrict__ ncore) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    bool is_core = false;
    if (i < item_count(a.nsitems, a.nsitems_dev)) {
        const int2 it = a.sitems[i];
        const int32_t p = it.x, u = it.y;
        const int32_t b = a.box_of_slot[p];
        {
            const JobParams<T>& jp = a.jobs[a.job_of_cell[u]];
            const uint32_t min_pts = jp.min_pts;
            uint32_t cnt = (a.btight[b] & 2) ? min_pts : 0u;
            if (cnt < min_pts) {
                const T eps2 = jp.eps2;
                const double gap2_max = jp.gap2_max;
                const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
                const OrdKey<T>* bbox = static_cast<const OrdKey<T>*>(a.bbox);
                T pc[D];
                load_row<T, D>(a.soa, a.ld, p, pc);
                // the own row first, from the own cell's first box on
                const int32_t r0 = static_cast<int32_t>(a.roff[u]);
                const int32_t nr = static_cast<int32_t>(a.roff[u + 1]) - r0;
                const int32_t center = a.rcenter[u];
                const int32_t own = a.cbox[u];
                for (int32_t k = -1; k < nr && cnt < min_pts; ++k) {
                    const int32_t r = k < 0 ? center : r0 + k;
                    if (k >= 0 && r == center) continue;
                    const int2 cs = a.rcells[r];
                    const int32_t e0 = a.cbox[cs.x], len = a.cbox[cs.y] - e0;
                    const int32_t rot = (k < 0 && own > e0 && own < e0 + len) ? own - e0 : 0;
                    for (int32_t t = 0; t < len && cnt < min_pts; ++t) {
                        const int32_t b2 = e0 + (t + rot < len ? t + rot : t + rot - len);
                        double gmin2, gmax2;
                        point_box<T, D>(bbox, b2, pc, gmin2, gmax2);
                        if (gmin2 > gap2_max) continue;
                        const int32_t q0 = a.bstart[b2], q1 = a.bstart[b2 + 1];
                        if (gmax2 <= full2) {
                            cnt += static_cast<uint32_t>(q1 - q0);
                            continue;
                        }
                        for (int32_t q = q0; q < q1 && cnt < min_pts; ++q) {
                            T qc[D];
                            load_row<T, D>(a.soa, a.ld, q, qc);
                            cnt += within<D>(pc, qc, eps2) ? 1u : 0u;
                        }
                    }
                }
            }
            is_core = cnt >= min_pts;
            core[p] = is_core ? 1 : 0;
            if (is_core && a.bmincore) atomicMin(&a.bmincore[b], static_cast<int32_t>(a.inrow[p]));
        }
    }
    (void)ncore;
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_hook(const DbscanArgs<T> a,
                                                   const uint8_t* __restrict__ core,
                                                   int32_t* __restrict__ parent) {
    Chunk c;
    if (!fetch_chunk(a.items, a.nitems, a.ustart, a.roff, c)) return;
    const T eps2 = a.jobs[uniform(a.job_of_cell[c.u])].eps2;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && core[p];
    if (!__any(active)) return;
    T pc[D];
    load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
    // any within-eps core before p is a valid parent (parent[x] < x keeps the
    // hooks a forest); the first one found is taken
    int32_t best = p;
    bool found = !active;
    auto body = [&](int32_t q, const T* qc, uint8_t cq, int32_t) {
        if (!cq) return;
        if (!found && q < p && within<D>(pc, qc, eps2)) {
            best = q;
            found = true;
        }
    };
    auto stop = [&] { return static_cast<bool>(__all(found)); };
    const int32_t pmax = uniform(min(c.first + kWave, c.end) - 1);
    // a cell's rows may be stored nearest first rather than ascending, so a
    // row that starts at or after pmax is skipped, not taken as the end
    for (int32_t r = c.r0; r < c.r1; ++r) {
        const int2 rg = a.ranges[r];
        const int32_t qb = uniform(rg.x);
        if (qb >= pmax) continue;
        if (visit_range<T, D>(a.soa, a.ld, core, nullptr, qb, uniform(min(rg.y, pmax)), body, stop))
            break;
    }
    if (active) parent[p] = best;
}

// Roots of the chunk's cores (path-halving find), stored back into parent,
// snapshot copy, and per-cell smallest / largest root.
__global__ __launch_bounds__(kThreads) void k_roots(const int2* __restrict__ items, int32_t nitems,
                                                    const int32_t* __restrict__ ustart,
                                                    const uint32_t* __restrict__ roff,
                                                    const uint8_t* __restrict__ core,
                                                    int32_t* parent, int32_t* __restrict__ rsnap,
                                                    int32_t* __restrict__ cmin,
                                                    int32_t* __restrict__ cmax) {
    Chunk c;
    if (!fetch_chunk(items, nitems, ustart, roff, c)) return;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && core[p];
    if (!__any(active)) return;
    int32_t r = INT_MAX;
    if (active) {
        r = uf_find(parent, p);
        st_agent(&parent[p], r);
        rsnap[p] = r;
    }
    const int32_t rmin = wave_min(active ? r : INT_MAX);
    const int32_t rmax = wave_max(active ? r : -1);
    if (lane_id() == 0) {
        atomicMin(&cmin[c.u], rmin);
        atomicMax(&cmax[c.u], rmax);
    }
}

// smallest sorted position of every slot-level component
__global__ __launch_bounds__(kThreads) void k_compmin(const uint8_t* __restrict__ core,
                                                      int32_t* parent, int64_t n,
                                                      const uint32_t* __restrict__ pos_of_slot,
                                                      int32_t* __restrict__ compmin) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n || !core[i]) return;
    const int32_t r = uf_find(parent, static_cast<int32_t>(i));
    const int32_t pos = pos_of_slot ? static_cast<int32_t>(pos_of_slot[i]) : static_cast<int32_t>(i);
    atomicMin(&compmin[r], pos);
}

// cluster of every core slot -> label, slot array of clusters, per-cell
// smallest / largest cluster
template <typename T>
__global__ __launch_bounds__(kThreads) void k_label_cores(
    const int2* __restrict__ items, int32_t nitems, const int32_t* __restrict__ ustart,
    const uint32_t* __restrict__ roff, const JobParams<T>* __restrict__ jobs,
    const int32_t* __restrict__ job_of_cell, const uint8_t* __restrict__ core,
    const int32_t* __restrict__ box_of_slot, const int32_t* __restrict__ bcluster, int32_t* parent,
    const int32_t* __restrict__ compmin, const uint32_t* __restrict__ inrow,
    int64_t* __restrict__ labels, int32_t* __restrict__ rootslot, int32_t* __restrict__ cmin,
    int32_t* __restrict__ cmax) {
    Chunk c;
    if (!fetch_chunk(items, nitems, ustart, roff, c)) return;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && core[p];
    if (!__any(active)) return;
    const int64_t first_pos = jobs[uniform(job_of_cell[c.u])].row_begin;
    int32_t r = INT_MAX;
    if (active) {
        r = bcluster ? bcluster[box_of_slot[p]] : compmin[uf_find(parent, p)];
        labels[inrow[p]] = -(static_cast<int64_t>(r) - first_pos + 1);
        rootslot[p] = r;
    }
    const int32_t rmin = wave_min(active ? r : INT_MAX);
    const int32_t rmax = wave_max(active ? r : -1);
    if (lane_id() == 0) {
        atomicMin(&cmin[c.u], rmin);
        atomicMax(&cmax[c.u], rmax);
    }
}

// Box mode: core labels one thread per slot (the cluster of a core is its
// box's); cmin / cmax are not needed there.
// 8 waves on each of the 1024 SIMDs
constexpr unsigned kLabelBlocks = 2048;

// skip_full (with *full_ok set): the rows of full boxes (btight bit 2) were
// labelled in input order (k_label_full_rows) and are skipped here.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_label_cores_boxes(
    int64_t n, const JobParams<T>* __restrict__ jobs, const int32_t* __restrict__ job_of_cell,
    const uint8_t* __restrict__ core, const int32_t* __restrict__ box_of_slot,
    const int32_t* __restrict__ bcell, const int32_t* __restrict__ bcluster,
    const uint32_t* __restrict__ inrow, int64_t* __restrict__ labels,
    int32_t* __restrict__ rootslot, unsigned long long* __restrict__ ncore, bool one_job,
    const uint8_t* __restrict__ skip_full, const unsigned int* __restrict__ full_ok) {
    if (skip_full && *full_ok == 0) skip_full = nullptr;
    // grid-stride over at most kLabelBlocks blocks, so that the core count
    // of the run statistics (folded in here) takes few atomics
    unsigned long long cores = 0;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; p < n;
         p += static_cast<int64_t>(gridDim.x) * kThreads) {
        if (!core[p]) continue;
        ++cores;  // every core, the row-order ones included
        const int32_t b = box_of_slot[p];
        if (skip_full && (skip_full[b] & 2)) continue;
        const int32_t r = bcluster[b];
        if (rootslot) rootslot[p] = r;
        // one job starts at row 0
        const int64_t first_pos = one_job ? 0 : jobs[job_of_cell[bcell[b]]].row_begin;
        labels[inrow[p]] = -(static_cast<int64_t>(r) - first_pos + 1);
    }
    if (ncore) block_add<kThreads>(cores, ncore);
}

// Row-order labels of full boxes.  A table indexed by cell key (the sort's
// keys start at 0 and are dense enough when the largest one is below the
// table's capacity) holds, for the key of every full box (tight with >=
// minPts rows: every row core, cluster = the box's), the label's magnitude
// bcluster - first row of the job + 1, and 0 for the keys of other boxes.
// Keys no box has are never looked up.  *ok = the largest key fits.
template <typename T, typename K>
__global__ __launch_bounds__(kThreads) void k_full_box_table(
    const K* __restrict__ keys, int64_t n, const int32_t* __restrict__ bstart, Count nbox,
    const int32_t* __restrict__ bcell, const int32_t* __restrict__ job_of_cell,
    const JobParams<T>* __restrict__ jobs, const uint8_t* __restrict__ btight,
    const int32_t* __restrict__ bcluster, int32_t* __restrict__ tab, int64_t cap,
    unsigned int* __restrict__ ok) {
    const bool fits = static_cast<uint64_t>(keys[n - 1]) < static_cast<uint64_t>(cap);
    const int64_t b = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b == 0) *ok = fits ? 1u : 0u;
    if (!fits || b >= nbox.get()) return;
    int32_t v = 0;
    if (btight[b] & 2)
        v = static_cast<int32_t>(static_cast<int64_t>(bcluster[b]) -
                                 jobs[job_of_cell[bcell[b]]].row_begin + 1);
    tab[static_cast<int64_t>(keys[bstart[b]])] = v;
}

// The labels of the rows of full boxes, written in input order: each row's
// key is computed again from its coordinates exactly as the sort's keys
// were (row_key), looked up in the table, and the label stored at the row's
// own place (coalesced, where labelling in slot order stores at permuted
// places).  Rows of other boxes (table value 0) are left to
// k_label_cores_boxes.  Nothing when the table did not fit.
template <typename T, int D, typename K>
__global__ __launch_bounds__(kThreads) void k_label_full_rows(
    const T* __restrict__ pts, int64_t n, const JobParams<T>* __restrict__ jobs, int32_t njobs,
    int sub_k, int sub_bits, const int32_t* __restrict__ tab, int64_t cap,
    const unsigned int* __restrict__ ok, int64_t* __restrict__ labels) {
    if (*ok == 0) return;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kThreads) {
        const JobParams<T>& jp = jobs[job_of_row_wave(jobs, njobs, i)];
        const K key = static_cast<K>(row_key<T, D>(pts + i * D, jp, sub_k, sub_bits));
        // (every row's key is a box's key, below cap when *ok is set)
        const int32_t v = static_cast<uint64_t>(key) < static_cast<uint64_t>(cap)
                              ? tab[static_cast<int64_t>(key)]
                              : 0;
        if (v != 0) labels[i] = -static_cast<int64_t>(v);
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_union(const DbscanArgs<T> a,
                                                    const uint8_t* __restrict__ core,
                                                    const int32_t* __restrict__ rsnap,
                                                    const int32_t* __restrict__ cmin,
                                                    const int32_t* __restrict__ cmax,
                                                    int32_t* parent) {
    Chunk c;
    if (!fetch_chunk(a.items, a.nitems, a.ustart, a.roff, c)) return;
    const T eps2 = a.jobs[uniform(a.job_of_cell[c.u])].eps2;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && core[p];
    if (!__any(active)) return;
    const int32_t rp = active ? rsnap[p] : -1;
    const int32_t rmin = wave_min(active ? rp : INT_MAX);
    const int32_t rmax = wave_max(active ? rp : -1);
    const int32_t R = uniform(rmin == rmax ? rmin : -1);  // the chunk's single root, if any
    T pc[D];
    load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
    // every core pair is handled by its larger position: only q < p
    const int32_t pmax = uniform(min(c.first + kWave, c.end) - 1);
    auto body = [&](int32_t q, const T* qc, uint8_t cq, int32_t rq) {
        if (!cq) return;
        const bool need = active && q < p && rq != rp;
        if (!__any(need)) return;
        if (need && within<D>(pc, qc, eps2)) uf_unite(parent, p, q);
    };
    auto never = [] { return false; };
    for (int32_t r = c.r0; r < c.r1; ++r) {
        const int2 rg = a.ranges[r];
        const int32_t qb = uniform(rg.x);
        if (qb >= pmax) continue;  // rows are not always ascending (nearest first)
        if (R >= 0 && cells_single_root(a.rcells[r], cmin, cmax, R)) continue;
        visit_range<T, D>(a.soa, a.ld, core, rsnap, qb, uniform(min(rg.y, pmax)), body, never);
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_border(const DbscanArgs<T> a,
                                                     const uint8_t* __restrict__ core,
                                                     const int32_t* __restrict__ root,
                                                     const int32_t* __restrict__ cmin,
                                                     const int32_t* __restrict__ cmax,
                                                     const uint32_t* __restrict__ order,
                                                     int64_t* __restrict__ labels,
                                                     unsigned long long* __restrict__ stats,
                                                     int2* __restrict__ amb_list) {
    Chunk c;
    if (!fetch_chunk(a.items, a.nitems, a.ustart, a.roff, c)) return;
    const JobParams<T>& jp = a.jobs[uniform(a.job_of_cell[c.u])];
    const T eps2 = jp.eps2;
    const int64_t first_pos = jp.row_begin;  // cluster ids are positions inside the job
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && !core[p];
    if (!__any(active)) return;
    // do the candidate cells hold cores of a single root only?
    int32_t R = -1;
    bool single = true;
    for (int32_t r = c.r0; r < c.r1 && single; ++r) {
        const int2 cs = a.rcells[r];
        for (int32_t cc = cs.x; cc < cs.y; ++cc) {
            const int32_t mn = cmin[cc], mx = cmax[cc];
            if (mn > mx) continue;
            if (mn != mx || (R >= 0 && R != mn)) {
                single = false;
                break;
            }
            R = mn;
        }
    }
    R = uniform(R);
    int32_t best = INT_MAX, worst = -1;
    if (!(single && R < 0)) {  // otherwise no core is near: all noise
        T pc[D];
        load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
        auto body = [&](int32_t, const T* qc, uint8_t cq, int32_t rq) {
            if (!cq) return;
            if (active && within<D>(pc, qc, eps2)) {
                best = min(best, rq);
                worst = max(worst, rq);
            }
        };
        auto stop = [&] {
            return single && static_cast<bool>(__all(!active || best != INT_MAX));
        };
        const int32_t center = uniform(a.rcenter[c.u]);
        for (int32_t i = -1; i < c.r1 - c.r0; ++i) {
            const int32_t r = i < 0 ? center : c.r0 + i;
            if (i >= 0 && r == center) continue;
            const int2 rg = a.ranges[r];
            if (visit_range<T, D>(a.soa, a.ld, core, root, uniform(rg.x), uniform(rg.y), body,
                                  stop))
                break;
        }
    }
    if (active) labels[order[p]] = best == INT_MAX ? 0 : static_cast<int64_t>(best) - first_pos + 1;
    const bool is_amb = active && best != INT_MAX && best != worst;
    const uint64_t amb_mask = ballot(is_amb);
    const unsigned long long amb = __popcll(amb_mask);
    unsigned long long base = 0;
    if (lane_id() == 0) {
        if (amb) base = atomicAdd(&stats[1], amb);
    }
    // record (slot, cell) of every border row next to two or more clusters
    if (amb_list && amb) {
        base = __shfl(base, 0, 64);
        if (is_amb)
            amb_list[base + static_cast<unsigned long long>(__popcll(amb_mask & lanemask_lt()))] =
                make_int2(p, c.u);
    }
}

// Box-mode border labels: every core of a box lies in the box's cluster
// (bcluster), so a non-core row only needs, per neighbour box whose cluster
// could change its answer, one within-eps core of that box (or a box within
// eps as a whole).  A box is skipped for a lane when its cluster is the
// lane's current one, or larger while the lane has not yet met two
// clusters; the smallest adjacent cluster and the two-cluster flag come out
// exact.
// Noise rows of the wave, added to one of kNoiseStripes counters (picked by
// the wave's index): one counter per wave would serialise the atomics of
// all waves with noise on one address.  k_sum_stripes adds them up.
__device__ __forceinline__ void count_noise_rows(bool is_noise, unsigned long long* noise) {
    if (!noise) return;
    const uint64_t nz = ballot(is_noise);
    if (lane_id() == 0 && nz) {
        const uint32_t w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
        atomicAdd(&noise[w & (kNoiseStripes - 1)], static_cast<unsigned long long>(__popcll(nz)));
    }
}

__global__ __launch_bounds__(kThreads) void k_sum_stripes(const unsigned long long* __restrict__ v,
                                                          unsigned long long* __restrict__ out) {
    unsigned long long t = 0;
    for (int i = threadIdx.x; i < kNoiseStripes; i += kThreads) t += v[i];
    block_add<kThreads>(t, out);
}

template <typename T, int D>
__device__ __forceinline__ void border_boxes_item(
    const DbscanArgs<T>& a, const uint8_t* __restrict__ core, const int32_t* __restrict__ bcluster,
    const uint32_t* __restrict__ order, int64_t* __restrict__ labels,
    unsigned long long* __restrict__ stats, int2* __restrict__ amb_list,
    unsigned long long* __restrict__ noise, int32_t wid) {
    Chunk c;
    chunk_at(a.ditems, wid, a.ustart, a.roff, c);
    const JobParams<T>& jp = a.jobs[uniform(a.job_of_cell[c.u])];
    const T eps2 = jp.eps2;
    const double gap2_max = jp.gap2_max;
    const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
    const int64_t first_pos = jp.row_begin;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && !core[p];
    if (!__any(active)) return;
    T pc[D];
    load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
    int32_t best = INT_MAX;
    bool amb = false;
    for_each_nbr_box<T, D>(a, c, bcluster, [&](int32_t, const double* lo, const double* hi,
                                               int32_t q0, int32_t q1, int32_t cl) {
        if (cl < 0) return false;  // no cores
        const bool want = active && cl != best && (cl < best || !amb);
        if (!__any(want)) return false;
        double gmin2, gmax2;
        point_box_lh<T, D>(lo, hi, pc, gmin2, gmax2);
        const bool near = want && gmin2 <= gap2_max;
        if (!__any(near)) return false;
        bool hit = near && gmax2 <= full2;
        const bool part = near && !hit;
        if (__any(part)) {
            auto body = [&](int32_t, const T* qc, uint8_t cq, int32_t) {
                if (part && !hit && cq && within<D>(pc, qc, eps2)) hit = true;
            };
            auto stop = [&] { return static_cast<bool>(__all(!part || hit)); };
            visit_range<T, D>(a.soa, a.ld, core, nullptr, q0, q1, body, stop);
        }
        if (hit) {
            if (best != INT_MAX) amb = true;
            best = min(best, cl);
        }
        return false;
    }, true);
    if (active) labels[order[p]] = best == INT_MAX ? 0 : static_cast<int64_t>(best) - first_pos + 1;
    count_noise_rows(active && best == INT_MAX, noise);
    const bool is_amb = active && amb;
    const uint64_t amb_mask = ballot(is_amb);
    const unsigned long long namb = __popcll(amb_mask);
    unsigned long long base = 0;
    if (lane_id() == 0) {
        if (namb) base = atomicAdd(&stats[1], namb);
    }
    if (amb_list && namb) {
        base = __shfl(base, 0, 64);
        if (is_amb)
            amb_list[base + static_cast<unsigned long long>(__popcll(amb_mask & lanemask_lt()))] =
                make_int2(p, c.u);
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads, 8) void k_border_boxes(
    const DbscanArgs<T> a, const uint8_t* __restrict__ core, const int32_t* __restrict__ bcluster,
    const uint32_t* __restrict__ order, int64_t* __restrict__ labels,
    unsigned long long* __restrict__ stats, int2* __restrict__ amb_list,
    unsigned long long* __restrict__ noise) {
    const int32_t wid = first_wave();
    if (wid < item_count(a.nditems, a.nditems_dev))
        border_boxes_item<T, D>(a, core, bcluster, order, labels, stats, amb_list, noise, wid);
}

// k_border_boxes with the waves walking items wid0, ... (as k_core_boxes_walk).
template <typename T, int D>
__global__ __launch_bounds__(kThreads, 4) void k_border_boxes_walk(
    const DbscanArgs<T> a, const uint8_t* __restrict__ core, const int32_t* __restrict__ bcluster,
    const uint32_t* __restrict__ order, int64_t* __restrict__ labels,
    unsigned long long* __restrict__ stats, int2* __restrict__ amb_list,
    unsigned long long* __restrict__ noise, int32_t wid0) {
    const int32_t nwork = item_count(a.nditems, a.nditems_dev);
    for (int32_t wid = wid0 + first_wave(); wid < nwork; wid += wave_stride())
        border_boxes_item<T, D>(a, core, bcluster, order, labels, stats, amb_list, noise, wid);
}

// Thread-per-row border labels for the non-core rows of sparse cells
// (a.sitems); same box-by-box search as k_border_boxes.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_border_sparse(const DbscanArgs<T> a,
                                                            const uint8_t* __restrict__ core,
                                                            const int32_t* __restrict__ bcluster,
                                                            int64_t* __restrict__ labels,
                                                            unsigned long long* __restrict__ stats,
                                                            int2* __restrict__ amb_list,
                                                            unsigned long long* __restrict__ noise) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    bool active = false, amb = false;
    int32_t best = INT_MAX, p = 0, u = 0;
    if (s < item_count(a.nsitems, a.nsitems_dev)) {
        const int2 it = a.sitems[s];
        p = it.x;
        u = it.y;
        active = !core[p];
    }
    if (active) {
        const JobParams<T>& jp = a.jobs[a.job_of_cell[u]];
        const T eps2 = jp.eps2;
        const double gap2_max = jp.gap2_max;
        const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
        const OrdKey<T>* bbox = static_cast<const OrdKey<T>*>(a.bbox);
        T pc[D];
        load_row<T, D>(a.soa, a.ld, p, pc);
        // the own row first, from the own cell's first box on: the own
        // cluster is found first and its other boxes are skipped
        const int32_t r0 = static_cast<int32_t>(a.roff[u]);
        const int32_t nr = static_cast<int32_t>(a.roff[u + 1]) - r0;
        const int32_t center = a.rcenter[u];
        const int32_t own = a.cbox[u];
        for (int32_t k = -1; k < nr; ++k) {
            const int32_t r = k < 0 ? center : r0 + k;
            if (k >= 0 && r == center) continue;
            const int2 cs = a.rcells[r];
            const int32_t e0 = a.cbox[cs.x], len = a.cbox[cs.y] - e0;
            const int32_t rot = (k < 0 && own > e0 && own < e0 + len) ? own - e0 : 0;
            for (int32_t t = 0; t < len; ++t) {
                const int32_t b2 = e0 + (t + rot < len ? t + rot : t + rot - len);
                const int32_t cl = bcluster[b2];
                if (cl < 0 || cl == best || (cl > best && amb)) continue;
                double gmin2, gmax2;
                point_box<T, D>(bbox, b2, pc, gmin2, gmax2);
                if (gmin2 > gap2_max) continue;
                bool hit = gmax2 <= full2;
                const int32_t q1 = a.bstart[b2 + 1];
                for (int32_t q = a.bstart[b2]; q < q1 && !hit; ++q) {
                    if (!core[q]) continue;
                    T qc[D];
                    load_row<T, D>(a.soa, a.ld, q, qc);
                    hit = within<D>(pc, qc, eps2);
                }
                if (hit) {
                    if (best != INT_MAX) amb = true;
                    best = min(best, cl);
                }
            }
        }
        labels[a.inrow[p]] =
            best == INT_MAX ? 0 : static_cast<int64_t>(best) - jp.row_begin + 1;
    }
    count_noise_rows(active && best == INT_MAX, noise);
    const bool is_amb = active && amb;
    const uint64_t amb_mask = ballot(is_amb);
    const unsigned long long namb = __popcll(amb_mask);
    unsigned long long base = 0;
    if (lane_id() == 0) {
        if (namb) base = atomicAdd(&stats[1], namb);
    }
    if (amb_list && namb) {
        base = __shfl(base, 0, 64);
        if (is_amb)
            amb_list[base + static_cast<unsigned long long>(__popcll(amb_mask & lanemask_lt()))] =
                make_int2(p, u);
    }
}

// ---------------------------------------------------------------------------
// The reference's border choice at one thread (HPD_BORDER_ONE_THREAD).
//
// Processing rows in sorted order, the reference gives core c the
// provisional label f(c) + 1 with f(c) = min(pos(c), f(q) over within-eps
// cores q before c) (hpdbscan.h:74-92 with spatial_index.h:757-763), i.e. the
// smallest position reachable from c along strictly decreasing positions.
// A border row keeps the smallest provisional label among its adjacent cores
// (the atomic_min of hpdbscan.h:91) and ends in that core's cluster.  Only
// border rows next to two or more clusters depend on it; f is computed for
// the cores of the clusters they touch, by in-place relaxation to the fixed
// point.
// ---------------------------------------------------------------------------

// clusters touched by the ambiguous border rows -> involved[cluster] = 1
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_ot_mark(const DbscanArgs<T> a,
                                                      const uint8_t* __restrict__ core,
                                                      const int32_t* __restrict__ rootslot,
                                                      const int2* __restrict__ amb, int32_t namb,
                                                      uint8_t* __restrict__ involved) {
    const int32_t i = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= namb) return;
    const int32_t p = amb[i].x, u = amb[i].y;
    const T eps2 = a.jobs[a.job_of_cell[u]].eps2;
    T pc[D];
    load_row<T, D>(a.soa, a.ld, p, pc);
    for (uint32_t r = a.roff[u]; r < a.roff[u + 1]; ++r) {
        const int2 rg = a.ranges[r];
        for (int32_t q = rg.x; q < rg.y; ++q) {
            if (!core[q]) continue;
            T qc[D];
            load_row<T, D>(a.soa, a.ld, q, qc);
            if (within<D>(pc, qc, eps2)) involved[rootslot[q]] = 1;
        }
    }
}

// one in-place relaxation sweep of f over the cores of involved clusters;
// f is indexed by slot and holds positions
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_ot_relax(const DbscanArgs<T> a,
                                                       const uint8_t* __restrict__ core,
                                                       const int32_t* __restrict__ rootslot,
                                                       const uint8_t* __restrict__ involved,
                                                       int32_t* f, unsigned int* __restrict__ changed) {
    Chunk c;
    if (!fetch_chunk(a.items, a.nitems, a.ustart, a.roff, c)) return;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && core[p] && involved[rootslot[p]];
    if (!__any(active)) return;
    const T eps2 = a.jobs[uniform(a.job_of_cell[c.u])].eps2;
    const int32_t pp = valid ? slot_pos(a.pos_of_slot, p) : 0;
    T pc[D];
    load_row<T, D>(a.soa, a.ld, valid ? p : c.first, pc);
    int32_t fmin = active ? ld_agent(&f[p]) : INT_MAX;
    const int32_t f0 = fmin;
    for (int32_t r = c.r0; r < c.r1; ++r) {
        const int2 rg = a.ranges[r];
        const int32_t qb = uniform(rg.x), qe = uniform(rg.y);
        // a row starting past the chunk's cell holds later positions only;
        // rows are not always ascending (nearest first), so skip, not stop
        if (qb >= c.end) continue;
        for (int32_t q = qb; q < qe; ++q) {
            if (!core[q]) continue;
            const int32_t qp = slot_pos(a.pos_of_slot, q);
            T qc[D];
            load_row<T, D>(a.soa, a.ld, q, qc);
            if (active && qp < pp && within<D>(pc, qc, eps2)) fmin = min(fmin, ld_agent(&f[q]));
        }
    }
    if (active && fmin < f0) {
        st_agent(&f[p], fmin);
        atomicOr(changed, 1u);
    }
}

// final choice for every ambiguous border row: the cluster of the adjacent
// core with the smallest f
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_ot_assign(const DbscanArgs<T> a,
                                                        const uint8_t* __restrict__ core,
                                                        const int32_t* __restrict__ rootslot,
                                                        const int2* __restrict__ amb, int32_t namb,
                                                        const int32_t* __restrict__ f,
                                                        int64_t* __restrict__ labels) {
    const int32_t i = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= namb) return;
    const int32_t p = amb[i].x, u = amb[i].y;
    const JobParams<T>& jp = a.jobs[a.job_of_cell[u]];
    T pc[D];
    load_row<T, D>(a.soa, a.ld, p, pc);
    int32_t best_f = INT_MAX, best_cluster = -1;
    for (uint32_t r = a.roff[u]; r < a.roff[u + 1]; ++r) {
        const int2 rg = a.ranges[r];
        for (int32_t q = rg.x; q < rg.y; ++q) {
            if (!core[q]) continue;
            T qc[D];
            load_row<T, D>(a.soa, a.ld, q, qc);
            if (within<D>(pc, qc, jp.eps2) && f[q] < best_f) {
                best_f = f[q];
                best_cluster = rootslot[q];
            }
        }
    }
    if (best_cluster >= 0)
        labels[a.inrow[p]] = static_cast<int64_t>(best_cluster) - jp.row_begin + 1;
}

__global__ __launch_bounds__(kThreads) void k_ot_init(const uint8_t* __restrict__ core,
                                                      const uint32_t* __restrict__ pos_of_slot,
                                                      int32_t* __restrict__ f, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) f[i] = core[i] ? slot_pos(pos_of_slot, static_cast<int32_t>(i)) : INT_MAX;
}

// Number of zero labels among labels[0, n): one atomic per block.
__global__ __launch_bounds__(kThreads) void k_count_zero(const int64_t* __restrict__ v, int64_t n,
                                                         unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[kWavesPerBlock];
    unsigned long long c = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kThreads)
        c += v[i] == 0 ? 1ull : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane_id() == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) t += part[w];
        if (t) atomicAdd(out, t);
    }
}

// Sum of n byte flags (0/1): one atomic per block.
__global__ __launch_bounds__(kThreads) void k_count_flags(const uint8_t* __restrict__ f, int64_t n,
                                                          unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[kWavesPerBlock];
    unsigned long long c = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kThreads)
        c += f[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane_id() == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) t += part[w];
        if (t) atomicAdd(out, t);
    }
}

__global__ __launch_bounds__(kThreads) void k_iota(int32_t* __restrict__ v, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) v[i] = static_cast<int32_t>(i);
}

__global__ __launch_bounds__(kThreads) void k_fill2(int32_t* __restrict__ a, int32_t va,
                                                    int32_t* __restrict__ b, int32_t vb, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) {
        a[i] = va;
        b[i] = vb;
    }
}

inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

// Waves for the chunk items of the core and border kernels.  A host-known
// count gets one wave per item.  A count that an earlier kernel wrote (the
// host holds a bound, n + n / 64 for chunks) gets one wave per item up to a
// likely count, n / 8 + n / 64 (cells of 8 rows on average; every wave past
// the count exits at once, and each costs a dispatch), and a small walking
// grid for the items beyond it, which exits at once when there are none.
constexpr int64_t kMaxItemWaves = int64_t{1} << 24;
constexpr int64_t kTailWaves = 16384;
struct ItemGrid {
    int64_t waves = 0;  // one-wave-per-item grid: items [0, waves)
    bool tail = false;  // a walking grid for items [waves, count)
};
inline ItemGrid item_grid(int64_t bound, bool device_count, int64_t n) {
    ItemGrid g;
    const int64_t likely = device_count ? n / 8 + n / 64 + 64 : bound;
    g.waves = std::min(bound, std::min(likely, kMaxItemWaves));
    g.tail = bound > g.waves;
    return g;
}


}  // namespace

// *out += number of nonzero flags among f[0, n)
void count_flags(const uint8_t* f, int64_t n, unsigned long long* out, hipStream_t s) {
    if (n <= 0) return;
    const int64_t want = (n + kThreads - 1) / kThreads;
    const unsigned grid = static_cast<unsigned>(want < 2048 ? want : 2048);
    hipLaunchKernelGGL(k_count_flags, dim3(grid), dim3(kThreads), 0, s, f, n, out);
    HPD_HIP(hipGetLastError());
}

// *out += number of zero labels (noise rows) among labels[0, n)
void count_noise(const int64_t* labels, int64_t n, unsigned long long* out, hipStream_t s) {
    if (n <= 0) return;
    const int64_t want = (n + kThreads - 1) / kThreads;
    const unsigned grid = static_cast<unsigned>(want < 2048 ? want : 2048);
    hipLaunchKernelGGL(k_count_zero, dim3(grid), dim3(kThreads), 0, s, labels, n, out);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_core(const DbscanArgs<T>& a, uint8_t* core, unsigned long long* ntests, hipStream_t s,
                 hipEvent_t count_start) {
    unsigned long long* ncore = nullptr;  // (the other core kernels count nothing)
    if (count_start) HPD_HIP(hipEventRecord(count_start, s));  // (no box pass: all of it)
    if (a.nitems <= 0) return;
    if constexpr (std::is_floating_point<T>::value) {
        if (a.bbox) {
            DbscanArgs<T> ac = a;
            if (!a.sparse_core) {  // the chunks of every cell that can hold cores
                ac.ditems = a.citems ? a.citems : a.items;
                ac.nditems = a.citems ? a.ncitems : a.nitems;
                ac.nditems_dev = a.citems ? a.ncitems_dev : nullptr;
                ac.nsitems = 0;
            }
            if (a.citems && !a.sparse_core && a.citems_need && a.bmincore) {
                // rows of tight boxes with >= minPts rows are core, every other
                // row starts non-core (no row of a cell outside citems is
                // core); only the chunks holding another row are counted
                launch_core_full(a.n, a.box_of_slot, a.btight, a.bstart, a.inrow, core, a.bmincore,
                                 s);
                launch_select_core_chunks(Count{a.ncitems, a.ncitems_dev}, a.citems, a.ustart,
                                          a.box_of_slot, a.btight, a.citems_need,
                                          a.ncitems_need_dev, s);
                ac.ditems = a.citems_need;
                ac.nditems_dev = a.ncitems_need_dev;
            } else if (a.citems && !a.sparse_core) {  // rows of the other cells are not core
                HPD_HIP(hipMemsetAsync(core, 0, static_cast<size_t>(a.n), s));
            }
            const ItemGrid g = item_grid(ac.nditems, ac.nditems_dev != nullptr, a.n);
            if (count_start) HPD_HIP(hipEventRecord(count_start, s));
            if (g.waves > 0) {
                if (ntests)
                    hipLaunchKernelGGL((k_core_boxes<T, D, true>),
                                       dim3(blocks_for(g.waves, kWavesPerBlock)), dim3(kThreads), 0,
                                       s, ac, core, ntests);
                else
                    hipLaunchKernelGGL((k_core_boxes<T, D>),
                                       dim3(blocks_for(g.waves, kWavesPerBlock)), dim3(kThreads), 0,
                                       s, ac, core, nullptr);
            }
            if (g.tail) {
                if (ntests)
                    hipLaunchKernelGGL((k_core_boxes_walk<T, D, true>),
                                       dim3(blocks_for(kTailWaves, kWavesPerBlock)), dim3(kThreads),
                                       0, s, ac, core, ntests, static_cast<int32_t>(g.waves));
                else
                    hipLaunchKernelGGL((k_core_boxes_walk<T, D>),
                                       dim3(blocks_for(kTailWaves, kWavesPerBlock)), dim3(kThreads),
                                       0, s, ac, core, nullptr, static_cast<int32_t>(g.waves));
            }
            if (ac.nsitems > 0)
                hipLaunchKernelGGL((k_core_sparse<T, D>), dim3(blocks_for(ac.nsitems, kThreads)),
                                   dim3(kThreads), 0, s, ac, core, ncore);
            return;
        }
    }
    hipLaunchKernelGGL((k_core<T, D>), dim3(blocks_for(a.nitems, kWavesPerBlock)), dim3(kThreads),
                       0, s, a, core, ncore);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D, typename K>
void launch_label_full_rows(const T* pts, const DbscanArgs<T>& a, int sub_k, int sub_bits,
                            const K* keys, Count nbox, const int32_t* bcluster, int32_t* tab,
                            int64_t cap, int64_t* labels, unsigned int* ok, hipStream_t s) {
    if (a.n <= 0 || nbox.host <= 0) return;
    hipLaunchKernelGGL((k_full_box_table<T, K>), dim3(blocks_for(nbox.host, kThreads)),
                       dim3(kThreads), 0, s, keys, a.n, a.bstart, nbox, a.bcell, a.job_of_cell,
                       a.jobs, a.btight, bcluster, tab, cap, ok);
    hipLaunchKernelGGL((k_label_full_rows<T, D, K>), dim3(blocks_for(a.n, kThreads)),
                       dim3(kThreads), 0, s, pts, a.n,
                       a.jobs, a.njobs, sub_k, sub_bits, static_cast<const int32_t*>(tab), cap,
                       static_cast<const unsigned int*>(ok), labels);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_union_points(const DbscanArgs<T>& a, const uint8_t* core, int32_t* parent,
                         int32_t* rsnap, int32_t* cmin, int32_t* cmax, int32_t* compmin,
                         hipStream_t s) {
    if (a.n <= 0) return;
    const unsigned gw = blocks_for(a.nitems, kWavesPerBlock);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(a.n, kThreads)), dim3(kThreads), 0, s, parent, a.n);
    hipLaunchKernelGGL((k_hook<T, D>), dim3(gw), dim3(kThreads), 0, s, a, core, parent);
    hipLaunchKernelGGL(k_fill2, dim3(blocks_for(a.m, kThreads)), dim3(kThreads), 0, s, cmin, INT_MAX,
                       cmax, -1, static_cast<int64_t>(a.m));
    hipLaunchKernelGGL(k_roots, dim3(gw), dim3(kThreads), 0, s, a.items, a.nitems, a.ustart, a.roff,
                       core, parent, rsnap, cmin, cmax);
    hipLaunchKernelGGL((k_union<T, D>), dim3(gw), dim3(kThreads), 0, s, a, core,
                       static_cast<const int32_t*>(rsnap), static_cast<const int32_t*>(cmin),
                       static_cast<const int32_t*>(cmax), parent);
    hipLaunchKernelGGL(k_fill2, dim3(blocks_for(a.n, kThreads)), dim3(kThreads), 0, s, compmin,
                       INT_MAX, rsnap, 0, a.n);
    hipLaunchKernelGGL(k_compmin, dim3(blocks_for(a.n, kThreads)), dim3(kThreads), 0, s, core,
                       parent, a.n, a.pos_of_slot, compmin);
    HPD_HIP(hipGetLastError());
}

template <typename T>
void launch_label_cores(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* bcluster,
                        int32_t* parent, const int32_t* compmin, int32_t* rootslot, int32_t* cmin,
                        int32_t* cmax, int64_t* labels, unsigned long long* ncore, hipStream_t s,
                        const unsigned int* full_ok) {
    if (a.n <= 0) return;
    if (bcluster && a.bcell) {
        const unsigned grid = std::min<unsigned>(blocks_for(a.n, kThreads), kLabelBlocks);
        hipLaunchKernelGGL(k_label_cores_boxes<T>, dim3(grid), dim3(kThreads), 0, s, a.n, a.jobs,
                           a.job_of_cell, core, a.box_of_slot, a.bcell, bcluster, a.inrow, labels,
                           rootslot, ncore, a.njobs == 1, full_ok ? a.btight : nullptr, full_ok);
        HPD_HIP(hipGetLastError());
        return;
    }
    const unsigned gw = blocks_for(a.nitems, kWavesPerBlock);
    count_flags(core, a.n, ncore, s);
    hipLaunchKernelGGL(k_fill2, dim3(blocks_for(a.m, kThreads)), dim3(kThreads), 0, s, cmin, INT_MAX,
                       cmax, -1, static_cast<int64_t>(a.m));
    hipLaunchKernelGGL(k_label_cores<T>, dim3(gw), dim3(kThreads), 0, s, a.items, a.nitems,
                       a.ustart, a.roff, a.jobs, a.job_of_cell, core, a.box_of_slot, bcluster,
                       parent, compmin, a.inrow, labels, rootslot, cmin, cmax);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_border(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* bcluster,
                   const int32_t* rootslot, const int32_t* cmin, const int32_t* cmax,
                   int64_t* labels, unsigned long long* stats, int2* amb_list,
                   unsigned long long* noise_stripes, hipStream_t s) {
    if (a.n <= 0) return;
    const unsigned gw = blocks_for(a.nitems, kWavesPerBlock);
    if constexpr (std::is_floating_point<T>::value) {
        if (bcluster && a.bbox) {
            const ItemGrid g = item_grid(a.nditems, a.nditems_dev != nullptr, a.n);
            if (g.waves > 0)
                hipLaunchKernelGGL((k_border_boxes<T, D>), dim3(blocks_for(g.waves, kWavesPerBlock)),
                                   dim3(kThreads), 0, s, a, core, bcluster, a.inrow, labels, stats,
                                   amb_list, noise_stripes);
            if (g.tail)
                hipLaunchKernelGGL((k_border_boxes_walk<T, D>),
                                   dim3(blocks_for(kTailWaves, kWavesPerBlock)), dim3(kThreads), 0,
                                   s, a, core, bcluster, a.inrow, labels, stats, amb_list,
                                   noise_stripes, static_cast<int32_t>(g.waves));
            if (a.nsitems > 0)
                hipLaunchKernelGGL((k_border_sparse<T, D>), dim3(blocks_for(a.nsitems, kThreads)),
                                   dim3(kThreads), 0, s, a, core, bcluster, labels, stats,
                                   amb_list, noise_stripes);
            if (noise_stripes)
                hipLaunchKernelGGL(k_sum_stripes, dim3(1), dim3(kThreads), 0, s,
                                   static_cast<const unsigned long long*>(noise_stripes), &stats[0]);
            else
                count_noise(labels, a.n, &stats[0], s);
            HPD_HIP(hipGetLastError());
            return;
        }
    }
    hipLaunchKernelGGL((k_border<T, D>), dim3(gw), dim3(kThreads), 0, s, a, core, rootslot, cmin,
                       cmax, a.inrow, labels, stats, amb_list);
    HPD_HIP(hipGetLastError());
    count_noise(labels, a.n, &stats[0], s);
}

template <typename T, int D>
void launch_one_thread_mark(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* rootslot,
                            const int2* amb, int32_t namb, uint8_t* involved, int32_t* f,
                            hipStream_t s) {
    hipLaunchKernelGGL((k_ot_mark<T, D>), dim3(blocks_for(namb, kThreads)), dim3(kThreads), 0, s,
                       a, core, rootslot, amb, namb, involved);
    hipLaunchKernelGGL(k_ot_init, dim3(blocks_for(a.n, kThreads)), dim3(kThreads), 0, s, core,
                       a.pos_of_slot, f, a.n);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_one_thread_relax(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* rootslot,
                             const uint8_t* involved, int32_t* f, unsigned int* changed,
                             hipStream_t s) {
    hipLaunchKernelGGL((k_ot_relax<T, D>), dim3(blocks_for(a.nitems, kWavesPerBlock)),
                       dim3(kThreads), 0, s, a, core, rootslot, involved, f, changed);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_one_thread_assign(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* rootslot,
                              const int2* amb, int32_t namb, const int32_t* f, int64_t* labels,
                              hipStream_t s) {
    hipLaunchKernelGGL((k_ot_assign<T, D>), dim3(blocks_for(namb, kThreads)), dim3(kThreads), 0, s,
                       a, core, rootslot, amb, namb, f, labels);
    HPD_HIP(hipGetLastError());
}

#define HPD_INST_DBSCAN(T, D)                                                                     \
    template void launch_core<T, D>(const DbscanArgs<T>&, uint8_t*, unsigned long long*,          \
                                    hipStream_t, hipEvent_t);                                     \
    template void launch_union_points<T, D>(const DbscanArgs<T>&, const uint8_t*, int32_t*,       \
                                            int32_t*, int32_t*, int32_t*, int32_t*, hipStream_t); \
    template void launch_border<T, D>(const DbscanArgs<T>&, const uint8_t*, const int32_t*,       \
                                      const int32_t*, const int32_t*, const int32_t*, int64_t*,   \
                                      unsigned long long*, int2*, unsigned long long*,            \
                                      hipStream_t);                                               \
    template void launch_one_thread_mark<T, D>(const DbscanArgs<T>&, const uint8_t*,              \
                                               const int32_t*, const int2*, int32_t, uint8_t*,    \
                                               int32_t*, hipStream_t);                            \
    template void launch_one_thread_relax<T, D>(const DbscanArgs<T>&, const uint8_t*,             \
                                                const int32_t*, const uint8_t*, int32_t*,         \
                                                unsigned int*, hipStream_t);                      \
    template void launch_one_thread_assign<T, D>(const DbscanArgs<T>&, const uint8_t*,            \
                                                 const int32_t*, const int2*, int32_t,            \
                                                 const int32_t*, int64_t*, hipStream_t);
#define HPD_INST_DBSCAN_ALL_D(T) \
    HPD_INST_DBSCAN(T, 1)        \
    HPD_INST_DBSCAN(T, 2)        \
    HPD_INST_DBSCAN(T, 3)        \
    HPD_INST_DBSCAN(T, 4)        \
    HPD_INST_DBSCAN(T, 5)        \
    HPD_INST_DBSCAN(T, 6)        \
    HPD_INST_DBSCAN(T, 7)        \
    HPD_INST_DBSCAN(T, 8)

HPD_FOR_EACH_DTYPE(HPD_INST_DBSCAN_ALL_D)

#define HPD_INST_LABEL_CORES(T)                                                                \
    template void launch_label_cores<T>(const DbscanArgs<T>&, const uint8_t*, const int32_t*,  \
                                        int32_t*, const int32_t*, int32_t*, int32_t*, int32_t*, \
                                        int64_t*, unsigned long long*, hipStream_t,      \
                                        const unsigned int*);
HPD_FOR_EACH_DTYPE(HPD_INST_LABEL_CORES)

#define HPD_INST_FULL_ROWS(T, D, K)                                                             \
    template void launch_label_full_rows<T, D, K>(const T*, const DbscanArgs<T>&, int, int,     \
                                                  const K*, Count, const int32_t*, int32_t*,     \
                                                  int64_t, int64_t*, unsigned int*, hipStream_t);
#define HPD_INST_FULL_ROWS_ALL(T)          \
    HPD_INST_FULL_ROWS(T, 1, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 2, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 3, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 4, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 5, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 6, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 7, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 8, uint32_t)     \
    HPD_INST_FULL_ROWS(T, 1, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 2, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 3, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 4, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 5, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 6, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 7, uint64_t)     \
    HPD_INST_FULL_ROWS(T, 8, uint64_t)
HPD_INST_FULL_ROWS_ALL(float)
HPD_INST_FULL_ROWS_ALL(double)


}  // namespace hpd
