// Sub-boxes of the eps-grid cells and the box-level union of core points.
//
// Every occupied cell is split into k^D sub-boxes along the same scaled
// coordinate the cell index uses, v = (x - min) * (1/eps):
//   sub_d = clamp(floor(v_d * k) - k * floor(v_d), 0, k - 1)
// with k = 2 for D <= 3 and k = 3 for D <= 8, so a box is at most eps/k wide
// and its diagonal at most 0.94 eps.  Rows are regrouped by (cell, box)
// inside each cell; the reference's sorted position of every row is kept
// beside it (the cluster ids are smallest positions, not slots).
//
// A box is "tight" when the reference's own squared-distance expression
// evaluated on the box extents, dist2(max - min), is <= EPS2.  Every
// rounding step of that expression (subtract, square, the summation tree,
// the flush of tiny values) is monotone in its inputs and |p_d - q_d| <=
// max_d - min_d for any two members, so every pair of members of a tight
// box passes the reference's within-eps test, and they share a cell, so they
// are each other's candidates.  Hence:
//   * a tight box with >= minPts rows makes all its rows core;
//   * all cores of a tight box lie in one cluster.
// The union of the cores is therefore computed on boxes: two boxes with
// cores are joined when some core pair across them is within eps (tested
// with the reference's expression, stopping at the first hit), unless they
// are already joined; a core row farther than eps from the other box's
// bounding box takes no part in the test.  Box pairs whose bounding boxes are farther apart than
// eps * (1 + 1e-5) cannot hold such a pair (the float result of the
// expression is within a few ulps of the exact squared distance) and are
// skipped.
#include <algorithm>
#include <climits>

#include "boxes.hpp"
#include "hpdbscan_capi.h"
#include "common.hpp"
#include "grid.hpp"
#include "index.hpp"
#include "internal.hpp"
#include "visit.hpp"

namespace hpd {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
// boxes of at most kSmallBox rows are hooked one thread per box, with at
// most kSmallTests distance tests per box pair
constexpr int kSmallBox = 32;
constexpr int kSmallTests = 512;


inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

// Rows in slot order: coordinates (SoA); with it the bounding box of every
// box (ordered keys).  A block of kThreads slots reduces its boxes in LDS
// (box ids are consecutive in slot order, so a box's LDS entry is its id less
// the block's first box): a box inside the block is stored by the block; the
// block's first box when it started in an earlier block, and its last box
// when it goes on past the block, leave the block's share in
// part[(2 * block + 0 / 1) * 2D ..] for k_box_span.  No global atomics: a
// box of a million rows would otherwise take thousands of them on one
// address.  With pts null the coordinates are already in soa (the sort
// moved them) and only the bounding boxes are made.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_gather_boxes(
    const T* __restrict__ pts, int64_t n, const uint32_t* __restrict__ inrow_of_slot,
    T* __restrict__ soa, int64_t ld, const int32_t* __restrict__ box_of_slot,
    const int32_t* __restrict__ bstart, OrdKey<T>* __restrict__ bbox, OrdKey<T>* __restrict__ part) {
    using U = OrdKey<T>;
    __shared__ U slo[kThreads][D];
    __shared__ U shi[kThreads][D];
    const int64_t blk0 = static_cast<int64_t>(blockIdx.x) * kThreads;
    const int64_t s = blk0 + threadIdx.x;
    const bool valid = s < n;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        slo[threadIdx.x][j] = static_cast<U>(~U(0));
        shi[threadIdx.x][j] = U(0);
    }
    const int32_t b_first = box_of_slot[blk0];  // the block's first box
    int32_t b = -1 - lane_id();  // distinct keys for lanes past the end
    T x[D];
    if (valid && !pts) {
        b = box_of_slot[s];
#pragma unroll
        for (int j = 0; j < D; ++j) x[j] = soa[static_cast<int64_t>(j) * ld + s];
    } else if (valid) {
        const uint32_t row = inrow_of_slot[s];
        b = box_of_slot[s];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            x[j] = pts[static_cast<int64_t>(row) * D + j];
            soa[static_cast<int64_t>(j) * ld + s] = x[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < D; ++j) x[j] = T(0);
    }
    __syncthreads();
    // every lane takes part in the shuffles; the head lane of each run in the
    // wave (lane 0 included) adds the run's extremes to the box's LDS entry
    const bool head = valid && run_head(b);
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const U k = ord_key(x[j]);
        const U lo = run_suffix_min<U>(k, b);
        const U hi = run_suffix_max<U>(k, b);
        if (head) {
            atomicMin(&slo[b - b_first][j], lo);
            atomicMax(&shi[b - b_first][j], hi);
        }
    }
    __syncthreads();
    const int64_t blk1 = blk0 + kThreads < n ? blk0 + kThreads : n;  // end of the block's slots
    const int32_t nruns = box_of_slot[blk1 - 1] - b_first + 1;
    const int t = threadIdx.x;
    if (t < nruns) {
        const int32_t bb = b_first + t;
        const bool starts = bstart[bb] >= blk0;
        const bool ends = bstart[bb + 1] <= blk1;
        U* dst = starts && ends ? &bbox[static_cast<int64_t>(bb) * D * 2]
                                : &part[(2 * static_cast<int64_t>(blockIdx.x) + (starts ? 1 : 0)) * D * 2];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            dst[2 * j] = slo[t][j];
            dst[2 * j + 1] = shi[t][j];
        }
    }
}

// Bounding boxes only, the coordinates being in slot order already (the
// sort carried them): k_gather_boxes' reduction over tiles of kBoundsRows
// wave-segments per thread (a tile of kThreads * kBoundsRows slots per
// block, all of a thread's slots loaded first), so that the LDS set-up,
// the barriers and the stores of a block are shared by four times the
// slots.  Rows of at most 8 bytes (LDS: 2 * tile * D ordered keys).
constexpr int kBoundsRows = 4;
constexpr int kBoundsTile = kThreads * kBoundsRows;

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_bounds(
    const T* __restrict__ soa, int64_t ld, int64_t n, const int32_t* __restrict__ box_of_slot,
    const int32_t* __restrict__ bstart, OrdKey<T>* __restrict__ bbox, OrdKey<T>* __restrict__ part) {
    using U = OrdKey<T>;
    static_assert(D * sizeof(T) <= 8, "rows of at most 8 bytes");
    __shared__ U slo[kBoundsTile][D];
    __shared__ U shi[kBoundsTile][D];
    const int64_t blk0 = static_cast<int64_t>(blockIdx.x) * kBoundsTile;
    for (int e = threadIdx.x; e < kBoundsTile; e += kThreads) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            slo[e][j] = static_cast<U>(~U(0));
            shi[e][j] = U(0);
        }
    }
    const int32_t b_first = box_of_slot[blk0];  // the block's first box
    const int w = static_cast<int>(threadIdx.x >> 6);
    // segment k of wave w: slots blk0 + (k * kWavesPerBlock + w) * 64 + lane
    int32_t b[kBoundsRows];
    T x[kBoundsRows][D];
#pragma unroll
    for (int k = 0; k < kBoundsRows; ++k) {
        const int64_t s = blk0 + (static_cast<int64_t>(k) * kWavesPerBlock + w) * kWave + lane_id();
        const bool valid = s < n;
        b[k] = valid ? box_of_slot[s] : -1 - static_cast<int32_t>(lane_id());
#pragma unroll
        for (int j = 0; j < D; ++j) x[k][j] = valid ? soa[static_cast<int64_t>(j) * ld + s] : T(0);
    }
    const int64_t blk1 = blk0 + kBoundsTile < n ? blk0 + kBoundsTile : n;
    __syncthreads();
    if (b_first == box_of_slot[blk1 - 1]) {
        // the whole tile is one box (large boxes): each thread's slots,
        // then one reduction per wave, then one LDS merge per wave
        U lo[D], hi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            lo[j] = static_cast<U>(~U(0));
            hi[j] = U(0);
        }
#pragma unroll
        for (int k = 0; k < kBoundsRows; ++k) {
            if (b[k] < 0) continue;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const U key = ord_key(x[k][j]);
                lo[j] = key < lo[j] ? key : lo[j];
                hi[j] = key > hi[j] ? key : hi[j];
            }
        }
#pragma unroll
        for (int j = 0; j < D; ++j) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const U a = static_cast<U>(__shfl_xor(lo[j], o, 64));
                const U z = static_cast<U>(__shfl_xor(hi[j], o, 64));
                lo[j] = a < lo[j] ? a : lo[j];
                hi[j] = z > hi[j] ? z : hi[j];
            }
            if (lane_id() == 0) {
                atomicMin(&slo[0][j], lo[j]);
                atomicMax(&shi[0][j], hi[j]);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kBoundsRows; ++k) {
            // the head lane of each run of the segment adds the run's extremes
            const bool head = b[k] >= 0 && run_head(b[k]);
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const U key = ord_key(x[k][j]);
                const U lo = run_suffix_min<U>(key, b[k]);
                const U hi = run_suffix_max<U>(key, b[k]);
                if (head) {
                    atomicMin(&slo[b[k] - b_first][j], lo);
                    atomicMax(&shi[b[k] - b_first][j], hi);
                }
            }
        }
    }
    __syncthreads();
    const int32_t nruns = box_of_slot[blk1 - 1] - b_first + 1;
    for (int t = threadIdx.x; t < nruns; t += kThreads) {
        const int32_t bb = b_first + t;
        const bool starts = bstart[bb] >= blk0;
        const bool ends = bstart[bb + 1] <= blk1;
        U* dst = starts && ends ? &bbox[static_cast<int64_t>(bb) * D * 2]
                                : &part[(2 * static_cast<int64_t>(blockIdx.x) + (starts ? 1 : 0)) * D * 2];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            dst[2 * j] = slo[t][j];
            dst[2 * j + 1] = shi[t][j];
        }
    }
}

// Bounding boxes of the boxes that span blocks of k_gather_boxes: one wave
// per block; the wave of the block where such a box starts combines the
// block's share (its last run) with the shares of the following blocks (their
// first runs) up to the box's last block.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_span(int64_t n, int64_t nblocks, int64_t tile,
                                                       const int32_t* __restrict__ box_of_slot,
                                                       const int32_t* __restrict__ bstart,
                                                       const OrdKey<T>* __restrict__ part,
                                                       OrdKey<T>* __restrict__ bbox) {
    using U = OrdKey<T>;
    const int64_t blk = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
    if (blk >= nblocks) return;
    const int64_t blk0 = blk * tile;
    const int64_t blk1 = blk0 + tile < n ? blk0 + tile : n;
    const int32_t b = box_of_slot[blk1 - 1];
    const int64_t b0 = bstart[b], b1 = bstart[b + 1];
    if (b0 < blk0 || b1 <= blk1) return;  // not a box that starts here and goes on
    const int64_t last = (b1 - 1) / tile;
    U lo[D], hi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        lo[j] = static_cast<U>(~U(0));
        hi[j] = U(0);
    }
    for (int64_t k = blk + lane_id(); k <= last; k += kWave) {
        const U* src = &part[(2 * k + (k == blk ? 1 : 0)) * D * 2];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            lo[j] = min(lo[j], src[2 * j]);
            hi[j] = max(hi[j], src[2 * j + 1]);
        }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo[j] = min(lo[j], static_cast<U>(__shfl_xor(lo[j], o, 64)));
            hi[j] = max(hi[j], static_cast<U>(__shfl_xor(hi[j], o, 64)));
        }
    }
    if (lane_id() == 0) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            bbox[(static_cast<int64_t>(b) * D + j) * 2] = lo[j];
            bbox[(static_cast<int64_t>(b) * D + j) * 2 + 1] = hi[j];
        }
    }
}

template <typename K>
__global__ __launch_bounds__(kThreads) void k_box_flags(const K* __restrict__ keys, int64_t n,
                                                        uint32_t* __restrict__ flags) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

template <typename K>
__global__ __launch_bounds__(kThreads) void k_box_table(const K* __restrict__ keys, int64_t n,
                                                        const uint32_t* __restrict__ flags,
                                                        const uint32_t* __restrict__ scan,
                                                        const uint32_t* __restrict__ cflags,
                                                        const uint32_t* __restrict__ cscan,
                                                        int32_t* __restrict__ bstart,
                                                        int32_t* __restrict__ bcell,
                                                        int32_t* __restrict__ box_of_slot,
                                                        const uint32_t* __restrict__ nbox_dev,
                                                        void* __restrict__ bbox, int bbox_keys,
                                                        int key_bytes, int shift,
                                                        uint32_t* __restrict__ cbflag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) {
        const uint32_t b = scan[i] + flags[i] - 1u;
        box_of_slot[i] = static_cast<int32_t>(b);
        if (flags[i]) {
            bstart[b] = static_cast<int32_t>(i);
            if (cflags)
                bcell[b] = static_cast<int32_t>(cscan[i] + cflags[i] - 1u);  // occupied-cell index
            else  // the box starts a cell: the cells come from the boxes
                cbflag[b] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1u : 0u;
            // empty bounding box: min keys all ones, max keys zero
            if (bbox) {
                const int64_t k0 = static_cast<int64_t>(b) * bbox_keys;
                for (int k = 0; k < bbox_keys; ++k) {
                    if (key_bytes == 8)
                        static_cast<unsigned long long*>(bbox)[k0 + k] = (k & 1) ? 0ull : ~0ull;
                    else
                        static_cast<uint32_t*>(bbox)[k0 + k] = (k & 1) ? 0u : ~0u;
                }
            }
        }
    }
    if (i == 0) bstart[*nbox_dev] = static_cast<int32_t>(n);
}

// The occupied cells from the boxes (a cell's first slot starts a box):
// cell u = the u-th box with cbflag set; ucell / ustart as the cell runs
// give them (ustart[m] = n), bcell per box.  The boxes walk a bounded grid.
template <typename K>
__global__ __launch_bounds__(kThreads) void k_cells_of_boxes(
    const K* __restrict__ keys, int64_t n, int shift, const int32_t* __restrict__ bstart,
    Count nbox_c, const uint32_t* __restrict__ cbflag, const uint32_t* __restrict__ cbscan,
    int32_t* __restrict__ bcell, K* __restrict__ ucell, int32_t* __restrict__ ustart,
    const uint32_t* __restrict__ ncells_dev) {
    const int64_t nbox = nbox_c.get();
    for (int64_t b = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; b < nbox;
         b += static_cast<int64_t>(gridDim.x) * kThreads) {
        const uint32_t f = cbflag[b], u = cbscan[b];
        bcell[b] = static_cast<int32_t>(u + f - 1u);
        if (f) {
            const int32_t s0 = bstart[b];
            ucell[u] = keys[s0] >> shift;
            ustart[u] = s0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ustart[*ncells_dev] = static_cast<int32_t>(n);
}

// first box of every cell; cbox[m] = number of boxes
__global__ __launch_bounds__(kThreads) void k_cell_boxes(const int32_t* __restrict__ bcell,
                                                         Count nbox_c, Count m_c,
                                                         int32_t* __restrict__ cbox) {
    const int32_t m = static_cast<int32_t>(m_c.get());
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b < nbox && (b == 0 || bcell[b] != bcell[b - 1])) cbox[bcell[b]] = b;
    if (b == 0) cbox[m] = nbox;
}

template <typename U>
__device__ __forceinline__ U wave_min_u(U v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const U w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
template <typename U>
__device__ __forceinline__ U wave_max_u(U v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const U w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// tight flag per box: dist2(extent) <= EPS2 in the reference's order
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_tight(const OrdKey<T>* __restrict__ bbox,
                                                        Count nbox_c,
                                                        const int32_t* __restrict__ bstart,
                                                        const int32_t* __restrict__ bcell,
                                                        const int32_t* __restrict__ job_of_cell,
                                                        const JobParams<T>* __restrict__ jobs,
                                                        uint8_t* __restrict__ btight,
                                                        unsigned int* __restrict__ nloose) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= nbox) return;
    const JobParams<T>& jp = jobs[job_of_cell[bcell[b]]];
    T lo[D], hi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        lo[j] = ord_val(bbox[(static_cast<int64_t>(b) * D + j) * 2]);
        hi[j] = ord_val(bbox[(static_cast<int64_t>(b) * D + j) * 2 + 1]);
    }
    const bool tight = within<D>(hi, lo, jp.eps2);
    const int32_t size = bstart[b + 1] - bstart[b];
    // bit 0: tight; bit 1: tight and at least minPts rows (all rows core)
    btight[b] = static_cast<uint8_t>((tight ? 1 : 0) |
                                     ((tight && static_cast<uint32_t>(size) >= jp.min_pts) ? 2 : 0));
    if (!tight && size > 1) atomicAdd(nloose, 1u);
}

// Box-level hook: every box with cores points at some box of a smaller
// index holding a core within eps of one of its cores (searched from the
// box's first 64-row chunk, stopping at the first hit; any such box will
// do, and a box without one stays a root).  parent
// stays <= the box index, so the hooks form a forest whose trees are
// connected sets of boxes.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_hook(
    const T* __restrict__ soa, int64_t ld, const int32_t* __restrict__ blist, Count nitems_c,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcell,
    const int32_t* __restrict__ cbox, const uint32_t* __restrict__ roff,
    const int2* __restrict__ rcells, const OrdKey<T>* __restrict__ bbox,
    const int32_t* __restrict__ bmincore, const uint8_t* __restrict__ core,
    const int32_t* __restrict__ job_of_cell, const JobParams<T>* __restrict__ jobs,
    const int32_t* __restrict__ rcenter, int32_t* __restrict__ bparent) {
    const int32_t nitems = static_cast<int32_t>(nitems_c.get());
    const int32_t stride = static_cast<int32_t>(gridDim.x) * kWavesPerBlock;
    // waves walk the items wid, wid + stride, ... (one each when the grid
    // covers them)
    for (int32_t wid = uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock +
                               static_cast<int32_t>(threadIdx.x >> 6));
         wid < nitems; wid += stride) [&] {
        // boxes with cores of more than kSmallBox rows; one hook per box is
        // enough: its first chunk looks for it (a missed link is found by
        // the exact union later)
        const int32_t b = uniform(blist[wid]);
        const int2 it = make_int2(b, uniform(bstart[b]));
        const int32_t end = uniform(bstart[b + 1]);
        const int32_t p = it.y + lane_id();
        const bool pcore = p < end && core[p];
        if (!__any(pcore)) return;
        const int32_t u = uniform(bcell[b]);
        const JobParams<T>& jp = jobs[uniform(job_of_cell[u])];
        const T eps2 = jp.eps2;
        const double gap2_max = jp.gap2_max;
        const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
        T pc[D];
        load_row<T, D>(soa, ld, pcore ? p : it.y, pc);
        const int32_t r0 = uniform(static_cast<int32_t>(roff[u]));
        const int32_t r1 = uniform(static_cast<int32_t>(roff[u + 1]));
        // the cell's own row first, from the own cell's first box: the box's
        // nearest neighbours are there, so the first hit comes soonest; boxes
        // before b only
        for_each_row_box<T, D>(
            r0, r1, uniform(rcenter[u]), rcells, cbox, bstart, bbox, bmincore,
            [&](int32_t b2, const double* lo, const double* hi, int32_t q0, int32_t q1,
                int32_t mincore) {
                if (mincore == INT_MAX) return false;  // no cores
                double gmin2, gmax2;
                point_box_lh<T, D>(lo, hi, pc, gmin2, gmax2);
                const bool act = pcore && gmin2 <= gap2_max;
                if (!__any(act)) return false;
                // a core row within eps of the whole box is within eps of its cores
                bool hit = act && gmax2 <= full2;
                if (!__any(hit)) {
                    auto body = [&](int32_t, const T* qc, uint8_t cq, int32_t) {
                        if (cq && act && within<D>(pc, qc, eps2)) hit = true;
                    };
                    auto stop = [&] { return static_cast<bool>(__any(hit)); };
                    if (!visit_range<T, D>(soa, ld, core, nullptr, q0, q1, body, stop)) return false;
                }
                if (lane_id() == 0) bparent[b] = b2;
                return true;
            },
            b, uniform(cbox[u]));
    }();
}

// Root join: every root of the hook forest (a box without an earlier
// within-eps core pair) looks through its whole neighbourhood for the first
// box of another tree holding a within-eps core pair and joins the two
// trees (union-find, smaller root on top).  One round merges most trees of
// a cluster, so far fewer pairs stay open for the union.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_root_join(
    const T* __restrict__ soa, int64_t ld, const int32_t* __restrict__ blist, Count nitems_c,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcell,
    const int32_t* __restrict__ cbox, const uint32_t* __restrict__ roff,
    const int2* __restrict__ rcells, const OrdKey<T>* __restrict__ bbox,
    const int32_t* __restrict__ rsnap, const uint8_t* __restrict__ core,
    const int32_t* __restrict__ job_of_cell, const JobParams<T>* __restrict__ jobs,
    const int32_t* __restrict__ rcenter, int32_t* bparent) {
    const int32_t nitems = static_cast<int32_t>(nitems_c.get());
    const int32_t stride = static_cast<int32_t>(gridDim.x) * kWavesPerBlock;
    // waves walk the items wid, wid + stride, ... (one each when the grid
    // covers them)
    for (int32_t wid = uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock +
                               static_cast<int32_t>(threadIdx.x >> 6));
         wid < nitems; wid += stride) [&] {
        // roots of the forest, each searched from its first chunk
        const int32_t b = uniform(blist[wid]);
        const int2 it = make_int2(b, uniform(bstart[b]));
        const int32_t end = uniform(bstart[b + 1]);
        const int32_t p = it.y + lane_id();
        const bool pcore = p < end && core[p];
        if (!__any(pcore)) return;
        const int32_t u = uniform(bcell[b]);
        const JobParams<T>& jp = jobs[uniform(job_of_cell[u])];
        const T eps2 = jp.eps2;
        const double gap2_max = jp.gap2_max;
        T pc[D];
        load_row<T, D>(soa, ld, pcore ? p : it.y, pc);
        const int32_t r0 = uniform(static_cast<int32_t>(roff[u]));
        const int32_t r1 = uniform(static_cast<int32_t>(roff[u + 1]));
        for_each_row_box<T, D>(
            r0, r1, uniform(rcenter[u]), rcells, cbox, bstart, bbox, rsnap,
            [&](int32_t b2, const double* lo, const double* hi, int32_t q0, int32_t q1,
                int32_t s2) {
                if (s2 < 0 || s2 == b) return false;  // no cores, or b's own tree
                double gmin2, gmax2;
                point_box_lh<T, D>(lo, hi, pc, gmin2, gmax2);
                const bool act = pcore && gmin2 <= gap2_max;
                if (!__any(act)) return false;
                bool hit = false;
                auto body = [&](int32_t, const T* qc, uint8_t cq, int32_t) {
                    if (cq && act && within<D>(pc, qc, eps2)) hit = true;
                };
                auto stop = [&] { return static_cast<bool>(__any(hit)); };
                if (!visit_range<T, D>(soa, ld, core, nullptr, q0, q1, body, stop)) return false;
                if (lane_id() == 0) uf_unite(bparent, b, b2);
                return true;
            });
    }();
}

// Box pairs the hook forest leaves open: one thread per box b finds every
// later box b2 with cores in a different hook tree whose bounding box is
// within eps of b's.  Count pass (list == nullptr): cnt[b] = their number;
// fill pass: the b2 go to list[off[b] ..].
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_scan(
    Count nbox_c, const int32_t* __restrict__ bcell, const int32_t* __restrict__ cbox,
    const uint32_t* __restrict__ roff, const int2* __restrict__ rcells,
    const OrdKey<T>* __restrict__ bbox, const int32_t* __restrict__ job_of_cell,
    const JobParams<T>* __restrict__ jobs, const int32_t* __restrict__ rsnap,
    uint32_t* __restrict__ cnt, int32_t* __restrict__ list, unsigned long long* __restrict__ total,
    const unsigned long long* __restrict__ total_chk, unsigned long long cap) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    // fill pass of the static pipeline: nothing to write when the pairs
    // exceed the list (the host path then takes over)
    if (list && total_chk && *total_chk > cap) return;
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    uint32_t k = (list && b < nbox) ? cnt[b] : 0u;
    // the fill pass walks only the boxes the count pass found pairs for
    const int32_t sb = b < nbox && (!list || cnt[b + 1] > k) ? rsnap[b] : -1;
    if (sb >= 0) {
        const int32_t u = bcell[b];
        const double gap2_max = jobs[job_of_cell[u]].gap2_max;
        double lo[D], hi[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            lo[j] = box_lo<T, D>(bbox, b, j);
            hi[j] = box_hi<T, D>(bbox, b, j);
        }
        for (uint32_t r = roff[u]; r < roff[u + 1]; ++r) {
            const int2 cs = rcells[r];
            const int32_t e1 = cbox[cs.y];
            for (int32_t b2 = max(cbox[cs.x], b + 1); b2 < e1; ++b2) {
                const int32_t s2 = rsnap[b2];
                if (s2 < 0 || s2 == sb) continue;
                double gap2 = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    const double g = fmax(0.0, fmax(box_lo<T, D>(bbox, b2, j) - hi[j],
                                                    lo[j] - box_hi<T, D>(bbox, b2, j)));
                    gap2 += g * g;
                }
                if (gap2 > gap2_max) continue;
                if (list) list[k] = b2;
                ++k;
            }
        }
    }
    if (!list) {
        if (b < nbox) cnt[b] = k;
        // 64-bit total, so that the caller can check it before a 32-bit scan
        block_add<kThreads>(k, total);
    }
}

// Union over the open pairs: one wave per 64-slot chunk of a box b; the
// chunk's core rows (lanes) meet the boxes of b's open list one by one.  A
// pair is skipped when the two boxes are joined by now (live roots) or no
// core of the chunk is within eps of the other box's bounding box;
// otherwise the other box's cores stream through the wave until the first
// within-eps pair, which joins the boxes.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_union_list(
    const T* __restrict__ soa, int64_t ld, const int2* __restrict__ bitems, Count nitems_c,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcell,
    const OrdKey<T>* __restrict__ bbox, const uint8_t* __restrict__ core,
    const int32_t* __restrict__ job_of_cell, const JobParams<T>* __restrict__ jobs,
    const uint32_t* __restrict__ off, const int32_t* __restrict__ list, int32_t* bparent,
    const unsigned long long* __restrict__ total_chk, unsigned long long cap) {
    if (total_chk && *total_chk > cap) return;  // no list was written
    const int32_t nitems = static_cast<int32_t>(nitems_c.get());
    const int32_t stride = static_cast<int32_t>(gridDim.x) * kWavesPerBlock;
    // waves walk the items wid, wid + stride, ... (one each when the grid
    // covers them)
    for (int32_t wid = uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock +
                               static_cast<int32_t>(threadIdx.x >> 6));
         wid < nitems; wid += stride) [&] {
        const int2 it = bitems[wid];
        const int32_t b = uniform(it.x);
        const uint32_t l0 = uniform(static_cast<int32_t>(off[b]));
        const uint32_t l1 = uniform(static_cast<int32_t>(off[b + 1]));
        if (l0 == l1) return;
        const int32_t end = uniform(bstart[b + 1]);
        const int32_t p = it.y + lane_id();
        const bool pcore = p < end && core[p];
        if (!__any(pcore)) return;
        const JobParams<T>& jp = jobs[uniform(job_of_cell[uniform(bcell[b])])];
        const T eps2 = jp.eps2;
        const double gap2_max = jp.gap2_max;
        T pc[D];
        load_row<T, D>(soa, ld, pcore ? p : it.y, pc);
        int32_t rb = uniform(uf_find(bparent, b));
        for (uint32_t L0 = l0; L0 < l1; L0 += kWave) {
            // lanes fetch 64 partners at once: box, bounding box, rows, and a
            // root (roots only merge, so an equal root is final)
            const uint32_t li = min(L0 + static_cast<uint32_t>(lane_id()), l1 - 1);
            const int32_t b2l = list[li];
            double lol[D], hil[D];
    #pragma unroll
            for (int j = 0; j < D; ++j) {
                lol[j] = box_lo<T, D>(bbox, b2l, j);
                hil[j] = box_hi<T, D>(bbox, b2l, j);
            }
            const int32_t q0l = bstart[b2l], q1l = bstart[b2l + 1];
            const int32_t r2l = uf_find(bparent, b2l);
            const int32_t nk = static_cast<int32_t>(min(static_cast<uint32_t>(kWave), l1 - L0));
            for (int32_t j = 0; j < nk; ++j) {
                double lo[D], hi[D];
    #pragma unroll
                for (int e = 0; e < D; ++e) {
                    lo[e] = bcast(lol[e], j);
                    hi[e] = bcast(hil[e], j);
                }
                double gmin2, gmax2;
                point_box_lh<T, D>(lo, hi, pc, gmin2, gmax2);
                const bool act = pcore && gmin2 <= gap2_max;
                if (!__any(act)) continue;
                int32_t rb2 = bcast(r2l, j);
                if (rb2 == rb) continue;
                rb = uniform(uf_find(bparent, rb));
                rb2 = uniform(uf_find(bparent, rb2));
                if (rb2 == rb) continue;
                bool hit = false;
                auto body = [&](int32_t, const T* qc, uint8_t cq, int32_t) {
                    if (cq && act && within<D>(pc, qc, eps2)) hit = true;
                };
                auto stop = [&] { return static_cast<bool>(__any(hit)); };
                if (visit_range<T, D>(soa, ld, core, nullptr, bcast(q0l, j), bcast(q1l, j), body,
                                      stop)) {
                    if (lane_id() == 0) uf_unite(bparent, rb, rb2);
                    rb = uniform(uf_find(bparent, b));
                }
            }
        }
    }();
}

// Union over the open pairs of the boxes of at most small_rows rows, one
// thread per box (a wave per box would leave most of its lanes idle): the
// box's core rows meet each listed box whose root differs -- a row within
// eps of the other box's whole bounding box joins at once, a row farther
// than eps from it skips it, and otherwise the rows are tested against the
// other box's cores until the first within-eps pair.  Every pair is decided
// exactly (no test budget), as in k_box_union_list.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_union_small(
    const T* __restrict__ soa, int64_t ld, Count nbox_c, const int32_t* __restrict__ bstart,
    const int32_t* __restrict__ bcell, const OrdKey<T>* __restrict__ bbox,
    const uint8_t* __restrict__ core, const int32_t* __restrict__ job_of_cell,
    const JobParams<T>* __restrict__ jobs, const uint32_t* __restrict__ off,
    const int32_t* __restrict__ list, int32_t* bparent,
    const unsigned long long* __restrict__ total_chk, unsigned long long cap, int32_t small_rows) {
    if (total_chk && *total_chk > cap) return;  // no list was written
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= static_cast<int32_t>(nbox_c.get())) return;
    const uint32_t l0 = off[b], l1 = off[b + 1];
    if (l0 == l1) return;
    const int32_t s0 = bstart[b], s1 = bstart[b + 1];
    if (s1 - s0 > small_rows) return;  // one wave per chunk instead
    const JobParams<T>& jp = jobs[job_of_cell[bcell[b]]];
    const T eps2 = jp.eps2;
    const double gap2_max = jp.gap2_max;
    const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
    for (uint32_t L = l0; L < l1; ++L) {
        const int32_t b2 = list[L];
        int32_t r = uf_find(bparent, b);
        const int32_t r2 = uf_find(bparent, b2);
        if (r == r2) continue;
        const int32_t t0 = bstart[b2], t1 = bstart[b2 + 1];
        bool hit = false;
        for (int32_t p = s0; p < s1 && !hit; ++p) {
            if (!core[p]) continue;
            T pc[D];
            load_row<T, D>(soa, ld, p, pc);
            double gmin2, gmax2;
            point_box<T, D>(bbox, b2, pc, gmin2, gmax2);
            if (gmin2 > gap2_max) continue;
            if (gmax2 <= full2) {  // within eps of all of b2, its cores included
                hit = true;
                break;
            }
            for (int32_t q = t0; q < t1; ++q) {
                if (!core[q]) continue;
                T qc[D];
                load_row<T, D>(soa, ld, q, qc);
                if (within<D>(pc, qc, eps2)) {
                    hit = true;
                    break;
                }
            }
        }
        if (hit) uf_unite(bparent, r, r2);
    }
}

// Hook of a small box (at most kSmallBox rows), one thread per box: the
// box's core rows meet the earlier boxes of its neighbourhood, own neighbour
// row first; a box pair gets at most kSmallTests distance tests (rows
// farther than eps from the other box's bounding box take none) before the
// search moves on: a missed hook only leaves work for the exact union.

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_hook_small(
    const T* __restrict__ soa, int64_t ld, const int32_t* __restrict__ blist, Count nlist_c,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcell,
    const int32_t* __restrict__ cbox, const uint32_t* __restrict__ roff,
    const int2* __restrict__ rcells, const int32_t* __restrict__ rcenter,
    const OrdKey<T>* __restrict__ bbox, const int32_t* __restrict__ bmincore,
    const uint8_t* __restrict__ core, const int32_t* __restrict__ job_of_cell,
    const JobParams<T>* __restrict__ jobs, int32_t* __restrict__ bparent) {
    const int32_t nlist = static_cast<int32_t>(nlist_c.get());
    const int32_t i = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= nlist) return;
    const int32_t b = blist[i];
    const int32_t s0 = bstart[b], s1 = bstart[b + 1];
    if (s1 - s0 > kSmallBox) return;  // k_box_hook's
    const int32_t u = bcell[b];
    const JobParams<T>& jp = jobs[job_of_cell[u]];
    const T eps2 = jp.eps2;
    const double gap2_max = jp.gap2_max;
    const double full2 = static_cast<double>(eps2) * (1.0 - 1e-5);
    double blo[D], bhi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        blo[j] = box_lo<T, D>(bbox, b, j);
        bhi[j] = box_hi<T, D>(bbox, b, j);
    }
    const int32_t r0 = static_cast<int32_t>(roff[u]), r1 = static_cast<int32_t>(roff[u + 1]);
    const int32_t center = rcenter[u];
    for (int32_t k = -1; k < r1 - r0; ++k) {
        const int32_t r = k < 0 ? center : r0 + k;
        if (k >= 0 && r == center) continue;
        const int2 cs = rcells[r];
        const int32_t e0 = cbox[cs.x];
        if (e0 >= b) continue;  // only boxes before b
        const int32_t e1 = min(cbox[cs.y], b);
        // the own row from the own cell's first box on, wrapping round
        const int32_t len = e1 - e0;
        const int32_t own = cbox[u];
        const int32_t rot = (k < 0 && own > e0 && own < e1) ? own - e0 : 0;
        for (int32_t t = 0; t < len; ++t) {
            const int32_t b2 = e0 + (t + rot < len ? t + rot : t + rot - len);
            if (bmincore[b2] == INT_MAX) continue;
            // the two bounding boxes: beyond eps of each other, nothing to
            // test; every pair within eps, the boxes' cores meet
            double gap2 = 0.0, far2 = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const double lo = box_lo<T, D>(bbox, b2, j), hi = box_hi<T, D>(bbox, b2, j);
                const double g = fmax(0.0, fmax(lo - bhi[j], blo[j] - hi));
                const double f = fmax(hi - blo[j], bhi[j] - lo);
                gap2 += g * g;
                far2 += f * f;
            }
            if (gap2 > gap2_max) continue;
            if (far2 <= full2) {
                bparent[b] = b2;
                return;
            }
            const int32_t t0 = bstart[b2], t1 = bstart[b2 + 1];
            int tests = 0;
            for (int32_t p = s0; p < s1 && tests < kSmallTests; ++p) {
                if (!core[p]) continue;
                T pc[D];
                load_row<T, D>(soa, ld, p, pc);
                double gmin2, gmax2;
                point_box<T, D>(bbox, b2, pc, gmin2, gmax2);
                if (gmin2 > gap2_max) continue;
                if (gmax2 <= full2) {  // within eps of all of b2, its cores included
                    bparent[b] = b2;
                    return;
                }
                for (int32_t q = t0; q < t1 && tests < kSmallTests; ++q) {
                    if (!core[q]) continue;
                    ++tests;
                    T qc[D];
                    load_row<T, D>(soa, ld, q, qc);
                    if (within<D>(pc, qc, eps2)) {
                        bparent[b] = b2;
                        return;
                    }
                }
            }
        }
    }
}

// Stream compaction with one atomic per block: block `blockIdx.x` takes
// the kSelectPer * kThreads consecutive indices from blockIdx.x *
// kSelectSpan, wave w the 64 * kSelectPer of them from w * 64 * kSelectPer,
// in rounds of 64.  The waves' counts meet in LDS, thread 0 reserves the
// block's range with one atomicAdd, and every taken index i is emitted at
// its place in the range (input order inside a block; blocks in the order
// of their reservations).  One reservation per wave instead serialises
// thousands of atomics on one address.
constexpr int kSelectPer = 8;
constexpr int kSelectSpan = kThreads * kSelectPer;

template <typename Take, typename Emit>
__device__ __forceinline__ void block_select(int64_t n, unsigned int* count, Take take,
                                             Emit emit) {
    __shared__ unsigned int wcount[kWavesPerBlock];
    __shared__ unsigned int bbase;
    const int w = static_cast<int>(threadIdx.x >> 6);
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kSelectSpan +
                       static_cast<int64_t>(w) * kWave * kSelectPer + lane_id();
    uint64_t masks[kSelectPer];
    unsigned int total = 0;
#pragma unroll
    for (int r = 0; r < kSelectPer; ++r) {
        const int64_t i = i0 + r * kWave;
        masks[r] = ballot(i < n && take(i));
        total += static_cast<unsigned int>(__popcll(masks[r]));
    }
    if (lane_id() == 0) wcount[w] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
#pragma unroll
        for (int k = 0; k < kWavesPerBlock; ++k) t += wcount[k];
        bbase = t ? atomicAdd(count, t) : 0u;
    }
    __syncthreads();
    unsigned int base = bbase;
    for (int k = 0; k < w; ++k) base += wcount[k];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < kSelectPer; ++r) {
        if ((masks[r] >> lane_id()) & 1ull)
            emit(i0 + r * kWave, base + static_cast<unsigned int>(__popcll(masks[r] & lt)));
        base += static_cast<unsigned int>(__popcll(masks[r]));
    }
}

// list[...] = every box b with (mode 0) cores (v = bmincore != INT_MAX) /
// (mode 1) v[b] == b (a root); *count += their number
__global__ __launch_bounds__(kThreads) void k_select_boxes(Count nbox_c, int mode,
                                                           const int32_t* __restrict__ v,
                                                           int32_t* __restrict__ list,
                                                           unsigned int* __restrict__ count,
                                                           const int32_t* __restrict__ bstart) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    block_select(
        nbox, count,
        [&](int64_t b) {
            if (mode == 1) return v[b] == static_cast<int32_t>(b);
            return v[b] != INT_MAX && (mode == 0 || bstart[b + 1] - bstart[b] > kSmallBox);
        },
        [&](int64_t b, unsigned int pos) { list[pos] = static_cast<int32_t>(b); });
}

// the 64-slot chunks of the boxes that have open pairs (off[b+1] > off[b])
__global__ __launch_bounds__(kThreads) void k_select_items(Count nitems_c,
                                                           const int2* __restrict__ items,
                                                           const uint32_t* __restrict__ off,
                                                           const int32_t* __restrict__ bstart,
                                                           int32_t min_rows,
                                                           int2* __restrict__ out,
                                                           unsigned int* __restrict__ count) {
    const int32_t nitems = static_cast<int32_t>(nitems_c.get());
    block_select(
        nitems, count,
        [&](int64_t i) {
            const int32_t b = items[i].x;
            return off[b + 1] > off[b] && (!bstart || bstart[b + 1] - bstart[b] > min_rows);
        },
        [&](int64_t i, unsigned int pos) { out[pos] = items[i]; });
}

__global__ __launch_bounds__(kThreads) void k_core_full(int64_t n,
                                                        const int32_t* __restrict__ box_of_slot,
                                                        const uint8_t* __restrict__ btight,
                                                        const int32_t* __restrict__ bstart,
                                                        const uint32_t* __restrict__ inrow,
                                                        uint8_t* __restrict__ core,
                                                        int32_t* __restrict__ bmincore) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (s >= n) return;
    const int32_t b = box_of_slot[s];
    const bool full = (btight[b] & 2) != 0;
    core[s] = full ? 1 : 0;
    if (full && bstart[b] == s) bmincore[b] = static_cast<int32_t>(inrow[s]);
}

// the chunks holding a row of a box that is not tight with >= minPts rows
__global__ __launch_bounds__(kThreads) void k_select_core_chunks(
    Count nitems_c, const int2* __restrict__ items, const int32_t* __restrict__ ustart,
    const int32_t* __restrict__ box_of_slot, const uint8_t* __restrict__ btight,
    int2* __restrict__ out, unsigned int* __restrict__ count) {
    const int32_t nitems = static_cast<int32_t>(nitems_c.get());
    block_select(
        nitems, count,
        [&](int64_t i) {
            const int2 it = items[i];
            const int32_t last = min(it.y + kWave, ustart[it.x + 1]) - 1;
            const int32_t b1 = box_of_slot[last];
            for (int32_t b = box_of_slot[it.y]; b <= b1; ++b)
                if (!(btight[b] & 2)) return true;
            return false;
        },
        [&](int64_t i, unsigned int pos) { out[pos] = items[i]; });
}

// root of every box's hook tree -> rsnap (-1 for boxes without cores)
__global__ __launch_bounds__(kThreads) void k_box_roots(Count nbox_c,
                                                        const int32_t* __restrict__ bmincore,
                                                        int32_t* bparent,
                                                        int32_t* __restrict__ rsnap) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= nbox) return;
    rsnap[b] = bmincore[b] == INT_MAX ? -1 : uf_find(bparent, b);
}

// Box-level union: one wave per 64-slot chunk of a box with cores; lanes
// own the chunk's rows.  Neighbour boxes come from the cell's neighbour rows
// (cells [a, b) -> boxes [cbox[a], cbox[b])); each unordered box pair is
// tested by the chunks of its lower box.  A pair is skipped when the boxes
// are already joined (live roots) or their bounding boxes are farther apart
// than eps; otherwise the lanes whose row is a core within eps of the other
// box's bounding box test the other box's cores as they stream through the
// wave, until the first within-eps pair, which joins the two boxes.
// One chunk of k_box_union.
template <typename T, int D>
__device__ __forceinline__ void box_union_item(
    const T* __restrict__ soa, int64_t ld, const int2* __restrict__ bitems,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcell,
    const int32_t* __restrict__ cbox, const uint32_t* __restrict__ roff,
    const int2* __restrict__ rcells, const OrdKey<T>* __restrict__ bbox,
    const int32_t* __restrict__ bmincore, const uint8_t* __restrict__ core,
    const int32_t* __restrict__ job_of_cell, const JobParams<T>* __restrict__ jobs,
    const int32_t* __restrict__ rsnap, int32_t* bparent, int32_t wid) {
    const int2 it = bitems[wid];
    const int32_t b = uniform(it.x);
    if (bmincore[b] == INT_MAX) return;
    const int32_t sb = uniform(rsnap[b]);
    const int32_t end = uniform(bstart[b + 1]);
    const int32_t p = it.y + lane_id();
    const bool pcore = p < end && core[p];
    if (!__any(pcore)) return;
    const int32_t u = uniform(bcell[b]);
    const JobParams<T>& jp = jobs[uniform(job_of_cell[u])];
    const T eps2 = jp.eps2;
    const double gap2_max = jp.gap2_max;
    T pc[D];
    load_row<T, D>(soa, ld, pcore ? p : it.y, pc);
    double lo[D], hi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        lo[j] = box_lo<T, D>(bbox, b, j);
        hi[j] = box_hi<T, D>(bbox, b, j);
    }
    int32_t rb = uniform(uf_find(bparent, b));
    const int32_t r0 = uniform(static_cast<int32_t>(roff[u]));
    const int32_t r1 = uniform(static_cast<int32_t>(roff[u + 1]));
    for (int32_t r = r0; r < r1; ++r) {
        const int2 cs = rcells[r];
        const int32_t e0 = uniform(cbox[uniform(cs.x)]);
        const int32_t e1 = uniform(cbox[uniform(cs.y)]);
        for (int32_t b2 = max(e0, b + 1); b2 < e1; ++b2) {
            // joined by the hook forest already, or no cores (rsnap -1)
            const int32_t s2 = uniform(rsnap[b2]);
            if (s2 == sb || s2 < 0) continue;
            double gap2 = 0.0, pgap2 = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const double lo2 = box_lo<T, D>(bbox, b2, j);
                const double hi2 = box_hi<T, D>(bbox, b2, j);
                const double g = fmax(0.0, fmax(lo2 - hi[j], lo[j] - hi2));
                gap2 += g * g;
                const double v = static_cast<double>(pc[j]);
                const double h = fmax(0.0, fmax(lo2 - v, v - hi2));
                pgap2 += h * h;
            }
            if (gap2 > gap2_max) continue;
            const bool act = pcore && pgap2 <= gap2_max;
            if (!__any(act)) continue;
            const int32_t rb2 = uniform(uf_find(bparent, b2));
            rb = uniform(uf_find(bparent, rb));
            if (rb2 == rb) continue;
            bool hit = false;
            auto body = [&](int32_t, const T* qc, uint8_t cq, int32_t) {
                if (cq && act && within<D>(pc, qc, eps2)) hit = true;
            };
            auto stop = [&] { return static_cast<bool>(__any(hit)); };
            if (visit_range<T, D>(soa, ld, core, nullptr, uniform(bstart[b2]),
                                  uniform(bstart[b2 + 1]), body, stop)) {
                if (lane_id() == 0) uf_unite(bparent, rb, rb2);
                rb = uniform(uf_find(bparent, b));
            }
        }
    }
}

// The full union walk: every box chunk against the later boxes of its
// neighbourhood.  The fallback when the open-pair lists do not fit, so it is
// launched on a small grid whose waves walk the chunks (only_over: run only
// when *only_over > cap, decided on the device).
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_box_union(
    const T* __restrict__ soa, int64_t ld, const int2* __restrict__ bitems, int32_t nitems,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcell,
    const int32_t* __restrict__ cbox, const uint32_t* __restrict__ roff,
    const int2* __restrict__ rcells, const OrdKey<T>* __restrict__ bbox,
    const int32_t* __restrict__ bmincore, const uint8_t* __restrict__ core,
    const int32_t* __restrict__ job_of_cell, const JobParams<T>* __restrict__ jobs,
    const int32_t* __restrict__ rsnap, int32_t* bparent,
    const unsigned long long* __restrict__ only_over, unsigned long long cap) {
    if (only_over && *only_over <= cap) return;
    const int32_t stride = static_cast<int32_t>(gridDim.x) * kWavesPerBlock;
    for (int32_t wid = uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock +
                               static_cast<int32_t>(threadIdx.x >> 6));
         wid < nitems; wid += stride)
        box_union_item<T, D>(soa, ld, bitems, bstart, bcell, cbox, roff, rcells, bbox, bmincore,
                             core, job_of_cell, jobs, rsnap, bparent, wid);
}

// Cluster ids.  A sorted position orders rows by (occupied cell, input
// row), so the smallest position of a component's cores is the smallest
// key (cell << 32 | row) of its boxes' smallest core rows (bmincore holds
// the input row); the position itself is the cell's first slot plus the
// number of the cell's rows with a smaller input row (every box of the
// cell holds its rows in input order, so a binary search per box counts).
__global__ __launch_bounds__(kThreads) void k_box_compmin(Count nbox_c,
                                                          const int32_t* __restrict__ bmincore,
                                                          const int32_t* __restrict__ bcell,
                                                          int32_t* bparent,
                                                          unsigned long long* __restrict__ cmin) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    const bool act = b < nbox && bmincore[b] != INT_MAX;
    // neighbouring boxes mostly share their root: lanes of a run of equal
    // roots combine first, and only the run's first lane takes the atomic
    const int32_t r = act ? uf_find(bparent, b) : -1 - lane_id();
    unsigned long long key = act ? (static_cast<unsigned long long>(bcell[b]) << 32) |
                                       static_cast<uint32_t>(bmincore[b])
                                 : ~0ull;
    key = run_suffix_min<unsigned long long>(key, r);
    const bool head = run_head(r);  // every lane takes part in the shuffles
    if (act && head) atomicMin(&cmin[r], key);
}

// number of entries < row in the ascending v[lo, hi)
__device__ __forceinline__ int32_t count_below(const uint32_t* __restrict__ v, int32_t lo,
                                               int32_t hi, uint32_t row) {
    const int32_t first = lo;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (v[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    return lo - first;
}

// position of every component root's smallest core, one thread per box
// (non-roots leave at once)
__global__ __launch_bounds__(kThreads) void k_box_resolve(
    Count nbox_c, const int32_t* __restrict__ bparent, const unsigned long long* __restrict__ cmin,
    const int32_t* __restrict__ cbox, const int32_t* __restrict__ bstart,
    const uint32_t* __restrict__ inrow, const int32_t* __restrict__ ustart,
    int32_t* __restrict__ pos) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t r = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (r >= nbox || bparent[r] != r) return;
    const unsigned long long key = cmin[r];
    if (key == ~0ull) return;  // no cores
    const int32_t u = static_cast<int32_t>(key >> 32);
    const uint32_t row = static_cast<uint32_t>(key & 0xffffffffu);
    int32_t below = 0;
    for (int32_t b2 = cbox[u]; b2 < cbox[u + 1]; ++b2)
        below += count_below(inrow, bstart[b2], bstart[b2 + 1], row);
    pos[r] = ustart[u] + below;
}

// cluster id (smallest core position of the component) per box with cores
__global__ __launch_bounds__(kThreads) void k_box_cluster(Count nbox_c,
                                                          const int32_t* __restrict__ bmincore,
                                                          int32_t* bparent,
                                                          const int32_t* __restrict__ compmin,
                                                          int32_t* __restrict__ bcluster) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= nbox) return;
    bcluster[b] = bmincore[b] == INT_MAX ? -1 : compmin[uf_find(bparent, b)];
}

__global__ __launch_bounds__(kThreads) void k_box_items_count(const int32_t* __restrict__ bstart,
                                                              Count nbox_c,
                                                              uint32_t* __restrict__ cnt) {
    const int32_t nbox = static_cast<int32_t>(nbox_c.get());
    const int32_t b = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= nbox) return;
    cnt[b] = static_cast<uint32_t>((bstart[b + 1] - bstart[b] + kWave - 1) / kWave);
}



__global__ __launch_bounds__(kThreads) void k_fill_u64(unsigned long long* __restrict__ a,
                                                       unsigned long long v, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kThreads) void k_fill_i32(int32_t* __restrict__ a, int32_t v,
                                                       int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kThreads) void k_iota_i32(int32_t* __restrict__ a, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) a[i] = static_cast<int32_t>(i);
}

}  // namespace

int box_split(int d) { return d <= 3 ? 2 : 3; }

int box_sub_bits(int d) { return box_sub_bits(d, box_split(d)); }

int box_sub_bits(int d, int k) {
    uint64_t v = 1;
    for (int j = 0; j < d; ++j) v *= static_cast<uint64_t>(k);
    const uint64_t mx = v - 1;
    return mx == 0 ? 0 : 64 - __builtin_clzll(mx);
}


template <typename T, int D>
void launch_gather_boxes(const T* pts, int64_t n, const uint32_t* inrow_of_slot, T* soa,
                         int64_t ld, const int32_t* box_of_slot, const int32_t* bstart, void* bbox,
                         void* part, hipStream_t s) {
    using U = OrdKey<T>;
    if (n <= 0) return;
    int64_t tile = kThreads;
    if constexpr (D * sizeof(T) <= 8) {
        if (!pts) tile = kBoundsTile;  // coordinates in place: bounding boxes only
    }
    const int64_t nb = (n + tile - 1) / tile;
    if (tile == kThreads) {
        hipLaunchKernelGGL((k_gather_boxes<T, D>), dim3(static_cast<unsigned>(nb)), dim3(kThreads),
                           0, s, pts, n, inrow_of_slot, soa, ld, box_of_slot, bstart,
                           static_cast<U*>(bbox), static_cast<U*>(part));
    } else {
        if constexpr (D * sizeof(T) <= 8)
            hipLaunchKernelGGL((k_box_bounds<T, D>), dim3(static_cast<unsigned>(nb)),
                               dim3(kThreads), 0, s, soa, ld, n, box_of_slot, bstart,
                               static_cast<U*>(bbox), static_cast<U*>(part));
    }
    hipLaunchKernelGGL((k_box_span<T, D>), dim3(blocks_for(nb, kWavesPerBlock)), dim3(kThreads), 0,
                       s, n, nb, tile, box_of_slot, bstart, static_cast<const U*>(part),
                       static_cast<U*>(bbox));
    HPD_HIP(hipGetLastError());
}

size_t gather_part_bytes(int64_t n, int d, int value_bytes) {
    const size_t nb = static_cast<size_t>((n + kThreads - 1) / kThreads);
    return 2 * nb * 2 * static_cast<size_t>(d) * static_cast<size_t>(value_bytes);
}

template <typename K>
void launch_box_table(const K* keys, int64_t n, const uint32_t* cflags, const uint32_t* cscan,
                      uint32_t* flags, uint32_t* scan, uint32_t* scan_tmp, int32_t* bstart,
                      int32_t* bcell, int32_t* box_of_slot, uint32_t* nbox_dev, void* bbox,
                      int bbox_keys, int key_bytes, hipStream_t s, int shift, uint32_t* cbflag) {
    if (!cflags && !cbflag)
        throw Error(HPD_EINVAL, "launch_box_table: cell flags or the per-box flag buffer needed");
    hipLaunchKernelGGL(k_box_flags<K>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, keys, n,
                       flags);
    exclusive_scan_u32(flags, scan, n, scan_tmp, nbox_dev, s);
    hipLaunchKernelGGL(k_box_table<K>, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, keys, n,
                       static_cast<const uint32_t*>(flags), static_cast<const uint32_t*>(scan),
                       cflags, cscan, bstart, bcell, box_of_slot,
                       static_cast<const uint32_t*>(nbox_dev), bbox, bbox_keys, key_bytes, shift,
                       cbflag);
    HPD_HIP(hipGetLastError());
}

template <typename K>
void launch_cells_of_boxes(const K* keys, int64_t n, int shift, const int32_t* bstart,
                           uint32_t* nbox_dev, uint32_t* cbflag, uint32_t* cbscan,
                           uint32_t* scan_tmp, int32_t* bcell, K* ucell, int32_t* ustart,
                           uint32_t* ncells_dev, hipStream_t s) {
    const Count nb{n, nbox_dev};
    exclusive_scan_u32_dev(cbflag, cbscan, nb, scan_tmp, s);
    copy_at_count(cbscan, nbox_dev, ncells_dev, s);
    const int64_t want = blocks_for(n, kThreads);
    hipLaunchKernelGGL(k_cells_of_boxes<K>, dim3(static_cast<unsigned>(want < 4096 ? want : 4096)),
                       dim3(kThreads), 0, s, keys, n, shift, bstart, nb,
                       static_cast<const uint32_t*>(cbflag), static_cast<const uint32_t*>(cbscan),
                       bcell, ucell, ustart, static_cast<const uint32_t*>(ncells_dev));
    HPD_HIP(hipGetLastError());
}

void launch_cell_boxes(const int32_t* bcell, Count nbox, Count m, int32_t* cbox, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_boxes, dim3(blocks_for(nbox.host > 0 ? nbox.host : 1, kThreads)),
                       dim3(kThreads), 0, s, bcell, nbox, m, cbox);
}

void launch_box_items(const int32_t* bstart, Count nbox, uint32_t* cnt, uint32_t* scan_tmp,
                      uint32_t* total_dev, hipStream_t s) {
    if (nbox.host > 0)
        hipLaunchKernelGGL(k_box_items_count, dim3(blocks_for(nbox.host, kThreads)), dim3(kThreads),
                           0, s, bstart, nbox, cnt);
    if (nbox.dev) {
        exclusive_scan_u32_dev(cnt, cnt, nbox, scan_tmp, s);
        copy_at_count(cnt, nbox.dev, total_dev, s);
    } else {
        exclusive_scan_u32(cnt, cnt, nbox.host, scan_tmp, total_dev, s);
    }
}

void launch_box_items_fill(const int32_t* bstart, Count nbox, const uint32_t* off, Count nitems,
                           int2* items, hipStream_t s) {
    launch_expand_chunks(off, nbox, bstart, nitems, items, s);
}

template <typename T, int D>
void launch_box_tight(const BoxArgs<T>& a, uint8_t* btight, unsigned int* nloose, hipStream_t s) {
    using U = OrdKey<T>;
    if (a.nbox <= 0) return;
    hipLaunchKernelGGL((k_box_tight<T, D>), dim3(blocks_for(a.nbox, kThreads)), dim3(kThreads), 0,
                       s, reinterpret_cast<const U*>(a.bbox), Count{a.nbox, a.nbox_dev}, a.bstart,
                       a.bcell, a.job_of_cell, a.jobs, btight, nloose);
}

// A grid of waves for a count of work items: one wave per item of a
// host-known count (the kernels' wave loops then run once), at most
// kMaxItemWaves, which walk a larger count.  For a count an earlier kernel
// wrote (the host holds a bound, n or n + n / 64) one wave per item up to an
// eighth of the bound, which walk the items beyond it: every wave past the
// count costs a dispatch.
constexpr int64_t kMaxItemWaves = int64_t{1} << 24;
inline unsigned wave_blocks(Count items) {
    int64_t w = items.dev && !items.tight ? std::min(items.host, items.host / 8 + 64) : items.host;
    w = std::min(w, kMaxItemWaves);
    return blocks_for(w > 0 ? w : 1, kWavesPerBlock);
}

template <typename T, int D>
void launch_box_hook(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                     const int32_t* bmincore, const uint8_t* core, const int32_t* blist,
                     Count nlist, int32_t* bparent, int32_t* big, unsigned int* nbig,
                     hipStream_t s) {
    using U = OrdKey<T>;
    if (a.nbox > 0)
        hipLaunchKernelGGL(k_iota_i32, dim3(blocks_for(a.nbox, kThreads)), dim3(kThreads), 0, s,
                           bparent, static_cast<int64_t>(a.nbox));
    if (nlist.host > 0) {
        // the larger boxes one wave each (first: their hooks are found
        // soonest), the small ones one thread each
        const Count nb{nlist.host, nbig};
        hipLaunchKernelGGL((k_box_hook<T, D>), dim3(wave_blocks(nb)), dim3(kThreads), 0, s, a.soa,
                           a.ld, big, nb, a.bstart, a.bcell, a.cbox, roff, rcells,
                           reinterpret_cast<const U*>(a.bbox), bmincore, core, a.job_of_cell, a.jobs,
                           a.rcenter, bparent);
        hipLaunchKernelGGL((k_box_hook_small<T, D>), dim3(blocks_for(nlist.host, kThreads)),
                           dim3(kThreads), 0, s, a.soa, a.ld, blist, nlist, a.bstart, a.bcell,
                           a.cbox, roff, rcells, a.rcenter, reinterpret_cast<const U*>(a.bbox),
                           bmincore, core, a.job_of_cell, a.jobs, bparent);
    }
}

void launch_box_roots(Count nbox, const int32_t* bmincore, int32_t* bparent, int32_t* rsnap,
                      hipStream_t s) {
    if (nbox.host <= 0) return;
    hipLaunchKernelGGL(k_box_roots, dim3(blocks_for(nbox.host, kThreads)), dim3(kThreads), 0, s,
                       nbox, bmincore, bparent, rsnap);
}

template <typename T, int D>
void launch_box_root_join(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                          const int32_t* rsnap, const uint8_t* core, const int32_t* blist,
                          Count nlist, int32_t* bparent, hipStream_t s) {
    using U = OrdKey<T>;
    if (nlist.host <= 0) return;
    hipLaunchKernelGGL((k_box_root_join<T, D>), dim3(wave_blocks(nlist)), dim3(kThreads), 0, s,
                       a.soa, a.ld, blist, nlist, a.bstart, a.bcell, a.cbox, roff, rcells,
                       reinterpret_cast<const U*>(a.bbox), rsnap, core, a.job_of_cell, a.jobs,
                       a.rcenter, bparent);
}

void launch_select_boxes(Count nbox, int mode, const int32_t* v, int32_t* list,
                         unsigned int* count, hipStream_t s, const int32_t* bstart) {
    if (nbox.host <= 0) return;
    if (mode == 2 && !bstart) throw Error(HPD_EINVAL, "launch_select_boxes: mode 2 needs bstart");
    hipLaunchKernelGGL(k_select_boxes, dim3(blocks_for(nbox.host, kSelectSpan)), dim3(kThreads), 0,
                       s, nbox, mode, v, list, count, bstart);
}

void launch_core_full(int64_t n, const int32_t* box_of_slot, const uint8_t* btight,
                      const int32_t* bstart, const uint32_t* inrow, uint8_t* core,
                      int32_t* bmincore, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_core_full, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, n,
                       box_of_slot, btight, bstart, inrow, core, bmincore);
}

void launch_select_core_chunks(Count nitems, const int2* items, const int32_t* ustart,
                               const int32_t* box_of_slot, const uint8_t* btight, int2* out,
                               unsigned int* count, hipStream_t s) {
    if (nitems.host <= 0) return;
    hipLaunchKernelGGL(k_select_core_chunks, dim3(blocks_for(nitems.host, kSelectSpan)),
                       dim3(kThreads), 0, s, nitems, items, ustart, box_of_slot, btight, out, count);
}

void launch_select_items(Count nitems, const int2* items, const uint32_t* off, int2* out,
                         unsigned int* count, hipStream_t s, const int32_t* bstart,
                         int32_t min_rows) {
    if (nitems.host <= 0) return;
    hipLaunchKernelGGL(k_select_items, dim3(blocks_for(nitems.host, kSelectSpan)), dim3(kThreads),
                       0, s, nitems, items, off, min_rows > 0 ? bstart : nullptr, min_rows, out,
                       count);
}

int32_t small_union_rows(int64_t n, int64_t nbox) {
    return nbox > 0 && n < kSmallUnionMean * nbox ? kSmallUnionRows : 0;
}

template <typename T, int D>
void launch_box_join(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                     const int32_t* bmincore, const uint8_t* core, const int32_t* rsnap,
                     int32_t* bparent, hipStream_t s, const unsigned long long* only_over,
                     unsigned long long cap) {
    using U = OrdKey<T>;
    if (a.nbitems <= 0) return;
    const int64_t waves = std::min<int64_t>(a.nbitems, 32768);
    hipLaunchKernelGGL((k_box_union<T, D>), dim3(blocks_for(waves, kWavesPerBlock)),
                       dim3(kThreads), 0, s, a.soa, a.ld, a.bitems, a.nbitems, a.bstart, a.bcell,
                       a.cbox, roff, rcells, reinterpret_cast<const U*>(a.bbox), bmincore, core,
                       a.job_of_cell, a.jobs, rsnap, bparent, only_over, cap);
}

template <typename T, int D>
void launch_box_scan(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                     const int32_t* rsnap, uint32_t* cnt, int32_t* list,
                     unsigned long long* total, hipStream_t s,
                     const unsigned long long* total_chk, unsigned long long cap) {
    using U = OrdKey<T>;
    if (a.nbox <= 0) return;
    hipLaunchKernelGGL((k_box_scan<T, D>), dim3(blocks_for(a.nbox, kThreads)), dim3(kThreads), 0,
                       s, Count{a.nbox, a.nbox_dev}, a.bcell, a.cbox, roff, rcells,
                       reinterpret_cast<const U*>(a.bbox), a.job_of_cell, a.jobs, rsnap, cnt, list,
                       total, total_chk, cap);
}

template <typename T, int D>
void launch_box_union_list(const BoxArgs<T>& a, const int2* items, Count nitems,
                           const uint8_t* core, const uint32_t* off, const int32_t* list,
                           int32_t* bparent, hipStream_t s, const unsigned long long* total_chk,
                           unsigned long long cap, int32_t small_rows) {
    using U = OrdKey<T>;
    if (a.nbox > 0 && small_rows > 0)
        hipLaunchKernelGGL((k_box_union_small<T, D>), dim3(blocks_for(a.nbox, kThreads)),
                           dim3(kThreads), 0, s, a.soa, a.ld, Count{a.nbox, a.nbox_dev}, a.bstart,
                           a.bcell, reinterpret_cast<const U*>(a.bbox), core, a.job_of_cell,
                           a.jobs, off, list, bparent, total_chk, cap, small_rows);
    if (nitems.host <= 0) return;
    hipLaunchKernelGGL((k_box_union_list<T, D>), dim3(wave_blocks(nitems)), dim3(kThreads), 0, s,
                       a.soa, a.ld, items, nitems, a.bstart, a.bcell,
                       reinterpret_cast<const U*>(a.bbox), core, a.job_of_cell, a.jobs, off, list,
                       bparent, total_chk, cap);
}

void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, p, v, n);
}

void launch_box_clusters(Count nbox, const int32_t* bmincore, const int32_t* bcell,
                         const int32_t* cbox, const int32_t* bstart, const uint32_t* inrow,
                         const int32_t* ustart, int32_t* bparent, unsigned long long* cmin,
                         int32_t* compmin, int32_t* bcluster, hipStream_t s) {
    if (nbox.host <= 0) return;
    const dim3 g(blocks_for(nbox.host, kThreads));
    hipLaunchKernelGGL(k_fill_u64, g, dim3(kThreads), 0, s, cmin, ~0ull, nbox.host);
    hipLaunchKernelGGL(k_box_compmin, g, dim3(kThreads), 0, s, nbox, bmincore, bcell, bparent, cmin);
    hipLaunchKernelGGL(k_box_resolve, g, dim3(kThreads), 0, s, nbox,
                       static_cast<const int32_t*>(bparent),
                       static_cast<const unsigned long long*>(cmin), cbox, bstart, inrow, ustart,
                       compmin);
    hipLaunchKernelGGL(k_box_cluster, g, dim3(kThreads), 0, s, nbox, bmincore, bparent,
                       static_cast<const int32_t*>(compmin), bcluster);
}

// Sorted positions of all slots (for the slot-level union and the
// one-thread border rule): a stable sort of the slots by occupied cell in
// input order gives the (cell, input row) order.
__global__ __launch_bounds__(kThreads) void k_pos_keys(int64_t n, const int32_t* __restrict__ box_of_slot,
                                                       const int32_t* __restrict__ bcell,
                                                       const uint32_t* __restrict__ inrow,
                                                       uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = inrow[s];
    keys[r] = static_cast<uint32_t>(bcell[box_of_slot[s]]);
    vals[r] = static_cast<uint32_t>(s);
}

__global__ __launch_bounds__(kThreads) void k_pos_scatter(int64_t n,
                                                          const uint32_t* __restrict__ slot_at,
                                                          uint32_t* __restrict__ pos_of_slot) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) pos_of_slot[slot_at[i]] = static_cast<uint32_t>(i);
}

void launch_positions(int64_t n, int32_t m, const int32_t* box_of_slot, const int32_t* bcell,
                      const uint32_t* inrow, uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b,
                      uint32_t* vals_b, uint32_t* hist, uint32_t* tmp, uint32_t* pos_of_slot,
                      hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pos_keys, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, n,
                       box_of_slot, bcell, inrow, keys_a, vals_a);
    const int bits = m > 1 ? 64 - __builtin_clzll(static_cast<unsigned long long>(m - 1)) : 0;
    const bool in_b = radix_sort_pairs<uint32_t>(keys_a, vals_a, keys_b, vals_b, n, bits, hist, tmp, s);
    hipLaunchKernelGGL(k_pos_scatter, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, n,
                       static_cast<const uint32_t*>(in_b ? vals_b : vals_a), pos_of_slot);
    HPD_HIP(hipGetLastError());
}

#define HPD_INST_BOX_KEYS(T, D)                                                                   \
    template void launch_gather_boxes<T, D>(const T*, int64_t, const uint32_t*, T*, int64_t,      \
                                            const int32_t*, const int32_t*, void*, void*,         \
                                            hipStream_t);                                         \
    template void launch_box_tight<T, D>(const BoxArgs<T>&, uint8_t*, unsigned int*,             \
                                         hipStream_t);                                            \
    template void launch_box_hook<T, D>(const BoxArgs<T>&, const uint32_t*, const int2*,          \
                                        const int32_t*, const uint8_t*, const int32_t*, Count,    \
                                        int32_t*, int32_t*, unsigned int*, hipStream_t);          \
    template void launch_box_root_join<T, D>(const BoxArgs<T>&, const uint32_t*, const int2*,     \
                                             const int32_t*, const uint8_t*, const int32_t*,      \
                                             Count, int32_t*, hipStream_t);                       \
    template void launch_box_join<T, D>(const BoxArgs<T>&, const uint32_t*, const int2*,          \
                                        const int32_t*, const uint8_t*, const int32_t*, int32_t*, \
                                        hipStream_t, const unsigned long long*,                   \
                                        unsigned long long);                                      \
    template void launch_box_scan<T, D>(const BoxArgs<T>&, const uint32_t*, const int2*,          \
                                        const int32_t*, uint32_t*, int32_t*,                      \
                                        unsigned long long*, hipStream_t,                         \
                                        const unsigned long long*, unsigned long long);           \
    template void launch_box_union_list<T, D>(const BoxArgs<T>&, const int2*, Count,               \
                                              const uint8_t*, const uint32_t*, const int32_t*,    \
                                              int32_t*, hipStream_t, const unsigned long long*,   \
                                              unsigned long long, int32_t);
#define HPD_INST_BOX_ALL_D(T) \
    HPD_INST_BOX_KEYS(T, 1)   \
    HPD_INST_BOX_KEYS(T, 2)   \
    HPD_INST_BOX_KEYS(T, 3)   \
    HPD_INST_BOX_KEYS(T, 4)   \
    HPD_INST_BOX_KEYS(T, 5)   \
    HPD_INST_BOX_KEYS(T, 6)   \
    HPD_INST_BOX_KEYS(T, 7)   \
    HPD_INST_BOX_KEYS(T, 8)
HPD_INST_BOX_ALL_D(float)
HPD_INST_BOX_ALL_D(double)

template void launch_box_table<uint32_t>(const uint32_t*, int64_t, const uint32_t*,
                                         const uint32_t*, uint32_t*, uint32_t*, uint32_t*,
                                         int32_t*, int32_t*, int32_t*, uint32_t*, void*, int, int,
                                         hipStream_t, int, uint32_t*);
template void launch_box_table<uint64_t>(const uint64_t*, int64_t, const uint32_t*,
                                         const uint32_t*, uint32_t*, uint32_t*, uint32_t*,
                                         int32_t*, int32_t*, int32_t*, uint32_t*, void*, int, int,
                                         hipStream_t, int, uint32_t*);
template void launch_cells_of_boxes<uint32_t>(const uint32_t*, int64_t, int, const int32_t*,
                                              uint32_t*, uint32_t*, uint32_t*, uint32_t*, int32_t*,
                                              uint32_t*, int32_t*, uint32_t*, hipStream_t);
template void launch_cells_of_boxes<uint64_t>(const uint64_t*, int64_t, int, const int32_t*,
                                              uint32_t*, uint32_t*, uint32_t*, uint32_t*, int32_t*,
                                              uint64_t*, int32_t*, uint32_t*, hipStream_t);

}  // namespace hpd
