// Launchers of the clustering kernels (dbscan.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace hpd {

template <typename T>
struct DbscanArgs {
    const T* soa = nullptr;          // sorted coordinates, [D][ld]
    int64_t n = 0;
    int64_t ld = 0;                  // row stride of soa: n rounded up to 16
    const int2* items = nullptr;     // (cell, first row) per 64-row chunk
    int32_t nitems = 0;
    const int32_t* ustart = nullptr; // first row of each occupied cell, [m+1]
    const uint32_t* roff = nullptr;  // CSR offsets into ranges, [m+1]
    const int2* ranges = nullptr;    // [first, end) rows of neighbour rows (ascending, or nearest first)
    const int2* rcells = nullptr;    // [first, end) occupied cells of the same rows
    const int32_t* rcenter = nullptr; // index into ranges of each cell's own row
    int32_t m = 0;                   // occupied cells
    const JobParams<T>* jobs = nullptr;      // eps2, minPts, first row per job
    int32_t njobs = 1;
    const int32_t* job_of_cell = nullptr;    // job of every occupied cell
    // slot order: rows sorted by cell, then (when boxes are used) by sub-box
    const uint32_t* pos_of_slot = nullptr;  // sorted position per slot (null: identity)
    const uint32_t* inrow = nullptr;        // input row per slot
    const int32_t* box_of_slot = nullptr;   // box per slot (null: no boxes)
    const uint8_t* btight = nullptr;        // box flags (bit 1: all rows core)
    // box tables (null without boxes): boxes of occupied cell u are
    // [cbox[u], cbox[u+1]), rows of box b are slots [bstart[b], bstart[b+1]),
    // bbox holds ordered-key min / max per box and dimension
    const int32_t* cbox = nullptr;
    const int32_t* bstart = nullptr;
    const int32_t* bcell = nullptr;         // occupied cell of every box
    const void* bbox = nullptr;
    // box mode: the core count and the border pass take the chunks of dense
    // cells from ditems (one wave each) and the rows of sparse cells from
    // sitems, (row, cell) per row (one thread each)
    const int2* ditems = nullptr;
    int32_t nditems = 0;
    const int2* sitems = nullptr;
    int32_t nsitems = 0;
    const uint32_t* nsitems_dev = nullptr;  // static pipeline: nsitems in device memory
    const uint32_t* ncitems_dev = nullptr;  // static pipeline: ncitems in device memory
    bool sparse_core = false;  // core count split too (else one wave per chunk of every cell)
    // chunks of the cells with >= minPts candidates: the only cells that can
    // hold cores (core flags of the others stay 0; null: every chunk)
    const int2* citems = nullptr;
    int32_t ncitems = 0;
    // box mode without the core split: scratch for the chunks of citems
    // holding a row outside the tight boxes of >= minPts rows (those rows
    // are core without counting), and a device counter (zeroed) for them
    int2* citems_need = nullptr;
    unsigned int* ncitems_need_dev = nullptr;
    // when not null, the number of work items of ditems is read from here
    // (nditems is then only the grid's bound)
    const unsigned int* nditems_dev = nullptr;
    // box mode: the core count also writes the smallest input row of a core
    // row per box (pre-filled with INT_MAX; stays INT_MAX without cores)
    int32_t* bmincore = nullptr;
};

// core flag of every slot; with box tables the candidates are visited box
// by box (boxes beyond eps skipped, boxes within eps counted whole).  The
// number of cores is added to *ncore by launch_label_cores.
template <typename T, int D>
// (count_start, when not null, is recorded right before the count kernel;
// with ntests not null the box path runs an instrumented count kernel that
// adds the row-row distance tests it executes to *ntests)
void launch_core(const DbscanArgs<T>& a, uint8_t* core, unsigned long long* ntests, hipStream_t s,
                 hipEvent_t count_start = nullptr);

// Components of the core-core graph on slots (used when boxes are not):
// hook forest, root snapshot, union of the pairs not yet connected, then
// compmin[root] = smallest sorted position of each component.
// parent, rsnap: ld entries; cmin, cmax: m; compmin: n.
template <typename T, int D>
void launch_union_points(const DbscanArgs<T>& a, const uint8_t* core, int32_t* parent,
                         int32_t* rsnap, int32_t* cmin, int32_t* cmax, int32_t* compmin,
                         hipStream_t s);

// Core labels: core slots get -(cluster + 1), where cluster comes from
// bcluster (box mode, bcluster != null) or compmin[find(parent, slot)];
// rootslot[slot] = cluster, cmin / cmax = smallest / largest cluster per
// occupied cell (not in box mode, where the border pass reads bcluster).
// Labels go to input order through a.inrow; *ncore += number of cores.
// rootslot: ld entries (may be null in box mode, where only the one-thread
// border rule reads it).
// full_ok (device word, box mode): when *full_ok is set,
// launch_label_full_rows labelled the rows of the full boxes, which are
// skipped here.
template <typename T>
void launch_label_cores(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* bcluster,
                        int32_t* parent, const int32_t* compmin, int32_t* rootslot, int32_t* cmin,
                        int32_t* cmax, int64_t* labels, unsigned long long* ncore, hipStream_t s,
                        const unsigned int* full_ok = nullptr);

// Box mode (floating types): labels of the rows of full boxes (tight with
// >= minPts rows, all core, cluster = bcluster of the box) written in input
// order.  tab (cap int32 entries, scratch) is indexed by the sorted keys;
// when the largest key is below cap, *ok = 1 and each row's label is looked
// up by its key (recomputed from pts with the same sub-box split);
// otherwise *ok = 0 and nothing is labelled.  launch_label_cores counts
// every core, these rows included.
template <typename T, int D, typename K>
void launch_label_full_rows(const T* pts, const DbscanArgs<T>& a, int sub_k, int sub_bits,
                            const K* keys, Count nbox, const int32_t* bcluster, int32_t* tab,
                            int64_t cap, int64_t* labels, unsigned int* ok, hipStream_t s);

// Border and noise labels of the non-core rows (in box mode box by box:
// all cores of a box share its cluster).  stats[0] += noise rows,
// stats[1] += ambiguous border rows; when amb_list is given it receives
// (slot, cell) of each ambiguous border row (capacity n).  noise_stripes
// (kNoiseStripes zeroed counters, or null) lets the box kernels count the
// noise rows as they label them instead of a pass over the labels.
constexpr int kNoiseStripes = 1024;
template <typename T, int D>
void launch_border(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* bcluster,
                   const int32_t* rootslot, const int32_t* cmin, const int32_t* cmax,
                   int64_t* labels, unsigned long long* stats, int2* amb_list,
                   unsigned long long* noise_stripes, hipStream_t s);

// The reference's one-thread border choice for the namb ambiguous rows:
// mark the touched clusters (involved, indexed by cluster position, zeroed
// by the caller) and initialise f; relax f until *changed stays 0; assign.
template <typename T, int D>
void launch_one_thread_mark(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* rootslot,
                            const int2* amb, int32_t namb, uint8_t* involved, int32_t* f,
                            hipStream_t s);
template <typename T, int D>
void launch_one_thread_relax(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* rootslot,
                             const uint8_t* involved, int32_t* f, unsigned int* changed,
                             hipStream_t s);
template <typename T, int D>
void launch_one_thread_assign(const DbscanArgs<T>& a, const uint8_t* core, const int32_t* rootslot,
                              const int2* amb, int32_t namb, const int32_t* f, int64_t* labels,
                              hipStream_t s);

}  // namespace hpd
