// Launchers of the sub-box kernels (boxes.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace hpd {

// boxes per cell and dimension (k) and bits of the sub-box digit
int box_split(int d);
int box_sub_bits(int d);
int box_sub_bits(int d, int k);  // bits of the sub-box digit with k parts per dimension

template <typename T>
struct BoxArgs {
    const T* soa = nullptr;            // coordinates in slot order, [D][ld]
    int64_t ld = 0;
    int32_t nbox = 0;
    const uint32_t* nbox_dev = nullptr;  // static pipeline: nbox in device memory (nbox a bound)
    const int32_t* bstart = nullptr;   // first slot of every box, [nbox + 1]
    const int32_t* bcell = nullptr;    // occupied cell of every box
    const int32_t* cbox = nullptr;     // first box of every cell, [m + 1]
    const int2* bitems = nullptr;      // (box, first slot) per 64-slot chunk
    int32_t nbitems = 0;
    void* bbox = nullptr;              // ordered-key min/max per box and dim
    const JobParams<T>* jobs = nullptr;
    const int32_t* job_of_cell = nullptr;
    const int32_t* rcenter = nullptr;  // index of every cell's own neighbour row
};

// coordinates in slot order (soa[j * ld + s] = pts[inrow_of_slot[s] * D + j];
// with pts null they are there already) and the bounding box of every box
// (bbox, 2 * D ordered keys per box, all written); part holds
// gather_part_bytes(n, D, sizeof(T)) bytes of scratch
template <typename T, int D>
void launch_gather_boxes(const T* pts, int64_t n, const uint32_t* inrow_of_slot, T* soa,
                         int64_t ld, const int32_t* box_of_slot, const int32_t* bstart, void* bbox,
                         void* part, hipStream_t s);
size_t gather_part_bytes(int64_t n, int d, int value_bytes);

// Boxes of keys sorted by (cell, sub-box digit): runs of equal keys.
// cflags / cscan are the cell-run flags and their exclusive scan (they give
// each box's occupied-cell index); without them (null) cbflag[b] is set to
// whether box b starts a cell (key >> shift changes), for
// launch_cells_of_boxes.  flags and scan (n words each) and scan_tmp are
// scratch; bstart: nbox + 1 entries; writes *nbox_dev; bbox (when not null,
// bbox_keys keys of key_bytes each per box) is set to empty boxes.
template <typename K>
void launch_box_table(const K* keys, int64_t n, const uint32_t* cflags, const uint32_t* cscan,
                      uint32_t* flags, uint32_t* scan, uint32_t* scan_tmp, int32_t* bstart,
                      int32_t* bcell, int32_t* box_of_slot, uint32_t* nbox_dev, void* bbox,
                      int bbox_keys, int key_bytes, hipStream_t s, int shift = 0,
                      uint32_t* cbflag = nullptr);
// The occupied cells from the boxes of launch_box_table (every cell starts
// a box): ucell, ustart (ustart[m] = n) and *ncells_dev as launch_cells
// gives them, and bcell per box, from one scan over the boxes instead of
// a flag pass and a scan over all n slots.  cbscan: n + 1 words.
template <typename K>
void launch_cells_of_boxes(const K* keys, int64_t n, int shift, const int32_t* bstart,
                           uint32_t* nbox_dev, uint32_t* cbflag, uint32_t* cbscan,
                           uint32_t* scan_tmp, int32_t* bcell, K* ucell, int32_t* ustart,
                           uint32_t* ncells_dev, hipStream_t s);
void launch_cell_boxes(const int32_t* bcell, Count nbox, Count m, int32_t* cbox, hipStream_t s);

// cnt[b] = 64-slot chunks of box b, scanned in place; *total_dev = chunks
void launch_box_items(const int32_t* bstart, Count nbox, uint32_t* cnt, uint32_t* scan_tmp,
                      uint32_t* total_dev, hipStream_t s);
void launch_box_items_fill(const int32_t* bstart, Count nbox, const uint32_t* off, Count nitems,
                           int2* items, hipStream_t s);

// btight[b]: bit 0 tight, bit 1 tight with >= minPts rows; *nloose counts
// boxes of two or more rows that are not tight
template <typename T, int D>
void launch_box_tight(const BoxArgs<T>& a, uint8_t* btight, unsigned int* nloose, hipStream_t s);

// Components of the boxes with cores:
//   launch_select_boxes  list of the boxes with cores (mode 0, v = bmincore),
//                        of those with more than kSmallBox rows (mode 2,
//                        with bstart) or of the roots (mode 1, v = rsnap);
//                        *count += n
//   launch_box_hook      hook forest over the listed boxes (bparent is
//                        reset first): every box points at some earlier box
//                        joined to it; big[0, *nbig) lists the boxes of
//                        more than kSmallBox rows (launch_select_boxes mode
//                        2), which take one wave each
//   launch_box_roots     root of every box -> rsnap (-1 without cores)
//   launch_box_root_join every listed root joins the first tree of another
//                        box it meets (kRootJoinRounds rounds, each followed
//                        by launch_box_roots)
//   then the open pairs (below), launch_box_clusters.
template <int D>
constexpr int kRootJoinRounds = D <= 3 ? 1 : 2;

void launch_select_boxes(Count nbox, int mode, const int32_t* v, int32_t* list,
                         unsigned int* count, hipStream_t s, const int32_t* bstart = nullptr);
template <typename T, int D>
void launch_box_hook(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                     const int32_t* bmincore, const uint8_t* core, const int32_t* blist,
                     Count nlist, int32_t* bparent, int32_t* big, unsigned int* nbig,
                     hipStream_t s);
void launch_box_roots(Count nbox, const int32_t* bmincore, int32_t* bparent, int32_t* rsnap,
                      hipStream_t s);
template <typename T, int D>
void launch_box_root_join(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                          const int32_t* rsnap, const uint8_t* core, const int32_t* blist,
                          Count nlist, int32_t* bparent, hipStream_t s);
// the box chunks (items) of the boxes with open pairs (off[b+1] > off[b])
// core[s] = 1 for the rows of tight boxes with >= minPts rows (btight bit 1),
// 0 for every other row; for those boxes bmincore[b] = inrow[bstart[b]]
// (every row is core and a box holds its rows in input order)
void launch_core_full(int64_t n, const int32_t* box_of_slot, const uint8_t* btight,
                      const int32_t* bstart, const uint32_t* inrow, uint8_t* core,
                      int32_t* bmincore, hipStream_t s);
// out[..] = the chunks (cell, first row) of items holding a row outside
// those boxes; *count += their number
void launch_select_core_chunks(Count nitems, const int2* items, const int32_t* ustart,
                               const int32_t* box_of_slot, const uint8_t* btight, int2* out,
                               unsigned int* count, hipStream_t s);
// (with min_rows > 0: only chunks of boxes of more than min_rows rows; the
// thread-per-box union of launch_box_union_list takes the others)
void launch_select_items(Count nitems, const int2* items, const uint32_t* off, int2* out,
                         unsigned int* count, hipStream_t s, const int32_t* bstart = nullptr,
                         int32_t min_rows = 0);
// Boxes of at most this many rows take their open pairs one thread per box
// (0: none): on when the boxes hold fewer than kSmallUnionMean rows on
// average (sparse data, where a wave per box leaves most lanes idle), off
// for denser data, whose small boxes meet large ones that a wave tests
// faster.
constexpr int kSmallUnionRows = 8;
constexpr int kSmallUnionMean = 8;
int32_t small_union_rows(int64_t n, int64_t nbox);
// only_over (optional): the join runs only when *only_over > cap, i.e. when
// the open-pair lists did not fit (decided on the device, no host check).
template <typename T, int D>
void launch_box_join(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                     const int32_t* bmincore, const uint8_t* core, const int32_t* rsnap,
                     int32_t* bparent, hipStream_t s, const unsigned long long* only_over = nullptr,
                     unsigned long long cap = 0);

// The pairs the hook forest leaves open.  launch_box_scan with list ==
// nullptr writes cnt[b] (nbox entries) and adds their sum to *total (64-bit,
// zeroed by the caller); with a list it writes every box's open partners to
// list[cnt[b] ..] (cnt scanned to offsets in between).
// launch_box_union_list joins the listed pairs that hold a within-eps core
// pair (boxes of at most small_rows rows one thread per box, the items --
// chunks of the other boxes -- one wave each); off has nbox + 1 entries.  When the list does not fit, the caller
// runs launch_box_join instead.  With total_chk (no host check in between),
// both do nothing when *total_chk > cap, and launch_box_join with the same
// check runs exactly then.
template <typename T, int D>
void launch_box_scan(const BoxArgs<T>& a, const uint32_t* roff, const int2* rcells,
                     const int32_t* rsnap, uint32_t* cnt, int32_t* list,
                     unsigned long long* total, hipStream_t s,
                     const unsigned long long* total_chk = nullptr, unsigned long long cap = 0);
template <typename T, int D>
void launch_box_union_list(const BoxArgs<T>& a, const int2* items, Count nitems,
                           const uint8_t* core, const uint32_t* off, const int32_t* list,
                           int32_t* bparent, hipStream_t s,
                           const unsigned long long* total_chk = nullptr,
                           unsigned long long cap = 0, int32_t small_rows = 0);
// p[0, n) = v
void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);

void launch_box_clusters(Count nbox, const int32_t* bmincore, const int32_t* bcell,
                         const int32_t* cbox, const int32_t* bstart, const uint32_t* inrow,
                         const int32_t* ustart, int32_t* bparent, unsigned long long* cmin,
                         int32_t* compmin, int32_t* bcluster, hipStream_t s);

// pos_of_slot[s] = sorted position of slot s (the reference's order: cell,
// then input row), by a stable radix sort of the slots by occupied cell in
// input order; keys_a / vals_a / keys_b / vals_b (n words each), hist and
// tmp are scratch sized as for radix_sort_pairs<uint32_t>
void launch_positions(int64_t n, int32_t m, const int32_t* box_of_slot, const int32_t* bcell,
                      const uint32_t* inrow, uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b,
                      uint32_t* vals_b, uint32_t* hist, uint32_t* tmp, uint32_t* pos_of_slot,
                      hipStream_t s);
}  // namespace hpd
