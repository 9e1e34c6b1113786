// Launchers of the occupied-cell and neighbour-range kernels (index.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace hpd {

// Cells of sorted keys whose cell id is key >> shift.  flags must hold
// 2*n + 1 uint32 (run flags, then their exclusive scan); writes ucell[0..m),
// ustart[0..m] with ustart[m] = n, and m to *ncells_dev.
template <typename K>
void launch_cells(const K* keys, int64_t n, int shift, uint32_t* flags, uint32_t* scan_tmp,
                  K* ucell, int32_t* ustart, uint32_t* ncells_dev, hipStream_t s);

// job_of_cell[u] = job of cell u; nranges[u] = number of non-empty
// neighbour rows of cell u; core_chunks[u] (when not null) = 64-row chunks
// of cell u if it has >= minPts candidates, else 0; adds
// Σ_u count(u) * |candidates(u)| to *sumcand and Σ_u nranges[u] to
// *rows_total (64-bit, unlike the scan of nranges).  With a stash (m * stash_rows
// entries, stash_rows >= the most neighbour rows a cell can have) the
// rows' cell ranges go to stash[u * stash_rows ..] and the own row's index
// among them to rcenter[u], for launch_neighbor_copy; with by_class (D <= 4)
// each row's offset class (the number of dimensions it is offset in) rides
// in the top bits of its two cell indices.
template <typename K, typename T>
void launch_neighbor_count(const K* ucell, const int32_t* ustart, Count m,
                           const JobParams<T>* jobs, int32_t njobs, int32_t* job_of_cell,
                           uint32_t* nranges, uint32_t* core_chunks, unsigned long long* sumcand,
                           unsigned long long* rows_total, int2* stash, int32_t stash_rows,
                           int32_t* rcenter, bool by_class, hipStream_t s,
                           const int32_t* ctab = nullptr);

// Dense table of every cell id below `ids`: ctab[id] = its occupied-cell
// index, -1 for an empty cell (ids int32 entries).  With it the neighbour
// kernels read a row's cells from the table instead of searching.
template <typename K>
void launch_cell_table(const K* ucell, Count m, int32_t* ctab, uint64_t ids, hipStream_t s);

// launch_neighbor_fill's output from the count pass's stash (roff scanned);
// with by_class the rows of each cell go nearest first (own row, then by
// the number of offset dimensions) instead of ascending
void launch_neighbor_copy(const int2* stash, int32_t stash_rows, const int32_t* ustart, int32_t m,
                          const uint32_t* roff, int2* ranges, int2* rcells, int32_t* rcenter,
                          bool by_class, hipStream_t s);

// ranges[roff[u] ..) = [first row, end row) of each non-empty neighbour row,
// ascending; rcells[..] = [first, end) occupied-cell index of the same row;
// rcenter[u] = index (into ranges) of the row holding cell u itself
template <typename K, typename T>
void launch_neighbor_fill(const K* ucell, const int32_t* ustart, Count m,
                          const JobParams<T>* jobs, const int32_t* job_of_cell,
                          const uint32_t* roff, int2* ranges, int2* rcells, int32_t* rcenter,
                          hipStream_t s, const int32_t* ctab = nullptr);

void launch_chunk_counts(const int32_t* ustart, Count m, uint32_t* nchunks, hipStream_t s);

// Work split of the neighbour kernels: cells with more than sparse_max rows
// give one item (cell, first row) per 64-row chunk (dcnt: chunks per cell),
// the others one item (row, cell) per row (scnt: rows per cell).  doff /
// soff are the exclusive scans of dcnt / scnt.
void launch_split_counts(const int32_t* ustart, Count m, int32_t sparse_max, uint32_t* dcnt,
                         uint32_t* scnt, hipStream_t s);
void launch_split_items(const int32_t* ustart, Count m, int32_t sparse_max, const uint32_t* doff,
                        const uint32_t* soff, int2* ditems, Count nditems, int2* sitems,
                        hipStream_t s);

// items[w] = (g, start[g] + 64 * (w - off[g])) for the `total` 64-row chunks
// of ngroups groups (off: exclusive scan of chunk counts, ngroups + 1 entries)
void launch_expand_chunks(const uint32_t* off, Count ngroups, const int32_t* start, Count total,
                          int2* items, hipStream_t s);
void launch_work_items(const int32_t* ustart, Count m, const uint32_t* coff, Count nitems,
                       int2* items, hipStream_t s);

}  // namespace hpd
