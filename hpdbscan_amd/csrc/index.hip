// Occupied-cell table and neighbour ranges on the sorted rows.
//
//   k_run_flags        1 where a sorted row starts a new cell
//   k_compact_cells    cell id and first row of every occupied cell
//                      (the reference's CellIndex map, spatial_index.h:167-182)
//   k_neighbor_rows    per occupied cell, the row ranges of its candidate
//                      cells (spatial_index.h:682-737), in CSR form
//   k_work_items       (cell, first row) of every 64-row chunk of a cell
//
// Neighbour enumeration.  The reference visits every cell whose index differs
// by at most one in each dimension and lies inside the grid.  Cells that are
// neighbours along swapped dimension 0 (stride 1) have consecutive ids, so
// the up-to-three cells (i0-1, i0, i0+1) of one "row" occupy one contiguous
// range of sorted rows.  A cell therefore has at most 3^(D-1) ranges, found
// by binary search of the row's first and last cell id in the occupied-cell
// table.  A cell's ranges are stored in ascending row order, or (D <= 4 with
// the count pass's stash) nearest first: the own row, then the rows offset in
// one, two, three dimensions.  The index of the row holding the cell itself
// is kept separately.  When the grid is small enough for a dense table of
// every cell id (ctab: occupied-cell index or -1), a row's cells are three
// table reads instead of a binary search.
#include "boxes.hpp"
#include "common.hpp"
#include "grid.hpp"
#include "index.hpp"
#include "internal.hpp"

namespace hpd {

namespace {

inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

template <typename K>
__global__ __launch_bounds__(256) void k_run_flags(const K* __restrict__ keys, int64_t n,
                                                   int shift, uint32_t* __restrict__ flags) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1u : 0u;
}

template <typename K>
__global__ __launch_bounds__(256) void k_compact_cells(const K* __restrict__ keys, int64_t n,
                                                       int shift,
                                                       const uint32_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ scan,
                                                       K* __restrict__ ucell,
                                                       int32_t* __restrict__ ustart,
                                                       const uint32_t* __restrict__ ncells_dev) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n && flags[i]) {
        const uint32_t u = scan[i];
        ucell[u] = keys[i] >> shift;
        ustart[u] = static_cast<int32_t>(i);
    }
    if (i == 0) ustart[*ncells_dev] = static_cast<int32_t>(n);
}

// first index u in [0, m) with ucell[u] >= key (m if none)
template <typename K>
__device__ __forceinline__ int32_t lower_bound_cells(const K* __restrict__ ucell, int32_t m,
                                                     uint64_t key, int32_t lo = 0) {
    int32_t hi = m;
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (static_cast<uint64_t>(ucell[mid]) < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Occupied cells [a, b) with ids in [lo, hi] (at most three consecutive
// ids): a by binary search from `from` (rows are visited in ascending
// order, so the previous row's end is a lower bound), b by a short scan.
template <typename K>
__device__ __forceinline__ int2 row_cells(const K* __restrict__ ucell, int32_t m, uint64_t lo,
                                          uint64_t hi, int32_t from) {
    const int32_t a = lower_bound_cells(ucell, m, lo, from);
    int32_t b = a;
    while (b < m && static_cast<uint64_t>(ucell[b]) <= hi) ++b;
    return make_int2(a, b);
}

// The same from the dense cell table: the occupied ones among the (at most
// three) ids lo..hi, which are consecutive in the occupied-cell order.
__device__ __forceinline__ int2 row_cells_tab(const int32_t* __restrict__ ctab, uint64_t lo,
                                              uint64_t hi) {
    int32_t a = INT32_MAX, b = -1;
    for (uint64_t c = lo; c <= hi; ++c) {
        const int32_t u = ctab[c];
        if (u >= 0) {
            a = min(a, u);
            b = max(b, u);
        }
    }
    return b >= 0 ? make_int2(a, b + 1) : make_int2(0, 0);
}

// ctab[id] = u for every occupied cell u (the table filled with -1 before)
template <typename K>
__global__ __launch_bounds__(256) void k_cell_table(const K* __restrict__ ucell, Count mc,
                                                    int32_t* __restrict__ ctab) {
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u < static_cast<int32_t>(mc.get())) ctab[static_cast<uint64_t>(ucell[u])] = u;
}

// Decode a cell id into per-swapped-dimension indices.
__device__ __forceinline__ void decode_cell(uint64_t key, const CellParams& cp, uint64_t* idx) {
#pragma unroll
    for (int k = 0; k < kMaxDims; ++k) {
        if (k < cp.d) {
            idx[k] = (key / cp.stride[k]) % cp.ncell[k];
        }
    }
}

// Visit the neighbour rows of a cell in ascending cell-id order.
// fn(row_lo_key, row_hi_key, is_center, nz) gets the inclusive cell-id
// bounds of each row that lies inside the grid and the number of dimensions
// 1..D-1 in which the row is offset from the own cell (0 for the own row).  Odometer order over the offsets of
// swapped dimensions 1..D-1, most significant dimension slowest, is ascending
// because the offsets of the lower dimensions never reach one stride of the
// next.
template <typename F>
__device__ __forceinline__ void for_each_row(uint64_t key, const CellParams& cp, F&& fn) {
    uint64_t idx[kMaxDims];
    decode_cell(key, cp, idx);
    const uint64_t lo0 = idx[0] > 0 ? 1 : 0;
    const uint64_t hi0 = idx[0] + 1 < cp.ncell[0] ? 1 : 0;
    int nrows = 1;
    for (int k = 1; k < cp.d; ++k) nrows *= 3;
    const int center_row = (nrows - 1) / 2;  // every digit 1: offset 0 in all dims
    for (int r = 0; r < nrows; ++r) {
        // digit k-1 of r in base 3 is the offset (0 -> -1, 1 -> 0, 2 -> +1)
        int rem = r, nz = 0;
        bool inside = true;
        uint64_t base = key;
        for (int k = 1; k < cp.d; ++k) {
            const int dgt = rem % 3;
            rem /= 3;
            if (dgt == 0) {
                if (idx[k] == 0) inside = false;
                base -= cp.stride[k];
                ++nz;
            } else if (dgt == 2) {
                if (idx[k] + 1 >= cp.ncell[k]) inside = false;
                base += cp.stride[k];
                ++nz;
            }
        }
        if (!inside) continue;
        fn(base - lo0, base + hi0, r == center_row, nz);
    }
}

// Job of an occupied cell: the last j with key_base <= key.  Jobs without
// rows get the key_base of the job after them, so that job is never one of
// them.
template <typename T>
__device__ __forceinline__ int32_t job_of_key(const JobParams<T>* __restrict__ jobs, int32_t njobs,
                                              uint64_t key) {
    int32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].key_base <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename K, typename T>
__global__ __launch_bounds__(256) void k_neighbor_rows_count(const K* __restrict__ ucell,
                                                             const int32_t* __restrict__ ustart,
                                                             Count mc,
                                                             const JobParams<T>* __restrict__ jobs,
                                                             int32_t njobs,
                                                             int32_t* __restrict__ job_of_cell,
                                                             uint32_t* __restrict__ nranges,
                                                             uint32_t* __restrict__ core_chunks,
                                                             unsigned long long* __restrict__ sumcand,
                                                             unsigned long long* __restrict__ rows_total,
                                                             int2* __restrict__ stash,
                                                             int32_t stash_rows,
                                                             int32_t* __restrict__ center_local,
                                                             bool by_class,
                                                             const int32_t* __restrict__ ctab) {
    const int32_t m = static_cast<int32_t>(mc.get());
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    uint64_t cand = 0;
    uint64_t count_u = 0;
    uint32_t nranges_u = 0;
    if (u < m) {
        const uint64_t key = static_cast<uint64_t>(ucell[u]);
        const int32_t j = job_of_key(jobs, njobs, key);
        job_of_cell[u] = j;
        const uint64_t base = jobs[j].key_base;
        uint32_t cnt = 0;
        int32_t from = 0, center = 0;
        int2* mine = stash ? stash + static_cast<int64_t>(u) * stash_rows : nullptr;
        for_each_row(key - base, jobs[j].cp, [&](uint64_t lo, uint64_t hi, bool is_center, int nz) {
            const int2 ab = ctab ? row_cells_tab(ctab, base + lo, base + hi)
                                 : row_cells(ucell, m, base + lo, base + hi, from);
            if (!ctab) from = ab.y;
            if (ab.y > ab.x) {
                if (mine) {
                    if (is_center) center = static_cast<int32_t>(cnt);
                    // with by_class, the row's offset class (nz <= 3) rides in
                    // the top bits of its cell indices (both below 2^31)
                    mine[cnt] = by_class ? make_int2(static_cast<int32_t>(
                                                         static_cast<uint32_t>(ab.x) |
                                                         (static_cast<uint32_t>(nz & 1) << 31)),
                                                     static_cast<int32_t>(
                                                         static_cast<uint32_t>(ab.y) |
                                                         (static_cast<uint32_t>((nz >> 1) & 1) << 31)))
                                         : ab;
                }
                ++cnt;
                cand += static_cast<uint64_t>(ustart[ab.y] - ustart[ab.x]);
            }
        });
        nranges[u] = cnt;
        nranges_u = cnt;
        if (mine) center_local[u] = center;
        count_u = static_cast<uint64_t>(ustart[u + 1] - ustart[u]);
        // the reference queries a row only with >= minPts candidates
        // (hpdbscan.h:74); with fewer, no row of the cell can be core
        if (core_chunks)
            core_chunks[u] = cand >= jobs[j].min_pts ? static_cast<uint32_t>((count_u + kWave - 1) / kWave)
                                                     : 0u;
    }
    // *sumcand += Σ over rows of the cell of |candidates| = count_u * cand;
    // *rows_total += the cell's neighbour rows (their 64-bit total is checked
    // on the host before the 32-bit offsets are used)
    block_add2<256>(static_cast<unsigned long long>(cand * count_u),
                    static_cast<unsigned long long>(nranges_u), sumcand, rows_total);
}

template <typename K, typename T>
__global__ __launch_bounds__(256) void k_neighbor_rows_fill(const K* __restrict__ ucell,
                                                            const int32_t* __restrict__ ustart,
                                                            Count mc,
                                                            const JobParams<T>* __restrict__ jobs,
                                                            const int32_t* __restrict__ job_of_cell,
                                                            const uint32_t* __restrict__ roff,
                                                            int2* __restrict__ ranges,
                                                            int2* __restrict__ rcells,
                                                            int32_t* __restrict__ rcenter,
                                                            const int32_t* __restrict__ ctab) {
    const int32_t m = static_cast<int32_t>(mc.get());
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= m) return;
    const int32_t j = job_of_cell[u];
    const uint64_t base = jobs[j].key_base;
    uint32_t w = roff[u];
    int32_t center = static_cast<int32_t>(w);
    int32_t from = 0;
    for_each_row(static_cast<uint64_t>(ucell[u]) - base, jobs[j].cp,
                 [&](uint64_t lo, uint64_t hi, bool is_center, int) {
                     const int2 ab = ctab ? row_cells_tab(ctab, base + lo, base + hi)
                                          : row_cells(ucell, m, base + lo, base + hi, from);
                     if (!ctab) from = ab.y;
                     if (ab.y > ab.x) {
                         if (is_center) center = static_cast<int32_t>(w);
                         rcells[w] = ab;
                         ranges[w++] = make_int2(ustart[ab.x], ustart[ab.y]);
                     }
                 });
    rcenter[u] = center;
}

// The fill from the count pass's stash: the same ranges, rcells and rcenter
// as k_neighbor_rows_fill, without searching again.
__global__ __launch_bounds__(256) void k_neighbor_rows_copy(const int2* __restrict__ stash,
                                                            int32_t stash_rows,
                                                            const int32_t* __restrict__ ustart,
                                                            int32_t m,
                                                            const uint32_t* __restrict__ roff,
                                                            int2* __restrict__ ranges,
                                                            int2* __restrict__ rcells,
                                                            int32_t* rcenter, bool by_class) {
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= m) return;
    const uint32_t w0 = roff[u], cnt = roff[u + 1] - w0;
    const int2* mine = stash + static_cast<int64_t>(u) * stash_rows;
    if (!by_class) {
        for (uint32_t k = 0; k < cnt; ++k) {
            const int2 ab = mine[k];
            rcells[w0 + k] = ab;
            ranges[w0 + k] = make_int2(ustart[ab.x], ustart[ab.y]);
        }
        rcenter[u] = static_cast<int32_t>(w0) + rcenter[u];  // local index from the count pass
        return;
    }
    // nearest rows first: the own row, then rows offset in one dimension,
    // then in two, then in three (a visit that stops early sees the rows
    // with the most within-eps neighbours first)
    uint32_t w = w0;
    for (uint32_t c = 0; c < 4; ++c) {
        for (uint32_t k = 0; k < cnt; ++k) {
            const int2 e = mine[k];
            const uint32_t x = static_cast<uint32_t>(e.x), y = static_cast<uint32_t>(e.y);
            if (((x >> 31) | ((y >> 31) << 1)) != c) continue;
            const int2 ab = make_int2(static_cast<int32_t>(x & 0x7fffffffu),
                                      static_cast<int32_t>(y & 0x7fffffffu));
            rcells[w] = ab;
            ranges[w] = make_int2(ustart[ab.x], ustart[ab.y]);
            ++w;
        }
    }
    rcenter[u] = static_cast<int32_t>(w0);  // the own row is the only one of class 0
}

// The by-class copy with at most R rows per cell (R = 3^(D-1) <= 27): the
// rows are read once into registers, counted per class, and written at
// their class's place.
template <int R>
__global__ __launch_bounds__(256) void k_neighbor_rows_copy_class(const int2* __restrict__ stash,
                                                                  const int32_t* __restrict__ ustart,
                                                                  int32_t m,
                                                                  const uint32_t* __restrict__ roff,
                                                                  int2* __restrict__ ranges,
                                                                  int2* __restrict__ rcells,
                                                                  int32_t* __restrict__ rcenter) {
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= m) return;
    const uint32_t w0 = roff[u], cnt = roff[u + 1] - w0;
    const int2* mine = stash + static_cast<int64_t>(u) * R;
    int2 e[R];
    uint32_t cls[R];
    uint32_t n1 = 0, n2 = 0;  // rows of class 1 and 2 (class 0 is the own row alone)
#pragma unroll
    for (int k = 0; k < R; ++k) {
        cls[k] = 4;
        if (static_cast<uint32_t>(k) < cnt) {
            const int2 v = mine[k];
            const uint32_t x = static_cast<uint32_t>(v.x), y = static_cast<uint32_t>(v.y);
            cls[k] = (x >> 31) | ((y >> 31) << 1);
            e[k] = make_int2(static_cast<int32_t>(x & 0x7fffffffu),
                             static_cast<int32_t>(y & 0x7fffffffu));
            n1 += cls[k] == 1 ? 1u : 0u;
            n2 += cls[k] == 2 ? 1u : 0u;
        }
    }
    uint32_t p0 = w0, p1 = w0 + 1, p2 = w0 + 1 + n1, p3 = w0 + 1 + n1 + n2;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (cls[k] > 3) continue;
        const uint32_t c = cls[k];
        const uint32_t w = c == 0 ? p0++ : c == 1 ? p1++ : c == 2 ? p2++ : p3++;
        rcells[w] = e[k];
        ranges[w] = make_int2(ustart[e[k].x], ustart[e[k].y]);
    }
    rcenter[u] = static_cast<int32_t>(w0);  // the own row is the only one of class 0
}

__global__ __launch_bounds__(256) void k_chunk_counts(const int32_t* __restrict__ ustart, Count mc,
                                                      uint32_t* __restrict__ nchunks) {
    const int32_t m = static_cast<int32_t>(mc.get());
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= m) return;
    const int32_t c = ustart[u + 1] - ustart[u];
    nchunks[u] = static_cast<uint32_t>((c + kWave - 1) / kWave);
}

// items[w] = (group, first row) of the w-th 64-row chunk, one thread per
// chunk: the group is the last one whose chunk offset is <= w (binary search
// over the exclusive scan off[0 .. ngroups]; empty groups are skipped)
__global__ __launch_bounds__(256) void k_expand_chunks(const uint32_t* __restrict__ off,
                                                       Count ngroups_c,
                                                       const int32_t* __restrict__ start,
                                                       Count total_c, int2* __restrict__ items) {
    const int32_t ngroups = static_cast<int32_t>(ngroups_c.get());
    const int32_t total = static_cast<int32_t>(total_c.get());
    const int32_t w = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (w >= total) return;
    int32_t lo = 0, hi = ngroups;  // off[lo] <= w < off[hi]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (off[mid] <= static_cast<uint32_t>(w)) lo = mid;
        else hi = mid;
    }
    items[w] = make_int2(lo, start[lo] + kWave * (w - static_cast<int32_t>(off[lo])));
}

// Per cell: 64-row chunks if the cell has more than `sparse_max` rows
// (dense), else its rows one by one (sparse).
__global__ __launch_bounds__(256) void k_split_counts(const int32_t* __restrict__ ustart, Count mc,
                                                      int32_t sparse_max,
                                                      uint32_t* __restrict__ dcnt,
                                                      uint32_t* __restrict__ scnt) {
    const int32_t m = static_cast<int32_t>(mc.get());
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= m) return;
    const int32_t c = ustart[u + 1] - ustart[u];
    const bool sparse = c <= sparse_max;
    dcnt[u] = sparse ? 0u : static_cast<uint32_t>((c + kWave - 1) / kWave);
    scnt[u] = sparse ? static_cast<uint32_t>(c) : 0u;
}

__global__ __launch_bounds__(256) void k_split_items(const int32_t* __restrict__ ustart, Count mc,
                                                     int32_t sparse_max,
                                                     const uint32_t* __restrict__ doff,
                                                     const uint32_t* __restrict__ soff,
                                                     int2* __restrict__ ditems,
                                                     int2* __restrict__ sitems) {
    const int32_t m = static_cast<int32_t>(mc.get());
    const int32_t u = static_cast<int32_t>(blockIdx.x) * 256 + threadIdx.x;
    if (u >= m) return;
    const int32_t s = ustart[u], e = ustart[u + 1];
    if (e - s <= sparse_max) {
        uint32_t w = soff[u];
        for (int32_t p = s; p < e; ++p) sitems[w++] = make_int2(p, u);
    }
    (void)doff;
    (void)ditems;
}

}  // namespace

void launch_split_counts(const int32_t* ustart, Count m, int32_t sparse_max, uint32_t* dcnt,
                         uint32_t* scnt, hipStream_t s) {
    if (m.host <= 0) return;
    hipLaunchKernelGGL(k_split_counts, dim3(blocks_for(m.host, 256)), dim3(256), 0, s, ustart, m,
                       sparse_max, dcnt, scnt);
    HPD_HIP(hipGetLastError());
}

void launch_split_items(const int32_t* ustart, Count m, int32_t sparse_max, const uint32_t* doff,
                        const uint32_t* soff, int2* ditems, Count nditems, int2* sitems,
                        hipStream_t s) {
    if (m.host <= 0) return;
    hipLaunchKernelGGL(k_split_items, dim3(blocks_for(m.host, 256)), dim3(256), 0, s, ustart, m,
                       sparse_max, doff, soff, ditems, sitems);
    launch_expand_chunks(doff, m, ustart, nditems, ditems, s);
    HPD_HIP(hipGetLastError());
}

void launch_expand_chunks(const uint32_t* off, Count ngroups, const int32_t* start, Count total,
                          int2* items, hipStream_t s) {
    if (total.host <= 0) return;
    hipLaunchKernelGGL(k_expand_chunks, dim3(blocks_for(total.host, 256)), dim3(256), 0, s, off,
                       ngroups, start, total, items);
    HPD_HIP(hipGetLastError());
}

template <typename K>
void launch_cells(const K* keys, int64_t n, int shift, uint32_t* flags, uint32_t* scan_tmp,
                  K* ucell, int32_t* ustart, uint32_t* ncells_dev, hipStream_t s) {
    hipLaunchKernelGGL(k_run_flags<K>, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, shift,
                       flags);
    exclusive_scan_u32(flags, flags + n, n, scan_tmp, ncells_dev, s);
    hipLaunchKernelGGL(k_compact_cells<K>, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n,
                       shift, static_cast<const uint32_t*>(flags),
                       static_cast<const uint32_t*>(flags + n), ucell, ustart,
                       static_cast<const uint32_t*>(ncells_dev));
    HPD_HIP(hipGetLastError());
}

template <typename K, typename T>
void launch_neighbor_count(const K* ucell, const int32_t* ustart, Count m,
                           const JobParams<T>* jobs, int32_t njobs, int32_t* job_of_cell,
                           uint32_t* nranges, uint32_t* core_chunks, unsigned long long* sumcand,
                           unsigned long long* rows_total, int2* stash, int32_t stash_rows,
                           int32_t* rcenter, bool by_class, hipStream_t s, const int32_t* ctab) {
    if (m.host <= 0) return;
    hipLaunchKernelGGL((k_neighbor_rows_count<K, T>), dim3(blocks_for(m.host, 256)), dim3(256), 0, s,
                       ucell, ustart, m, jobs, njobs, job_of_cell, nranges, core_chunks, sumcand,
                       rows_total, stash, stash_rows, rcenter, by_class, ctab);
    HPD_HIP(hipGetLastError());
}

template <typename K>
void launch_cell_table(const K* ucell, Count m, int32_t* ctab, uint64_t ids, hipStream_t s) {
    if (m.host <= 0) return;
    fill_i32(ctab, -1, static_cast<int64_t>(ids), s);
    hipLaunchKernelGGL(k_cell_table<K>, dim3(blocks_for(m.host, 256)), dim3(256), 0, s, ucell, m,
                       ctab);
    HPD_HIP(hipGetLastError());
}

void launch_neighbor_copy(const int2* stash, int32_t stash_rows, const int32_t* ustart, int32_t m,
                          const uint32_t* roff, int2* ranges, int2* rcells, int32_t* rcenter,
                          bool by_class, hipStream_t s) {
    const dim3 grid(blocks_for(m, 256));
    if (by_class && stash_rows == 9)
        hipLaunchKernelGGL(k_neighbor_rows_copy_class<9>, grid, dim3(256), 0, s, stash, ustart, m,
                           roff, ranges, rcells, rcenter);
    else if (by_class && stash_rows == 27)
        hipLaunchKernelGGL(k_neighbor_rows_copy_class<27>, grid, dim3(256), 0, s, stash, ustart, m,
                           roff, ranges, rcells, rcenter);
    else
        hipLaunchKernelGGL(k_neighbor_rows_copy, grid, dim3(256), 0, s, stash, stash_rows, ustart, m,
                           roff, ranges, rcells, rcenter, by_class);
    HPD_HIP(hipGetLastError());
}

template <typename K, typename T>
void launch_neighbor_fill(const K* ucell, const int32_t* ustart, Count m,
                          const JobParams<T>* jobs, const int32_t* job_of_cell,
                          const uint32_t* roff, int2* ranges, int2* rcells, int32_t* rcenter,
                          hipStream_t s, const int32_t* ctab) {
    if (m.host <= 0) return;
    hipLaunchKernelGGL((k_neighbor_rows_fill<K, T>), dim3(blocks_for(m.host, 256)), dim3(256), 0, s,
                       ucell, ustart, m, jobs, job_of_cell, roff, ranges, rcells, rcenter, ctab);
    HPD_HIP(hipGetLastError());
}

void launch_chunk_counts(const int32_t* ustart, Count m, uint32_t* nchunks, hipStream_t s) {
    if (m.host <= 0) return;
    hipLaunchKernelGGL(k_chunk_counts, dim3(blocks_for(m.host, 256)), dim3(256), 0, s, ustart, m,
                       nchunks);
    HPD_HIP(hipGetLastError());
}

void launch_work_items(const int32_t* ustart, Count m, const uint32_t* coff, Count nitems,
                       int2* items, hipStream_t s) {
    launch_expand_chunks(coff, m, ustart, nitems, items, s);
}

#define HPD_INST_INDEX(K)                                                                     \
    template void launch_cells<K>(const K*, int64_t, int, uint32_t*, uint32_t*, K*, int32_t*, \
                                  uint32_t*, hipStream_t);                                    \
    template void launch_cell_table<K>(const K*, Count, int32_t*, uint64_t, hipStream_t);
HPD_INST_INDEX(uint32_t)
HPD_INST_INDEX(uint64_t)

#define HPD_INST_ROWS(T)                                                                        \
    template void launch_neighbor_count<uint32_t, T>(const uint32_t*, const int32_t*, Count,    \
                                                     const JobParams<T>*, int32_t, int32_t*,    \
                                                     uint32_t*, uint32_t*, unsigned long long*, \
                                                     unsigned long long*, int2*, int32_t,       \
                                                     int32_t*, bool, hipStream_t,               \
                                                     const int32_t*);                           \
    template void launch_neighbor_count<uint64_t, T>(const uint64_t*, const int32_t*, Count,    \
                                                     const JobParams<T>*, int32_t, int32_t*,    \
                                                     uint32_t*, uint32_t*, unsigned long long*, \
                                                     unsigned long long*, int2*, int32_t,       \
                                                     int32_t*, bool, hipStream_t,               \
                                                     const int32_t*);                           \
    template void launch_neighbor_fill<uint32_t, T>(const uint32_t*, const int32_t*, Count,     \
                                                    const JobParams<T>*, const int32_t*,        \
                                                    const uint32_t*, int2*, int2*, int32_t*,    \
                                                    hipStream_t, const int32_t*);               \
    template void launch_neighbor_fill<uint64_t, T>(const uint64_t*, const int32_t*, Count,     \
                                                    const JobParams<T>*, const int32_t*,        \
                                                    const uint32_t*, int2*, int2*, int32_t*,    \
                                                    hipStream_t, const int32_t*);
HPD_FOR_EACH_DTYPE(HPD_INST_ROWS)

}  // namespace hpd
