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
// so the kernels compute exactly that:
//   k_core      count within-eps candidates per row, stop at minPts
//   k_union     union-find over core-core pairs (hook larger root under the
//               smaller, so a root is the smallest position)
//   k_flatten   root of every core; writes -(root+1) for core rows
//   k_border    smallest adjacent root for non-core rows (0 = noise)
//
// Work decomposition: one wave per 64-row chunk of an occupied cell; lanes
// own the chunk's rows, the candidate rows stream through the wave with
// wave-uniform addresses (the hardware broadcasts them), in the order of the
// cell's neighbour ranges.
#include <climits>

#include "common.hpp"
#include "dbscan.hpp"
#include "grid.hpp"
#include "internal.hpp"

namespace hpd {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;

template <typename T, int D>
__device__ __forceinline__ void load_row(const T* __restrict__ soa, int64_t ld, int64_t i, T* c) {
#pragma unroll
    for (int j = 0; j < D; ++j) c[j] = soa[static_cast<int64_t>(j) * ld + i];
}

// Candidates are visited in aligned batches of kBatch<T> rows.  Every batch
// is fetched with wave-uniform addresses before it is used, which lets the
// compiler issue them as a few wide scalar loads and wait once.
template <typename T>
constexpr int kBatch = sizeof(T) == 8 ? 8 : 16;

template <typename U, int B>
struct alignas(B * sizeof(U) >= 4 ? B * sizeof(U) : 4) Batch {
    U v[B];
};

template <typename U, int B>
__device__ __forceinline__ Batch<U, B> load_batch(const U* __restrict__ base, int64_t a) {
    return *reinterpret_cast<const Batch<U, B>*>(base + a);
}

// Visit candidate rows [qb, qe).  body(q, qc, core_q) runs per candidate
// (core_q is the candidate's core flag when `flags` is given, else 1);
// stop() is checked after each batch and ends the visit when true.  Returns
// true if stopped.
template <typename T, int D, typename Body, typename Stop>
__device__ __forceinline__ bool visit_range(const T* __restrict__ soa, int64_t ld,
                                            const uint8_t* __restrict__ flags, int32_t qb,
                                            int32_t qe, Body&& body, Stop&& stop) {
    constexpr int B = kBatch<T>;
    for (int32_t a = qb & ~(B - 1); a < qe; a += B) {
        Batch<T, B> b[D];
#pragma unroll
        for (int j = 0; j < D; ++j) b[j] = load_batch<T, B>(soa + static_cast<int64_t>(j) * ld, a);
        Batch<uint8_t, B> f;
        if (flags) f = load_batch<uint8_t, B>(flags, a);
        if (a >= qb && a + B <= qe) {
#pragma unroll
            for (int t = 0; t < B; ++t) {
                T qc[D];
#pragma unroll
                for (int j = 0; j < D; ++j) qc[j] = b[j].v[t];
                body(a + t, qc, flags ? f.v[t] : uint8_t(1));
            }
        } else {
            const int lo = qb > a ? qb - a : 0;
            const int hi = qe - a < B ? qe - a : B;
#pragma unroll
            for (int t = 0; t < B; ++t) {
                if (t < lo || t >= hi) continue;
                T qc[D];
#pragma unroll
                for (int j = 0; j < D; ++j) qc[j] = b[j].v[t];
                body(a + t, qc, flags ? f.v[t] : uint8_t(1));
            }
        }
        if (stop()) return true;
    }
    return false;
}

struct Chunk {
    int32_t u;      // occupied cell
    int32_t first;  // first row of the chunk
    int32_t end;    // end row of the cell
    int32_t r0, r1; // its neighbour ranges
};

__device__ __forceinline__ bool fetch_chunk(const int2* __restrict__ items, int32_t nitems,
                                            const int32_t* __restrict__ ustart,
                                            const uint32_t* __restrict__ roff, Chunk& c) {
    const int32_t wid =
        uniform(static_cast<int32_t>(blockIdx.x) * kWavesPerBlock + static_cast<int32_t>(threadIdx.x >> 6));
    if (wid >= nitems) return false;
    const int2 it = items[wid];
    c.u = uniform(it.x);
    c.first = uniform(it.y);
    c.end = uniform(ustart[c.u + 1]);
    c.r0 = uniform(static_cast<int32_t>(roff[c.u]));
    c.r1 = uniform(static_cast<int32_t>(roff[c.u + 1]));
    return true;
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_core(const T* __restrict__ soa, int64_t ld,
                                                   const int2* __restrict__ items, int32_t nitems,
                                                   const int32_t* __restrict__ ustart,
                                                   const uint32_t* __restrict__ roff,
                                                   const int2* __restrict__ ranges, T eps2,
                                                   uint32_t min_pts, uint8_t* __restrict__ core,
                                                   unsigned long long* __restrict__ ncore) {
    Chunk c;
    if (!fetch_chunk(items, nitems, ustart, roff, c)) return;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    T pc[D];
    load_row<T, D>(soa, ld, valid ? p : c.first, pc);
    uint32_t cnt = 0;
    auto body = [&](int32_t, const T* qc, uint8_t) { cnt += within<D>(pc, qc, eps2) ? 1u : 0u; };
    auto stop = [&] { return static_cast<bool>(__all(!valid || cnt >= min_pts)); };
    bool done = stop();
    for (int32_t r = c.r0; r < c.r1 && !done; ++r) {
        const int2 rg = ranges[r];
        done = visit_range<T, D>(soa, ld, nullptr, uniform(rg.x), uniform(rg.y), body, stop);
    }
    const bool is_core = valid && cnt >= min_pts;
    if (valid) core[p] = is_core ? 1 : 0;
    const unsigned long long nc = __popcll(ballot(is_core));
    if (lane_id() == 0 && nc) atomicAdd(ncore, nc);
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_union(const T* __restrict__ soa, int64_t ld,
                                                    const int2* __restrict__ items, int32_t nitems,
                                                    const int32_t* __restrict__ ustart,
                                                    const uint32_t* __restrict__ roff,
                                                    const int2* __restrict__ ranges, T eps2,
                                                    const uint8_t* __restrict__ core,
                                                    int32_t* parent) {
    Chunk c;
    if (!fetch_chunk(items, nitems, ustart, roff, c)) return;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && core[p];
    if (!__any(active)) return;
    T pc[D];
    load_row<T, D>(soa, ld, valid ? p : c.first, pc);
    // every core pair is handled by its larger position: only q < p
    const int32_t pmax = uniform(min(c.first + kWave, c.end) - 1);
    auto body = [&](int32_t q, const T* qc, uint8_t cq) {
        if (!cq) return;
        const bool hit = active && q < p && within<D>(pc, qc, eps2);
        if (hit) uf_unite(parent, p, q);
    };
    auto never = [] { return false; };
    for (int32_t r = c.r0; r < c.r1; ++r) {
        const int2 rg = ranges[r];
        const int32_t qb = uniform(rg.x);
        const int32_t qe = uniform(min(rg.y, pmax));
        if (qb < qe) visit_range<T, D>(soa, ld, core, qb, qe, body, never);
    }
}

__global__ __launch_bounds__(kThreads) void k_parent_init(int32_t* __restrict__ parent, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) parent[i] = static_cast<int32_t>(i);
}

__global__ __launch_bounds__(kThreads) void k_flatten(const uint8_t* __restrict__ core,
                                                      int32_t* parent, int64_t n,
                                                      const uint32_t* __restrict__ order,
                                                      int64_t* __restrict__ labels) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n || !core[i]) return;
    const int32_t r = uf_find(parent, static_cast<int32_t>(i));
    parent[i] = r;
    labels[order[i]] = -(static_cast<int64_t>(r) + 1);
}

template <typename T, int D>
__global__ __launch_bounds__(kThreads) void k_border(const T* __restrict__ soa, int64_t ld,
                                                     const int2* __restrict__ items, int32_t nitems,
                                                     const int32_t* __restrict__ ustart,
                                                     const uint32_t* __restrict__ roff,
                                                     const int2* __restrict__ ranges, T eps2,
                                                     const uint8_t* __restrict__ core,
                                                     const int32_t* __restrict__ root,
                                                     const uint32_t* __restrict__ order,
                                                     int64_t* __restrict__ labels,
                                                     unsigned long long* __restrict__ stats) {
    Chunk c;
    if (!fetch_chunk(items, nitems, ustart, roff, c)) return;
    const int32_t p = c.first + lane_id();
    const bool valid = p < c.end;
    const bool active = valid && !core[p];
    if (!__any(active)) return;
    T pc[D];
    load_row<T, D>(soa, ld, valid ? p : c.first, pc);
    int32_t best = INT_MAX, worst = -1;
    auto body = [&](int32_t q, const T* qc, uint8_t cq) {
        if (!cq) return;
        if (active && within<D>(pc, qc, eps2)) {
            const int32_t rq = root[q];
            best = min(best, rq);
            worst = max(worst, rq);
        }
    };
    auto never = [] { return false; };
    for (int32_t r = c.r0; r < c.r1; ++r) {
        const int2 rg = ranges[r];
        visit_range<T, D>(soa, ld, core, uniform(rg.x), uniform(rg.y), body, never);
    }
    if (active) labels[order[p]] = best == INT_MAX ? 0 : static_cast<int64_t>(best) + 1;
    const unsigned long long noise = __popcll(ballot(active && best == INT_MAX));
    const unsigned long long amb = __popcll(ballot(active && best != INT_MAX && best != worst));
    if (lane_id() == 0) {
        if (noise) atomicAdd(&stats[0], noise);
        if (amb) atomicAdd(&stats[1], amb);
    }
}

inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

template <typename T, int D>
void launch_core(const DbscanArgs<T>& a, uint8_t* core, unsigned long long* ncore, hipStream_t s) {
    if (a.nitems <= 0) return;
    hipLaunchKernelGGL((k_core<T, D>), dim3(blocks_for(a.nitems, kWavesPerBlock)), dim3(kThreads),
                       0, s, a.soa, a.ld, a.items, a.nitems, a.ustart, a.roff, a.ranges, a.eps2,
                       a.min_pts, core, ncore);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_union(const DbscanArgs<T>& a, const uint8_t* core, int32_t* parent, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_parent_init, dim3(blocks_for(a.n, kThreads)), dim3(kThreads), 0, s, parent,
                       a.n);
    if (a.nitems > 0)
        hipLaunchKernelGGL((k_union<T, D>), dim3(blocks_for(a.nitems, kWavesPerBlock)),
                           dim3(kThreads), 0, s, a.soa, a.ld, a.items, a.nitems, a.ustart, a.roff,
                           a.ranges, a.eps2, core, parent);
    HPD_HIP(hipGetLastError());
}

template <typename T, int D>
void launch_labels(const DbscanArgs<T>& a, const uint8_t* core, int32_t* parent,
                   const uint32_t* order, int64_t* labels, unsigned long long* stats,
                   hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_flatten, dim3(blocks_for(a.n, kThreads)), dim3(kThreads), 0, s, core,
                       parent, a.n, order, labels);
    if (a.nitems > 0)
        hipLaunchKernelGGL((k_border<T, D>), dim3(blocks_for(a.nitems, kWavesPerBlock)),
                           dim3(kThreads), 0, s, a.soa, a.ld, a.items, a.nitems, a.ustart, a.roff,
                           a.ranges, a.eps2, core, static_cast<const int32_t*>(parent), order,
                           labels, stats);
    HPD_HIP(hipGetLastError());
}

#define HPD_INST_DBSCAN(T, D)                                                                  \
    template void launch_core<T, D>(const DbscanArgs<T>&, uint8_t*, unsigned long long*,       \
                                    hipStream_t);                                              \
    template void launch_union<T, D>(const DbscanArgs<T>&, const uint8_t*, int32_t*,           \
                                     hipStream_t);                                             \
    template void launch_labels<T, D>(const DbscanArgs<T>&, const uint8_t*, int32_t*,          \
                                      const uint32_t*, int64_t*, unsigned long long*,          \
                                      hipStream_t);
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

}  // namespace hpd
