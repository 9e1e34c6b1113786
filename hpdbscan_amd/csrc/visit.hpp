// Streaming candidate rows through a wave: the rows are read in aligned
// batches with wave-uniform addresses (a few wide scalar loads per batch),
// every lane tests its own row against each candidate of the batch.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <type_traits>

namespace hpd {

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

// Visit candidate rows [qb, qe).  body(q, qc, core_q, aux_q) runs per
// candidate; core_q is flags[q] (1 without flags), aux_q is aux[q] (0 without
// aux).  stop() is checked after each batch and ends the visit when true.
// Returns true if stopped.  soa, flags and aux are read up to the next
// multiple of the batch size past qe (callers pad them to ld rows).
template <typename T, int D, typename Body, typename Stop>
__device__ __forceinline__ bool visit_range(const T* __restrict__ soa, int64_t ld,
                                            const uint8_t* __restrict__ flags,
                                            const int32_t* __restrict__ aux, int32_t qb, int32_t qe,
                                            Body&& body, Stop&& stop) {
    constexpr int B = kBatch<T>;
    for (int32_t a = qb & ~(B - 1); a < qe; a += B) {
        Batch<T, B> b[D];
#pragma unroll
        for (int j = 0; j < D; ++j) b[j] = load_batch<T, B>(soa + static_cast<int64_t>(j) * ld, a);
        Batch<uint8_t, B> f;
        if (flags) f = load_batch<uint8_t, B>(flags, a);
        Batch<int32_t, B> x;
        if (aux) x = load_batch<int32_t, B>(aux, a);
        const bool full = a >= qb && a + B <= qe;
        const int lo = qb > a ? qb - a : 0;
        const int hi = qe - a < B ? qe - a : B;
#pragma unroll
        for (int t = 0; t < B; ++t) {
            if (!full && (t < lo || t >= hi)) continue;
            T qc[D];
#pragma unroll
            for (int j = 0; j < D; ++j) qc[j] = b[j].v[t];
            body(a + t, qc, flags ? f.v[t] : uint8_t(1), aux ? x.v[t] : 0);
        }
        if (stop()) return true;
    }
    return false;
}

// order-preserving integer images of floating-point values (for atomics)
__device__ __forceinline__ uint32_t ord_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_val(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ unsigned long long ord_key(double f) {
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(f));
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord_val(unsigned long long k) {
    return __longlong_as_double(static_cast<long long>((k & 0x8000000000000000ull)
                                                           ? (k & 0x7fffffffffffffffull)
                                                           : ~k));
}

template <typename T>
using OrdKey = std::conditional_t<sizeof(T) == 8, unsigned long long, uint32_t>;

template <typename T, int D>
__device__ __forceinline__ double box_lo(const OrdKey<T>* __restrict__ bbox, int32_t b, int j) {
    return static_cast<double>(ord_val(bbox[(static_cast<int64_t>(b) * D + j) * 2]));
}
template <typename T, int D>
__device__ __forceinline__ double box_hi(const OrdKey<T>* __restrict__ bbox, int32_t b, int j) {
    return static_cast<double>(ord_val(bbox[(static_cast<int64_t>(b) * D + j) * 2 + 1]));
}

// Squared distances from a row to a box's bounding box, in double: gmin2
// to its nearest point, gmax2 to its farthest corner.  Every member q of the
// box satisfies gmin2 <= |p - q|^2 <= gmax2.
template <typename T, int D>
__device__ __forceinline__ void point_box(const OrdKey<T>* __restrict__ bbox, int32_t b,
                                          const T* pc, double& gmin2, double& gmax2) {
    gmin2 = 0.0;
    gmax2 = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double lo = box_lo<T, D>(bbox, b, j);
        const double hi = box_hi<T, D>(bbox, b, j);
        const double v = static_cast<double>(pc[j]);
        const double g = fmax(0.0, fmax(lo - v, v - hi));
        const double f = fmax(v - lo, hi - v);
        gmin2 += g * g;
        gmax2 += f * f;
    }
}

// Wave-uniform copy of lane j's value (j wave-uniform): v_readlane.
__device__ __forceinline__ int32_t bcast(int32_t v, int j) {
    return __builtin_amdgcn_readlane(v, j);
}
__device__ __forceinline__ uint64_t bcast64(uint64_t v, int j) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), j));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), j));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ double bcast(double v, int j) {
    const long long x = __double_as_longlong(v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), j));
    const uint32_t hi =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x >> 32), j));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// Squared distances from a row to the box [lo, hi] (double): gmin2 to its
// nearest point, gmax2 to its farthest corner.
template <int D>
__device__ __forceinline__ void point_box_dd(const double* lo, const double* hi, const double* v,
                                             double& gmin2, double& gmax2) {
    gmin2 = 0.0;
    gmax2 = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double g = fmax(0.0, fmax(lo[j] - v[j], v[j] - hi[j]));
        const double f = fmax(v[j] - lo[j], hi[j] - v[j]);
        gmin2 += g * g;
        gmax2 += f * f;
    }
}
template <typename T, int D>
__device__ __forceinline__ void point_box_lh(const double* lo, const double* hi, const T* pc,
                                             double& gmin2, double& gmax2) {
    double v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = static_cast<double>(pc[j]);
    point_box_dd<D>(lo, hi, v, gmin2, gmax2);
}

// Visit the boxes of the neighbour rows [r0, r1) of an occupied cell, the
// cell's own row (index `center`) first.  The box table is fetched up to 64
// boxes at a time, one box per lane (the rows' box ranges likewise, one row
// per lane), and handed to fn box by box with readlane, so a box costs no
// dependent memory access of its own.  fn(b2, lo, hi, q0, q1, tag) gets the
// box, its bounding box (double), its rows [q0, q1) and tag[b2] (0 without
// tag), all wave-uniform; it returns true to stop the visit.  Only boxes
// below bmax are visited.  With first >= 0 (a box of the own row), the own
// row's boxes are visited from `first` on, wrapping round to the row's
// earlier boxes: a visit that stops early sees the own cell's boxes first.
template <typename T, int D, typename Fn>
__device__ __forceinline__ void for_each_row_box(int32_t r0, int32_t r1, int32_t center,
                                                 const int2* __restrict__ rcells,
                                                 const int32_t* __restrict__ cbox,
                                                 const int32_t* __restrict__ bstart,
                                                 const OrdKey<T>* __restrict__ bbox,
                                                 const int32_t* __restrict__ tag, Fn&& fn,
                                                 int32_t bmax = INT_MAX, int32_t first = -1) {
    const int32_t nrows = r1 - r0;
    const int lane = __lane_id();
    for (int32_t g0 = 0; g0 < nrows; g0 += 64) {
        const int32_t i = g0 + lane;  // visiting order: own row, then the others
        int32_t rb0 = 0, rcnt = 0;
        if (i < nrows) {
            const int32_t r = i == 0 ? center : (r0 + i - 1 < center ? r0 + i - 1 : r0 + i);
            const int2 cs = rcells[r];
            rb0 = cbox[cs.x];
            rcnt = max(0, min(cbox[cs.y], bmax) - rb0);  // boxes below bmax only
        }
        int32_t incl = rcnt;  // inclusive prefix over the lanes
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        const int32_t excl = incl - rcnt;
        const int32_t total = bcast(incl, 63);
        const int32_t grows = min(64, nrows - g0);
        // rotation of the own row (row 0 of the first group, lane 0)
        int32_t rot = 0, rot_n = 0;
        if (g0 == 0 && first >= 0) {
            const int32_t b0 = bcast(rb0, 0);
            rot_n = bcast(rcnt, 0);
            rot = (first >= b0 && first < b0 + rot_n) ? first - b0 : 0;
        }
        for (int32_t k0 = 0; k0 < total; k0 += 64) {
            const int32_t kk = min(k0 + lane, total - 1);
            int32_t row = 0;  // last row whose first box index is <= kk
            for (int32_t t = 1; t < grows; ++t) row = bcast(excl, t) <= kk ? t : row;
            int32_t local = kk - __shfl(excl, row, 64);
            if (row == 0 && rot > 0) local = local + rot < rot_n ? local + rot : local + rot - rot_n;
            const int32_t bl = __shfl(rb0, row, 64) + local;
            double lol[D], hil[D];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                lol[j] = box_lo<T, D>(bbox, bl, j);
                hil[j] = box_hi<T, D>(bbox, bl, j);
            }
            const int32_t q0l = bstart[bl], q1l = bstart[bl + 1];
            const int32_t tagl = tag ? tag[bl] : 0;
            const int32_t nk = min(64, total - k0);
            for (int32_t j = 0; j < nk; ++j) {
                double lo[D], hi[D];
#pragma unroll
                for (int e = 0; e < D; ++e) {
                    lo[e] = bcast(lol[e], j);
                    hi[e] = bcast(hil[e], j);
                }
                if (fn(bcast(bl, j), lo, hi, bcast(q0l, j), bcast(q1l, j), bcast(tagl, j))) return;
            }
        }
    }
}

// The same visit as for_each_row_box, handed to fn a group of up to 64 boxes
// at a time with one box per lane: fn(nk, bl, lol, hil, q0l, q1l, tagl) gets
// the group's size nk (wave-uniform) and, on lanes below nk, the lane's box,
// its bounding box, rows and tag (lanes at or above nk repeat the group's
// last box); it returns true to stop the visit.  Boxes keep the visiting
// order of for_each_row_box (lane order inside a group).  Lets a kernel
// classify a whole group against the bounding box of its rows in parallel
// and walk only the boxes that need row tests.
// The boxes of the neighbour rows g0 + lane of the visiting order (own row
// first): the first box rb0 and the number rcnt below bmax (0 past the rows).
__device__ __forceinline__ void row_boxes(int32_t r0, int32_t r1, int32_t center, int32_t g0,
                                          const int2* __restrict__ rcells,
                                          const int32_t* __restrict__ cbox, int32_t bmax,
                                          int32_t& rb0, int32_t& rcnt) {
    const int32_t i = g0 + __lane_id();
    rb0 = 0;
    rcnt = 0;
    if (i < r1 - r0) {
        const int32_t r = i == 0 ? center : (r0 + i - 1 < center ? r0 + i - 1 : r0 + i);
        const int2 cs = rcells[r];
        rb0 = cbox[cs.x];
        rcnt = max(0, min(cbox[cs.y], bmax) - rb0);
    }
}

// pre (optional): row_boxes of the first group, loaded by the caller ahead
// of its own loads so that the two overlap.
struct RowBoxes {
    int32_t rb0, rcnt;
};

template <typename T, int D, typename Fn>
__device__ __forceinline__ void for_each_row_box_group(int32_t r0, int32_t r1, int32_t center,
                                                       const int2* __restrict__ rcells,
                                                       const int32_t* __restrict__ cbox,
                                                       const int32_t* __restrict__ bstart,
                                                       const OrdKey<T>* __restrict__ bbox,
                                                       const int32_t* __restrict__ tag, Fn&& fn,
                                                       int32_t bmax = INT_MAX,
                                                       int32_t first = -1,
                                                       const RowBoxes* pre = nullptr) {
    const int32_t nrows = r1 - r0;
    const int lane = __lane_id();
    for (int32_t g0 = 0; g0 < nrows; g0 += 64) {
        int32_t rb0, rcnt;
        if (g0 == 0 && pre) {
            rb0 = pre->rb0;
            rcnt = pre->rcnt;
        } else {
            row_boxes(r0, r1, center, g0, rcells, cbox, bmax, rb0, rcnt);
        }
        int32_t incl = rcnt;  // inclusive prefix over the lanes
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        const int32_t excl = incl - rcnt;
        const int32_t total = bcast(incl, 63);
        const int32_t grows = min(64, nrows - g0);
        int32_t rot = 0, rot_n = 0;
        if (g0 == 0 && first >= 0) {
            const int32_t b0 = bcast(rb0, 0);
            rot_n = bcast(rcnt, 0);
            rot = (first >= b0 && first < b0 + rot_n) ? first - b0 : 0;
        }
        for (int32_t k0 = 0; k0 < total; k0 += 64) {
            const int32_t kk = min(k0 + lane, total - 1);
            int32_t row = 0;  // last row whose first box index is <= kk
            for (int32_t t = 1; t < grows; ++t) row = bcast(excl, t) <= kk ? t : row;
            int32_t local = kk - __shfl(excl, row, 64);
            if (row == 0 && rot > 0) local = local + rot < rot_n ? local + rot : local + rot - rot_n;
            const int32_t bl = __shfl(rb0, row, 64) + local;
            double lol[D], hil[D];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                lol[j] = box_lo<T, D>(bbox, bl, j);
                hil[j] = box_hi<T, D>(bbox, bl, j);
            }
            const int32_t q0l = bstart[bl], q1l = bstart[bl + 1];
            const int32_t tagl = tag ? tag[bl] : 0;
            if (fn(min(64, total - k0), bl, lol, hil, q0l, q1l, tagl)) return;
        }
    }
}

// Box-to-box squared distances (double): gmin2 between the nearest points
// of [alo, ahi] and [blo, bhi], gmax2 between the farthest.  For every row p
// inside box a, gmin2 <= gmin2(p, b) and gmax2(p, b) <= gmax2, also as
// computed in double: every step (subtract, max, square, add) rounds
// monotonically.
template <int D>
__device__ __forceinline__ void box_box(const double* alo, const double* ahi, const double* blo,
                                        const double* bhi, double& gmin2, double& gmax2) {
    gmin2 = 0.0;
    gmax2 = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double g = fmax(0.0, fmax(blo[j] - ahi[j], alo[j] - bhi[j]));
        const double f = fmax(bhi[j] - alo[j], ahi[j] - blo[j]);
        gmin2 += g * g;
        gmax2 += f * f;
    }
}

// A wave-uniform double held in scalar registers (v_readfirstlane).
__device__ __forceinline__ double uniform_d(double v) {
    const long long x = __double_as_longlong(v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x >> 32)));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// Wave-uniform bounding box (double) of the rows of the lanes where `on`.
template <typename T, int D>
__device__ __forceinline__ void rows_bbox(const T* pc, bool on, double* lo, double* hi) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double a = on ? static_cast<double>(pc[j]) : __builtin_huge_val();
        double b = on ? static_cast<double>(pc[j]) : -__builtin_huge_val();
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a = fmin(a, __shfl_xor(a, o, 64));
            b = fmax(b, __shfl_xor(b, o, 64));
        }
        lo[j] = uniform_d(a);
        hi[j] = uniform_d(b);
    }
}

// Runs of equal `key` over consecutive lanes (boxes are contiguous slot
// ranges): afterwards every lane holds the minimum (maximum) of v over the
// lanes from itself to the end of its run.
template <typename U>
__device__ __forceinline__ U run_suffix_min(U v, int32_t key) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const U w = __shfl_down(v, o, 64);
        const int32_t k = __shfl_down(key, o, 64);
        if (__lane_id() + o < 64 && k == key) v = w < v ? w : v;
    }
    return v;
}
template <typename U>
__device__ __forceinline__ U run_suffix_max(U v, int32_t key) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const U w = __shfl_down(v, o, 64);
        const int32_t k = __shfl_down(key, o, 64);
        if (__lane_id() + o < 64 && k == key) v = w > v ? w : v;
    }
    return v;
}

// True on the lane holding the first slot of its run of equal `key`.
__device__ __forceinline__ bool run_head(int32_t key) {
    const int32_t prev = __shfl_up(key, 1, 64);
    return __lane_id() == 0 || prev != key;
}

}  // namespace hpd
