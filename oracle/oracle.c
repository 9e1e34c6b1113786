/*
 * CPU oracle for HPDBSCAN's single-node clustering path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a sequential restatement of the
 * reference algorithm (Markus-Goetz/hpdbscan, run with one OpenMP thread) used
 * as the checker for the HIP implementation.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  The product never links it
 * and never falls back to it.
 *
 * What each step restates (file:line in the reference):
 *   min/max per dimension ............ spatial_index.h:99-122, initial values
 *                                      :531-532 (max starts at
 *                                      numeric_limits<T>::min(), i.e. FLT_MIN /
 *                                      DBL_MIN for floating types)
 *   cells per dimension .............. spatial_index.h:124-131
 *   dimension order .................. spatial_index.h:133-142 (std::sort of
 *                                      <= 16 keys is an insertion sort, so ties
 *                                      keep index order)
 *   cell id per point ................ spatial_index.h:144-165
 *   cell index + stable sort ......... spatial_index.h:167-216
 *   neighbour candidate list ......... spatial_index.h:682-737
 *   region query ..................... spatial_index.h:739-768
 *   local DBSCAN, one thread ......... hpdbscan.h:51-101
 *   rules (label merge) .............. rules.h:26-71
 *   apply rules ...................... hpdbscan.h:198-211
 *   scatter back to input order ...... spatial_index.h:852-872
 *
 * Arithmetic as the reference's build (CMakeLists.txt:12: -Ofast, x86-64
 * baseline, g++ 11.4) executes it:
 *   - `x / eps` inside loops becomes `x * (1/eps)` with 1/eps rounded once
 *     (-freciprocal-math): float for float and integer data, double for
 *     double data (eps widened first).
 *   - -Ofast links crtfastmath.o, which turns on flush-to-zero and
 *     denormals-are-zero; this file sets the same MXCSR bits while it runs.
 *   - The squared distance of spatial_index.h:751-754 is summed by the SSE2
 *     vectoriser: float uses four partial sums over d = j (mod 4) reduced as
 *     (a0 + a2) + (a1 + a3), then a two-wide remainder pair, then a last
 *     scalar; double uses two partial sums over d = j (mod 2) reduced as
 *     a0 + a1, then a last scalar.  Dimensions 1..3 therefore add left to
 *     right.  No fused multiply-add (SSE2 has none).
 *   - Integer data computes `T distance = p - q` and `offset += distance *
 *     distance` in T, i.e. modulo 2^bits(T); EPS2 = (T)(float)(eps * eps).
 *
 * The labels are those of the reference at one thread.  With several threads
 * the reference may give a border point that touches cores of two different
 * clusters either cluster; `border_ambiguous` flags those points.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <xmmintrin.h>

#define HPD_NOT_VISITED INT64_MAX
#define HPD_NOISE (INT64_MAX - 1)

enum {
    HPD_F32 = 0,
    HPD_F64 = 1,
    HPD_I8 = 2,
    HPD_I16 = 3,
    HPD_I32 = 4,
    HPD_I64 = 5,
    HPD_U8 = 6,
    HPD_U16 = 7,
    HPD_U32 = 8,
    HPD_U64 = 9,
};

/* ------------------------------------------------------------------------ */
/* per-dtype arithmetic                                                      */
/* ------------------------------------------------------------------------ */

/* Squared distance, float, as vectorised by the reference build. */
static float dist2_f32(const float* p, const float* q, int d) {
    float s = 0.0f;
    int base = 0;
    if (d >= 4) {
        const int k4 = d / 4;
        float a[4];
        for (int j = 0; j < 4; ++j) {
            const float t = p[j] - q[j];
            a[j] = t * t;
        }
        for (int k = 1; k < k4; ++k) {
            for (int j = 0; j < 4; ++j) {
                const float t = p[4 * k + j] - q[4 * k + j];
                a[j] = a[j] + t * t;
            }
        }
        s = (a[0] + a[2]) + (a[1] + a[3]);
        base = 4 * k4;
    }
    if (d - base >= 2) {
        const float t0 = p[base] - q[base];
        const float t1 = p[base + 1] - q[base + 1];
        s = s + (t0 * t0 + t1 * t1);
        base += 2;
    }
    if (d - base == 1) {
        const float t = p[base] - q[base];
        s = s + t * t;
    }
    return s;
}

/* Squared distance, double, as vectorised by the reference build. */
static double dist2_f64(const double* p, const double* q, int d) {
    double s = 0.0;
    int base = 0;
    if (d >= 2) {
        const int k2 = d / 2;
        double a0, a1;
        {
            const double t0 = p[0] - q[0];
            const double t1 = p[1] - q[1];
            a0 = t0 * t0;
            a1 = t1 * t1;
        }
        for (int k = 1; k < k2; ++k) {
            const double t0 = p[2 * k] - q[2 * k];
            const double t1 = p[2 * k + 1] - q[2 * k + 1];
            a0 = a0 + t0 * t0;
            a1 = a1 + t1 * t1;
        }
        s = a0 + a1;
        base = 2 * k2;
    }
    if (d - base == 1) {
        const double t = p[base] - q[base];
        s = s + t * t;
    }
    return s;
}

/* Integer distance: every step wraps modulo 2^bits(T) (two's complement). */
#define HPD_INT_DIST(NAME, T, UT)                                           \
    static T NAME(const T* p, const T* q, int d) {                          \
        UT s = 0;                                                           \
        for (int i = 0; i < d; ++i) {                                       \
            const UT t = (UT)((UT)p[i] - (UT)q[i]);                         \
            s = (UT)(s + (UT)((uint64_t)t * (uint64_t)t));                  \
        }                                                                   \
        return (T)s;                                                        \
    }
HPD_INT_DIST(dist2_i8, int8_t, uint8_t)
HPD_INT_DIST(dist2_i16, int16_t, uint16_t)
HPD_INT_DIST(dist2_i32, int32_t, uint32_t)
HPD_INT_DIST(dist2_i64, int64_t, uint64_t)
HPD_INT_DIST(dist2_u8, uint8_t, uint8_t)
HPD_INT_DIST(dist2_u16, uint16_t, uint16_t)
HPD_INT_DIST(dist2_u32, uint32_t, uint32_t)
HPD_INT_DIST(dist2_u64, uint64_t, uint64_t)

static size_t dtype_size(int dt) {
    switch (dt) {
        case HPD_F32: return 4;
        case HPD_F64: return 8;
        case HPD_I8: case HPD_U8: return 1;
        case HPD_I16: case HPD_U16: return 2;
        case HPD_I32: case HPD_U32: return 4;
        case HPD_I64: case HPD_U64: return 8;
        default: return 0;
    }
}

/* Float -> integer conversions as the x86-64 build performs them for
 * out-of-range values (cvttss2si yields the "integer indefinite" value). */
static int64_t f32_to_i64_x86(float v) {
    if (!(v >= -9223372036854775808.0f && v < 9223372036854775808.0f)) return INT64_MIN;
    return (int64_t)v;
}
static int32_t f32_to_i32_x86(float v) {
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
    return (int32_t)v;
}
static uint64_t f32_to_u64_x86(float v) {
    if (v < 9223372036854775808.0f) return (uint64_t)f32_to_i64_x86(v);
    return ((uint64_t)f32_to_i64_x86(v - 9223372036854775808.0f)) ^ 0x8000000000000000ull;
}
static uint64_t f64_to_u64_x86(double v) {
    if (v < 9223372036854775808.0) {
        if (!(v >= -9223372036854775808.0)) return 0x8000000000000000ull;
        return (uint64_t)(int64_t)v;
    }
    const double w = v - 9223372036854775808.0;
    int64_t r = (w < 9223372036854775808.0) ? (int64_t)w : INT64_MIN;
    return ((uint64_t)r) ^ 0x8000000000000000ull;
}

/* The typed view of one coordinate as the float the cell arithmetic uses
 * ((x - min) in T's promoted type, converted to float). */
static float int_diff_as_f32(int dt, const void* base, size_t idx, const void* minv) {
    switch (dt) {
        case HPD_I8: return (float)((int)((const int8_t*)base)[idx] - (int)*(const int8_t*)minv);
        case HPD_U8: return (float)((int)((const uint8_t*)base)[idx] - (int)*(const uint8_t*)minv);
        case HPD_I16: return (float)((int)((const int16_t*)base)[idx] - (int)*(const int16_t*)minv);
        case HPD_U16: return (float)((int)((const uint16_t*)base)[idx] - (int)*(const uint16_t*)minv);
        case HPD_I32:
            return (float)(int32_t)((uint32_t)((const int32_t*)base)[idx] - (uint32_t)*(const int32_t*)minv);
        case HPD_U32: return (float)(uint32_t)(((const uint32_t*)base)[idx] - *(const uint32_t*)minv);
        case HPD_I64:
            return (float)(int64_t)((uint64_t)((const int64_t*)base)[idx] - (uint64_t)*(const int64_t*)minv);
        case HPD_U64: return (float)(uint64_t)(((const uint64_t*)base)[idx] - *(const uint64_t*)minv);
        default: return 0.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* spatial index                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t cell;
    int64_t index;
} cell_item;

static int cmp_cell_item(const void* a, const void* b) {
    const cell_item* x = (const cell_item*)a;
    const cell_item* y = (const cell_item*)b;
    if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
    return x->index < y->index ? -1 : (x->index > y->index ? 1 : 0);
}

typedef struct {
    int64_t n;
    int d;
    int dt;
    size_t esz;
    float eps;
    uint64_t cells_per_dim[64];
    int swapped[64];
    uint64_t total_cells;
    /* sorted state */
    uint64_t* cell_sorted;   /* [n] */
    int64_t* orig_of_sorted; /* [n] */
    unsigned char* pts;      /* [n*d*esz] sorted coordinates */
    /* occupied cell table, ascending */
    int64_t n_cells;
    uint64_t* ucell;   /* [n_cells] */
    int64_t* ustart;   /* [n_cells] */
    int64_t* ucount;   /* [n_cells] */
} spatial;

static int64_t find_cell(const spatial* s, uint64_t cell) {
    int64_t lo = 0, hi = s->n_cells;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (s->ucell[mid] < cell) lo = mid + 1;
        else hi = mid;
    }
    if (lo < s->n_cells && s->ucell[lo] == cell) return lo;
    return -1;
}

/* Returns 0 on success, 2 when out of memory. */
static int build_spatial(spatial* s, const void* data) {
    const int64_t n = s->n;
    const int d = s->d;
    const int dt = s->dt;
    const size_t esz = s->esz;

    /* min / max with the reference's initial values (spatial_index.h:531-532) */
    unsigned char minv[64 * 8], maxv[64 * 8];
    for (int j = 0; j < d; ++j) {
        switch (dt) {
            case HPD_F32: ((float*)minv)[j] = FLT_MAX; ((float*)maxv)[j] = FLT_MIN; break;
            case HPD_F64: ((double*)minv)[j] = DBL_MAX; ((double*)maxv)[j] = DBL_MIN; break;
            case HPD_I8: ((int8_t*)minv)[j] = INT8_MAX; ((int8_t*)maxv)[j] = INT8_MIN; break;
            case HPD_I16: ((int16_t*)minv)[j] = INT16_MAX; ((int16_t*)maxv)[j] = INT16_MIN; break;
            case HPD_I32: ((int32_t*)minv)[j] = INT32_MAX; ((int32_t*)maxv)[j] = INT32_MIN; break;
            case HPD_I64: ((int64_t*)minv)[j] = INT64_MAX; ((int64_t*)maxv)[j] = INT64_MIN; break;
            case HPD_U8: ((uint8_t*)minv)[j] = UINT8_MAX; ((uint8_t*)maxv)[j] = 0; break;
            case HPD_U16: ((uint16_t*)minv)[j] = UINT16_MAX; ((uint16_t*)maxv)[j] = 0; break;
            case HPD_U32: ((uint32_t*)minv)[j] = UINT32_MAX; ((uint32_t*)maxv)[j] = 0; break;
            case HPD_U64: ((uint64_t*)minv)[j] = UINT64_MAX; ((uint64_t*)maxv)[j] = 0; break;
        }
    }
#define HPD_MINMAX(T)                                                         \
    {                                                                         \
        const T* x = (const T*)data;                                          \
        T* mn = (T*)minv;                                                     \
        T* mx = (T*)maxv;                                                     \
        for (int64_t i = 0; i < n; ++i)                                       \
            for (int j = 0; j < d; ++j) {                                     \
                const T c = x[i * d + j];                                     \
                if (c < mn[j]) mn[j] = c;                                     \
                if (mx[j] < c) mx[j] = c;                                     \
            }                                                                 \
    }
    switch (dt) {
        case HPD_F32: HPD_MINMAX(float) break;
        case HPD_F64: HPD_MINMAX(double) break;
        case HPD_I8: HPD_MINMAX(int8_t) break;
        case HPD_I16: HPD_MINMAX(int16_t) break;
        case HPD_I32: HPD_MINMAX(int32_t) break;
        case HPD_I64: HPD_MINMAX(int64_t) break;
        case HPD_U8: HPD_MINMAX(uint8_t) break;
        case HPD_U16: HPD_MINMAX(uint16_t) break;
        case HPD_U32: HPD_MINMAX(uint32_t) break;
        case HPD_U64: HPD_MINMAX(uint64_t) break;
    }
#undef HPD_MINMAX

    /* cells per dimension (spatial_index.h:124-131), reciprocal form */
    const float rcp_f = 1.0f / s->eps;
    const double rcp_d = 1.0 / (double)s->eps;
    s->total_cells = 1;
    for (int j = 0; j < d; ++j) {
        uint64_t cells;
        if (dt == HPD_F32) {
            const float v = (((float*)maxv)[j] - ((float*)minv)[j]) * rcp_f;
            cells = f32_to_u64_x86(ceilf(v)) + 1;
        } else if (dt == HPD_F64) {
            const double v = (((double*)maxv)[j] - ((double*)minv)[j]) * rcp_d;
            cells = f64_to_u64_x86(ceil(v)) + 1;
        } else {
            const float v = int_diff_as_f32(dt, maxv, (size_t)j, minv + (size_t)j * esz) * rcp_f;
            cells = f32_to_u64_x86(ceilf(v)) + 1;
        }
        s->cells_per_dim[j] = cells;
        s->total_cells *= cells;
    }

    /* dimension order: stable ascending by cell count (spatial_index.h:133-142) */
    for (int j = 0; j < d; ++j) s->swapped[j] = j;
    for (int j = 1; j < d; ++j) {
        const int v = s->swapped[j];
        int k = j - 1;
        while (k >= 0 && s->cells_per_dim[v] < s->cells_per_dim[s->swapped[k]]) {
            s->swapped[k + 1] = s->swapped[k];
            --k;
        }
        s->swapped[k + 1] = v;
    }

    /* cell id per point (spatial_index.h:144-165) */
    cell_item* items = (cell_item*)malloc(sizeof(cell_item) * (size_t)(n > 0 ? n : 1));
    if (!items) return 2;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t cell = 0, acc = 1;
        for (int k = 0; k < d; ++k) {
            const int j = s->swapped[k];
            uint64_t idx;
            if (dt == HPD_F32) {
                const float v = (((const float*)data)[i * d + j] - ((float*)minv)[j]) * rcp_f;
                idx = f32_to_u64_x86(floorf(v));
            } else if (dt == HPD_F64) {
                const double v = (((const double*)data)[i * d + j] - ((double*)minv)[j]) * rcp_d;
                idx = f64_to_u64_x86(floor(v));
            } else {
                const float v = int_diff_as_f32(dt, data, (size_t)(i * d + j), minv + (size_t)j * esz) * rcp_f;
                idx = f32_to_u64_x86(floorf(v));
            }
            cell += idx * acc;
            acc *= s->cells_per_dim[j];
        }
        items[i].cell = cell;
        items[i].index = i;
    }

    /* stable sort by cell (spatial_index.h:184-216): ties keep input order */
    qsort(items, (size_t)n, sizeof(cell_item), cmp_cell_item);

    s->cell_sorted = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    s->orig_of_sorted = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    s->pts = (unsigned char*)malloc(esz * (size_t)d * (size_t)(n > 0 ? n : 1) + 1);
    s->ucell = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    s->ustart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    s->ucount = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!s->cell_sorted || !s->orig_of_sorted || !s->pts || !s->ucell || !s->ustart || !s->ucount) {
        free(items);
        return 2;
    }
    s->n_cells = 0;
    for (int64_t i = 0; i < n; ++i) {
        s->cell_sorted[i] = items[i].cell;
        s->orig_of_sorted[i] = items[i].index;
        memcpy(s->pts + (size_t)i * d * esz, (const unsigned char*)data + (size_t)items[i].index * d * esz,
               (size_t)d * esz);
        if (i == 0 || items[i].cell != items[i - 1].cell) {
            s->ucell[s->n_cells] = items[i].cell;
            s->ustart[s->n_cells] = i;
            s->ucount[s->n_cells] = 0;
            s->n_cells++;
        }
        s->ucount[s->n_cells - 1]++;
    }
    free(items);
    return 0;
}

static void free_spatial(spatial* s) {
    free(s->cell_sorted);
    free(s->orig_of_sorted);
    free(s->pts);
    free(s->ucell);
    free(s->ustart);
    free(s->ucount);
}

/* Candidate list of a cell (spatial_index.h:682-737): the cell itself, then
 * for each dimension in swapped order the lower / upper neighbour of every
 * entry so far; cells without points contribute nothing.  Writes the point
 * ranges [start, start+count) into `rs`/`rc` (capacity 3^d) and returns the
 * number of ranges. */
static int64_t neighbor_ranges(const spatial* s, uint64_t cell, uint64_t* scratch_cells, int64_t* rs,
                               int64_t* rc) {
    int64_t ncell = 0;
    scratch_cells[ncell++] = cell;
    uint64_t lower = 1, current = 1;
    for (int k = 0; k < s->d; ++k) {
        const int j = s->swapped[k];
        current *= s->cells_per_dim[j];
        const int64_t end = ncell;
        for (int64_t i = 0; i < end; ++i) {
            const uint64_t c = scratch_cells[i];
            if (c % current >= lower) scratch_cells[ncell++] = c - lower;
            if (c % current < current - lower) scratch_cells[ncell++] = c + lower;
        }
        lower = current;
    }
    int64_t nr = 0;
    for (int64_t i = 0; i < ncell; ++i) {
        const int64_t u = find_cell(s, scratch_cells[i]);
        if (u < 0) continue;
        rs[nr] = s->ustart[u];
        rc[nr] = s->ucount[u];
        ++nr;
    }
    return nr;
}

/* exact within-eps test of two sorted positions, in T */
typedef struct {
    float f32;
    double f64;
    int64_t i64; /* EPS2 for signed integer T, already converted to T */
    uint64_t u64;
} eps2_t;

static int within(const spatial* s, int64_t p, int64_t q, const eps2_t* e) {
    const int d = s->d;
    const size_t row = (size_t)d * s->esz;
    const void* a = s->pts + (size_t)p * row;
    const void* b = s->pts + (size_t)q * row;
    switch (s->dt) {
        case HPD_F32: return dist2_f32((const float*)a, (const float*)b, d) <= e->f32;
        case HPD_F64: return dist2_f64((const double*)a, (const double*)b, d) <= e->f64;
        case HPD_I8: return (int64_t)dist2_i8((const int8_t*)a, (const int8_t*)b, d) <= e->i64;
        case HPD_I16: return (int64_t)dist2_i16((const int16_t*)a, (const int16_t*)b, d) <= e->i64;
        case HPD_I32: return (int64_t)dist2_i32((const int32_t*)a, (const int32_t*)b, d) <= e->i64;
        case HPD_I64: return dist2_i64((const int64_t*)a, (const int64_t*)b, d) <= e->i64;
        case HPD_U8: return (uint64_t)dist2_u8((const uint8_t*)a, (const uint8_t*)b, d) <= e->u64;
        case HPD_U16: return (uint64_t)dist2_u16((const uint16_t*)a, (const uint16_t*)b, d) <= e->u64;
        case HPD_U32: return (uint64_t)dist2_u32((const uint32_t*)a, (const uint32_t*)b, d) <= e->u64;
        case HPD_U64: return dist2_u64((const uint64_t*)a, (const uint64_t*)b, d) <= e->u64;
    }
    return 0;
}

/* EPS2 = (float)(eps * eps) (hpdbscan.h:53), passed to region_query as T
 * (spatial_index.h:739). */
static eps2_t make_eps2(int dt, float eps) {
    eps2_t e;
    const float f = eps * eps;
    e.f32 = f;
    e.f64 = (double)f;
    e.i64 = 0;
    e.u64 = 0;
    switch (dt) {
        case HPD_I8: e.i64 = (int8_t)f32_to_i32_x86(f); break;
        case HPD_I16: e.i64 = (int16_t)f32_to_i32_x86(f); break;
        case HPD_I32: e.i64 = f32_to_i32_x86(f); break;
        case HPD_I64: e.i64 = f32_to_i64_x86(f); break;
        case HPD_U8: e.u64 = (uint8_t)f32_to_i32_x86(f); break;
        case HPD_U16: e.u64 = (uint16_t)f32_to_i32_x86(f); break;
        case HPD_U32: e.u64 = (uint32_t)f32_to_i64_x86(f); break;
        case HPD_U64: e.u64 = f32_to_u64_x86(f); break;
    }
    return e;
}

/* ------------------------------------------------------------------------ */
/* rules (rules.h:26-71) on a dense array: link[l] for labels 1..n, 0 = none */
/* ------------------------------------------------------------------------ */

static void rules_update(int64_t* link, int64_t n, int64_t first, int64_t second) {
    for (;;) {
        if (first <= second || first >= HPD_NOISE) return;
        if (first > n) return; /* labels never exceed n */
        const int64_t r = link[first];
        if (r == 0) {
            link[first] = second;
            return;
        }
        if (r > second) {
            /* update(r, second) touches only labels below `first`, so the
             * order of the two steps in rules.h:60-62 does not matter */
            link[first] = second;
            first = r;
        } else {
            const int64_t f2 = second;
            second = r;
            first = f2;
        }
    }
}

static int64_t rules_rule(const int64_t* link, int64_t n, int64_t label) {
    if (label == HPD_NOISE) return 0;
    if (label >= 1 && label <= n && link[label] != 0) return link[label];
    return HPD_NOT_VISITED;
}

/* ------------------------------------------------------------------------ */
/* public entry points                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t sum_candidates; /* Σ_p |candidates(cell(p))|, the reference's work */
    uint64_t occupied_cells;
    uint64_t total_cells;
    uint64_t core_points;
    uint64_t noise_points;
    uint64_t ambiguous_borders;
    int32_t swapped[16];
    uint64_t cells_per_dim[16];
} hpd_oracle_stats;

/*
 * Cluster n points of dimension d (row-major, dtype `dt`).
 *   labels[n]            reference labels in input order (0 noise, -k core of
 *                        cluster k, +k border of cluster k)
 *   sorted_pos[n]        optional: sorted position of every input row
 *   cell_of_sorted[n]    optional: cell id at every sorted position
 *   core[n]              optional: 1 where the input row is a core point
 *   border_ambiguous[n]  optional: 1 where a border input row lies within eps
 *                        of cores of two or more clusters
 *   stats                optional
 * Returns 0 on success, 1 on invalid arguments (eps <= 0 like
 * hpdbscan.h:295-297, or d outside 1..10), 2 when out of memory.
 */
int hpd_oracle_cluster(int dt, const void* data, int64_t n, int32_t d, float eps, uint64_t min_pts,
                       int64_t* labels, int64_t* sorted_pos, uint64_t* cell_of_sorted, uint8_t* core,
                       uint8_t* border_ambiguous, hpd_oracle_stats* stats) {
    if (!(eps > 0.0f)) return 1;
    if (d < 1 || d > 10 || n < 0) return 1;
    const size_t esz = dtype_size(dt);
    if (esz == 0) return 1;

    const unsigned int saved_csr = _mm_getcsr();
    _mm_setcsr(saved_csr | 0x8040); /* FTZ | DAZ, as crtfastmath.o sets them */

    int rc = 0;
    spatial s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.d = d;
    s.dt = dt;
    s.esz = esz;
    s.eps = eps;

    int64_t* clusters = NULL;
    int64_t* link = NULL;
    int64_t* cand = NULL;
    uint64_t* scratch_cells = NULL;
    int64_t *rs = NULL, *rcn = NULL;
    int64_t* area = NULL;
    uint8_t* core_sorted = NULL;

    if (build_spatial(&s, data) != 0) {
        rc = 2;
        goto done;
    }

    int64_t n3 = 1;
    for (int j = 0; j < d; ++j) n3 *= 3;
    clusters = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    link = (int64_t*)calloc((size_t)(n + 2), sizeof(int64_t));
    cand = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    area = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    scratch_cells = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n3);
    rs = (int64_t*)malloc(sizeof(int64_t) * (size_t)n3);
    rcn = (int64_t*)malloc(sizeof(int64_t) * (size_t)n3);
    core_sorted = (uint8_t*)calloc((size_t)(n + 1), 1);
    if (!clusters || !link || !cand || !area || !scratch_cells || !rs || !rcn || !core_sorted) {
        rc = 2;
        goto done;
    }

    const eps2_t e2 = make_eps2(dt, eps);
    for (int64_t i = 0; i < n; ++i) clusters[i] = HPD_NOT_VISITED;

    /* local DBSCAN with one thread (hpdbscan.h:51-101) */
    uint64_t sum_cand = 0;
    uint64_t prev_cell = 0;
    int have_prev = 0;
    int64_t ncand = 0;
    for (int64_t p = 0; p < n; ++p) {
        const uint64_t cell = s.cell_sorted[p];
        if (!have_prev || cell != prev_cell) {
            const int64_t nr = neighbor_ranges(&s, cell, scratch_cells, rs, rcn);
            ncand = 0;
            for (int64_t r = 0; r < nr; ++r)
                for (int64_t q = rs[r]; q < rs[r] + rcn[r]; ++q) cand[ncand++] = q;
            prev_cell = cell;
            have_prev = 1;
        }
        sum_cand += (uint64_t)ncand;

        int64_t narea = 0;
        int64_t label = HPD_NOISE;
        if ((uint64_t)ncand >= min_pts) {
            /* region_query (spatial_index.h:739-768) */
            label = p + 1;
            for (int64_t c = 0; c < ncand; ++c) {
                const int64_t q = cand[c];
                if (within(&s, p, q, &e2)) {
                    const int64_t nl = clusters[q];
                    area[narea++] = q;
                    if (nl != HPD_NOT_VISITED && nl < 0 && -nl < label) label = -nl;
                }
            }
        }
        if ((uint64_t)narea >= min_pts) {
            core_sorted[p] = 1;
            if (-label < clusters[p]) clusters[p] = -label;
            for (int64_t a = 0; a < narea; ++a) {
                const int64_t q = area[a];
                const int64_t ql = clusters[q] < 0 ? -clusters[q] : clusters[q];
                if (clusters[q] < 0) {
                    const int64_t lo = label < ql ? label : ql;
                    const int64_t hi = label < ql ? ql : label;
                    rules_update(link, n, hi, lo);
                }
                if (label < clusters[q]) clusters[q] = label;
            }
        } else if (clusters[p] == HPD_NOT_VISITED) {
            clusters[p] = HPD_NOISE;
        }
    }

    /* apply rules (hpdbscan.h:198-211) */
    for (int64_t i = 0; i < n; ++i) {
        const int is_core = clusters[i] < 0;
        int64_t c = is_core ? -clusters[i] : clusters[i];
        int64_t m = rules_rule(link, n, c);
        while (m < HPD_NOISE) {
            c = m;
            m = rules_rule(link, n, m);
        }
        clusters[i] = is_core ? -c : c;
    }

    /* border points touching cores of more than one cluster */
    uint64_t n_amb = 0, n_core = 0, n_noise = 0;
    if (border_ambiguous) memset(border_ambiguous, 0, (size_t)n);
    have_prev = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (core_sorted[p]) ++n_core;
        if (clusters[p] == 0) ++n_noise;
        if (core_sorted[p] || clusters[p] == 0) continue;
        const uint64_t cell = s.cell_sorted[p];
        const int64_t nr = neighbor_ranges(&s, cell, scratch_cells, rs, rcn);
        int64_t seen = 0;
        int amb = 0;
        for (int64_t r = 0; r < nr && !amb; ++r)
            for (int64_t q = rs[r]; q < rs[r] + rcn[r]; ++q) {
                if (!core_sorted[q] || !within(&s, p, q, &e2)) continue;
                const int64_t k = -clusters[q];
                if (seen == 0) seen = k;
                else if (k != seen) {
                    amb = 1;
                    break;
                }
            }
        if (amb) {
            ++n_amb;
            if (border_ambiguous) border_ambiguous[s.orig_of_sorted[p]] = 1;
        }
    }

    /* scatter back to input order (spatial_index.h:852-872) */
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = s.orig_of_sorted[i];
        labels[o] = clusters[i];
        if (sorted_pos) sorted_pos[o] = i;
        if (core) core[o] = core_sorted[i];
        if (cell_of_sorted) cell_of_sorted[i] = s.cell_sorted[i];
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->sum_candidates = sum_cand;
        stats->occupied_cells = (uint64_t)s.n_cells;
        stats->total_cells = s.total_cells;
        stats->core_points = n_core;
        stats->noise_points = n_noise;
        stats->ambiguous_borders = n_amb;
        for (int j = 0; j < d; ++j) {
            stats->swapped[j] = s.swapped[j];
            stats->cells_per_dim[j] = s.cells_per_dim[j];
        }
    }

done:
    free(clusters);
    free(link);
    free(cand);
    free(area);
    free(scratch_cells);
    free(rs);
    free(rcn);
    free(core_sorted);
    free_spatial(&s);
    _mm_setcsr(saved_csr);
    return rc;
}

/*
 * Checker for labels produced under any processing order.  Given candidate
 * labels `got` (input order) and the oracle's one-thread labels `want`, counts
 *   mismatched_exact: rows where got != want and the row is not an ambiguous
 *                     border (core, noise and one-cluster border rows must be
 *                     identical),
 *   bad_ambiguous:    ambiguous border rows whose label is not +k for some
 *                     cluster k that owns a core within eps of the row.
 * Returns 0 on success, 1/2 like hpd_oracle_cluster.
 */
int hpd_oracle_check(int dt, const void* data, int64_t n, int32_t d, float eps, uint64_t min_pts,
                     const int64_t* got, uint64_t* mismatched_exact, uint64_t* bad_ambiguous) {
    if (!(eps > 0.0f)) return 1;
    if (d < 1 || d > 10 || n < 0) return 1;
    int64_t* want = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    uint8_t* amb = (uint8_t*)malloc((size_t)(n + 1));
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!want || !amb || !pos) {
        free(want);
        free(amb);
        free(pos);
        return 2;
    }
    int rc = hpd_oracle_cluster(dt, data, n, d, eps, min_pts, want, pos, NULL, NULL, amb, NULL);
    if (rc != 0) {
        free(want);
        free(amb);
        free(pos);
        return rc;
    }
    uint64_t mism = 0, bad = 0;
    const unsigned int saved_csr = _mm_getcsr();
    _mm_setcsr(saved_csr | 0x8040);
    spatial s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.d = d;
    s.dt = dt;
    s.esz = dtype_size(dt);
    s.eps = eps;
    int have_index = 0;
    int64_t n3 = 1;
    for (int j = 0; j < d; ++j) n3 *= 3;
    uint64_t* scratch_cells = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n3);
    int64_t* rs = (int64_t*)malloc(sizeof(int64_t) * (size_t)n3);
    int64_t* rcn = (int64_t*)malloc(sizeof(int64_t) * (size_t)n3);
    int64_t* sorted_to_orig = NULL;
    const eps2_t e2 = make_eps2(dt, eps);
    for (int64_t i = 0; i < n; ++i) {
        if (!amb[i]) {
            if (got[i] != want[i]) ++mism;
            continue;
        }
        if (!have_index) {
            if (build_spatial(&s, data) != 0 || !scratch_cells || !rs || !rcn) {
                rc = 2;
                break;
            }
            have_index = 1;
            sorted_to_orig = s.orig_of_sorted;
        }
        /* got[i] must be +k with a core of cluster k within eps */
        int ok = 0;
        if (got[i] > 0) {
            const int64_t p = pos[i];
            const int64_t nr = neighbor_ranges(&s, s.cell_sorted[p], scratch_cells, rs, rcn);
            for (int64_t r = 0; r < nr && !ok; ++r)
                for (int64_t q = rs[r]; q < rs[r] + rcn[r]; ++q) {
                    const int64_t wq = want[sorted_to_orig[q]];
                    if (wq < 0 && -wq == got[i] && within(&s, p, q, &e2)) {
                        ok = 1;
                        break;
                    }
                }
        }
        if (!ok) ++bad;
    }
    _mm_setcsr(saved_csr);
    if (have_index) free_spatial(&s);
    free(scratch_cells);
    free(rs);
    free(rcn);
    free(want);
    free(amb);
    free(pos);
    if (mismatched_exact) *mismatched_exact = mism;
    if (bad_ambiguous) *bad_ambiguous = bad;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* parallel checker of the closed form (large inputs)                        */
/* ------------------------------------------------------------------------ */

/* Lock-free union-find over sorted positions: a root is always the smallest
 * position of its tree (the larger root is hooked under the smaller). */
static int64_t uf_find_par(int64_t* parent, int64_t x) {
    for (;;) {
        const int64_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
        if (p == x) return x;
        const int64_t g = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
        if (g != p) __atomic_compare_exchange_n(&parent[x], &(int64_t){p}, g, 0, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED);
        x = g;
    }
}

static void uf_unite_par(int64_t* parent, int64_t a, int64_t b) {
    for (;;) {
        a = uf_find_par(parent, a);
        b = uf_find_par(parent, b);
        if (a == b) return;
        if (a < b) {
            const int64_t t = a;
            a = b;
            b = t;
        }
        int64_t expect = a;
        if (__atomic_compare_exchange_n(&parent[a], &expect, b, 0, __ATOMIC_ACQ_REL,
                                        __ATOMIC_RELAXED))
            return;
    }
}

/*
 * Checker of candidate labels `got` (input order) against the closed form of
 * the reference's result, computed with `threads` OpenMP threads (0: the
 * default), for inputs too large for the one-thread restatement above:
 *   core(p)  <=>  |{q in candidates(p) : dist2(p, q) <= EPS2}| >= minPts
 *   label of a core = -(1 + smallest sorted position of a core of its
 *                     eps-connected component of cores)
 *   non-core row: 0 without a core within eps; +k if the cores within eps
 *                 all lie in cluster k; else (ambiguous) any +k of them.
 * Counts like hpd_oracle_check: mismatched_exact (core, noise, core labels,
 * one-cluster borders) and bad_ambiguous.  Optional: number of cores and of
 * ambiguous rows.  The candidates, the distance test and the sorted order
 * are those of the restatement above (build_spatial, neighbor_ranges,
 * within); the core count stops at minPts, the union tests each core pair
 * once.  Returns 0, or 1/2 like hpd_oracle_cluster.
 */
int hpd_oracle_check_parallel(int dt, const void* data, int64_t n, int32_t d, float eps,
                              uint64_t min_pts, const int64_t* got, int threads,
                              uint64_t* mismatched_exact, uint64_t* bad_ambiguous,
                              uint64_t* cores_out, uint64_t* ambiguous_out) {
    if (!(eps > 0.0f)) return 1;
    if (d < 1 || d > 10 || n < 0) return 1;
    const size_t esz = dtype_size(dt);
    if (esz == 0) return 1;
    if (threads > 0) omp_set_num_threads(threads);
    const unsigned int saved_csr = _mm_getcsr();
    _mm_setcsr(saved_csr | 0x8040);
    spatial s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.d = d;
    s.dt = dt;
    s.esz = esz;
    s.eps = eps;
    int rc = 0;
    uint8_t* core = NULL;
    int64_t* parent = NULL;
    uint64_t mism = 0, bad = 0, ncore = 0, namb = 0;
    if (build_spatial(&s, data) != 0) {
        rc = 2;
        goto done;
    }
    core = (uint8_t*)calloc((size_t)(n + 1), 1);
    parent = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!core || !parent) {
        rc = 2;
        goto done;
    }
    for (int64_t i = 0; i < n; ++i) parent[i] = i;
    const eps2_t e2 = make_eps2(dt, eps);
    int64_t n3 = 1;
    for (int j = 0; j < d; ++j) n3 *= 3;
    const int64_t ncells = s.n_cells;
    int oom = 0;

    /* pass 1: core flags; pass 2: union of core pairs; pass 3: labels */
    for (int pass = 1; pass <= 3 && !oom; ++pass) {
#pragma omp parallel reduction(+ : mism, bad, ncore, namb) reduction(| : oom)
        {
            const unsigned int thread_csr = _mm_getcsr();
            _mm_setcsr(thread_csr | 0x8040);
            uint64_t* sc = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n3);
            int64_t* rs = (int64_t*)malloc(sizeof(int64_t) * (size_t)n3);
            int64_t* rcn = (int64_t*)malloc(sizeof(int64_t) * (size_t)n3);
            if (!sc || !rs || !rcn) oom = 1;
#pragma omp for schedule(dynamic, 1)
            for (int64_t u = 0; u < ncells; ++u) {
                if (!sc || !rs || !rcn) continue;
                const int64_t nr = neighbor_ranges(&s, s.ucell[u], sc, rs, rcn);
                const int64_t p0 = s.ustart[u], p1 = p0 + s.ucount[u];
                for (int64_t p = p0; p < p1; ++p) {
                    if (pass == 1) {
                        uint64_t cnt = 0;
                        for (int64_t r = 0; r < nr && cnt < min_pts; ++r)
                            for (int64_t q = rs[r]; q < rs[r] + rcn[r] && cnt < min_pts; ++q)
                                cnt += within(&s, p, q, &e2) ? 1u : 0u;
                        core[p] = cnt >= min_pts;
                    } else if (pass == 2) {
                        if (!core[p]) continue;
                        for (int64_t r = 0; r < nr; ++r)
                            for (int64_t q = rs[r]; q < rs[r] + rcn[r] && q < p; ++q) {
                                if (!core[q]) continue;
                                if (uf_find_par(parent, q) == uf_find_par(parent, p)) continue;
                                if (within(&s, p, q, &e2)) uf_unite_par(parent, p, q);
                            }
                    } else {
                        const int64_t g = got[s.orig_of_sorted[p]];
                        if (core[p]) {
                            ++ncore;
                            if (g != -(uf_find_par(parent, p) + 1)) ++mism;
                            continue;
                        }
                        int64_t k0 = -1;   /* first adjacent cluster (root) */
                        int many = 0;      /* two or more adjacent clusters */
                        int got_adj = 0;   /* g = +k for an adjacent cluster k */
                        for (int64_t r = 0; r < nr; ++r)
                            for (int64_t q = rs[r]; q < rs[r] + rcn[r]; ++q) {
                                if (!core[q] || !within(&s, p, q, &e2)) continue;
                                const int64_t k = uf_find_par(parent, q);
                                if (k0 < 0) k0 = k;
                                else if (k != k0) many = 1;
                                if (g == k + 1) got_adj = 1;
                            }
                        if (!many) {
                            if (g != (k0 < 0 ? 0 : k0 + 1)) ++mism;
                        } else {
                            ++namb;
                            if (!got_adj) ++bad;
                        }
                    }
                }
            }
            free(sc);
            free(rs);
            free(rcn);
            _mm_setcsr(thread_csr);
        }
        if (pass == 2) {
            /* flatten: every core points at its root */
            for (int64_t i = 0; i < n; ++i)
                if (core[i]) parent[i] = uf_find_par(parent, i);
        }
    }
    if (oom) rc = 2;
done:
    free(core);
    free(parent);
    free_spatial(&s);
    _mm_setcsr(saved_csr);
    if (mismatched_exact) *mismatched_exact = mism;
    if (bad_ambiguous) *bad_ambiguous = bad;
    if (cores_out) *cores_out = ncore;
    if (ambiguous_out) *ambiguous_out = namb;
    return rc;
}
