// Host-side declarations shared by the translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

namespace hpd {

constexpr int kMaxDims = 8;  // dimensions handled by the device kernels

// An error carrying an hpd_status code; the C-ABI layer turns it into the
// return value and the thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what);
#define HPD_HIP(call) ::hpd::check_hip((call), #call)

// Grow-only device arena.  Allocations are 256-byte aligned bump pointers
// into one hipMalloc'd block; reset() recycles the block for the next call.
// An arena attached to memory it does not own never allocates: reserve()
// beyond its capacity throws HPD_EUNSUPPORTED.
class Arena {
public:
    Arena() = default;
    ~Arena();
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;

    // Use [base, base + cap) (256-byte aligned) instead of allocating.
    void attach(void* base, size_t cap) {
        base_ = static_cast<char*>(base);
        cap_ = cap;
        used_ = 0;
        owned_ = false;
    }
    // Make sure at least `bytes` are available in one block (drops contents).
    void reserve(size_t bytes);
    void reset() { used_ = 0; }
    void* take(size_t bytes);
    template <typename T>
    T* take_n(size_t n) {
        return static_cast<T*>(take(n * sizeof(T)));
    }
    size_t capacity() const { return cap_; }
    const void* base() const { return base_; }
    size_t used() const { return used_; }

private:
    char* base_ = nullptr;
    size_t cap_ = 0;
    size_t used_ = 0;
    bool owned_ = true;
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// A number of items known on the host, or (dev not null) written into device
// memory by an earlier kernel, `host` being then an upper bound that sizes
// the launch (tight: a bound close to the count, such as the number of boxes
// for a list of boxes, rather than a worst case).  The static pipeline (no
// host read-backs) passes the latter throughout, the plain one for lists
// whose length it does not wait for.
struct Count {
    int64_t host = 0;
    const uint32_t* dev = nullptr;
    bool tight = false;  // host is close to the count: one wave per item of it
    __device__ __forceinline__ int64_t get() const {
        return dev ? static_cast<int64_t>(*dev) : host;
    }
};
inline Count host_count(int64_t n) { return Count{n, nullptr}; }

// Exclusive prefix sum of n uint32 values (in may equal out).  `tmp` must
// hold scan_tmp_words(n) uint32.  The total is written to *total_dev when
// total_dev is not null (device memory).
size_t scan_tmp_words(int64_t n);
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                        uint32_t* total_dev, hipStream_t s);
// The same over a length that an earlier kernel wrote (n.dev, at most
// n.host, which sizes the launches); the total goes to out[length].
void exclusive_scan_u32_dev(const uint32_t* in, uint32_t* out, Count n, uint32_t* tmp,
                            hipStream_t s);
// *dst = a[*idx_dev] (a scan's total at a device-written length, moved to a
// fixed place)
void copy_at_count(const uint32_t* a, const uint32_t* idx_dev, uint32_t* dst, hipStream_t s);
// *dst[i] = *src[i] for i < k (device words, one launch)
struct WordCopies {
    const uint32_t* src[8];
    uint32_t* dst[8];
    int k = 0;
    void add(const uint32_t* from, uint32_t* to) {
        src[k] = from;
        dst[k++] = to;
    }
};
void copy_words(const WordCopies& c, hipStream_t s);

// In-place exclusive scans of k <= kMaxScanSet arrays of one length n, each
// array's total to its entry n (every array holds n + 1 entries), in two
// launches for all of them; tmp holds tmp_words uint32 (separate scans
// where that is too few or n exceeds 4096 * 4096).
constexpr int kMaxScanSet = 8;
struct ScanSet {
    uint32_t* a[kMaxScanSet];
};
void exclusive_scan_set_u32(const ScanSet& set, int k, int64_t n, uint32_t* tmp, size_t tmp_words,
                            hipStream_t s, const uint32_t* n_dev = nullptr);

// Stable LSD radix sort of (key, value) pairs by the low `bits` bits of the
// key.  Returns true when the result is in (keys_b, vals_b), false when it is
// in (keys_a, vals_a).  `hist` must hold radix_hist_words(n) uint32 and `tmp`
// scan_tmp_words(radix_hist_words(n)).
size_t radix_hist_words(int64_t n);
// The same with the key's significant bits in device memory (*bits_dev <=
// max_bits, 8-bit digits; passes beyond the bits do nothing): the result is
// always in (keys_a, vals_a).
template <typename K>
void radix_sort_pairs_dev(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                          int max_bits, const int* bits_dev, uint32_t* hist, uint32_t* tmp,
                          hipStream_t s);
template <typename K>
bool radix_sort_pairs(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                      int bits, uint32_t* hist, uint32_t* tmp, hipStream_t s);
// radix_sort_pairs (n > 1, bits > 0, pairs of row index i -> row i) that
// also moves the rows: row i of pts (D values) lands where its pair does,
// in soa (D planes of stride ld); rows_a and rows_b (n * D values each; b
// only with three passes or more) hold the rows between passes.  Reading
// each tile's rows where it lies replaces a gather of random rows after
// the sort.  Instantiated for rows of 4 or 8 bytes (float, D <= 2; double,
// D = 1): wider rows would take LDS that the sort's occupancy needs.
template <typename K, typename T, int D>
bool radix_sort_pairs_rows(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, int64_t n,
                           int bits, uint32_t* hist, uint32_t* tmp, const T* pts, T* rows_a,
                           T* rows_b, T* soa, int64_t ld, hipStream_t s);
int radix_pass_count(int bits);  // passes of radix_sort_pairs over `bits` key bits
template <typename T, int D>
constexpr bool sort_moves_rows() {
    return (std::is_same<T, float>::value && D <= 2) || (std::is_same<T, double>::value && D == 1);
}

// Grid geometry of the eps-grid, computed on the host from the min/max of
// the data exactly as the reference does (spatial_index.h:124-142).
struct Geometry {
    int d = 0;
    uint64_t cells[kMaxDims] = {};   // cells per input dimension
    int swapped[kMaxDims] = {};      // dimension order, fewest cells first
    uint64_t stride[kMaxDims] = {};  // stride of swapped dimension k
    uint64_t ncell[kMaxDims] = {};   // cells of swapped dimension k
    uint64_t total = 1;              // product of cells (wrapping)
    bool overflow = false;           // product exceeded 2^64
    int key_bits = 0;                // bits of the largest cell id
};

// Device-side parameter blocks.
struct CellParams {
    int d;
    int swapped[kMaxDims];
    uint64_t stride[kMaxDims];
    uint64_t ncell[kMaxDims];
};

// Everything a kernel needs to know about one clustering job (one point
// cloud with its own eps and minPts).  A plain call is a batch of one job.
// Jobs are laid out back to back: rows [row_begin, row_end) of the
// concatenated input, cell keys [key_base, key_base + total cells).  After
// the sort by key, a job's sorted positions are again [row_begin, row_end).
template <typename T>
struct JobParams {
    CellParams cp;
    uint64_t key_base = 0;
    int64_t row_begin = 0;
    int64_t row_end = 0;
    T minv[kMaxDims] = {};  // per-dimension minimum: the grid origin
    T eps2{};               // EPS2 converted to T
    float rcp_f = 0.0f;     // 1/eps in float (float and integer data)
    double rcp_d = 0.0;     // 1/eps in double (double data)
    double gap2_max = 0.0;  // box pairs farther apart than this cannot connect
    uint32_t min_pts = 0;
};

// job index of a row (rows sorted by job): last j with row_begin <= row
template <typename T>
__device__ __forceinline__ int32_t job_of_row(const JobParams<T>* __restrict__ jobs, int32_t njobs,
                                              int64_t row) {
    int32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].row_begin <= row) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The same for the rows of a wave that are consecutive and ascending over
// its lanes: one search for the first active lane's row (wave-uniform,
// scalar loads), then each lane steps past the jobs that end before its
// row (jobs hold many rows, so rarely more than one step).
template <typename T>
__device__ __forceinline__ int32_t job_of_row_wave(const JobParams<T>* __restrict__ jobs,
                                                   int32_t njobs, int64_t row) {
    if (njobs == 1) return 0;
    const int64_t row0 = static_cast<int64_t>(__builtin_amdgcn_readfirstlane(
                             static_cast<int>(row & 0xffffffff))) |
                         (static_cast<int64_t>(__builtin_amdgcn_readfirstlane(
                              static_cast<int>(row >> 32)))
                          << 32);
    int32_t j = job_of_row(jobs, njobs, row0);
    while (j + 1 < njobs && jobs[j].row_end <= row) ++j;
    return j;
}

}  // namespace hpd
