// Segmented-sort ORACLE: a CPU restatement of the reference's sequential path.
//
// THIS IS TEST INFRASTRUCTURE. It is the checker for the HIP product library and
// the CPU baseline timed beside it; it is never the thing measured or shipped.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// What it restates (reference file:line, paths relative to the reference root):
//   * per-segment std::stable_sort ............ src/mseq.cu:44-49
//   * stable pair sort by key only ............ src/bb_segsort/main_BKP.cu:122-144
//   * radix key order (bit twiddle) ........... lib/cub/util_type.cuh:904-911 (unsigned),
//                                               :947-955 (signed), :1017-1027 (float/double)
//   * segment semantics (end<=begin is empty,
//     gaps untouched, inputs unmodified) ...... lib/cub/device/device_segmented_radix_sort.cuh:87,140-153;
//                                               lib/cub/device/dispatch/dispatch_radix_sort.cuh:380-385
//   * [begin_bit,end_bit) sub-range ........... dispatch_radix_sort.cuh:1359-1366
//   * descending order (bit-flipped digits,
//     ties kept in input order) ............... device_segmented_radix_sort.cuh:302-327
//   * fix-pass (segment id packed above the
//     key, one global stable sort) ............ src/fixseq.cu:61-83, widened to a 64-bit
//                                               composite so that it cannot overflow
//   * temp-storage size of the cub path ....... lib/cub/util_device.cuh:58-101
//
// Pinning (see tests/test_oracle.py): cub's documented known-answer vectors, the
// reference's committed src/cub.out dump and src/mergesort2 inputs, exhaustive
// permutations of tiny multisets, and the float closed forms (-0.0 < +0.0,
// denormals by magnitude) checked against numpy's own ordering.
//
// Build: oracle/Makefile -> oracle/liboracle.so (ctypes), g++ -O2 -fopenmp.

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>

namespace {

enum KeyKind { K_U32 = 0, K_I32 = 1, K_F32 = 2, K_U64 = 3, K_I64 = 4, K_F64 = 5 };

inline bool is64(int kk) { return kk >= K_U64; }

// cub Traits<T>::TwiddleIn (util_type.cuh:904-911, 947-955, 1017-1027).
inline uint64_t twiddle_in(uint64_t bits, int kk) {
    switch (kk) {
    case K_U32: case K_U64: return bits;
    case K_I32: return bits ^ 0x80000000ull;
    case K_I64: return bits ^ 0x8000000000000000ull;
    case K_F32: {
        uint32_t b = (uint32_t)bits;
        uint32_t mask = (b & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;
        return (uint64_t)(b ^ mask);
    }
    case K_F64: {
        uint64_t mask = (bits & 0x8000000000000000ull) ? ~0ull : 0x8000000000000000ull;
        return bits ^ mask;
    }
    }
    return bits;
}

inline uint64_t load_key(const void* keys, int64_t i, int kk) {
    if (is64(kk)) return ((const uint64_t*)keys)[i];
    return ((const uint32_t*)keys)[i];
}

inline void store_key(void* keys, int64_t i, uint64_t v, int kk) {
    if (is64(kk)) ((uint64_t*)keys)[i] = v;
    else ((uint32_t*)keys)[i] = (uint32_t)v;
}

inline void copy_val(void* dst, int64_t di, const void* src, int64_t si, int vbytes) {
    if (vbytes == 4) ((uint32_t*)dst)[di] = ((const uint32_t*)src)[si];
    else if (vbytes == 8) ((uint64_t*)dst)[di] = ((const uint64_t*)src)[si];
}

// The comparison key the reference's radix path orders by: twiddled bits,
// bit-flipped when descending, restricted to [begin_bit, end_bit).
inline uint64_t order_key(uint64_t raw, int kk, int begin_bit, int end_bit, int descending) {
    int kbits = is64(kk) ? 64 : 32;
    uint64_t t = twiddle_in(raw, kk);
    if (descending) t = ~t & (kbits == 64 ? ~0ull : 0xFFFFFFFFull);
    int nb = end_bit - begin_bit;
    if (nb <= 0) return 0;
    uint64_t mask = (nb >= 64) ? ~0ull : ((1ull << nb) - 1ull);
    return (t >> begin_bit) & mask;
}

void sort_one_segment(int kk, int vbytes, const void* keys_in, const void* vals_in, void* keys_out,
                      void* vals_out, int64_t b, int64_t e, int begin_bit, int end_bit, int descending,
                      std::vector<std::pair<uint64_t, int64_t>>& buf) {
    int64_t n = e - b;
    if (n <= 0) return;  // dispatch_radix_sort.cuh:383-385: empty segment, nothing written
    buf.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        buf[(size_t)i] = {order_key(load_key(keys_in, b + i, kk), kk, begin_bit, end_bit, descending), b + i};
    // mseq.cu:47 / main_BKP.cu:135: std::stable_sort, comparing the key only.
    std::stable_sort(buf.begin(), buf.end(),
                     [](const std::pair<uint64_t, int64_t>& x, const std::pair<uint64_t, int64_t>& y) {
                         return x.first < y.first;
                     });
    for (int64_t i = 0; i < n; ++i) {
        int64_t src = buf[(size_t)i].second;
        store_key(keys_out, b + i, load_key(keys_in, src, kk), kk);
        if (vbytes && vals_out) copy_val(vals_out, b + i, vals_in, src, vbytes);
    }
}

}  // namespace

extern "C" {

int oracle_version() { return 1; }

// Per-segment stable sort: the semantics of cub::DeviceSegmentedRadixSort::SortPairs /
// SortKeys (device_segmented_radix_sort.cuh:136-177, 532-570) restated with
// std::stable_sort per segment (mseq.cu:44-49). vbytes = 0 means keys only.
// Positions outside every segment are left as the caller filled them.
// threads <= 0 uses every OpenMP thread; returns the thread count used.
int oracle_segmented_sort(int key_kind, int vbytes, const void* keys_in, const void* vals_in,
                          void* keys_out, void* vals_out, int64_t num_items, int num_segments,
                          const int* begin, const int* end, int begin_bit, int end_bit,
                          int descending, int threads) {
    (void)num_items;
    int nt = threads > 0 ? threads : omp_get_max_threads();
    if (nt == 1) {
        std::vector<std::pair<uint64_t, int64_t>> buf;
        for (int s = 0; s < num_segments; ++s)
            sort_one_segment(key_kind, vbytes, keys_in, vals_in, keys_out, vals_out, begin[s], end[s],
                             begin_bit, end_bit, descending, buf);
        return 1;
    }
#pragma omp parallel num_threads(nt)
    {
        std::vector<std::pair<uint64_t, int64_t>> buf;
#pragma omp for schedule(dynamic, 16)
        for (int s = 0; s < num_segments; ++s)
            sort_one_segment(key_kind, vbytes, keys_in, vals_in, keys_out, vals_out, begin[s], end[s],
                             begin_bit, end_bit, descending, buf);
    }
    return nt;
}

// CPU fix-pass (fixseq.cu:61-83), widened: the reference adds seg<<msb to a
// 32-bit key, which overflows once msb + log2(S) > 32 (fixcub.cu:52-53). Here
// the composite is (segment id << 32 | twiddled 32-bit key) in 64 bits, one
// std::stable_sort over all items of the CSR range [offsets[0], offsets[S]),
// then the segment id is stripped. 32-bit keys only, ascending, full bit range.
// An independent second path to cross-check oracle_segmented_sort.
int oracle_fixpass_sort(int key_kind, int vbytes, const void* keys_in, const void* vals_in,
                        void* keys_out, void* vals_out, int num_segments, const int* offsets) {
    if (is64(key_kind)) return -1;
    int64_t lo = offsets[0], hi = offsets[num_segments];
    int64_t n = hi - lo;
    if (n <= 0) return 0;
    std::vector<std::pair<uint64_t, int64_t>> comp((size_t)n);
    for (int s = 0; s < num_segments; ++s)
        for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i)
            comp[(size_t)(i - lo)] = {((uint64_t)s << 32) | twiddle_in(load_key(keys_in, i, key_kind), key_kind), i};
    std::stable_sort(comp.begin(), comp.end(),
                     [](const std::pair<uint64_t, int64_t>& x, const std::pair<uint64_t, int64_t>& y) {
                         return x.first < y.first;
                     });
    for (int64_t i = 0; i < n; ++i) {
        int64_t src = comp[(size_t)i].second;
        store_key(keys_out, lo + i, load_key(keys_in, src, key_kind), key_kind);
        if (vbytes && vals_out) copy_val(vals_out, lo + i, vals_in, src, vbytes);
    }
    return 1;
}

// Order key of one raw key (twiddle + descending flip + bit range); exposed so
// tests can pin the twiddle closed forms directly.
uint64_t oracle_order_key(uint64_t raw, int key_kind, int begin_bit, int end_bit, int descending) {
    return order_key(raw, key_kind, begin_bit, end_bit, descending);
}

// Temp bytes the cub segmented path would request for (num_items, K, V):
// two 256-aligned ping-pong buffers + 255 slack (dispatch_radix_sort.cuh:1343-1356,
// util_device.cuh:66-80); reported for the driver-compatibility tests only.
int64_t oracle_cub_temp_bytes(int64_t num_items, int key_bytes, int vbytes) {
    auto r = [](int64_t x) { return (x + 255) & ~int64_t(255); };
    int64_t b = r(num_items * key_bytes) + r(num_items * (int64_t)vbytes) + 255;
    return b == 0 ? 1 : b;
}

// Backward of the differentiable pair sort: scatter upstream gradients through the
// saved per-row permutation, gin[off[r] + perm[i]] = gout[i] for i in row r.
void oracle_perm_scatter_f32(const float* gout, const int* perm, const int* offsets, int rows, float* gin) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int r = 0; r < rows; ++r)
        for (int i = offsets[r]; i < offsets[r + 1]; ++i) gin[offsets[r] + perm[i]] = gout[i];
}

// Wall time in seconds of oracle_segmented_sort on the given arrays, steady clock.
double oracle_time_segmented_sort(int key_kind, int vbytes, const void* keys_in, const void* vals_in,
                                  void* keys_out, void* vals_out, int64_t num_items, int num_segments,
                                  const int* begin, const int* end, int threads) {
    auto t0 = std::chrono::steady_clock::now();
    oracle_segmented_sort(key_kind, vbytes, keys_in, vals_in, keys_out, vals_out, num_items, num_segments,
                          begin, end, 0, is64(key_kind) ? 64 : 32, 0, threads);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int oracle_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
