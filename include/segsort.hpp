// segsort.hpp - header-only C++ facade with cub::DeviceSegmentedRadixSort's
// signatures over the libsegsort.so C-ABI (segsort.h), so that code written
// against cub (lib/cub/device/device_segmented_radix_sort.cuh:136-855) compiles
// against this library by changing the namespace.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <type_traits>

#include "segsort.h"

namespace segsort {

// cub::DoubleBuffer (lib/cub/util_type.cuh): two buffers and a selector naming
// the one that holds the current data.
template <typename T>
struct DoubleBuffer {
    T* d_buffers[2];
    int selector;
    DoubleBuffer() : d_buffers{nullptr, nullptr}, selector(0) {}
    DoubleBuffer(T* d_current, T* d_alternate) : d_buffers{d_current, d_alternate}, selector(0) {}
    T* Current() { return d_buffers[selector]; }
    T* Alternate() { return d_buffers[selector ^ 1]; }
};

template <typename K>
constexpr int key_type_of() {
    static_assert(sizeof(K) == 4 || sizeof(K) == 8, "keys of 4 or 8 bytes");
    if (std::is_floating_point<K>::value) return sizeof(K) == 4 ? SEGSORT_KEY_F32 : SEGSORT_KEY_F64;
    if (std::is_signed<K>::value) return sizeof(K) == 4 ? SEGSORT_KEY_I32 : SEGSORT_KEY_I64;
    return sizeof(K) == 4 ? SEGSORT_KEY_U32 : SEGSORT_KEY_U64;
}

namespace detail {
inline hipError_t finish(int rc, hipStream_t stream, bool debug_synchronous, const char* what) {
    hipError_t e = (hipError_t)rc;
    if (e == hipSuccess && debug_synchronous) e = hipStreamSynchronize(stream);
    if (debug_synchronous && e != hipSuccess) fprintf(stderr, "%s: %s\n", what, segsort_last_error());
    return e;
}

template <typename K, typename V>
hipError_t sort(void* d_temp, size_t& temp_bytes, const K* ki, K* ko, const V* vi, V* vo, int num_items,
                int num_segments, const int* d_begin, const int* d_end, int begin_bit, int end_bit, bool descending,
                hipStream_t stream, bool debug_synchronous, int algo, int max_segment_length) {
    constexpr int vb = std::is_void<V>::value ? 0 : (int)sizeof(typename std::conditional<std::is_void<V>::value,
                                                                                          char, V>::type);
    static_assert(vb == 0 || vb == 4 || vb == 8, "values of 4 or 8 bytes");
    const int rc = max_segment_length > 0
                       ? segsort_sort_hinted(d_temp, &temp_bytes, ki, ko, vi, vo, num_items, num_segments, d_begin,
                                             d_end, begin_bit, end_bit, key_type_of<K>(), vb, descending ? 1 : 0, algo,
                                             max_segment_length, (void*)stream)
                       : segsort_sort(d_temp, &temp_bytes, ki, ko, vi, vo, num_items, num_segments, d_begin, d_end,
                                      begin_bit, end_bit, key_type_of<K>(), vb, descending ? 1 : 0, algo,
                                      (void*)stream);
    return finish(rc, stream, debug_synchronous && d_temp != nullptr, "segsort");
}
}  // namespace detail

// cub::DeviceSegmentedRadixSort with the same argument order and meaning; the
// trailing `algo` selects the batched (default), plain or fix-pass form, and the
// optional `max_segment_length` (0: none) is the caller's bound on the segment
// lengths (segsort_sort_hinted: the tuning for that length distribution, and no
// MSD launches when every segment fits a tile; a wrong bound is sorted correctly,
// more slowly).
struct DeviceSegmentedRadixSort {
    template <typename KeyT, typename ValueT>
    static hipError_t SortPairs(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out, int num_items,
                                int num_segments, const int* d_begin_offsets, const int* d_end_offsets,
                                int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = 0,
                                bool debug_synchronous = false, int algo = SEGSORT_ALGO_BATCHED,
                                int max_segment_length = 0) {
        return detail::sort<KeyT, ValueT>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in,
                                          d_values_out, num_items, num_segments, d_begin_offsets, d_end_offsets,
                                          begin_bit, end_bit, false, stream, debug_synchronous,
                                          algo, max_segment_length);
    }

    template <typename KeyT, typename ValueT>
    static hipError_t SortPairsDescending(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                          KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out,
                                          int num_items, int num_segments, const int* d_begin_offsets,
                                          const int* d_end_offsets, int begin_bit = 0,
                                          int end_bit = sizeof(KeyT) * 8, hipStream_t stream = 0,
                                          bool debug_synchronous = false, int algo = SEGSORT_ALGO_BATCHED,
                                          int max_segment_length = 0) {
        return detail::sort<KeyT, ValueT>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in,
                                          d_values_out, num_items, num_segments, d_begin_offsets, d_end_offsets,
                                          begin_bit, end_bit, true, stream, debug_synchronous,
                                          algo, max_segment_length);
    }

    template <typename KeyT>
    static hipError_t SortKeys(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                               KeyT* d_keys_out, int num_items, int num_segments, const int* d_begin_offsets,
                               const int* d_end_offsets, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                               hipStream_t stream = 0, bool debug_synchronous = false,
                               int algo = SEGSORT_ALGO_BATCHED, int max_segment_length = 0) {
        return detail::sort<KeyT, void>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, nullptr, nullptr,
                                        num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit, end_bit,
                                        false, stream, debug_synchronous, algo, max_segment_length);
    }

    template <typename KeyT>
    static hipError_t SortKeysDescending(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                         KeyT* d_keys_out, int num_items, int num_segments,
                                         const int* d_begin_offsets, const int* d_end_offsets, int begin_bit = 0,
                                         int end_bit = sizeof(KeyT) * 8, hipStream_t stream = 0,
                                         bool debug_synchronous = false, int algo = SEGSORT_ALGO_BATCHED,
                                         int max_segment_length = 0) {
        return detail::sort<KeyT, void>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, nullptr, nullptr,
                                        num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit, end_bit,
                                        true, stream, debug_synchronous, algo, max_segment_length);
    }

    // DoubleBuffer forms (device_segmented_radix_sort.cuh:246-260, 445-459,
    // 630-643, 811-825): the sorted data ends in the alternate buffer and the
    // selector is flipped.
    template <typename KeyT, typename ValueT>
    static hipError_t SortPairs(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                                DoubleBuffer<ValueT>& d_values, int num_items, int num_segments,
                                const int* d_begin_offsets, const int* d_end_offsets, int begin_bit = 0,
                                int end_bit = sizeof(KeyT) * 8, hipStream_t stream = 0,
                                bool debug_synchronous = false, int algo = SEGSORT_ALGO_BATCHED,
                                int max_segment_length = 0) {
        return pairs_db(d_temp_storage, temp_storage_bytes, d_keys, d_values, num_items, num_segments,
                        d_begin_offsets, d_end_offsets, begin_bit, end_bit, false, stream, debug_synchronous,
                        algo, max_segment_length);
    }

    template <typename KeyT>
    static hipError_t SortKeys(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                               int num_items, int num_segments, const int* d_begin_offsets, const int* d_end_offsets,
                               int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = 0,
                               bool debug_synchronous = false, int algo = SEGSORT_ALGO_BATCHED,
                               int max_segment_length = 0) {
        return keys_db(d_temp_storage, temp_storage_bytes, d_keys, num_items, num_segments, d_begin_offsets,
                       d_end_offsets, begin_bit, end_bit, false, stream, debug_synchronous, algo, max_segment_length);
    }

    // device_segmented_radix_sort.cuh:445-459
    template <typename KeyT, typename ValueT>
    static hipError_t SortPairsDescending(void* d_temp_storage, size_t& temp_storage_bytes,
                                          DoubleBuffer<KeyT>& d_keys, DoubleBuffer<ValueT>& d_values, int num_items,
                                          int num_segments, const int* d_begin_offsets, const int* d_end_offsets,
                                          int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = 0,
                                          bool debug_synchronous = false, int algo = SEGSORT_ALGO_BATCHED,
                                          int max_segment_length = 0) {
        return pairs_db(d_temp_storage, temp_storage_bytes, d_keys, d_values, num_items, num_segments,
                        d_begin_offsets, d_end_offsets, begin_bit, end_bit, true, stream, debug_synchronous,
                        algo, max_segment_length);
    }

    // device_segmented_radix_sort.cuh:811-825
    template <typename KeyT>
    static hipError_t SortKeysDescending(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                                         int num_items, int num_segments, const int* d_begin_offsets,
                                         const int* d_end_offsets, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                                         hipStream_t stream = 0, bool debug_synchronous = false,
                                         int algo = SEGSORT_ALGO_BATCHED, int max_segment_length = 0) {
        return keys_db(d_temp_storage, temp_storage_bytes, d_keys, num_items, num_segments, d_begin_offsets,
                       d_end_offsets, begin_bit, end_bit, true, stream, debug_synchronous, algo, max_segment_length);
    }

private:
    template <typename KeyT, typename ValueT>
    static hipError_t pairs_db(void* d_temp, size_t& temp_bytes, DoubleBuffer<KeyT>& d_keys,
                               DoubleBuffer<ValueT>& d_values, int num_items, int num_segments, const int* d_begin,
                               const int* d_end, int begin_bit, int end_bit, bool descending, hipStream_t stream,
                               bool debug_synchronous, int algo, int max_segment_length) {
        hipError_t e = detail::sort<KeyT, ValueT>(d_temp, temp_bytes, d_keys.Current(), d_keys.Alternate(),
                                                  d_values.Current(), d_values.Alternate(), num_items, num_segments,
                                                  d_begin, d_end, begin_bit, end_bit, descending, stream,
                                                  debug_synchronous, algo, max_segment_length);
        if (e == hipSuccess && d_temp) {
            d_keys.selector ^= 1;
            d_values.selector ^= 1;
        }
        return e;
    }

    template <typename KeyT>
    static hipError_t keys_db(void* d_temp, size_t& temp_bytes, DoubleBuffer<KeyT>& d_keys, int num_items,
                              int num_segments, const int* d_begin, const int* d_end, int begin_bit, int end_bit,
                              bool descending, hipStream_t stream, bool debug_synchronous, int algo,
                              int max_segment_length) {
        hipError_t e = detail::sort<KeyT, void>(d_temp, temp_bytes, d_keys.Current(), d_keys.Alternate(), nullptr,
                                                nullptr, num_items, num_segments, d_begin, d_end, begin_bit, end_bit,
                                                descending, stream, debug_synchronous, algo, max_segment_length);
        if (e == hipSuccess && d_temp) d_keys.selector ^= 1;
        return e;
    }
};

}  // namespace segsort
