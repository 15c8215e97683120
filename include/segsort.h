/*
 * segsort.h - C-ABI of the MI355X (gfx950) segmented sort library libsegsort.so.
 *
 * Every entry point takes plain device pointers, sizes and an opaque stream
 * handle (a hipStream_t passed as void*, NULL = the null stream), and returns
 * an int holding a hipError_t value (0 = hipSuccess). Nothing throws or exits
 * across this boundary.
 *
 * The sort entry points replace cub::DeviceSegmentedRadixSort as the
 * reference's drivers call it:
 *   SortPairs  lib/cub/device/device_segmented_radix_sort.cuh:136-177
 *   SortKeys   lib/cub/device/device_segmented_radix_sort.cuh:532-570
 *   SortPairsDescending / SortKeysDescending  :302-350, :685-726
 *   callers    src/radixseg.cu:116-124, src/fixcub.cu:141-165,
 *              src/mergeseg.cu:118, src/bb_segsort/main.cu:102
 * with cub's temp-storage protocol (util_device.cuh:58-101,
 * dispatch_radix_sort.cuh:1349-1356): d_temp == NULL writes the required size
 * to *temp_bytes and returns success; a too-small *temp_bytes returns
 * hipErrorInvalidValue. Inputs are never modified; output positions outside
 * every segment are never written; a segment with end <= begin is empty.
 *
 * The caller owns all device memory, including temp; no entry point except
 * the in-place adapters (segsort_bb_segsort, segsort_mgpu_segmented_sort)
 * allocates. Calls are stream-ordered and asynchronous, except the PLAIN
 * algorithm, which reads its class counts back to the host (as the
 * reference's bb_segsort does, src/bb_segsort/bb_bin.h:46).
 */
#ifndef SEGSORT_H
#define SEGSORT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Key types (the cub key types of the reference's drivers are u32). */
enum segsort_key_type {
    SEGSORT_KEY_U32 = 0,
    SEGSORT_KEY_I32 = 1,
    SEGSORT_KEY_F32 = 2,
    SEGSORT_KEY_U64 = 3,
    SEGSORT_KEY_I64 = 4,
    SEGSORT_KEY_F64 = 5
};

/* Algorithms. BATCHED: device binning into tiers, no host sync, capturable.
 * PLAIN: one launch per segment-length class after a host read-back of the
 * class counts, one workgroup per segment beyond the largest LDS tile.
 * FIXPASS: segment id packed above the key in a 64-bit composite and one
 * global LSD radix sort (src/fixcub.cu:140-166, widened); 32-bit keys only,
 * segments must tile [d_begin[0], d_end[num_segments-1]) in order. */
enum segsort_algo {
    SEGSORT_ALGO_BATCHED = 0,
    SEGSORT_ALGO_PLAIN = 1,
    SEGSORT_ALGO_FIXPASS = 2
};

/* Stable segmented sort of (key, value) pairs or of keys alone.
 * value_bytes: 0 (keys only; vals_in/vals_out ignored), 4 or 8.
 * begin_bit/end_bit: key bits [begin_bit, end_bit) of the radix order, as in
 * cub (0 and 8*sizeof(key) for the whole key). descending: 0 or 1.
 * Replaces DeviceSegmentedRadixSort::SortPairs/SortKeys(/Descending),
 * device_segmented_radix_sort.cuh:136-177, 302-350, 532-570, 685-726. */
int segsort_sort(void* d_temp, size_t* temp_bytes,
                 const void* d_keys_in, void* d_keys_out,
                 const void* d_vals_in, void* d_vals_out,
                 int num_items, int num_segments,
                 const int* d_begin_offsets, const int* d_end_offsets,
                 int begin_bit, int end_bit,
                 int key_type, int value_bytes, int descending, int algo,
                 void* stream);

/* segsort_sort with the caller's bound on the segment lengths: every segment
 * has end - begin <= max_segment_length (0: no bound, which is segsort_sort).
 * No counterpart in cub, whose calls cannot see the offsets on the host
 * either; callers that hold their offsets on the host (the reference's drivers
 * read them from stdin, src/radixseg.cu:77-90) know the bound for free.
 * The bound does two things. It selects the tuning measured for that length
 * distribution (the table's entries carry ceil(log2 max length) as a third
 * feature beside round(log2 num_items) and round(log2 mean length)). And when
 * it is at most the largest LDS tile of that tuning, the batched form launches
 * no multi-workgroup MSD kernels (about 8 us of launches that would find no
 * work). A wrong bound never gives a wrong result: a segment longer than the
 * largest tile is then sorted by one workgroup (all digit passes over global
 * memory, as the plain form's per-segment LSD), only slower. The temp query
 * must pass the same bound as the call (the tuning it selects sizes temp).
 * Other arguments as segsort_sort. */
int segsort_sort_hinted(void* d_temp, size_t* temp_bytes,
                        const void* d_keys_in, void* d_keys_out,
                        const void* d_vals_in, void* d_vals_out,
                        int num_items, int num_segments,
                        const int* d_begin_offsets, const int* d_end_offsets,
                        int begin_bit, int end_bit,
                        int key_type, int value_bytes, int descending, int algo,
                        int max_segment_length, void* stream);

/* segsort_sort with the batched algorithm and no temp argument, for callers
 * that cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference
 * (cub's calls always take caller temp). The temp storage is a static device
 * array of segsort_noalloc_max_bytes() bytes (64 MiB), one per GPU, held until
 * the process ends; its address is looked up once per device on the first
 * call (hipGetSymbolAddress). A call whose temp storage would exceed it
 * returns hipErrorInvalidValue (segsort_last_error says how much it needs;
 * about 2^21 u32 pairs fit). The call allocates nothing and creates no
 * stream, event or graph: kernels are launched directly on `stream`, on that
 * stream alone. Every call on one GPU uses the same array, which keeps the
 * last call's contents (each call clears what must start at zero): calls that
 * overlap on one GPU (other streams, other threads) race. Arguments as in
 * segsort_sort. */
int segsort_sort_noalloc(const void* d_keys_in, void* d_keys_out, const void* d_vals_in, void* d_vals_out,
                         int num_items, int num_segments, const int* d_begin_offsets, const int* d_end_offsets,
                         int begin_bit, int end_bit, int key_type, int value_bytes, int descending, void* stream);
size_t segsort_noalloc_max_bytes(void);

/* Typed shorthands of segsort_sort for the reference's u32/u32 drivers
 * (src/radixseg.cu:116-124 calls SortPairs<uint, uint>). */
int segsort_sort_pairs_u32(void* d_temp, size_t* temp_bytes,
                           const uint32_t* d_keys_in, uint32_t* d_keys_out,
                           const uint32_t* d_vals_in, uint32_t* d_vals_out,
                           int num_items, int num_segments,
                           const int* d_begin_offsets, const int* d_end_offsets,
                           int algo, void* stream);
int segsort_sort_keys_u32(void* d_temp, size_t* temp_bytes,
                          const uint32_t* d_keys_in, uint32_t* d_keys_out,
                          int num_items, int num_segments,
                          const int* d_begin_offsets, const int* d_end_offsets,
                          int algo, void* stream);

/* In-place adapter for the bb_segsort driver: int bb_segsort(K*, T*, int n,
 * int* d_segs, int length) (src/bb_segsort/bb_segsort.h:36). d_segs holds
 * `num_segments` begin offsets; the last segment ends at num_items. Sorts
 * keys and values in place, allocating and freeing its own temp (as the
 * reference does, bb_segsort.h:39-56). Stable, unlike the reference.
 * Returns 1 on success and 0 on failure, as the reference does. */
int segsort_bb_segsort(void* d_keys, void* d_vals, int num_items, const int* d_segs, int num_segments,
                       int key_type, int value_bytes, void* stream);

/* In-place adapter for mgpu::segmented_sort(keys, vals, count, segments,
 * num_segments, less, context) (lib/moderngpu/kernel_segsort.hxx:405-416,
 * called from src/mergeseg.cu:118): begin offsets only, the last segment
 * ends at count; d_vals may be NULL (keys only, kernel_segsort.hxx:432-439).
 * Allocates and frees its own temp. Returns a hipError_t value. */
int segsort_mgpu_segmented_sort(void* d_keys, void* d_vals, int count, const int* d_segments, int num_segments,
                                int key_type, int value_bytes, void* stream);

/* Comparators of the moderngpu adapters: mgpu::less_t / mgpu::greater_t. */
enum segsort_comp { SEGSORT_LESS = 0, SEGSORT_GREATER = 1 };

/* mgpu::segmented_sort with its comparator argument (kernel_segsort.hxx:405-416):
 * comp SEGSORT_LESS (ascending) or SEGSORT_GREATER (descending; equal keys
 * keep their input order, as in mgpu's stable merge sort). Keys compare in
 * the radix order (floats: -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf <
 * +NaN), a total order where a comparator on NaN is undefined. d_vals may be
 * NULL. Allocates and frees its own temp. Returns a hipError_t value. */
int segsort_mgpu_segmented_sort_comp(void* d_keys, void* d_vals, int count, const int* d_segments, int num_segments,
                                     int key_type, int value_bytes, int comp, void* stream);

/* mgpu::segmented_sort_indices (kernel_segsort.hxx:418-429): sorts the keys in
 * place and writes to d_indices[i] the input position of the key now at i
 * (positions outside every segment get their own index). Begin offsets only,
 * the last segment ends at count. Allocates and frees its own temp. */
int segsort_mgpu_segmented_sort_indices(void* d_keys, int* d_indices, int count, const int* d_segments,
                                        int num_segments, int key_type, int comp, void* stream);

/* Per-segment multi-stream baseline: every segment sorted by its own
 * device-wide sort, the segments dealt over num_streams (1..32) streams round
 * robin (schedule 0) or each to the stream with the fewest items so far
 * (schedule 1). Replaces the loops of src/ncub.cu:113-145
 * (cub::DeviceRadixSort::SortKeys per segment over NUM_STREAMS streams) and
 * src/newthrust.cu:91-167 (thrust::sort per segment, round robin or SCHEDULE).
 * h_offsets is a HOST array of num_segments + 1 ascending offsets (segment i
 * is [h_offsets[i], h_offsets[i+1])), needed by the temp query as well; items
 * outside every segment are not written. Temp protocol as segsort_sort. The
 * call synchronises `stream` once (it uploads the segment lengths) and so
 * cannot be captured; the sorts themselves run on the fan-out streams, joined
 * back into `stream` before it returns. A comparison point, not a fast path:
 * it launches one batched sequence per non-empty segment. */
int segsort_multistream_sort(void* d_temp, size_t* temp_bytes, const void* d_keys_in, void* d_keys_out,
                             const void* d_vals_in, void* d_vals_out, int num_items, int num_segments,
                             const int* h_offsets, int begin_bit, int end_bit, int key_type, int value_bytes,
                             int descending, int num_streams, int schedule, void* stream);

/* Differentiable pair sort, forward helper: d_perm[i] = i - d_begin[r] for
 * every i of row r (the row-local index vector that a stable pair sort then
 * carries as values). One workgroup per row. */
int segsort_row_iota(const int* d_begin_offsets, const int* d_end_offsets, int rows, int* d_perm, void* stream);

/* Differentiable pair sort, backward: d_grad_in[d_begin[r] + d_perm[i]] =
 * d_grad_out[i] for every i of row r; positions outside every row are left
 * untouched. d_perm holds the row-local stable permutation of the forward
 * (each row's entries a permutation of 0..n-1; rows of at most 8192 items are
 * permuted in LDS and written out in order). */
int segsort_perm_scatter_f32(const float* d_grad_out, const int* d_perm,
                             const int* d_begin_offsets, const int* d_end_offsets, int rows,
                             float* d_grad_in, void* stream);

/* Tuning of the batched and plain forms. Every call picks its tuning inside
 * the library, as cub picks its policy inside the call
 * (dispatch_radix_sort.cuh:441-661): from a table built in at compile time
 * (the autotuner's measurements, one entry per signature key width, value
 * width, round(log2 num_items), round(log2 mean segment length); the nearest
 * entry of the same widths when the signature is not in it), host arithmetic
 * only, no device read. segsort_set_tuning installs an override for every call
 * with the given key and value widths (the autotuner times candidates this
 * way); segsort_reset_tuning removes it (key_width 0: every override).
 * wave_cap: ranges of 2..wave_cap items use the wave tier (2 <= wave_cap <= 512).
 * tile_enable: bit c (c = 0..3) enables LDS tile class c (capacities given
 * by segsort_tile_caps); bit 4 + g (g = 0..3) enables the wave-group class
 * that sorts 64 / (4 << g) ranges of at most 4 << g items per wave. Ranges go
 * to the smallest enabled class that holds them (groups, then the wave tier
 * if wave_cap is larger, then the tiles); ranges above the largest enabled
 * class go to the multi-workgroup MSD tier (batched) or the
 * one-workgroup-per-segment LSD (plain). Bit 8 (0x100): the batched form's
 * 16-item tile classes take their ranges longest first (length buckets
 * filled by the binning pass); it pays off for mixed lengths and costs a few
 * percent for equal ones. Bit 9 (0x200): the 4096-item tile class without its
 * half-depth build for ranges of at most 2048 items (faster when that class
 * holds full-size ranges). Bit 10 (0x400): with wave_cap above 256, ranges of
 * at most 256 items and the longer wave-tier ranges are two classes with
 * builds and launches of their own. Default 0xEF.
 * grid_mult: persistent workgroups per CU per resident slot (1..8).
 * concurrent: 1 sorts the ranges that fit a class on a second stream while
 * the multi-workgroup tier partitions the longer segments on the caller's
 * stream (joined back before the call's work ends); 2 does so only for the
 * ranges of the 16-item tile classes (256x16, 512x16, 1024x16), the smaller
 * classes following the multi-workgroup tier on the caller's stream; 0 keeps
 * every launch on the caller's stream. Under stream capture (the caller's graph or
 * the library's own replay) the fork and join are recorded too, as two
 * parallel branches of the graph.
 * key_width: 4 or 8; value_bytes: 0, 4 or 8. Return a hipError_t value.
 * segsort_get_tuning writes the slot's override, or the built-in default
 * (256, 0xEF, 1, 0) when none is installed. */
int segsort_set_tuning(int key_width, int value_bytes, int wave_cap, int tile_enable, int grid_mult,
                       int concurrent);
int segsort_reset_tuning(int key_width, int value_bytes);
int segsort_get_tuning(int key_width, int value_bytes, int* wave_cap, int* tile_enable, int* grid_mult,
                       int* concurrent);
/* The tuning a call with these sizes would run with (host only); *from_table
 * is 1 when it comes from the built-in table, 0 when from an override. */
int segsort_lookup_tuning(int num_items, int num_segments, int key_width, int value_bytes, int* wave_cap,
                          int* tile_enable, int* grid_mult, int* concurrent, int* from_table);
/* The same for a call of segsort_sort_hinted with this length bound (0: none). */
int segsort_lookup_tuning_hinted(int num_items, int num_segments, int max_segment_length, int key_width,
                                 int value_bytes, int* wave_cap, int* tile_enable, int* grid_mult, int* concurrent,
                                 int* from_table);
/* Capacities (items) of the LDS tile classes for a key width / value size;
 * writes up to 4 values and returns how many. */
int segsort_tile_caps(int key_width, int value_bytes, int* caps);

/* Graph replay of the batched form (no counterpart in the reference, whose
 * calls are single kernel launches per pass). 0 (the default, or SEGSORT_GRAPH
 * unset/0): every call launches its kernels on the caller's stream. 1 (or
 * SEGSORT_GRAPH=1): the first call with a given argument set is captured into a
 * hipGraph on a private stream and later calls replay it (16 graphs cached).
 * On MI355X with ROCm 7.2 direct launches measured faster than replay. Either
 * way the batched call is capture-safe: a caller may capture it into its own
 * graph (no host sync, no allocation, no memset). Returns the previous mode. */
int segsort_set_graph_replay(int on);

/* Per-launch timing (no counterpart in the reference, whose drivers time the
 * whole call with cudaEvents, src/radixseg.cu:121-139). While enabled, every
 * kernel launch is bracketed by HIP events on its stream and, after each
 * batched/plain call, the call's per-class element counts are read back (a
 * host sync), so it is for measurement runs only. Enabling clears the log. */
int segsort_profile_enable(int on);
/* Writes "name<TAB>launches<TAB>total_ms<TAB>algorithmic_bytes" lines, one per
 * kernel name, into buf (NUL-terminated, truncated to buflen) after waiting
 * for the logged events; returns the full report size including the NUL. */
size_t segsort_profile_report(char* buf, size_t buflen);

/* Human-readable text of the last error this thread's calls returned. */
const char* segsort_last_error(void);
/* Library version (major*10000 + minor*100 + patch). */
int segsort_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SEGSORT_H */
