// Kernels of the three forms (batched, plain per-class, fix-pass) and of the
// differentiable pair sort. All scheduling state lives in device memory, so
// the batched sequence needs no host synchronisation and can be captured.
#pragma once

#include <type_traits>

#include "tile.hpp"

namespace ssort {

constexpr int MAX_TILE_CLASSES = 9;  // wave-group classes, the wave tier, the LDS tiles
// Queue q < MAX_TILE_CLASSES holds class q's ranges; when a call splits them,
// MAX_TILE_CLASSES + c holds the ranges that the MSD tier queues for class c
// (so the binned ranges can be sorted while the MSD tier still runs).
constexpr int MAX_QUEUES = 2 * MAX_TILE_CLASSES;
constexpr int MAX_LEVELS = 8;        // MSD levels of the large tier (64-bit keys)
// Length buckets of the binned ranges of the large LDS tile classes: bucket b
// of class c holds the ranges whose length falls in the b-th of LPT_BUCKETS
// equal parts of the class's length span, and the class kernel takes the
// buckets longest first. A persistent grid over a class of a few hundred or
// thousand ranges is a list schedule, and taking the longest ranges first
// keeps the last ranges short: the kernel no longer ends on one long range
// that started late while most CUs idle.
constexpr int LPT_BUCKETS = 8;

// Device-resident counters of one call; cleared at the start of the call (by
// the binning kernel's first workgroup, which then opens `gate`; k_init in the
// fix-pass form). Everything before `gate` is cleared.
struct Counters {
    uint32_t tile_count[MAX_QUEUES];
    uint32_t pad0[2];
    unsigned long long level[MAX_LEVELS + 1];  // (ranges << 32) | chunks, per MSD level
    unsigned long long tile_items[MAX_QUEUES];  // elements queued per queue
    unsigned long long level_items[MAX_LEVELS + 1];   // elements queued per MSD level
    uint32_t claim;                                   // next range of the depth-first tier
    uint32_t bcount[MAX_TILE_CLASSES][LPT_BUCKETS];   // binned ranges per length bucket (bucketed classes)
    uint32_t bin_done;                                // binning workgroups past the gate
    uint32_t pad1;
    unsigned long long gate;  // the call's token once the counters above are clear; 0 again after binning
};

// MSD levels of the batched form: two multi-workgroup levels (the long
// segments, then their children of more than MSD_BIG_TILES partition tiles),
// then the depth-first tier for every other child larger than a tile.
constexpr int MSD_LEVEL0 = 0, MSD_LEVEL1 = 1, MSD_LOCAL = 2;
constexpr int MSD_BIG_TILES = 8;
constexpr int MSD_MAX_CHUNKS = 256;  // chunks per range of a multi-workgroup level

// Length caps of the classes of one call: class c holds ranges with
// cap[c-1] < n <= cap[c] (cap[-1] = 0); n > cap[nclass-1] is large.
struct ClassCaps {
    int cap[MAX_TILE_CLASSES];
    int nclass;
};

// What the MSD level's first pass records per chunk: the OR and AND of the
// chunk's order keys inside the range's bits, and the highest digit position
// whose bits vary inside the chunk (-1: every key equal); the chunk's digit
// histogram is kept for that position only.
struct ChunkInfo {
    unsigned long long kor;
    unsigned long long kand;
    int lp;
    int pad;
};

// Where the queues of one call live (all inside the caller's temp storage).
struct Sched {
    Counters* ctr;
    TileItem* tq[MAX_QUEUES];
    LargeItem* lq[MAX_LEVELS + 1];
    uint32_t* hist[2];      // per multi-workgroup level: [chunk][RADIX], the histogram of the chunk's highest varying digit
    ChunkInfo* cinfo[2];    // per multi-workgroup level: [chunk]
    int chunk;       // partition tile of the MSD levels (items); chunks are multiples of it
    // Capacities: a push beyond them is dropped and readers clamp their counts,
    // so overlapping (invalid) segments cannot write outside the temp storage.
    uint32_t tq_cap[MAX_QUEUES];
    uint32_t lq_cap;
    uint32_t chunk_cap;
    int child_q[MAX_TILE_CLASSES];  // queue of class c's MSD children (c: the binned ranges' queue)
    // Bucketed classes (nullptr: none): class c's binned ranges in length bucket
    // b are at bq[c][b * bcap[c] + i]; its queue tq[c] then holds the MSD
    // children only.
    TileItem* bq[MAX_TILE_CLASSES];
    uint32_t bcap[MAX_TILE_CLASSES];
};

// Length bucket of a range of n items in a class of the lengths lo..hi (a tile
// class: hi <= 16384, so (n - lo) * LPT_BUCKETS fits 32 bits).
__device__ __forceinline__ int lpt_bucket(int lo, int hi, int n) {
    const uint32_t b = ((uint32_t)(n - lo) * LPT_BUCKETS) / (uint32_t)(hi - lo + 1);
    return b >= LPT_BUCKETS ? LPT_BUCKETS - 1 : (int)b;
}

// The ranges a bucketed class's kernel takes before its queue tq[cls]: the
// bucket counts (clamped to the capacity), wave-uniform, and their total.
struct BucketView {
    uint32_t cnt[LPT_BUCKETS];
    int total;
};

__device__ __forceinline__ BucketView bucket_view(const Sched& sc, int cls) {
    BucketView v;
    v.total = 0;
#pragma unroll
    for (int b = 0; b < LPT_BUCKETS; ++b) v.cnt[b] = 0;
    if (cls >= MAX_TILE_CLASSES || sc.bq[cls] == nullptr) return v;  // queue cls alone (MSD children, plain form)
    const uint32_t cap = sc.bcap[cls];
#pragma unroll
    for (int b = 0; b < LPT_BUCKETS; ++b) {
        const uint32_t c = sc.ctr->bcount[cls][b];
        v.cnt[b] = (uint32_t)__builtin_amdgcn_readfirstlane((int)(c < cap ? c : cap));
        v.total += (int)v.cnt[b];
    }
    return v;
}

// Entry i < v.total of a bucketed class, the longest bucket first.
__device__ __forceinline__ const TileItem* bucket_entry(const Sched& sc, const BucketView& v, int cls, int i) {
    int b = LPT_BUCKETS - 1;
    uint32_t r = (uint32_t)i;
    while (b > 0 && r >= v.cnt[b]) {
        r -= v.cnt[b];
        --b;
    }
    return sc.bq[cls] + (size_t)b * sc.bcap[cls] + r;
}

// Every lane of a wave reads the same count and the same queue entry: taken
// into scalar registers, the range's bounds and its buffer pointers are
// wave-uniform, so global addresses are a scalar base plus a 32-bit lane offset.
__device__ __forceinline__ int tile_count(const Sched& sc, int cls) {
    const uint32_t c = sc.ctr->tile_count[cls];
    return __builtin_amdgcn_readfirstlane((int)(c < sc.tq_cap[cls] ? c : sc.tq_cap[cls]));
}

__device__ __forceinline__ TileItem queue_item(const TileItem* q, int i) {
    TileItem t = q[i];
    t.begin = __builtin_amdgcn_readfirstlane(t.begin);
    t.n_src = __builtin_amdgcn_readfirstlane(t.n_src);
    return t;
}

__device__ __forceinline__ void level_counts(const Sched& sc, int level, int& count, int& chunks) {
    const unsigned long long packed = sc.ctr->level[level];
    const uint32_t c = (uint32_t)(packed >> 32), g = (uint32_t)(packed & 0xFFFFFFFFull);
    count = (int)(c < sc.lq_cap ? c : sc.lq_cap);
    chunks = (int)(g < sc.chunk_cap ? g : sc.chunk_cap);
}

__device__ __forceinline__ int class_of(const ClassCaps& cc, int n) {
    for (int c = 0; c < cc.nclass; ++c)
        if (n <= cc.cap[c]) return c;
    return -1;
}

__global__ void k_init(Counters* ctr) {
    uint32_t* p = (uint32_t*)ctr;
    for (int i = threadIdx.x; i < (int)(sizeof(Counters) / 4); i += blockDim.x) p[i] = 0;
}

// The binning kernel clears the counters itself (no launch before it): its
// workgroup 0 clears them and then stores the call's token (a 64-bit value the
// host never repeats) into `gate`; every other workgroup does its loads and
// LDS work first and waits for the token only before its first global atomic.
// The last workgroup past the gate sets it back to 0, so a graph replaying the
// call (same token) waits again. Workgroup 0 is dispatched first and never
// waits, so the wait ends; it is bounded all the same (about 0.3 s).
constexpr int GATE_SPIN_MAX = 1 << 21;

// The counters are only ever read and updated by device-scope atomics after the
// clear (and by later kernels), so the clear is made of device-scope atomic
// stores, acknowledged by every wave before the barrier and so before the
// gate's store: no cache write-back or invalidation on either side.
__device__ __forceinline__ void clear_counters_and_open(Counters* ctr, unsigned long long token) {
    uint32_t* p = (uint32_t*)ctr;
    for (int i = threadIdx.x; i < (int)(offsetof(Counters, gate) / 4); i += blockDim.x)
        __hip_atomic_store(p + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // every wave's clearing stores acknowledged (gfx9: vmcnt counts stores)
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&ctr->gate, token, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wait_gate(Counters* ctr, unsigned long long token) {
    for (int i = 0; i < GATE_SPIN_MAX; ++i) {
        if (__hip_atomic_load(&ctr->gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == token) return;
        __builtin_amdgcn_s_sleep(2);
    }
}

// Queue slots taken by a whole workgroup at once: per-wave LDS counts, then
// one global atomic per class for the slots (and one for the class's element
// total), so that the few global queue counters see few atomics.
struct BlockPush {
    uint32_t cnt[MAX_TILE_CLASSES];
    uint32_t base[MAX_TILE_CLASSES];
    unsigned long long items[MAX_TILE_CLASSES];
};

// Every thread of the workgroup calls this (it holds barriers); threads with
// cls < 0 queue nothing.
__device__ __forceinline__ void push_tiles_block(const Sched& sc, BlockPush& bp, int cls, TileItem it) {
    if (threadIdx.x < MAX_TILE_CLASSES) {
        bp.cnt[threadIdx.x] = 0;
        bp.items[threadIdx.x] = 0;
    }
    __syncthreads();
    uint32_t slot = 0;
#pragma unroll
    for (int c = 0; c < MAX_TILE_CLASSES; ++c) {
        const uint64_t m = __ballot(cls == c);
        if (m == 0) continue;
        const uint32_t items = wave_sum(cls == c ? (uint32_t)ti_n(it) : 0u);
        const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
        uint32_t base = 0;
        if (lane_id() == leader) {
            base = atomicAdd(&bp.cnt[c], (uint32_t)__popcll(m));
            atomicAdd(&bp.items[c], (unsigned long long)items);
        }
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader);
        if (cls == c) slot = base + lanes_below(m);
    }
    __syncthreads();
    if (threadIdx.x < MAX_TILE_CLASSES) {
        const int c = threadIdx.x;
        const uint32_t k = bp.cnt[c];
        bp.base[c] = k ? atomicAdd(&sc.ctr->tile_count[sc.child_q[c]], k) : 0u;
        if (k) atomicAdd(&sc.ctr->tile_items[sc.child_q[c]], bp.items[c]);
    }
    __syncthreads();
    if (cls >= 0) {
        const int q = sc.child_q[cls];
        const uint32_t g = bp.base[cls] + slot;
        if (g < sc.tq_cap[q]) sc.tq[q][g] = it;
    }
}

// Queue a range for MSD level `level`. For the multi-workgroup levels
// (chunked) the range is cut into at most MSD_MAX_CHUNKS chunks of a whole
// number of partition tiles, which take consecutive ids of the level's chunk
// arrays.
__device__ __forceinline__ void push_large(const Sched& sc, int level, int begin, int n, int src, int hi_bit,
                                           bool chunked) {
    int chunk_len = 0, nch = 0;
    if (chunked) {
        const int tiles = (n + sc.chunk - 1) / sc.chunk;
        chunk_len = ((tiles + MSD_MAX_CHUNKS - 1) / MSD_MAX_CHUNKS) * sc.chunk;
        nch = (n + chunk_len - 1) / chunk_len;
    }
    const unsigned long long old = atomicAdd(&sc.ctr->level[level], (1ull << 32) | (unsigned long long)nch);
    atomicAdd(&sc.ctr->level_items[level], (unsigned long long)n);
    LargeItem li;
    li.begin = begin;
    li.n = n;
    li.src = src;
    li.hi_bit = hi_bit;
    li.chunk_base = (int)(old & 0xFFFFFFFFull);
    li.chunk_len = chunk_len;
    li.nch = nch;
    li.pad = 0;
    if ((old >> 32) < sc.lq_cap) {
        // past the chunk capacity (only overlapping segments can get there) the
        // record stays, empty, so that no later range maps its chunks to a stale one
        if (chunked && (old & 0xFFFFFFFFull) + nch > sc.chunk_cap) li.n = li.nch = 0;
        sc.lq[level][old >> 32] = li;
    }
}

// A segment longer than the largest tile in a call whose caller promised that
// no segment is (a wrong max_segment_length): queued for the one-workgroup-per-
// segment LSD, which such calls launch in place of the MSD tier.
__device__ __forceinline__ void push_oversize(const Sched& sc, int begin, int n) {
    const uint32_t slot = (uint32_t)atomicAdd(&sc.ctr->level[MAX_LEVELS], 1ull);
    atomicAdd(&sc.ctr->level_items[MAX_LEVELS], (unsigned long long)n);
    if (slot < sc.lq_cap) {
        LargeItem li;
        li.begin = begin;
        li.n = n;
        li.src = BUF_IN;
        li.hi_bit = 0;
        li.chunk_base = li.chunk_len = li.nch = li.pad = 0;
        sc.lq[MAX_LEVELS][slot] = li;
    }
}

// Bin every segment by length: length 1 is copied here, the rest go to the
// queue of their class (batched and plain forms) or to MSD level 0 (large).
// Each block bins spt * BIN_NT segments (spt <= BIN_SPT, chosen by the host so
// that the grid keeps about a hundred blocks: the global queue counters see one
// atomic per block and class); slots are taken from LDS counters
// (one LDS atomic per wave and class) and then from the global counters with
// one atomic per block and class, so the hot global counters see few atomics.
// LPT: the classes with length buckets (sc.bq[c] set) file their ranges by
// bucket instead, with an LDS atomic per range. At most four segments per
// thread: the per-segment steps are unrolled, and the kernel's code size counts
// (it starts every call with a cold instruction cache; measured against eight:
// ragged 2^24 0.1916 -> 0.1901 ms, 2^20 segments of 1..32 items 0.1547 ->
// 0.1520 ms; two per thread doubles the blocks at 2^20 segments and is slower).
constexpr int BIN_NT = 1024;
constexpr int BIN_SPT = 4;

template <typename U, typename V, bool HAS_V, bool LPT>
__global__ __launch_bounds__(BIN_NT) void k_bin(const int* __restrict__ d_begin, const int* __restrict__ d_end,
                                                int num_segments, Bufs b, Sched sc, ClassCaps cc, int end_bit,
                                                int large_to_level0, int spt, unsigned long long token) {
    constexpr int NB = LPT ? MAX_TILE_CLASSES * LPT_BUCKETS : 1;
    constexpr int NC = LPT ? MAX_TILE_CLASSES : 1;
    __shared__ uint32_t lcount[MAX_TILE_CLASSES];
    __shared__ uint32_t lbase[MAX_TILE_CLASSES];
    __shared__ unsigned long long litems[MAX_TILE_CLASSES];
    // per-class tables that the per-range steps index by the range's class (an
    // LDS read instead of a select over the kernel arguments)
    __shared__ TileItem* ltq[MAX_TILE_CLASSES];
    __shared__ uint32_t ltq_cap[MAX_TILE_CLASSES];
    __shared__ uint32_t lwbase[BIN_NT / WAVE][MAX_TILE_CLASSES];  // each wave's slot base per class
    __shared__ uint32_t lbcount[NB];  // LPT: per class and length bucket
    __shared__ uint32_t lbbase[NB];
    __shared__ TileItem* lbq[NC];     // LPT: bucket regions (nullptr: class not bucketed)
    __shared__ uint32_t lbcap[NC];
    __shared__ int lcap[NC];
    if (threadIdx.x < MAX_TILE_CLASSES) {
        const int c = threadIdx.x;
        lcount[c] = 0;
        litems[c] = 0;
        ltq[c] = sc.tq[c];
        ltq_cap[c] = sc.tq_cap[c];
        if constexpr (LPT) {
            lbq[c] = sc.bq[c];
            lbcap[c] = sc.bcap[c];
            lcap[c] = cc.cap[c];
        }
    }
    if (LPT && threadIdx.x < NB) lbcount[threadIdx.x] = 0;
    int beg[BIN_SPT], n[BIN_SPT], cls[BIN_SPT];
    uint32_t slot[BIN_SPT];
    int bkt[BIN_SPT];  // LPT: length bucket of a bucketed class's range, else -1
    const int s0 = blockIdx.x * (BIN_NT * spt) + threadIdx.x;
    // every offset load issued before any is used (in workgroup 0 also before the
    // counters are cleared, so that the clearing's wait covers them too)
#pragma unroll
    for (int j = 0; j < BIN_SPT; ++j) {
        const int s = s0 + j * BIN_NT;
        beg[j] = 0;
        n[j] = 0;
        if (j < spt && s < num_segments) {
            beg[j] = d_begin[s];
            n[j] = d_end[s];
        }
    }
    if (blockIdx.x == 0) clear_counters_and_open(sc.ctr, token);  // includes a barrier
    else __syncthreads();
    // length-1 segments are copied: their loads are issued here and their stores
    // come last, so that no wave waits for the loads in between
    U one_k[BIN_SPT] = {};
    V one_v[BIN_SPT] = {};
#pragma unroll
    for (int j = 0; j < BIN_SPT; ++j) {
        n[j] = n[j] > beg[j] ? n[j] - beg[j] : 0;
        if (n[j] == 1) {
            one_k[j] = ((const U*)b.k[BUF_IN])[beg[j]];
            if constexpr (HAS_V) one_v[j] = ((const V*)b.v[BUF_IN])[beg[j]];
        }
        cls[j] = n[j] >= 2 ? class_of(cc, n[j]) : -2;
        bkt[j] = -1;
    }
    if constexpr (LPT) {
        // ranges of a bucketed class take a slot of their length bucket: an LDS
        // atomic per range (the long ranges of a mixed call spread over the buckets)
#pragma unroll
        for (int j = 0; j < BIN_SPT; ++j) {
            if (cls[j] >= 0 && lbq[cls[j]] != nullptr) {
                const int c = cls[j];
                bkt[j] = lpt_bucket(c > 0 ? lcap[c - 1] + 1 : 2, lcap[c], n[j]);
                slot[j] = atomicAdd(&lbcount[c * LPT_BUCKETS + bkt[j]], 1u);
            }
        }
    }
    // per class: the wave's slots counted with ballots in registers, then one LDS
    // atomic per class and wave for the wave's base inside the block (a loop over
    // the call's classes, not unrolled: this kernel starts every call with a cold
    // instruction cache)
    const int w = (int)(threadIdx.x / WAVE);
#pragma unroll 1
    for (int c = 0; c < cc.nclass; ++c) {
        uint32_t run = 0, items = 0;
#pragma unroll
        for (int j = 0; j < BIN_SPT; ++j) {
            const bool mine = cls[j] == c;
            const bool queued = mine && bkt[j] < 0;  // bucketed ranges have their slot
            const uint64_t m = __ballot(queued);
            if (mine) items += (uint32_t)n[j];
            if (queued) slot[j] = run + lanes_below(m);
            run += (uint32_t)__popcll(m);
        }
        if (__ballot(items != 0) == 0) continue;  // no range of the class in this wave
        items = wave_sum(items);
        if (lane_id() == 0) {
            lwbase[w][c] = run ? atomicAdd(&lcount[c], run) : 0u;
            atomicAdd(&litems[c], (unsigned long long)items);
        }
    }
    // (the wave's own LDS writes above complete before its reads here)
#pragma unroll
    for (int j = 0; j < BIN_SPT; ++j)
        if (cls[j] >= 0 && bkt[j] < 0) slot[j] += lwbase[w][cls[j]];
    if (threadIdx.x == 0 && blockIdx.x != 0) wait_gate(sc.ctr, token);
    __syncthreads();
    // the block's global slots: every atomic issued before any result is used
    // (the long segments' queue atomics below included)
    uint32_t g0 = 0, g1 = 0;
    if (threadIdx.x < (unsigned)cc.nclass && lcount[threadIdx.x])
        g0 = atomicAdd(&sc.ctr->tile_count[threadIdx.x], lcount[threadIdx.x]);
    if (LPT && threadIdx.x < (unsigned)(cc.nclass * LPT_BUCKETS) && lbcount[threadIdx.x])
        g1 = atomicAdd(&sc.ctr->bcount[threadIdx.x / LPT_BUCKETS][threadIdx.x % LPT_BUCKETS], lbcount[threadIdx.x]);
    if (threadIdx.x < (unsigned)cc.nclass && litems[threadIdx.x])
        atomicAdd(&sc.ctr->tile_items[threadIdx.x], litems[threadIdx.x]);
    if (threadIdx.x == BIN_NT - 1 && atomicAdd(&sc.ctr->bin_done, 1u) == gridDim.x - 1)
        __hip_atomic_store(&sc.ctr->gate, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int j = 0; j < BIN_SPT; ++j) {
        if (cls[j] == -1 && large_to_level0 == 1) push_large(sc, MSD_LEVEL0, beg[j], n[j], BUF_IN, end_bit, true);
        if (cls[j] == -1 && large_to_level0 == 2) push_oversize(sc, beg[j], n[j]);
    }
    if (threadIdx.x < (unsigned)cc.nclass) lbase[threadIdx.x] = g0;
    if (LPT && threadIdx.x < (unsigned)(cc.nclass * LPT_BUCKETS)) lbbase[threadIdx.x] = g1;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < BIN_SPT; ++j) {
        if (cls[j] >= 0 && bkt[j] < 0) {
            const uint32_t g = lbase[cls[j]] + slot[j];
            if (g < ltq_cap[cls[j]]) ltq[cls[j]][g] = make_ti(beg[j], n[j], BUF_IN);
        } else if (LPT && cls[j] >= 0) {
            const uint32_t g = lbbase[cls[j] * LPT_BUCKETS + bkt[j]] + slot[j];
            const uint32_t cap = lbcap[cls[j]];
            if (g < cap) lbq[cls[j]][(size_t)bkt[j] * cap + g] = make_ti(beg[j], n[j], BUF_IN);
        }
    }
#pragma unroll
    for (int j = 0; j < BIN_SPT; ++j) {
        if (n[j] == 1) {
            buf_ptr<U>(b.k[BUF_OUT])[beg[j]] = one_k[j];
            if constexpr (HAS_V) buf_ptr<V>(b.v[BUF_OUT])[beg[j]] = one_v[j];
        }
    }
}

// Plain form: class ids listed per segment for the one-workgroup-per-segment
// global LSD (ranges larger than the biggest LDS tile).
__global__ void k_bin_large_ids(const int* __restrict__ d_begin, const int* __restrict__ d_end, int num_segments,
                                Sched sc, ClassCaps cc) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int n = 0, beg = 0;
    if (s < num_segments) {
        beg = d_begin[s];
        n = d_end[s] - beg;
    }
    const bool big = s < num_segments && n > cc.cap[cc.nclass - 1];
    const uint64_t m = __ballot(big);
    if (m == 0) return;
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = (uint32_t)atomicAdd(&sc.ctr->level[MAX_LEVELS], (unsigned long long)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)leader, WAVE);
    if (big) atomicAdd(&sc.ctr->level_items[MAX_LEVELS], (unsigned long long)n);
    if (big) {
        LargeItem li;
        li.begin = beg;
        li.n = n;
        li.src = BUF_IN;
        li.hi_bit = 0;
        li.chunk_base = li.chunk_len = li.nch = li.pad = 0;
        const uint32_t slot = base + lanes_below(m);
        if (slot < sc.lq_cap) sc.lq[MAX_LEVELS][slot] = li;
    }
}

// ---- tile tiers --------------------------------------------------------------

// Wave tier: each wave sorts one queued range of n <= MAXN items (compare-count
// up to rank_max items, a register bitonic network above). For MAXN = 512 the
// per-wave LDS area allows 7 waves per SIMD with 32-bit keys (6 with 64-bit
// keys, whose bucket path also keeps a position array) and the register budget
// is held to the same (72 VGPRs spill about 23 per lane with 32-bit keys, and
// still measured faster than 6 waves: one 2^24 segment 0.373 -> 0.366 ms);
// ranges of at most 256 items need half the area and fewer registers, and aim
// at 8.
template <typename U, typename V, bool HAS_V, int MAXN>
__global__ __launch_bounds__(256, MAXN <= 256 ? 8 : sizeof(U) == 4 ? 7 : 6) void k_wave_tier(Bufs b, Sched sc, int cls, KeyMode km, int persistent,
                                                   int rank_max) {
    const int waves_per_block = blockDim.x / WAVE;
    const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + (threadIdx.x >> 6));
    const int stride = persistent ? gridDim.x * waves_per_block : 0x7FFFFFFF;
    const int count = tile_count(sc, cls);
    __shared__ WaveSmem<U, V, MAXN> wsm[4];  // one region per wave (blocks of 4 waves)
    WaveSmem<U, V, MAXN>& my = wsm[threadIdx.x >> 6];
    for (int i = wid; i < count; i += stride) {
        const TileItem it = queue_item(sc.tq[cls], i);
        const int src = ti_src(it);
        wave_sort<U, V, HAS_V, MAXN>((const U*)b.k[src], (const V*)b.v[src], buf_ptr<U>(b.k[BUF_OUT]),
                               buf_ptr<V>(b.v[BUF_OUT]), it.begin, ti_n(it), km, rank_max, my);
        if (!persistent) break;
    }
}

// Wave-group tier: each wave sorts 64/G queued ranges of n <= G items at once,
// one range per group of G lanes (a bitonic network confined to the group).
// The loop of one wave over its share of queue cls; no barriers.
template <int G, typename U, typename V, bool HAS_V>
__device__ __forceinline__ void group_loop(const Bufs& b, const Sched& sc, int cls, const KeyMode& km) {
    constexpr int PER_WAVE = WAVE / G;
    const int waves_per_block = blockDim.x / WAVE;
    const int wid = blockIdx.x * waves_per_block + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * waves_per_block;
    const int count = tile_count(sc, cls);
    const int g = (int)(lane_id() / G);
    const int step = nwaves * PER_WAVE;
    int base = wid * PER_WAVE;
    if (base >= count) return;  // whole wave
    // pipeline: queue entries two batches ahead, keys and values one batch ahead
    auto entry = [&](int at) {
        TileItem it = make_ti(0, 0, BUF_IN);
        if (at + g < count) it = sc.tq[cls][at + g];
        return it;
    };
    GroupItems<U, V> cur, nxt;
    group_load<G, U, V, HAS_V>(b, entry(base), base + g < count, cur);
    TileItem it_next = entry(base + step);
    for (;;) {
        const int next = base + step;
        if (next < count) {
            group_load<G, U, V, HAS_V>(b, it_next, next + g < count, nxt);
            it_next = entry(next + step);
        }
        group_sort_loaded<G, U, V, HAS_V>(b, cur, km);
        if (next >= count) break;
        cur = nxt;
        base = next;
    }
}

template <int G, typename U, typename V, bool HAS_V>
__global__ __launch_bounds__(256) void k_group_tier(Bufs b, Sched sc, int cls, KeyMode km) {
    group_loop<G, U, V, HAS_V>(b, sc, cls, km);
}

// The tile tier's loop for tiles of at most 4 items per thread: software
// pipeline, the next range's keys are loaded while this one is written out.
template <int NT, int IPT, typename U, typename V, bool HAS_V>
__device__ __forceinline__ void tile_loop_pipelined(typename RadixTile<NT, IPT, U, V, HAS_V>::Smem& sm, const Bufs& b,
                                                    const Sched& sc, int cls, const KeyMode& km, int persistent) {
    using T = RadixTile<NT, IPT, U, V, HAS_V>;
    const int count = tile_count(sc, cls);
    int i = blockIdx.x;
    if (i >= count) return;
    TileItem it = queue_item(sc.tq[cls], i);
    U raw[IPT];
    T::load_raw((const U*)b.k[ti_src(it)], it.begin, ti_n(it), raw);
    for (;;) {
        const TileItem cur = it;
        U keys[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) keys[j] = raw[j];
        const int next = persistent ? i + (int)gridDim.x : count;
        const int src = ti_src(cur);
        T::sort_loaded(sm, keys, (const U*)b.k[src], (const V*)b.v[src], buf_ptr<U>(b.k[BUF_OUT]),
                       buf_ptr<V>(b.v[BUF_OUT]), cur.begin, ti_n(cur), km, [&] {
                           if (next < count) {
                               it = queue_item(sc.tq[cls], next);
                               T::load_raw((const U*)b.k[ti_src(it)], it.begin, ti_n(it), raw);
                           }
                       });
        if (next >= count) break;
        i = next;
    }
}

// LDS tier: each workgroup sorts one queued range of n <= NT*IPT items.
// At least 4 waves per SIMD (16 per CU) so that LDS and memory latency can be
// hidden; this caps the tile kernels at 128 VGPRs.
// LPT (16-item tiles): the class's length buckets are taken longest first,
// then its queue. A separate build: the bucket walk's state costs the kernel
// scalar registers (spills) that equal-length segments gain nothing for.
// HALF (tiles of at most 4096 items): ranges of at most half the tile use the
// half-depth build below. Also a separate build: the two builds in one kernel
// make it spill (256x16: 12 bytes per lane) and slow full-size ranges.
template <int NT, int IPT, typename U, typename V, bool HAS_V, bool LPT, bool HALF>
__global__ __launch_bounds__(NT, 4) void k_tile_tier(Bufs b, Sched sc, int cls, KeyMode km, int persistent) {
    using T = RadixTile<NT, IPT, U, V, HAS_V>;
    if constexpr (IPT > 4) {
        // One range at a time (no register room for a second set of keys;
        // measured: pipelining the 16-item tiles spills and is slower). In
        // tiles of up to 4096 items a range of at most half the tile is sorted
        // by the build with IPT / 2 items per thread, so that every wave takes
        // part and each wave's serial ranking steps halve (in larger tiles the
        // second build made the code spill and slowed full-size ranges).
        using T2 = RadixTile<NT, IPT / 2, U, V, HAS_V>;
        __shared__ union {
            typename T::Smem full;
            typename T2::Smem half;
        } sm;
        const int count = tile_count(sc, cls);
        BucketView bv;
        bv.total = 0;
        if constexpr (LPT) bv = bucket_view(sc, cls);  // the binned ranges, longest first, then the queue
        const int total = bv.total + count;
        for (int i = blockIdx.x; i < total; i += persistent ? (int)gridDim.x : total) {
            TileItem cur;
            if constexpr (LPT) {
                cur = i < bv.total ? queue_item(bucket_entry(sc, bv, cls, i), 0) : queue_item(sc.tq[cls], i - bv.total);
            } else {
                cur = queue_item(sc.tq[cls], i);
            }
            const int src = ti_src(cur);
            const int n = ti_n(cur);
            const U* ks = (const U*)b.k[src];
            const V* vs = (const V*)b.v[src];
            U* ko = buf_ptr<U>(b.k[BUF_OUT]);
            V* vo = buf_ptr<V>(b.v[BUF_OUT]);
            if (HALF && NT * IPT <= 4096 && n <= T2::TILE) T2::sort(sm.half, ks, vs, ko, vo, cur.begin, n, km);
            else T::sort(sm.full, ks, vs, ko, vo, cur.begin, n, km);
        }
        return;
    }
    __shared__ typename T::Smem sm;
    tile_loop_pipelined<NT, IPT, U, V, HAS_V>(sm, b, sc, cls, km, persistent);
}

// Wave tier and the 256 x IPT (IPT <= 4) tile tier in one persistent launch
// (and, when enabled, the wave-group classes before them): each wave first
// sorts its share of the group queues and of the wave-tier queue qw on its own
// (no barriers), then the workgroup sorts ranges of the tile queue qt together.
// One launch instead of two, and a workgroup's tile ranges start while other
// workgroups still sort their last wave-tier ranges. Both tiers run at 8 waves
// per SIMD with about 15 KB of LDS per workgroup, so neither loses occupancy.
// (For 64-bit keys the 256 x 4 tile then spills a few dozen bytes per lane;
// measured, that beats 5 or 6 waves per SIMD without spills.)
// Queues of the wave-group classes (G = 4, 8, 16, 32) sorted first by the same
// launch, or -1.
struct GroupQueues {
    int q[4];
};

template <typename U, typename V, bool HAS_V, int MAXN, int IPT>
__global__ __launch_bounds__(256, 8) void k_wave_tile_tier(Bufs b, Sched sc, int qw, int qt, KeyMode km,
                                                           int rank_max, GroupQueues gq) {
    using T = RadixTile<256, IPT, U, V, HAS_V>;
    static_assert(IPT <= 4, "the pipelined tile loop");
    __shared__ union {
        WaveSmem<U, V, MAXN> wsm[4];
        typename T::Smem tsm;
    } sm;
    // wave-group classes first (per wave, no barriers, no LDS); 32-bit keys only
    // (the host keeps their own launches for 64-bit keys)
    if constexpr (sizeof(U) == 4) {
        if (gq.q[0] >= 0) group_loop<4, U, V, HAS_V>(b, sc, gq.q[0], km);
        if (gq.q[1] >= 0) group_loop<8, U, V, HAS_V>(b, sc, gq.q[1], km);
        if (gq.q[2] >= 0) group_loop<16, U, V, HAS_V>(b, sc, gq.q[2], km);
        if (gq.q[3] >= 0) group_loop<32, U, V, HAS_V>(b, sc, gq.q[3], km);
    }
    {
        const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
        const int count = tile_count(sc, qw);
        WaveSmem<U, V, MAXN>& my = sm.wsm[threadIdx.x >> 6];
        for (int i = wid; i < count; i += (int)gridDim.x * 4) {
            const TileItem it = queue_item(sc.tq[qw], i);
            const int src = ti_src(it);
            wave_sort<U, V, HAS_V, MAXN>((const U*)b.k[src], (const V*)b.v[src], buf_ptr<U>(b.k[BUF_OUT]),
                                         buf_ptr<V>(b.v[BUF_OUT]), it.begin, ti_n(it), km, rank_max, my);
        }
    }
    __syncthreads();  // the wave regions become the tile's LDS
    tile_loop_pipelined<256, IPT, U, V, HAS_V>(sm.tsm, b, sc, qt, km, 1);
}

// ---- large tier (MSD partition over many workgroups) --------------------------

// The last record whose first chunk is at or before g (chunk_base grows with the
// record index). Wave-parallel: each round the 64 lanes test 64 evenly spaced
// records of the remaining span at once, so a level of up to 4096 ranges takes
// two dependent loads (a binary search took log2(count), each a round trip).
__device__ __forceinline__ int find_item(const LargeItem* items, int count, int g) {
    int lo = 0, hi = count - 1;  // items[lo].chunk_base <= g throughout
    while (lo < hi) {
        const int step = (hi - lo + WAVE) / WAVE;  // ceil((hi - lo + 1) / WAVE)
        const int idx = lo + (int)lane_id() * step;
        const uint64_t ok = __ballot(idx <= hi && items[idx].chunk_base <= g);
        lo += (63 - __builtin_clzll(ok)) * step;  // lane 0 (idx = lo) always passes
        hi = min(lo + step - 1, hi);
    }
    return lo;
}

// The range record owning chunk g of a multi-workgroup level, in scalar registers.
__device__ __forceinline__ LargeItem chunk_owner(const LargeItem* items, int count, int g) {
    LargeItem it = items[find_item(items, count, g)];
    it.begin = __builtin_amdgcn_readfirstlane(it.begin);
    it.n = __builtin_amdgcn_readfirstlane(it.n);
    it.src = __builtin_amdgcn_readfirstlane(it.src);
    it.hi_bit = __builtin_amdgcn_readfirstlane(it.hi_bit);
    it.chunk_base = __builtin_amdgcn_readfirstlane(it.chunk_base);
    it.chunk_len = __builtin_amdgcn_readfirstlane(it.chunk_len);
    it.nch = __builtin_amdgcn_readfirstlane(it.nch);
    return it;
}

__device__ __forceinline__ int chunk_items(const LargeItem& it, int c) {
    const int left = it.n - c * it.chunk_len;
    return left < it.chunk_len ? left : it.chunk_len;
}

__device__ __forceinline__ int npos_of(int hi_bit, int begin_bit) {
    return (hi_bit - begin_bit + RADIX_BITS - 1) / RADIX_BITS;
}

// Pass 1 of a level: per chunk, the OR and AND of the order keys (inside the
// range's bits [begin_bit, hi_bit)) and the digit histogram of the chunk's
// highest varying digit position. The level's digit is the range's highest
// varying position, which no chunk exceeds; a chunk whose own top is lower
// holds a single digit value there, known from its AND, so one histogram per
// chunk is all the partition needs. The first read of a chunk counts the
// digits of the range's top position as it goes, which is the chunk's own
// top in the common case (keys that vary in their high bits), so the chunk is
// read once; a chunk whose top digit turns out constant is counted again at
// its own top (re-read when it spans more than one tile; a one-tile chunk's
// keys stay in registers).
template <int NT, int IPT, typename U, typename V, bool HAS_V>
__global__ __launch_bounds__(NT) void k_large_hist(Bufs b, Sched sc, int level, KeyMode km) {
    constexpr int NW = NT / WAVE, TILE = NT * IPT;
    __shared__ uint32_t hist[RADIX];
    __shared__ U red_or[NW];
    __shared__ U red_and[NW];
    const int t = threadIdx.x;
    int count, chunks;
    level_counts(sc, level, count, chunks);
    const LargeItem* items = sc.lq[level];
    // a contiguous span of chunks per workgroup: consecutive chunks of one range
    // share its record, which is looked up once
    const int per = (chunks + (int)gridDim.x - 1) / (int)gridDim.x;
    const int g_stop = min(((int)blockIdx.x + 1) * per, chunks);
    LargeItem it;
    it.chunk_base = it.nch = 0;
    for (int g = (int)blockIdx.x * per; g < g_stop; ++g) {
        if (g >= it.chunk_base + it.nch) it = chunk_owner(items, count, g);
        const int c = g - it.chunk_base;
        const int len = chunk_items(it, c);
        const U rmask = range_mask<U>(km.begin_bit, it.hi_bit);
        const U* ks = (const U*)b.k[it.src] + ((int64_t)it.begin + (int64_t)c * it.chunk_len);
        U key[IPT], raw[IPT];
        auto fetch = [&](int t0) {  // raw keys of the tile at t0 (slots past the end read 0)
            const int tn = (len - t0) < TILE ? (len - t0) : TILE;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = j * NT + t;
                raw[j] = i < tn ? at32(ks + t0, i) : U(0);
            }
        };
        auto convert = [&]() {
#pragma unroll
            for (int j = 0; j < IPT; ++j) key[j] = U(twiddle_in<U>(raw[j], km.kind, km.desc) & rmask);
        };
        // digit counts of the tile in registers at position shift sh (digit mask dm)
        auto count_digits = [&](int tn, int sh, U dm) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if (j * NT + (t & ~63) >= tn) break;  // this and later rows: whole wave past the end
                const bool v = j * NT + t < tn;
                const uint32_t d = (uint32_t)((key[j] >> sh) & dm);
                const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
                const uint64_t act = __ballot(v);
                if (__ballot(v && d != d0) == 0) {
                    if ((t & 63) == 0) atomicAdd(&hist[d0], (uint32_t)__popcll(act));
                } else if (v) {
                    atomicAdd(&hist[d], 1u);
                }
            }
        };
        auto digit_mask = [&](int sh) {
            return U((U(1) << ((it.hi_bit - sh) < RADIX_BITS ? (it.hi_bit - sh) : RADIX_BITS)) - 1);
        };
        const int top = npos_of(it.hi_bit, km.begin_bit) - 1;
        const int tsh = km.begin_bit + top * RADIX_BITS;
        const U tdm = digit_mask(tsh);
        for (int i = t; i < RADIX; i += NT) hist[i] = 0;
        __syncthreads();
        U o = 0, a = Bits<U>::ALL;
        int tn = 0;
        // the next tile's loads are issued before this one is counted, where the
        // registers allow (16 keys of 64 bits per thread do not)
        constexpr bool PREFETCH = sizeof(U) * IPT <= 64;
        if (PREFETCH) fetch(0);
        for (int t0 = 0; t0 < len; t0 += TILE) {
            tn = (len - t0) < TILE ? (len - t0) : TILE;
            if (!PREFETCH) fetch(t0);
            convert();
            if (PREFETCH && t0 + TILE < len) fetch(t0 + TILE);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if (j * NT + t < tn) {
                    o |= key[j];
                    a &= key[j];
                }
            }
            count_digits(tn, tsh, tdm);
        }
        o = wave_or(o);
        a = wave_and(a);
        if ((t & 63) == 0) {
            red_or[t >> 6] = o;
            red_and[t >> 6] = a;
        }
        __syncthreads();
        o = 0;
        a = Bits<U>::ALL;
        for (int w = 0; w < NW; ++w) {
            o |= red_or[w];
            a &= red_and[w];
        }
        const U var = U((o ^ a) & rmask);
        const int lp = var ? (bit_length<U>(var) - 1 - km.begin_bit) / RADIX_BITS : -1;
        if (lp >= 0 && lp != top) {
            // the top digit is constant in this chunk: count its own top instead
            for (int i = t; i < RADIX; i += NT) hist[i] = 0;
            __syncthreads();
            const int sh = km.begin_bit + lp * RADIX_BITS;
            const U dm = digit_mask(sh);
            for (int t0 = 0; t0 < len; t0 += TILE) {
                if (len > TILE) {  // a one-tile chunk still holds its keys
                    tn = (len - t0) < TILE ? (len - t0) : TILE;
                    fetch(t0);
                    convert();
                }
                count_digits(tn, sh, dm);
            }
        }
        __syncthreads();
        if (t == 0) sc.cinfo[level][g] = ChunkInfo{(unsigned long long)o, (unsigned long long)a, lp, 0};
        if (lp >= 0)
            for (int i = t; i < RADIX; i += NT) sc.hist[level][(size_t)g * RADIX + i] = hist[i];
        __syncthreads();
    }
}

// Pass 2 of a level: the stable partition of every chunk by the range's
// highest varying digit. A workgroup takes a contiguous span of chunks and
// derives the destinations of the span's part in each range itself from the
// range's chunk records (no separate scan launch): the range's OR/AND give the
// digit, and NT/256 groups of 256 threads (one per digit) total the digit over
// a share of the range's chunks each, separately for the chunks before the
// part; the digit bases are the exclusive scan of the totals. The part is then
// partitioned tile by tile into the child buffer (or straight into the output
// when no lower digit varies), with running destinations in LDS. The workgroup
// holding the range's first chunk queues the children: to the tile classes, to
// the next multi-workgroup level (children of level 0 with more than `big`
// items), or to the depth-first tier. A range whose keys are all equal is
// copied to the output.
template <int NT, int IPT, typename U, typename V, bool HAS_V>
__global__ __launch_bounds__(NT) void k_large_scatter(Bufs b, Sched sc, int level, KeyMode km, ClassCaps cc,
                                                      int tile_max, int big) {
    using T = RadixTile<NT, IPT, U, V, HAS_V>;
    constexpr int NW = NT / WAVE;
    constexpr int G = NT / RADIX;  // digit groups of the offset scan
    static_assert(NT % RADIX == 0 && 2 * G * RADIX <= NW * RADIX, "partial totals fit the wave-count words");
    __shared__ typename T::Smem sm;
    __shared__ uint32_t run[RADIX];     // running destination of each digit
    __shared__ uint32_t cstart[RADIX + 1];  // range-relative start of each digit's child
    __shared__ uint32_t wsum[RADIX / WAVE];
    __shared__ unsigned long long red_or[NW];
    __shared__ unsigned long long red_and[NW];
    __shared__ int csel[MSD_MAX_CHUNKS];  // per chunk of the range: -1 or its one digit (below)
    __shared__ BlockPush bp;
    static_assert(MSD_MAX_CHUNKS <= NT, "one chunk record per thread");
    const int t = threadIdx.x;
    const int d = t & (RADIX - 1);
    const int grp = t / RADIX;
    int count, chunks;
    level_counts(sc, level, count, chunks);
    const LargeItem* items = sc.lq[level];
    const int child_buf = 1 + (level & 1);  // BUF_TMP_A below level 0, BUF_TMP_B below level 1
    // Each workgroup partitions a contiguous span of the level's chunks. The part
    // of the span inside one range is partitioned as one region (the running
    // digit destinations carry from tile to tile), so the range lookup, the chunk
    // records and the digit totals are read once per range part, not per chunk.
    const int per = (chunks + (int)gridDim.x - 1) / (int)gridDim.x;
    const int g_stop = min(((int)blockIdx.x + 1) * per, chunks);
    int g_next = 0;
    for (int g = (int)blockIdx.x * per; g < g_stop; g = g_next) {
        const LargeItem it = chunk_owner(items, count, g);
        const int c = g - it.chunk_base;  // chunks [c, c_stop) of this range
        const int c_stop = min(g_stop - it.chunk_base, it.nch);
        g_next = max(g + 1, it.chunk_base + c_stop);
        const int len = (c_stop == it.nch ? it.n : c_stop * it.chunk_len) - c * it.chunk_len;
        if (len <= 0) continue;  // an emptied record of a level past its chunk capacity
        const int64_t start = (int64_t)it.begin + (int64_t)c * it.chunk_len;
        const U* ks = (const U*)b.k[it.src];
        const V* vs = (const V*)b.v[it.src];
        // the range's OR / AND over its chunks (thread x holds chunk x's record)
        unsigned long long o = 0, a = ~0ull, my_and = 0;
        int my_lp = 0;
        if (t < it.nch) {
            const ChunkInfo ci = sc.cinfo[level][it.chunk_base + t];
            o = ci.kor;
            a = ci.kand;
            my_and = ci.kand;
            my_lp = ci.lp;
        }
        o = wave_or(o);
        a = wave_and(a);
        if ((t & 63) == 0) {
            red_or[t >> 6] = o;
            red_and[t >> 6] = a;
        }
        __syncthreads();
        o = 0;
        a = ~0ull;
        for (int w = 0; w < NW; ++w) {
            o |= red_or[w];
            a &= red_and[w];
        }
        const unsigned long long var = o ^ a;  // already inside [begin_bit, hi_bit)
        if (var == 0) {
            U* ko = buf_ptr<U>(b.k[BUF_OUT]);
            V* vo = buf_ptr<V>(b.v[BUF_OUT]);
            for (int i = t; i < len; i += NT) {
                ko[start + i] = ks[start + i];
                if constexpr (HAS_V) vo[start + i] = vs[start + i];
            }
            __syncthreads();
            continue;
        }
        const int pstar = (bit_length<unsigned long long>(var) - 1 - km.begin_bit) / RADIX_BITS;
        const int shift = km.begin_bit + pstar * RADIX_BITS;
        const int bits = (it.hi_bit - shift) < RADIX_BITS ? (it.hi_bit - shift) : RADIX_BITS;
        const unsigned long long dmask = (1ull << bits) - 1;
        const bool final_ = (var & ((1ull << shift) - 1)) == 0;
        // each chunk's digit selector: -1 when its histogram counts this level's
        // digit position, else the one digit value all of its items hold there
        if (t < it.nch) csel[t] = my_lp == pstar ? -1 : (int)((my_and >> shift) & dmask);
        __syncthreads();
        // digit d's total over the range and over the chunks before this one:
        // the group's share of the histograms is read XB chunks at a time, all
        // loads of a batch in flight together (a share of 64 chunks takes two
        // round trips with XB = 32; a short share takes the 8-chunk batches)
        uint32_t tot = 0, pre = 0;
        {
            const int x0 = __builtin_amdgcn_readfirstlane((it.nch * grp) / G);
            const int x1 = __builtin_amdgcn_readfirstlane((it.nch * (grp + 1)) / G);
            auto totals = [&](auto xbc) {
                constexpr int XB = decltype(xbc)::value;
                for (int xb = x0; xb < x1; xb += XB) {
                    uint32_t h[XB];
#pragma unroll
                    for (int q = 0; q < XB; ++q) {
                        const size_t gc = (size_t)it.chunk_base + (size_t)(xb + q < x1 ? xb + q : x1 - 1);
                        h[q] = sc.hist[level][gc * RADIX + d];  // read even when unused: one batch in flight
                    }
#pragma unroll
                    for (int q = 0; q < XB; ++q) {
                        const int x = xb + q;
                        if (x < x1) {
                            const int sel = csel[x];
                            const uint32_t k =
                                sel < 0 ? h[q] : (sel == d ? (uint32_t)chunk_items(it, x) : 0u);
                            tot += k;
                            pre += x < c ? k : 0u;
                        }
                    }
                }
            };
            if (x1 - x0 > 8) totals(std::integral_constant<int, 32>{});
            else totals(std::integral_constant<int, 8>{});
        }
        uint32_t* part = sm.wcnt;  // free until the partition's first ranking pass
        part[grp * RADIX + d] = tot;
        part[(G + grp) * RADIX + d] = pre;
        __syncthreads();
        uint32_t total = 0, before = 0, incl = 0;
        if (t < RADIX) {
#pragma unroll
            for (int q = 0; q < G; ++q) {
                total += part[q * RADIX + t];
                before += part[(G + q) * RADIX + t];
            }
            incl = wave_incl_scan(total);
            if ((t & 63) == 63) wsum[t >> 6] = incl;
        }
        __syncthreads();
        if (t < RADIX) {
            uint32_t base = incl - total;
            for (int w = 0; w < (t >> 6); ++w) base += wsum[w];
            cstart[t] = base;
            if (t == RADIX - 1) cstart[RADIX] = base + total;
            run[t] = (uint32_t)it.begin + base + before;
        }
        __syncthreads();
        if (c == 0 && !final_) {
            const uint32_t cb = t < RADIX ? cstart[t] : 0u;
            const int cn = t < RADIX ? (int)(cstart[t + 1] - cb) : 0;
            const int cls = (cn >= 1 && cn <= tile_max) ? class_of(cc, cn) : -1;
            push_tiles_block(sc, bp, cls, make_ti(it.begin + (int)cb, cls >= 0 ? cn : 0, child_buf));
            if (cn > tile_max) {
                if (level == MSD_LEVEL0 && cn > big)
                    push_large(sc, MSD_LEVEL1, it.begin + (int)cb, cn, child_buf, shift, true);
                else
                    push_large(sc, MSD_LOCAL, it.begin + (int)cb, cn, child_buf, shift, false);
            }
        }
        const int dbuf = final_ ? BUF_OUT : child_buf;
        T::partition_range(sm, ks, vs, buf_ptr<U>(b.k[dbuf]), buf_ptr<V>(b.v[dbuf]), start, len, km, shift, bits, run);
    }
}

// The depth-first tier: one workgroup takes a range that a multi-workgroup level
// queued and partitions it, and every descendant still larger than a tile,
// itself, depth first, so that no workgroup ever waits for another. Each
// partition histograms the range for every digit position below its hi_bit,
// picks the highest position whose digit takes more than one value, and
// scatters the range chunk by chunk (running destinations in LDS) into the
// other temporary buffer, or into the output when no lower digit varies.
// Children that fit a tile go to the tile queues; a range whose keys are all
// equal within the bit range is copied out as it is.
template <int NT, int IPT, typename U, typename V, bool HAS_V>
__global__ __launch_bounds__(NT) void k_large_local(Bufs b, Sched sc, int level, KeyMode km, ClassCaps cc,
                                                    int tile_max) {
    using T = RadixTile<NT, IPT, U, V, HAS_V>;
    constexpr int NPOS = int(sizeof(U));  // 8-bit digit positions of a key
    constexpr int NW = NT / WAVE;
    __shared__ union {
        typename T::Smem sm;
        struct {  // the histogram phase, before the tile's LDS is used
            uint32_t hist[RADIX];
            U red_or[NW];
            U red_and[NW];
        };
    } u;
    __shared__ uint32_t cstart[NPOS][RADIX + 1];  // child starts (range-relative) at each depth
    __shared__ uint32_t run[RADIX];               // running destinations of a scatter
    __shared__ uint32_t wsum[RADIX / WAVE];
    __shared__ int fb[NPOS + 1], fn[NPOS + 1], fsrc[NPOS + 1], fhi[NPOS + 1], fdst[NPOS + 1], fshift[NPOS + 1],
        fnext[NPOS + 1];
    __shared__ int sel;
    __shared__ BlockPush bp;
    const int t = threadIdx.x;
    int count, chunks;
    level_counts(sc, level, count, chunks);
    // workgroups beyond the range count (all of them when the tier has no work)
    // leave without a claim: workgroup 0 keeps claiming until the ranges run out
    if ((int)blockIdx.x >= count) return;
    const LargeItem* items = sc.lq[level];
    for (;;) {
        if (t == 0) sel = (int)atomicAdd(&sc.ctr->claim, 1u);
        __syncthreads();
        const int i = sel;
        __syncthreads();
        if (i >= count) break;
        if (t == 0) {
            const LargeItem it = items[i];
            fb[0] = it.begin;
            fn[0] = it.n;
            fsrc[0] = it.src;
            fhi[0] = it.hi_bit;
        }
        __syncthreads();
        int depth = 0;
        bool fresh = true;
        while (depth >= 0) {
            if (fresh) {
                const int beg = fb[depth], n = fn[depth], src = fsrc[depth], hi = fhi[depth];
                const int npos = npos_of(hi, km.begin_bit);
                const U* ks = (const U*)b.k[src];
                const V* vs = (const V*)b.v[src];
                // the histogram of the range's top digit position, with the OR and
                // AND of its order keys: their highest differing bit names the
                // position to partition on, which is the top one unless the keys
                // share their top digit (then that position is counted again)
                const U rmask = range_mask<U>(km.begin_bit, hi);
                auto digit_mask = [&](int sh) {
                    return U((U(1) << ((hi - sh) < RADIX_BITS ? (hi - sh) : RADIX_BITS)) - 1);
                };
                const int top = npos - 1;
                for (int x = t; x < RADIX; x += NT) u.hist[x] = 0;
                __syncthreads();
                U o = 0, a = Bits<U>::ALL;
                {
                    const int sh = km.begin_bit + top * RADIX_BITS;
                    T::histogram_or_and(u.hist, ks, beg, n, km, rmask, sh, digit_mask(sh), o, a);
                }
                o = wave_or(o);
                a = wave_and(a);
                if ((t & 63) == 0) {
                    u.red_or[t >> 6] = o;
                    u.red_and[t >> 6] = a;
                }
                __syncthreads();
                o = 0;
                a = Bits<U>::ALL;
                for (int w = 0; w < NW; ++w) {
                    o |= u.red_or[w];
                    a &= u.red_and[w];
                }
                const U var = U(o ^ a);
                const int pstar = var ? (bit_length<U>(var) - 1 - km.begin_bit) / RADIX_BITS : -1;
                if (pstar >= 0 && pstar != top) {
                    for (int x = t; x < RADIX; x += NT) u.hist[x] = 0;
                    __syncthreads();
                    U o2 = 0, a2 = Bits<U>::ALL;
                    const int sh = km.begin_bit + pstar * RADIX_BITS;
                    T::histogram_or_and(u.hist, ks, beg, n, km, rmask, sh, digit_mask(sh), o2, a2);
                    __syncthreads();
                }
                if (pstar < 0) {
                    U* ko = buf_ptr<U>(b.k[BUF_OUT]);
                    V* vo = buf_ptr<V>(b.v[BUF_OUT]);
                    for (int x = t; x < n; x += NT) {
                        ko[beg + x] = ks[beg + x];
                        if constexpr (HAS_V) vo[beg + x] = vs[beg + x];
                    }
                    __syncthreads();
                    --depth;
                    fresh = false;
                    continue;
                }
                const int shift = km.begin_bit + pstar * RADIX_BITS;
                const bool final_ = (var & U((U(1) << shift) - 1)) == 0;  // no lower bit varies
                const int bits = (hi - shift) < RADIX_BITS ? (hi - shift) : RADIX_BITS;
                const int dst = final_ ? BUF_OUT : (src == BUF_TMP_A ? BUF_TMP_B : BUF_TMP_A);
                uint32_t mine = 0, incl = 0;
                if (t < RADIX) {
                    mine = u.hist[t];
                    incl = wave_incl_scan(mine);
                    if (lane_id() == 63) wsum[t >> 6] = incl;
                }
                __syncthreads();
                if (t < RADIX) {
                    uint32_t excl = incl - mine;
                    for (int ww = 0; ww < (t >> 6); ++ww) excl += wsum[ww];
                    cstart[depth][t] = excl;
                    if (t == RADIX - 1) cstart[depth][RADIX] = excl + mine;
                    run[t] = (uint32_t)beg + excl;
                }
                if (t == 0) {
                    fshift[depth] = shift;
                    fdst[depth] = dst;
                    fnext[depth] = 0;
                }
                __syncthreads();  // the histogram (aliased with the tile) is dead from here
                T::template partition_range<false>(u.sm, ks, vs, buf_ptr<U>(b.k[dst]), buf_ptr<V>(b.v[dst]),
                                                   (int64_t)beg, n, km, shift, bits, run);
                if (!final_) {
                    const uint32_t cb = t < RADIX ? cstart[depth][t] : 0u;
                    const uint32_t cn = t < RADIX ? cstart[depth][t + 1] - cb : 0u;
                    const int cls = (cn >= 1 && (int)cn <= tile_max) ? class_of(cc, (int)cn) : -1;
                    push_tiles_block(sc, bp, cls, make_ti(beg + (int)cb, cls >= 0 ? (int)cn : 0, dst));
                }
                __syncthreads();  // the scatter's writes are visible to the workgroup before a child is read
                fresh = false;
                if (final_) {
                    --depth;
                    continue;
                }
            }
            // the next child of this depth's range that is larger than a tile
            const int from = fnext[depth];
            if (t == 0) sel = RADIX;
            __syncthreads();
            if (t < RADIX && t >= from && (int)(cstart[depth][t + 1] - cstart[depth][t]) > tile_max)
                atomicMin(&sel, t);
            __syncthreads();
            const int dn = sel;
            __syncthreads();
            if (dn >= RADIX) {
                --depth;
                continue;
            }
            if (t == 0) {
                fnext[depth] = dn + 1;
                fb[depth + 1] = fb[depth] + (int)cstart[depth][dn];
                fn[depth + 1] = (int)(cstart[depth][dn + 1] - cstart[depth][dn]);
                fsrc[depth + 1] = fdst[depth];
                fhi[depth + 1] = fshift[depth];
            }
            __syncthreads();
            ++depth;
            fresh = true;
        }
    }
}

// ---- plain form: one workgroup per large segment, global LSD ---------------

// All digit passes of one segment by one workgroup (the reference's one-CTA-
// per-segment radix, dispatch_radix_sort.cuh:313-426, with 8-bit digits):
// histogram the whole segment, scan, then rank and scatter tile by tile. A
// workgroup takes queue entries blockIdx.x, blockIdx.x + gridDim.x, ...
template <int NT, int IPT, typename U, typename V, bool HAS_V>
__global__ __launch_bounds__(NT) void k_seg_lsd(Bufs b, Sched sc, KeyMode km) {
    using T = RadixTile<NT, IPT, U, V, HAS_V>;
    __shared__ typename T::Smem sm;
    __shared__ uint32_t hist[RADIX];
    __shared__ uint32_t run[RADIX];
    __shared__ uint32_t wsum[RADIX / WAVE > NT / WAVE ? RADIX / WAVE : NT / WAVE];
    const uint32_t nl = (uint32_t)sc.ctr->level[MAX_LEVELS];
    const int count = (int)(nl < sc.lq_cap ? nl : sc.lq_cap);
    const int nbits = km.end_bit - km.begin_bit;
    const int passes = (nbits + RADIX_BITS - 1) / RADIX_BITS;
    for (int q = blockIdx.x; q < count; q += gridDim.x) {
        if (q != (int)blockIdx.x) __syncthreads();  // the previous segment's last partition is done with LDS
        const LargeItem it = sc.lq[MAX_LEVELS][q];
        for (int ps = 0; ps < passes; ++ps) {
            // pass ps reads IN (ps==0) or the ping buffer, writes OUT on the last pass
            const int sbuf = ps == 0 ? BUF_IN : ((ps & 1) ? BUF_TMP_A : BUF_TMP_B);
            const int dbuf = ps == passes - 1 ? BUF_OUT : ((ps & 1) ? BUF_TMP_B : BUF_TMP_A);
            const int shift = km.begin_bit + ps * RADIX_BITS;
            const int bits = (km.end_bit - shift) < RADIX_BITS ? (km.end_bit - shift) : RADIX_BITS;
            const U* ks = (const U*)b.k[sbuf];
            const V* vs = (const V*)b.v[sbuf];
            for (int i = threadIdx.x; i < RADIX; i += NT) hist[i] = 0;
            __syncthreads();
            T::histogram(hist, ks, it.begin, it.n, km, shift, 1, shift + bits);
            __syncthreads();
            // exclusive scan of hist -> run (begin-relative global offsets)
            uint32_t mine = 0;
            if ((int)threadIdx.x < RADIX) mine = hist[threadIdx.x];
            const uint32_t incl = wave_incl_scan(mine);
            if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
            __syncthreads();
            if ((int)threadIdx.x < RADIX) {
                uint32_t base = incl - mine;
                for (int ww = 0; ww < (int)(threadIdx.x >> 6); ++ww) base += wsum[ww];
                run[threadIdx.x] = (uint32_t)it.begin + base;
            }
            __syncthreads();
            T::partition_range(sm, ks, vs, buf_ptr<U>(b.k[dbuf]), buf_ptr<V>(b.v[dbuf]), (int64_t)it.begin, it.n, km,
                               shift, bits, run);
        }
    }
}

// ---- fix-pass: one global LSD over (segment id, key) composites ---------------

// Max of the order keys over [begin[0], end[S-1]) (fixcub.cu:141-143 takes the
// max to size the segment-id shift). IPT loads per thread are in flight at
// once; the clamped tail loads repeat the last key, which leaves the max as is.
template <typename U, int NT, int IPT>
__global__ __launch_bounds__(NT) void k_fix_max(const U* __restrict__ keys, const int* d_begin, const int* d_end,
                                                int S, KeyMode km, unsigned long long* dmax) {
    const int lo = d_begin[0];
    const int hi = d_end[S - 1];
    const U rm = range_mask<U>(km.begin_bit, km.end_bit);
    unsigned long long m = 0;
    for (int64_t b0 = lo + (int64_t)blockIdx.x * NT * IPT; b0 < hi; b0 += (int64_t)gridDim.x * NT * IPT) {
        U x[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int64_t i = b0 + k * NT + threadIdx.x;
            x[k] = keys[i < hi ? i : hi - 1];
        }
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const unsigned long long okey = (unsigned long long)((twiddle_in<U>(x[k], km.kind, km.desc) & rm) >> km.begin_bit);
            m = okey > m ? okey : m;
        }
    }
    for (int o = 1; o < WAVE; o <<= 1) {
        const unsigned long long y = __shfl_xor(m, o, WAVE);
        m = y > m ? y : m;
    }
    // one atomic per workgroup: thousands of same-address atomics serialise
    __shared__ unsigned long long wm[NT / WAVE];
    if (lane_id() == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / WAVE; ++w) m = wm[w] > m ? wm[w] : m;
        if (m) atomicMax(dmax, m);
    }
}

// composite[i - lo] = (segment id << kbits) | order key, and a payload per element.
// Segments must tile [begin[0], end[S-1]) in order (CSR offsets), as the
// reference's fix-pass assumes (fixcub.cu:103-108); an element belongs to the
// last segment whose begin is at or before it. Each workgroup takes NT * IPT
// consecutive elements: two binary searches give the segments that start in
// it, each marks its first element in LDS (the largest id wins, so empty
// segments sharing a begin give way to the one that holds the elements), and
// a running max over the marks gives every element its segment.
//
// The payload is the input index (FIX_PAY_INDEX: the last step gathers keys and
// values by it), the input value itself (FIX_PAY_VALUE), or nothing
// (FIX_PAY_NONE, keys only). With a value or no payload the downsweeps write
// the output themselves; when no pass runs at all (one segment, every order
// key 0) the input is already in order and is copied to the output here.
constexpr int FIX_PAY_INDEX = 0, FIX_PAY_VALUE = 1, FIX_PAY_NONE = 2;

template <typename U, int NT, int IPT>
__global__ __launch_bounds__(NT) void k_fix_composite(const U* __restrict__ keys, const int* __restrict__ d_begin,
                                                      const int* d_end, int S, KeyMode km,
                                                      const unsigned long long* dmax, int segbits, int pay_mode,
                                                      const uint32_t* __restrict__ vals, unsigned long long* comp,
                                                      uint32_t* pay, U* keys_out, uint32_t* vals_out) {
    constexpr int CH = NT * IPT;
    constexpr int PAD = CH + CH / 16;  // one spare word per 16 keeps the per-thread runs off one bank
    __shared__ uint32_t head[PAD];
    __shared__ uint32_t wmax[NT / WAVE];
    __shared__ int srange[2];
    auto at = [](int p) { return p + (p >> 4); };
    const int lo = d_begin[0];
    const int hi = d_end[S - 1];
    const int m = hi - lo;
    const U rm = range_mask<U>(km.begin_bit, km.end_bit);
    const int kbits = bit_length<unsigned long long>(*dmax);
    const int t = threadIdx.x;
    if (pay_mode != FIX_PAY_INDEX && kbits + segbits == 0) {
        for (int64_t i = lo + (int64_t)blockIdx.x * NT + t; i < hi; i += (int64_t)gridDim.x * NT) {
            keys_out[i] = keys[i];
            if (pay_mode == FIX_PAY_VALUE) vals_out[i] = vals[i];
        }
        return;
    }
    for (int64_t c0 = (int64_t)blockIdx.x * CH; c0 < m; c0 += (int64_t)gridDim.x * CH) {
        const int base = lo + (int)c0;
        const int n = (m - (int)c0) < CH ? (m - (int)c0) : CH;
        // keys (and values) first, unconditionally at clamped positions, so that
        // the loads are in flight during the segment search below
        U kx[IPT];
        uint32_t vx[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int p = k * NT + t < n ? k * NT + t : n - 1;
            kx[k] = keys[base + p];
            vx[k] = pay_mode == FIX_PAY_VALUE ? vals[base + p] : 0u;
        }
        for (int p = t; p < CH; p += NT) head[at(p)] = 0;
        if (t < 2) {  // t 0: last segment with begin <= base; t 1: last with begin < base + n
            const int key = t == 0 ? base : base + n - 1;
            int a = 0, z = S - 1;
            while (a < z) {
                const int mid = (a + z + 1) >> 1;
                if (d_begin[mid] <= key) a = mid;
                else z = mid - 1;
            }
            srange[t] = a;
        }
        __syncthreads();
        const int s0 = srange[0], s1 = srange[1];
        for (int sg = s0 + 1 + t; sg <= s1; sg += NT) {
            const int p = d_begin[sg] - base;  // in [1, n) for non-decreasing begins; others are ignored
            if (p >= 0 && p < n) atomicMax(&head[at(p)], (uint32_t)sg + 1u);
        }
        __syncthreads();
        // running max over this thread's IPT consecutive positions, then across threads
        uint32_t run = 0;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t h = head[at(t * IPT + j)];
            run = h > run ? h : run;
        }
        const uint32_t incl = wave_incl_scan_op(run, 0u, [](uint32_t x, uint32_t y) { return x > y ? x : y; });
        if ((t & 63) == 63) wmax[t >> 6] = incl;
        uint32_t prev = (uint32_t)__shfl_up((int)incl, 1, WAVE);
        if (lane_id() == 0) prev = 0;
        __syncthreads();
        for (int w = 0; w < (t >> 6); ++w) prev = wmax[w] > prev ? wmax[w] : prev;
        uint32_t cur = prev > (uint32_t)s0 + 1u ? prev : (uint32_t)s0 + 1u;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t h = head[at(t * IPT + j)];
            cur = h > cur ? h : cur;
            head[at(t * IPT + j)] = cur - 1u;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int p = k * NT + t;
            if (p < n) {
                const unsigned long long okey =
                    (unsigned long long)((twiddle_in<U>(kx[k], km.kind, km.desc) & rm) >> km.begin_bit);
                comp[c0 + p] = ((unsigned long long)head[at(p)] << kbits) | okey;
                if (pay_mode == FIX_PAY_INDEX) pay[c0 + p] = (uint32_t)(c0 + p);
                else if (pay_mode == FIX_PAY_VALUE) pay[c0 + p] = vx[k];
            }
        }
        __syncthreads();
    }
}

// Upsweep of one fix-pass digit pass: per-chunk digit counts, digit-major.
// The passes over segment-id bits see long runs of one digit (a chunk covers
// few segments), so a wave whose valid lanes share one digit adds its count
// with a single LDS atomic instead of 64 to one address.
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_fix_upsweep(const unsigned long long* __restrict__ comp, const int* d_begin,
                                                    const int* d_end, int S, const unsigned long long* dmax,
                                                    int segbits, int pass, int chunk, int nchunks_max,
                                                    uint32_t* hist) {
    __shared__ uint32_t h[RADIX];
    const int m = d_end[S - 1] - d_begin[0];
    const int total_bits = bit_length<unsigned long long>(*dmax) + segbits;
    const int shift = pass * RADIX_BITS;
    if (shift >= total_bits) return;
    const int nch = (m + chunk - 1) / chunk;
    const int t = threadIdx.x;
    for (int g = blockIdx.x; g < nch; g += gridDim.x) {
        for (int i = t; i < RADIX; i += NT) h[i] = 0;
        __syncthreads();
        const int start = g * chunk;
        const int n = (m - start) < chunk ? (m - start) : chunk;
        for (int i0 = 0; i0 < n; i0 += NT * IPT) {
            uint32_t dg[IPT];
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = i0 + j * NT + t;
                dg[j] = (uint32_t)(comp[start + (i < n ? i : n - 1)] >> shift) & 0xFFu;
            }
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const bool v = i0 + j * NT + t < n;
                const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)dg[j]);
                const uint64_t act = __ballot(v);
                if (__ballot(v && dg[j] != d0) == 0) {
                    if (lane_id() == 0 && act) atomicAdd(&h[d0], (uint32_t)__popcll(act));
                } else if (v) {
                    atomicAdd(&h[dg[j]], 1u);
                }
            }
        }
        __syncthreads();
        for (int i = t; i < RADIX; i += NT) hist[(size_t)i * nchunks_max + g] = h[i];
        __syncthreads();
    }
}

// Spine of one fix-pass digit pass, one workgroup per digit: the exclusive
// scan of the digit's per-chunk counts over the chunks (in place), and the
// digit's total, from which each downsweep workgroup forms the digit bases.
template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_fix_spine(const int* d_begin, const int* d_end, int S,
                                                  const unsigned long long* dmax, int segbits, int pass, int chunk,
                                                  int nchunks_max, uint32_t* hist, uint32_t* totals) {
    __shared__ uint32_t wsum[NT / WAVE];
    const int m = d_end[S - 1] - d_begin[0];
    const int total_bits = bit_length<unsigned long long>(*dmax) + segbits;
    if (pass * RADIX_BITS >= total_bits) return;
    const int nch = (m + chunk - 1) / chunk;
    uint32_t* h = hist + (size_t)blockIdx.x * nchunks_max;
    const int t = threadIdx.x;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < nch; c0 += NT * IPT) {
        const int first = c0 + t * IPT;
        uint32_t x[IPT];
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            x[j] = first + j < nch ? h[first + j] : 0u;
            sum += x[j];
        }
        const uint32_t incl = wave_incl_scan(sum);
        if ((t & 63) == 63) wsum[t >> 6] = incl;
        __syncthreads();
        uint32_t base = carry + incl - sum, blk = 0;
        for (int w = 0; w < NT / WAVE; ++w) {
            const uint32_t ws = wsum[w];
            if (w < (t >> 6)) base += ws;
            blk += ws;
        }
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (first + j < nch) h[first + j] = base;
            base += x[j];
        }
        carry += blk;
        __syncthreads();
    }
    if (t == 0) totals[blockIdx.x] = carry;
}

// Downsweep of one fix-pass digit pass: stable partition of each chunk of
// composites (with their payload: the input index, or the value itself) by the
// pass digit into the other ping-pong buffer. With keys_final set (the order
// key is the whole 32-bit key, so a composite's low kbits bits, kbits the bit
// length of the key max, are the twiddled key) the last pass that runs writes
// the sorted keys and values to the output
// instead; which pass is last depends on the key max, so each pass checks.
template <int NT, int IPT, bool HAS_P>
__global__ __launch_bounds__(NT) void k_fix_downsweep(const unsigned long long* comp_in, const uint32_t* pay_in,
                                                      unsigned long long* comp_out, uint32_t* pay_out,
                                                      const int* d_begin, const int* d_end, int S,
                                                      const unsigned long long* dmax, int segbits, int pass,
                                                      int nchunks_max, const uint32_t* hist,
                                                      const uint32_t* totals, uint32_t* keys_final,
                                                      uint32_t* vals_final, KeyMode kmu) {
    using T = RadixTile<NT, IPT, unsigned long long, uint32_t, HAS_P>;
    static_assert(NT >= RADIX, "one thread per digit for the digit bases");
    __shared__ typename T::Smem sm;
    __shared__ uint32_t dbase[RADIX];
    __shared__ uint32_t tws[RADIX / WAVE];
    const int lo = d_begin[0];
    const int m = d_end[S - 1] - lo;
    const int total_bits = bit_length<unsigned long long>(*dmax) + segbits;
    const int shift = pass * RADIX_BITS;
    if (shift >= total_bits) return;
    // digit bases: exclusive scan of the spine's digit totals
    const int t = threadIdx.x;
    if (t < RADIX) {
        const uint32_t tt = totals[t];
        const uint32_t inc = wave_incl_scan(tt);
        if ((t & 63) == 63) tws[t >> 6] = inc;
        dbase[t] = inc - tt;
    }
    __syncthreads();
    if (t < RADIX) {
        uint32_t add = 0;
        for (int w = 0; w < (t >> 6); ++w) add += tws[w];
        dbase[t] += add;
    }
    const int bits = (total_bits - shift) < RADIX_BITS ? (total_bits - shift) : RADIX_BITS;
    const bool last = keys_final != nullptr && shift + RADIX_BITS >= total_bits;
    const unsigned long long kmask = (1ull << (total_bits - segbits)) - 1;  // order-key bits of a composite
    const int nch = (m + T::TILE - 1) / T::TILE;
    KeyMode km;
    km.kind = KIND_UNSIGNED;
    km.desc = 0;
    km.begin_bit = 0;
    km.end_bit = 64;
    for (int g = blockIdx.x; g < nch; g += gridDim.x) {
        // thread d wrote dbase[d] itself; partition_with passes barriers before it reads goff
        for (int d = t; d < RADIX; d += NT) sm.goff[d] = hist[(size_t)d * nchunks_max + g] + dbase[d];
        const int start = g * T::TILE;
        const int n = (m - start) < T::TILE ? (m - start) : T::TILE;
        if (last) {
            T::partition_with(sm, comp_in, pay_in, start, n, km, shift, bits, sm.goff,
                              [&](uint32_t gi, unsigned long long key, uint32_t val) {
                                  keys_final[lo + gi] = twiddle_out<uint32_t>((uint32_t)(key & kmask), kmu.kind,
                                                                              kmu.desc);
                                  if constexpr (HAS_P) vals_final[lo + gi] = val;
                              });
        } else {
            T::partition_with(sm, comp_in, pay_in, start, n, km, shift, bits, sm.goff,
                              [&](uint32_t gi, unsigned long long key, uint32_t val) {
                                  comp_out[gi] = key;
                                  if constexpr (HAS_P) pay_out[gi] = val;
                              });
        }
    }
}

// Final fix-pass gather: sorted position k takes the key and value of input
// element lo + idx[k]; the data sits in ping-pong buffer (passes run) & 1.
template <typename U, typename V, bool HAS_V>
__global__ void k_fix_gather(Bufs b, const uint32_t* idx0, const uint32_t* idx1, const int* d_begin, const int* d_end,
                             int S, const unsigned long long* dmax, int segbits) {
    const int lo = d_begin[0];
    const int m = d_end[S - 1] - lo;
    const int total_bits = bit_length<unsigned long long>(*dmax) + segbits;
    const int passes = (total_bits + RADIX_BITS - 1) / RADIX_BITS;
    const uint32_t* idx = (passes & 1) ? idx1 : idx0;
    const U* ki = (const U*)b.k[BUF_IN];
    const V* vi = (const V*)b.v[BUF_IN];
    U* ko = buf_ptr<U>(b.k[BUF_OUT]);
    V* vo = buf_ptr<V>(b.v[BUF_OUT]);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const uint32_t src = idx[k];
        ko[lo + k] = ki[lo + src];
        if constexpr (HAS_V) vo[lo + k] = vi[lo + src];
    }
}

// ---- differentiable pair sort helpers ---------------------------------------

// p[i] = i for i < n (the input positions an index sort carries as values).
__global__ void k_iota(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

// vals[i] = i - begin[r] for every i of row r (one workgroup per row).
__global__ void k_row_iota(const int* d_begin, const int* d_end, int rows, int* vals) {
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int b0 = d_begin[r], e0 = d_end[r];
        for (int i = b0 + threadIdx.x; i < e0; i += blockDim.x) vals[i] = i - b0;
    }
}

// Backward of the pair sort: gin[begin[r] + perm[i]] = gout[i] for every i of
// row r, perm holding row-local source indices (one workgroup per row). A row
// of at most PERM_LDS items is permuted inside LDS: gout and perm are read in
// order, each value is stored at its row position in LDS, and the row is
// written out in order (a direct scatter writes 4 bytes at a time to scattered
// addresses and runs at a fraction of the bandwidth). Longer rows scatter
// straight to global memory. Positions a row's perm does not name are left
// unwritten in the direct scatter; perm must be a permutation of 0..n-1 (as
// the forward sort makes it) for the two paths to agree.
constexpr int PERM_LDS = 8192;

template <int NT>
__global__ __launch_bounds__(NT) void k_perm_scatter_f32(const float* __restrict__ gout, const int* __restrict__ perm,
                                                         const int* d_begin, const int* d_end, int rows,
                                                         float* __restrict__ gin) {
    constexpr int U = 4;  // loads in flight per thread
    __shared__ float buf[PERM_LDS];
    const int t = threadIdx.x;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int b0 = d_begin[r], e0 = d_end[r];
        const int n = e0 - b0;
        if (n > PERM_LDS) {
            for (int i = b0 + t; i < e0; i += NT) gin[b0 + perm[i]] = gout[i];
            continue;
        }
        for (int i0 = 0; i0 < n; i0 += NT * U) {
            float g[U];
            int p[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * NT + t;
                const int ic = i < n ? i : 0;
                g[u] = gout[b0 + ic];
                p[u] = i < n ? perm[b0 + ic] : -1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if ((unsigned)p[u] < (unsigned)n) buf[p[u]] = g[u];
        }
        __syncthreads();
        for (int j = t; j < n; j += NT) gin[b0 + j] = buf[j];
        __syncthreads();
    }
}

}  // namespace ssort
