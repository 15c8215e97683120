// Workgroup- and wave-level building blocks:
//  * RadixTile: stable LSD radix sort of one range in LDS (digit ranking with
//    64-bit ballots), the stable one-digit partition of one chunk (used by the
//    multi-workgroup MSD tier, the plain form's per-segment LSD and the
//    fix-pass LSD) and chunk histograms;
//  * wave sorts of short ranges: a bitonic network over packed sort words held
//    in registers, and a compare-count rank for the shortest ranges.
#pragma once

#include "common.hpp"

#include <type_traits>

namespace ssort {

// Element i of an array at a wave-uniform base: the byte offset is formed in 32
// bits, so the access is a scalar base plus a 32-bit lane offset (one VALU
// operation for the address instead of a 64-bit add). For accesses inside one
// range, chunk or tile (i * sizeof(T) < 2^32).
template <typename T>
__device__ __forceinline__ T& at32(T* base, uint32_t i) {
    using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return *reinterpret_cast<T*>(reinterpret_cast<C*>(base) + (uint32_t)(i * (uint32_t)sizeof(T)));
}

// NT threads hold IPT items each in "wave-striped" order: item j of lane l of
// wave w is tile position w*64*IPT + j*64 + l, which is the input order, so
// ranking items in (w, j, l) order is stable. Only the first
// wact = ceil(n / (64*IPT)) waves hold items; the others skip the per-item
// work, so the cost follows n with a granularity of 64*IPT items.
template <int NT, int IPT, typename U, typename V, bool HAS_V>
struct RadixTile {
    static constexpr int NW = NT / WAVE;
    static constexpr int TILE = NT * IPT;
    static constexpr int WITEMS = WAVE * IPT;
    static constexpr int NDT = NT < RADIX ? NT : RADIX;  // threads in the digit scan
    static constexpr int DPT = RADIX / NDT;               // digits per scanning thread
    static constexpr int NDW = (NDT + WAVE - 1) / WAVE;   // waves in the digit scan
    // Bucket finish (every tile size): one pass puts each item in its bucket of
    // the top varying bits (up to NW*256 buckets, about four items each), then
    // a compare-count of sort words inside the bucket gives its stable rank;
    // taken while no bucket exceeds BUCKET_GUARD items.
    static constexpr int BUCKET_GUARD = 32;
    // tiles of at most 4 waves and 1024 items: 8-bit buckets, counted and
    // scanned by one wave, no sample check (measured faster there than wider
    // buckets, whose larger code made the shared wave/tile kernel spill)
    static constexpr bool SMALL_BUCKET = NW <= 4 && TILE <= 1024;
    static constexpr int SMALL_GUARD = 24;
    // (when the sort words -- 8 bytes per item over keys[] and idx[] -- fit in LDS)
    static constexpr bool BUCKET_FINISH = size_t(TILE) * (sizeof(U) + 4) + size_t(NW) * RADIX * 4 +
                                              size_t(NT) * 4 + 4096 <= size_t(160 * 1024);
    static constexpr int NBK_LOG = NW <= 1 ? 8 : NW <= 2 ? 9 : NW <= 4 ? 10 : NW <= 8 ? 11 : 12;  // log2(NW*RADIX)
    static_assert((1 << NBK_LOG) == NW * RADIX, "bucket counters are the wave-count words");
    // larger tiles whose buckets overflow (skewed keys): two passes on the top
    // 16 varying bits, then a compare-count inside each run of equal top bits
    // (runs of at most TOP_GUARD items); keys only: when a 16-bit place per
    // slot still fits in LDS
    static constexpr bool TOP2_FINISH =
        TILE > 1024 && (HAS_V || size_t(TILE) * (sizeof(U) + 2) + size_t(NW) * RADIX * 4 + size_t(NT) * 4 + 4096 <=
                                     size_t(160 * 1024));
    static constexpr int TOP_GUARD = 16;
    static_assert(IPT <= 32, "one bit per slot");
    static_assert(NT % WAVE == 0, "whole waves");
    static_assert(RADIX % NDT == 0, "digit split");
    static_assert(TILE <= 65536, "16-bit source indices");

    struct alignas(16) Smem {
        U keys[TILE];
        union {
            V vals[HAS_V ? TILE : 1];                       // partition(): values ride along
            // sort(): source index per item, ping-pong; the top-16 finish keeps
            // each slot's final place in the last array. The bucket finish reads
            // keys[] and this array as one array of TILE 64-bit sort words
            // (32-bit keys), or keeps each slot's input position in idx[0]
            // (64-bit keys).
            uint16_t idx[(HAS_V || BUCKET_FINISH) ? 2 : 1][(HAS_V || TOP2_FINISH || BUCKET_FINISH) ? TILE : 1];
        };
        uint32_t wcnt[NW * RADIX];  // per-wave digit counts, then scatter bases
        uint32_t dummy[NT];         // store target of the lanes that do not advance a count
        uint32_t dstart[RADIX + 1]; // tile-exclusive start of each digit (dstart[RADIX]: slots counted)
        uint32_t maxcnt;            // largest digit count of the last pass over real items (NW <= 4)
        uint32_t goff[RADIX];       // global destination of each digit (partition)
        uint32_t wsum[NDW];
        uint32_t bsum[NW];          // bucket scan: per-wave totals
        uint32_t bmax[NW];          // bucket scan: per-wave largest bucket
        U red_or[NW];
        U red_and[NW];
    };
    static_assert(!BUCKET_FINISH || sizeof(U) == 8 || (sizeof(U) * TILE) % 8 == 0,
                  "sort words overlay keys[] and idx[]");

    __device__ static __forceinline__ int wave() { return threadIdx.x >> 6; }
    __device__ static __forceinline__ int pos(int j) { return wave() * WITEMS + j * WAVE + (threadIdx.x & 63); }
    // pos(0) tied to the range size, so that the per-slot addresses derived
    // from it are computed per range instead of being hoisted out of a
    // persistent caller's loop (where they would outlive the range in
    // registers, or spill)
    __device__ static __forceinline__ int lane_base(int n) {
        int b = wave() * WITEMS + (threadIdx.x & 63);
        __asm__("" : "+v"(b) : "s"(n));
        return b;
    }
    __device__ static __forceinline__ int active_waves(int n) { return (n + WITEMS - 1) / WITEMS; }

    // Lanes of this wave whose item has the same NBITS-bit digit d as the
    // calling lane: for every bit, AND in NOT((ballot of the bit) XOR (own bit
    // replicated)). v_bitop3 does that in one instruction per 32-lane half
    // (truth table 0x90 = src0 & ~(src1 ^ src2)). The replicated bits are
    // hidden from the optimiser so that each ballot compares one with zero
    // (one instruction) instead of re-deriving the bit, and all bits are
    // extracted and compared before they are combined, so that the compares'
    // results are not read right after they are written.
    template <int NBITS>
    __device__ static __forceinline__ uint64_t match_digit(uint32_t d) {
        uint32_t t[NBITS];
        uint64_t bal[NBITS];
#pragma unroll
        for (int b = 0; b < NBITS; ++b) {
            t[b] = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);
            __asm__("" : "+v"(t[b]));
        }
#pragma unroll
        for (int b = 0; b < NBITS; ++b) bal[b] = __ballot(t[b] != 0);
        uint32_t acc_lo = ~0u, acc_hi = ~0u;
#pragma unroll
        for (int b = 0; b < NBITS; ++b) {
            acc_lo = __builtin_amdgcn_bitop3_b32(acc_lo, (uint32_t)bal[b], t[b], 0x90);
            acc_hi = __builtin_amdgcn_bitop3_b32(acc_hi, (uint32_t)(bal[b] >> 32), t[b], 0x90);
        }
        return (uint64_t(acc_hi) << 32) | uint64_t(acc_lo);
    }

    __device__ static __forceinline__ uint64_t match_digit_rt(uint32_t d, int bits) {
        uint32_t acc_lo = ~0u, acc_hi = ~0u;
        for (int b = 0; b < bits; ++b) {
            uint32_t t = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);
            __asm__("" : "+v"(t));
            const uint64_t bal = __ballot(t != 0);
            acc_lo = __builtin_amdgcn_bitop3_b32(acc_lo, (uint32_t)bal, t, 0x90);
            acc_hi = __builtin_amdgcn_bitop3_b32(acc_hi, (uint32_t)(bal >> 32), t, 0x90);
        }
        return (uint64_t(acc_hi) << 32) | uint64_t(acc_lo);
    }

    // Per-wave counting of one item per lane: every lane reads the wave's
    // running count of its digit (lanes of one digit read the same word), and
    // the lowest lane of each digit group stores the count advanced by the
    // group's size; the other lanes store into a private dummy word, so the
    // code has no branch. LDS operations of one wave complete in order, so the
    // next item's read sees this item's store.
    // The count is read before the peers are matched, so that the match's
    // instructions cover the LDS round trip.
    template <int NBITS>
    __device__ static __forceinline__ uint32_t count_one(uint32_t* wc, uint32_t* dummy, uint32_t d, int bits) {
        const uint32_t base = wc[d];
        __builtin_amdgcn_sched_barrier(0);  // keep the read ahead of the match
        uint64_t peers;
        if constexpr (NBITS > 0) peers = match_digit<NBITS>(d);
        else peers = match_digit_rt(d, bits);
        const uint32_t below = lanes_below(peers);
        uint32_t* dst = below == 0 ? &wc[d] : dummy;
        *dst = base + (uint32_t)__popcll(peers);
        return base + below;
    }

    // Stable rank of every valid item (position < n) by digit
    // (key >> shift) & (2^bits - 1): dst[j] is the item's tile position after a
    // stable counting sort; sm.dstart[d] the first position of digit d. Slots
    // at or past n hold the all-ones key and come after every valid item in
    // rank order, so they are counted like any item without disturbing the
    // ranks of valid items (they only add to the counts of the last digit, in
    // the last active wave, after every valid item of that digit).
    __device__ static void rank_digits(Smem& sm, const U (&k)[IPT], int n, int shift, int bits,
                                       uint32_t (&dst)[IPT]) {
        const int t = threadIdx.x;
        const int w = wave();
        const uint32_t l = t & 63;
        const int wact = active_waves(n);
        uint32_t* wc = sm.wcnt + w * RADIX;
        uint32_t* dummy = sm.dummy + t;
        const U dmask = U((U(1) << bits) - 1);
        if (w < wact) {
#pragma unroll
            for (int q = 0; q < RADIX / WAVE; ++q) wc[q * WAVE + l] = 0;
            if (bits == RADIX_BITS) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const uint32_t d = (uint32_t)((k[j] >> shift) & U(RADIX - 1));
                    dst[j] = count_one<RADIX_BITS>(wc, dummy, d, RADIX_BITS);
                }
            } else {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const uint32_t d = (uint32_t)((k[j] >> shift) & dmask);
                    dst[j] = count_one<0>(wc, dummy, d, bits);
                }
            }
        }
        __syncthreads();
        if constexpr (NW <= 4) {
            // one wave scans the counts: lane l takes digits l, l+64, l+128, l+192
            if (w == 0) {
                // all counts read first (one LDS round trip), four independent scans,
                // then all bases written
                constexpr int G = RADIX / WAVE;
                uint32_t c[G][NW];
                uint32_t s[G], incl[G];
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    s[i] = 0;
#pragma unroll
                    for (int ww = 0; ww < NW; ++ww) {
                        c[i][ww] = ww < wact ? sm.wcnt[ww * RADIX + i * WAVE + (int)l] : 0u;
                        s[i] += c[i][ww];
                    }
                }
#pragma unroll
                for (int i = 0; i < G; ++i) incl[i] = wave_incl_scan(s[i]);
                uint32_t carry = 0, big = 0;
                // slots past n hold the all-ones key: they are counted in the last digit
                const uint32_t pad = (uint32_t)(wact * WITEMS - n);
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    const int d = i * WAVE + (int)l;
                    uint32_t base = carry + incl[i] - s[i];
                    carry += (uint32_t)__builtin_amdgcn_readlane((int)incl[i], 63);
                    sm.dstart[d] = base;
                    const uint32_t real = (uint32_t)d == dmask ? s[i] - pad : s[i];
                    big = real > big ? real : big;
#pragma unroll
                    for (int ww = 0; ww < NW; ++ww) {
                        if (ww < wact) sm.wcnt[ww * RADIX + d] = base;
                        base += c[i][ww];
                    }
                }
                big = wave_max(big);
                if (l == 0) {
                    sm.dstart[RADIX] = carry;
                    sm.maxcnt = big;
                }
            }
        } else {
            // thread d < RADIX owns digit d: all of its per-wave counts are read
            // at once (one LDS round trip), summed, scanned across the digits,
            // and the per-wave bases written back
            static_assert(NDT == RADIX && DPT == 1, "one digit per scanning thread");
            uint32_t c[NW];
            uint32_t s = 0;
            if (t < RADIX) {
#pragma unroll
                for (int ww = 0; ww < NW; ++ww) c[ww] = ww < wact ? sm.wcnt[ww * RADIX + t] : 0u;
#pragma unroll
                for (int ww = 0; ww < NW; ++ww) s += c[ww];
            }
            const uint32_t incl = wave_incl_scan(s);
            if (t < RADIX && l == 63) sm.wsum[w] = incl;
            __syncthreads();
            if (t < RADIX) {
                uint32_t base = incl - s;
#pragma unroll
                for (int ww = 0; ww < NDW; ++ww) base += ww < w ? sm.wsum[ww] : 0u;
                sm.dstart[t] = base;
#pragma unroll
                for (int ww = 0; ww < NW; ++ww) {
                    if (ww < wact) sm.wcnt[ww * RADIX + t] = base;
                    base += c[ww];
                }
            }
        }
        __syncthreads();
        if (w < wact) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const uint32_t d = (uint32_t)((k[j] >> shift) & dmask);
                dst[j] += wc[d];
            }
        }
    }

    // Twiddled keys of the wave's items; slots at or past n read the last
    // element (always in bounds) and are then marked with the all-ones key.
    __device__ static __forceinline__ void load_keys(const U* ks, int64_t begin, int n, const KeyMode& km,
                                                     U (&k)[IPT]) {
        const int lb = lane_base(n);
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int p = (lb + j * WAVE);
            const U raw = at32(ks + begin, (p < n ? p : n - 1));
            k[j] = p < n ? twiddle_in<U>(raw, km.kind, km.desc) : Bits<U>::ALL;
        }
    }

    // OR and AND of the valid keys over the whole tile (every thread gets them).
    __device__ static void reduce_or_and(Smem& sm, const U (&k)[IPT], int n, U& o, U& a) {
        const int lb = lane_base(n);
        U vo = 0, va = Bits<U>::ALL;
        if (wave() < active_waves(n)) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const bool v = (lb + j * WAVE) < n;
                vo |= v ? k[j] : U(0);
                va &= v ? k[j] : Bits<U>::ALL;
            }
        }
        vo = wave_or(vo);
        va = wave_and(va);
        const int w = wave();
        if ((threadIdx.x & 63) == 0) {
            sm.red_or[w] = vo;
            sm.red_and[w] = va;
        }
        __syncthreads();
        vo = 0;
        va = Bits<U>::ALL;
        for (int ww = 0; ww < NW; ++ww) {
            vo |= sm.red_or[ww];
            va &= sm.red_and[ww];
        }
        o = vo;
        a = va;
    }

    // Stable sort of n <= TILE items at [begin, begin+n) of (ks, vs) into
    // (ko, vo) at the same offsets. Only digits whose bits vary inside the tile
    // are ranked, so a low-entropy tile costs fewer passes. Values do not move
    // through the passes: a 16-bit source index does, and each value is read
    // once from its source at the end.
    __device__ static __forceinline__ void sort(Smem& sm, const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n,
                                const KeyMode& km) {
        U raw[IPT];
        load_raw(ks, begin, n, raw);
        sort_loaded(sm, raw, ks, vs, ko, vo, begin, n, km, [] {});
    }

    // The raw keys of this thread's slots (clamped, always in bounds), as
    // sort_loaded takes them. A persistent caller issues the loads of its next
    // range from sort_loaded's `prefetch` hook, which runs once the current
    // keys are no longer needed in registers (after the last pass), so they
    // overlap the current range's write-out. Workgroup barriers wait on LDS
    // only, so the loads stay in flight across them.
    __device__ static __forceinline__ void load_raw(const U* ks, int64_t begin, int n, U (&raw)[IPT]) {
        const int lb = lane_base(n);
        if (wave() < active_waves(n)) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                raw[j] = at32(ks + begin, (p < n ? p : n - 1));
            }
        }
    }

    // Buckets of a range of n items with varying order-key bits [lo, hi): the
    // top nb varying bits, about 2^nb >= n / 2 buckets, at most NW*256.
    __device__ static __forceinline__ int bucket_bits(int n, int hi, int lo) {
        int nb = bit_length<uint32_t>((uint32_t)n) - 1;
        nb = nb < NBK_LOG ? nb : NBK_LOG;
        nb = nb > RADIX_BITS ? nb : RADIX_BITS;
        return nb < hi - lo ? nb : hi - lo;
    }

    // Whether the bucket finish is likely to fit: among the range's first 64
    // keys (the sample in sm.keys), the pairs that share their bucket estimate
    // how many items share an item's bucket, (pairs / (64*63/2)) * n; the
    // finish is tried while that stays at most BUCKET_GUARD / 4. Skewed keys
    // (floats of a normal distribution) fail it and go to the top-16 passes.
    __device__ static __forceinline__ bool bucket_promising(Smem& sm, int n, const KeyMode& km, int hi, int lo) {
        const int nb = bucket_bits(n, hi, lo);
        const U rm = range_mask<U>(km.begin_bit, km.end_bit);
        const uint32_t l = threadIdx.x & 63;
        const int ns = n < WAVE ? n : WAVE;
        const U kk = sm.keys[l < (uint32_t)ns ? l : 0];
        const uint32_t b = (uint32_t)(U((kk & rm) >> (hi - nb)) & U((1u << nb) - 1));
        const uint64_t peers = match_digit_rt(b, nb) & __ballot(l < (uint32_t)ns);
        const uint32_t pairs = wave_sum(l < (uint32_t)ns ? (uint32_t)__popcll(peers) - 1u : 0u) / 2u;
        const uint32_t all_pairs = (uint32_t)(ns * (ns - 1) / 2);
        return (uint64_t)pairs * (uint64_t)n * 4u <= (uint64_t)BUCKET_GUARD * all_pairs + 4u * (uint64_t)n;
    }

    // The bucket finish. Every item goes to a slot of its bucket (the top nb
    // varying order-key bits) through LDS atomics on the block's bucket
    // counters (unordered inside a bucket), as a sort word (masked order key,
    // input position); a scan of the counters gives the bucket starts. When
    // no bucket holds more than BUCKET_GUARD items, the item at each slot is
    // ranked inside its bucket by counting the smaller sort words -- its stable
    // rank -- and written to its final place (threads take consecutive slots,
    // so the lanes of a wave share buckets: coherent loops, broadcast reads);
    // keys and values are gathered by input position. Returns false, with
    // nothing written to global memory and the prefetch hook not run, when a
    // bucket overflows.
    template <typename Prefetch>
    __device__ static __forceinline__ bool bucket_finish(Smem& sm, const U (&k)[IPT], const U* ks, const V* vs,
                                                         U* ko, V* vo, int64_t begin, int n, const KeyMode& km,
                                                         int hi, int lo, Prefetch& prefetch) {
        constexpr int CPT = (NW * RADIX) / NT;  // bucket counters per thread in the scan (4)
        const int lb = lane_base(n);
        const bool act = wave() < active_waves(n);
        const int nb = bucket_bits(n, hi, lo);
        const int shift = hi - nb;
        const uint32_t bm = (1u << nb) - 1u;
        const U rm = range_mask<U>(km.begin_bit, km.end_bit);
        uint32_t* cnt = sm.wcnt;  // NW * RADIX counters
        const int t = threadIdx.x;
        const int w = wave();
        const uint32_t l = t & 63;
#pragma unroll
        for (int i = 0; i < CPT; ++i) cnt[i * NT + t] = 0;
        __syncthreads();
        uint32_t rk[IPT];
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                if (p < n) rk[j] = atomicAdd(&cnt[(uint32_t)(k[j] >> shift) & bm], 1u);
            }
        }
        __syncthreads();
        // exclusive scan of the counters: thread t owns counters t*CPT .. t*CPT+CPT-1
        {
            uint32_t c[CPT], sum = 0, big = 0;
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                c[i] = cnt[t * CPT + i];
                sum += c[i];
                big = c[i] > big ? c[i] : big;
            }
            const uint32_t incl = wave_incl_scan(sum);
            big = wave_max(big);
            if (l == 63) sm.bsum[w] = incl;
            if (l == 0) sm.bmax[w] = big;
            __syncthreads();
            uint32_t base = incl - sum;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) base += ww < w ? sm.bsum[ww] : 0u;
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                cnt[t * CPT + i] = base;
                base += c[i];
            }
            if (t == 0) {
                uint32_t m = 0;
#pragma unroll
                for (int ww = 0; ww < NW; ++ww) m = sm.bmax[ww] > m ? sm.bmax[ww] : m;
                sm.maxcnt = m;
            }
        }
        __syncthreads();
        if (sm.maxcnt > (uint32_t)BUCKET_GUARD) return false;
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                if (p < n) {
                    const uint32_t at = cnt[(uint32_t)(k[j] >> shift) & bm] + rk[j];
                    if constexpr (sizeof(U) == 4) {
                        reinterpret_cast<uint64_t*>(sm.keys)[at] = (uint64_t(U(k[j] & rm)) << 32) | uint64_t(p);
                    } else {
                        sm.keys[at] = U(k[j] & rm);
                        sm.idx[0][at] = (uint16_t)p;
                    }
                }
            }
        }
        __syncthreads();
        prefetch();
        const bool full = rm == Bits<U>::ALL;
        const int nbk = 1 << nb;
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int q0 = (lb + j * WAVE);
                if (q0 < n) {
                    U mk;
                    uint32_t p, c = 0;
                    int s0, e0;
                    if constexpr (sizeof(U) == 4) {
                        const uint64_t* wd = reinterpret_cast<const uint64_t*>(sm.keys);
                        const uint64_t mine = wd[q0];
                        mk = U(mine >> 32);
                        p = (uint32_t)mine;
                        const uint32_t d = (uint32_t)(mk >> shift) & bm;
                        s0 = (int)cnt[d];
                        e0 = (int)d + 1 < nbk ? (int)cnt[d + 1] : n;
                        for (int q = s0; q < e0; ++q) c += wd[q] < mine;
                    } else {
                        mk = sm.keys[q0];
                        p = sm.idx[0][q0];
                        const uint32_t d = (uint32_t)(mk >> shift) & bm;
                        s0 = (int)cnt[d];
                        e0 = (int)d + 1 < nbk ? (int)cnt[d + 1] : n;
                        for (int q = s0; q < e0; ++q) {
                            const U kq = sm.keys[q];
                            c += (kq < mk) | ((kq == mk) & ((uint32_t)sm.idx[0][q] < p));
                        }
                    }
                    const int r = s0 + (int)c;
                    // loads before stores: the memory counter completes in order, so a
                    // load issued after the key store would also wait for that store
                    V val{};
                    if constexpr (HAS_V) val = at32(vs + begin, p);
                    const U kout = full ? twiddle_out<U>(mk, km.kind, km.desc) : at32(ks + begin, p);
                    at32(ko + begin, r) = kout;
                    if constexpr (HAS_V) at32(vo + begin, r) = val;
                }
            }
        }
        __syncthreads();
        return true;
    }

    // Whether runs of equal top 16 varying bits are likely short enough for the
    // top-16 finish. Among the range's first 64 keys (a sample of unsorted
    // input), the pairs with equal top 12 bits estimate sum(p^2) over the
    // 12-bit values, and 1/16 of that (the next 4 bits spread inside a 12-bit
    // value) the same for 16 bits; n times it is the expected number of items
    // that share an item's top 16 bits. The finish is taken while that stays
    // at most 2: pairs * n <= 2 * 16 * (64 * 63 / 2). Uniform keys see 0.5
    // sample pairs on average (a 16K tile goes elsewhere with probability
    // about 0.2%); skewed keys, such as floats of a normal distribution (whose
    // top 16 bits are the sign, the exponent and 7 mantissa bits), see tens
    // and keep the full LSD instead of overrunning the finish's guard. Every
    // wave reads the same sample, so the answer is uniform over the tile.
    __device__ static __forceinline__ bool top2_promising(Smem& sm, int n, const KeyMode& km, int hi) {
        constexpr int SAMPLE_BITS = 12;
        const U rm = range_mask<U>(km.begin_bit, km.end_bit);
        const uint32_t l = threadIdx.x & 63;
        const U kk = sm.keys[l];
        const uint32_t t = (uint32_t)(U((kk & rm) >> (hi - SAMPLE_BITS)) & U((1u << SAMPLE_BITS) - 1));
        const uint64_t peers = match_digit<SAMPLE_BITS>(t);
        const uint32_t pairs = wave_sum((uint32_t)__popcll(peers) - 1u) / 2u;  // every pair counted twice
        return (uint64_t)pairs * (uint64_t)n <= 2u * 16u * (64u * 63u / 2u);
    }

    // Two stable passes on the varying bits [hi - 16, hi) order the tile by its
    // top 16 varying bits; items with equal top bits then form short runs, and
    // each item's final place is its run's start plus the number of run items
    // with a smaller key or an equal key at an earlier position (the passes
    // are stable). Returns false, with nothing written to global memory and
    // the prefetch hook not run, when some run is longer than TOP_GUARD items.
    template <typename Prefetch>
    __device__ static __forceinline__ bool top2_finish(Smem& sm, U (&k)[IPT], const V* vs, U* ko, V* vo, int64_t begin, int n,
                                       const KeyMode& km, int hi, Prefetch& prefetch) {
        const int lb = lane_base(n);
        const bool act = wave() < active_waves(n);
        const int sh = hi - 2 * RADIX_BITS;
        uint32_t dst[IPT];
        rank_digits(sm, k, n, sh, RADIX_BITS, dst);
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                if (p < n) {
                    sm.keys[dst[j]] = k[j];
                    if constexpr (HAS_V) sm.idx[1][dst[j]] = (uint16_t)p;
                }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                k[j] = p < n ? sm.keys[p] : Bits<U>::ALL;
            }
        }
        rank_digits(sm, k, n, sh + RADIX_BITS, RADIX_BITS, dst);
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                if (p < n) {
                    sm.keys[dst[j]] = k[j];
                    if constexpr (HAS_V) sm.idx[0][dst[j]] = sm.idx[1][p];
                }
            }
        }
        __syncthreads();
        // the order covers only the key bits [begin_bit, end_bit); above hi
        // those bits are the same in every item
        const U rm = range_mask<U>(km.begin_bit, km.end_bit);
        auto top = [&](U x) { return U((x & rm) >> sh); };
        // bit j of `in_run`: the item in slot j shares its top bits with a
        // neighbour; most items do not, and their place is their slot. The
        // neighbours' top bits come from the adjacent lanes (slot j's lane 63
        // is followed by slot j + 1's lane 0), and from LDS only across waves.
        const uint32_t l = threadIdx.x & 63;
        uint16_t* const place = sm.idx[HAS_V ? 1 : 0];  // free after the passes
        uint32_t in_run = 0;
        if (act) {
            U t[IPT];
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                t[j] = top(sm.keys[p < n ? p : n - 1]);
            }
            const int w0 = lb - (int)l;  // first slot of the wave
            const U before_wave = top(sm.keys[w0 > 0 ? w0 - 1 : 0]);
            const int w1 = w0 + IPT * WAVE;  // first slot after the wave
            const U after_wave = top(sm.keys[w1 < n ? w1 : n - 1]);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                const U prev_last = j > 0 ? read_lane<U>(t[j > 0 ? j - 1 : 0], 63) : before_wave;
                const U next_first = j + 1 < IPT ? read_lane<U>(t[j + 1 < IPT ? j + 1 : j], 0) : after_wave;
                const U tl = lane_shift<DPP_WAVE_SHR1>(t[j], prev_last);
                const U tr = lane_shift<DPP_WAVE_SHL1>(t[j], next_first);
                const bool valid = p < n;
                in_run |= uint32_t(valid & (((p > 0) & (tl == t[j])) | ((p + 1 < n) & (tr == t[j])))) << j;
            }
        }
        // Items in runs are queued per wave (in the wave's digit-count words,
        // free now: QCAP 16-bit entries) and ranked one per lane; the place
        // of slot p goes to place[p]. A run longer than TOP_GUARD stops the
        // finish before anything is written to global memory. LDS operations
        // of one wave complete in order, so the queue needs no barrier.
        int long_run = 0;
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                if (p < n && !((in_run >> j) & 1u)) place[p] = (uint16_t)p;
            }
            constexpr uint32_t QCAP = 2 * RADIX;
            uint16_t* queue = reinterpret_cast<uint16_t*>(sm.wcnt + wave() * RADIX);
            // which lanes queue slot j: wave-uniform masks and counts
            uint64_t mq[IPT];
            uint32_t total = 0;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                mq[j] = __ballot((lb + j * WAVE) < n && ((in_run >> j) & 1u));
                total += (uint32_t)__popcll(mq[j]);
            }
            auto rank_queue = [&](uint32_t cnt) {
                for (uint32_t e0 = 0; e0 < cnt; e0 += WAVE) {
                    if (e0 + l < cnt) {
                        const int p = queue[e0 + l];
                        const U kk = sm.keys[p];
                        const U km_ = U(kk & rm);
                        const U tk = top(kk);
                        // the run (contiguous, holds p) seen through a window of
                        // +-3 slots; runs reaching the window's edge are walked
                        constexpr int WIN = 3;
                        uint32_t before = 0, c = 0;
                        bool edge = false;
#pragma unroll
                        for (int d = -WIN; d <= WIN; ++d) {
                            if (d == 0) continue;
                            const int q = p + d;
                            const bool in = q >= 0 && q < n;
                            const U kq = sm.keys[in ? q : p];
                            const U kqm = U(kq & rm);
                            const bool same = in & (top(kq) == tk);
                            before += uint32_t(same & (d < 0));
                            c += uint32_t(same & ((kqm < km_) | ((kqm == km_) & (d < 0))));
                            if (d == -WIN || d == WIN) edge |= same;
                        }
                        uint32_t r = (uint32_t)p - before + c;
                        if (edge) {
                            // bounded walks: a run longer than TOP_GUARD stops them
                            int s = p;
                            while (s > 0 && top(sm.keys[s - 1]) == tk) {
                                --s;
                                if (p - s > TOP_GUARD) break;
                            }
                            int e = p + 1;
                            while (e < n && top(sm.keys[e]) == tk) {
                                ++e;
                                if (e - s > TOP_GUARD) break;
                            }
                            if (e - s > TOP_GUARD) {
                                long_run = 1;
                                r = (uint32_t)p;
                            } else {
                                c = 0;
                                for (int q = s; q < e; ++q) {
                                    const U kq = U(sm.keys[q] & rm);
                                    c += (kq < km_) | ((kq == km_) & (q < p));
                                }
                                r = (uint32_t)s + c;
                            }
                        }
                        place[p] = (uint16_t)r;
                    }
                }
            };
            auto push = [&](int j, uint32_t& at) {
                if ((mq[j] >> l) & 1u) queue[at + lanes_below(mq[j])] = (uint16_t)(lb + j * WAVE);
                at += (uint32_t)__popcll(mq[j]);
            };
            if (total <= QCAP) {
                uint32_t at = 0;
#pragma unroll
                for (int j = 0; j < IPT; ++j) push(j, at);
                rank_queue(total);
            } else {
                // more queued items than entries: QCAP / 64 slots at a time
                constexpr int QS = int(QCAP / WAVE);
#pragma unroll
                for (int h = 0; h < IPT; h += QS) {
                    uint32_t at = 0;
#pragma unroll
                    for (int j = h; j < h + QS && j < IPT; ++j) push(j, at);
                    rank_queue(at);
                }
            }
        }
        if (__syncthreads_or(long_run)) return false;
        prefetch();
        constexpr bool STAGE = HAS_V && sizeof(V) <= sizeof(U);
        V v[IPT];
        if (act && STAGE) {
            // the values' global read overlaps the keys' write-out
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                v[j] = at32(vs + begin, (p < n ? p : n - 1));  // unconditional: the loads issue together
            }
        }
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                const int pc = p < n ? p : n - 1;
                const U kk = sm.keys[pc];
                const uint32_t r = place[pc];
                if (p < n) at32(ko + begin, r) = twiddle_out<U>(kk, km.kind, km.desc);
            }
        }
        if constexpr (STAGE) {
            // values: staged in the key area of LDS (free once every run is
            // ranked), then moved to their final places
            __syncthreads();
            V* stage = reinterpret_cast<V*>(sm.keys);
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = lb + j * WAVE;
                    if (p < n) stage[p] = v[j];
                }
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = lb + j * WAVE;
                    if (p < n) at32(vo + begin, place[p]) = stage[sm.idx[0][p]];
                }
            }
        } else if constexpr (HAS_V) {
            __syncthreads();
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = lb + j * WAVE;
                    if (p < n) at32(vo + begin, place[p]) = at32(vs + begin, sm.idx[0][p]);
                }
            }
        }
        __syncthreads();
        return true;
    }

    template <typename Prefetch>
    __device__ static __forceinline__ void sort_loaded(Smem& sm, const U (&raw)[IPT], const U* ks, const V* vs, U* ko, V* vo,
                                       int64_t begin, int n, const KeyMode& km, Prefetch prefetch) {
        const int lb = lane_base(n);
        const bool act = wave() < active_waves(n);
        U k[IPT];
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) k[j] = (lb + j * WAVE) < n ? twiddle_in<U>(raw[j], km.kind, km.desc) : Bits<U>::ALL;
        }
        // the range's first 64 keys, read back by every wave after the barrier
        // in reduce_or_and (the key area is not used before the first pass)
        if constexpr (TOP2_FINISH || (BUCKET_FINISH && !SMALL_BUCKET)) {
            if (wave() == 0) sm.keys[threadIdx.x] = k[0];
        }
        U o, a;
        reduce_or_and(sm, k, n, o, a);
        const U varying = U((o ^ a) & range_mask<U>(km.begin_bit, km.end_bit));
        if (varying == 0) {
            prefetch();
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    if (p < n) {
                        at32(ko + begin, p) = twiddle_out<U>(k[j], km.kind, km.desc);
                        if constexpr (HAS_V) at32(vo + begin, p) = at32(vs + begin, p);
                    }
                }
            }
            __syncthreads();
            return;
        }
        const int lo = lowest_bit<U>(varying);
        const int hi = bit_length<U>(varying);
        if constexpr (SMALL_BUCKET) {
            // One pass places every item in its bucket of the top 8 varying bits
            // (LDS atomics, unordered inside a bucket) as a sort word (masked
            // order key, original position); when no bucket holds more than
            // SMALL_GUARD items, the item at each slot is ranked inside its bucket by
            // comparing sort words -- the count of smaller words is its stable
            // rank -- and written to its final place, instead of three more passes.
            if (hi - lo > RADIX_BITS) {
                static_assert(NW <= 4, "one wave scans the bucket counts");
                const int shift = hi - RADIX_BITS;
                const U rm = range_mask<U>(km.begin_bit, km.end_bit);
                uint32_t* cnt = sm.wcnt;  // RADIX counters (any wave's count area is free here)
                const int t = threadIdx.x;
                for (int d = t; d < RADIX; d += NT) cnt[d] = 0;
                __syncthreads();
                uint32_t rk[IPT];
                if (act) {
#pragma unroll
                    for (int j = 0; j < IPT; ++j) {
                        const int p = (lb + j * WAVE);
                        if (p < n) rk[j] = atomicAdd(&cnt[(uint32_t)(k[j] >> shift) & (RADIX - 1)], 1u);
                    }
                }
                __syncthreads();
                if (wave() == 0) {
                    const uint32_t l = t & 63;
                    uint32_t c[RADIX / WAVE];
#pragma unroll
                    for (int i = 0; i < RADIX / WAVE; ++i) c[i] = cnt[i * WAVE + (int)l];
                    uint32_t carry = 0, big = 0;
#pragma unroll
                    for (int i = 0; i < RADIX / WAVE; ++i) {
                        const uint32_t incl = wave_incl_scan(c[i]);
                        sm.dstart[i * WAVE + (int)l] = carry + incl - c[i];
                        carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                        big = c[i] > big ? c[i] : big;
                    }
                    big = wave_max(big);
                    if (l == 0) {
                        sm.dstart[RADIX] = (uint32_t)n;
                        sm.maxcnt = big;
                    }
                }
                __syncthreads();
                if (sm.maxcnt <= (uint32_t)SMALL_GUARD) {
                    if (act) {
#pragma unroll
                        for (int j = 0; j < IPT; ++j) {
                            const int p = (lb + j * WAVE);
                            const uint32_t d = (uint32_t)(k[j] >> shift) & (RADIX - 1);
                            if (p < n) {
                                const uint32_t at = sm.dstart[d] + rk[j];
                                if constexpr (sizeof(U) == 4) {
                                    reinterpret_cast<uint64_t*>(sm.keys)[at] =
                                        (uint64_t(U(k[j] & rm)) << 32) | uint64_t(p);
                                } else {
                                    sm.keys[at] = U(k[j] & rm);
                                    sm.idx[0][at] = (uint16_t)p;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    prefetch();
                    // each thread finishes the items at its slots of the bucket
                    // order (neighbouring lanes share buckets: coherent loops and
                    // broadcast reads); the sort word names the item's input position
                    const bool full = rm == Bits<U>::ALL;
                    if (act) {
                        // every value gather issued before any store (the memory
                        // counter completes in order); the positions are read
                        // from the sort words again below
                        V vg[IPT];
                        if constexpr (HAS_V && sizeof(U) == 4) {
                            const uint64_t* w = reinterpret_cast<const uint64_t*>(sm.keys);
#pragma unroll
                            for (int j = 0; j < IPT; ++j) {
                                const int q0 = (lb + j * WAVE);
                                vg[j] = at32(vs + begin, q0 < n ? (uint32_t)w[q0] : 0u);
                            }
                        }
#pragma unroll
                        for (int j = 0; j < IPT; ++j) {
                            const int q0 = (lb + j * WAVE);
                            if (q0 < n) {
                                U mk;
                                uint32_t p;
                                uint32_t c = 0;
                                int s0;
                                if constexpr (sizeof(U) == 4) {
                                    const uint64_t* w = reinterpret_cast<const uint64_t*>(sm.keys);
                                    const uint64_t mine = w[q0];
                                    mk = U(mine >> 32);
                                    p = (uint32_t)mine;
                                    const uint32_t d = (uint32_t)(mk >> shift) & (RADIX - 1);
                                    s0 = (int)sm.dstart[d];
                                    const int e0 = (int)sm.dstart[d + 1];
                                    for (int q = s0; q < e0; ++q) c += w[q] < mine;
                                } else {
                                    mk = sm.keys[q0];
                                    p = sm.idx[0][q0];
                                    const uint32_t d = (uint32_t)(mk >> shift) & (RADIX - 1);
                                    s0 = (int)sm.dstart[d];
                                    const int e0 = (int)sm.dstart[d + 1];
                                    for (int q = s0; q < e0; ++q) {
                                        const U kq = sm.keys[q];
                                        c += (kq < mk) | ((kq == mk) & ((uint32_t)sm.idx[0][q] < p));
                                    }
                                }
                                const int r = s0 + (int)c;
                                V val{};  // loads before stores (in-order memory counter)
                                if constexpr (HAS_V && sizeof(U) == 4) val = vg[j];
                                else if constexpr (HAS_V) val = at32(vs + begin, p);
                                const U kout =
                                    full ? twiddle_out<U>(mk, km.kind, km.desc) : at32(ks + begin, p);
                                at32(ko + begin, r) = kout;
                                if constexpr (HAS_V) at32(vo + begin, r) = val;
                            }
                        }
                    }
                    __syncthreads();
                    return;
                }
            }
        }
        if constexpr (BUCKET_FINISH && !SMALL_BUCKET) {
            if (hi - lo > RADIX_BITS && bucket_promising(sm, n, km, hi, lo)) {
                if (bucket_finish(sm, k, ks, vs, ko, vo, begin, n, km, hi, lo, prefetch)) return;
            }
        }
        if constexpr (TOP2_FINISH) {
            if (hi - lo > 2 * RADIX_BITS && top2_promising(sm, n, km, hi)) {
                if (top2_finish(sm, k, vs, ko, vo, begin, n, km, hi, prefetch)) return;
                // a run of equal top bits is too long: the full LSD below, from
                // the original keys (read again; the registers hold the passes')
                if (act) load_keys(ks, begin, n, km, k);
            }
        }
        int cur = 0;  // idx buffer holding the source index of each tile position
        for (int shift = lo; shift < hi; shift += RADIX_BITS) {
            const int bits = (hi - shift) < RADIX_BITS ? (hi - shift) : RADIX_BITS;
            uint32_t dst[IPT];
            rank_digits(sm, k, n, shift, bits, dst);
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    if (p < n) {
                        sm.keys[dst[j]] = k[j];
                        if constexpr (HAS_V) sm.idx[cur ^ 1][dst[j]] = shift == lo ? (uint16_t)p : sm.idx[cur][p];
                    }
                }
            }
            cur ^= 1;
            __syncthreads();
            if (act && shift + RADIX_BITS < hi) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    k[j] = p < n ? sm.keys[p] : Bits<U>::ALL;  // slots past n stay last
                }
            }
        }
        prefetch();
        if constexpr (HAS_V && sizeof(V) <= sizeof(U)) {
            // Values: one coalesced read of the range into the key area of LDS
            // (free once the sorted keys are out), then the permutation is
            // applied from LDS. A gather straight from global memory would pull
            // a whole cache line for every value.
            V v[IPT];
            uint32_t srcs[IPT];
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    v[j] = at32(vs + begin, (p < n ? p : n - 1));  // unconditional: the loads issue together
                }
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    if (p < n) {
                        srcs[j] = sm.idx[cur][p];
                        at32(ko + begin, p) = twiddle_out<U>(sm.keys[p], km.kind, km.desc);
                    }
                }
            }
            __syncthreads();
            V* stage = reinterpret_cast<V*>(sm.keys);
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j)
                    if ((lb + j * WAVE) < n) stage[(lb + j * WAVE)] = v[j];
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    if (p < n) at32(vo + begin, p) = stage[srcs[j]];
                }
            }
        } else {
            if (act) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    if (p < n) {
                        at32(ko + begin, p) = twiddle_out<U>(sm.keys[p], km.kind, km.desc);
                        if constexpr (HAS_V) at32(vo + begin, p) = at32(vs + begin, sm.idx[cur][p]);
                    }
                }
            }
        }
        __syncthreads();
    }

    // Stable partition by digit (twiddled >> shift) & (2^bits - 1) of n <= TILE
    // items: the item with digit d and rank r among the tile's digit-d items
    // goes to global index goff[d] + r. The tile is ordered by digit in LDS
    // first, so each digit run is written contiguously. `out(g, key, value)`
    // stores one item: g is its global index, key the (twiddled) key as it
    // sits in LDS.
    template <typename Out>
    __device__ static void partition_with(Smem& sm, const U* ks, const V* vs, int64_t begin, int n,
                                          const KeyMode& km, int shift, int bits, const uint32_t* goff, Out out) {
        const int lb = lane_base(n);
        const bool act = wave() < active_waves(n);
        U k[IPT];
        V v[IPT];
        if (act) {
            load_keys(ks, begin, n, km, k);
            if constexpr (HAS_V) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = (lb + j * WAVE);
                    v[j] = at32(vs + begin, (p < n ? p : n - 1));
                }
            }
        }
        partition_loaded(sm, k, v, n, shift, bits, goff, out);
    }

    // The partition of n <= TILE items whose twiddled keys (all-ones past n)
    // and values the threads already hold, in the slot layout of load_keys.
    template <typename Out>
    __device__ static void partition_loaded(Smem& sm, const U (&k)[IPT], const V (&v)[IPT], int n, int shift,
                                            int bits, const uint32_t* goff, Out out) {
        const int lb = lane_base(n);
        const bool act = wave() < active_waves(n);
        uint32_t dst[IPT];
        rank_digits(sm, k, n, shift, bits, dst);  // contains barriers
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j)
                if ((lb + j * WAVE) < n) {
                    sm.keys[dst[j]] = k[j];
                    if constexpr (HAS_V) sm.vals[dst[j]] = v[j];
                }
        }
        __syncthreads();
        const U dmask = U((U(1) << bits) - 1);
        if (act) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = (lb + j * WAVE);
                if (p < n) {
                    const U key = sm.keys[p];
                    const uint32_t d = (uint32_t)((key >> shift) & dmask);
                    const uint32_t g = goff[d] + ((uint32_t)p - sm.dstart[d]);
                    if constexpr (HAS_V) out(g, key, sm.vals[p]);
                    else out(g, key, V(0));
                }
            }
        }
        __syncthreads();
    }

    // Stable partition of the len items at ks[begin..) (values alongside) into
    // kd/vd, tile by tile, with running destinations run[d] in LDS that advance
    // by each tile's count of digit d (the workgroup keeps them across the
    // tiles of one range). Software pipelined: the next tile's keys are loaded
    // into registers before this tile is ranked, and this tile's values right
    // before that, so that both are in flight during the ranking (barriers wait
    // for LDS only). Every thread loads every slot of a tile (slots past its
    // end read its last item again), so each thread's loads keep one shape.
    template <bool PIPELINE = true>
    __device__ static void partition_range(Smem& sm, const U* ks, const V* vs, U* kd, V* vd, int64_t begin, int len,
                                           const KeyMode& km, int shift, int bits, uint32_t* run) {
        if (len <= 0) return;
        U rk[IPT];
        auto fetch = [&](int64_t b0, int n) {
            const int lb = lane_base(n);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                rk[j] = at32(ks + b0, (p < n ? p : n - 1));
            }
        };
        // (no prefetch where a second set of keys does not fit the registers:
        // 16 keys of 64 bits per thread)
        constexpr bool PREFETCH = PIPELINE && sizeof(U) * IPT <= 64;
        if (PREFETCH) fetch(begin, len < TILE ? len : TILE);
        for (int t0 = 0; t0 < len; t0 += TILE) {
            const int n = (len - t0) < TILE ? (len - t0) : TILE;
            if (!PREFETCH) fetch(begin + t0, n);
            U k[IPT];
            V v[IPT];
            const int lb = lane_base(n);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = lb + j * WAVE;
                if constexpr (HAS_V) v[j] = at32(vs + begin + t0, (p < n ? p : n - 1));
                k[j] = p < n ? twiddle_in<U>(rk[j], km.kind, km.desc) : Bits<U>::ALL;
            }
            const int nn = len - t0 - TILE;
            if (PREFETCH && nn > 0) fetch(begin + t0 + TILE, nn < TILE ? nn : TILE);
            partition_loaded(sm, k, v, n, shift, bits, run, [&](uint32_t g, U key, V val) {
                kd[g] = twiddle_out<U>(key, km.kind, km.desc);
                if constexpr (HAS_V) vd[g] = val;
            });
            for (int d = threadIdx.x; d < RADIX; d += NT) {
                const uint32_t end = d + 1 < RADIX ? sm.dstart[d + 1] : (uint32_t)n;
                run[d] += end - sm.dstart[d];
            }
            __syncthreads();
        }
    }

    // Digit histograms of the n <= TILE items at ks[begin..) for `npos` digit
    // positions pos_p = pos0 + 8p (the last may be narrower than 8 bits, up to
    // hi_bit); counts are accumulated into the LDS array hist[npos][RADIX].
    // Keys are read a tile at a time, all IPT loads of a thread issued before
    // any is counted. A digit that is the same in every active lane (a
    // constant high digit of small or duplicate-heavy keys) is added once by
    // one lane instead of by every lane into the same LDS word.
    __device__ static void histogram(uint32_t* hist, const U* ks, int64_t begin, int n, const KeyMode& km,
                                     int pos0, int npos, int hi_bit) {
        const int t = threadIdx.x;
        for (int c0 = 0; c0 < n; c0 += TILE) {
            U key[IPT];
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = c0 + j * NT + t;
                key[j] = i < n ? at32(ks + begin + c0, j * NT + t) : U(0);  // offsets inside one tile
            }
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = c0 + j * NT + t;
                if (c0 + j * NT + (t & ~63) >= n) break;  // this and later rows: whole wave past the end
                const bool v = i < n;
                const U k = twiddle_in<U>(key[j], km.kind, km.desc);
                const uint64_t act = __ballot(v);
                for (int p = 0; p < npos; ++p) {
                    const int sh = pos0 + p * RADIX_BITS;
                    const int nb = (hi_bit - sh) < RADIX_BITS ? (hi_bit - sh) : RADIX_BITS;
                    const uint32_t d = (uint32_t)((k >> sh) & U((U(1) << nb) - 1));
                    const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
                    if (__ballot(v && d != d0) == 0) {
                        if ((t & 63) == 0) atomicAdd(&hist[p * RADIX + d0], (uint32_t)__popcll(act));
                    } else if (v) {
                        atomicAdd(&hist[p * RADIX + d], 1u);
                    }
                }
            }
        }
    }

    // Digit histogram of the n items at ks[begin..) at one position (shift sh,
    // digit mask dm) of their order keys (twiddled, masked by rmask), counted
    // into the LDS array hist[RADIX], with this thread's OR and AND of those
    // order keys accumulated into o and a (the caller reduces them over the
    // workgroup). One LDS atomic per item, or one per wave when the wave's
    // digits are all equal.
    __device__ static void histogram_or_and(uint32_t* hist, const U* ks, int64_t begin, int n, const KeyMode& km,
                                            U rmask, int sh, U dm, U& o, U& a) {
        const int t = threadIdx.x;
        for (int c0 = 0; c0 < n; c0 += TILE) {
            U key[IPT];
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = c0 + j * NT + t;
                key[j] = i < n ? at32(ks + begin + c0, j * NT + t) : U(0);  // offsets inside one tile
            }
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = c0 + j * NT + t;
                if (c0 + j * NT + (t & ~63) >= n) break;  // this and later rows: whole wave past the end
                const bool v = i < n;
                const U k = U(twiddle_in<U>(key[j], km.kind, km.desc) & rmask);
                if (v) {
                    o |= k;
                    a &= k;
                }
                const uint64_t act = __ballot(v);
                const uint32_t d = (uint32_t)((k >> sh) & dm);
                const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
                if (__ballot(v && d != d0) == 0) {
                    if ((t & 63) == 0) atomicAdd(&hist[d0], (uint32_t)__popcll(act));
                } else if (v) {
                    atomicAdd(&hist[d], 1u);
                }
            }
        }
    }
};

// ---- wave sorts of short ranges ------------------------------------------------

// A sort word: the masked order key with the item's position in the range as
// tie-break, so that all words differ and any sorting network sorts stably.
// 32-bit keys pack into one u64; 64-bit keys use a (key, position) pair.
template <typename U>
struct SortWord;

template <>
struct SortWord<uint32_t> {
    uint64_t w;
    __device__ static __forceinline__ SortWord make(uint32_t okey, uint32_t p) {
        return {(uint64_t(okey) << 32) | p};
    }
    __device__ static __forceinline__ SortWord inf() { return {~0ull}; }
    __device__ __forceinline__ bool lt(const SortWord& o) const { return w < o.w; }
    __device__ __forceinline__ uint32_t pos() const { return (uint32_t)w; }
    __device__ __forceinline__ uint32_t key() const { return (uint32_t)(w >> 32); }
    template <typename F>
    __device__ __forceinline__ SortWord map(F f) const {
        return {(uint64_t(f((uint32_t)(w >> 32))) << 32) | uint64_t(f((uint32_t)w))};
    }
    __device__ static __forceinline__ SortWord select(bool c, const SortWord& a, const SortWord& b) {
        return {c ? a.w : b.w};
    }
};

template <>
struct SortWord<uint64_t> {
    uint64_t k;
    uint32_t p;
    __device__ static __forceinline__ SortWord make(uint64_t okey, uint32_t pp) { return {okey, pp}; }
    __device__ static __forceinline__ SortWord inf() { return {~0ull, ~0u}; }
    __device__ __forceinline__ bool lt(const SortWord& o) const { return k < o.k || (k == o.k && p < o.p); }
    __device__ __forceinline__ uint32_t pos() const { return p; }
    __device__ __forceinline__ uint64_t key() const { return k; }
    template <typename F>
    __device__ __forceinline__ SortWord map(F f) const {
        return {(uint64_t(f((uint32_t)(k >> 32))) << 32) | uint64_t(f((uint32_t)k)), f(p)};
    }
    // field by field: a select of the whole struct goes through scratch memory
    __device__ static __forceinline__ SortWord select(bool c, const SortWord& a, const SortWord& b) {
        return {c ? a.k : b.k, c ? a.p : b.p};
    }
};

// Value of lane (lane ^ M) for a compile-time M: DPP quad permutes for 1 and
// 2, ds_swizzle (bitmask mode) for 4, 8 and 16, ds_bpermute for 32.
template <int M>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (M < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (M << 10));
    else return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane_id() ^ 32u) << 2), (int)v);
}

template <typename W>
__device__ __forceinline__ void compare_exchange(W& a, W& b, bool ascending) {
    const bool swap = ascending ? b.lt(a) : a.lt(b);
    const W x = W::select(swap, b, a);
    const W y = W::select(swap, a, b);
    a = x;
    b = y;
}

// One cross-lane stage of the bitonic network at lane distance M (element
// distance M*R): the lower partner keeps the smaller word when the run is
// ascending, the upper partner the larger one.
template <int R, int M, typename W>
__device__ __forceinline__ void bitonic_cross(W (&w)[R], int k) {
    const uint32_t l = lane_id();
    const bool lower = (l & M) == 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const W o = w[r].map([](uint32_t x) { return xor_lane<M>(x); });
        const bool asc = (((int)l * R + r) & k) == 0;
        const bool keep_small = lower == asc;
        const bool o_smaller = o.lt(w[r]);
        w[r] = W::select(keep_small == o_smaller, o, w[r]);
    }
}

// Bitonic sort of 64*R words held R per lane; afterwards lane l holds the
// words of ranks l*R .. l*R+R-1 in order.
template <int R, typename W>
__device__ __forceinline__ void bitonic_wave(W (&w)[R]) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int k = 2; k <= WAVE * R; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < R) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if ((r & j) == 0) {
                        const bool asc = (((int)l * R + r) & k) == 0;
                        compare_exchange(w[r], w[r | j], asc);
                    }
                }
            } else {
                const int m = j / R;
                if (m == 1) bitonic_cross<R, 1>(w, k);
                else if (m == 2) bitonic_cross<R, 2>(w, k);
                else if (m == 4) bitonic_cross<R, 4>(w, k);
                else if (m == 8) bitonic_cross<R, 8>(w, k);
                else if (m == 16) bitonic_cross<R, 16>(w, k);
                else bitonic_cross<R, 32>(w, k);
            }
        }
    }
}

// Bitonic sort of the G words held one per lane by each aligned group of G
// lanes (G <= 32): every exchange distance is below G, so groups never mix,
// and the direction bit is taken from the lane's index inside its group.
template <int G, typename W>
__device__ __forceinline__ void bitonic_group(W& w) {
    const uint32_t i = lane_id() & (G - 1);
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            W o;
            if (j == 1) o = w.map([](uint32_t x) { return xor_lane<1>(x); });
            else if (j == 2) o = w.map([](uint32_t x) { return xor_lane<2>(x); });
            else if (j == 4) o = w.map([](uint32_t x) { return xor_lane<4>(x); });
            else if (j == 8) o = w.map([](uint32_t x) { return xor_lane<8>(x); });
            else o = w.map([](uint32_t x) { return xor_lane<16>(x); });
            const bool asc = (i & (uint32_t)k) == 0;
            const bool lower = (i & (uint32_t)j) == 0;
            const bool keep_small = lower == asc;
            w = W::select(keep_small == o.lt(w), o, w);
        }
    }
}

// The range of one group of G lanes and, in lane i of the group, item i's key
// and value (i < n), as loaded ahead of the group sort.
template <typename U, typename V>
struct GroupItems {
    int64_t begin;
    int n;
    int src;
    U key;
    V val;
};

template <int G, typename U, typename V, bool HAS_V>
__device__ __forceinline__ void group_load(const Bufs& b, const TileItem& it, bool valid, GroupItems<U, V>& gi) {
    const int i = (int)(lane_id() & (G - 1));
    gi.begin = it.begin;
    gi.n = valid ? ti_n(it) : 0;
    gi.src = ti_src(it);
    if (i < gi.n) {
        gi.key = ((const U*)b.k[gi.src])[gi.begin + i];
        if constexpr (HAS_V) gi.val = ((const V*)b.v[gi.src])[gi.begin + i];
    }
}

// Value of `x` in lane `lane` (all lanes take part).
template <typename T>
__device__ __forceinline__ T lane_fetch(T x, int lane) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = (uint64_t)x;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(lane << 2, (int)(uint32_t)u);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(lane << 2, (int)(uint32_t)(u >> 32));
        return (T)((uint64_t(hi) << 32) | lo);
    } else {
        return (T)(uint32_t)__builtin_amdgcn_ds_bpermute(lane << 2, (int)(uint32_t)x);
    }
}

// Each aligned group of G lanes sorts its own range (n <= G; n == 0 for a
// group without a range): the group's network sorts (order key, i) words and
// lane i writes the i-th smallest, taking its key and value from the lane
// that loaded them (cross-lane permute, no second read of global memory).
template <int G, typename U, typename V, bool HAS_V>
__device__ __forceinline__ void group_sort_loaded(const Bufs& b, const GroupItems<U, V>& gi, const KeyMode& km) {
    using W = SortWord<U>;
    const int lane = (int)lane_id();
    const int i = lane & (G - 1);
    const U rm = range_mask<U>(km.begin_bit, km.end_bit);
    W w = W::inf();
    if (i < gi.n) w = W::make(U(twiddle_in<U>(gi.key, km.kind, km.desc) & rm), (uint32_t)i);
    bitonic_group<G>(w);
    const int from = (lane & ~(G - 1)) + (int)(w.pos() & (G - 1));
    const U key = lane_fetch<U>(gi.key, from);
    V val{};
    if constexpr (HAS_V) val = lane_fetch<V>(gi.val, from);
    if (i < gi.n) {
        buf_ptr<U>(b.k[BUF_OUT])[gi.begin + i] = key;
        if constexpr (HAS_V) buf_ptr<V>(b.v[BUF_OUT])[gi.begin + i] = val;
    }
}

// One wave sorts a range of n <= 64*R items: sort words are loaded with
// coalesced reads (slot (lane, r) takes item r*64 + lane), sorted by the
// network, and written in rank order; keys are rebuilt from the words when
// the whole key takes part in the order, else read back by position, and
// values are read by position.
// Values are read with the keys and staged in the wave's LDS region `stage`
// (64*R entries) after the network, so the permutation is applied from LDS
// (LDS operations of one wave complete in order: no barrier is needed).
template <int R, typename U, typename V, bool HAS_V>
__device__ __forceinline__ void wave_bitonic_sort(const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n,
                                                  const KeyMode& km, V* stage) {
    using W = SortWord<U>;
    const uint32_t l = lane_id();
    const U rm = range_mask<U>(km.begin_bit, km.end_bit);
    W w[R];
    V v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int p = r * WAVE + (int)l;
        const int pc = p < n ? p : n - 1;
        const U raw = at32(ks + begin, pc);
        if constexpr (HAS_V) v[r] = at32(vs + begin, pc);
        w[r] = p < n ? W::make(U(twiddle_in<U>(raw, km.kind, km.desc) & rm), (uint32_t)p) : W::inf();
    }
    bitonic_wave<R>(w);
    if constexpr (HAS_V) {
#pragma unroll
        for (int r = 0; r < R; ++r) stage[r * WAVE + (int)l] = v[r];
    }
    const bool full = rm == Bits<U>::ALL;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int q = (int)l * R + r;
        if (q < n) {
            const uint32_t src = w[r].pos();
            at32(ko + begin, q) = full ? twiddle_out<U>(U(w[r].key()), km.kind, km.desc) : at32(ks + begin, src);
            if constexpr (HAS_V) at32(vo + begin, q) = stage[src];
        }
    }
}

// ---- one wave, one counting pass and a compare-count inside the buckets ----------

constexpr int WAVE_SORT_MAX = 512;  // items of one wave-tier range
// Wave bucket finish: lane l ranks the items at slots r*64 + l of the bucket
// order (neighbouring lanes share buckets: coherent loops, broadcast reads; the
// keys and values are then gathered by input position) when the range has more
// than two items per lane, else its own input items (keys and values straight
// from registers). Measured on MI355X: 2^16 segments of 256 items 0.116 ->
// 0.107 ms by slot, 2^18 segments of 64 items 0.116 -> 0.125 ms by slot.
template <int R>
constexpr bool wave_finish_by_slot() { return R >= 4; }

// The wave's LDS region for wave_msd_sort, for ranges of at most MAXN items.
// The bucket path of 32-bit keys reads keys[] and vals[] together as one
// array of MAXN 64-bit sort words (they are adjacent and 16-byte aligned).
// Ranges of at most 256 items with 32-bit keys also stage their values there
// (by original position) for the bucket finish, which then gathers them from
// LDS instead of reading them from global memory a second time.
template <typename U, int MAXN>
constexpr bool wave_stage_values() { return sizeof(U) == 4 && MAXN <= 256; }

template <typename U, typename V, int MAXN = WAVE_SORT_MAX>
struct alignas(16) WaveSmem {
    U keys[MAXN];        // raw keys in bucket order; 64-bit keys: order keys of the bucket path
    V vals[MAXN];        // values by original position
    uint16_t pos[sizeof(U) == 8 ? MAXN : 1];  // 64-bit keys: original position of each bucket-ordered key
    V vstage[wave_stage_values<U, MAXN>() ? MAXN : 1];  // values by original position (bucket finish)
    uint32_t cnt[RADIX + 1];      // running digit counts, then bucket starts (cnt[buckets] = n)
    uint32_t dummy[WAVE];         // store target of the lanes that do not advance a count
    static_assert(sizeof(U) == 8 || sizeof(U) * MAXN + sizeof(V) * MAXN >= 8 * MAXN, "sort words overlay");
};

// Lanes whose `bits`-bit digit d equals the calling lane's (see
// RadixTile::match_digit): per bit, one ballot and a v_bitop3 per 32-lane half.
__device__ __forceinline__ uint64_t wave_match_rt(uint32_t d, int bits) {
    uint32_t acc_lo = ~0u, acc_hi = ~0u;
    for (int b = 0; b < bits; ++b) {
        uint32_t t = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);
        __asm__("" : "+v"(t));
        const uint64_t bal = __ballot(t != 0);
        acc_lo = __builtin_amdgcn_bitop3_b32(acc_lo, (uint32_t)bal, t, 0x90);
        acc_hi = __builtin_amdgcn_bitop3_b32(acc_hi, (uint32_t)(bal >> 32), t, 0x90);
    }
    return (uint64_t(acc_hi) << 32) | uint64_t(acc_lo);
}

// Exclusive scan of the wave's first `nbk` (<= RADIX) counters in place: lane l
// takes counters l, l+64, l+128, l+192; cnt[nbk] = n afterwards. Returns the
// largest count to every lane. LDS operations of one wave complete in order,
// so every lane's reads precede every lane's writes.
template <typename U, typename V, int MAXN>
__device__ __forceinline__ uint32_t wave_scan_counts(WaveSmem<U, V, MAXN>& ws, int nbk, int n) {
    const int l = (int)lane_id();
    uint32_t c[RADIX / WAVE];
#pragma unroll
    for (int i = 0; i < RADIX / WAVE; ++i) c[i] = i * WAVE + l < nbk ? ws.cnt[i * WAVE + l] : 0u;
    uint32_t carry = 0, big = 0;
#pragma unroll
    for (int i = 0; i < RADIX / WAVE; ++i) {
        if (i * WAVE >= nbk) break;  // wave-uniform
        const uint32_t incl = wave_incl_scan(c[i]);
        if (i * WAVE + l < nbk) ws.cnt[i * WAVE + l] = carry + incl - c[i];
        carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        big = c[i] > big ? c[i] : big;
    }
    if (l == 0) ws.cnt[nbk] = (uint32_t)n;
    return wave_max(big);
}

// One wave sorts a range of n <= 64*R items. Order keys are the twiddled keys
// masked to the bit range; the range's varying bits are [lo, hi).
//  * hi - lo <= 8: one stable counting pass on those bits orders the range
//    (digit ranks from 64-bit-ballot peer masks), written straight out;
//  * else: a bucket pass on the top `nb` varying bits (nb about log2(2n), at
//    most 8) places every item in its bucket through LDS atomics (unordered
//    inside the bucket), then each lane ranks its own items inside their
//    bucket by comparing sort words (order key, original position): the count
//    of smaller words is the item's stable rank. Keys and values go from the
//    lane's registers straight to their final places.
// Returns false, having written nothing, when a bucket exceeds `guard` items
// (then a network sort is cheaper). Lanes of one wave need no barrier: its LDS
// operations complete in order.
template <int R, typename U, typename V, bool HAS_V, int MAXN>
__device__ bool wave_msd_sort(const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n, const KeyMode& km,
                              WaveSmem<U, V, MAXN>& ws, int guard) {
    const int l = (int)lane_id();
    const U rm = range_mask<U>(km.begin_bit, km.end_bit);
    U raw[R], ok[R];
    V v[R];
    U o = 0, a = Bits<U>::ALL;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int p = r * WAVE + l;
        const int pc = p < n ? p : n - 1;
        raw[r] = at32(ks + begin, pc);
        if constexpr (HAS_V) v[r] = at32(vs + begin, pc);
        ok[r] = U(twiddle_in<U>(raw[r], km.kind, km.desc) & rm);
        if (p < n) {
            o |= ok[r];
            a &= ok[r];
        }
    }
    o = wave_or(o);
    a = wave_and(a);
    const U var = U(o ^ a);
    if (var == 0) {  // every key equal within the bit range: already in order
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int p = r * WAVE + l;
            if (p < n) {
                at32(ko + begin, p) = raw[r];
                if constexpr (HAS_V) at32(vo + begin, p) = v[r];
            }
        }
        return true;
    }
    const int hi = bit_length<U>(var), lo = lowest_bit<U>(var);
    if (hi - lo <= RADIX_BITS) {
        // one stable counting pass holds every varying bit
        const int bits = hi - lo;
        const uint32_t dm = (1u << bits) - 1u;
#pragma unroll
        for (int i = 0; i < RADIX / WAVE; ++i) ws.cnt[i * WAVE + l] = 0;
        uint32_t rk[R], dg[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool valid = r * WAVE + l < n;
            dg[r] = (uint32_t)(ok[r] >> lo) & dm;
            const uint32_t base = ws.cnt[dg[r]];
            const uint64_t peers = wave_match_rt(dg[r], bits) & __ballot(valid);
            const uint32_t below = lanes_below(peers);
            uint32_t* dst = (valid && below == 0) ? &ws.cnt[dg[r]] : &ws.dummy[l];
            *dst = base + (uint32_t)__popcll(peers);
            rk[r] = base + below;
        }
        wave_scan_counts(ws, 1 << bits, n);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r * WAVE + l < n) {
                const uint32_t d = ws.cnt[dg[r]] + rk[r];
                at32(ko + begin, d) = raw[r];
                if constexpr (HAS_V) at32(vo + begin, d) = v[r];
            }
        }
        return true;
    }
    // bucket pass on the top nb varying bits: about two buckets per item
    int nb = bit_length<uint32_t>((uint32_t)n) + 1;
    nb = nb < RADIX_BITS ? nb : RADIX_BITS;
    const int shift = hi - nb;
    const int nbk = 1 << nb;
    const uint32_t dm = (uint32_t)nbk - 1u;
#pragma unroll
    for (int i = 0; i < RADIX / WAVE; ++i)
        if (i * WAVE + l < nbk) ws.cnt[i * WAVE + l] = 0;
    uint32_t rk[R], dg[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dg[r] = (uint32_t)(ok[r] >> shift) & dm;
        if (r * WAVE + l < n) rk[r] = atomicAdd(&ws.cnt[dg[r]], 1u);  // place inside the bucket (unordered)
    }
    if (wave_scan_counts(ws, nbk, n) > (uint32_t)guard) return false;
    // sort words in bucket order: (order key, original position)
    constexpr bool STAGE_V = HAS_V && wave_finish_by_slot<R>() && wave_stage_values<U, MAXN>();
    uint32_t st[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int p = r * WAVE + l;
        st[r] = ws.cnt[dg[r]];
        if (p < n) {
            if constexpr (STAGE_V) ws.vstage[p] = v[r];
            const uint32_t at = st[r] + rk[r];
            if constexpr (sizeof(U) == 4) {
                reinterpret_cast<uint64_t*>(ws.keys)[at] = (uint64_t(ok[r]) << 32) | uint64_t(p);
            } else {
                ws.keys[at] = ok[r];
                ws.pos[at] = (uint16_t)p;
            }
        }
    }
    if constexpr (wave_finish_by_slot<R>()) {
        // lane l finishes the items at slots r*64 + l of the bucket order
        const bool full = rm == Bits<U>::ALL;
        // values not staged in LDS: every gather issued before any store (the
        // memory counter completes in order), positions from the sort words
        constexpr bool GATHER_FIRST = HAS_V && !STAGE_V && sizeof(U) == 4;
        V vg[R];
        if constexpr (GATHER_FIRST) {
            const uint64_t* w = reinterpret_cast<const uint64_t*>(ws.keys);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int q0 = r * WAVE + l;
                vg[r] = at32(vs + begin, q0 < n ? (uint32_t)w[q0] : 0u);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int q0 = r * WAVE + l;
            if (q0 < n) {
                U mk;
                uint32_t p, c = 0;
                int s;
                if constexpr (sizeof(U) == 4) {
                    const uint64_t* w = reinterpret_cast<const uint64_t*>(ws.keys);
                    const uint64_t mine = w[q0];
                    mk = U(mine >> 32);
                    p = (uint32_t)mine;
                    const uint32_t d = (uint32_t)(mk >> shift) & dm;
                    s = (int)ws.cnt[d];
                    const int e = (int)ws.cnt[d + 1];
                    for (int q = s; q < e; ++q) c += w[q] < mine;
                } else {
                    mk = ws.keys[q0];
                    p = ws.pos[q0];
                    const uint32_t d = (uint32_t)(mk >> shift) & dm;
                    s = (int)ws.cnt[d];
                    const int e = (int)ws.cnt[d + 1];
                    for (int q = s; q < e; ++q) {
                        const U kq = ws.keys[q];
                        c += (kq < mk) | ((kq == mk) & ((uint32_t)ws.pos[q] < p));
                    }
                }
                V val{};  // loads before stores (in-order memory counter)
                if constexpr (STAGE_V) val = ws.vstage[p];
                else if constexpr (GATHER_FIRST) val = vg[r];
                else if constexpr (HAS_V) val = at32(vs + begin, p);
                const U kout = full ? twiddle_out<U>(mk, km.kind, km.desc) : at32(ks + begin, p);
                at32(ko + begin, s + (int)c) = kout;
                if constexpr (HAS_V) at32(vo + begin, s + (int)c) = val;
            }
        }
        return true;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int p = r * WAVE + l;
        if (p < n) {
            const int s = (int)st[r], e = (int)ws.cnt[dg[r] + 1];
            uint32_t c = 0;
            if constexpr (sizeof(U) == 4) {
                const uint64_t* w = reinterpret_cast<const uint64_t*>(ws.keys);
                const uint64_t mine = (uint64_t(ok[r]) << 32) | uint64_t(p);
                for (int q = s; q < e; ++q) c += w[q] < mine;
            } else {
                for (int q = s; q < e; ++q) {
                    const U kq = ws.keys[q];
                    c += (kq < ok[r]) | ((kq == ok[r]) & ((int)ws.pos[q] < p));
                }
            }
            at32(ko + begin, s + (int)c) = raw[r];
            if constexpr (HAS_V) at32(vo + begin, s + (int)c) = v[r];
        }
    }
    return true;
}

// One wave sorts a range of n <= 64 items stably: the rank of item l is the
// number of items whose masked order key is smaller, or equal and earlier.
template <typename U, typename V, bool HAS_V>
__device__ __forceinline__ void wave_rank_sort(const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n,
                                               const KeyMode& km) {
    const int l = (int)lane_id();
    const bool valid = l < n;
    const U raw = valid ? at32(ks + begin, l) : U(0);
    V val = V(0);
    if constexpr (HAS_V) val = valid ? at32(vs + begin, l) : V(0);
    const U okey = U(twiddle_in<U>(raw, km.kind, km.desc) & range_mask<U>(km.begin_bit, km.end_bit));
    uint32_t rank = 0;
    for (int j = 0; j < n; ++j) {
        const U kj = read_lane<U>(okey, j);
        rank += (kj < okey) | ((kj == okey) & (j < l));
    }
    if (valid) {
        at32(ko + begin, rank) = raw;
        if constexpr (HAS_V) at32(vo + begin, rank) = val;
    }
}

// Sort one range of n <= MAXN (256 or 512) items with one wave: compare-count
// for the shortest ranges, a bitonic network up to 64 items, above that one
// counting pass plus a compare-count inside the buckets (the network on the
// smallest power of two of words per lane when a bucket is too large).
template <typename U, typename V, bool HAS_V, int MAXN>
__device__ __forceinline__ void wave_sort(const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n,
                                          const KeyMode& km, int rank_max, WaveSmem<U, V, MAXN>& ws) {
    static_assert(MAXN == 256 || MAXN == 512, "wave-tier range limit");
    constexpr int GUARD = 24;  // largest bucket the compare-count finish takes on
    if (n <= rank_max) wave_rank_sort<U, V, HAS_V>(ks, vs, ko, vo, begin, n, km);
    else if (n <= 64) {
        if (!wave_msd_sort<1, U, V, HAS_V, MAXN>(ks, vs, ko, vo, begin, n, km, ws, GUARD))
            wave_bitonic_sort<1, U, V, HAS_V>(ks, vs, ko, vo, begin, n, km, ws.vals);
    } else if (n <= 128) {
        if (!wave_msd_sort<2, U, V, HAS_V, MAXN>(ks, vs, ko, vo, begin, n, km, ws, GUARD))
            wave_bitonic_sort<2, U, V, HAS_V>(ks, vs, ko, vo, begin, n, km, ws.vals);
    } else if (MAXN == 256 || n <= 256) {
        if (!wave_msd_sort<4, U, V, HAS_V, MAXN>(ks, vs, ko, vo, begin, n, km, ws, GUARD))
            wave_bitonic_sort<4, U, V, HAS_V>(ks, vs, ko, vo, begin, n, km, ws.vals);
    } else if constexpr (MAXN > 256) {
        if (!wave_msd_sort<8, U, V, HAS_V, MAXN>(ks, vs, ko, vo, begin, n, km, ws, GUARD))
            wave_bitonic_sort<8, U, V, HAS_V>(ks, vs, ko, vo, begin, n, km, ws.vals);
    }
}

}  // namespace ssort
