// Workgroup-level building blocks: stable digit ranking with 64-bit ballots,
// the LDS-resident LSD radix sort of one tile, the stable partition of one
// chunk by one digit (used by the large-range tier, the per-segment global
// LSD of the plain form and the fix-pass LSD), and the wave sort of tiny ranges.
#pragma once

#include "common.hpp"

namespace ssort {

// NT threads hold IPT items each, in "wave-striped" order: item j of lane l of
// wave w is tile position w*64*IPT + j*64 + l. This order is the input order,
// so ranking items in (w, j, l) order is stable.
template <int NT, int IPT, typename U, typename V, bool HAS_V>
struct RadixTile {
    static constexpr int NW = NT / WAVE;
    static constexpr int TILE = NT * IPT;
    static constexpr int NDT = NT < RADIX ? NT : RADIX;  // threads in the digit scan
    static constexpr int DPT = RADIX / NDT;               // digits per scanning thread
    static constexpr int NDW = (NDT + WAVE - 1) / WAVE;   // waves in the digit scan
    static_assert(NT % WAVE == 0, "whole waves");
    static_assert(RADIX % NDT == 0, "digit split");

    struct Smem {
        U keys[TILE];
        V vals[HAS_V ? TILE : 1];
        uint32_t wcnt[NW * RADIX];  // per-wave digit counts, then scatter bases
        uint32_t dstart[RADIX];     // tile-exclusive start of each digit
        uint32_t goff[RADIX];       // global destination of each digit (partition)
        uint32_t wsum[NDW];
        U red_or[NW];
        U red_and[NW];
    };

    __device__ static __forceinline__ int pos(int j) {
        const int t = threadIdx.x;
        return (t >> 6) * (WAVE * IPT) + j * WAVE + (t & 63);
    }

    // Stable rank of every valid item (position < n) by digit
    // (key >> shift) & (2^bits - 1): dst[j] is the item's tile position after a
    // stable counting sort; sm.dstart[d] the first position of digit d.
    __device__ static void rank_digits(Smem& sm, const U (&k)[IPT], int n, int shift, int bits,
                                       uint32_t (&dst)[IPT]) {
        const int t = threadIdx.x;
        const int w = t >> 6;
        const uint32_t l = t & 63;
        uint32_t* wc = sm.wcnt + w * RADIX;
#pragma unroll
        for (int q = 0; q < RADIX / WAVE; ++q) wc[q * WAVE + l] = 0;
        const U dmask = U((U(1) << bits) - 1);
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const bool valid = pos(j) < n;
            const uint32_t d = (uint32_t)((k[j] >> shift) & dmask);
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < RADIX_BITS; ++b) {
                if (b < bits) {
                    const bool bit = (d >> b) & 1u;
                    const uint64_t bb = __ballot(bit);
                    peers &= bit ? bb : ~bb;
                }
            }
            const uint32_t below = lanes_below(peers);
            const uint32_t base = wc[d];
            dst[j] = base + below;
            if (valid && below == 0) wc[d] = base + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        uint32_t tot[DPT];
        uint32_t run = 0;
        if (t < NDT) {
#pragma unroll
            for (int i = 0; i < DPT; ++i) {
                const int d = t * DPT + i;
                uint32_t s = 0;
                for (int ww = 0; ww < NW; ++ww) {
                    const uint32_t c = sm.wcnt[ww * RADIX + d];
                    sm.wcnt[ww * RADIX + d] = s;
                    s += c;
                }
                tot[i] = run;
                run += s;
            }
        }
        const uint32_t incl = wave_incl_scan(run);
        if (t < NDT && l == 63) sm.wsum[w] = incl;
        __syncthreads();
        if (t < NDT) {
            uint32_t excl = incl - run;
            for (int ww = 0; ww < w; ++ww) excl += sm.wsum[ww];
#pragma unroll
            for (int i = 0; i < DPT; ++i) {
                const int d = t * DPT + i;
                const uint32_t base = excl + tot[i];
                sm.dstart[d] = base;
                for (int ww = 0; ww < NW; ++ww) sm.wcnt[ww * RADIX + d] += base;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t d = (uint32_t)((k[j] >> shift) & dmask);
            dst[j] += wc[d];
        }
    }

    // Load n <= TILE items from src + begin (twiddled keys) into registers.
    __device__ static __forceinline__ void load(const U* ks, const V* vs, int64_t begin, int n, const KeyMode& km,
                                                U (&k)[IPT], V (&v)[IPT]) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int p = pos(j);
            if (p < n) {
                k[j] = twiddle_in<U>(ks[begin + p], km.kind, km.desc);
                if constexpr (HAS_V) v[j] = vs[begin + p];
            } else {
                k[j] = Bits<U>::ALL;
                if constexpr (HAS_V) v[j] = V(0);
            }
        }
    }

    // OR and AND of the valid keys over the whole tile (every thread gets them).
    __device__ static void reduce_or_and(Smem& sm, const U (&k)[IPT], int n, U& o, U& a) {
        U vo = 0, va = Bits<U>::ALL;
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (pos(j) < n) {
                vo |= k[j];
                va &= k[j];
            }
        vo = wave_or(vo);
        va = wave_and(va);
        const int w = threadIdx.x >> 6;
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
    // are ranked, so a low-entropy tile costs fewer passes.
    __device__ static void sort(Smem& sm, const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n,
                                const KeyMode& km) {
        U k[IPT];
        V v[IPT];
        load(ks, vs, begin, n, km, k, v);
        U o, a;
        reduce_or_and(sm, k, n, o, a);
        const U varying = U((o ^ a) & range_mask<U>(km.begin_bit, km.end_bit));
        if (varying == 0) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int p = pos(j);
                if (p < n) {
                    ko[begin + p] = twiddle_out<U>(k[j], km.kind, km.desc);
                    if constexpr (HAS_V) vo[begin + p] = v[j];
                }
            }
            __syncthreads();
            return;
        }
        const int lo = lowest_bit<U>(varying);
        const int hi = bit_length<U>(varying);
        for (int shift = lo; shift < hi; shift += RADIX_BITS) {
            const int bits = (hi - shift) < RADIX_BITS ? (hi - shift) : RADIX_BITS;
            uint32_t dst[IPT];
            rank_digits(sm, k, n, shift, bits, dst);
#pragma unroll
            for (int j = 0; j < IPT; ++j)
                if (pos(j) < n) {
                    sm.keys[dst[j]] = k[j];
                    if constexpr (HAS_V) sm.vals[dst[j]] = v[j];
                }
            __syncthreads();
            if (shift + RADIX_BITS < hi) {
#pragma unroll
                for (int j = 0; j < IPT; ++j) {
                    const int p = pos(j);
                    if (p < n) {
                        k[j] = sm.keys[p];
                        if constexpr (HAS_V) v[j] = sm.vals[p];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int p = pos(j);
            if (p < n) {
                ko[begin + p] = twiddle_out<U>(sm.keys[p], km.kind, km.desc);
                if constexpr (HAS_V) vo[begin + p] = sm.vals[p];
            }
        }
        __syncthreads();
    }

    // Stable partition of the n <= TILE items at ks[begin..) by digit
    // (twiddled >> shift) & (2^bits - 1). Item with digit d and rank r among
    // the chunk's digit-d items goes to global index goff[d] + r. The chunk is
    // ordered by digit in LDS first, so each digit run is written contiguously.
    __device__ static void partition(Smem& sm, const U* ks, const V* vs, U* kd, V* vd, int64_t begin, int n,
                                     const KeyMode& km, int shift, int bits, const uint32_t* goff_global) {
        U k[IPT];
        V v[IPT];
        load(ks, vs, begin, n, km, k, v);
        for (int d = threadIdx.x; d < RADIX; d += NT) sm.goff[d] = goff_global[d];
        uint32_t dst[IPT];
        rank_digits(sm, k, n, shift, bits, dst);  // contains barriers: goff visible after
#pragma unroll
        for (int j = 0; j < IPT; ++j)
            if (pos(j) < n) {
                sm.keys[dst[j]] = k[j];
                if constexpr (HAS_V) sm.vals[dst[j]] = v[j];
            }
        __syncthreads();
        const U dmask = U((U(1) << bits) - 1);
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int p = pos(j);
            if (p < n) {
                const U key = sm.keys[p];
                const uint32_t d = (uint32_t)((key >> shift) & dmask);
                const uint32_t g = sm.goff[d] + ((uint32_t)p - sm.dstart[d]);
                kd[g] = twiddle_out<U>(key, km.kind, km.desc);
                if constexpr (HAS_V) vd[g] = sm.vals[p];
            }
        }
        __syncthreads();
    }

    // Digit histograms of the n <= TILE items at ks[begin..) for `npos` digit
    // positions pos_p = pos0 + 8p (the last may be narrower than 8 bits, up to
    // hi_bit); counts are accumulated into the LDS array hist[npos][RADIX].
    __device__ static void histogram(uint32_t* hist, const U* ks, int64_t begin, int n, const KeyMode& km,
                                     int pos0, int npos, int hi_bit) {
        for (int i = threadIdx.x; i < n; i += NT) {
            const U key = twiddle_in<U>(ks[begin + i], km.kind, km.desc);
            for (int p = 0; p < npos; ++p) {
                const int sh = pos0 + p * RADIX_BITS;
                const int nb = (hi_bit - sh) < RADIX_BITS ? (hi_bit - sh) : RADIX_BITS;
                const uint32_t d = (uint32_t)((key >> sh) & U((U(1) << nb) - 1));
                atomicAdd(&hist[p * RADIX + d], 1u);
            }
        }
    }
};

// One wave sorts a range of n <= 64 items stably: the rank of item l is the
// number of items whose masked order key is smaller, or equal and earlier.
template <typename U, typename V, bool HAS_V>
__device__ __forceinline__ void wave_rank_sort(const U* ks, const V* vs, U* ko, V* vo, int64_t begin, int n,
                                               const KeyMode& km) {
    const int l = (int)lane_id();
    const bool valid = l < n;
    const U raw = valid ? ks[begin + l] : U(0);
    V val = V(0);
    if constexpr (HAS_V) val = valid ? vs[begin + l] : V(0);
    const U okey = U(twiddle_in<U>(raw, km.kind, km.desc) & range_mask<U>(km.begin_bit, km.end_bit));
    uint32_t rank = 0;
    for (int j = 0; j < n; ++j) {
        const U kj = read_lane<U>(okey, j);
        rank += (kj < okey) | ((kj == okey) & (j < l));
    }
    if (valid) {
        ko[begin + rank] = raw;
        if constexpr (HAS_V) vo[begin + rank] = val;
    }
}

}  // namespace ssort
