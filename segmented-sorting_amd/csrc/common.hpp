// Shared device primitives for the gfx950 segmented sort: key twiddling
// (radix order), wave64 helpers and the buffer/work-item records that the
// kernels exchange through device memory.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ssort {

constexpr int WAVE = 64;
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;

// Key classes, as the radix path orders them (lib/cub/util_type.cuh:889-1036).
enum : int { KIND_UNSIGNED = 0, KIND_SIGNED = 1, KIND_FLOAT = 2 };

template <typename U>
struct Bits {
    static constexpr int N = int(sizeof(U) * 8);
    static constexpr U HIGH = U(U(1) << (N - 1));
    static constexpr U ALL = U(~U(0));
};

// Map a raw key to the unsigned pattern whose integer order is the sort order:
// unsigned identity, signed ^ sign bit, float (neg ? ~b : b ^ sign), and every
// bit flipped for a descending sort (util_type.cuh:904-911, 947-955, 1017-1027).
// Written as one XOR with a mask built from the (uniform) kind and direction,
// without branches, so that loads feeding it can all be in flight at once.
template <typename U>
__device__ __forceinline__ U twiddle_in(U b, int kind, bool desc) {
    const U neg = U(U(0) - U(b >> (Bits<U>::N - 1)));  // all ones when the sign bit is set
    const U fmask = U(neg | Bits<U>::HIGH);
    const U m = kind == KIND_FLOAT ? fmask : (kind == KIND_SIGNED ? Bits<U>::HIGH : U(0));
    return U(b ^ m ^ (desc ? Bits<U>::ALL : U(0)));
}

template <typename U>
__device__ __forceinline__ U twiddle_out(U r, int kind, bool desc) {
    const U a = U(r ^ (desc ? Bits<U>::ALL : U(0)));
    const U neg = U(U(0) - U(a >> (Bits<U>::N - 1)));
    const U fmask = U(U(~neg) | Bits<U>::HIGH);  // sign set: the sign bit, else all ones
    const U m = kind == KIND_FLOAT ? fmask : (kind == KIND_SIGNED ? Bits<U>::HIGH : U(0));
    return U(a ^ m);
}

// How the keys of one call are ordered: key class, direction and the
// [begin_bit, end_bit) range of twiddled bits that takes part in the order.
struct KeyMode {
    int kind;
    int desc;
    int begin_bit;
    int end_bit;
};

template <typename U>
__device__ __forceinline__ U range_mask(int begin_bit, int end_bit) {
    int nb = end_bit - begin_bit;
    if (nb <= 0) return U(0);
    U m = (nb >= Bits<U>::N) ? Bits<U>::ALL : U((U(1) << nb) - 1);
    return U(m << begin_bit);
}

template <typename U>
__device__ __forceinline__ int bit_length(U x) {
    if (x == 0) return 0;
    if constexpr (sizeof(U) == 8) return 64 - __clzll((long long)x);
    else return 32 - __clz((int)x);
}

template <typename U>
__device__ __forceinline__ int lowest_bit(U x) {
    if constexpr (sizeof(U) == 8) return __ffsll((long long)x) - 1;
    else return __ffs((int)x) - 1;
}

// ---- wave64 helpers --------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of m at lanes strictly below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// DPP controls: row_shr:n shifts within rows of 16 lanes, row_bcast:15 and
// row_bcast:31 pass lane 15 of each row (lane 31) on to the next row(s).
constexpr int DPP_ROW_SHR = 0x110;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane l reads lane l + 1 (lane 63 keeps `old`)
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane l reads lane l - 1 (lane 0 keeps `old`)

// The value of the neighbouring lane (l - 1 for SHR, l + 1 for SHL) of a 32-
// or 64-bit word; the lane without a neighbour gets `edge`.
template <int CTRL, typename W>
__device__ __forceinline__ W lane_shift(W x, W edge) {
    if constexpr (sizeof(W) == 4) {
        return (W)__builtin_amdgcn_update_dpp((int)edge, (int)x, CTRL, 0xF, 0xF, false);
    } else {
        const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)edge, (int)(uint32_t)x, CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(edge >> 32), (int)(uint32_t)(x >> 32), CTRL, 0xF,
                                                   0xF, false);
        return (W(uint32_t(hi)) << 32) | W(uint32_t(lo));
    }
}

// One step of a DPP scan: the value of the source lane, or `identity` where
// the control has no source lane or the row is masked off.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_src(uint32_t x, uint32_t identity) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)x, CTRL, ROWS, 0xF, false);
}

// Inclusive scan over the 64 lanes with cross-lane DPP moves (no LDS):
// Hillis-Steele inside each row of 16, then rows 1 and 3 add lane 15 of the
// row before them, and rows 2 and 3 add lane 31.
template <typename Op>
__device__ __forceinline__ uint32_t wave_incl_scan_op(uint32_t x, uint32_t identity, Op op) {
    x = op(x, dpp_src<DPP_ROW_SHR + 1, 0xF>(x, identity));
    x = op(x, dpp_src<DPP_ROW_SHR + 2, 0xF>(x, identity));
    x = op(x, dpp_src<DPP_ROW_SHR + 4, 0xF>(x, identity));
    x = op(x, dpp_src<DPP_ROW_SHR + 8, 0xF>(x, identity));
    x = op(x, dpp_src<DPP_ROW_BCAST15, 0xA>(x, identity));
    x = op(x, dpp_src<DPP_ROW_BCAST31, 0xC>(x, identity));
    return x;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    return wave_incl_scan_op(x, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

// Maximum of x over the wave, returned to every lane.
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
    x = wave_incl_scan_op(x, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Sum of x over the wave, returned to every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(x), 63);
}

// OR / AND of x over the wave, returned to every lane (lane 63 of the
// inclusive scan holds the whole wave's value).
template <typename U>
__device__ __forceinline__ U wave_or(U x) {
    auto f = [](uint32_t v) {
        v = wave_incl_scan_op(v, 0u, [](uint32_t a, uint32_t b) { return a | b; });
        return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
    };
    if constexpr (sizeof(U) == 8) return (U(f((uint32_t)(x >> 32))) << 32) | U(f((uint32_t)x));
    else return (U)f((uint32_t)x);
}

template <typename U>
__device__ __forceinline__ U wave_and(U x) {
    auto f = [](uint32_t v) {
        v = wave_incl_scan_op(v, ~0u, [](uint32_t a, uint32_t b) { return a & b; });
        return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
    };
    if constexpr (sizeof(U) == 8) return (U(f((uint32_t)(x >> 32))) << 32) | U(f((uint32_t)x));
    else return (U)f((uint32_t)x);
}

template <typename U>
__device__ __forceinline__ U read_lane(U v, int l) {
    if constexpr (sizeof(U) == 8) {
        uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
        uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
        return (U(hi) << 32) | U(lo);
    } else {
        return (U)(uint32_t)__builtin_amdgcn_readlane((int)v, l);
    }
}

// ---- buffers and work items -----------------------------------------------

// The four places a range of keys/values can live in during one call.
enum : int { BUF_IN = 0, BUF_TMP_A = 1, BUF_TMP_B = 2, BUF_OUT = 3 };

struct Bufs {
    const void* k[4];
    const void* v[4];
};

template <typename T>
__device__ __forceinline__ T* buf_ptr(const void* p) {
    return (T*)const_cast<void*>(p);
}

// A range that one tile (or one wave) sorts from buffer `src` into BUF_OUT.
// n < 2^30; the two high bits of n_src hold the source buffer id.
struct TileItem {
    int begin;
    int n_src;
};

__device__ __forceinline__ int ti_n(TileItem t) { return t.n_src & 0x3FFFFFFF; }
__device__ __forceinline__ int ti_src(TileItem t) { return (int)((uint32_t)t.n_src >> 30); }
__device__ __forceinline__ TileItem make_ti(int begin, int n, int src) {
    TileItem t;
    t.begin = begin;
    t.n_src = n | (src << 30);
    return t;
}

// A range larger than one tile. In the multi-workgroup MSD levels it is cut
// into nch chunks of chunk_len items (a multiple of the partition tile, so
// that a range has at most a few hundred chunks), numbered from chunk_base in
// its level's chunk arrays; the depth-first tier and the plain form's
// per-segment LSD use only the first four fields.
struct LargeItem {
    int begin;       // first element (same index in every buffer)
    int n;           // element count
    int src;         // buffer holding the range now
    int hi_bit;      // bits at and above hi_bit are equal within the range
    int chunk_base;  // first chunk id of this range in its level
    int chunk_len;   // items per chunk (the last one may be shorter)
    int nch;         // chunks of the range
    int pad;
};

}  // namespace ssort
