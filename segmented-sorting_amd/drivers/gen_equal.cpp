// gen_equal: equal-length segments in the drivers' input format, the
// reference's utils/gen_equal.cpp (arguments: <segments> <segment size>,
// EXP_BITS_SIZE 12 by default) with a seed instead of time(NULL).
//   --bits B   keys in [0, 2^B) (default 12)
//   --seed S   stream seed (default 1)
//   --mode M   rand (default) | randminmax | sortasc | sortdesc, the
//              generator's compile-time variants (gen_equal.cpp:18-95):
//              uniform keys; each segment holding 0 and 2^B-1 among its
//              uniform keys, shuffled by swapping every slot with a random one;
//              all keys sorted ascending; all keys sorted descending.
// Keys of --mode rand are word i of stream (seed, 1) >> (64 - B).
#include <algorithm>
#include <functional>

#include "gen_common.hpp"

int main(int argc, char** argv) {
    gen::Args o;
    o.bits = 12;
    if (!gen::parse(argc, argv, o) || o.a < 1) {
        fprintf(stderr, "usage: gen_equal <number of segments> <size of each segment> [--bits B] [--seed S] "
                        "[--mode rand|randminmax|sortasc|sortdesc]\n");
        return 1;
    }
    const long long S = o.a, L = o.b, N = S * L;
    if (N >= (1ll << 31)) {
        fprintf(stderr, "gen_equal: more than 2^31 - 1 keys\n");
        return 1;
    }
    std::vector<long long> offsets((size_t)S + 1);
    for (long long i = 0; i <= S; ++i) offsets[(size_t)i] = i * L;
    std::vector<uint32_t> keys((size_t)N);
    gen::Stream ks(o.seed, 1);
    for (auto& k : keys) k = ks.bits(o.bits);
    const uint32_t top = o.bits == 32 ? 0xFFFFFFFFu : (o.bits == 0 ? 0u : (1u << o.bits) - 1u);
    if (o.mode == "randminmax") {
        gen::Stream sh(o.seed, 3);
        for (long long s = 0; s < S && L > 0; ++s) {
            uint32_t* seg = keys.data() + s * L;
            seg[0] = 0;
            if (L > 1) seg[1] = top;
            for (long long i = 0; i < L; ++i) std::swap(seg[i], seg[sh.below((uint64_t)L)]);
        }
    } else if (o.mode == "sortasc") {
        std::sort(keys.begin(), keys.end());
    } else if (o.mode == "sortdesc") {
        std::sort(keys.begin(), keys.end(), std::greater<uint32_t>());
    } else if (o.mode != "rand") {
        fprintf(stderr, "gen_equal: unknown mode %s\n", o.mode.c_str());
        return 1;
    }
    gen::write(offsets, keys);
    return 0;
}
