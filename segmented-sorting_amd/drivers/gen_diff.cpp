// gen_diff: segments of random lengths in the drivers' input format, the
// reference's utils/gen_diff.cpp (arguments: <segments> <elements>,
// EXP_BITS_SIZE 10 by default) with a seed instead of time(NULL): offset 0,
// then segments - 1 distinct random positions in [1, elements) (drawn until
// distinct, as the reference does, from stream (seed, 2)), sorted, then
// `elements`; no segment is empty.
//   --bits B   keys in [0, 2^B) (default 10), word i of stream (seed, 1) >> (64 - B)
//   --seed S   stream seed (default 1)
#include <algorithm>

#include "gen_common.hpp"

int main(int argc, char** argv) {
    gen::Args o;
    o.bits = 10;
    if (!gen::parse(argc, argv, o) || o.a < 1 || o.b < o.a) {
        fprintf(stderr, "usage: gen_diff <number of segments> <number of elements> [--bits B] [--seed S]\n");
        return 1;
    }
    const long long S = o.a, N = o.b;
    if (N >= (1ll << 31)) {
        fprintf(stderr, "gen_diff: more than 2^31 - 1 keys\n");
        return 1;
    }
    std::vector<char> taken((size_t)N, 0);
    std::vector<long long> offsets;
    offsets.reserve((size_t)S + 1);
    offsets.push_back(0);
    taken[0] = 1;
    gen::Stream ps(o.seed, 2);
    while ((long long)offsets.size() < S) {
        const long long x = (long long)ps.below((uint64_t)N);
        if (!taken[(size_t)x]) {
            taken[(size_t)x] = 1;
            offsets.push_back(x);
        }
    }
    offsets.push_back(N);
    std::sort(offsets.begin(), offsets.end());
    std::vector<uint32_t> keys((size_t)N);
    gen::Stream ks(o.seed, 1);
    for (auto& k : keys) k = ks.bits(o.bits);
    gen::write(offsets, keys);
    return 0;
}
