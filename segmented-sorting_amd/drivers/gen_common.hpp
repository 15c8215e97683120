// Seeded input generators for the drivers (host only). The reference's
// generators seed rand() with time(NULL) (utils/gen_equal.cpp:107,
// utils/gen_diff.cpp:62), so their files cannot be made again; these draw
// from a counter-based splitmix64 stream keyed by (seed, stream), the same
// stream as the Python package's gen.u64, so a file can be regenerated
// anywhere and its keys recomputed by the test harness.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace gen {

inline uint64_t mix(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

// Word i (0-based) of the stream keyed by (seed, stream).
struct Stream {
    uint64_t base;
    uint64_t i = 0;
    Stream(uint64_t seed, uint64_t stream) : base(mix(seed * 0x1000193ull + stream * 0x7F4A7C15ull + 1ull)) {}
    uint64_t next() { return mix((++i) * 0x9E3779B97F4A7C15ull + base); }
    // uniform in [0, 2^bits), from the word's top bits
    uint32_t bits(int b) { return b == 0 ? 0u : (uint32_t)(next() >> (64 - b)); }
    // uniform in [0, m) (m > 0), by the word's value modulo m
    uint64_t below(uint64_t m) { return next() % m; }
};

// The reference's file format: "S\n" + S+1 offsets + "\nN\n" + N keys + "\n"
// (each number followed by a space, as utils/gen_*.cpp print them).
inline void write(const std::vector<long long>& offsets, const std::vector<uint32_t>& keys) {
    std::string out;
    out.reserve(offsets.size() * 11 + keys.size() * 11 + 64);
    char buf[32];
    int len = snprintf(buf, sizeof(buf), "%zu\n", offsets.size() - 1);
    out.append(buf, (size_t)len);
    for (long long o : offsets) out.append(buf, (size_t)snprintf(buf, sizeof(buf), "%lld ", o));
    len = snprintf(buf, sizeof(buf), "\n%zu\n", keys.size());
    out.append(buf, (size_t)len);
    for (uint32_t k : keys) out.append(buf, (size_t)snprintf(buf, sizeof(buf), "%u ", k));
    out += "\n";
    fwrite(out.data(), 1, out.size(), stdout);
}

struct Args {
    long long a = -1, b = -1;
    int bits;
    uint64_t seed = 1;
    std::string mode = "rand";
};

inline bool parse(int argc, char** argv, Args& o) {
    int pos = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--bits") && i + 1 < argc) o.bits = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--seed") && i + 1 < argc) o.seed = strtoull(argv[++i], nullptr, 0);
        else if (!strcmp(argv[i], "--mode") && i + 1 < argc) o.mode = argv[++i];
        else if (pos == 0) o.a = atoll(argv[i]), ++pos;
        else if (pos == 1) o.b = atoll(argv[i]), ++pos;
        else return false;
    }
    return pos == 2 && o.a >= 0 && o.b >= 0 && o.bits >= 0 && o.bits <= 32;
}

}  // namespace gen
