// Shared I/O of the reference-compatible drivers: the stdin format
// "S \n S+1 offsets \n N \n N keys" and the stdout format "\n" + "k " per key
// + "\n" (src/radixseg.cu:54-60, 77-90, 154-159), or one line with the mean
// milliseconds of the timed sort calls when run with --time.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace drv {

struct Options {
    bool time = false;     // print mean ms instead of the keys (-DELAPSED_TIME=1 in the reference)
    bool values = false;   // also print the sorted values, as a second line in the keys' format
    int executions = 1;    // timed repetitions (-DEXECUTIONS)
    int warmup = 0;        // untimed calls before them (0, as the reference: its first call is timed too)
    int algo = 0;          // segsort_algo
    bool bound = false;    // pass the longest segment (known from the host offsets) as the length bound
    bool resident = false; // upload the inputs once instead of before every call (radixseg)
};

inline Options parse(int argc, char** argv, int default_algo) {
    Options o;
    o.algo = default_algo;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--time")) o.time = true;
        else if (!strcmp(argv[i], "--values")) o.values = true;
        else if (!strcmp(argv[i], "--bound")) o.bound = true;
        else if (!strcmp(argv[i], "--resident")) o.resident = true;
        else if (!strcmp(argv[i], "--executions") && i + 1 < argc) o.executions = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--warmup") && i + 1 < argc) o.warmup = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--algo") && i + 1 < argc) {
            const char* a = argv[++i];
            o.algo = !strcmp(a, "plain") ? 1 : !strcmp(a, "fixpass") ? 2 : 0;
        }
    }
    if (o.executions < 1) o.executions = 1;
    if (o.warmup < 0) o.warmup = 0;
    return o;
}

// Reads the reference's input format. Keys are parsed as unsigned 32-bit
// (values >= 2^31 are kept; the reference's scanf("%d") wraps them, same bits).
inline bool read_input(std::vector<int>& offsets, std::vector<unsigned>& keys) {
    long long s = 0, n = 0;
    if (scanf("%lld", &s) != 1 || s < 0) return false;
    offsets.resize((size_t)s + 1);
    for (long long i = 0; i <= s; ++i) {
        long long v;
        if (scanf("%lld", &v) != 1) return false;
        offsets[(size_t)i] = (int)v;
    }
    if (scanf("%lld", &n) != 1 || n < 0) return false;
    keys.resize((size_t)n);
    for (long long i = 0; i < n; ++i) {
        long long v;
        if (scanf("%lld", &v) != 1) return false;
        keys[(size_t)i] = (unsigned)v;
    }
    return true;
}

template <typename T>
inline void print_line(const T* h, size_t n, bool lead) {
    std::string out;
    out.reserve(n * 11 + 2);
    if (lead) out += "\n";
    char buf[32];
    for (size_t i = 0; i < n; ++i) {
        int len = snprintf(buf, sizeof(buf), "%lld ", (long long)h[i]);
        out.append(buf, (size_t)len);
    }
    out += "\n";
    fwrite(out.data(), 1, out.size(), stdout);
}

// "\n" + "k " per key + "\n", the reference's output (src/cub.out).
template <typename T>
inline void print_keys(const T* h, size_t n) {
    print_line(h, n, true);
}

// With --values: the sorted values follow as one more line, "v " per value + "\n".
template <typename T>
inline void print_values(const T* h, size_t n) {
    print_line(h, n, false);
}

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        fprintf(stderr, "%s failed: %s (code %d)\n", what, hipGetErrorString(e), (int)e);
        exit(EXIT_FAILURE);
    }
}

}  // namespace drv
