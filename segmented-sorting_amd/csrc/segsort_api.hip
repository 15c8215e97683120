// C-ABI of libsegsort.so: argument checks, temp-storage layout, tuning table
// and the launch sequences of the batched, plain and fix-pass forms.
#include <hip/hip_ext.h>

#include "kernels.hpp"
#include "profile.hpp"
#include "../../include/segsort.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

using namespace ssort;

namespace {

thread_local std::string g_last_error = "no error";

// Set while segsort_sort_noalloc runs: the call launches its kernels directly on
// the caller's stream (no graph capture, no second stream), so that it creates
// and allocates nothing.
thread_local bool t_no_alloc = false;

// The caller's bound on the segment lengths of the call in progress
// (segsort_sort_hinted), 0 when none was given.
thread_local int t_max_len = 0;

int fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
}

int ok() {
    g_last_error = "no error";
    return 0;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }


// Workgroups of the oversize fallback (k_seg_lsd over the segments that break
// a caller's length bound): empty in every correct call, so small.
constexpr int FALLBACK_GRID = 32;

// The binning kernel's gate token of one call: a process-wide counter, from a
// random start, through a bijective 64-bit mix (splitmix64's finaliser), so no
// two calls of the process share one and stale temp memory matches one only
// by chance (2^-64).
unsigned long long next_gate_token() {
    static std::atomic<unsigned long long> counter{((unsigned long long)std::random_device{}() << 32) ^
                                                   (unsigned long long)std::random_device{}()};
    unsigned long long z = counter.fetch_add(1, std::memory_order_relaxed) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z ? z : 1;  // 0 is the closed gate
}

// Segments per binning thread: one while that keeps >= 64 blocks, at most BIN_SPT.
int bin_spt(int S) { return std::min(BIN_SPT, std::max(1, (S + BIN_NT * 64 - 1) / (BIN_NT * 64))); }
int bin_grid(int S) { return (S + BIN_NT * bin_spt(S) - 1) / (BIN_NT * bin_spt(S)); }

// Ranges up to this length use the wave's compare-count rank (about 4n VALU
// operations per item); longer ones the register bitonic network.
constexpr int WAVE_RANK_MAX = 24;

// ---- tuning ------------------------------------------------------------------

struct Tuning {
    int wave_cap;     // ranges of 2..wave_cap items go to the wave tier
    int tile_enable;  // bit i < 4: LDS tile class i on; bit 4 + g: wave-group class of 4 << g items on;
                      // TUNE_LPT: the 16-item tile classes take their ranges longest first
    int grid_mult;    // persistent workgroups per resident slot
    int concurrent;   // 1: binned ranges sorted on a second stream while the MSD tier runs; 2: only those of the
                      // 16-item tile classes
};

// The built-in tuning table (generated from tuning.json by build.py): one entry
// per signature (key width, value width, round(log2 N), round(log2 mean
// segment length), and ceil(log2 max segment length) for the calls that state a
// bound on their segment lengths, -1 for those that do not), as the autotuner
// measured them. A call takes the entry of its own signature, else the nearest
// one with the same key and value widths (distance |d log2 N| + 2 |d log2 mean|
// + 2 |d log2 max|; a call with a bound looks at the bounded entries first),
// else DEFAULT_TUNING, the way cub picks its tuning policy inside the call
// (dispatch_radix_sort.cuh:441-661): host arithmetic only, no device read.
struct TableEntry {
    int key_width, value_bytes, log2n, log2m, log2x;
    Tuning tu;
};
#include "tuning_table.inc"

// tile_enable bit 8: the batched form's 16-item tile classes (256x16, 512x16,
// 1024x16) take their binned ranges longest first, from length buckets. Measured
// (2^24 items in 2^16 lognormal segments): 0.209 -> 0.202 ms, the 1024x16 and
// 256x16 kernels no longer ending on long ranges started last. Off for equal
// lengths (one bucket) and tiny segments: the bucketed builds cost 2-9% there
// (2^10 segments of 2^14: 0.159 -> 0.168 ms), so the table sets it only for the
// signatures where it measured faster, and the default leaves it off.
constexpr int TUNE_LPT = 0x100;

// tile_enable bit 9: the 4096-item tile (256x16) without its half-depth build
// (8 items per thread for ranges of at most 2048 items), which otherwise makes
// the kernel spill. Measured: 2^12 equal segments of 4096 0.1181 -> 0.1137 ms
// without it, ragged 2^24 0.1953 -> 0.2029 ms (its 1025..2048-item ranges want it).
constexpr int TUNE_NO_HALF = 0x200;

// tile_enable bit 10: with a wave cap above 256, the wave tier is two classes:
// ranges of at most 256 items in the 256-item build (8 waves per SIMD, no
// spills) and the longer ones in the 512-item build, in separate launches,
// instead of every range in the 512-item build (7 waves per SIMD, about 23
// VGPRs spilled per lane). For the ranges of 256 +- 16 items that an 8-bit
// partition of 2^16-item ranges leaves. Measured: 2^8 equal segments of 2^16
// 0.2202 -> 0.2177 ms, one 2^24 segment unchanged (the 512-item build's time is
// in its longer ranges themselves); the table sets it for the former only.
constexpr int TUNE_SPLIT_WAVE = 0x400;
constexpr int TUNE_BITS = 0xFF | TUNE_LPT | TUNE_NO_HALF | TUNE_SPLIT_WAVE;

constexpr Tuning DEFAULT_TUNING = {256, 0xEF, 1, 0};

// Overrides installed by segsort_set_tuning, per (key width, value width) slot:
// index (key_width == 8) * 3 + value-bytes slot (0, 4, 8 -> 0, 1, 2).
Tuning g_override[6];
bool g_override_on[6] = {false, false, false, false, false, false};
std::mutex g_tuning_mu;

int tuning_slot(int key_width, int value_bytes) {
    const int vs = value_bytes == 0 ? 0 : value_bytes == 4 ? 1 : value_bytes == 8 ? 2 : -1;
    if (vs < 0 || (key_width != 4 && key_width != 8)) return -1;
    return (key_width == 8 ? 3 : 0) + vs;
}

int rounded_log2(double x) { return (int)std::lround(std::log2(x < 1.0 ? 1.0 : x)); }

// Bits needed for a length: the smallest x with max_len <= 2^x.
int ceil_log2(long long v) {
    int x = 0;
    while (x < 62 && (1ll << x) < v) ++x;
    return x;
}

const TableEntry* table_lookup(int key_width, int value_bytes, long long num_items, long long num_segments,
                               long long max_len) {
    const int ln = rounded_log2((double)(num_items > 1 ? num_items : 1));
    const double mean = (double)(num_items > 1 ? num_items : 1) / (double)(num_segments > 1 ? num_segments : 1);
    const int lm = rounded_log2(mean);
    const int lx = max_len > 0 ? ceil_log2(max_len) : -1;
    for (int pass = lx >= 0 ? 0 : 1; pass < 2; ++pass) {
        // pass 0: the bounded entries (calls with a bound); pass 1: the others
        const TableEntry* best = nullptr;
        int dist = 0;
        for (int i = 0; i < TUNING_TABLE_SIZE; ++i) {
            const TableEntry& e = TUNING_TABLE[i];
            if (e.key_width != key_width || e.value_bytes != value_bytes) continue;
            if ((pass == 0) != (e.log2x >= 0)) continue;
            const int d = std::abs(e.log2n - ln) + 2 * std::abs(e.log2m - lm) + (pass == 0 ? 2 * std::abs(e.log2x - lx) : 0);
            if (!best || d < dist) {
                best = &e;
                dist = d;
            }
        }
        if (best) return best;
    }
    return nullptr;
}

// The tuning a call of this shape runs with: the override of its slot if one is
// installed, else the table's.
Tuning choose_tuning(int key_width, int value_bytes, long long num_items, long long num_segments, long long max_len,
                     bool* from_override = nullptr) {
    const int slot = tuning_slot(key_width, value_bytes);
    {
        std::lock_guard<std::mutex> lk(g_tuning_mu);
        if (from_override) *from_override = g_override_on[slot];
        if (g_override_on[slot]) return g_override[slot];
    }
    const TableEntry* e = table_lookup(key_width, value_bytes, num_items, num_segments, max_len);
    return e ? e->tu : DEFAULT_TUNING;
}

// Per-device CU count and per-kernel residency, queried once and cached.
std::mutex g_occ_mu;
std::map<std::pair<int, const void*>, int> g_occ;

int current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    return dev;
}

int device_cus() {
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(g_occ_mu);
    auto it = g_occ.find({dev, nullptr});
    if (it != g_occ.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    g_occ[{dev, nullptr}] = cus;
    return cus;
}

template <typename K>
int resident_blocks(K kernel, int threads) {
    const int cus = device_cus();
    const int dev = current_device();
    const void* key = (const void*)kernel;
    {
        std::lock_guard<std::mutex> lk(g_occ_mu);
        auto it = g_occ.find({dev, key});
        if (it != g_occ.end()) return it->second * cus;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    std::lock_guard<std::mutex> lk(g_occ_mu);
    g_occ[{dev, key}] = per_cu;
    return per_cu * cus;
}

// Second stream and fork/join events of the batched form's `concurrent` mode,
// one set per (host thread, device, caller stream), so that independent
// caller streams never share (and so serialise on) one second stream. The
// streams keep the default priority. A thread's sets are destroyed when the
// thread ends.
struct SideStreams {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t join = nullptr;
};

// Streams of the per-segment baseline (segsort_multistream_sort), per device:
// a fork event, and one stream and join event per fan-out slot.
struct FanStreams {
    hipEvent_t fork = nullptr;
    std::vector<hipStream_t> s;
    std::vector<hipEvent_t> join;
};

struct SidePool {
    std::map<std::pair<int, hipStream_t>, SideStreams> sets;
    std::map<int, hipStream_t> capture;  // the graph-replay mode's private capture stream, per device
    std::map<int, FanStreams> fan;
    ~SidePool() {
        for (auto& kv : sets) {
            if (kv.second.s) (void)hipStreamDestroy(kv.second.s);
            if (kv.second.fork) (void)hipEventDestroy(kv.second.fork);
            if (kv.second.join) (void)hipEventDestroy(kv.second.join);
        }
        for (auto& kv : capture)
            if (kv.second) (void)hipStreamDestroy(kv.second);
        for (auto& kv : fan) {
            if (kv.second.fork) (void)hipEventDestroy(kv.second.fork);
            for (hipStream_t s : kv.second.s) (void)hipStreamDestroy(s);
            for (hipEvent_t e : kv.second.join) (void)hipEventDestroy(e);
        }
    }
};

SidePool& side_pool() {
    static thread_local SidePool pool;
    return pool;
}

constexpr size_t SIDE_SETS_MAX = 32;  // per thread; beyond this a call runs on the caller's stream alone

// The second stream for a call on `st`, or nullptr when the call must run on
// st alone: creation failed, the thread already holds SIDE_SETS_MAX sets, or
// the second stream is still part of a stream capture that is not st's own
// (a stream that joined a capture stays in it until that capture ends).
SideStreams* side_streams(hipStream_t st, bool capturing) {
    SidePool& pool = side_pool();
    const std::pair<int, hipStream_t> key{current_device(), st};
    auto it = pool.sets.find(key);
    if (it == pool.sets.end()) {
        if (pool.sets.size() >= SIDE_SETS_MAX) return nullptr;
        SideStreams ss;
        const bool made = hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) == hipSuccess &&
                          hipEventCreateWithFlags(&ss.join, hipEventDisableTiming) == hipSuccess &&
                          hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking) == hipSuccess;
        if (!made) {
            if (ss.s) (void)hipStreamDestroy(ss.s);
            if (ss.fork) (void)hipEventDestroy(ss.fork);
            if (ss.join) (void)hipEventDestroy(ss.join);
            return nullptr;
        }
        it = pool.sets.emplace(key, ss).first;
    }
    SideStreams* ss = &it->second;
    hipStreamCaptureStatus side_cs = hipStreamCaptureStatusNone;
    unsigned long long side_id = 0;
    if (hipStreamGetCaptureInfo(ss->s, &side_cs, &side_id) != hipSuccess) return nullptr;
    if (side_cs == hipStreamCaptureStatusNone) return ss;
    if (!capturing) return nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    unsigned long long id = 0;
    if (hipStreamGetCaptureInfo(st, &cs, &id) != hipSuccess || cs != hipStreamCaptureStatusActive) return nullptr;
    return id == side_id ? ss : nullptr;
}

constexpr int FAN_MAX = 32;  // streams of the per-segment baseline (the reference's loops use up to 32)

// The calling thread's fan-out streams on the current device, at least n of
// them (n <= FAN_MAX), created on first use with the default priority.
FanStreams* fan_streams(int n) {
    FanStreams& f = side_pool().fan[current_device()];
    if (!f.fork && hipEventCreateWithFlags(&f.fork, hipEventDisableTiming) != hipSuccess) {
        f.fork = nullptr;
        return nullptr;
    }
    while ((int)f.s.size() < n) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(s);
            return nullptr;
        }
        f.s.push_back(s);
        f.join.push_back(e);
    }
    return &f;
}

// ---- hipGraph capture of the batched sequence ---------------------------------

// Everything a captured sequence depends on; an identical call replays the graph.
struct GraphKey {
    const void* p[8];  // temp, keys in/out, values in/out, begin, end, stream
    int n, s, kw, vb, dev;
    int bounded;  // the caller's length bound replaced the MSD tier by the oversize fallback
    ssort::KeyMode km;
    Tuning tu;
};

struct GraphEntry {
    GraphKey key;
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipEvent_t done;  // recorded after every replay: eviction waits on it before destroying
    uint64_t last_use;
};

hipError_t replay(GraphEntry& g, hipStream_t st) {
    hipError_t e = hipGraphLaunch(g.exec, st);
    if (e != hipSuccess) return e;
    return hipEventRecord(g.done, st);
}

std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs;
uint64_t g_graph_clock = 0;
int g_graph_mode = -1;  // -1: read SEGSORT_GRAPH once (default off), 0 off, 1 on
constexpr size_t GRAPH_CACHE = 16;

bool graphs_enabled() {
    if (g_graph_mode < 0) {
        const char* e = getenv("SEGSORT_GRAPH");
        g_graph_mode = (e && e[0] == '1') ? 1 : 0;
    }
    return g_graph_mode == 1;
}

bool same_key(const GraphKey& a, const GraphKey& b) { return memcmp(&a, &b, sizeof(GraphKey)) == 0; }

// Run `enqueue` on st: replayed from a cached graph when possible, captured on a
// private stream on first use, launched directly while profiling or when the
// caller's stream is itself being captured.
template <typename F>
hipError_t run_graph(const GraphKey& key, hipStream_t st, F&& enqueue) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (st) {
        hipError_t e = hipStreamIsCapturing(st, &cs);
        if (e != hipSuccess) return e;
    }
    if (!graphs_enabled() || Profiler::get().on() || cs != hipStreamCaptureStatusNone)
        return enqueue(st, cs != hipStreamCaptureStatusNone);
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (auto& g : g_graphs) {
        if (same_key(g.key, key)) {
            g.last_use = ++g_graph_clock;
            return replay(g, st);
        }
    }
    // one capture stream per thread and device (a stream belongs to the device
    // that was current when it was created)
    hipStream_t& cap = side_pool().capture[key.dev];
    if (!cap) {
        hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
        if (e != hipSuccess) {
            cap = nullptr;
            return e;
        }
    }
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = enqueue(cap, true);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cap, &graph);
    if (le != hipSuccess || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return le != hipSuccess ? le : e;
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return e;
    }
    hipEvent_t done = nullptr;
    e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return e;
    }
    if (g_graphs.size() >= GRAPH_CACHE) {
        auto victim = std::min_element(g_graphs.begin(), g_graphs.end(),
                                       [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
        (void)hipEventSynchronize(victim->done);  // its last replay may still be running
        (void)hipGraphExecDestroy(victim->exec);
        (void)hipGraphDestroy(victim->graph);
        (void)hipEventDestroy(victim->done);
        g_graphs.erase(victim);
    }
    g_graphs.push_back({key, graph, exec, done, ++g_graph_clock});
    return replay(g_graphs.back(), st);
}

// ---- per-(key, value) configuration -------------------------------------------

// LDS tile classes: (threads, items per thread). The largest tile also sets the
// chunk of the large tier and of the plain form's per-segment LSD.
template <typename U, typename V, bool HAS_V>
struct Cfg {
    static constexpr int ITEM_BYTES = int(sizeof(U)) + (HAS_V ? int(sizeof(V)) : 0);
    static constexpr int NT0 = 256, IPT0 = 4;
    static constexpr int NT1 = 256, IPT1 = 16;
    static constexpr int NT2 = 512, IPT2 = (sizeof(U) == 4 && ITEM_BYTES <= 8) ? 16 : 8;
    static constexpr int NT3 = 1024, IPT3 = ITEM_BYTES <= 8 ? 16 : 8;
    static constexpr int CAP[4] = {NT0 * IPT0, NT1 * IPT1, NT2 * IPT2, NT3 * IPT3};
    static constexpr int TILE_MAX = NT3 * IPT3;
    // chunks of the multi-workgroup MSD level (the largest tile: fewer chunks keep the per-range scan short)
    static constexpr int NTC = NT3, IPTC = IPT3, CHUNK = NTC * IPTC;
    static constexpr int NPOS = int(sizeof(U));  // 8-bit digit positions of a key
};

constexpr int KIND_WAVE = 0;   // one range per wave
constexpr int KIND_TILE = 1;   // + i: LDS tile class i (0..3)
constexpr int KIND_GROUP = 5;  // + g: 64 / (4 << g) ranges of <= 4 << g items per wave (g = 0..3)

struct Plan {
    ClassCaps cc;
    int kind[MAX_TILE_CLASSES];
    int nclass;
    int tile_max;
    bool lpt;   // batched form: length buckets for the 16-item tile classes (TUNE_LPT)
    bool half;  // the 4096-item tile's half-depth build (not TUNE_NO_HALF)
};

// Classes in increasing capacity: the enabled wave-group classes, the wave
// tier (if its cap is above them), then the enabled LDS tile classes.
template <typename C>
Plan make_plan(const Tuning& tu) {
    Plan p;
    p.nclass = 0;
    p.lpt = (tu.tile_enable & TUNE_LPT) != 0;
    p.half = (tu.tile_enable & TUNE_NO_HALF) == 0;
    int last = 1;
    for (int g = 0; g < 4; ++g) {
        if (!(tu.tile_enable & (0x10 << g))) continue;
        p.cc.cap[p.nclass] = 4 << g;
        p.kind[p.nclass++] = KIND_GROUP + g;
        last = 4 << g;
    }
    const int wave_cap = std::min(512, std::max(2, tu.wave_cap));
    const int enable = tu.tile_enable & 0xF;
    if (wave_cap > last) {
        int tiles = 0;  // tile classes above the wave cap
        for (int i = 0; i < 4; ++i) tiles += (enable & (1 << i)) && C::CAP[i] > wave_cap;
        if ((tu.tile_enable & TUNE_SPLIT_WAVE) && wave_cap > 256 && last < 256 &&
            p.nclass + 2 + tiles <= MAX_TILE_CLASSES) {
            p.cc.cap[p.nclass] = 256;
            p.kind[p.nclass++] = KIND_WAVE;
        }
        p.cc.cap[p.nclass] = wave_cap;
        p.kind[p.nclass++] = KIND_WAVE;
        last = wave_cap;
    }
    for (int i = 0; i < 4; ++i) {
        if (!(enable & (1 << i))) continue;
        if (C::CAP[i] <= last) continue;
        p.cc.cap[p.nclass] = C::CAP[i];
        p.kind[p.nclass++] = KIND_TILE + i;
    }
    p.cc.nclass = p.nclass;
    for (int c = p.nclass; c < MAX_TILE_CLASSES; ++c) p.cc.cap[c] = 0;
    p.tile_max = p.cc.cap[p.nclass - 1];  // ranges above this go to the MSD (or per-segment LSD) tier
    return p;
}

struct Layout {
    size_t ctr, tq[MAX_TILE_CLASSES], lq[MAX_LEVELS + 1], hist[2], cinfo[2], tk[2], tv[2];
    size_t bq[MAX_TILE_CLASSES], bcap[MAX_TILE_CLASSES];  // length buckets of the large tile classes
    size_t total;
    size_t tq_cap[MAX_TILE_CLASSES];
    size_t lq_cap;
    size_t chunks;
};

template <typename C, typename U, typename V, bool HAS_V>
Layout make_layout(const Plan& p, int num_items, int num_segments, int levels, bool plain) {
    Layout L;
    memset(&L, 0, sizeof(L));
    size_t at = 0;
    auto take = [&](size_t bytes) {
        size_t o = at;
        at += align256(bytes);
        return o;
    };
    const size_t N = (size_t)num_items, S = (size_t)num_segments;
    L.lq_cap = N / (size_t)(p.tile_max + 1) + 1;
    L.chunks = N / (size_t)C::CHUNK + L.lq_cap + 1;
    // every partition (level 0 or deeper) queues at most RADIX tile children; the
    // ranges partitioned at one depth are disjoint and larger than a tile, and
    // their children are disjoint non-empty ranges (at most N of them)
    const size_t bucket_bound =
        plain ? 0 : std::min((size_t)RADIX * L.lq_cap, std::max<size_t>(N, 1)) * (size_t)C::NPOS;
    L.ctr = take(sizeof(Counters));
    for (int c = 0; c < p.nclass; ++c) {
        L.tq_cap[c] = S + bucket_bound;
        L.tq[c] = take(L.tq_cap[c] * sizeof(TileItem));
    }
    // TUNE_LPT: the batched form's large tile classes (16 items per thread) take
    // their binned ranges longest first: LPT_BUCKETS length buckets, each able to
    // hold every binned range of the class (at most min(S, N / shortest length))
    for (int c = 0; c < p.nclass && !plain && p.lpt; ++c) {
        if (p.kind[c] < KIND_TILE + 1 || p.kind[c] > KIND_TILE + 3) continue;
        const size_t lo = c > 0 ? (size_t)p.cc.cap[c - 1] + 1 : 2;
        L.bcap[c] = std::min(S, N / lo + 1);
        L.bq[c] = take(L.bcap[c] * LPT_BUCKETS * sizeof(TileItem));
    }
    if (plain) {
        L.lq_cap = std::max<size_t>(std::min(S, L.lq_cap), 1);
        L.lq[MAX_LEVELS] = take(L.lq_cap * sizeof(LargeItem));
    } else if (levels > 0) {
        // the two multi-workgroup levels and the depth-first tier; a level's
        // ranges are disjoint, so its chunks (a whole number of partition tiles
        // each, at most MSD_MAX_CHUNKS per range) number at most N / CHUNK plus one
        // per range
        L.lq[MSD_LEVEL0] = take(L.lq_cap * sizeof(LargeItem));
        L.lq[MSD_LEVEL1] = take(L.lq_cap * sizeof(LargeItem));
        L.lq[MSD_LOCAL] = take(L.lq_cap * sizeof(LargeItem));
        // segments above the largest tile in a call that bounded its lengths below
        // it (a wrong bound), for the one-workgroup-per-segment fallback
        L.lq[MAX_LEVELS] = take(L.lq_cap * sizeof(LargeItem));
        for (int l = 0; l < 2; ++l) {
            L.hist[l] = take(L.chunks * RADIX * sizeof(uint32_t));
            L.cinfo[l] = take(L.chunks * sizeof(ChunkInfo));
        }
    }
    const bool need_tmp = plain ? (N > (size_t)p.tile_max) : (levels > 0);
    for (int i = 0; i < 2; ++i) {
        L.tk[i] = need_tmp ? take(N * sizeof(U)) : 0;
        L.tv[i] = (need_tmp && HAS_V) ? take(N * sizeof(V)) : 0;
    }
    L.total = at + 255;  // slack for aligning the caller's pointer
    return L;
}

template <typename U, typename V, bool HAS_V>
struct Launcher {
    using C = Cfg<U, V, HAS_V>;

    static constexpr double PAIR_BYTES = 2.0 * C::ITEM_BYTES;  // one read + one write per element

    template <int NT, int IPT>
    static void launch_tile(Bufs b, Sched sc, int cls, KeyMode km, hipStream_t st, int grid, bool persistent,
                            int grid_mult, bool lpt, bool half) {
        // the length-bucket builds exist for the 16-item tiles, the half-depth
        // ones for those of at most 4096 items
        auto k = k_tile_tier<NT, IPT, U, V, HAS_V, false, false>;
        if constexpr (IPT > 4 && NT * IPT <= 4096) {
            k = lpt ? (half ? k_tile_tier<NT, IPT, U, V, HAS_V, true, true> : k_tile_tier<NT, IPT, U, V, HAS_V, true, false>)
                    : (half ? k_tile_tier<NT, IPT, U, V, HAS_V, false, true> : k_tile_tier<NT, IPT, U, V, HAS_V, false, false>);
        } else if constexpr (IPT > 4) {
            if (lpt) k = k_tile_tier<NT, IPT, U, V, HAS_V, true, false>;
        }
        // persistent: every resident slot, but no more workgroups than the class
        // can have ranges (`grid` > 0: that bound)
        int g = persistent ? resident_blocks(k, NT) * grid_mult : grid;
        if (persistent && grid > 0) g = std::min(g, grid);
        char name[32];
        snprintf(name, sizeof(name), "tile_%dx%d", NT, IPT);
        if (g > 0) SEGSORT_LAUNCH(name, st, cls, PAIR_BYTES, 0.0, (k<<<g, NT, 0, st>>>(b, sc, cls, km, persistent)));
    }

    // Persistent grid, or (plain form) enough 256-thread blocks for `grid` ranges.
    template <int G>
    static void launch_group(Bufs b, Sched sc, int cls, KeyMode km, hipStream_t st, int grid, bool persistent,
                             int grid_mult) {
        auto k = k_group_tier<G, U, V, HAS_V>;
        constexpr int PER_BLOCK = 4 * (WAVE / G);
        int g = persistent ? resident_blocks(k, 256) * grid_mult : (grid + PER_BLOCK - 1) / PER_BLOCK;
        if (persistent && grid > 0) g = std::min(g, (grid + PER_BLOCK - 1) / PER_BLOCK);
        char name[32];
        snprintf(name, sizeof(name), "group_%d", G);
        if (g > 0) SEGSORT_LAUNCH(name, st, cls, PAIR_BYTES, 0.0, (k<<<g, 256, 0, st>>>(b, sc, cls, km)));
    }

    // wave_max: the class's longest range (for the wave tier's build). grid: the
    // ranges to sort (non-persistent, the plain form), or (persistent) a bound on
    // them that caps the grid (0: none).
    static void launch_class(int kind, Bufs b, Sched sc, int cls, KeyMode km, hipStream_t st, int grid, bool persistent,
                             int grid_mult, int wave_max, bool lpt, bool half) {
        switch (kind) {
        case 0: {
            // ranges of at most 256 items: the smaller-LDS, higher-occupancy build
            auto k = wave_max <= 256 ? k_wave_tier<U, V, HAS_V, 256> : k_wave_tier<U, V, HAS_V, 512>;
            int g = persistent ? resident_blocks(k, 256) * grid_mult : (grid + 3) / 4;
            if (persistent && grid > 0) g = std::min(g, (grid + 3) / 4);
            if (g > 0)
                SEGSORT_LAUNCH("wave_tier", st, cls, PAIR_BYTES, 0.0,
                               (k<<<g, 256, 0, st>>>(b, sc, cls, km, persistent ? 1 : 0, WAVE_RANK_MAX)));
            break;
        }
        case 1: launch_tile<C::NT0, C::IPT0>(b, sc, cls, km, st, grid, persistent, grid_mult, lpt, half); break;
        case 2: launch_tile<C::NT1, C::IPT1>(b, sc, cls, km, st, grid, persistent, grid_mult, lpt, half); break;
        case 3: launch_tile<C::NT2, C::IPT2>(b, sc, cls, km, st, grid, persistent, grid_mult, lpt, half); break;
        case 4: launch_tile<C::NT3, C::IPT3>(b, sc, cls, km, st, grid, persistent, grid_mult, lpt, half); break;
        case KIND_GROUP + 0: launch_group<4>(b, sc, cls, km, st, grid, persistent, grid_mult); break;
        case KIND_GROUP + 1: launch_group<8>(b, sc, cls, km, st, grid, persistent, grid_mult); break;
        case KIND_GROUP + 2: launch_group<16>(b, sc, cls, km, st, grid, persistent, grid_mult); break;
        case KIND_GROUP + 3: launch_group<32>(b, sc, cls, km, st, grid, persistent, grid_mult); break;
        }
    }

    // The classes of one queue set (queue q0 + c for class c), smallest first;
    // the length buckets belong to the binned ranges' set (q0 = 0).
    // A wave tier of ranges up to 256 items followed by the 256 x 4 tile class
    // runs as one launch (k_wave_tile_tier), for 32-bit keys together with the
    // wave-group classes (all enabled ones come first in the plan). Measured:
    // 2^20 segments of U[1,32] f32 keys about 6% faster with the groups inside;
    // with 64-bit keys the shared kernel's register budget made the groups
    // slower (9.6% at 5 waves per SIMD), so those keep their own launches.
    //
    // A class whose ranges hold at least lo items has at most N / lo of them, so
    // its persistent grid is capped there: a small call launches few workgroups
    // (an empty 16K-workgroup launch takes about 4.5 us here, a 2K one 2.5 us).
    static int max_ranges(const Plan& p, int c, long long n_items) {
        const long long lo = c > 0 ? p.cc.cap[c - 1] + 1 : 2;
        return (int)std::min<long long>(n_items / lo + 1, 0x7FFFFFFF);
    }

    static void launch_classes(const Plan& p, Bufs b, Sched sc, int q0, KeyMode km, hipStream_t st, int grid_mult,
                               long long n_items, uint32_t only = ~0u) {  // only: bit c launches class c
        int w = 0;  // the wave class, after the group classes it takes in
        if (sizeof(U) == 4)
            while (w < p.nclass && p.kind[w] >= KIND_GROUP) ++w;
        else
            while (w < p.nclass && p.kind[w] != KIND_WAVE) ++w;
        const int g0 = sizeof(U) == 4 ? 0 : w;  // first class of the shared launch
        const bool merged = w + 1 < p.nclass && p.kind[w] == KIND_WAVE && p.cc.cap[w] <= 256 &&
                            p.kind[w + 1] == KIND_TILE + 0;
        for (int c = 0; c < p.nclass; ++c) {
            if (!(only >> c & 1)) continue;
            if (merged && c >= g0 && c < w) continue;  // sorted by the shared launch below
            if (merged && c == w) {
                GroupQueues gq = {{-1, -1, -1, -1}};
                uint32_t more = 1u << (q0 + c + 1);
                for (int cg = g0; cg < w; ++cg) {
                    gq.q[p.kind[cg] - KIND_GROUP] = q0 + cg;
                    more |= 1u << (q0 + cg);
                }
                auto k = k_wave_tile_tier<U, V, HAS_V, 256, C::IPT0>;
                // workgroups the shared launch can use: four waves per block for the
                // wave tier, 4 * 64 / G ranges per block for a group class, one range
                // per block for the tiles
                int cap = std::max((max_ranges(p, w, n_items) + 3) / 4, max_ranges(p, w + 1, n_items));
                for (int cg = g0; cg < w; ++cg) {
                    const int per = 4 * (WAVE / (4 << (p.kind[cg] - KIND_GROUP)));
                    cap = std::max(cap, (max_ranges(p, cg, n_items) + per - 1) / per);
                }
                const int g = std::min(resident_blocks(k, 256) * grid_mult, cap);
                SEGSORT_LAUNCHM("wave_tile", st, q0 + c, more, PAIR_BYTES,
                                (k<<<g, 256, 0, st>>>(b, sc, q0 + c, q0 + c + 1, km, WAVE_RANK_MAX, gq)));
                ++c;
                continue;
            }
            launch_class(p.kind[c], b, sc, q0 + c, km, st, max_ranges(p, c, n_items), true, grid_mult, p.cc.cap[c],
                         p.lpt && q0 == 0, p.half);
        }
    }

    // In profiling mode: read the call's element counts back and attribute bytes.
    static void profile_attribute(size_t from, const Sched& sc, int nclass, hipStream_t st) {
        if (!Profiler::get().on()) return;
        Counters h;
        if (hipMemcpyAsync(&h, sc.ctr, sizeof(Counters), hipMemcpyDeviceToHost, st) != hipSuccess) return;
        if (hipStreamSynchronize(st) != hipSuccess) return;
        (void)nclass;
        Profiler::get().attribute(from, h.tile_items, MAX_QUEUES, h.level_items, MAX_LEVELS + 1);
    }

    // split (bit c: class c): the MSD tier queues class c's children behind its
    // binned ranges (from entry S on) as a separate queue, so that the binned
    // ranges can be sorted while the MSD tier runs; the other classes' children
    // join their binned ranges' queue.
    static Sched make_sched(char* base, const Layout& L, const Plan& p, bool plain, uint32_t split = 0,
                            size_t S = 0) {
        Sched sc;
        memset(&sc, 0, sizeof(sc));
        sc.ctr = (Counters*)(base + L.ctr);
        for (int c = 0; c < p.nclass; ++c) sc.tq[c] = (TileItem*)(base + L.tq[c]);
        if (plain) {
            sc.lq[MAX_LEVELS] = (LargeItem*)(base + L.lq[MAX_LEVELS]);
        } else {
            for (int l = 0; l <= MAX_LEVELS; ++l) sc.lq[l] = L.lq[l] ? (LargeItem*)(base + L.lq[l]) : nullptr;
            for (int l = 0; l < 2; ++l) {
                sc.hist[l] = L.hist[l] ? (uint32_t*)(base + L.hist[l]) : nullptr;
                sc.cinfo[l] = L.cinfo[l] ? (ChunkInfo*)(base + L.cinfo[l]) : nullptr;
            }
        }
        sc.chunk = C::CHUNK;
        for (int c = 0; c < p.nclass; ++c) sc.tq_cap[c] = (uint32_t)std::min<size_t>(L.tq_cap[c], 0xFFFFFFFFu);
        for (int c = 0; c < p.nclass; ++c) {
            sc.bq[c] = L.bcap[c] ? (TileItem*)(base + L.bq[c]) : nullptr;
            sc.bcap[c] = (uint32_t)L.bcap[c];
        }
        for (int c = 0; c < MAX_TILE_CLASSES; ++c) sc.child_q[c] = c;
        for (int c = 0; c < p.nclass; ++c) {
            if (split >> c & 1) {
                sc.child_q[c] = MAX_TILE_CLASSES + c;
                const size_t own = std::min(S, L.tq_cap[c]);
                sc.tq[MAX_TILE_CLASSES + c] = sc.tq[c] + own;
                sc.tq_cap[MAX_TILE_CLASSES + c] = (uint32_t)std::min<size_t>(L.tq_cap[c] - own, 0xFFFFFFFFu);
                sc.tq_cap[c] = (uint32_t)own;
            }
        }
        sc.lq_cap = (uint32_t)L.lq_cap;
        sc.chunk_cap = (uint32_t)L.chunks;
        return sc;
    }

    static Bufs make_bufs(char* base, const Layout& L, const void* ki, void* ko, const void* vi, void* vo) {
        Bufs b;
        b.k[BUF_IN] = ki;
        b.k[BUF_OUT] = ko;
        b.v[BUF_IN] = vi;
        b.v[BUF_OUT] = vo;
        for (int i = 0; i < 2; ++i) {
            b.k[BUF_TMP_A + i] = L.tk[i] ? base + L.tk[i] : nullptr;
            b.v[BUF_TMP_A + i] = L.tv[i] ? base + L.tv[i] : nullptr;
        }
        return b;
    }

    // MSD levels a call launches: 0 when no segment can exceed the largest tile,
    // 1 when one can (multi-workgroup level 0 and the depth-first tier), 2 when
    // in addition the mean segment length is at least a quarter of MSD_BIG_TILES
    // partition tiles, so that a child of level 0 holding more than MSD_BIG_TILES
    // tiles (keys skewed in their top digit) is plausible. Below that mean such a
    // child is still sorted, by the depth-first tier, and calls whose long
    // segments are a small share (the ragged mixes) do not pay for two launches
    // that would find nothing to do (measured: 2^24 items in 2^16 lognormal
    // segments 0.203 -> 0.219 ms with them).
    static int levels_for(const KeyMode& km, int num_items, int num_segments, const Plan& p) {
        (void)km;
        if (num_items <= p.tile_max) return 0;
        const long long big = (long long)MSD_BIG_TILES * C::CHUNK;
        const long long s = num_segments > 0 ? num_segments : 1;
        return (num_items > big && 4ll * num_items >= big * s) ? 2 : 1;
    }

    // Launch the batched sequence (bin, MSD levels, tile tiers) on st.
    // With `fallback` (levels is then 0) the call's caller bounded its segment
    // lengths by the largest tile: the MSD tier is not launched, and a segment
    // that breaks the bound is sorted by one workgroup (k_seg_lsd) at the end.
    static hipError_t enqueue_batched(char* base, const Layout& L, const Plan& p, const Tuning& tu, int levels,
                                      const void* ki, void* ko, const void* vi, void* vo, int N, int S, const int* db,
                                      const int* de, KeyMode km, hipStream_t st, bool fallback = false) {
        // With `concurrent` and long segments possible, the
        // binned ranges are sorted on a second stream while the MSD tier (a few
        // hundred workgroups at most, latency-bound) partitions the long ones on
        // the caller's stream; the MSD tier's children are then sorted on the
        // caller's stream and the second stream is joined back. Measured on
        // MI355X back to back (ragged 2^24, two persistent workgroups per slot):
        // 0.1852 ms on one stream, 0.1821 with `concurrent` 1, 0.1808 with 2; but
        // 2^16 equal segments of 256 (no segment long) 0.0722, 0.0885 and 0.0834
        // ms, because the (then empty) MSD launches wait for CUs behind the
        // persistent class grids, and calls synchronised one at a time are slower
        // with it. The host cannot tell the cases apart, so the tuning table
        // chooses per signature (every stored entry keeps one stream).
        // Under stream capture the fork and join are recorded as well: the second
        // stream joins the capture and the graph gets parallel branches.
        // `concurrent` 2 forks only the 16-item tile classes: their binned ranges
        // are sorted on the second stream, and the smaller classes (binned ranges
        // and MSD children in one queue each) on the caller's stream after the
        // MSD tier, so that an empty MSD tier waits behind fewer workgroups (the
        // empty launches of both streams still contend, as measured above).
        uint32_t fork_mask = 0;
        if (levels > 0 && tu.concurrent == 1) fork_mask = (1u << p.nclass) - 1u;
        if (levels > 0 && tu.concurrent == 2)
            for (int c = 0; c < p.nclass; ++c)
                if (p.kind[c] >= KIND_TILE + 1 && p.kind[c] <= KIND_TILE + 3) fork_mask |= 1u << c;
        SideStreams* ss = nullptr;
        bool capturing = false;
        if (fork_mask) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(st, &cs) == hipSuccess) {
                capturing = cs != hipStreamCaptureStatusNone;
                ss = side_streams(st, capturing);  // nullptr: one stream it is
            }
        }
        const uint32_t split = ss != nullptr ? fork_mask : 0u;
        const Sched sc = make_sched(base, L, p, false, split, (size_t)S);
        const Bufs b = make_bufs(base, L, ki, ko, vi, vo);
        const size_t mark = Profiler::get().mark();
        const unsigned long long token = next_gate_token();  // the binning kernel clears the counters
        const int large_mode = fallback ? 2 : 1;  // k_bin: segments above the largest tile to MSD level 0 or the fallback
        auto kbin = p.lpt ? k_bin<U, V, HAS_V, true> : k_bin<U, V, HAS_V, false>;
        if (split && !capturing) {
            // the fork event completes with the binning kernel itself (no separate marker on the caller's stream)
            SEGSORT_LAUNCH("bin", st, -1, 0.0, 8.0 * S,
                           hipExtLaunchKernelGGL(kbin, dim3(bin_grid(S)), dim3(BIN_NT), 0, st, nullptr, ss->fork, 0,
                                                 db, de, S, b, sc, p.cc, km.end_bit, large_mode, bin_spt(S), token));
        } else {
            SEGSORT_LAUNCH("bin", st, -1, 0.0, 8.0 * S,
                           (kbin<<<bin_grid(S), BIN_NT, 0, st>>>(db, de, S, b, sc, p.cc, km.end_bit, large_mode,
                                                                 bin_spt(S), token)));
        }
        hipError_t e = hipSuccess;
        if (split) {
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if (capturing && (e = hipEventRecord(ss->fork, st)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(ss->s, ss->fork, 0)) != hipSuccess) return e;
            launch_classes(p, b, sc, 0, km, ss->s, tu.grid_mult, N, split);
            if ((e = hipEventRecord(ss->join, ss->s)) != hipSuccess) return e;
        }
        if (levels > 0) {
            // level 0 over the long segments, level 1 over their children of more
            // than MSD_BIG_TILES tiles (launched only when such a child can exist),
            // then the depth-first tier
            auto kh = k_large_hist<C::NTC, C::IPTC, U, V, HAS_V>;
            auto ks = k_large_scatter<C::NTC, C::IPTC, U, V, HAS_V>;
            auto kl = k_large_local<C::NT3, C::IPT3, U, V, HAS_V>;
            // children of level 0 above `big` items go to level 1, which runs only
            // when levels > 1; otherwise every child larger than a tile goes to the
            // depth-first tier
            const int big = levels > 1 ? MSD_BIG_TILES * C::CHUNK : 0x7FFFFFFF;
            // one workgroup per resident slot (measured: twice or three times as
            // many, mostly finding no chunk, only wait for CUs under the overlap)
            // (no more workgroups than chunks, or than depth-first ranges)
            const int nch = (int)std::min<size_t>(L.chunks, 0x7FFFFFFF);
            const int gh = std::min(resident_blocks(kh, C::NTC), nch);
            const int gs = std::min(resident_blocks(ks, C::NTC), nch);
            const int gl = std::min(resident_blocks(kl, C::NT3), (int)std::min<size_t>(L.lq_cap, 0x7FFFFFFF));
            SEGSORT_LAUNCH("large_hist", st, 100 + MSD_LEVEL0, (double)sizeof(U), 0.0,
                           (kh<<<gh, C::NTC, 0, st>>>(b, sc, MSD_LEVEL0, km)));
            SEGSORT_LAUNCH("large_scatter", st, 100 + MSD_LEVEL0, PAIR_BYTES, 0.0,
                           (ks<<<gs, C::NTC, 0, st>>>(b, sc, MSD_LEVEL0, km, p.cc, p.tile_max, big)));
            if (levels > 1) {
                // the full grids: keys skewed in their top digit leave most of a
                // segment to this level (one 2^24 segment of N(0,1) floats:
                // 0.625 -> 0.533 ms against half the scatter grid), and an empty
                // launch costs the same either way
                SEGSORT_LAUNCH("large_hist", st, 100 + MSD_LEVEL1, (double)sizeof(U), 0.0,
                               (kh<<<gh, C::NTC, 0, st>>>(b, sc, MSD_LEVEL1, km)));
                SEGSORT_LAUNCH("large_scatter", st, 100 + MSD_LEVEL1, PAIR_BYTES, 0.0,
                               (ks<<<gs, C::NTC, 0, st>>>(b, sc, MSD_LEVEL1, km, p.cc, p.tile_max, big)));
            }
            SEGSORT_LAUNCH("large_local", st, 100 + MSD_LOCAL, PAIR_BYTES + sizeof(U), 0.0,
                           (kl<<<gl, C::NT3, 0, st>>>(b, sc, MSD_LOCAL, km, p.cc, p.tile_max)));
        }
        if (split) {
            // the classes that did not fork (their queues hold the MSD children
            // too), then the forked classes' MSD children
            launch_classes(p, b, sc, 0, km, st, tu.grid_mult, N, ~split);
            launch_classes(p, b, sc, MAX_TILE_CLASSES, km, st, tu.grid_mult, N, split);
        } else {
            launch_classes(p, b, sc, 0, km, st, tu.grid_mult, N);
        }
        if (fallback) {
            SEGSORT_LAUNCH("oversize_lsd", st, 100 + MAX_LEVELS, PAIR_BYTES, 0.0,
                           (k_seg_lsd<C::NT3, C::IPT3, U, V, HAS_V><<<FALLBACK_GRID, C::NT3, 0, st>>>(b, sc, km)));
        }
        if (split && (e = hipStreamWaitEvent(st, ss->join, 0)) != hipSuccess) return e;
        profile_attribute(mark, sc, p.nclass, st);
        return hipGetLastError();
    }

    static int batched(void* d_temp, size_t* temp_bytes, const void* ki, void* ko, const void* vi, void* vo, int N,
                       int S, const int* db, const int* de, KeyMode km, hipStream_t st) {
        const Tuning tu = choose_tuning(sizeof(U), HAS_V ? sizeof(V) : 0, N, S, t_max_len);
        const Plan p = make_plan<C>(tu);
        const int levels = levels_for(km, N, S, p);
        // the caller's bound keeps every segment inside the largest tile: no MSD
        // launches (the layout keeps its buffers, for the fallback)
        const bool fallback = levels > 0 && t_max_len > 0 && t_max_len <= p.tile_max;
        const Layout L = make_layout<C, U, V, HAS_V>(p, N, S, levels, false);
        if (!d_temp) {
            *temp_bytes = L.total;
            return ok();
        }
        if (*temp_bytes < L.total) return fail(hipErrorInvalidValue, "segsort_sort: temp storage too small");
        if (N == 0 || S == 0) return ok();
        char* base = (char*)((((uintptr_t)d_temp) + 255) & ~uintptr_t(255));
        GraphKey key;
        memset(&key, 0, sizeof(key));
        const void* ptrs[] = {d_temp, ki, ko, vi, vo, db, de, (const void*)st};
        for (int i = 0; i < 8; ++i) key.p[i] = ptrs[i];
        key.n = N;
        key.s = S;
        key.km = km;
        key.kw = (int)sizeof(U);
        key.vb = HAS_V ? (int)sizeof(V) : 0;
        key.tu = tu;
        key.dev = current_device();
        key.bounded = fallback ? 1 : 0;
        const int run_levels = fallback ? 0 : levels;
        hipError_t e;
        if (t_no_alloc) {
            Tuning one = tu;
            one.concurrent = 0;
            e = enqueue_batched(base, L, p, one, run_levels, ki, ko, vi, vo, N, S, db, de, km, st, fallback);
        } else {
            e = run_graph(key, st, [&](hipStream_t cs, bool captured) {
                (void)captured;
                return enqueue_batched(base, L, p, tu, run_levels, ki, ko, vi, vo, N, S, db, de, km, cs, fallback);
            });
        }
        if (e != hipSuccess) return fail(e, "segsort_sort(batched): launch");
        return ok();
    }

    // The per-segment baseline: every segment sorted by its own single-segment
    // batched sequence, the segments dealt over `ns` streams round robin
    // (ncub.cu:113-145 over NUM_STREAMS streams, newthrust.cu:141-160) or each
    // to the stream with the fewest items so far (newthrust.cu:91-107, SCHEDULE).
    // Each stream owns one temp region sized for the longest segment; the
    // segment lengths are uploaded into the temp storage (one synchronisation of
    // st, so that the host array may be released), after which the streams fork
    // from st and join back into it.
    static int multistream(void* d_temp, size_t* temp_bytes, const void* ki, void* ko, const void* vi, void* vo,
                           int N, int S, const int* h_off, KeyMode km, int ns, int schedule, hipStream_t st) {
        int maxlen = 0;
        for (int i = 0; i < S; ++i) {
            const int a = h_off[i], e = h_off[i + 1];
            if (a < 0 || e < a || e > N)
                return fail(hipErrorInvalidValue, "segsort_multistream_sort: offsets not ascending inside [0, N]");
            maxlen = std::max(maxlen, e - a);
        }
        Tuning tu = choose_tuning(sizeof(U), HAS_V ? sizeof(V) : 0, std::max(maxlen, 1), 1, 0);
        tu.concurrent = 0;
        const Plan p = make_plan<C>(tu);
        const int levels = levels_for(km, std::max(maxlen, 1), 1, p);
        const Layout L = make_layout<C, U, V, HAS_V>(p, std::max(maxlen, 1), 1, levels, false);
        const int used = std::max(1, std::min(ns, S));
        const size_t lens_bytes = align256(sizeof(int) * ((size_t)S + 1));
        const size_t per = align256(L.total);
        if (!d_temp) {
            *temp_bytes = 255 + lens_bytes + (size_t)used * per;
            return ok();
        }
        if (*temp_bytes < 255 + lens_bytes + (size_t)used * per)
            return fail(hipErrorInvalidValue, "segsort_multistream_sort: temp storage too small");
        if (N == 0 || S == 0) return ok();
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (st && hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return fail(hipErrorStreamCaptureUnsupported, "segsort_multistream_sort: not capturable (it synchronises)");
        FanStreams* f = fan_streams(used);
        if (!f) return fail(hipErrorOutOfMemory, "segsort_multistream_sort: stream creation");
        char* base = (char*)((((uintptr_t)d_temp) + 255) & ~uintptr_t(255));
        int* d_lens = (int*)base;  // d_lens[0] = 0, d_lens[1 + i] = length of segment i
        std::vector<int> lens((size_t)S + 1);
        lens[0] = 0;
        for (int i = 0; i < S; ++i) lens[(size_t)i + 1] = h_off[i + 1] - h_off[i];
        hipError_t e = hipMemcpyAsync(d_lens, lens.data(), sizeof(int) * lens.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess) e = hipEventRecord(f->fork, st);
        for (int s = 0; e == hipSuccess && s < used; ++s) e = hipStreamWaitEvent(f->s[(size_t)s], f->fork, 0);
        if (e != hipSuccess) return fail(e, "segsort_multistream_sort: fork");
        std::vector<long long> load((size_t)used, 0);
        for (int i = 0; i < S; ++i) {
            const int n = lens[(size_t)i + 1];
            int s = i % used;
            if (schedule) s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[(size_t)s] += n;
            if (n == 0) continue;
            const size_t o = (size_t)h_off[i];
            char* bs = base + lens_bytes + (size_t)s * per;
            e = enqueue_batched(bs, L, p, tu, levels_for(km, n, 1, p), (const char*)ki + o * sizeof(U),
                                (char*)ko + o * sizeof(U), HAS_V ? (const char*)vi + o * sizeof(V) : nullptr,
                                HAS_V ? (char*)vo + o * sizeof(V) : nullptr, n, 1, d_lens, d_lens + 1 + i, km,
                                f->s[(size_t)s]);
            if (e != hipSuccess) return fail(e, "segsort_multistream_sort: launch");
        }
        for (int s = 0; e == hipSuccess && s < used; ++s) {
            e = hipEventRecord(f->join[(size_t)s], f->s[(size_t)s]);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, f->join[(size_t)s], 0);
        }
        if (e != hipSuccess) return fail(e, "segsort_multistream_sort: join");
        return ok();
    }

    static int plain(void* d_temp, size_t* temp_bytes, const void* ki, void* ko, const void* vi, void* vo, int N,
                     int S, const int* db, const int* de, KeyMode km, hipStream_t st) {
        const Tuning tu = choose_tuning(sizeof(U), HAS_V ? sizeof(V) : 0, N, S, t_max_len);
        const Plan p = make_plan<C>(tu);
        const Layout L = make_layout<C, U, V, HAS_V>(p, N, S, 0, true);
        if (!d_temp) {
            *temp_bytes = L.total;
            return ok();
        }
        if (*temp_bytes < L.total) return fail(hipErrorInvalidValue, "segsort_sort: temp storage too small");
        if (N == 0 || S == 0) return ok();
        char* base = (char*)((((uintptr_t)d_temp) + 255) & ~uintptr_t(255));
        const Sched sc = make_sched(base, L, p, true);
        const Bufs b = make_bufs(base, L, ki, ko, vi, vo);
        const size_t mark = Profiler::get().mark();
        SEGSORT_LAUNCH("bin", st, -1, 0.0, 8.0 * S,
                       (k_bin<U, V, HAS_V, false><<<bin_grid(S), BIN_NT, 0, st>>>(db, de, S, b, sc, p.cc, km.end_bit, 0,
                                                                                  bin_spt(S), next_gate_token())));
        SEGSORT_LAUNCH("bin_large", st, -1, 0.0, 8.0 * S,
                       (k_bin_large_ids<<<(S + 255) / 256, 256, 0, st>>>(db, de, S, sc, p.cc)));
        Counters h;
        hipError_t e = hipMemcpyAsync(&h, sc.ctr, sizeof(Counters), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return fail(e, "segsort_sort(plain): count read-back");
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(e, "segsort_sort(plain): sync");
        for (int c = 0; c < p.nclass; ++c)
            if (h.tile_count[c])
                launch_class(p.kind[c], b, sc, c, km, st, (int)h.tile_count[c], false, 1, p.cc.cap[c], false, p.half);
        const int nlarge = (int)h.level[MAX_LEVELS];
        if (nlarge > 0) {
            if (km.end_bit > km.begin_bit)
                SEGSORT_LAUNCH("seg_lsd", st, 100 + MAX_LEVELS, PAIR_BYTES, 0.0,
                               (k_seg_lsd<C::NT3, C::IPT3, U, V, HAS_V><<<nlarge, C::NT3, 0, st>>>(b, sc, km)));
        }
        profile_attribute(mark, sc, p.nclass, st);
        e = hipGetLastError();
        if (e != hipSuccess) return fail(e, "segsort_sort(plain): launch");
        return ok();
    }
};

// Fix-pass temp: counters | comp[2][M] | idx[2][M] | hist[RADIX][chunks]
constexpr int FIX_NT = 1024, FIX_IPT = 8, FIX_TILE = FIX_NT * FIX_IPT;

template <typename U, typename V, bool HAS_V>
int fixpass(void* d_temp, size_t* temp_bytes, const void* ki, void* ko, const void* vi, void* vo, int N, int S,
            const int* db, const int* de, KeyMode km, hipStream_t st) {
    const size_t M = (size_t)N;
    const size_t nch = M / FIX_TILE + 2;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        size_t o = at;
        at += align256(bytes);
        return o;
    };
    const size_t o_ctr = take(sizeof(Counters));
    const size_t o_c0 = take(M * 8), o_c1 = take(M * 8);
    const size_t o_i0 = take(M * 4), o_i1 = take(M * 4);
    const size_t o_h = take(nch * RADIX * 4);
    const size_t o_t = take(RADIX * 4);
    const size_t total = at + 255;
    if (!d_temp) {
        *temp_bytes = total;
        return ok();
    }
    if (*temp_bytes < total) return fail(hipErrorInvalidValue, "segsort_sort(fixpass): temp storage too small");
    if (N == 0 || S == 0) return ok();
    char* base = (char*)((((uintptr_t)d_temp) + 255) & ~uintptr_t(255));
    Counters* ctr = (Counters*)(base + o_ctr);
    unsigned long long* dmax = &ctr->level[0];
    unsigned long long* comp[2] = {(unsigned long long*)(base + o_c0), (unsigned long long*)(base + o_c1)};
    uint32_t* idx[2] = {(uint32_t*)(base + o_i0), (uint32_t*)(base + o_i1)};
    uint32_t* hist = (uint32_t*)(base + o_h);
    uint32_t* totals = (uint32_t*)(base + o_t);
    int segbits = 0;
    while (segbits < 31 && (1ll << segbits) < (long long)S) ++segbits;  // bit length of S-1
    const int cus = device_cus();
    SEGSORT_LAUNCH("init", st, -1, 0.0, 0.0, (k_init<<<1, 64, 0, st>>>(ctr)));
    SEGSORT_LAUNCH("fix_max", st, -1, 0.0, (double)N * sizeof(U),
                   (k_fix_max<U, 1024, 8><<<cus * 2, 1024, 0, st>>>((const U*)ki, db, de, S, km, dmax)));
    // With the whole 32-bit key as the order key, a composite's order-key bits
    // are the twiddled key: the passes carry the values (or nothing, keys only) and the
    // last one writes the output. A bit sub-range carries input indices and
    // gathers keys and values by them at the end. 64-bit values always gather.
    const bool direct = km.begin_bit == 0 && km.end_bit == 32 && (!HAS_V || sizeof(V) == 4);
    const int pay_mode = !direct ? FIX_PAY_INDEX : HAS_V ? FIX_PAY_VALUE : FIX_PAY_NONE;
    const int pay_bytes = pay_mode == FIX_PAY_NONE ? 0 : 4;
    SEGSORT_LAUNCH("fix_composite", st, -1, 0.0, (double)N * (sizeof(U) + 8 + 2 * pay_bytes),
                   (k_fix_composite<U, 256, 16><<<cus * 4, 256, 0, st>>>(
                       (const U*)ki, db, de, S, km, dmax, segbits, pay_mode, (const uint32_t*)vi, comp[0], idx[0],
                       (U*)ko, (uint32_t*)vo)));
    const int kb = km.end_bit - km.begin_bit;
    const int max_passes = (kb + segbits + RADIX_BITS - 1) / RADIX_BITS;
    auto kd = pay_mode == FIX_PAY_NONE ? k_fix_downsweep<FIX_NT, FIX_IPT, false> : k_fix_downsweep<FIX_NT, FIX_IPT, true>;
    const int gd = resident_blocks(kd, FIX_NT);
    uint32_t* kfin = direct ? (uint32_t*)ko : nullptr;
    uint32_t* vfin = direct && HAS_V ? (uint32_t*)vo : nullptr;
    for (int ps = 0; ps < max_passes; ++ps) {
        const int s = ps & 1;
        SEGSORT_LAUNCH("fix_upsweep", st, -1, 0.0, (double)N * 8,
                       (k_fix_upsweep<256, 8><<<cus * 4, 256, 0, st>>>(comp[s], db, de, S, dmax, segbits, ps,
                                                                       FIX_TILE, (int)nch, hist)));
        SEGSORT_LAUNCH("fix_spine", st, -1, 0.0, 0.0,
                       (k_fix_spine<256, 8><<<RADIX, 256, 0, st>>>(db, de, S, dmax, segbits, ps, FIX_TILE, (int)nch,
                                                                   hist, totals)));
        SEGSORT_LAUNCH("fix_downsweep", st, -1, 0.0, (double)N * (16 + 2 * pay_bytes),
                       (kd<<<gd, FIX_NT, 0, st>>>(comp[s], idx[s], comp[s ^ 1], idx[s ^ 1], db, de, S, dmax, segbits,
                                                 ps, (int)nch, hist, totals, kfin, vfin, km)));
    }
    if (direct) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(e, "segsort_sort(fixpass): launch");
        return ok();
    }
    Bufs b;
    memset(&b, 0, sizeof(b));
    b.k[BUF_IN] = ki;
    b.k[BUF_OUT] = ko;
    b.v[BUF_IN] = vi;
    b.v[BUF_OUT] = vo;
    SEGSORT_LAUNCH("fix_gather", st, -1, 0.0, (double)N * (4 + 2 * (sizeof(U) + (HAS_V ? sizeof(V) : 0))),
                   (k_fix_gather<U, V, HAS_V><<<cus * 8, 256, 0, st>>>(b, idx[0], idx[1], db, de, S, dmax, segbits)));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(e, "segsort_sort(fixpass): launch");
    return ok();
}

template <typename U, typename V, bool HAS_V>
int dispatch_algo(int algo, void* d_temp, size_t* temp_bytes, const void* ki, void* ko, const void* vi, void* vo,
                  int N, int S, const int* db, const int* de, KeyMode km, hipStream_t st) {
    switch (algo) {
    case SEGSORT_ALGO_BATCHED:
        return Launcher<U, V, HAS_V>::batched(d_temp, temp_bytes, ki, ko, vi, vo, N, S, db, de, km, st);
    case SEGSORT_ALGO_PLAIN:
        return Launcher<U, V, HAS_V>::plain(d_temp, temp_bytes, ki, ko, vi, vo, N, S, db, de, km, st);
    case SEGSORT_ALGO_FIXPASS:
        if constexpr (sizeof(U) == 4) return fixpass<U, V, HAS_V>(d_temp, temp_bytes, ki, ko, vi, vo, N, S, db, de, km, st);
        return fail(hipErrorInvalidValue, "segsort_sort(fixpass): 32-bit keys only");
    }
    return fail(hipErrorInvalidValue, "segsort_sort: unknown algo");
}

}  // namespace

extern "C" {

int segsort_version(void) { return 100; }

int segsort_set_graph_replay(int on) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    const int before = graphs_enabled() ? 1 : 0;
    g_graph_mode = on ? 1 : 0;
    return before;
}

int segsort_profile_enable(int on) {
    Profiler::get().enable(on != 0);
    return ok();
}

size_t segsort_profile_report(char* buf, size_t buflen) {
    const std::string r = Profiler::get().report();
    if (buf && buflen) {
        const size_t n = std::min(buflen - 1, r.size());
        memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return r.size() + 1;
}

const char* segsort_last_error(void) { return g_last_error.c_str(); }

int segsort_sort(void* d_temp, size_t* temp_bytes, const void* d_keys_in, void* d_keys_out, const void* d_vals_in,
                 void* d_vals_out, int num_items, int num_segments, const int* d_begin_offsets,
                 const int* d_end_offsets, int begin_bit, int end_bit, int key_type, int value_bytes, int descending,
                 int algo, void* stream) {
    if (!temp_bytes) return fail(hipErrorInvalidValue, "segsort_sort: temp_bytes is NULL");
    if (num_items < 0 || num_segments < 0) return fail(hipErrorInvalidValue, "segsort_sort: negative size");
    if (key_type < SEGSORT_KEY_U32 || key_type > SEGSORT_KEY_F64)
        return fail(hipErrorInvalidValue, "segsort_sort: bad key_type");
    if (value_bytes != 0 && value_bytes != 4 && value_bytes != 8)
        return fail(hipErrorInvalidValue, "segsort_sort: value_bytes must be 0, 4 or 8");
    const int kw = key_type >= SEGSORT_KEY_U64 ? 8 : 4;
    if (begin_bit < 0 || end_bit > kw * 8 || begin_bit > end_bit)
        return fail(hipErrorInvalidValue, "segsort_sort: bad bit range");
    if (d_temp && num_items > 0 && num_segments > 0) {
        if (!d_keys_in || !d_keys_out || !d_begin_offsets || !d_end_offsets)
            return fail(hipErrorInvalidValue, "segsort_sort: NULL buffer");
        if (value_bytes && (!d_vals_in || !d_vals_out)) return fail(hipErrorInvalidValue, "segsort_sort: NULL values");
    }
    KeyMode km;
    km.kind = (key_type % 3 == 0) ? KIND_UNSIGNED : (key_type % 3 == 1) ? KIND_SIGNED : KIND_FLOAT;
    km.desc = descending ? 1 : 0;
    km.begin_bit = begin_bit;
    km.end_bit = end_bit;
    hipStream_t st = (hipStream_t)stream;
#define SEGSORT_ARGS algo, d_temp, temp_bytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, num_items, num_segments, \
                     d_begin_offsets, d_end_offsets, km, st
    if (kw == 4) {
        if (value_bytes == 0) return dispatch_algo<uint32_t, uint32_t, false>(SEGSORT_ARGS);
        if (value_bytes == 4) return dispatch_algo<uint32_t, uint32_t, true>(SEGSORT_ARGS);
        return dispatch_algo<uint32_t, uint64_t, true>(SEGSORT_ARGS);
    }
    if (value_bytes == 0) return dispatch_algo<uint64_t, uint32_t, false>(SEGSORT_ARGS);
    if (value_bytes == 4) return dispatch_algo<uint64_t, uint32_t, true>(SEGSORT_ARGS);
    return dispatch_algo<uint64_t, uint64_t, true>(SEGSORT_ARGS);
#undef SEGSORT_ARGS
}

int segsort_multistream_sort(void* d_temp, size_t* temp_bytes, const void* d_keys_in, void* d_keys_out,
                             const void* d_vals_in, void* d_vals_out, int num_items, int num_segments,
                             const int* h_offsets, int begin_bit, int end_bit, int key_type, int value_bytes,
                             int descending, int num_streams, int schedule, void* stream) {
    if (!temp_bytes) return fail(hipErrorInvalidValue, "segsort_multistream_sort: temp_bytes is NULL");
    if (num_items < 0 || num_segments < 0) return fail(hipErrorInvalidValue, "segsort_multistream_sort: negative size");
    if (num_segments > 0 && !h_offsets) return fail(hipErrorInvalidValue, "segsort_multistream_sort: NULL offsets");
    if (num_streams < 1 || num_streams > FAN_MAX)
        return fail(hipErrorInvalidValue, "segsort_multistream_sort: num_streams must be 1..32");
    if (key_type < SEGSORT_KEY_U32 || key_type > SEGSORT_KEY_F64)
        return fail(hipErrorInvalidValue, "segsort_multistream_sort: bad key_type");
    if (value_bytes != 0 && value_bytes != 4 && value_bytes != 8)
        return fail(hipErrorInvalidValue, "segsort_multistream_sort: value_bytes must be 0, 4 or 8");
    const int kw = key_type >= SEGSORT_KEY_U64 ? 8 : 4;
    if (begin_bit < 0 || end_bit > kw * 8 || begin_bit > end_bit)
        return fail(hipErrorInvalidValue, "segsort_multistream_sort: bad bit range");
    if (d_temp && num_items > 0 && num_segments > 0) {
        if (!d_keys_in || !d_keys_out) return fail(hipErrorInvalidValue, "segsort_multistream_sort: NULL buffer");
        if (value_bytes && (!d_vals_in || !d_vals_out))
            return fail(hipErrorInvalidValue, "segsort_multistream_sort: NULL values");
    }
    KeyMode km;
    km.kind = (key_type % 3 == 0) ? KIND_UNSIGNED : (key_type % 3 == 1) ? KIND_SIGNED : KIND_FLOAT;
    km.desc = descending ? 1 : 0;
    km.begin_bit = begin_bit;
    km.end_bit = end_bit;
    hipStream_t st = (hipStream_t)stream;
#define SEGSORT_ARGS d_temp, temp_bytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, num_items, num_segments, \
                     h_offsets, km, num_streams, schedule, st
    if (kw == 4) {
        if (value_bytes == 0) return Launcher<uint32_t, uint32_t, false>::multistream(SEGSORT_ARGS);
        if (value_bytes == 4) return Launcher<uint32_t, uint32_t, true>::multistream(SEGSORT_ARGS);
        return Launcher<uint32_t, uint64_t, true>::multistream(SEGSORT_ARGS);
    }
    if (value_bytes == 0) return Launcher<uint64_t, uint32_t, false>::multistream(SEGSORT_ARGS);
    if (value_bytes == 4) return Launcher<uint64_t, uint32_t, true>::multistream(SEGSORT_ARGS);
    return Launcher<uint64_t, uint64_t, true>::multistream(SEGSORT_ARGS);
#undef SEGSORT_ARGS
}

}  // extern "C"

namespace {

// Scratch of segsort_sort_noalloc: one static array per GPU (each device holds
// its own copy for the life of the process). In host code `scratch` names a
// host-side stand-in, not device memory: only the address that
// hipGetSymbolAddress returns for the current device is passed to kernels.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> device address of `scratch`

hipError_t scratch_address(void** out) {
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        it = g_scratch_ptr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return hipSuccess;
}

struct NoAllocScope {
    NoAllocScope() { t_no_alloc = true; }
    ~NoAllocScope() { t_no_alloc = false; }
};

}  // namespace

extern "C" {

size_t segsort_noalloc_max_bytes(void) { return SCRATCH_BYTES; }

int segsort_sort_hinted(void* d_temp, size_t* temp_bytes, const void* d_keys_in, void* d_keys_out,
                        const void* d_vals_in, void* d_vals_out, int num_items, int num_segments,
                        const int* d_begin_offsets, const int* d_end_offsets, int begin_bit, int end_bit, int key_type,
                        int value_bytes, int descending, int algo, int max_segment_length, void* stream) {
    if (max_segment_length < 0) return fail(hipErrorInvalidValue, "segsort_sort_hinted: negative max_segment_length");
    struct Scope {
        int prev;
        explicit Scope(int m) : prev(t_max_len) { t_max_len = m; }
        ~Scope() { t_max_len = prev; }
    } scope(max_segment_length);
    return segsort_sort(d_temp, temp_bytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, num_items, num_segments,
                        d_begin_offsets, d_end_offsets, begin_bit, end_bit, key_type, value_bytes, descending, algo,
                        stream);
}
int segsort_sort_noalloc(const void* d_keys_in, void* d_keys_out, const void* d_vals_in, void* d_vals_out,
                         int num_items, int num_segments, const int* d_begin_offsets, const int* d_end_offsets,
                         int begin_bit, int end_bit, int key_type, int value_bytes, int descending, void* stream) {
    NoAllocScope scope;
    size_t need = 0;
    int rc = segsort_sort(nullptr, &need, d_keys_in, d_keys_out, d_vals_in, d_vals_out, num_items, num_segments,
                          d_begin_offsets, d_end_offsets, begin_bit, end_bit, key_type, value_bytes, descending,
                          SEGSORT_ALGO_BATCHED, stream);
    if (rc) return rc;
    if (need > SCRATCH_BYTES) {
        char msg[160];
        snprintf(msg, sizeof(msg), "segsort_sort_noalloc: the call needs %zu bytes of scratch, more than the %zu of "
                 "the static scratch array", need, SCRATCH_BYTES);
        g_last_error = msg;
        return (int)hipErrorInvalidValue;
    }
    if (num_items == 0 || num_segments == 0) return ok();
    void* ptr = nullptr;
    const hipError_t e = scratch_address(&ptr);
    if (e != hipSuccess) return fail(e, "segsort_sort_noalloc: hipGetSymbolAddress(scratch)");
    size_t tb = SCRATCH_BYTES;
    return segsort_sort(ptr, &tb, d_keys_in, d_keys_out, d_vals_in, d_vals_out, num_items, num_segments,
                        d_begin_offsets, d_end_offsets, begin_bit, end_bit, key_type, value_bytes, descending,
                        SEGSORT_ALGO_BATCHED, stream);
}

int segsort_sort_pairs_u32(void* d_temp, size_t* temp_bytes, const uint32_t* d_keys_in, uint32_t* d_keys_out,
                           const uint32_t* d_vals_in, uint32_t* d_vals_out, int num_items, int num_segments,
                           const int* d_begin_offsets, const int* d_end_offsets, int algo, void* stream) {
    return segsort_sort(d_temp, temp_bytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, num_items, num_segments,
                        d_begin_offsets, d_end_offsets, 0, 32, SEGSORT_KEY_U32, 4, 0, algo, stream);
}

int segsort_sort_keys_u32(void* d_temp, size_t* temp_bytes, const uint32_t* d_keys_in, uint32_t* d_keys_out,
                          int num_items, int num_segments, const int* d_begin_offsets, const int* d_end_offsets,
                          int algo, void* stream) {
    return segsort_sort(d_temp, temp_bytes, d_keys_in, d_keys_out, nullptr, nullptr, num_items, num_segments,
                        d_begin_offsets, d_end_offsets, 0, 32, SEGSORT_KEY_U32, 0, 0, algo, stream);
}

int segsort_row_iota(const int* d_begin_offsets, const int* d_end_offsets, int rows, int* d_perm, void* stream) {
    if (rows < 0) return fail(hipErrorInvalidValue, "segsort_row_iota: negative rows");
    if (rows == 0) return ok();
    k_row_iota<<<std::min(rows, 65535), 256, 0, (hipStream_t)stream>>>(d_begin_offsets, d_end_offsets, rows, d_perm);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ok() : fail(e, "segsort_row_iota");
}

int segsort_perm_scatter_f32(const float* d_grad_out, const int* d_perm, const int* d_begin_offsets,
                             const int* d_end_offsets, int rows, float* d_grad_in, void* stream) {
    if (rows < 0) return fail(hipErrorInvalidValue, "segsort_perm_scatter_f32: negative rows");
    if (rows == 0) return ok();
    k_perm_scatter_f32<256><<<std::min(rows, 65535), 256, 0, (hipStream_t)stream>>>(
        d_grad_out, d_perm, d_begin_offsets, d_end_offsets, rows, d_grad_in);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ok() : fail(e, "segsort_perm_scatter_f32");
}

int segsort_set_tuning(int key_width, int value_bytes, int wave_cap, int tile_enable, int grid_mult, int concurrent) {
    const int s = tuning_slot(key_width, value_bytes);
    if (s < 0 || wave_cap < 2 || wave_cap > 512 || tile_enable < 0 || (tile_enable & ~TUNE_BITS) != 0 || grid_mult < 1 ||
        grid_mult > 8 || concurrent < 0 || concurrent > 2)
        return fail(hipErrorInvalidValue, "segsort_set_tuning: out of range");
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    g_override[s] = {wave_cap, tile_enable, grid_mult, concurrent};
    g_override_on[s] = true;
    return ok();
}

int segsort_reset_tuning(int key_width, int value_bytes) {
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    if (key_width == 0) {
        for (bool& on : g_override_on) on = false;
        return ok();
    }
    const int s = tuning_slot(key_width, value_bytes);
    if (s < 0) return fail(hipErrorInvalidValue, "segsort_reset_tuning: bad slot");
    g_override_on[s] = false;
    return ok();
}

int segsort_get_tuning(int key_width, int value_bytes, int* wave_cap, int* tile_enable, int* grid_mult,
                       int* concurrent) {
    const int s = tuning_slot(key_width, value_bytes);
    if (s < 0) return fail(hipErrorInvalidValue, "segsort_get_tuning: bad slot");
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    const Tuning t = g_override_on[s] ? g_override[s] : DEFAULT_TUNING;
    if (wave_cap) *wave_cap = t.wave_cap;
    if (tile_enable) *tile_enable = t.tile_enable;
    if (grid_mult) *grid_mult = t.grid_mult;
    if (concurrent) *concurrent = t.concurrent;
    return ok();
}

int segsort_lookup_tuning(int num_items, int num_segments, int key_width, int value_bytes, int* wave_cap,
                          int* tile_enable, int* grid_mult, int* concurrent, int* from_table) {
    return segsort_lookup_tuning_hinted(num_items, num_segments, 0, key_width, value_bytes, wave_cap, tile_enable,
                                        grid_mult, concurrent, from_table);
}
int segsort_lookup_tuning_hinted(int num_items, int num_segments, int max_segment_length, int key_width,
                                 int value_bytes, int* wave_cap, int* tile_enable, int* grid_mult, int* concurrent,
                                 int* from_table) {
    if (tuning_slot(key_width, value_bytes) < 0 || num_items < 0 || num_segments < 0 || max_segment_length < 0)
        return fail(hipErrorInvalidValue, "segsort_lookup_tuning: bad arguments");
    bool ovr = false;
    const Tuning t = choose_tuning(key_width, value_bytes, num_items, num_segments, max_segment_length, &ovr);
    if (wave_cap) *wave_cap = t.wave_cap;
    if (tile_enable) *tile_enable = t.tile_enable;
    if (grid_mult) *grid_mult = t.grid_mult;
    if (concurrent) *concurrent = t.concurrent;
    if (from_table) *from_table = ovr ? 0 : 1;
    return ok();
}

int segsort_tile_caps(int key_width, int value_bytes, int* caps) {
    const int s = tuning_slot(key_width, value_bytes);
    if (s < 0 || !caps) return fail(hipErrorInvalidValue, "segsort_tile_caps: bad arguments");
    const int* src = nullptr;
    if (key_width == 4)
        src = value_bytes == 0 ? Cfg<uint32_t, uint32_t, false>::CAP
              : value_bytes == 4 ? Cfg<uint32_t, uint32_t, true>::CAP
                                 : Cfg<uint32_t, uint64_t, true>::CAP;
    else
        src = value_bytes == 0 ? Cfg<uint64_t, uint32_t, false>::CAP
              : value_bytes == 4 ? Cfg<uint64_t, uint32_t, true>::CAP
                                 : Cfg<uint64_t, uint64_t, true>::CAP;
    for (int i = 0; i < 4; ++i) caps[i] = src[i];
    return 4;
}

// In-place adapters: sort into a scratch copy, then copy back. With `iota`
// the values are not read: d_vals receives each sorted key's input position
// (mgpu's segmented_sort_indices, kernel_segsort.hxx:418-429), and positions
// outside every segment their own index.
static int inplace_sort(void* d_keys, void* d_vals, int n, const int* d_segs, int S, int key_type, int value_bytes,
                        int descending, bool iota, hipStream_t st) {
    if (n <= 0 || S <= 0) return ok();
    if (key_type < SEGSORT_KEY_U32 || key_type > SEGSORT_KEY_F64)
        return fail(hipErrorInvalidValue, "adapter: bad key_type");
    if (!d_keys || !d_segs || (iota && !d_vals)) return fail(hipErrorInvalidValue, "adapter: NULL buffer");
    const int kw = key_type >= SEGSORT_KEY_U64 ? 8 : 4;
    const int vb = iota ? 4 : (d_vals ? value_bytes : 0);
    if (vb != 0 && vb != 4 && vb != 8) return fail(hipErrorInvalidValue, "adapter: value_bytes must be 0, 4 or 8");
    // end offsets: d_segs[1..S-1] then n
    int* d_end = nullptr;
    void *kout = nullptr, *vout = nullptr, *vin = nullptr, *temp = nullptr;
    const void* vsrc = d_vals;
    size_t tb = 0;
    int rc = 0;
    hipError_t e;
    if ((e = hipMalloc(&d_end, sizeof(int) * (size_t)S)) != hipSuccess) return fail(e, "adapter: alloc");
    if (S > 1 && (e = hipMemcpyAsync(d_end, d_segs + 1, sizeof(int) * (size_t)(S - 1), hipMemcpyDeviceToDevice, st)) !=
                     hipSuccess) {
        rc = fail(e, "adapter: copy offsets");
        goto done;
    }
    if ((e = hipMemcpyAsync(d_end + (S - 1), &n, sizeof(int), hipMemcpyHostToDevice, st)) != hipSuccess) {
        rc = fail(e, "adapter: copy n");
        goto done;
    }
    if ((e = hipMalloc(&kout, (size_t)n * kw)) != hipSuccess) {
        rc = fail(e, "adapter: alloc keys");
        goto done;
    }
    if (vb && (e = hipMalloc(&vout, (size_t)n * vb)) != hipSuccess) {
        rc = fail(e, "adapter: alloc values");
        goto done;
    }
    if (iota) {
        // the values to carry: 0..n-1 (one row covering everything)
        if ((e = hipMalloc(&vin, (size_t)n * 4)) != hipSuccess) {
            rc = fail(e, "adapter: alloc indices");
            goto done;
        }
        k_iota<<<std::min((n + 1023) / 1024, 4096), 256, 0, st>>>((int*)vin, n);
        if ((e = hipGetLastError()) != hipSuccess) {
            rc = fail(e, "adapter: iota");
            goto done;
        }
        vsrc = vin;
    }
    rc = segsort_sort(nullptr, &tb, d_keys, kout, vsrc, vout, n, S, d_segs, d_end, 0, kw * 8, key_type, vb, descending,
                      SEGSORT_ALGO_BATCHED, st);
    if (rc) goto done;
    if ((e = hipMalloc(&temp, tb)) != hipSuccess) {
        rc = fail(e, "adapter: alloc temp");
        goto done;
    }
    // positions before d_segs[0] are not in any segment: keep them
    if ((e = hipMemcpyAsync(kout, d_keys, (size_t)n * kw, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
        rc = fail(e, "adapter: prefill keys");
        goto done;
    }
    if (vb && (e = hipMemcpyAsync(vout, vsrc, (size_t)n * vb, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
        rc = fail(e, "adapter: prefill values");
        goto done;
    }
    rc = segsort_sort(temp, &tb, d_keys, kout, vsrc, vout, n, S, d_segs, d_end, 0, kw * 8, key_type, vb, descending,
                      SEGSORT_ALGO_BATCHED, st);
    if (rc) goto done;
    if ((e = hipMemcpyAsync(d_keys, kout, (size_t)n * kw, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
        rc = fail(e, "adapter: copy back keys");
        goto done;
    }
    if (vb && (e = hipMemcpyAsync(d_vals, vout, (size_t)n * vb, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
        rc = fail(e, "adapter: copy back values");
        goto done;
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) rc = fail(e, "adapter: sync");
done:
    (void)hipStreamSynchronize(st);
    if (temp) (void)hipFree(temp);
    if (vin) (void)hipFree(vin);
    if (vout) (void)hipFree(vout);
    if (kout) (void)hipFree(kout);
    if (d_end) (void)hipFree(d_end);
    return rc;
}

int segsort_bb_segsort(void* d_keys, void* d_vals, int num_items, const int* d_segs, int num_segments, int key_type,
                       int value_bytes, void* stream) {
    return inplace_sort(d_keys, d_vals, num_items, d_segs, num_segments, key_type, value_bytes, 0, false,
                        (hipStream_t)stream) == 0
               ? 1
               : 0;
}

int segsort_mgpu_segmented_sort(void* d_keys, void* d_vals, int count, const int* d_segments, int num_segments,
                                int key_type, int value_bytes, void* stream) {
    return inplace_sort(d_keys, d_vals, count, d_segments, num_segments, key_type, value_bytes, 0, false,
                        (hipStream_t)stream);
}

int segsort_mgpu_segmented_sort_comp(void* d_keys, void* d_vals, int count, const int* d_segments, int num_segments,
                                     int key_type, int value_bytes, int comp, void* stream) {
    if (comp != SEGSORT_LESS && comp != SEGSORT_GREATER)
        return fail(hipErrorInvalidValue, "segsort_mgpu_segmented_sort_comp: comp must be SEGSORT_LESS or SEGSORT_GREATER");
    return inplace_sort(d_keys, d_vals, count, d_segments, num_segments, key_type, value_bytes,
                        comp == SEGSORT_GREATER ? 1 : 0, false, (hipStream_t)stream);
}

int segsort_mgpu_segmented_sort_indices(void* d_keys, int* d_indices, int count, const int* d_segments,
                                        int num_segments, int key_type, int comp, void* stream) {
    if (comp != SEGSORT_LESS && comp != SEGSORT_GREATER)
        return fail(hipErrorInvalidValue,
                    "segsort_mgpu_segmented_sort_indices: comp must be SEGSORT_LESS or SEGSORT_GREATER");
    return inplace_sort(d_keys, d_indices, count, d_segments, num_segments, key_type, 4,
                        comp == SEGSORT_GREATER ? 1 : 0, true, (hipStream_t)stream);
}

}  // extern "C"
