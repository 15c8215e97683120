// Optional per-launch timing of the library's kernels with HIP events on the
// launch stream. Off by default; while on, every launch is bracketed by two
// events and the read-out synchronises them (so it is never used under graph
// capture or inside a timed region that measures the whole call).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace ssort {

struct ProfRecord {
    std::string name;
    hipEvent_t a, b;
    double bytes;  // algorithmic bytes of this launch (filled after the call)
    int tag;       // tile class c, 100 + MSD level, or -1
    uint32_t more;  // bit q: tile queue q is also sorted by this launch
    double per_item;  // algorithmic bytes per element of the tagged queue
};

class Profiler {
public:
    static Profiler& get() {
        static Profiler p;
        return p;
    }
    bool on() const { return on_; }
    void enable(bool e) {
        std::lock_guard<std::mutex> lk(mu_);
        on_ = e;
        if (e) clear_locked();
    }
    // Returns the record index (or -1 when off) after recording the start event.
    int start(const char* name, hipStream_t st, int tag, double per_item, double bytes, uint32_t more = 0) {
        if (!on_) return -1;
        std::lock_guard<std::mutex> lk(mu_);
        ProfRecord r;
        r.name = name;
        r.tag = tag;
        r.more = more;
        r.per_item = per_item;
        r.bytes = bytes;
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        (void)hipEventRecord(r.a, st);
        recs_.push_back(r);
        return (int)recs_.size() - 1;
    }
    void stop(int idx, hipStream_t st) {
        if (idx < 0) return;
        std::lock_guard<std::mutex> lk(mu_);
        (void)hipEventRecord(recs_[idx].b, st);
    }
    size_t mark() {
        std::lock_guard<std::mutex> lk(mu_);
        return recs_.size();
    }
    // Attribute element counts of the call whose records start at `from`.
    void attribute(size_t from, const unsigned long long* tile_items, int nclass, const unsigned long long* level_items,
                   int nlevels) {
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t i = from; i < recs_.size(); ++i) {
            ProfRecord& r = recs_[i];
            if (r.tag >= 0 && r.tag < nclass) {
                r.bytes = r.per_item * (double)tile_items[r.tag];
                for (int q = 0; q < nclass && q < 32; ++q)
                    if ((r.more >> q) & 1u) r.bytes += r.per_item * (double)tile_items[q];
            }
            else if (r.tag >= 100 && r.tag - 100 < nlevels) r.bytes = r.per_item * (double)level_items[r.tag - 100];
        }
    }
    // "name<TAB>launches<TAB>total_ms<TAB>bytes" per kernel name.
    std::string report() {
        std::lock_guard<std::mutex> lk(mu_);
        struct Agg {
            int n = 0;
            double ms = 0, bytes = 0;
        };
        std::map<std::string, Agg> agg;
        for (auto& r : recs_) {
            (void)hipEventSynchronize(r.b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            Agg& a = agg[r.name];
            a.n += 1;
            a.ms += ms;
            a.bytes += r.bytes;
        }
        std::string out;
        char line[256];
        for (auto& kv : agg) {
            snprintf(line, sizeof(line), "%s\t%d\t%.6f\t%.0f\n", kv.first.c_str(), kv.second.n, kv.second.ms,
                     kv.second.bytes);
            out += line;
        }
        return out;
    }

private:
    void clear_locked() {
        for (auto& r : recs_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        recs_.clear();
    }
    bool on_ = false;
    std::mutex mu_;
    std::vector<ProfRecord> recs_;
};

// The same for a launch that also sorts the tile queues in the bit mask `more`.
#define SEGSORT_LAUNCHM(name, st, tag, more, per_item, stmt)                                      \
    do {                                                                                          \
        const int _pi = ::ssort::Profiler::get().start(name, st, tag, per_item, 0.0, more);      \
        stmt;                                                                                     \
        ::ssort::Profiler::get().stop(_pi, st);                                                   \
    } while (0)

// Brackets one launch statement with profiler events.
#define SEGSORT_LAUNCH(name, st, tag, per_item, bytes, stmt)                                      \
    do {                                                                                          \
        const int _pi = ::ssort::Profiler::get().start(name, st, tag, per_item, bytes);          \
        stmt;                                                                                     \
        ::ssort::Profiler::get().stop(_pi, st);                                                   \
    } while (0)

}  // namespace ssort
