// Optional per-kernel timing with HIP events recorded on the launch stream.
// When a job enables profiling, every launch wrapped in KSCOPE is bracketed
// by two events; after the phase synchronizes, durations are summed per
// kernel name.  Off by default: the wrapper then costs one branch.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <string>
#include <vector>

namespace cdbg_amd {

struct KernelStat {
    uint64_t launches = 0;
    double ms = 0;
};

struct Profiler {
    bool on = false;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::map<std::string, KernelStat> stats;

    void begin(const std::string& name, hipStream_t st) {
        Rec r{name, nullptr, nullptr};
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        (void)hipEventRecord(r.a, st);
        pending.push_back(r);
    }
    void end(hipStream_t st) { (void)hipEventRecord(pending.back().b, st); }
    // call after the stream has synchronized
    void collect() {
        for (auto& r : pending) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            auto& s = stats[r.name];
            s.launches += 1;
            s.ms += ms;
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        pending.clear();
    }
    // drops the events of launches that did not complete (a failed run)
    void discard() {
        for (auto& r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        pending.clear();
    }
};

// The profiler of the job running on this thread (null when none).
inline thread_local Profiler* g_prof = nullptr;

// Makes `p` the profiler of this thread for a scope.  On every exit the
// previous profiler is restored and events still pending are dropped, so a
// run that throws leaves neither a pointer to its job's profiler nor live
// events behind (a successful run collects its events before the scope ends).
struct ProfScope {
    Profiler* prev;
    Profiler* p;
    explicit ProfScope(Profiler* q) : prev(g_prof), p(q) { g_prof = q; }
    ~ProfScope() {
        if (p) p->discard();
        g_prof = prev;
    }
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;
};

struct KScope {
    hipStream_t st;
    bool on;
    KScope(const std::string& name, hipStream_t s) : st(s), on(g_prof && g_prof->on) {
        if (on) g_prof->begin(name, st);
    }
    ~KScope() {
        if (on) g_prof->end(st);
    }
};

// Device bytes currently held by DevBuf allocations of this process.
inline std::atomic<uint64_t> g_device_bytes{0};
// Stream that the calling thread's device buffers are used on: released
// buffers go to the block cache tagged with it (no hipFree, which would wait
// for the whole device).  Null: plain hipMalloc / hipFree.
inline thread_local hipStream_t t_alloc_stream = nullptr;

// Sets t_alloc_stream for a scope.
struct AllocStream {
    hipStream_t prev;
    explicit AllocStream(hipStream_t s) : prev(t_alloc_stream) { t_alloc_stream = s; }
    ~AllocStream() { t_alloc_stream = prev; }
    AllocStream(const AllocStream&) = delete;
    AllocStream& operator=(const AllocStream&) = delete;
};

}  // namespace cdbg_amd

#define KSCOPE(name, st) ::cdbg_amd::KScope _kscope_##__LINE__(name, st)
