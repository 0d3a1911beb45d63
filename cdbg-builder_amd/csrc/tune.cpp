// Autotuner of phase 2's performance knobs.
//
// A trial is phase 2 over the job's first buckets only (a chunk of about
// `sample` k-mer occurrences), timed with events on the job's stream; the
// best of two runs counts.  Knobs are tried one at a time from the defaults
// (coordinate descent): cell size, radix digit width, persistent cell sorts,
// graph capture of that stage, colour-set table slack; a value is kept when
// it is at least 1% faster than the best so far.  No knob changes a result,
// so the tuner needs no reference output.
#include <algorithm>
#include <vector>

#include "job.h"

namespace cdbg_amd {

void check_tuning(const cdbg_gpu_tuning& t) {
    if (t.cell_target != 0.0 && !(t.cell_target >= 16.0 && t.cell_target <= 65536.0))
        throw Error("tuning: cell_target must be 0 or within [16, 65536]");
    if (t.digit_bits > 11) throw Error("tuning: digit_bits must be 0 or within [1, 11]");
    if (t.set_slack != 0.0 && !(t.set_slack >= 1.25 && t.set_slack <= 16.0))
        throw Error("tuning: set_slack must be 0 or within [1.25, 16]");
    if (t.persistent > 1 || t.graph > 1) throw Error("tuning: persistent and graph are 0 or 1");
    if (t.gzip_device > 2) throw Error("tuning: gzip_device is 0, 1 or 2");
}

cdbg_gpu_tune_report autotune(Job& J, uint64_t sample) {
    if (!J.phase1_done) throw Error("autotune needs phase 1 first");
    if (sample == 0) sample = uint64_t(1) << 26;
    struct Restore {
        Job& J;
        uint64_t chunk;
        uint32_t max_chunks;
        bool prof;
        ~Restore() {
            J.chunk_tuples = chunk;
            J.p2_max_chunks = max_chunks;
            J.prof.on = prof;
            J.phase2_done = false;  // the trials leave no result
        }
    } restore{J, J.chunk_tuples, J.p2_max_chunks, J.prof.on};
    const uint32_t saved_gzip = J.tune.gzip_device;
    J.chunk_tuples = sample;
    J.p2_max_chunks = 1;
    J.prof.on = false;  // the trials stay out of the job's kernel times
    cdbg_gpu_tune_report rep{};
    rep.sample_occurrences = sample;
    auto time_of = [&](const cdbg_gpu_tuning& t) {
        J.tune = t;
        double best = 1e300;
        for (int r = 0; r < 2; ++r) {
            run_phase2(J);
            best = std::min(best, double(J.stats.ms_phase2));
            ++rep.trials;
        }
        return best;
    };
    cdbg_gpu_tuning cur{};
    cur.gzip_device = saved_gzip;  // not a phase-2 knob: kept
    J.tune = cur;
    run_phase2(J);  // warm: kernels loaded, buffers cached
    double best = time_of(cur);
    rep.default_ms = best;
    const bool wide = J.k > 31;
    const std::vector<double> cells = wide ? std::vector<double>{350, 500, 1000}
                                           : std::vector<double>{500, 700, 1400, 2000};
    for (double c : cells) {
        cdbg_gpu_tuning t = cur;
        t.cell_target = c;
        const double ms = time_of(t);
        if (ms < 0.99 * best) best = ms, cur = t;
    }
    for (uint32_t d : {6u, 11u}) {
        cdbg_gpu_tuning t = cur;
        t.digit_bits = d;
        const double ms = time_of(t);
        if (ms < 0.99 * best) best = ms, cur = t;
    }
    {
        cdbg_gpu_tuning t = cur;
        t.persistent = 1;
        t.graph = 0;
        const double ms = time_of(t);
        if (ms < 0.99 * best) best = ms, cur = t;
        t.graph = 1;
        const double mg = time_of(t);
        if (mg < 0.99 * best) best = mg, cur = t;
    }
    for (double sl : {1.5, 4.0}) {
        cdbg_gpu_tuning t = cur;
        t.set_slack = sl;
        const double ms = time_of(t);
        if (ms < 0.99 * best) best = ms, cur = t;
    }
    J.tune = cur;
    rep.best = cur;
    rep.best_ms = best;
    return rep;
}

}  // namespace cdbg_amd
