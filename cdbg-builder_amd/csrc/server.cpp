// Served mode: a request loop over build jobs (sequences + k) on one GPU.
//
// The reference builds one genome list per process run; its phase-2
// scheduler admits buckets against a RAM budget (24 x the bucket's
// uncompressed bytes, live-RSS ceiling: bucket_walker.hpp:623-671).  Here the
// unit of admission is a whole job and the budget is HBM:
//
//   * submit() queues the request and returns an id; the sequence bytes are
//     copied, or only referenced when the server borrows its callers' inputs
//     (a host copy of a large request costs about as much as building it);
//   * `slots` worker threads each run one job at a time on a HIP stream of
//     their own, so jobs on different slots overlap on the device; a job's
//     buffers come from the process-wide block cache and go back to it when
//     the job ends (the cache keeps them on the device for the next job);
//   * the queue head is admitted (FIFO) when a slot is free and its HBM
//     estimate, the running jobs' estimates and the kept results fit the
//     budget;
//   * phase 2 of a job is chunked so that its estimate fits one slot's share
//     (results do not depend on chunking);
//   * a finished job's phase-2 artifacts are copied into device buffers of
//     the job (they count against the budget until release()); poll()
//     reports the state, the stats and the submit / start / done times
//     (steady clock, seconds), result() copies the artifacts to the host.
//   * a job that runs out of device memory anyway runs again once, alone
//     (the other slots finish their jobs first) and from an empty cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cdbg_gpu.h"
#include "common.h"
#include "hbm_model.h"
#include "job.h"

namespace cdbg_amd {

int set_error(const std::string& msg);

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

uint64_t hbm_estimate(uint64_t bases, uint32_t k, uint32_t m, uint64_t chunk_occ) {
    const uint64_t all = uint64_t(occ_per_base(k, m) * double(bases));
    const uint64_t occ = chunk_occ ? std::min(all, chunk_occ) : all;
    return HBM_FIXED_BYTES + uint64_t(p1_per_base(k, m) * double(bases)) + uint64_t(p2_per_occ(k) * double(occ));
}

// phase-2 chunk (k-mer occurrences) that keeps a job's estimate within cap
uint64_t chunk_for_cap(uint64_t bases, uint32_t k, uint32_t m, uint64_t cap) {
    const double p1 = double(HBM_FIXED_BYTES) + p1_per_base(k, m) * double(bases);
    const double left = double(cap) - p1;
    const uint64_t occ = left > 0 ? uint64_t(left / p2_per_occ(k)) : 0;
    return std::max<uint64_t>(occ, 1u << 20);
}

struct SJob {
    uint64_t id = 0;
    std::vector<uint8_t> bytes;   // owned copy of the sequences (empty when borrowed)
    const uint8_t* in = nullptr;  // the sequences: bytes.data() or the caller's buffer
    std::vector<uint64_t> off;
    uint32_t nc = 0, k = 0, m = 0, B = 0;
    uint64_t est = 0, chunk = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, -1 failed
    uint32_t attempts = 0;
    std::string error;
    double t_submit = 0, t_start = 0, t_done = 0;
    cdbg_gpu_stats stats{};
    DevBuf frags, words, offsets;  // phase-2 artifacts, kept on the device until released
    uint64_t res_bytes = 0;
};

struct Slot {
    std::unique_ptr<Job> job;  // while running: the job's device buffers and stream (slot thread only)
    uint64_t reserve = 0;      // while running: the job's estimate
    bool busy = false;
};

struct Server {
    cdbg_gpu_ctx* ctx = nullptr;
    uint64_t budget = 0;
    uint32_t n_slots = 1;
    bool keep_results = true;
    bool borrow = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SJob*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<SJob>> jobs;
    std::vector<Slot> slots;
    std::vector<std::thread> threads;
    hipStream_t res_st = nullptr;  // stream of the result copies (lives as long as the server)
    uint64_t result_bytes = 0;     // device bytes of finished jobs' results not yet released
    uint64_t next_id = 1;
    bool stop = false;
    int excl = -1;  // slot that reruns a job alone after running out of memory

    uint64_t share() const { return budget / n_slots; }

    // the queue head fits next to the running jobs' estimates and the kept results
    bool admissible(uint32_t me, const SJob& j) const {
        uint64_t total = result_bytes;
        for (uint32_t s = 0; s < n_slots; ++s)
            if (s != me && slots[s].busy) total += slots[s].reserve;
        // a job larger than the whole budget still runs once it is alone
        return total + j.est <= budget || total == 0;
    }

    // Waits until this slot may run alone: no other slot runs a job, and no
    // new job starts until it is done; then empties the block cache.
    void run_alone(uint32_t me) {
        (void)hipGetLastError();  // the failed call's error is not this attempt's
        slots[me].job.reset();    // this slot's buffers back to the cache
        {
            std::unique_lock<std::mutex> lk(mu);
            slots[me].busy = false;  // a slot waiting here holds nothing (two may wait at once)
            cv_work.notify_all();
            cv_work.wait(lk, [&] { return stop || excl < 0; });
            if (stop) throw Error("server stopped");
            excl = int(me);
            cv_work.wait(lk, [&] {
                if (stop) return true;
                for (uint32_t s = 0; s < n_slots; ++s)
                    if (s != me && slots[s].busy) return false;
                return true;
            });
            if (stop) {
                excl = -1;
                throw Error("server stopped");
            }
            slots[me].busy = true;
        }
        g_block_cache.trim();
    }

    void worker(uint32_t me) {
        (void)hipSetDevice(ctx->device);
        for (;;) {
            SJob* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || (excl < 0 && !queue.empty() && admissible(me, *queue.front())); });
                if (stop) break;
                j = queue.front();
                queue.pop_front();
                j->state = 1;
                j->t_start = now_s();
                slots[me].busy = true;
                slots[me].reserve = j->est;
            }
            cv_work.notify_all();  // the next head may fit another slot
            std::string err;
            for (int attempt = 0; attempt < 2; ++attempt) {
                err.clear();
                j->attempts = uint32_t(attempt + 1);
                try {
                    // The job's buffers come from the block cache and go back
                    // to it when the job ends, so the next job of any slot
                    // reuses them.  A job that still runs out of memory (the
                    // estimate is a model; cached blocks may be up to twice
                    // what they serve) runs again alone: the other slots
                    // finish their jobs and the cache is emptied.
                    if (attempt > 0) run_alone(me);
                    slots[me].job = std::make_unique<Job>();
                    Job& J = *slots[me].job;
                    J.tune = ctx->tune;
                    job_set_input(J, ctx, j->in, j->off.data(), j->nc, j->k, j->m, j->B);
                    J.chunk_tuples = j->chunk;
                    run_phase1(J);
                    run_phase2(J);
                    J.phase2_done = true;
                    j->stats = J.stats;
                    if (keep_results) {
                        AllocStream on_result_stream(res_st);
                        uint8_t* f = j->frags.ensure<uint8_t>(J.stats.frag_bytes);
                        uint64_t* w = j->words.ensure<uint64_t>((J.stats.cs_bits + 63) / 64);
                        uint64_t* o = j->offsets.ensure<uint64_t>(J.stats.num_color_sets);
                        copy_phase2_results_device(J, f, w, o, res_st);
                        j->res_bytes = j->frags.cap + j->words.cap + j->offsets.cap;
                    }
                } catch (std::exception& e) {
                    err = e.what();
                } catch (...) {
                    err = "unknown C++ exception";
                }
                if (err.find("out of memory") == std::string::npos) break;
            }
            slots[me].job.reset();
            {
                std::lock_guard<std::mutex> lk(mu);
                slots[me].busy = false;
                slots[me].reserve = 0;
                if (excl == int(me)) excl = -1;
                j->t_done = now_s();
                j->state = err.empty() ? 2 : -1;
                j->error = err;
                result_bytes += j->res_bytes;
                j->bytes = std::vector<uint8_t>();  // the input is no longer needed
                j->in = nullptr;
                j->off = std::vector<uint64_t>();
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
        slots[me].job.reset();
    }
};

}  // namespace cdbg_amd

using namespace cdbg_amd;

struct cdbg_gpu_server {
    Server S;
};

extern "C" {

CDBG_API uint64_t cdbg_gpu_hbm_estimate(uint64_t bases, uint32_t k, uint32_t m, uint64_t chunk_occurrences) {
    if (k < 2 || m < 2 || m >= k) return 0;
    return hbm_estimate(bases, k, m, chunk_occurrences);
}

CDBG_API uint64_t cdbg_gpu_chunk_budget(uint64_t usable_bytes, uint32_t k) {
    if (k < 2 || k > 63) return 0;
    return chunk_budget(usable_bytes, k);
}

CDBG_API int cdbg_gpu_server_create(cdbg_gpu_ctx* ctx, const cdbg_gpu_server_params* params, cdbg_gpu_server** out) {
    CDBG_TRY({
        if (!ctx || !params || !out) throw Error("null argument");
        if (params->slots == 0 || params->slots > 16) throw Error("slots must be in [1, 16]");
        CDBG_HIP(hipSetDevice(ctx->device));
        uint64_t budget = params->hbm_budget;
        if (budget == 0) {
            size_t free_b = 0, total_b = 0;
            CDBG_HIP(hipMemGetInfo(&free_b, &total_b));
            free_b += size_t(g_block_cache.cached.load());  // cached blocks are free for the server
            budget = uint64_t(0.75 * double(free_b));
        }
        auto* h = new cdbg_gpu_server();
        Server& S = h->S;
        S.ctx = ctx;
        S.budget = budget;
        S.n_slots = params->slots;
        S.keep_results = params->keep_results != 0;
        S.borrow = params->borrow_inputs != 0;
        S.slots.resize(S.n_slots);
        S.res_st = acquire_stream(ctx->device);
        for (uint32_t s = 0; s < S.n_slots; ++s) S.threads.emplace_back([&S, s] { S.worker(s); });
        *out = h;
    })
}

CDBG_API int cdbg_gpu_server_submit(cdbg_gpu_server* srv, const uint8_t* bytes, const uint64_t* color_off,
                                    uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets,
                                    uint64_t* job_id) {
    CDBG_TRY({
        if (!srv || !color_off || !job_id || (!bytes && num_colors)) throw Error("null argument");
        if (k < 2 || k > 63) throw Error("k must satisfy 2 <= k <= 63");
        if (m < 2 || m >= k) throw Error("m must satisfy 2 <= m < k");
        if (num_buckets == 0) throw Error("num_buckets must be >= 1");
        if (num_colors == 0) throw Error("no colours");
        for (uint32_t c = 0; c < num_colors; ++c)
            if (color_off[c + 1] < color_off[c]) throw Error("colour offsets must be non-decreasing");
        Server& S = srv->S;
        auto j = std::make_unique<SJob>();
        const uint64_t n = color_off[num_colors] - color_off[0];
        if (S.borrow) {
            j->in = bytes + color_off[0];
        } else {
            j->bytes.assign(bytes + color_off[0], bytes + color_off[0] + n);
            j->in = j->bytes.data();
        }
        j->off.resize(num_colors + 1);
        for (uint32_t c = 0; c <= num_colors; ++c) j->off[c] = color_off[c] - color_off[0];
        j->nc = num_colors;
        j->k = k;
        j->m = m;
        j->B = num_buckets;
        j->chunk = chunk_for_cap(n, k, m, S.share());
        j->est = hbm_estimate(n, k, m, j->chunk);
        j->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lk(S.mu);
            j->id = S.next_id++;
            *job_id = j->id;
            S.queue.push_back(j.get());
            S.jobs.emplace(j->id, std::move(j));
        }
        S.cv_work.notify_all();
    })
}

CDBG_API int cdbg_gpu_server_poll(cdbg_gpu_server* srv, uint64_t job_id, int wait, cdbg_gpu_job_status* out) {
    CDBG_TRY({
        if (!srv || !out) throw Error("null argument");
        Server& S = srv->S;
        std::unique_lock<std::mutex> lk(S.mu);
        auto it = S.jobs.find(job_id);
        if (it == S.jobs.end()) throw Error("unknown job id " + std::to_string(job_id));
        SJob& j = *it->second;
        if (wait) S.cv_done.wait(lk, [&] { return j.state == 2 || j.state == -1; });
        out->state = j.state;
        out->attempts = j.attempts;
        out->hbm_estimate = j.est;
        out->chunk_occurrences = j.chunk;
        out->submit_s = j.t_submit;
        out->start_s = j.t_start;
        out->done_s = j.t_done;
        out->stats = j.stats;
        if (j.state == -1) throw Error("job " + std::to_string(job_id) + " failed: " + j.error);
    })
}

CDBG_API int cdbg_gpu_server_result(cdbg_gpu_server* srv, uint64_t job_id, uint8_t* frags, uint64_t* words,
                                    uint64_t* offsets) {
    CDBG_TRY({
        if (!srv) throw Error("null argument");
        Server& S = srv->S;
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.jobs.find(job_id);
        if (it == S.jobs.end()) throw Error("unknown job id " + std::to_string(job_id));
        SJob& j = *it->second;
        if (j.state != 2) throw Error("job " + std::to_string(job_id) + " has not finished");
        if (!S.keep_results) throw Error("the server does not keep results");
        const cdbg_gpu_stats& st = j.stats;
        if (frags && st.frag_bytes) CDBG_HIP(hipMemcpy(frags, j.frags.p, st.frag_bytes, hipMemcpyDeviceToHost));
        if (words && st.cs_bits)
            CDBG_HIP(hipMemcpy(words, j.words.p, (st.cs_bits + 63) / 64 * 8, hipMemcpyDeviceToHost));
        if (offsets && st.num_color_sets)
            CDBG_HIP(hipMemcpy(offsets, j.offsets.p, st.num_color_sets * 8, hipMemcpyDeviceToHost));
    })
}

CDBG_API int cdbg_gpu_server_release(cdbg_gpu_server* srv, uint64_t job_id) {
    CDBG_TRY({
        if (!srv) throw Error("null argument");
        Server& S = srv->S;
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.jobs.find(job_id);
        if (it == S.jobs.end()) throw Error("unknown job id " + std::to_string(job_id));
        if (it->second->state != 2 && it->second->state != -1) throw Error("job is still queued or running");
        S.result_bytes -= it->second->res_bytes;
        S.jobs.erase(it);
        S.cv_work.notify_all();  // the freed results may let the queue head in
    })
}

CDBG_API void cdbg_gpu_server_destroy(cdbg_gpu_server* srv) {
    if (!srv) return;
    Server& S = srv->S;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        S.stop = true;
    }
    S.cv_work.notify_all();
    for (auto& t : S.threads) t.join();
    S.jobs.clear();  // results free in order on res_st, which goes last
    hipStream_t rs = S.res_st;
    const int dev = S.ctx->device;
    delete srv;
    if (rs) release_stream(dev, rs);
}

}  // extern "C"
