// C ABI entry points (include/cdbg_gpu.h).  Exceptions never cross this
// boundary: each entry point converts them into a status code plus the
// thread-local message returned by cdbg_gpu_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cdbg_gpu.h"
#include "common.h"
#include "hbm_model.h"
#include "job.h"

namespace cdbg_amd {

static thread_local std::string g_last_error;

int set_error(const std::string& msg) {
    g_last_error = msg;
    return -1;
}
const char* last_error() { return g_last_error.c_str(); }

namespace {
std::mutex g_stream_mu;
std::vector<std::pair<int, hipStream_t>> g_idle_streams;  // live until the process ends
}  // namespace

hipStream_t acquire_stream(int device) {
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (size_t i = 0; i < g_idle_streams.size(); ++i)
            if (g_idle_streams[i].first == device) {
                hipStream_t s = g_idle_streams[i].second;
                g_idle_streams.erase(g_idle_streams.begin() + i);
                return s;
            }
    }
    hipStream_t s = nullptr;
    CDBG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void release_stream(int device, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    g_idle_streams.emplace_back(device, s);
}

StreamOwner::~StreamOwner() {
    if (s) release_stream(device, s);
}

Job::~Job() { free_phase2(*this); }  // the stream goes last, after every buffer freed on it

void job_reset(Job& J, cdbg_gpu_ctx* ctx, uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets) {
    if (!ctx) throw Error("null argument");
    if (k < 2 || k > 63) throw Error("k must satisfy 2 <= k <= 63");
    if (m < 2 || m >= k) throw Error("m must satisfy 2 <= m < k");
    if (num_buckets == 0) throw Error("num_buckets must be >= 1");
    if (num_colors == 0) throw Error("no colours");
    J.ctx = ctx;
    J.k = k;
    J.m = m;
    J.B = num_buckets;
    J.num_colors = num_colors;
    J.phase1_done = J.phase2_done = J.from_files = false;
    J.n_supers = 0;
    J.auto_chunk = 0;
    J.n_chunks = 0;
    J.h_bucket_off.clear();
    J.stats = cdbg_gpu_stats{};
    J.p1_auto_bases = 0;
    if (!J.st) {
        J.st = acquire_stream(ctx->device);
        J.stream_owner.s = J.st;
        J.stream_owner.device = ctx->device;
    }
}

void job_set_input(Job& J, cdbg_gpu_ctx* ctx, const uint8_t* bytes, const uint64_t* color_off, uint32_t num_colors,
                   uint32_t k, uint32_t m, uint32_t num_buckets) {
    if (!ctx || !color_off || (!bytes && num_colors)) throw Error("null argument");
    if (num_colors == 0) throw Error("no colours");
    for (uint32_t c = 0; c < num_colors; ++c)
        if (color_off[c + 1] < color_off[c]) throw Error("colour offsets must be non-decreasing");
    job_reset(J, ctx, num_colors, k, m, num_buckets);
    J.n_bytes = color_off[num_colors] - color_off[0];
    // Pad with a non-ACGT byte so tiles and 32-byte groups need no tail case.
    J.n_pad = ((J.n_bytes + 64) / 64) * 64 + 64;
    AllocStream alloc_on_job_stream(J.st);
    uint8_t* d = J.bytes.ensure<uint8_t>(J.n_pad);
    CDBG_HIP(hipMemsetAsync(d + J.n_bytes, '\n', J.n_pad - J.n_bytes, J.st));
    CDBG_HIP(hipMemcpyAsync(d, bytes + color_off[0], J.n_bytes, hipMemcpyHostToDevice, J.st));
    std::vector<uint64_t>& off = J.h_color_off;
    off.resize(num_colors + 1);
    for (uint32_t c = 0; c <= num_colors; ++c) off[c] = color_off[c] - color_off[0];
    uint64_t* doff = J.color_off.ensure<uint64_t>(num_colors + 1);
    CDBG_HIP(hipMemcpyAsync(doff, off.data(), off.size() * 8, hipMemcpyHostToDevice, J.st));
    CDBG_HIP(hipStreamSynchronize(J.st));  // the host arrays are pageable and the caller's
}

__global__ void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

CDBG_API const char* cdbg_gpu_last_error(void) { return last_error(); }

CDBG_API int cdbg_gpu_ctx_create(int device, cdbg_gpu_ctx** out) {
    CDBG_TRY({
        if (!out) throw Error("null output pointer");
        int n = 0;
        CDBG_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw Error("no HIP device " + std::to_string(device));
        CDBG_HIP(hipSetDevice(device));
        auto* c = new cdbg_gpu_ctx();
        c->device = device;
        *out = c;
    })
}

CDBG_API void cdbg_gpu_ctx_destroy(cdbg_gpu_ctx* ctx) { delete ctx; }

CDBG_API int cdbg_gpu_job_create(cdbg_gpu_ctx* ctx, const uint8_t* bytes, const uint64_t* color_off,
                                 uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets,
                                 cdbg_gpu_job** out) {
    CDBG_TRY({
        if (!ctx || !out) throw Error("null argument");
        CDBG_HIP(hipSetDevice(ctx->device));
        auto* h = new cdbg_gpu_job();
        h->J.tune = ctx->tune;
        try {
            job_set_input(h->J, ctx, bytes, color_off, num_colors, k, m, num_buckets);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    })
}

CDBG_API int cdbg_gpu_job_create_from_files(cdbg_gpu_ctx* ctx, const char* const* files, uint32_t n_files, uint32_t k,
                                            uint32_t m, uint32_t num_buckets, uint32_t threads, cdbg_gpu_job** out) {
    CDBG_TRY({
        if (!ctx || !files || !out) throw Error("null argument");
        CDBG_HIP(hipSetDevice(ctx->device));
        std::vector<std::string> paths;
        for (uint32_t i = 0; i < n_files; ++i) {
            if (!files[i]) throw Error("null file name");
            paths.emplace_back(files[i]);
        }
        auto* h = new cdbg_gpu_job();
        h->J.tune = ctx->tune;
        try {
            Job& J = h->J;
            job_reset(J, ctx, n_files, k, m, num_buckets);
            J.io_threads = std::max<uint32_t>(1, threads);
            J.input_on_device = load_inputs_device(J, paths, threads);
            if (!J.input_on_device) {  // FASTQ quality lines: the host reader
                void* seqs = cdbg_load_fasta(files, n_files, threads);
                if (!seqs) throw Error(last_error());
                struct Free {
                    void* p;
                    ~Free() { cdbg_seqs_free(p); }
                } fr{seqs};
                job_set_input(J, ctx, cdbg_seqs_bytes(seqs), cdbg_seqs_offsets(seqs), n_files, k, m, num_buckets);
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    })
}

CDBG_API int64_t cdbg_gpu_job_input(cdbg_gpu_job* job, uint8_t* bytes, uint64_t cap, uint64_t* color_off) {
    try {
        if (!job) throw Error("null argument");
        Job& J = job->J;
        if (bytes) {
            if (cap < J.n_bytes) throw Error("input buffer too small");
            CDBG_HIP(hipSetDevice(J.ctx->device));
            CDBG_HIP(hipStreamSynchronize(J.st));
            if (J.n_bytes) CDBG_HIP(hipMemcpy(bytes, J.bytes.p, J.n_bytes, hipMemcpyDeviceToHost));
        }
        if (color_off)
            for (uint32_t c = 0; c <= J.num_colors; ++c) color_off[c] = J.h_color_off[c];
        return int64_t(J.input_on_device ? J.n_bytes | (uint64_t(1) << 62) : J.n_bytes);
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API void cdbg_gpu_job_destroy(cdbg_gpu_job* job) { delete job; }

CDBG_API int cdbg_gpu_job_set_tuning(cdbg_gpu_job* job, const cdbg_gpu_tuning* t) {
    CDBG_TRY({
        if (!job || !t) throw Error("null argument");
        check_tuning(*t);
        job->J.tune = *t;
    })
}

CDBG_API int cdbg_gpu_job_get_tuning(cdbg_gpu_job* job, cdbg_gpu_tuning* t) {
    CDBG_TRY({
        if (!job || !t) throw Error("null argument");
        *t = job->J.tune;
    })
}

CDBG_API int cdbg_gpu_ctx_set_tuning(cdbg_gpu_ctx* ctx, const cdbg_gpu_tuning* t) {
    CDBG_TRY({
        if (!ctx || !t) throw Error("null argument");
        check_tuning(*t);
        ctx->tune = *t;
    })
}

CDBG_API int cdbg_gpu_job_autotune(cdbg_gpu_job* job, uint64_t sample_occurrences, cdbg_gpu_tune_report* out) {
    CDBG_TRY({
        if (!job) throw Error("null argument");
        CDBG_HIP(hipSetDevice(job->J.ctx->device));
        const cdbg_gpu_tune_report r = autotune(job->J, sample_occurrences);
        if (out) *out = r;
    })
}

CDBG_API int cdbg_gpu_job_set_bucket_map(cdbg_gpu_job* job, const uint64_t* cuts, uint32_t n_cuts) {
    CDBG_TRY({
        if (!job) throw Error("null job");
        Job& J = job->J;
        if (!cuts) {
            J.h_cuts.clear();
        } else {
            if (n_cuts + 1 != J.B) throw Error("bucket map: n_cuts must be num_buckets - 1");
            for (uint32_t i = 0; i < n_cuts; ++i) {
                if (cuts[i] >> 32) throw Error("bucket map: cuts must be below 2^32");
                if (i && cuts[i] < cuts[i - 1]) throw Error("bucket map: cuts must be ascending");
            }
            J.h_cuts.assign(cuts, cuts + n_cuts);
        }
        J.phase1_done = J.phase2_done = false;
    })
}

CDBG_API int cdbg_gpu_job_set_phase1_chunk_bases(cdbg_gpu_job* job, uint64_t max_bases) {
    if (!job) return set_error("null job");
    job->J.p1_chunk_bases = max_bases;
    job->J.phase1_done = job->J.phase2_done = false;
    return 0;
}

CDBG_API int64_t cdbg_gpu_job_num_phase1_ranges(cdbg_gpu_job* job) {
    if (!job) return set_error("null job");
    return job->J.p1_ranges;
}

CDBG_API int cdbg_gpu_job_phase1(cdbg_gpu_job* job) {
    CDBG_TRY({
        if (!job) throw Error("null job");
        if (job->J.from_files) throw Error("this job was loaded from bucket files: it has no sequences for phase 1");
        CDBG_HIP(hipSetDevice(job->J.ctx->device));
        job->J.phase2_done = false;
        ProfScope prof(&job->J.prof);
        run_phase1(job->J);
        job->J.prof.collect();
    })
}

CDBG_API int cdbg_gpu_job_phase2(cdbg_gpu_job* job) {
    CDBG_TRY({
        if (!job) throw Error("null job");
        if (!job->J.phase1_done) throw Error("phase 2 requires phase 1");
        CDBG_HIP(hipSetDevice(job->J.ctx->device));
        ProfScope prof(&job->J.prof);
        run_phase2(job->J);
        job->J.prof.collect();
    })
}

CDBG_API int cdbg_gpu_job_run(cdbg_gpu_job* job) {
    if (job && job->J.from_files) return cdbg_gpu_job_phase2(job);
    int rc = cdbg_gpu_job_phase1(job);
    if (rc) return rc;
    return cdbg_gpu_job_phase2(job);
}

CDBG_API int cdbg_gpu_job_stats(cdbg_gpu_job* job, cdbg_gpu_stats* out) {
    CDBG_TRY({
        if (!job || !out) throw Error("null argument");
        *out = job->J.stats;
    })
}

CDBG_API int cdbg_gpu_job_supers(cdbg_gpu_job* job, int which, uint64_t* start, uint32_t* len, uint8_t* flags,
                                 uint32_t* bucket, uint32_t* color) {
    CDBG_TRY({
        if (!job) throw Error("null job");
        Job& J = job->J;
        if (!J.phase1_done) throw Error("phase 1 has not run");
        const uint64_t n = J.n_supers;
        Phase1Bufs& P = J.p1;
        auto cp = [&](void* dst, const DevBuf& src, size_t elem) {
            if (dst && n) CDBG_HIP(hipMemcpyAsync(dst, src.p, n * elem, hipMemcpyDeviceToHost, J.st));
        };
        if (which == 0) {
            // position order: the bucket order re-sorted by (start, length);
            // two supers share a start only when the first is the one-(k-1)-mer
            // run at the start of a run, and it is also the shorter one
            std::vector<uint64_t> s(n);
            std::vector<uint32_t> l(n), b(n), c(n);
            std::vector<uint8_t> f(n);
            cp(s.data(), P.b_start, 8);
            cp(l.data(), P.b_len, 4);
            cp(f.data(), P.b_flags, 1);
            cp(b.data(), P.key, 4);
            cp(c.data(), P.b_color, 4);
            CDBG_HIP(hipStreamSynchronize(J.st));
            std::vector<uint32_t> perm(n);
            for (uint64_t i = 0; i < n; ++i) perm[i] = uint32_t(i);
            std::sort(perm.begin(), perm.end(),
                      [&](uint32_t x, uint32_t y) { return s[x] < s[y] || (s[x] == s[y] && l[x] < l[y]); });
            for (uint64_t i = 0; i < n; ++i) {
                if (start) start[i] = s[perm[i]];
                if (len) len[i] = l[perm[i]];
                if (flags) flags[i] = f[perm[i]];
                if (bucket) bucket[i] = b[perm[i]];
                if (color) color[i] = c[perm[i]];
            }
        } else {
            cp(start, P.b_start, 8);
            cp(len, P.b_len, 4);
            cp(flags, P.b_flags, 1);
            cp(bucket, P.key, 4);
            cp(color, P.b_color, 4);
        }
        CDBG_HIP(hipStreamSynchronize(J.st));
    })
}

CDBG_API int cdbg_gpu_job_set_profiling(cdbg_gpu_job* job, int on) {
    CDBG_TRY({
        if (!job) throw Error("null job");
        job->J.prof.on = on != 0;
        job->J.prof.stats.clear();
    })
}

CDBG_API int64_t cdbg_gpu_job_kernel_times(cdbg_gpu_job* job, char* buf, uint64_t cap) {
    if (!job) return set_error("null job");
    std::string txt;
    for (auto& kv : job->J.prof.stats)
        txt += kv.first + " " + std::to_string(kv.second.launches) + " " + std::to_string(kv.second.ms) + "\n";
    if (buf && cap) {
        size_t n = std::min<size_t>(txt.size(), cap - 1);
        std::memcpy(buf, txt.data(), n);
        buf[n] = 0;
    }
    return int64_t(txt.size());
}

CDBG_API int cdbg_gpu_job_write_bucket_files(cdbg_gpu_job* job, const char* dir, uint32_t threads,
                                             uint64_t* bucket_unc_bytes) {
    CDBG_TRY({
        if (!job || !dir) throw Error("null argument");
        CDBG_HIP(hipSetDevice(job->J.ctx->device));
        ProfScope prof(&job->J.prof);
        write_bucket_files(job->J, dir, threads, bucket_unc_bytes);
    })
}

static int job_from_bucket_files(cdbg_gpu_ctx* ctx, const char* dir, uint32_t num_buckets, uint32_t k,
                                 uint32_t num_colors, uint32_t threads, bool delete_consumed, cdbg_gpu_job** out) {
    CDBG_TRY({
        if (!ctx || !dir || !out) throw Error("null argument");
        if (k < 2 || k > 63) throw Error("k must satisfy 2 <= k <= 63");
        if (num_buckets == 0 || num_colors == 0) throw Error("num_buckets and num_colors must be >= 1");
        CDBG_HIP(hipSetDevice(ctx->device));
        auto* h = new cdbg_gpu_job();
        try {
            Job& J = h->J;
            J.ctx = ctx;
            J.tune = ctx->tune;
            J.k = k;
            J.B = num_buckets;
            J.num_colors = num_colors;
            J.st = acquire_stream(ctx->device);
            J.stream_owner.s = J.st;
            J.stream_owner.device = ctx->device;
            J.io_threads = std::max<uint32_t>(1, threads);
            load_bucket_files(J, dir, threads, delete_consumed);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    })
}

CDBG_API int cdbg_gpu_job_from_bucket_files(cdbg_gpu_ctx* ctx, const char* dir, uint32_t num_buckets, uint32_t k,
                                            uint32_t num_colors, uint32_t threads, cdbg_gpu_job** out) {
    return job_from_bucket_files(ctx, dir, num_buckets, k, num_colors, threads, false, out);
}

CDBG_API int cdbg_gpu_job_write_phase2_files(cdbg_gpu_job* job, const char* frag_path, const char* color_sets_path,
                                             cdbg_gpu_color_state* state) {
    CDBG_TRY({
        if (!job || !color_sets_path) throw Error("null argument");
        CDBG_HIP(hipSetDevice(job->J.ctx->device));
        write_phase2_files(job->J, frag_path ? frag_path : "", color_sets_path, state);
    })
}

CDBG_API int cdbg_gpu_bucket_write(cdbg_gpu_ctx* ctx, const char* const* files, uint32_t n_files, uint32_t k,
                                   uint32_t m, uint32_t num_buckets, const char* tmp_dir, uint32_t threads,
                                   uint64_t* bucket_unc_bytes) {
    cdbg_gpu_job* job = nullptr;
    int rc = cdbg_gpu_job_create_from_files(ctx, files, n_files, k, m, num_buckets, threads, &job);
    if (rc == 0) rc = cdbg_gpu_job_phase1(job);
    if (rc == 0) rc = cdbg_gpu_job_write_bucket_files(job, tmp_dir, threads, bucket_unc_bytes);
    if (job) cdbg_gpu_job_destroy(job);
    return rc;
}

CDBG_API int cdbg_gpu_bucket_process_ex(cdbg_gpu_ctx* ctx, const char* tmp_dir, uint32_t num_buckets, uint32_t k,
                                        uint32_t num_colors, uint32_t threads, const char* frag_path,
                                        const char* color_sets_path, cdbg_gpu_color_state* state, uint64_t* done,
                                        uint64_t mem_budget_bytes, int delete_consumed_buckets) {
    cdbg_gpu_job* job = nullptr;
    int rc = job_from_bucket_files(ctx, tmp_dir, num_buckets, k, num_colors, threads, delete_consumed_buckets != 0,
                                   &job);
    if (rc == 0) {
        job->J.progress = done;
        // the device working set of a phase-2 chunk within the budget (one
        // bucket at least, as the reference admits one bucket whatever its size)
        if (mem_budget_bytes)
            job->J.chunk_tuples = std::max<uint64_t>(1, uint64_t(0.6 * double(mem_budget_bytes) / p2_per_occ(k)));
        rc = cdbg_gpu_job_phase2(job);
    }
    if (rc == 0) rc = cdbg_gpu_job_write_phase2_files(job, frag_path, color_sets_path, state);
    if (job) cdbg_gpu_job_destroy(job);
    return rc;
}

CDBG_API int cdbg_gpu_bucket_process(cdbg_gpu_ctx* ctx, const char* tmp_dir, uint32_t num_buckets, uint32_t k,
                                     uint32_t num_colors, uint32_t threads, const char* frag_path,
                                     const char* color_sets_path, cdbg_gpu_color_state* state) {
    return cdbg_gpu_bucket_process_ex(ctx, tmp_dir, num_buckets, k, num_colors, threads, frag_path, color_sets_path,
                                      state, nullptr, 0, 0);
}

CDBG_API uint64_t cdbg_gpu_device_bytes(void) { return g_device_bytes.load(); }

CDBG_API void cdbg_gpu_inject_oom(int64_t n) { g_block_cache.inject_oom = n < 0 ? -1 : n; }

CDBG_API void cdbg_gpu_alloc_stats(uint64_t out[8]) {
    const BlockCache& c = g_block_cache;
    const uint64_t v[8] = {c.cached.load(), c.n_hit.load(),      c.n_malloc.load(), c.malloc_bytes.load(),
                           c.malloc_ns.load(), c.n_trim.load(), c.trim_bytes.load(), c.trim_ns.load()};
    for (int i = 0; i < 8; ++i) out[i] = v[i];
}

CDBG_API void cdbg_gpu_trim_cache(void) { g_block_cache.trim(); }

CDBG_API int cdbg_gpu_copy_probe(cdbg_gpu_ctx* ctx, uint64_t bytes, int iters, double* gbps) {
    CDBG_TRY({
        if (!ctx || !gbps || iters < 1) throw Error("bad argument");
        CDBG_HIP(hipSetDevice(ctx->device));
        const uint64_t n = bytes / 16;
        DevBuf a, b;
        a.ensure<uint4>(n);
        b.ensure<uint4>(n);
        CDBG_HIP(hipMemset(a.p, 1, n * 16));
        hipEvent_t e0, e1;
        CDBG_HIP(hipEventCreate(&e0));
        CDBG_HIP(hipEventCreate(&e1));
        k_copy_probe<<<4096, 256>>>(a.as<uint4>(), b.as<uint4>(), n);
        CDBG_HIP(hipEventRecord(e0, nullptr));
        for (int i = 0; i < iters; ++i) k_copy_probe<<<4096, 256>>>(a.as<uint4>(), b.as<uint4>(), n);
        CDBG_HIP(hipEventRecord(e1, nullptr));
        CDBG_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CDBG_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = 2.0 * double(n) * 16.0 * iters / (ms * 1e-3) / 1e9;
    })
}

CDBG_API int cdbg_gpu_job_bucket_offsets(cdbg_gpu_job* job, uint64_t* out) {
    CDBG_TRY({
        if (!job || !out) throw Error("null argument");
        if (!job->J.phase1_done) throw Error("phase 1 has not run");
        std::memcpy(out, job->J.h_bucket_off.data(), job->J.h_bucket_off.size() * 8);
    })
}

}  // extern "C"
