// One build job resident in HBM: inputs, phase-1 records, phase-2 results.
// Buffers only grow, so repeated runs of the same job allocate nothing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "cdbg_gpu.h"
#include "prims.h"

namespace cdbg_amd {

struct Phase1Bufs {
    DevBuf packed, valid;              // 2-bit bases (16 per u32), ACGT bitmap (32 per u32)
    DevBuf tile_cnt, tile_off, totals; // per-tile start/end counts and their scans
    DevBuf st_s, st_e;                 // per-tile super starts (4 B) and ends (2 B)
    DevBuf sk_rec, sk_rec_t;           // super records (16 B), bucket sort payload
    DevBuf key, key_t;                 // bucket sort keys (bucket of each super)
    DevBuf b_start, b_len, b_flags, b_color;                // supers, bucket-major
    DevBuf bucket_off;                 // B+1
    DevBuf rank0;                      // per record: stream rank of its first k-mer (bucket files only)
    DevBuf cuts;                       // bucket map (B - 1 cuts), when one is set
    DevBuf rkey_t, range_tab;          // colour ranges: sort buffer, merge table
    std::vector<std::unique_ptr<DevBuf>> range_rec;  // colour ranges: every range's records, sorted by bucket
};

struct Phase2Bufs;  // defined in phase2.hip

// Streams are created once per device and handed from finished jobs to new
// ones: creating a stream costs tens of milliseconds here, and a new job on
// a recycled stream reuses the blocks the previous one freed on it, in order.
hipStream_t acquire_stream(int device);
void release_stream(int device, hipStream_t s);

// Holds a job's HIP stream; a Job declares it first, so it is given back
// after every device buffer of the job has been freed in order on it.
struct StreamOwner {
    hipStream_t s = nullptr;
    int device = 0;
    StreamOwner() = default;
    StreamOwner(const StreamOwner&) = delete;
    StreamOwner& operator=(const StreamOwner&) = delete;
    ~StreamOwner();
};

struct Job {
    StreamOwner stream_owner;
    cdbg_gpu_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
    uint32_t k = 0, m = 0, B = 0, num_colors = 0;
    uint64_t n_bytes = 0;        // input bytes (unpadded)
    uint64_t n_pad = 0;          // padded to a multiple of 64
    DevBuf bytes, color_off;
    Phase1Bufs p1;
    RadixScratch rs;
    Phase2Bufs* p2 = nullptr;
    bool phase1_done = false, phase2_done = false;
    bool from_files = false;  // phase-1 records were read from bucket files
    uint64_t n_supers = 0;
    std::vector<uint64_t> h_bucket_off;  // host copy (B+1)
    std::vector<uint64_t> h_cuts;        // bucket map: B - 1 ascending cuts of (minimizer >> 1), or empty
    cdbg_gpu_stats stats{};
    Profiler prof;  // per-kernel event timing, off unless enabled
    uint64_t chunk_tuples = 0;  // phase-2 chunk budget in k-mer occurrences (0 = from free HBM)
    uint64_t* progress = nullptr;  // caller's counter of buckets finished by phase 2 (added to atomically)
    uint64_t auto_chunk = 0;    // budget derived from free HBM on the first run, then kept
    uint32_t n_chunks = 0;      // chunks used by the last phase-2 run
    std::vector<uint64_t> h_color_off;  // colour offsets (relative to the first colour), host copy
    uint64_t p1_chunk_bases = 0;  // phase-1 colour range limit in bases (0 = from free HBM, ~0 = one range)
    uint64_t p1_auto_bases = 0;   // limit derived from free HBM on the first run, then kept (~0: one range)
    uint32_t p1_ranges = 0;       // colour ranges used by the last phase-1 run
    bool frags_in_host = false;   // the last phase-2 run kept its fragment stream in host memory
    cdbg_gpu_tuning tune{};       // phase-2 performance knobs (zero fields: defaults)
    uint32_t p2_max_chunks = 0;   // stop phase 2 after this many chunks (0: all; autotuner trials)
    bool input_on_device = false; // the input was decoded and parsed on the device
    uint32_t io_threads = 1;      // host threads for the job's file reads and writes
    ~Job();
};

// (Re)loads a job's input: validates the arguments, uploads the sequences and
// colour offsets on the job's stream, keeps the device buffers it already has.
void job_set_input(Job& J, cdbg_gpu_ctx* ctx, const uint8_t* bytes, const uint64_t* color_off, uint32_t num_colors,
                   uint32_t k, uint32_t m, uint32_t num_buckets);
// Validates the job parameters and clears its state (no input yet).
void job_reset(Job& J, cdbg_gpu_ctx* ctx, uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets);
// inflate.hip: the files (colour = position) into J's input on the device;
// false (J's input untouched) when the host reader must take them.
bool load_inputs_device(Job& J, const std::vector<std::string>& files, uint32_t threads);
void run_phase1(Job& J);
void phase1_core(Job& J);  // run_phase1 without timing or allocation scope
void run_phase2(Job& J);
void check_tuning(const cdbg_gpu_tuning& t);  // throws on an invalid knob
cdbg_gpu_tune_report autotune(Job& J, uint64_t sample_occurrences);
void free_phase2(Job& J);
// Host copies of the phase-2 results (any pointer may be null).
void copy_phase2_results(Job& J, uint8_t* frags, uint64_t* words, uint64_t* offsets);
// Where the fragment stream of the last phase-2 run is (n bytes): on the
// device, or in phase 2's own pinned host copy.
const uint8_t* phase2_frags_where(Job& J, bool& on_device, uint64_t& n);
// Device copies of the same results on stream st (waits for the job first).
void copy_phase2_results_device(Job& J, void* frags, void* words, void* offsets, hipStream_t st);
// bucket_io.hip
void write_bucket_files(Job& J, const std::string& dir, uint32_t threads, uint64_t* unc_bytes);
void load_bucket_files(Job& J, const std::string& dir, uint32_t threads, bool delete_consumed = false);
// frag_path empty: the colour-set files only
void write_phase2_files(Job& J, const std::string& frag_path, const std::string& cs_path,
                        cdbg_gpu_color_state* state);
// stitch_gpu.hip: phase 3 over the job's fragment stream where it lies
struct Unitigs;
Unitigs* stitch_job_gpu(Job& J, uint32_t threads);

}  // namespace cdbg_amd

struct cdbg_gpu_ctx {
    int device = 0;
    cdbg_gpu_tuning tune{};  // default tuning of jobs created on this context
};

struct cdbg_gpu_job {
    cdbg_amd::Job J;
};
