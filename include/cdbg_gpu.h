/*
 * cdbg_gpu.h — C ABI of the MI355X-native bucket-write (phase 1) and
 * bucket-process (phase 2) of cdbg-builder.
 *
 * Plain pointers and sizes only; no C++ types cross this boundary.  Every
 * call returns 0 on success and a negative status on failure, with the
 * message available from cdbg_gpu_last_error() (thread-local).  The reference
 * reports failures as C++ exceptions caught in main (src/main.cpp:82-88);
 * nothing here throws.
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   cdbg_gpu_job_create + cdbg_gpu_job_phase1
 *       -> ingest_bucketed(files, k, m, num_buckets, writer, T, ...)
 *          include/phase1_bucket_write/bucket_ingester.hpp:183-215,
 *          called from include/builder.hpp:180-181
 *   cdbg_gpu_job_phase2
 *       -> process_buckets(writer, k, num_colors, T, sink, global_dict, ...)
 *          include/phase2_bucket_process/bucket_walker.hpp:607-779,
 *          called from include/builder.hpp:269-271
 *   cdbg_gpu_bucket_write / cdbg_gpu_bucket_process
 *       -> the same two phases with their on-disk artifacts: bucket_<b>.bin
 *          (bucket_io.hpp:19-40), frag_unitigs.bin (unitig_spill.hpp:491-522),
 *          <out>.color_sets + .tmp_offsets (streaming_color_set_dict.hpp:86-182)
 */
#ifndef CDBG_GPU_H
#define CDBG_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDBG_API __attribute__((visibility("default")))

typedef struct cdbg_gpu_ctx cdbg_gpu_ctx;
typedef struct cdbg_gpu_job cdbg_gpu_job;

/* Statistics of one job.  num_kmers equals the reference's metadata
 * num_kmers (builder.hpp:405-406): every distinct canonical k-mer is primary
 * in exactly one bucket. */
typedef struct {
    uint64_t num_bases;      /* input bytes scanned by phase 1 */
    uint64_t num_supers;     /* super-k-mer occurrences = bucket records */
    uint64_t num_occ;        /* k-mer occurrences over all records */
    uint64_t num_tuples;     /* (k-mer, colour) tuples sorted in phase 2 */
    uint64_t num_nodes;      /* per-bucket nodes, foreign copies included */
    uint64_t num_kmers;      /* distinct canonical k-mers */
    uint64_t num_color_sets; /* distinct colour sets */
    uint64_t num_frags;      /* unitig fragments */
    uint64_t frag_bytes;     /* bytes of the frag_unitigs.bin stream */
    uint64_t cs_bits;        /* bits of the colour-set stream */
    uint64_t total_integers; /* sum of colour-set sizes */
    double ms_phase1;        /* device wall time of the last phase-1 run */
    double ms_phase2;        /* device wall time of the last phase-2 run */
} cdbg_gpu_stats;

/* Error string of the last failing call on this thread. */
CDBG_API const char* cdbg_gpu_last_error(void);

/* One context per device.  `device` is a HIP device ordinal. */
CDBG_API int cdbg_gpu_ctx_create(int device, cdbg_gpu_ctx** out);
CDBG_API void cdbg_gpu_ctx_destroy(cdbg_gpu_ctx* ctx);

/* Stage one build job in HBM.  `bytes` holds the decoded sequences of all
 * colours; colour c is bytes[color_off[c] .. color_off[c+1]), its records
 * separated by any non-ACGT byte (the layout cdbg_gpu_load_fasta produces).
 * k in [2, 63] (k <= 31 uses 64-bit k-mers, 32..63 uses 128-bit), m in
 * [2, k-1], num_buckets >= 1.  Mirrors the inputs of ingest_bucketed
 * (bucket_ingester.hpp:183-187): colour = index of the input file. */
CDBG_API int cdbg_gpu_job_create(cdbg_gpu_ctx* ctx, const uint8_t* bytes, const uint64_t* color_off,
                                 uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets,
                                 cdbg_gpu_job** out);
CDBG_API void cdbg_gpu_job_destroy(cdbg_gpu_job* job);

/* Phase 1 on the device: 2-bit packing, canonical ntHash minimizers,
 * super-k-mers with BEGIN/END/OWNS flags, bucket scatter
 * (bucket_ingester.hpp:36-127). */
CDBG_API int cdbg_gpu_job_phase1(cdbg_gpu_job* job);
/* Phase 2 on the device over the phase-1 records: per-bucket canonical
 * k-mers, colour sets, global colour-set ids, unitig fragments
 * (bucket_walker.hpp:166-566,713-761).  Requires phase 1. */
CDBG_API int cdbg_gpu_job_phase2(cdbg_gpu_job* job);
/* Both phases back to back. */
CDBG_API int cdbg_gpu_job_run(cdbg_gpu_job* job);

CDBG_API int cdbg_gpu_job_stats(cdbg_gpu_job* job, cdbg_gpu_stats* out);

/* Phase-1 records.  which = 0: position order; which = 1: bucket-major
 * (position order inside each bucket).  Arrays hold num_supers entries.
 * bucket_offsets (num_buckets + 1 entries) delimits the buckets of which=1. */
CDBG_API int cdbg_gpu_job_supers(cdbg_gpu_job* job, int which, uint64_t* start, uint32_t* len,
                                 uint8_t* flags, uint32_t* bucket, uint32_t* color);
CDBG_API int cdbg_gpu_job_bucket_offsets(cdbg_gpu_job* job, uint64_t* out);

/* Per-kernel timing with HIP events on the job's stream (off by default).
 * Enabling resets the totals.  cdbg_gpu_job_kernel_times writes lines
 * "<kernel> <launches> <total ms>" into buf (NUL-terminated, truncated to cap)
 * and returns the full text length.  No reference counterpart: it replaces
 * the reference's in-process phase profilers (util.hpp:433-585). */
CDBG_API int cdbg_gpu_job_set_profiling(cdbg_gpu_job* job, int on);
CDBG_API int64_t cdbg_gpu_job_kernel_times(cdbg_gpu_job* job, char* buf, uint64_t cap);

/* Device bytes currently held by all jobs of this process (the HBM gauge an
 * admission policy reads; the reference's analogue is the live RSS probe,
 * util.hpp:113-215). */
CDBG_API uint64_t cdbg_gpu_device_bytes(void);

/* Counters of the process-wide device block cache (released job buffers are
 * kept for reuse instead of freed): out[0] bytes cached now, [1] requests
 * served from the cache, [2] hipMalloc calls, [3] their bytes, [4] their
 * nanoseconds, [5] trims (cached blocks handed back to the driver), [6]
 * bytes trimmed, [7] nanoseconds in trims.  No counterpart in the reference
 * (host RAM there is plain malloc). */
CDBG_API void cdbg_gpu_alloc_stats(uint64_t out[8]);
/* Hands every cached block back to the driver (waits for the device).  For
 * a process that is done with its jobs, or that starts child processes which
 * need the HBM.  No counterpart in the reference. */
CDBG_API void cdbg_gpu_trim_cache(void);
/* Testing aid: the device allocation after `n` more from the block cache
 * fails as out of memory (n < 0: none).  No counterpart in the reference. */
CDBG_API void cdbg_gpu_inject_oom(int64_t n);

/* Device copy of `bytes` bytes repeated `iters` times; *gbps = read+write
 * GB/s.  Measurement aid (achievable HBM bandwidth, counter calibration); no
 * reference counterpart. */
CDBG_API int cdbg_gpu_copy_probe(cdbg_gpu_ctx* ctx, uint64_t bytes, int iters, double* gbps);

/* Phase 2 works on chunks of consecutive buckets whose k-mer occurrences fit
 * a budget, so HBM use is bounded by the chunk.  0 (default) derives the
 * budget from free HBM on the first run.  Results do not depend on it.  The
 * reference bounds host RAM per in-flight bucket instead
 * (bucket_walker.hpp:623-671). */
CDBG_API int cdbg_gpu_job_set_chunk_tuples(cdbg_gpu_job* job, uint64_t max_tuples);
CDBG_API int64_t cdbg_gpu_job_num_chunks(cdbg_gpu_job* job);

/* Phase 1 runs over ranges of whole colours when the job's phase-1 working
 * set (tile staging, records and their sort buffers) does not fit free HBM:
 * each range is sorted by bucket on its own and the ranges are merged bucket
 * by bucket, so the records are those of one pass.  max_bases caps a range
 * (0, the default: from free HBM on the first run; UINT64_MAX: one range).
 * The reference streams its input file by file through per-thread buffers
 * (bucket_ingester.hpp:183-215, bucket_io.hpp:605-647). */
CDBG_API int cdbg_gpu_job_set_phase1_chunk_bases(cdbg_gpu_job* job, uint64_t max_bases);
CDBG_API int64_t cdbg_gpu_job_num_phase1_ranges(cdbg_gpu_job* job);

/* Bucket map for phase 1: a super-k-mer whose minimizer hash is h goes to
 * the bucket b with cuts[b-1] <= x < cuts[b], x = the low 32 bits of h >> 1
 * (n_cuts = num_buckets - 1 ascending values below 2^32; cuts = NULL restores
 * (h >> 1) % num_buckets).  Every
 * minimizer keeps one bucket, so the graph is the same; bucket sizes follow
 * the map (skewed-bucket workloads).  No counterpart in the reference, which
 * always uses the modulo (bucket_ingester.hpp:67).  Takes effect at the next
 * phase 1. */
CDBG_API int cdbg_gpu_job_set_bucket_map(cdbg_gpu_job* job, const uint64_t* cuts, uint32_t n_cuts);

/* Allocation-free phase 1.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same kernels as
 * cdbg_gpu_job_phase1 run over d_bytes, a DEVICE array laid out as
 * cdbg_gpu_job_create's `bytes` (colour c = d_bytes[color_off[c] ..
 * color_off[c+1]), color_off on the host), and write the super-k-mer records
 * bucket-major, as cdbg_gpu_job_supers(which = 1) returns them, into the
 * caller's device arrays (max_supers entries each) and the num_buckets + 1
 * bucket offsets into d_bucket_off; *n_supers (host) receives the count.
 * Every intermediate buffer lives in a file-scope device array of
 * CDBG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: no
 * hipMalloc or hipFree is called.  Inputs above CDBG_NOALLOC_MAX_BASES bytes,
 * more than max_supers records, or intermediates beyond the array fail with a
 * message and write nothing beyond max_supers.  No bucket map.  The work is
 * ordered on `stream` (a hipStream_t; NULL: the null stream) and finished
 * when the call returns.  Calls in one process share each GPU's array and are
 * serialised by a lock; overlapping them in any other way would race. */
#define CDBG_NOALLOC_SCRATCH_BYTES (64u << 20)
#define CDBG_NOALLOC_MAX_BASES (4u << 20)
CDBG_API int cdbg_gpu_phase1_noalloc(cdbg_gpu_ctx* ctx, const uint8_t* d_bytes, const uint64_t* color_off,
                                     uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets,
                                     uint64_t max_supers, uint64_t* d_start, uint32_t* d_len, uint8_t* d_flags,
                                     uint32_t* d_bucket, uint32_t* d_color, uint64_t* d_bucket_off,
                                     uint64_t* n_supers, void* stream);

/* Phase-2 artifacts: the frag_unitigs.bin byte stream (frag_bytes bytes),
 * the colour-set bit stream (ceil(cs_bits/64) words) and the bit offset of
 * every class (num_color_sets words). */
CDBG_API int cdbg_gpu_job_frags(cdbg_gpu_job* job, uint8_t* dst);
CDBG_API int cdbg_gpu_job_color_sets(cdbg_gpu_job* job, uint64_t* words, uint64_t* offsets);

/* State of the colour-set stream at the end of phase 2: what the host's
 * finalize step (streaming_color_set_dict.hpp:221-294) needs besides the
 * files: the trailing partial word (num_bits % 64 bits) and the totals. */
typedef struct {
    uint32_t num_colors, sparse_threshold, dense_threshold;
    uint64_t num_color_sets;
    uint64_t num_bits;
    uint64_t total_integers;
    uint64_t trailing_word;
    uint64_t num_kmers;
    uint64_t num_fragments;
} cdbg_gpu_color_state;

/* Performance knobs of phase 2 (results never depend on them).  Zero fields
 * take the defaults.  cell_target: mean k-mer occurrences per LDS sort cell
 * (default 1000 for k <= 31, 700 above; 16 .. 65536); digit_bits: largest
 * digit of the global tuple radix passes (default 8; 1 .. 11); set_slack:
 * colour-set table slots per multi-colour set (default 2; 1.25 .. 16);
 * persistent: 1 = the cell sorts run as persistent workgroups that take
 * cells from a device queue, with cell counts read on the device;
 * graph: 1 (with persistent) = that stage is captured once per chunk in a
 * hipGraph and replayed while its arguments are unchanged; gzip_device:
 * where cdbg_gpu_job_create_from_files inflates gzip files (0 = where a
 * model of the measured rates says it is faster: one file per wavefront at
 * ~6 MB/s, two per CU, against ~350 MB/s per host thread; 1 = the device;
 * 2 = host threads).  No counterpart in the reference (its knobs are compile-time
 * constants and -t). */
typedef struct {
    double cell_target;
    uint32_t digit_bits;
    uint32_t persistent;
    double set_slack;
    uint32_t graph;
    uint32_t gzip_device;
} cdbg_gpu_tuning;

/* What cdbg_gpu_job_autotune measured: the best tuning, the trials run and
 * the phase-2 time of the sample with the defaults and with the best. */
typedef struct {
    cdbg_gpu_tuning best;
    uint32_t trials;
    uint32_t reserved;
    uint64_t sample_occurrences;
    double default_ms;
    double best_ms;
} cdbg_gpu_tune_report;

/* Drop-in phase 1: reads the FASTA/FASTQ files (colour = index in `files`),
 * runs phase 1 on the device and writes <tmp_dir>/bucket_<b>.bin for every
 * bucket (bucket_io.hpp:19-40, records super_kmer.hpp:105-117).  As the
 * reference's compacting writer does (bucket_io.hpp:127-277), records with
 * equal bases (and here equal ownership bits) in one bucket are merged into
 * one record with the sorted unique union of their colours, in order of
 * first occurrence; the stream goes out in LZ4 frames of at most 64 KiB,
 * compressed on the device.
 * bucket_unc_bytes (num_buckets entries, may be NULL) receives each bucket's
 * uncompressed size, as bucket_writer::close records it (bucket_io.hpp:389-415).
 * Replaces ingest_bucketed (bucket_ingester.hpp:183-215). */
CDBG_API int cdbg_gpu_bucket_write(cdbg_gpu_ctx* ctx, const char* const* files, uint32_t n_files, uint32_t k,
                                   uint32_t m, uint32_t num_buckets, const char* tmp_dir, uint32_t threads,
                                   uint64_t* bucket_unc_bytes);
/* Drop-in phase 2: reads <tmp_dir>/bucket_<b>.bin, runs phase 2 on the device
 * and writes frag_path (frag_unitigs.bin), color_sets_path (28-byte zero
 * header + complete words) and color_sets_path.tmp_offsets.  Replaces
 * process_buckets (bucket_walker.hpp:607-779) with a frag_unitig_writer sink
 * and a streaming_color_set_dict. */
CDBG_API int cdbg_gpu_bucket_process(cdbg_gpu_ctx* ctx, const char* tmp_dir, uint32_t num_buckets, uint32_t k,
                                     uint32_t num_colors, uint32_t threads, const char* frag_path,
                                     const char* color_sets_path, cdbg_gpu_color_state* state);
/* cdbg_gpu_bucket_process with the remaining parameters of process_buckets
 * (bucket_walker.hpp:607-612):
 *   done              (may be NULL) a counter the call adds the number of
 *                     finished buckets to, atomically, chunk by chunk, as the
 *                     reference's std::atomic<uint64_t>* done progress gauge;
 *   mem_budget_bytes  (0: from free device memory) the device working set of
 *                     a phase-2 chunk, by the library's per-occurrence model;
 *                     a chunk holds one bucket at least, as the reference
 *                     admits one bucket whatever its size (:646-654);
 *   delete_consumed_buckets  nonzero: the bucket files are removed once they
 *                     have all been read (:684-691).
 * The outputs are the same whatever the budget. */
CDBG_API int cdbg_gpu_bucket_process_ex(cdbg_gpu_ctx* ctx, const char* tmp_dir, uint32_t num_buckets, uint32_t k,
                                        uint32_t num_colors, uint32_t threads, const char* frag_path,
                                        const char* color_sets_path, cdbg_gpu_color_state* state, uint64_t* done,
                                        uint64_t mem_budget_bytes, int delete_consumed_buckets);

/* Tuning of one job, and the default for jobs created later on a context
 * (served jobs included).  Invalid values are an error. */
CDBG_API int cdbg_gpu_job_set_tuning(cdbg_gpu_job* job, const cdbg_gpu_tuning* t);
CDBG_API int cdbg_gpu_job_get_tuning(cdbg_gpu_job* job, cdbg_gpu_tuning* t);
CDBG_API int cdbg_gpu_ctx_set_tuning(cdbg_gpu_ctx* ctx, const cdbg_gpu_tuning* t);
/* Autotuner: after phase 1, times phase 2 over a sample of the job's first
 * buckets (about sample_occurrences k-mer occurrences; 0 = 2^26) under
 * candidate knobs, one knob at a time (cell size, digit width, persistent
 * cells, graph capture, table slack), keeps the fastest and sets it on the
 * job.  Phase 2 must then be run again (the trials leave no result). */
CDBG_API int cdbg_gpu_job_autotune(cdbg_gpu_job* job, uint64_t sample_occurrences, cdbg_gpu_tune_report* out);

/* A job over input files (colour = position in `files`), decoded on the
 * device: gzip members are inflated one wavefront per file and the FASTA
 * records parsed there with the rules of kseq_read as seq_reader uses them
 * (seq_reader.hpp:44-77, external/kseq.h:161-201), into the same input
 * cdbg_gpu_job_create takes.  Inputs with FASTQ quality lines are read by the
 * host reader (cdbg_load_fasta) instead.  Corrupt gzip data (a bad code,
 * length, CRC-32 or size field) is an error.  Replaces the reading half of
 * ingest_bucketed (bucket_ingester.hpp:129-177, seq_reader.hpp:44-77). */
CDBG_API int cdbg_gpu_job_create_from_files(cdbg_gpu_ctx* ctx, const char* const* files, uint32_t n_files, uint32_t k,
                                            uint32_t m, uint32_t num_buckets, uint32_t threads, cdbg_gpu_job** out);
/* The job's input bytes (to `bytes`, cap bytes of room, may be NULL) and its
 * num_colors + 1 colour offsets (may be NULL).  Returns the input size, with
 * bit 62 set when the input was decoded on the device; negative on error.
 * No counterpart in the reference (an inspection aid). */
CDBG_API int64_t cdbg_gpu_job_input(cdbg_gpu_job* job, uint8_t* bytes, uint64_t cap, uint64_t* color_off);

/* Job-level pieces of the two calls above. */
CDBG_API int cdbg_gpu_job_write_bucket_files(cdbg_gpu_job* job, const char* dir, uint32_t threads,
                                             uint64_t* bucket_unc_bytes);
CDBG_API int cdbg_gpu_job_from_bucket_files(cdbg_gpu_ctx* ctx, const char* dir, uint32_t num_buckets, uint32_t k,
                                            uint32_t num_colors, uint32_t threads, cdbg_gpu_job** out);
/* frag_path may be NULL: the colour-set files and state only (the fragment
 * stream then goes to cdbg_gpu_job_stitch). */
CDBG_API int cdbg_gpu_job_write_phase2_files(cdbg_gpu_job* job, const char* frag_path, const char* color_sets_path,
                                             cdbg_gpu_color_state* state);

/* Served mode: a request loop over build jobs (sequences + k).  The reference
 * runs one genome list per process; its phase-2 scheduler admits buckets
 * against a RAM budget (bucket_walker.hpp:623-671, cost 24 x uncompressed
 * bucket bytes).  Here `slots` worker threads each run one job at a time
 * on a HIP stream of their own, so jobs on different slots overlap on the
 * device; the queue head is admitted (FIFO) when a slot is free and its
 * estimate (cdbg_gpu_hbm_estimate), the running jobs' estimates and the kept
 * results fit hbm_budget.  Job buffers come from the process-wide block cache
 * (cdbg_gpu_alloc_stats) and go back to it; a job that runs out of device
 * memory anyway runs again once, alone.  A job's results are the same as
 * those of cdbg_gpu_job_run on the same input.  Thread-safe. */
typedef struct cdbg_gpu_server cdbg_gpu_server;
typedef struct {
    uint64_t hbm_budget;   /* device bytes the server may hold (0: 75% of free HBM at creation) */
    uint32_t slots;        /* jobs on the device at once, 1..16 */
    uint32_t keep_results; /* 1: keep each finished job's artifacts in device memory (counted against
                              hbm_budget) until cdbg_gpu_server_release */
    uint32_t borrow_inputs; /* 1: submit keeps the caller's sequence bytes instead of copying them; they
                               must stay valid and unchanged until poll reports the job done or failed */
} cdbg_gpu_server_params;
typedef struct {
    int32_t state;              /* 0 queued, 1 running, 2 done, -1 failed */
    uint32_t attempts;          /* runs of the job: 2 when it ran out of memory and ran again alone */
    uint64_t hbm_estimate;      /* bytes admitted for the job */
    uint64_t chunk_occurrences; /* phase-2 chunk that keeps the estimate within one slot's share */
    double submit_s, start_s, done_s; /* steady-clock seconds: submitted, admitted, artifacts complete */
    cdbg_gpu_stats stats;
} cdbg_gpu_job_status;
/* HBM model: device bytes of a job over `bases` input bytes with k-mer
 * length k and minimizer length m whose phase 2 runs in chunks of
 * `chunk_occurrences` k-mer occurrences (0: one chunk); 0 for invalid k, m. */
CDBG_API uint64_t cdbg_gpu_hbm_estimate(uint64_t bases, uint32_t k, uint32_t m, uint64_t chunk_occurrences);
/* Phase-2 chunk (k-mer occurrences) that the automatic budget plans when
 * `usable_bytes` of device memory are free for phase 2 at k-mer length k
 * (the per-occurrence model of cdbg_gpu_hbm_estimate: 166 B with 64-bit
 * k-mers, 236 B with 128-bit ones); 0 for invalid k.  The reference sizes its
 * admission by 24 x uncompressed bucket bytes (bucket_walker.hpp:631,647). */
CDBG_API uint64_t cdbg_gpu_chunk_budget(uint64_t usable_bytes, uint32_t k);
CDBG_API int cdbg_gpu_server_create(cdbg_gpu_ctx* ctx, const cdbg_gpu_server_params* params,
                                    cdbg_gpu_server** out);
/* Queues a request (layout of cdbg_gpu_job_create).  The colour offsets are
 * copied; the sequence bytes are copied too unless the server borrows inputs. */
CDBG_API int cdbg_gpu_server_submit(cdbg_gpu_server* srv, const uint8_t* bytes, const uint64_t* color_off,
                                    uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets,
                                    uint64_t* job_id);
/* State of a job; wait != 0 blocks until it is done or failed.  Returns -1
 * (message in cdbg_gpu_last_error) for an unknown or failed job. */
CDBG_API int cdbg_gpu_server_poll(cdbg_gpu_server* srv, uint64_t job_id, int wait, cdbg_gpu_job_status* out);
/* Copies a finished job's artifacts to the host: stats.frag_bytes bytes,
 * ceil(stats.cs_bits/64) words, stats.num_color_sets offsets (any pointer may
 * be NULL). */
CDBG_API int cdbg_gpu_server_result(cdbg_gpu_server* srv, uint64_t job_id, uint8_t* frags, uint64_t* words,
                                    uint64_t* offsets);
/* Forgets a finished or failed job and its host copies. */
CDBG_API int cdbg_gpu_server_release(cdbg_gpu_server* srv, uint64_t job_id);
/* Stops the workers after their running jobs; queued jobs are dropped. */
CDBG_API void cdbg_gpu_server_destroy(cdbg_gpu_server* srv);

/* Phase 3, stitch, on the host (replaces stitch_unitigs_extmem_file_streaming,
 * phase3_stitch/stitch_extmem.hpp:1012-1129, called from builder.hpp): reads
 * frag_unitigs.bin records, joins open fragment ends that share their
 * canonical boundary k-mer with compatible orientation, keeps the shared
 * k-mer once with its owner's colour set, and cuts the joined tigs into
 * monochromatic unitigs, dropping foreign placeholder runs at tig ends
 * (stitch_extmem.hpp:112-355).  Returns a handle of the unitigs in emit order
 * (ascending colour-set id), or NULL (message in cdbg_gpu_last_error). */
typedef struct {
    uint64_t num_fragments;   /* fragment records read */
    uint64_t open_ends;       /* open fragment sides */
    uint64_t joins;           /* pairs of open sides joined */
    uint64_t unmatched_ends;  /* open sides without a partner (ends of the graph) */
    uint64_t cycles;          /* tigs that close on themselves */
    uint64_t num_unitigs;     /* monochromatic unitigs */
    uint64_t unitig_kmers;    /* k-mers over all unitigs (equals num_kmers) */
    uint64_t foreign_dropped; /* foreign placeholder k-mers dropped at tig ends */
} cdbg_stitch_stats;
CDBG_API void* cdbg_stitch(const uint8_t* frags, uint64_t n_bytes, uint32_t k, uint32_t threads);
CDBG_API void* cdbg_stitch_file(const char* frag_path, uint32_t k, uint32_t threads);
/* Phase 3 on the device (the same replacement of
 * stitch_unitigs_extmem_file_streaming, stitch_extmem.hpp:1012-1129): the
 * same unitigs, in the same order, with the same statistics and errors as
 * cdbg_stitch / cdbg_stitch_file.  The host parses the record offsets and
 * walks tigs that close on themselves (`threads` threads); joins, chain
 * ranking, unitig cutting, ordering and the bases run on ctx's device.
 * Returns the same kind of handle. */
CDBG_API void* cdbg_gpu_stitch(cdbg_gpu_ctx* ctx, const uint8_t* frags, uint64_t n_bytes, uint32_t k,
                               uint32_t threads);
CDBG_API void* cdbg_gpu_stitch_file(cdbg_gpu_ctx* ctx, const char* frag_path, uint32_t k, uint32_t threads);
/* The same stitch over a job's phase-2 fragment stream where it lies (in HBM,
 * or in host memory when phase 2 kept it there): no frag_unitigs.bin is
 * written or read.  No counterpart in the reference, whose phases 2 and 3
 * meet at that file (builder.hpp:239-394); for callers of the resident job
 * API that want unitigs. */
CDBG_API void* cdbg_gpu_job_stitch(cdbg_gpu_job* job, uint32_t threads);
/* Start of every record of a frag_unitigs.bin stream (frag_record_parse,
 * unitig_spill.hpp:646-665, applied record after record), parsed by up to
 * `threads` threads over ranges of at least 64 KiB: *count = the number of
 * records, the first min(count, cap) offsets go to out; fails on a record
 * that does not fit the stream or when cap < count. */
CDBG_API int cdbg_frag_offsets(const uint8_t* frags, uint64_t n_bytes, uint32_t threads, uint64_t* out, uint64_t cap,
                               uint64_t* count);
CDBG_API int cdbg_unitigs_stats(void* unitigs, cdbg_stitch_stats* out);
/* Total bases of the unitigs, and the unitigs themselves in emit order:
 * sequences concatenated into seq (cdbg_unitigs_bases bytes), offsets
 * (num_unitigs + 1 entries) and colour-set ids (num_unitigs); any pointer may
 * be NULL. */
CDBG_API uint64_t cdbg_unitigs_bases(void* unitigs);
CDBG_API int cdbg_unitigs_export(void* unitigs, char* seq, uint64_t* offsets, uint64_t* cids);
CDBG_API void cdbg_unitigs_free(void* unitigs);

/* Phase 4, emit, on the host (replaces emit_fasta and emit_colors,
 * phase4_emit/emit.hpp:96-180, and builder::write_metadata_, builder.hpp:752-764):
 * <out>.fa (">cid\n<seq>\n" in ascending cid) and <out>.u2c (bit i set iff
 * unitig i ends a run of equal cids); the finished <out>.color_sets from the
 * placeholder, the complete words and <out>.color_sets.tmp_offsets that
 * cdbg_gpu_bucket_process wrote plus its colour state (trailing word,
 * Elias-Fano of the class offsets, header; streaming_color_set_dict.hpp:221-294;
 * the sidecar is removed); <out>.metadata.txt. */
CDBG_API int cdbg_emit_fasta(void* unitigs, const char* out_basename);
CDBG_API int cdbg_finalize_color_sets(const char* color_sets_path, const cdbg_gpu_color_state* state);
CDBG_API int cdbg_write_metadata(const char* out_basename, uint32_t k, uint64_t num_kmers, uint32_t num_colors,
                                 uint64_t num_unitigs, uint64_t num_color_sets);

/* Host input: FASTA/FASTQ (plain or gzip) with kseq record semantics
 * (external/kseq.h:161-201).  Returns a handle holding the job layout
 * (bytes + colour offsets) or NULL. */
CDBG_API void* cdbg_load_fasta(const char* const* files, uint32_t n_files, uint32_t threads);
CDBG_API void* cdbg_load_fasta_list(const char* list_path, uint32_t threads);
CDBG_API uint64_t cdbg_seqs_total_bytes(void* h);
CDBG_API uint32_t cdbg_seqs_num_colors(void* h);
CDBG_API const uint8_t* cdbg_seqs_bytes(void* h);
CDBG_API const uint64_t* cdbg_seqs_offsets(void* h);
CDBG_API void cdbg_seqs_free(void* h);

/* LZ4 block codec of the bucket-file frames (public block format).  Return
 * the output length, or a negative status. */
CDBG_API int64_t cdbg_lz4_compress(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap);
CDBG_API int64_t cdbg_lz4_decompress(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap);

/* The same block format on the device, one lane per frame: the codec of the
 * bucket files' frames (LZ4_compress_default / LZ4_decompress_safe at
 * bucket_io.hpp:255-265,695-696).  compress: frame f = src[frame_off[f],
 * frame_off[f+1]) (at most 65536 bytes each) becomes dst[comp_off[f],
 * comp_off[f+1]) (comp_off has n_frames + 1 entries; dst_cap bytes of room).
 * decompress: block f = comp[comp_off[f], comp_off[f+1]) must decode to
 * exactly dst_off[f+1] - dst_off[f] bytes at dst + dst_off[f]; a malformed
 * block is an error naming the frame. */
CDBG_API int cdbg_gpu_lz4_compress_frames(cdbg_gpu_ctx* ctx, const uint8_t* src, const uint64_t* frame_off,
                                          uint32_t n_frames, uint8_t* dst, uint64_t dst_cap, uint64_t* comp_off);
CDBG_API int cdbg_gpu_lz4_decompress_frames(cdbg_gpu_ctx* ctx, const uint8_t* comp, const uint64_t* comp_off,
                                            uint32_t n_frames, uint8_t* dst, const uint64_t* dst_off);

/* Synthetic pangenome generator (seeded std::mt19937_64, explicit modulo). */
typedef struct {
    uint64_t seed;
    uint32_t num_colors;
    uint64_t genome_len;
    double snp_rate;
    double indel_rate;
    uint32_t indel_max;
    double pav_frac;
    uint32_t pav_min, pav_max;
    uint32_t n_inversions, inv_min, inv_max;
} cdbg_synth_params;

CDBG_API void* cdbg_synth_create(const cdbg_synth_params* p, uint32_t threads);
CDBG_API uint64_t cdbg_synth_total_bytes(void* h);
CDBG_API const uint8_t* cdbg_synth_bytes(void* h);
CDBG_API const uint64_t* cdbg_synth_offsets(void* h);
CDBG_API void cdbg_synth_free(void* h);
CDBG_API int cdbg_synth_write_fasta(void* h, uint32_t color, const char* path, int gz);

#ifdef __cplusplus
}
#endif

#endif /* CDBG_GPU_H */
