// cdbg-build: driver of the four phases, with the reference's command line
// (src/main.cpp:13-74) and configuration rules (builder.hpp:550-649).  Phase 1
// runs on the device and writes the bucket files, phase 2 runs on the device
// over those files and leaves frag_unitigs.bin in the scratch directory and
// <out>.color_sets / <out>.color_sets.tmp_offsets as the reference leaves
// them after bucket-process; the stitch (phase 3, on the device; the host
// stitch with --host-stitch) and the host emit (phase 4) then write <out>.fa,
// <out>.u2c, the finished <out>.color_sets and <out>.metadata.txt
// (builder.hpp:72-79 order).  --phase2-only stops after phase 2 and reports
// the state the reference's own phases 3-4 need in <out>.phase2.txt.
//
// Deviations: -b accepts [1, 16] (the reference's [10, 16] rejects the
// 64-bucket configuration); k runs up to 63 in every build (64- or 128-bit
// k-mers chosen at run time instead of the CDBG_LARGE_K compile flag).
// --resident keeps the phase-1 records in HBM instead of writing bucket files.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "cdbg_gpu.h"

namespace {

struct Config {
    std::string list, out, tmp;
    uint32_t k = 0, t = 1, m = 0, b = 0;
    double g = 0, alpha = 2.0, beta = 7.0;
    uint64_t flush = 0, spill = 0;
    bool verbose = false, keep_tmp = false, resident = false, phase2_only = false, host_stitch = false;
};

void usage(const char* p) {
    std::fprintf(stderr,
                 "usage: %s -i <filenames_list> -o <out_basename> -k <k> [-t threads] [-m m] [-b buckets_log2]\n"
                 "          [-d tmp_dir] [-g max_ram_gb] [--alpha a] [--beta b] [--flush n] [--spill n]\n"
                 "          [--verbose] [--keep-tmp] [--resident] [--phase2-only] [--host-stitch]\n",
                 p);
}

bool parse(int argc, char** argv, Config& c) {
    bool has_i = false, has_o = false, has_k = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const char* name) -> std::string {
            if (i + 1 >= argc) throw std::runtime_error(std::string("missing value for ") + name);
            return argv[++i];
        };
        if (a == "-i") c.list = val("-i"), has_i = true;
        else if (a == "-o") c.out = val("-o"), has_o = true;
        else if (a == "-k") c.k = std::stoul(val("-k")), has_k = true;
        else if (a == "-t") c.t = std::stoul(val("-t"));
        else if (a == "-m") c.m = std::stoul(val("-m"));
        else if (a == "-b") c.b = std::stoul(val("-b"));
        else if (a == "-d") c.tmp = val("-d");
        else if (a == "-g") c.g = std::stod(val("-g"));
        else if (a == "--alpha") c.alpha = std::stod(val("--alpha"));
        else if (a == "--beta") c.beta = std::stod(val("--beta"));
        else if (a == "--flush") c.flush = std::stoull(val("--flush"));
        else if (a == "--spill") c.spill = std::stoull(val("--spill"));
        else if (a == "--verbose") c.verbose = true;
        else if (a == "--keep-tmp") c.keep_tmp = true;
        else if (a == "--resident") c.resident = true;
        else if (a == "--phase2-only") c.phase2_only = true;
        else if (a == "--host-stitch") c.host_stitch = true;
        else if (a == "-h" || a == "--help") return false;
        else throw std::runtime_error("unknown option " + a);
    }
    if (!has_i || !has_o || !has_k) return false;
    return true;
}

// compute_best_m (phase1_bucket_write/minimizer.hpp:65-74)
uint32_t best_m(uint32_t k) {
    if (k <= 13) return (k / 2 > k - 4) ? (k / 2) : (k > 4 ? (k - 4) : 1);
    if (k <= 15) return 9;
    if (k <= 21) return 10;
    if (k <= 30) return 11;
    if (k <= 37) return 12;
    if (k <= 42) return 13;
    if (k <= 64) return 14;
    return (k + 2) / 4;
}

// most buckets phase 1 supports (16-bit bucket ids)
constexpr uint64_t MAX_BUCKETS = 65536;

// validate_and_resolve_config / auto_bucket_count_ (builder.hpp:550-649)
uint32_t resolve(Config& c) {
    if (c.k < 2 || c.k > 63) throw std::runtime_error("k must satisfy 2 <= k <= 63");
    if (c.t == 0) c.t = 1;
    if (c.g > 0 && c.g < 4.0) throw std::runtime_error("-g must be at least 4 GiB");
    if (c.m == 0) c.m = best_m(c.k);
    if (c.m < 2 || c.m >= c.k) throw std::runtime_error("invalid m=" + std::to_string(c.m) + " (need 2 <= m < k)");
    const double flush = c.flush ? double(c.flush) : 4096.0, spill = c.spill ? double(c.spill) : 65536.0;
    const double M = 0.5 * c.g * 1024.0 * 1024.0 * 1024.0;
    const double per_bucket = c.alpha * c.t * flush + c.beta * spill;
    if (c.g > 0 && c.b == 0 && 64.0 * per_bucket > M) {
        // the per-thread bucket buffers of 64 buckets exceed half of -g
        const double room = M / 64.0 - c.beta * spill;
        const long t_max = room > 0 ? long(room / (c.alpha * flush)) : 0;
        throw std::runtime_error("RAM budget: bucket-write per-thread buffers at -t " + std::to_string(c.t) +
                                 " need " + std::to_string(uint64_t(64.0 * per_bucket)) + " bytes > " +
                                 std::to_string(uint64_t(M)) + ". " +
                                 (t_max >= 1 ? "Reduce -t to <= " + std::to_string(t_max) + ", or raise -g."
                                             : std::string("Raise -g.")));
    }
    if (c.b) {
        if (c.b < 1 || c.b > 16) throw std::runtime_error("-b must be in [1, 16]");
        return 1u << c.b;
    }
    if (c.g <= 0) return 1024;
    const double b = M / per_bucket;
    uint64_t n = b < 1.0 ? 1 : uint64_t(b);
    if (n < 64) n = 64;
    if (n > MAX_BUCKETS) {
        // bucket ids are 16-bit on the device; the reference clamps its own
        // count to the file-descriptor limit the same way (builder.hpp:654-675)
        std::cerr << "warning: -g " << c.g << " sizes " << n << " buckets; clamped to " << MAX_BUCKETS << "\n";
        n = MAX_BUCKETS;
    }
    return uint32_t(n);
}

std::string scratch_dir(const Config& c) {
    namespace fs = std::filesystem;
    if (c.tmp.empty()) {
        std::string tmpl = (fs::temp_directory_path() / "cdbg_buckets_XXXXXX").string();
        std::vector<char> buf(tmpl.begin(), tmpl.end());
        buf.push_back(0);
        if (!mkdtemp(buf.data())) throw std::runtime_error("mkdtemp failed");
        return buf.data();
    }
    std::error_code ec;
    if (fs::exists(c.tmp, ec)) {
        if (!fs::is_directory(c.tmp, ec)) throw std::runtime_error("-d " + c.tmp + " exists but is not a directory");
        if (!fs::is_empty(c.tmp, ec)) throw std::runtime_error("-d " + c.tmp + " is not empty");
    } else {
        fs::create_directories(c.tmp, ec);
        if (ec) throw std::runtime_error("cannot create -d " + c.tmp + ": " + ec.message());
    }
    return c.tmp;
}

void check(int rc) {
    if (rc != 0) throw std::runtime_error(cdbg_gpu_last_error());
}

}  // namespace

// the device stitch numbers fragment sides with 32 bits
static bool too_many(uint64_t fragments) { return fragments >= (uint64_t(1) << 30); }

int main(int argc, char** argv) {
    Config c;
    try {
        if (!parse(argc, argv, c)) {
            usage(argv[0]);
            return 1;
        }
        const uint32_t B = resolve(c);
        std::vector<std::string> files;
        {
            std::ifstream in(c.list);
            if (!in) throw std::runtime_error("cannot open filenames list: " + c.list);
            std::string line;
            while (std::getline(in, line))
                if (!line.empty()) files.push_back(line);
        }
        if (files.empty()) throw std::runtime_error("no input files");
        std::vector<const char*> fp;
        for (auto& f : files) fp.push_back(f.c_str());
        const uint32_t nc = uint32_t(files.size());
        std::cout << "k = " << c.k << ", m = " << c.m << ", num_colors = " << nc << ", num_threads = " << c.t
                  << ", num_buckets = " << B << "\n";
        const std::string tmp = scratch_dir(c);
        std::cout << "  tmp_dir = " << tmp << "\n";

        cdbg_gpu_ctx* ctx = nullptr;
        check(cdbg_gpu_ctx_create(0, &ctx));
        const std::string frag = tmp + "/frag_unitigs.bin", cs = c.out + ".color_sets";
        cdbg_gpu_color_state st{};
        cdbg_gpu_stats s{};
        void* U = nullptr;  // unitigs stitched from the resident job's fragments
        bool device_failed = false;
        if (c.resident) {
            cdbg_gpu_job* job = nullptr;
            check(cdbg_gpu_job_create_from_files(ctx, fp.data(), nc, c.k, c.m, B, c.t, &job));
            check(cdbg_gpu_job_run(job));
            check(cdbg_gpu_job_stats(job, &s));
            // the fragment stream stays in HBM for the stitch unless a file is wanted
            const bool frag_file = c.keep_tmp || c.phase2_only || c.host_stitch || too_many(s.num_frags);
            check(cdbg_gpu_job_write_phase2_files(job, frag_file ? frag.c_str() : nullptr, cs.c_str(), &st));
            std::cout << "  phase 1: " << s.ms_phase1 << " ms, phase 2: " << s.ms_phase2 << " ms (device)\n";
            if (!c.phase2_only && !c.host_stitch && !too_many(s.num_frags)) {
                U = cdbg_gpu_job_stitch(job, c.t);
                if (!U) {  // the host stitch needs the stream as a file
                    std::cerr << "warning: device stitch failed (" << cdbg_gpu_last_error()
                              << "); stitching on the host\n";
                    device_failed = true;
                    if (!frag_file)
                        check(cdbg_gpu_job_write_phase2_files(job, frag.c_str(), cs.c_str(), &st));
                }
            }
            cdbg_gpu_job_destroy(job);
        } else {
            std::vector<uint64_t> unc(B);
            check(cdbg_gpu_bucket_write(ctx, fp.data(), nc, c.k, c.m, B, tmp.c_str(), c.t, unc.data()));
            uint64_t tot = 0;
            for (uint64_t u : unc) tot += u;
            std::cout << "  bucket-write: " << tot << " uncompressed bucket bytes\n";
            uint64_t done = 0;
            check(cdbg_gpu_bucket_process_ex(ctx, tmp.c_str(), B, c.k, nc, c.t, frag.c_str(), cs.c_str(), &st, &done,
                                             0, c.keep_tmp ? 0 : 1));
            std::cout << "  bucket-process: " << done << " of " << B << " buckets\n";
        }
        std::cout << "  bucket fragments: " << st.num_fragments << "\n"
                  << "  distinct color sets: " << st.num_color_sets << "\n"
                  << "  distinct k-mers: " << st.num_kmers << "\n";
        if (c.phase2_only) {
            std::ofstream meta(c.out + ".phase2.txt");
            meta << "k=" << c.k << "\nnum_kmers=" << st.num_kmers << "\nnum_colors=" << nc
                 << "\nnum_color_sets=" << st.num_color_sets << "\nnum_fragments=" << st.num_fragments
                 << "\ncolor_set_bits=" << st.num_bits << "\ncolor_set_total_integers=" << st.total_integers
                 << "\ncolor_set_trailing_word=" << st.trailing_word << "\nfrag_unitigs=" << frag << "\n";
            cdbg_gpu_ctx_destroy(ctx);
            std::cout << "done. wrote " << cs << ", " << cs << ".tmp_offsets, " << frag << " and " << c.out
                      << ".phase2.txt\n";
            return 0;
        }
        // phase 3: stitch (on the device unless --host-stitch); phase 4: emit
        const bool host = c.host_stitch || too_many(st.num_fragments) || device_failed;
        if (host && !c.host_stitch)
            std::cout << "  stitch: " << st.num_fragments
                      << " fragments, more than the device stitch indexes (2^30): stitching on the host\n";
        if (!U && !host) {
            U = cdbg_gpu_stitch_file(ctx, frag.c_str(), c.k, c.t);
            if (!U) {  // e.g. a stream whose working set does not fit the HBM left
                std::cerr << "warning: device stitch failed (" << cdbg_gpu_last_error()
                          << "); stitching on the host\n";
                cdbg_gpu_trim_cache();
            }
        }
        if (!U) U = cdbg_stitch_file(frag.c_str(), c.k, c.t);
        cdbg_gpu_ctx_destroy(ctx);
        if (!U) throw std::runtime_error(cdbg_gpu_last_error());
        cdbg_stitch_stats ss{};
        int rc = cdbg_unitigs_stats(U, &ss);
        if (rc == 0) rc = cdbg_emit_fasta(U, c.out.c_str());
        cdbg_unitigs_free(U);
        check(rc);
        std::cout << "  stitch: " << ss.joins << " joins, " << ss.cycles << " cycles; num_unitigs = " << ss.num_unitigs
                  << "\n";
        check(cdbg_finalize_color_sets(cs.c_str(), &st));
        check(cdbg_write_metadata(c.out.c_str(), c.k, st.num_kmers, nc, ss.num_unitigs, st.num_color_sets));
        if (!c.keep_tmp) {
            std::error_code ec;
            std::filesystem::remove(frag, ec);  // absent when the resident stitch used the HBM copy
        }
        std::cout << "done. wrote " << c.out << ".fa, " << c.out << ".u2c, " << cs << " and " << c.out
                  << ".metadata.txt\n";
        return 0;
    } catch (std::exception& e) {
        std::cerr << "error: " << e.what() << '\n';
        return 1;
    }
    return 0;
}
