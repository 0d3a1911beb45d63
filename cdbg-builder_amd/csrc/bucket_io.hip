// On-disk artifacts of phases 1 and 2 (the drop-in boundary).
//
// Bucket files (bucket_io.hpp:19-40, super_kmer.hpp:43-68,105-117):
//   frames [u32 uncompressed][u32 compressed][LZ4 block] ... [u32 0];
//   decompressed stream = records
//     varint(n_colors) varint(c0) varint(c1-c0)... varint((len<<4)|flags)
//     ceil(len/4) bytes, base i in bits 2(i%4) of byte i/4.
// Phase 1 writes one record per super-k-mer occurrence (one colour), in
// colour-major position order inside each bucket: a valid bucket stream
// (the reference's online compaction only merges identical records, which
// "is purely a memory bound, never a correctness one", algorithm.md:264-266).
// Records are serialized on the device; LZ4 framing and file writes run on
// host threads.
//
// Reading a bucket file (bucket_reader, bucket_io.hpp:651-720) builds the
// phase-2 input: the records' packed bases, and one (record, colour) pair
// per colour of every record with len >= k (load_bucket skips shorter ones,
// bucket_walker.hpp:179), ordered colour-major inside the bucket.  Each pair
// carries the stream rank of its record's first k-mer, so first-occurrence
// order is that of the file.
//
// Phase-2 files: frag_unitigs.bin (unitig_spill.hpp:491-522), the colour-set
// stream <out>.color_sets as the reference leaves it at the end of phase 2 (a
// 28-byte zero header, then every complete 64-bit word;
// streaming_color_set_dict.hpp:93-97,350-365) and <out>.color_sets.tmp_offsets
// (one u64 bit offset per class, :170-175).  The trailing partial word and
// the totals are returned for the host's finalize step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <thread>

#include "job.h"
#include "kmer_ops.h"
#include "lz4_block.h"

namespace cdbg_amd {

namespace {

constexpr int TPB = 256;
constexpr size_t FRAME_BYTES = 256 * 1024;

inline uint32_t blocks_for(uint64_t n) { return uint32_t(std::max<uint64_t>(1, (n + TPB - 1) / TPB)); }

__host__ __device__ __forceinline__ uint32_t varint_len(uint64_t v) {
    uint32_t n = 1;
    while (v >= 0x80) {
        v >>= 7;
        ++n;
    }
    return n;
}
__device__ __forceinline__ uint8_t* put_varint(uint8_t* p, uint64_t v) {
    while (v >= 0x80) {
        *p++ = uint8_t((v & 0x7f) | 0x80);
        v >>= 7;
    }
    *p++ = uint8_t(v);
    return p;
}

__global__ void k_record_size(uint64_t n, const uint32_t* __restrict__ len, const uint8_t* __restrict__ flags,
                              const uint32_t* __restrict__ color, uint64_t* __restrict__ size) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t lf = (uint64_t(len[i]) << 4) | (flags[i] & 15u);
    size[i] = varint_len(1) + varint_len(color[i]) + varint_len(lf) + (uint64_t(len[i]) + 3) / 4;
}

__device__ __forceinline__ uint64_t window64(const uint32_t* __restrict__ packed, uint64_t p) {
    const uint64_t w = p >> 4;
    const uint32_t sh = 2 * uint32_t(p & 15);
    const uint64_t lo = uint64_t(packed[w]) | (uint64_t(packed[w + 1]) << 32);
    uint64_t v = lo >> sh;
    if (sh) v |= uint64_t(packed[w + 2]) << (64 - sh);
    return v;
}

__global__ void k_record_write(uint64_t n, const uint64_t* __restrict__ start, const uint32_t* __restrict__ len,
                               const uint8_t* __restrict__ flags, const uint32_t* __restrict__ color,
                               const uint32_t* __restrict__ packed, const uint64_t* __restrict__ off,
                               uint8_t* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t* p = out + off[i];
    p = put_varint(p, 1);
    p = put_varint(p, color[i]);
    const uint32_t L = len[i];
    p = put_varint(p, (uint64_t(L) << 4) | (flags[i] & 15u));
    const uint64_t s = start[i];
    const uint32_t nb = (L + 3) / 4;
    for (uint32_t j = 0; j < nb; ++j) {
        uint32_t b = uint32_t(window64(packed, s + 4ull * j)) & 0xffu;
        const uint32_t rem = L - 4 * j;
        if (rem < 4) b &= (1u << (2 * rem)) - 1;  // unused base slots stay zero
        p[j] = uint8_t(b);
    }
}

std::string bucket_path(const std::string& dir, uint32_t b) { return dir + "/bucket_" + std::to_string(b) + ".bin"; }

void write_frames(const std::string& path, const uint8_t* data, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error("cannot open bucket file: " + path);
    std::vector<uint8_t> comp;
    bool ok = true;
    for (size_t o = 0; o < n && ok; o += FRAME_BYTES) {
        const size_t u = std::min(FRAME_BYTES, n - o);
        lz4_compress_block(data + o, u, comp);
        const uint32_t hu = uint32_t(u), hc = uint32_t(comp.size());
        ok = std::fwrite(&hu, 4, 1, f) == 1 && std::fwrite(&hc, 4, 1, f) == 1 &&
             std::fwrite(comp.data(), 1, comp.size(), f) == comp.size();
    }
    const uint32_t eof = 0;
    ok = ok && std::fwrite(&eof, 4, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw Error("short write to " + path);
}

std::vector<uint8_t> read_decoded(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error("cannot open " + path);
    std::vector<uint8_t> dec, comp;
    for (;;) {
        uint32_t u = 0, c = 0;
        if (std::fread(&u, 4, 1, f) != 1) {
            if (std::feof(f)) break;  // a missing EOF marker is tolerated (bucket_io.hpp:675-679)
            std::fclose(f);
            throw Error("short read of frame header on " + path);
        }
        if (u == 0) break;
        if (std::fread(&c, 4, 1, f) != 1) {
            std::fclose(f);
            throw Error("short read of compressed_size on " + path);
        }
        comp.resize(c);
        if (std::fread(comp.data(), 1, c, f) != c) {
            std::fclose(f);
            throw Error("short read of compressed payload on " + path);
        }
        const size_t off = dec.size();
        dec.resize(off + u);
        size_t got = 0;
        if (!lz4_decompress_block(comp.data(), c, dec.data() + off, u, got) || got != u) {
            std::fclose(f);
            throw Error("LZ4 decode failed on " + path);
        }
    }
    std::fclose(f);
    return dec;
}

uint64_t get_varint(const uint8_t* buf, size_t n, size_t& pos) {  // super_kmer.hpp:71-82
    uint64_t v = 0;
    unsigned sh = 0;
    while (pos < n) {
        const uint8_t b = buf[pos++];
        v |= uint64_t(b & 0x7f) << sh;
        if (!(b & 0x80)) return v;
        sh += 7;
        if (sh >= 64) break;
    }
    return v;
}

// One bucket's parsed records, ready to concatenate.
struct ParsedBucket {
    std::vector<uint8_t> bases;  // packed payloads of records with len >= k
    std::vector<uint64_t> rec_byte;  // payload byte offset inside `bases`
    std::vector<uint32_t> rec_len;
    std::vector<uint8_t> rec_flags;
    std::vector<uint32_t> rec_rank0;
    std::vector<uint32_t> pair_rec, pair_color;  // colour-major
    uint64_t records_read = 0;
};

void parse_bucket(const std::string& path, uint32_t k, ParsedBucket& P) {
    const std::vector<uint8_t> d = read_decoded(path);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;  // (colour, record)
    std::vector<uint32_t> colors;
    size_t q = 0;
    uint64_t rank = 0;
    // read_super_kmer (super_kmer.hpp:121-143): stop at the first malformed record
    while (q < d.size()) {
        const uint64_t nc = get_varint(d.data(), d.size(), q);
        colors.clear();
        uint32_t prev = 0;
        for (uint64_t i = 0; i < nc && q < d.size(); ++i) {
            prev += uint32_t(get_varint(d.data(), d.size(), q));
            colors.push_back(prev);
        }
        if (q >= d.size() || colors.size() != nc) break;
        const uint64_t lf = get_varint(d.data(), d.size(), q);
        const uint64_t len = lf >> 4;
        const size_t nbytes = size_t((len + 3) / 4);
        if (q + nbytes > d.size()) break;
        ++P.records_read;
        if (len >= k) {
            const uint32_t r = uint32_t(P.rec_len.size());
            P.rec_byte.push_back(P.bases.size());
            P.bases.insert(P.bases.end(), d.begin() + q, d.begin() + q + nbytes);
            P.rec_len.push_back(uint32_t(len));
            P.rec_flags.push_back(uint8_t(lf & 15));
            P.rec_rank0.push_back(uint32_t(rank));
            rank += len - k + 1;
            if (rank >= (1ull << 32)) throw Error("bucket holds more than 2^32 k-mer occurrences: " + path);
            for (uint32_t c : colors) pairs.push_back({c, r});
        }
        q += nbytes;
    }
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                         return a.first < b.first;
                     });
    P.pair_rec.resize(pairs.size());
    P.pair_color.resize(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) {
        P.pair_color[i] = pairs[i].first;
        P.pair_rec[i] = pairs[i].second;
    }
}

template <class F>
void parallel_for(uint32_t n, uint32_t threads, F&& f) {
    std::atomic<uint32_t> next{0};
    std::string err;
    std::atomic<bool> failed{false};
    auto work = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n || failed) break;
            try {
                f(i);
            } catch (std::exception& e) {
                if (!failed.exchange(true)) err = e.what();
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < std::max(1u, std::min(threads, n)); ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
    if (failed) throw Error(err);
}

}  // namespace

void write_bucket_files(Job& J, const std::string& dir, uint32_t threads, uint64_t* unc_bytes) {
    AllocStream alloc_on_job_stream(J.st);
    if (!J.phase1_done || !J.bytes.p) throw Error("bucket files need phase 1 of a sequence job");
    hipStream_t st = J.st;
    Phase1Bufs& Q = J.p1;
    const uint64_t n = J.n_supers;
    DevBuf d_size, d_off, d_total, d_out, scan_tmp;
    uint64_t* size = d_size.ensure<uint64_t>(n);
    uint64_t* off = d_off.ensure<uint64_t>(n + 1);
    if (n) k_record_size<<<blocks_for(n), TPB, 0, st>>>(n, Q.b_len.as<uint32_t>(), Q.b_flags.as<uint8_t>(), Q.b_color.as<uint32_t>(), size);
    exclusive_scan<uint64_t, uint64_t>(size, off, n, off + n, scan_tmp, st);
    uint64_t total = 0;
    CDBG_HIP(hipMemcpyAsync(&total, off + n, 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    uint8_t* out = d_out.ensure<uint8_t>(total);
    if (n) k_record_write<<<blocks_for(n), TPB, 0, st>>>(n, Q.b_start.as<uint64_t>(), Q.b_len.as<uint32_t>(), Q.b_flags.as<uint8_t>(), Q.b_color.as<uint32_t>(), Q.packed.as<uint32_t>(), off, out);
    CDBG_HIP(hipGetLastError());
    std::vector<uint8_t> h(total);
    std::vector<uint64_t> h_off(n + 1);
    if (total) CDBG_HIP(hipMemcpyAsync(h.data(), out, total, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipMemcpyAsync(h_off.data(), off, (n + 1) * 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    std::filesystem::create_directories(dir);
    const auto& bo = J.h_bucket_off;
    parallel_for(J.B, threads, [&](uint32_t b) {
        const uint64_t a = h_off[bo[b]], e = h_off[bo[b + 1]];
        write_frames(bucket_path(dir, b), h.data() + a, e - a);
        if (unc_bytes) unc_bytes[b] = e - a;
    });
}

void load_bucket_files(Job& J, const std::string& dir, uint32_t threads) {
    AllocStream alloc_on_job_stream(J.st);
    const uint32_t B = J.B, k = J.k;
    std::vector<ParsedBucket> pb(B);
    parallel_for(B, threads, [&](uint32_t b) { parse_bucket(bucket_path(dir, b), k, pb[b]); });
    // concatenate in bucket order
    uint64_t n_bytes = 0, n_rec = 0, n_pair = 0;
    std::vector<uint64_t> byte_base(B), rec_base(B);
    for (uint32_t b = 0; b < B; ++b) {
        byte_base[b] = n_bytes;
        rec_base[b] = n_rec;
        n_bytes += pb[b].bases.size();
        n_rec += pb[b].rec_len.size();
        n_pair += pb[b].pair_rec.size();
    }
    if (n_pair >= (1ull << 32)) throw Error("more than 2^32 (record, colour) pairs in one job");
    std::vector<uint8_t> bases(n_bytes + 64, 0);
    std::vector<uint64_t> start(n_pair);
    std::vector<uint32_t> len(n_pair), color(n_pair), bucket(n_pair), rank0(n_pair);
    std::vector<uint8_t> flags(n_pair);
    J.h_bucket_off.assign(B + 1, 0);
    uint64_t pi = 0;
    for (uint32_t b = 0; b < B; ++b) {
        ParsedBucket& P = pb[b];
        if (!P.bases.empty()) std::memcpy(bases.data() + byte_base[b], P.bases.data(), P.bases.size());
        J.h_bucket_off[b] = pi;
        for (size_t i = 0; i < P.pair_rec.size(); ++i, ++pi) {
            const uint32_t r = P.pair_rec[i];
            start[pi] = 4 * (byte_base[b] + P.rec_byte[r]);
            len[pi] = P.rec_len[r];
            flags[pi] = P.rec_flags[r];
            rank0[pi] = P.rec_rank0[r];
            color[pi] = P.pair_color[i];
            bucket[pi] = b;
            if (color[pi] >= J.num_colors) throw Error("bucket " + std::to_string(b) + " names colour " +
                                                       std::to_string(color[pi]) + " >= num_colors");
        }
        J.stats.num_supers += P.records_read;
        ParsedBucket().bases.swap(P.bases);
    }
    J.h_bucket_off[B] = pi;
    hipStream_t st = J.st;
    Phase1Bufs& Q = J.p1;
    auto up = [&](DevBuf& d, const void* src, size_t bytes) {
        d.ensure<uint8_t>(bytes);
        if (bytes) CDBG_HIP(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, st));
    };
    up(Q.packed, bases.data(), bases.size());
    up(Q.b_start, start.data(), n_pair * 8);
    up(Q.b_len, len.data(), n_pair * 4);
    up(Q.b_flags, flags.data(), n_pair);
    up(Q.b_color, color.data(), n_pair * 4);
    up(Q.key, bucket.data(), n_pair * 4);
    up(Q.rank0, rank0.data(), n_pair * 4);
    up(Q.bucket_off, J.h_bucket_off.data(), (B + 1) * 8);
    CDBG_HIP(hipStreamSynchronize(st));  // host vectors are pageable and local
    J.n_supers = n_pair;
    J.phase1_done = true;
    J.from_files = true;
}

void write_phase2_files(Job& J, const std::string& frag_path, const std::string& cs_path,
                        cdbg_gpu_color_state* state) {
    AllocStream alloc_on_job_stream(J.st);
    if (!J.phase2_done) throw Error("phase 2 has not run");
    const cdbg_gpu_stats& s = J.stats;
    std::vector<uint8_t> frags(s.frag_bytes);
    const uint64_t nw = (s.cs_bits + 63) / 64;
    std::vector<uint64_t> words(nw), offs(s.num_color_sets);
    copy_phase2_results(J, frags.data(), words.data(), offs.data());
    auto write_all = [](const std::string& path, const void* p, size_t n, const char* mode) {
        FILE* f = std::fopen(path.c_str(), mode);
        if (!f) throw Error("cannot open " + path);
        const bool ok = (n == 0 || std::fwrite(p, 1, n, f) == n);
        if (std::fclose(f) != 0 || !ok) throw Error("short write to " + path);
    };
    write_all(frag_path, frags.data(), frags.size(), "wb");
    // colour-set stream as left by phase 2: zero header + complete words
    std::vector<uint8_t> cs(28, 0);
    const uint64_t full = s.cs_bits / 64;
    cs.insert(cs.end(), (const uint8_t*)words.data(), (const uint8_t*)(words.data() + full));
    write_all(cs_path, cs.data(), cs.size(), "wb");
    write_all(cs_path + ".tmp_offsets", offs.data(), offs.size() * 8, "wb");
    if (state) {
        state->num_colors = J.num_colors;
        state->sparse_threshold = J.num_colors / 4;
        state->dense_threshold = uint32_t(uint64_t(J.num_colors) * 3 / 4);
        state->num_color_sets = s.num_color_sets;
        state->num_bits = s.cs_bits;
        state->total_integers = s.total_integers;
        state->trailing_word = (s.cs_bits % 64) ? words[full] : 0;
        state->num_kmers = s.num_kmers;
        state->num_fragments = s.num_frags;
    }
}

}  // namespace cdbg_amd
