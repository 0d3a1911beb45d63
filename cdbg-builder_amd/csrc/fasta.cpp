// FASTA/FASTQ input (plain or gzip through zlib's gzread), with the record
// semantics of kseq_read (external/kseq.h:161-201) as used by seq_reader
// (seq_reader.hpp:44-77):
//   * records start at '>' or '@'; the header line is skipped;
//   * sequence = following lines until a line starting with '>', '@' or '+';
//     '\n' is dropped, every other byte ('\r', 'N', lower case) is kept;
//   * after a '+' line, quality bytes in [33, 127] are consumed until as many
//     as the sequence has, plus the byte that follows them; a truncated
//     quality string ends the file (kseq_read returns -2, which seq_reader
//     treats as end of input).
// A job's colour c holds its records, each followed by one '\n' separator.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "cdbg_gpu.h"
#include "common.h"
#include "lz4_block.h"

namespace cdbg_amd {

namespace {

struct GzStream {
    gzFile f = nullptr;
    std::vector<uint8_t> buf;
    size_t beg = 0, end = 0;
    bool eof = false;
    explicit GzStream(const std::string& path) : buf(128 * 1024) {
        f = gzopen(path.c_str(), "rb");
        if (!f) throw Error("could not open input file: " + path);
        gzbuffer(f, 128 * 1024);
    }
    ~GzStream() {
        if (f) gzclose(f);
    }
    int getc() {
        if (beg >= end) {
            if (eof) return -1;
            int n = gzread(f, buf.data(), unsigned(buf.size()));
            if (n <= 0) {
                eof = true;
                return -1;
            }
            beg = 0;
            end = size_t(n);
        }
        return buf[beg++];
    }
    // Appends bytes up to (excluding) the next '\n' to out; consumes the '\n'.
    void line(std::vector<uint8_t>* out) {
        for (;;) {
            if (beg >= end) {
                if (eof) return;
                int n = gzread(f, buf.data(), unsigned(buf.size()));
                if (n <= 0) {
                    eof = true;
                    return;
                }
                beg = 0;
                end = size_t(n);
            }
            const uint8_t* p = buf.data() + beg;
            const uint8_t* nl = (const uint8_t*)std::memchr(p, '\n', end - beg);
            const size_t take = nl ? size_t(nl - p) : end - beg;
            if (out) out->insert(out->end(), p, p + take);
            beg += take;
            if (nl) {
                ++beg;
                return;
            }
        }
    }
};

// Appends every record of one file to `out` (each followed by '\n').
void read_file(const std::string& path, std::vector<uint8_t>& out) {
    GzStream s(path);
    int c;
    // to the first header
    while ((c = s.getc()) != -1 && c != '>' && c != '@') {
    }
    if (c == -1) return;
    for (;;) {
        s.line(nullptr);  // header
        const size_t start = out.size();
        while ((c = s.getc()) != -1 && c != '>' && c != '+' && c != '@') {
            if (c == '\n') continue;
            out.push_back(uint8_t(c));
            s.line(&out);
        }
        const size_t len = out.size() - start;
        out.push_back('\n');
        if (c == '+') {
            // skip the rest of the '+' line; EOF there truncates the record
            while ((c = s.getc()) != -1 && c != '\n') {
            }
            bool ok = c != -1;
            if (ok) {
                size_t q = 0;
                while ((c = s.getc()) != -1 && q < len) {
                    if (c >= 33 && c <= 127) ++q;
                }
                ok = q == len;
            }
            if (!ok) {  // truncated quality: the record is dropped, input ends
                out.resize(start);
                return;
            }
            // the byte after the qualities is consumed; seek the next header
            while ((c = s.getc()) != -1 && c != '>' && c != '@') {
            }
        }
        if (c == -1) return;
    }
}

struct SeqJob {
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> offsets;
};

}  // namespace
}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

// Reads the input files (colour = position in `files`) with `threads` threads.
CDBG_API void* cdbg_load_fasta(const char* const* files, uint32_t n_files, uint32_t threads) {
    try {
        std::vector<std::vector<uint8_t>> per(n_files);
        std::atomic<uint32_t> next{0};
        std::string err;
        std::atomic<bool> failed{false};
        auto work = [&]() {
            for (;;) {
                uint32_t i = next.fetch_add(1);
                if (i >= n_files || failed) break;
                try {
                    read_file(files[i], per[i]);
                } catch (std::exception& e) {
                    if (!failed.exchange(true)) err = e.what();
                }
            }
        };
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < std::max(1u, threads); ++t) th.emplace_back(work);
        for (auto& t : th) t.join();
        if (failed) throw Error(err);
        auto* J = new SeqJob();
        J->offsets.resize(n_files + 1, 0);
        uint64_t total = 0;
        for (uint32_t i = 0; i < n_files; ++i) {
            J->offsets[i] = total;
            total += per[i].size();
        }
        J->offsets[n_files] = total;
        J->bytes.resize(total);
        for (uint32_t i = 0; i < n_files; ++i) {
            if (!per[i].empty()) std::memcpy(J->bytes.data() + J->offsets[i], per[i].data(), per[i].size());
            std::vector<uint8_t>().swap(per[i]);
        }
        return J;
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

// Reads a filenames list (one path per line, empty lines skipped: the line's
// index among non-empty lines is the colour, builder.hpp:98-113) and loads it.
CDBG_API void* cdbg_load_fasta_list(const char* list_path, uint32_t threads) {
    try {
        std::ifstream in(list_path);
        if (!in) throw Error(std::string("cannot open filenames list: ") + list_path);
        std::vector<std::string> files;
        std::string line;
        while (std::getline(in, line))
            if (!line.empty()) files.push_back(line);
        if (files.empty()) throw Error("no input files");
        std::vector<const char*> ptrs;
        for (auto& f : files) ptrs.push_back(f.c_str());
        return cdbg_load_fasta(ptrs.data(), uint32_t(ptrs.size()), threads);
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

CDBG_API uint64_t cdbg_seqs_total_bytes(void* h) { return ((SeqJob*)h)->bytes.size(); }
CDBG_API uint32_t cdbg_seqs_num_colors(void* h) { return uint32_t(((SeqJob*)h)->offsets.size() - 1); }
CDBG_API const uint8_t* cdbg_seqs_bytes(void* h) { return ((SeqJob*)h)->bytes.data(); }
CDBG_API const uint64_t* cdbg_seqs_offsets(void* h) { return ((SeqJob*)h)->offsets.data(); }
CDBG_API void cdbg_seqs_free(void* h) { delete (SeqJob*)h; }

CDBG_API int64_t cdbg_lz4_compress(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap) {
    std::vector<uint8_t> out;
    lz4_compress_block(src, n, out);
    if (out.size() > cap) return set_error("lz4: output capacity too small");
    std::memcpy(dst, out.data(), out.size());
    return int64_t(out.size());
}

CDBG_API int64_t cdbg_lz4_decompress(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap) {
    size_t len = 0;
    if (!lz4_decompress_block(src, n, dst, cap, len)) return set_error("lz4: malformed block");
    return int64_t(len);
}

}  // extern "C"
