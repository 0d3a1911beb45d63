// Phase 4 (emit) on the host, and the C ABI of phases 3 and 4.
//
// <out>.fa: the unitigs in ascending colour-set id, each as ">cid\n<seq>\n"
// (emit_fasta, phase4_emit/emit.hpp:96-165); <out>.u2c: a bit vector of one
// bit per unitig, bit i set iff unitig i ends a run of equal ids, stored as
// [u64 num_bits][u64 num_words][words] (emit.hpp:27-82,131-166);
// <out>.color_sets: the 28-byte placeholder and the complete words phase 2
// wrote are finished with the trailing partial word, the Elias-Fano sequence
// of the num_sets + 1 class bit offsets, and the header [u32 num_colors][u32
// sparse][u32 dense][u64 num_bits][u64 num_words]
// (streaming_color_set_dict.hpp:221-294); <out>.metadata.txt: k, num_kmers,
// num_colors, num_unitigs, num_color_sets, one "key=value" per line
// (builder.hpp:752-764).
//
// The Elias-Fano tail is written in the layout of jermp/bits'
// elias_fano<false, false> serialised by essentials' generic saver (a POD as
// its bytes, a vector as a u64 length and its elements): u64 back, the high
// bits as a bit vector, the select index of the high bits (darray over the
// ones: u64 positions, then block inventory i64[], subblock inventory u16[]
// and overflow positions u64[]; blocks of 1024 ones, subblocks of 32), an
// empty index of the zeros, then the low bits as a compact vector (u64 size,
// u64 width, u64 mask, u64[] words with one spare word).  That library is
// not in the reference tree, so this layout is restated from its published
// design, not checked against it.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "stitch.h"

namespace cdbg_amd {

int set_error(const std::string& msg);

namespace {

struct File {
    std::FILE* f = nullptr;
    std::string path;
    File(const std::string& p, const char* mode) : path(p) {
        f = std::fopen(p.c_str(), mode);
        if (!f) throw Error("cannot open " + p);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) throw Error("short write to " + path);
    }
    template <class T>
    void pod(const T& v) {
        write(&v, sizeof(T));
    }
    template <class T>
    void vec(const std::vector<T>& v) {
        pod(uint64_t(v.size()));
        write(v.data(), v.size() * sizeof(T));
    }
    void close() {
        if (f && std::fclose(f) != 0) {
            f = nullptr;
            throw Error("cannot close " + path);
        }
        f = nullptr;
    }
};

inline uint32_t msb64(uint64_t x) { return 63u - uint32_t(__builtin_clzll(x)); }

// select index over the one bits of `words` (num_bits bits)
struct DArray {
    uint64_t positions = 0;
    std::vector<int64_t> blocks;
    std::vector<uint16_t> subblocks;
    std::vector<uint64_t> overflow;
    static constexpr size_t BLOCK = 1024, SUB = 32, MAX_SPAN = 1u << 16;

    void flush(std::vector<uint64_t>& cur) {
        if (cur.empty()) return;
        if (cur.back() - cur.front() < MAX_SPAN) {
            blocks.push_back(int64_t(cur.front()));
            for (size_t i = 0; i < cur.size(); i += SUB) subblocks.push_back(uint16_t(cur[i] - cur.front()));
        } else {
            blocks.push_back(-int64_t(overflow.size()) - 1);
            overflow.insert(overflow.end(), cur.begin(), cur.end());
            for (size_t i = 0; i < cur.size(); i += SUB) subblocks.push_back(uint16_t(0xFFFF));
        }
        cur.clear();
    }
    void build(const std::vector<uint64_t>& words, uint64_t num_bits) {
        std::vector<uint64_t> cur;
        for (uint64_t w = 0; w < words.size(); ++w) {
            uint64_t x = words[w];
            while (x) {
                const uint64_t pos = (w << 6) + uint64_t(__builtin_ctzll(x));
                x &= x - 1;
                if (pos >= num_bits) break;
                cur.push_back(pos);
                ++positions;
                if (cur.size() == BLOCK) flush(cur);
            }
        }
        flush(cur);
    }
    void save(File& f) const {
        f.pod(positions);
        f.vec(blocks);
        f.vec(subblocks);
        f.vec(overflow);
    }
};

// Elias-Fano of the non-decreasing v[0..n) below universe u (written after them)
void save_elias_fano(File& f, const std::vector<uint64_t>& v, uint64_t u) {
    const uint64_t n = v.size();
    const uint32_t l = (n && u / n) ? msb64(u / n) : 0;
    const uint64_t hb_bits = n + (u >> l) + 1;
    std::vector<uint64_t> high((hb_bits + 63) / 64, 0);
    const uint64_t low_bits = n * l;
    std::vector<uint64_t> low((low_bits + 63) / 64 + 1, 0);
    const uint64_t mask = l == 64 ? ~0ull : (uint64_t(1) << l) - 1;
    uint64_t last = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t x = v[i];
        if (i && x < last) throw Error("color_sets: class offsets are not sorted");
        if (x > u) throw Error("color_sets: a class offset lies beyond the stream");
        if (l) {
            const uint64_t lo = x & mask, pos = i * l;
            low[pos >> 6] |= lo << (pos & 63);
            if ((pos & 63) + l > 64) low[(pos >> 6) + 1] |= lo >> (64 - (pos & 63));
        }
        const uint64_t hp = (x >> l) + i;
        high[hp >> 6] |= uint64_t(1) << (hp & 63);
        last = x;
    }
    DArray d1, d0;
    d1.build(high, hb_bits);
    f.pod(last);  // back
    f.pod(hb_bits);
    f.vec(high);
    d1.save(f);
    d0.save(f);  // no index of the zeros
    f.pod(n);
    f.pod(uint64_t(l));
    f.pod(mask);
    f.vec(low);
}

char* put_u64(char* p, uint64_t v) { return std::to_chars(p, p + 20, v).ptr; }

// fn(begin, end) over up to `threads` contiguous ranges of [0, n)
template <class F>
void parallel_ranges(uint64_t n, uint32_t threads, F&& fn) {
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads ? threads : 1, n));
    if (T == 1) {
        fn(uint64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < T; ++t) th.emplace_back([&, t] { fn(n * t / T, n * (t + 1) / T); });
    for (auto& x : th) x.join();
}

}  // namespace

void emit_fasta(Unitigs& U, const std::string& out) {
    U.build_order();
    const uint64_t n = U.size();
    // .u2c: bit i set iff unitig i ends a run of equal ids
    {
        File u2c(out + ".u2c", "wb");
        u2c.pod(n);
        const uint64_t nw = (n + 63) / 64;
        u2c.pod(nw);
        std::vector<uint64_t> words(nw, 0);
        parallel_ranges(nw, U.threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t w = b; w < e; ++w) {
                uint64_t word = 0;
                const uint64_t i1 = std::min(n, 64 * w + 64);
                for (uint64_t i = 64 * w; i < i1; ++i)
                    if (i + 1 == n || U.item(i + 1).cid != U.item(i).cid) word |= uint64_t(1) << (i & 63);
                words[w] = word;
            }
        });
        u2c.write(words.data(), nw * 8);
        u2c.close();
    }
    // .fa: ">cid\n<seq>\n" per unitig; contiguous ranges of unitigs are sized,
    // placed by a prefix sum and formatted and written in parallel
    const std::string fa_path = out + ".fa";
    const int fd = ::open(fa_path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) throw Error("cannot open " + fa_path);
    try {
        const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(U.threads, n / 4096 + 1));
        std::vector<uint64_t> at(T + 1, 0);
        auto digits = [](uint64_t v) {
            uint32_t d = 1;
            while (v >= 10) {
                v /= 10;
                ++d;
            }
            return d;
        };
        parallel_ranges(T, T, [&](uint64_t b, uint64_t e) {
            for (uint64_t t = b; t < e; ++t) {
                uint64_t bytes = 0;
                for (uint64_t i = n * t / T; i < n * (t + 1) / T; ++i) bytes += 3 + digits(U.item(i).cid) + U.item(i).len;
                at[t + 1] = bytes;
            }
        });
        for (uint32_t t = 0; t < T; ++t) at[t + 1] += at[t];
        // the file's extent up front (measured: 3.3 GB in 0.28 s of parallel
        // writes, against 0.33-0.55 s growing it write by write); a file
        // system that cannot allocate it is written all the same
        if (at[T]) (void)posix_fallocate(fd, 0, off_t(at[T]));
        std::atomic<bool> failed{false};
        parallel_ranges(T, T, [&](uint64_t b, uint64_t e) {
            for (uint64_t t = b; t < e; ++t) {
                std::vector<char> buf(size_t(1) << 22);
                size_t pos = 0;
                uint64_t where = at[t];
                auto flush = [&]() {
                    size_t done = 0;
                    while (done < pos) {
                        const ssize_t w = ::pwrite(fd, buf.data() + done, pos - done, off_t(where + done));
                        if (w <= 0) {
                            failed = true;
                            return;
                        }
                        done += size_t(w);
                    }
                    where += pos;
                    pos = 0;
                };
                for (uint64_t i = n * t / T; i < n * (t + 1) / T && !failed; ++i) {
                    const auto& it = U.item(i);
                    if (pos + 24 > buf.size()) flush();
                    buf[pos++] = '>';
                    pos = size_t(put_u64(buf.data() + pos, it.cid) - buf.data());
                    buf[pos++] = '\n';
                    const char* sq = U.seq(i);
                    uint64_t left = it.len;
                    while (left) {
                        if (pos == buf.size()) flush();
                        const size_t take = std::min<uint64_t>(left, buf.size() - pos);
                        std::memcpy(buf.data() + pos, sq, take);
                        pos += take;
                        sq += take;
                        left -= take;
                    }
                    if (pos == buf.size()) flush();
                    buf[pos++] = '\n';
                }
                flush();
            }
        });
        if (failed) throw Error("short write to " + fa_path);
        if (::close(fd) != 0) throw Error("cannot close " + fa_path);
    } catch (...) {
        ::close(fd);
        throw;
    }
}

void finalize_color_sets(const std::string& cs_path, const cdbg_gpu_color_state& s) {
    namespace fs = std::filesystem;
    const std::string off_path = cs_path + ".tmp_offsets";
    const uint64_t full = s.num_bits / 64, words = (s.num_bits + 63) / 64;
    std::error_code ec;
    const uint64_t have = fs::file_size(cs_path, ec);
    if (ec || have != 28 + 8 * full)
        throw Error(cs_path + ": expected the 28-byte placeholder and " + std::to_string(full) +
                    " complete words from phase 2");
    const uint64_t osz = fs::file_size(off_path, ec);
    if (ec || osz != 8 * s.num_color_sets) throw Error(off_path + ": expected one offset per colour set");
    std::vector<uint64_t> offs(s.num_color_sets + 1);
    {
        std::FILE* f = std::fopen(off_path.c_str(), "rb");
        if (!f) throw Error("cannot open " + off_path);
        const size_t got = s.num_color_sets ? std::fread(offs.data(), 8, s.num_color_sets, f) : 0;
        std::fclose(f);
        if (got != s.num_color_sets) throw Error("short read of " + off_path);
    }
    offs[s.num_color_sets] = s.num_bits;
    File f(cs_path, "r+b");
    if (std::fseek(f.f, 0, SEEK_END) != 0) throw Error("cannot seek " + cs_path);
    if (words > full) f.pod(s.trailing_word);
    save_elias_fano(f, offs, s.num_bits);
    if (std::fseek(f.f, 0, SEEK_SET) != 0) throw Error("cannot seek " + cs_path);
    f.pod(s.num_colors);
    f.pod(s.sparse_threshold);
    f.pod(s.dense_threshold);
    f.pod(s.num_bits);
    f.pod(words);
    f.close();
    fs::remove(off_path, ec);
}

void write_metadata(const std::string& out, uint32_t k, uint64_t num_kmers, uint32_t num_colors,
                    uint64_t num_unitigs, uint64_t num_color_sets) {
    std::ofstream m(out + ".metadata.txt");
    if (!m) throw Error("cannot open " + out + ".metadata.txt");
    m << "k=" << k << "\nnum_kmers=" << num_kmers << "\nnum_colors=" << num_colors << "\nnum_unitigs=" << num_unitigs
      << "\nnum_color_sets=" << num_color_sets << "\n";
    if (!m) throw Error("failed writing " + out + ".metadata.txt");
}

}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

CDBG_API void* cdbg_stitch(const uint8_t* frags, uint64_t n_bytes, uint32_t k, uint32_t threads) {
    try {
        if (!frags && n_bytes) throw Error("null argument");
        return stitch_fragments(frags, n_bytes, k, threads);
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

CDBG_API void* cdbg_stitch_file(const char* frag_path, uint32_t k, uint32_t threads) {
    try {
        if (!frag_path) throw Error("null argument");
        uint64_t n = 0;
        const std::unique_ptr<uint8_t[]> buf = read_file(frag_path, threads, n);
        return stitch_fragments(buf.get(), n, k, threads);
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

CDBG_API int cdbg_frag_offsets(const uint8_t* frags, uint64_t n_bytes, uint32_t threads, uint64_t* out, uint64_t cap,
                               uint64_t* count) {
    try {
        if ((!frags && n_bytes) || !count || (!out && cap)) throw Error("null argument");
        const std::vector<uint64_t> off = fragment_offsets(frags, n_bytes, threads);
        *count = off.size();
        if (off.size() > cap) throw Error("cdbg_frag_offsets: output capacity too small");
        if (!off.empty()) std::memcpy(out, off.data(), off.size() * 8);
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API int cdbg_unitigs_stats(void* h, cdbg_stitch_stats* out) {
    if (!h || !out) return set_error("null argument");
    *out = ((Unitigs*)h)->stats;
    return 0;
}

CDBG_API uint64_t cdbg_unitigs_bases(void* h) {
    uint64_t b = 0;
    for (auto& l : ((Unitigs*)h)->lists)
        for (uint64_t i = 0; i < l.size(); ++i) b += l.item(i).len;
    return b;
}

CDBG_API int cdbg_unitigs_export(void* h, char* seq, uint64_t* offsets, uint64_t* cids) {
    try {
        if (!h) throw Error("null argument");
        Unitigs& U = *(Unitigs*)h;
        U.build_order();
        uint64_t at = 0;
        const uint64_t n = U.size();
        for (uint64_t i = 0; i < n; ++i) {
            const auto& it = U.item(i);
            if (seq) std::memcpy(seq + at, U.seq(i), it.len);
            if (offsets) offsets[i] = at;
            if (cids) cids[i] = it.cid;
            at += it.len;
        }
        if (offsets) offsets[n] = at;
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API void cdbg_unitigs_free(void* h) { delete (Unitigs*)h; }

CDBG_API int cdbg_emit_fasta(void* h, const char* out_basename) {
    try {
        if (!h || !out_basename) throw Error("null argument");
        emit_fasta(*(Unitigs*)h, out_basename);
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API int cdbg_finalize_color_sets(const char* color_sets_path, const cdbg_gpu_color_state* state) {
    try {
        if (!color_sets_path || !state) throw Error("null argument");
        finalize_color_sets(color_sets_path, *state);
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API int cdbg_write_metadata(const char* out_basename, uint32_t k, uint64_t num_kmers, uint32_t num_colors,
                                 uint64_t num_unitigs, uint64_t num_color_sets) {
    try {
        if (!out_basename) throw Error("null argument");
        write_metadata(out_basename, k, num_kmers, num_colors, num_unitigs, num_color_sets);
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

}  // extern "C"
