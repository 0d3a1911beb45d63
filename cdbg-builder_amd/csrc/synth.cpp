// Seeded synthetic pangenome generator for the build's workloads.
//
// A base genome of uniform random ACGT is mutated once per colour.  Every
// random draw is a raw std::mt19937_64 word reduced with an explicit modulo,
// so the output is identical across standard libraries (the std::uniform_*
// distributions are implementation-defined).  Mutations are applied in a
// fixed order: inversions (in place, reverse complement), presence/absence
// deletions, SNPs (in place), then small insertions/deletions.
//
// Layout of the generated job: colour c occupies bytes
// [offsets[c], offsets[c+1]) of one buffer: its sequence followed by one '\n'
// separator, so no ACGT run crosses a colour boundary.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "cdbg_gpu.h"
#include "common.h"

namespace cdbg_amd {

namespace {

inline uint64_t uni(std::mt19937_64& r, uint64_t n) { return n ? r() % n : 0; }
inline uint64_t uni_range(std::mt19937_64& r, uint64_t lo, uint64_t hi) {  // [lo, hi]
    return lo + uni(r, hi - lo + 1);
}
const char ACGT[4] = {'A', 'C', 'G', 'T'};

inline char comp(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        default: return 'A';
    }
}

std::string mutate(const std::string& base, const cdbg_synth_params& p, uint32_t color) {
    std::mt19937_64 rng(p.seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL * (color + 1));
    std::string g = base;
    // Inversions of U[inv_min, inv_max] bp at uniform start positions.
    for (uint32_t i = 0; i < p.n_inversions && g.size() > p.inv_max; ++i) {
        uint64_t len = uni_range(rng, p.inv_min, p.inv_max);
        uint64_t s = uni(rng, g.size() - len + 1);
        std::reverse(g.begin() + s, g.begin() + s + len);
        for (uint64_t j = s; j < s + len; ++j) g[j] = comp(g[j]);
    }
    // Presence/absence deletions until `pav_frac` of the base length is gone.
    if (p.pav_frac > 0) {
        uint64_t target = (uint64_t)(p.pav_frac * (double)base.size());
        std::vector<std::pair<uint64_t, uint64_t>> iv;
        uint64_t removed = 0;
        while (removed < target && g.size() > p.pav_max * 2) {
            uint64_t len = uni_range(rng, p.pav_min, p.pav_max);
            uint64_t s = uni(rng, g.size() - len + 1);
            iv.push_back({s, s + len});
            removed += len;
        }
        std::sort(iv.begin(), iv.end());
        std::string out;
        out.reserve(g.size());
        uint64_t cur = 0;
        for (auto& e : iv) {
            if (e.first > cur) out.append(g, cur, e.first - cur);
            cur = std::max(cur, e.second);
        }
        if (cur < g.size()) out.append(g, cur, std::string::npos);
        g.swap(out);
    }
    // SNPs: substitute a uniformly chosen different base.
    uint64_t n_snp = (uint64_t)(p.snp_rate * (double)g.size() + 0.5);
    for (uint64_t i = 0; i < n_snp; ++i) {
        uint64_t pos = uni(rng, g.size());
        int cur = g[pos] == 'A' ? 0 : g[pos] == 'C' ? 1 : g[pos] == 'G' ? 2 : 3;
        g[pos] = ACGT[(cur + 1 + uni(rng, 3)) & 3];
    }
    // Indels: half insertions, half deletions, lengths U[1, indel_max].
    uint64_t n_indel = (uint64_t)(p.indel_rate * (double)g.size() + 0.5);
    struct Ev {
        uint64_t pos;
        uint32_t len;
        bool ins;
        uint64_t seed;
    };
    std::vector<Ev> ev(n_indel);
    for (auto& e : ev) {
        e.pos = uni(rng, g.size());
        e.ins = (rng() & 1) != 0;
        e.len = (uint32_t)uni_range(rng, 1, p.indel_max ? p.indel_max : 1);
        e.seed = rng();
    }
    std::sort(ev.begin(), ev.end(), [](const Ev& a, const Ev& b) { return a.pos < b.pos; });
    std::string out;
    out.reserve(g.size() + n_indel * 6);
    uint64_t cur = 0;
    for (auto& e : ev) {
        if (e.pos < cur) continue;  // swallowed by a previous deletion
        out.append(g, cur, e.pos - cur);
        cur = e.pos;
        if (e.ins) {
            std::mt19937_64 r2(e.seed);
            for (uint32_t j = 0; j < e.len; ++j) out.push_back(ACGT[r2() & 3]);
        } else {
            cur = std::min<uint64_t>(g.size(), cur + e.len);
        }
    }
    if (cur < g.size()) out.append(g, cur, std::string::npos);
    return out;
}

}  // namespace

struct SynthJob {
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> offsets;
};

}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

CDBG_API void* cdbg_synth_create(const cdbg_synth_params* p, uint32_t threads) {
    try {
        if (!p || p->num_colors == 0 || p->genome_len == 0) throw std::runtime_error("bad synth params");
        std::mt19937_64 rng(p->seed);
        std::string base(p->genome_len, 'A');
        for (auto& c : base) c = ACGT[rng() & 3];
        std::vector<std::string> g(p->num_colors);
        std::atomic<uint32_t> next{0};
        auto work = [&]() {
            for (;;) {
                uint32_t c = next.fetch_add(1);
                if (c >= p->num_colors) break;
                g[c] = mutate(base, *p, c);
            }
        };
        if (threads < 1) threads = 1;
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < threads; ++t) th.emplace_back(work);
        for (auto& t : th) t.join();
        auto* J = new SynthJob();
        J->offsets.resize(p->num_colors + 1, 0);
        uint64_t total = 0;
        for (uint32_t c = 0; c < p->num_colors; ++c) {
            J->offsets[c] = total;
            total += g[c].size() + 1;
        }
        J->offsets[p->num_colors] = total;
        J->bytes.resize(total);
        for (uint32_t c = 0; c < p->num_colors; ++c) {
            std::memcpy(J->bytes.data() + J->offsets[c], g[c].data(), g[c].size());
            J->bytes[J->offsets[c + 1] - 1] = '\n';
            std::string().swap(g[c]);
        }
        return J;
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

CDBG_API uint64_t cdbg_synth_total_bytes(void* h) { return ((SynthJob*)h)->bytes.size(); }
CDBG_API const uint8_t* cdbg_synth_bytes(void* h) { return ((SynthJob*)h)->bytes.data(); }
CDBG_API const uint64_t* cdbg_synth_offsets(void* h) { return ((SynthJob*)h)->offsets.data(); }
CDBG_API void cdbg_synth_free(void* h) { delete (SynthJob*)h; }

// Write colour c as FASTA (80-column lines, header ">color_<c>") to `path`,
// gzip-compressed at level 6 when `gz` is non-zero.
CDBG_API int cdbg_synth_write_fasta(void* h, uint32_t c, const char* path, int gz) {
    SynthJob* J = (SynthJob*)h;
    if (c + 1 >= J->offsets.size()) return set_error("colour out of range");
    const uint8_t* s = J->bytes.data() + J->offsets[c];
    uint64_t n = J->offsets[c + 1] - J->offsets[c] - 1;
    std::string txt = ">color_" + std::to_string(c) + "\n";
    txt.reserve(n + n / 80 + 64);
    for (uint64_t i = 0; i < n; i += 80) {
        txt.append((const char*)s + i, std::min<uint64_t>(80, n - i));
        txt.push_back('\n');
    }
    if (gz) {
        gzFile f = gzopen(path, "wb6");
        if (!f) return set_error(std::string("cannot open ") + path);
        size_t off = 0;
        while (off < txt.size()) {
            unsigned chunk = (unsigned)std::min<size_t>(1u << 30, txt.size() - off);
            if (gzwrite(f, txt.data() + off, chunk) != (int)chunk) {
                gzclose(f);
                return set_error(std::string("gzwrite failed on ") + path);
            }
            off += chunk;
        }
        gzclose(f);
    } else {
        FILE* f = fopen(path, "wb");
        if (!f) return set_error(std::string("cannot open ") + path);
        if (fwrite(txt.data(), 1, txt.size(), f) != txt.size()) {
            fclose(f);
            return set_error(std::string("short write to ") + path);
        }
        fclose(f);
    }
    return 0;
}

}  // extern "C"
