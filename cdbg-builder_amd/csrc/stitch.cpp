// Phase 3 (stitch) on the host: joins the open ends of the phase-2 fragments
// into maximal tigs and cuts every tig into monochromatic unitigs.
//
// Semantics (stitch_extmem.hpp:112-355): an open fragment side continues into
// another bucket, and the two fragments that meet there both hold the full
// boundary k-mer, so open sides are keyed on their canonical boundary k-mer;
// two sides with the same key and compatible orientation join
// (ext_pair_compatible :296-300), the shared k-mer is kept once, and its
// colour comes from the side that owns it (ext_concat_runs :170-226).  A tig
// is cut at colour-set changes; foreign placeholder runs at a tig's ends are
// dropped, their k-mers being owned and emitted elsewhere
// (ext_split_monochromatic :112-151).
//
// Method (this file's own): every open side's key is computed in parallel,
// the keys are bucketed by hash and sorted per bucket, equal keys pair up,
// and each chain of joined fragments is walked once from the end with the
// smaller (fragment, side), found first by following the pairing alone; the
// walks run in parallel over contiguous ranges of chains.  Chains that close
// on themselves are what is left: they are walked in fragment order, kept as
// one cycle (the shared k-mer once) and, when they carry several colour
// sets, rotated to start at a colour change before they are cut.  The result
// is the reference's unitig set (its round-based doubling join reaches the
// same pairs); ids and order follow this file's rule: ascending colour-set
// id, then chain order.
#include "stitch.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace cdbg_amd {

int set_error(const std::string& msg);

namespace {

using u128 = unsigned __int128;
constexpr uint64_t NONE = ~0ull;

inline uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint64_t rd64(const uint8_t* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

struct Run {
    uint64_t cid;
    uint32_t n;
};

// One frag_unitigs.bin record: [u8 open][u32 nruns]{[u64 cid][u32 n]}[u32 len][seq]
// (unitig_spill.hpp:491-522, parser :646-665).
struct FragView {
    uint8_t open;
    uint32_t nruns;
    const uint8_t* runs;
    uint32_t len;
    const char* seq;
    Run run(uint32_t i) const { return Run{rd64(runs + 12 * i), rd32(runs + 12 * i + 8)}; }
};

inline FragView view_at(const uint8_t* buf, uint64_t off) {
    FragView f;
    f.open = buf[off];
    f.nruns = rd32(buf + off + 1);
    f.runs = buf + off + 5;
    f.len = rd32(buf + off + 5 + 12 * uint64_t(f.nruns));
    f.seq = (const char*)(buf + off + 9 + 12 * uint64_t(f.nruns));
    return f;
}

// 2-bit code of an ACGT byte, -1 for any other byte
struct CodeTable {
    int8_t v[256];
    CodeTable() {
        for (int i = 0; i < 256; ++i) v[i] = -1;
        v[int('A')] = 0;
        v[int('C')] = 1;
        v[int('G')] = 2;
        v[int('T')] = 3;
    }
};
const CodeTable g_code;
inline int code_of(char c) { return g_code.v[uint8_t(c)]; }

inline char comp(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        default: return 'A';
    }
}

// canonical k-mer of seq[0, k): min of the 2-bit forward and reverse-complement words
inline u128 canonical_kmer(const char* s, uint32_t k, bool& fwd_is_canonical) {
    u128 f = 0, r = 0;
    int bad = 0;
    for (uint32_t i = 0; i < k; ++i) {
        const int c = code_of(s[i]), d = code_of(s[k - 1 - i]);
        bad |= c | d;
        f = (f << 2) | u128(c & 3);
        r = (r << 2) | u128(3 - (d & 3));
    }
    if (bad < 0) throw Error("stitch: non-ACGT base in a fragment sequence");
    fwd_is_canonical = f <= r;
    return fwd_is_canonical ? f : r;
}

inline uint64_t mix(u128 x) {
    uint64_t z = uint64_t(x) ^ (uint64_t(x >> 64) * 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct End {
    u128 key;
    uint64_t id;  // frag << 2 | side << 1 | fwd_is_canonical   (side 0 left, 1 right)
};

// fn(begin, end, thread) over `threads` contiguous ranges of [0, n); ranges
// hold at least `grain` items
template <class F>
void parallel_for(uint64_t n, uint32_t threads, F&& fn, uint64_t grain = 4096) {
    threads = std::max<uint32_t>(1, std::min<uint64_t>(threads, std::max<uint64_t>(n / grain, 1)));
    if (threads == 1) {
        fn(uint64_t(0), n, 0u);
        return;
    }
    std::vector<std::thread> th;
    std::string err;
    std::atomic<bool> failed{false};
    for (uint32_t t = 0; t < threads; ++t)
        th.emplace_back([&, t] {
            try {
                fn(n * t / threads, n * (t + 1) / threads, t);
            } catch (std::exception& e) {
                if (!failed.exchange(true)) err = e.what();
            }
        });
    for (auto& x : th) x.join();
    if (failed) throw Error(err);
}

inline bool is_base(uint8_t c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

// Size of the record at p, 0 when it does not fit in [p, n).
inline uint64_t record_size(const uint8_t* buf, uint64_t n, uint64_t p) {
    if (n - p < 9) return 0;
    const uint64_t nr = rd32(buf + p + 1);
    if (n - p < 9 + 12 * nr) return 0;
    const uint64_t len = rd32(buf + p + 5 + 12 * nr);
    if (n - p - 9 - 12 * nr < len) return 0;
    return 9 + 12 * nr + len;
}

// Whether p looks like a record start inside a stream: an open-flag byte
// (0-3) after the last base of a sequence, opening a chain of a few records
// that fit, each with a flag byte and ACGT at both ends of its sequence.
inline bool plausible_start(const uint8_t* buf, uint64_t n, uint64_t p) {
    if (p == 0 || buf[p] > 3 || !is_base(buf[p - 1])) return false;
    for (int r = 0; r < 4 && p < n; ++r) {
        if (buf[p] > 3) return false;
        const uint64_t s = record_size(buf, n, p);
        if (!s) return false;
        const uint64_t seq = p + 9 + 12 * uint64_t(rd32(buf + p + 1));  // s = seq - p + len
        if (seq < p + s && (!is_base(buf[seq]) || !is_base(buf[p + s - 1]))) return false;
        p += s;
    }
    return true;
}

// A tig under construction: bases and the colour-set run of every k-mer.
struct Tig {
    std::string seq;
    std::vector<Run> runs;
    void push(uint64_t cid, uint32_t n) {
        if (!n) return;
        if (!runs.empty() && runs.back().cid == cid)
            runs.back().n += n;
        else
            runs.push_back(Run{cid, n});
    }
};

struct Stitcher {
    const uint8_t* buf;
    uint64_t n_bytes;
    uint32_t k;
    uint32_t threads;
    std::vector<uint64_t> off;      // record start of every fragment
    std::vector<uint64_t> partner;  // (frag << 1 | side) -> partner's (frag << 1 | side), NONE
    std::vector<uint8_t> used;
    cdbg_stitch_stats st{};

    FragView frag(uint64_t i) const { return view_at(buf, off[i]); }

    void index() {
        off = fragment_offsets(buf, n_bytes, threads);
        st.num_fragments = off.size();
        // every record's runs cover its k-mers
        parallel_for(off.size(), threads, [&](uint64_t b, uint64_t e, uint32_t) {
            for (uint64_t i = b; i < e; ++i) {
                const FragView f = frag(i);
                if (f.len < k) throw Error("stitch: a fragment is shorter than k");
                uint64_t s = 0;
                for (uint32_t r = 0; r < f.nruns; ++r) s += f.run(r).n;
                if (s != uint64_t(f.len) - k + 1) throw Error("stitch: colour runs do not cover a fragment's k-mers");
            }
        });
    }

    void match() {
        const uint64_t nf = off.size();
        partner.resize(2 * nf);
        // open sides, in fragment order
        std::vector<std::vector<End>> part(threads);
        parallel_for(nf, threads, [&](uint64_t b, uint64_t e, uint32_t t) {
            std::fill(partner.begin() + 2 * b, partner.begin() + 2 * e, NONE);
            auto& v = part[t];
            v.reserve(2 * (e - b));
            for (uint64_t i = b; i < e; ++i) {
                const FragView f = frag(i);
                for (uint32_t side = 0; side < 2; ++side) {
                    if (!(f.open & (side ? FRAG_OPEN_RIGHT : FRAG_OPEN_LEFT))) continue;
                    bool fwd;
                    const u128 key = canonical_kmer(f.seq + (side ? f.len - k : 0), k, fwd);
                    v.push_back(End{key, (i << 2) | (uint64_t(side) << 1) | uint64_t(fwd)});
                }
            }
        });
        // hash buckets of ends, each sorted by key
        constexpr uint32_t HB = 12, NB = 1u << HB;
        std::vector<std::vector<uint64_t>> hcnt(part.size(), std::vector<uint64_t>(NB + 1, 0));
        parallel_for(part.size(), threads, [&](uint64_t b, uint64_t e, uint32_t) {
            for (uint64_t t = b; t < e; ++t)
                for (const End& x : part[t]) ++hcnt[t][mix(x.key) >> (64 - HB)];
        }, 1);
        std::vector<uint64_t> boff(NB + 1, 0);
        std::vector<std::vector<uint64_t>> at(part.size(), std::vector<uint64_t>(NB, 0));
        uint64_t run = 0;
        for (uint32_t h = 0; h < NB; ++h) {
            boff[h] = run;
            for (size_t t = 0; t < part.size(); ++t) {
                at[t][h] = run;
                run += hcnt[t][h];
            }
        }
        boff[NB] = run;
        st.open_ends = run;
        std::unique_ptr<End[]> ends(new End[run]);  // every entry is written below
        parallel_for(part.size(), threads, [&](uint64_t b, uint64_t e, uint32_t) {
            for (uint64_t t = b; t < e; ++t) {
                for (const End& x : part[t]) ends[at[t][mix(x.key) >> (64 - HB)]++] = x;
                std::vector<End>().swap(part[t]);
            }
        }, 1);
        std::atomic<uint64_t> joins{0}, unmatched{0};
        std::atomic<uint32_t> next{0};
        parallel_for(threads, threads, [&](uint64_t, uint64_t, uint32_t) {
            for (;;) {
                const uint32_t h = next.fetch_add(1);
                if (h >= NB) break;
                End* a = ends.get() + boff[h];
                End* z = ends.get() + boff[h + 1];
                std::sort(a, z, [](const End& x, const End& y) { return x.key < y.key || (x.key == y.key && x.id < y.id); });
                uint64_t nj = 0, nu = 0;  // this bucket's joins and unmatched ends (one atomic add each)
                for (End* p = a; p < z;) {
                    End* q = p + 1;
                    while (q < z && q->key == p->key) ++q;
                    if (q - p > 2) throw Error("stitch: more than two open fragment ends share a boundary k-mer");
                    bool paired = false;
                    if (q - p == 2) {
                        const uint64_t sa = (p[0].id >> 1) & 1, sb = (p[1].id >> 1) & 1;
                        const bool fa = p[0].id & 1, fb = p[1].id & 1;
                        const bool ok = sa != sb ? fa == fb : fa != fb;
                        if (ok && (p[0].id >> 1) != (p[1].id >> 1)) {
                            partner[p[0].id >> 1] = p[1].id >> 1;
                            partner[p[1].id >> 1] = p[0].id >> 1;
                            paired = true;
                            ++nj;
                        }
                    }
                    if (!paired) nu += uint64_t(q - p);
                    p = q;
                }
                joins.fetch_add(nj);
                unmatched.fetch_add(nu);
            }
        }, 1);
        st.joins = joins;
        st.unmatched_ends = unmatched;
    }

    // fragment i entered at `side` (0 left, 1 right): appended to t in the
    // walk's orientation, reversed (reverse complement) when entered at its
    // right side; `join` drops the shared first k-mer of the fragment
    void append(Tig& t, uint64_t i, uint32_t side, bool join) const {
        const FragView f = frag(i);
        const bool rev = side == 1;
        auto run_at = [&](uint32_t j) { return f.run(rev ? f.nruns - 1 - j : j); };  // in walk order
        uint32_t skip = 0;  // k-mers of the fragment already in t
        if (join) {
            // the shared k-mer is t's last and the fragment's first: keep one
            // copy, with the owning side's colour (foreign when neither owns it)
            const uint64_t a = t.runs.back().cid, b = run_at(0).cid;
            const uint64_t real = a != COLOR_FOREIGN ? a : b;
            if (real != a) {
                if (--t.runs.back().n == 0) t.runs.pop_back();
                t.push(real, 1);
            }
            skip = 1;
        }
        for (uint32_t j = 0; j < f.nruns; ++j) {
            const Run x = run_at(j);
            const uint32_t take = x.n > skip ? x.n - skip : 0;
            skip = x.n > skip ? 0 : skip - x.n;
            t.push(x.cid, take);
        }
        const uint64_t from = join ? k : 0;
        if (!rev) {
            t.seq.append(f.seq + from, f.len - from);
        } else {
            for (uint64_t j = from; j < f.len; ++j) t.seq.push_back(comp(f.seq[f.len - 1 - j]));
        }
    }

    // other end of the chain that enters fragment i at `side`
    uint64_t far_end(uint64_t i, uint32_t side) const {
        uint64_t cur = (i << 1) | (1 - side);
        for (uint64_t steps = 0;; ++steps) {
            const uint64_t p = partner[cur];
            if (p == NONE) return cur;
            if (steps > off.size()) throw Error("stitch: a chain does not end");
            cur = p ^ 1;  // leave the partner fragment by its other side
        }
    }

    // walks the chain entering fragment `start` at `side`; returns the tig and
    // whether it closed on itself
    bool walk(uint64_t start, uint32_t side, Tig& t, std::vector<uint8_t>& mark) const {
        t.seq.clear();
        t.runs.clear();
        append(t, start, side, false);
        mark[start] = 1;
        uint64_t cur = (start << 1) | (1 - side);
        for (;;) {
            const uint64_t p = partner[cur];
            if (p == NONE) return false;
            const uint64_t nxt = p >> 1;
            const uint32_t nside = uint32_t(p & 1);
            if (mark[nxt]) {
                if (nxt != start || nside != side) throw Error("stitch: a chain re-enters a fragment");
                // closed: the last k-mer is the first one again; keep it once,
                // with the owner's colour
                const uint64_t last = t.runs.back().cid;
                if (--t.runs.back().n == 0) t.runs.pop_back();
                if (t.runs.front().cid == COLOR_FOREIGN && last != COLOR_FOREIGN) {
                    if (--t.runs.front().n == 0) t.runs.erase(t.runs.begin());
                    if (!t.runs.empty() && t.runs.front().cid == last)
                        ++t.runs.front().n;
                    else
                        t.runs.insert(t.runs.begin(), Run{last, 1});
                }
                t.seq.pop_back();
                return true;
            }
            append(t, nxt, nside, true);
            mark[nxt] = 1;
            cur = (nxt << 1) | (1 - nside);
        }
    }
};

// Cuts a linear tig into monochromatic unitigs (ext_split_monochromatic).
void cut_linear(const Tig& t, uint32_t k, UnitigList& out, cdbg_stitch_stats& st) {
    size_t lo = 0, hi = t.runs.size();
    uint64_t base = 0;
    if (lo < hi && t.runs[lo].cid == COLOR_FOREIGN) {
        st.foreign_dropped += t.runs[lo].n;
        base += t.runs[lo].n;
        ++lo;
    }
    if (lo < hi && t.runs[hi - 1].cid == COLOR_FOREIGN) {
        st.foreign_dropped += t.runs[hi - 1].n;
        --hi;
    }
    for (size_t i = lo; i < hi; ++i) {
        const Run& r = t.runs[i];
        if (r.cid == COLOR_FOREIGN) throw Error("stitch: a foreign k-mer inside a tig (its owner never joined)");
        UnitigList::Item it{r.cid, out.bases.size(), uint64_t(r.n) + k - 1};
        out.bases.append(t.seq, base, it.len);
        out.items.push_back(it);
        st.unitig_kmers += r.n;
        base += r.n;
    }
}

// Cuts a cycle of n k-mers whose sequence holds n + k - 1 bases (the last
// k - 1 repeat the first): one circular unitig when it has one colour set,
// else rotated to start at a colour change and cut like a linear tig.
void cut_cycle(Tig& t, uint32_t k, UnitigList& out, cdbg_stitch_stats& st) {
    uint64_t n = 0;
    for (const Run& r : t.runs) {
        if (r.cid == COLOR_FOREIGN) throw Error("stitch: a foreign k-mer in a closed tig");
        n += r.n;
    }
    if (t.runs.size() == 1) {
        UnitigList::Item it{t.runs[0].cid, out.bases.size(), t.seq.size()};
        out.bases.append(t.seq);
        out.items.push_back(it);
        st.unitig_kmers += n;
        return;
    }
    if (t.runs.front().cid == t.runs.back().cid) {
        // the seam lies inside a run: start the cycle at that run's first
        // k-mer (a colour change), so every piece is a whole run
        const uint64_t r = n - t.runs.back().n;
        std::string seq(n + k - 1, 'A');
        for (uint64_t i = 0; i < n + k - 1; ++i) seq[i] = t.seq[(r + i) % n];
        t.seq.swap(seq);
        t.runs.front().n += t.runs.back().n;
        t.runs.pop_back();
    }
    cut_linear(t, k, out, st);
}

}  // namespace

std::vector<uint64_t> fragment_offsets(const uint8_t* buf, uint64_t n, uint32_t threads) {
    // thread t guesses the first record start at or after n t / T and follows
    // the chain of records from there past the end of its range (ranges of
    // at least 64 KiB)
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads, n >> 16));
    struct Part {
        uint64_t lo, hi, guess = ~0ull, exit = 0;
        bool failed = false;
        std::vector<uint64_t> starts;
    };
    std::vector<Part> parts(T);
    for (uint32_t t = 0; t < T; ++t) {
        parts[t].lo = n * t / T;
        parts[t].hi = n * (t + 1) / T;
    }
    parallel_for(T, T, [&](uint64_t b, uint64_t e, uint32_t) {
        for (uint64_t t = b; t < e; ++t) {
            Part& P = parts[t];
            uint64_t p = P.lo;
            if (t > 0)
                while (p < P.hi && !plausible_start(buf, n, p)) ++p;
            if (p >= P.hi && t > 0) continue;  // no start guessed in this range
            P.guess = p;
            P.starts.reserve((P.hi - P.lo) / 48 + 16);  // records average well above 48 bytes
            while (p < P.hi) {
                const uint64_t s = record_size(buf, n, p);
                if (!s) {
                    P.failed = true;
                    break;
                }
                P.starts.push_back(p);
                p += s;
            }
            P.exit = p;
        }
    }, 1);
    // the true chain enters range t at e: the guess is kept when it is e, the
    // range is parsed again from e otherwise (or when its chain broke)
    std::vector<uint64_t> off;
    uint64_t total = 0;
    for (auto& P : parts) total += P.starts.size();
    off.reserve(total + 16);
    uint64_t e = 0;
    for (uint32_t t = 0; t < T; ++t) {
        Part& P = parts[t];
        if (e >= P.hi) continue;  // a record spans the whole range
        if (P.guess == e && !P.failed) {
            off.insert(off.end(), P.starts.begin(), P.starts.end());
            e = P.exit;
        } else {
            while (e < P.hi) {
                const uint64_t s = record_size(buf, n, e);
                if (!s) throw Error("stitch: truncated fragment record");
                off.push_back(e);
                e += s;
            }
        }
        std::vector<uint64_t>().swap(P.starts);
    }
    return off;
}

uint64_t file_size_of(const char* path) {
    struct stat sb;
    if (::stat(path, &sb) != 0) throw Error(std::string("cannot stat ") + path);
    return uint64_t(sb.st_size);
}

void read_file_into(const char* path, uint8_t* dst, uint64_t n, uint32_t threads) {
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) throw Error(std::string("cannot open ") + path);
    // ranges of at least 64 MiB per thread
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads ? threads : 1, n >> 26));
    std::atomic<bool> short_read{false};
    parallel_for(T, T, [&](uint64_t b, uint64_t e, uint32_t) {
        for (uint64_t t = b; t < e; ++t) {
            uint64_t p = n * t / T;
            const uint64_t end = n * (t + 1) / T;
            while (p < end) {
                const ssize_t got = ::pread(fd, dst + p, size_t(std::min<uint64_t>(end - p, 1ull << 30)), off_t(p));
                if (got <= 0) {
                    short_read = true;
                    break;
                }
                p += uint64_t(got);
            }
        }
    }, 1);
    ::close(fd);
    if (short_read) throw Error(std::string("short read of ") + path);
}

void write_file_parallel(const char* path, const void* src, uint64_t n, uint32_t threads) {
    const int fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) throw Error(std::string("cannot open ") + path);
    // ranges of at least 64 MiB per thread, written at their offsets
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads ? threads : 1, n >> 26));
    std::atomic<bool> failed{false};
    parallel_for(T, T, [&](uint64_t b, uint64_t e, uint32_t) {
        for (uint64_t t = b; t < e; ++t) {
            uint64_t p = n * t / T;
            const uint64_t end = n * (t + 1) / T;
            while (p < end && !failed) {
                const ssize_t w = ::pwrite(fd, (const uint8_t*)src + p, size_t(std::min<uint64_t>(end - p, 1ull << 30)),
                                           off_t(p));
                if (w <= 0) {
                    failed = true;
                    break;
                }
                p += uint64_t(w);
            }
        }
    }, 1);
    const bool closed = ::close(fd) == 0;
    if (failed || !closed) throw Error(std::string("short write to ") + path);
}

std::unique_ptr<uint8_t[]> read_file(const char* path, uint32_t threads, uint64_t& n) {
    n = file_size_of(path);
    std::unique_ptr<uint8_t[]> buf(new uint8_t[n + 1]);
    read_file_into(path, buf.get(), n, threads);
    return buf;
}

void parallel_copy(void* dst, const void* src, uint64_t n, uint32_t threads) {
    parallel_for(n, threads, [&](uint64_t b, uint64_t e, uint32_t) {
        std::memcpy((uint8_t*)dst + b, (const uint8_t*)src + b, e - b);
    }, uint64_t(1) << 24);
}

void stitch_cycles(const uint8_t* buf, uint64_t n, const std::vector<uint64_t>& rec, uint32_t k,
                   const std::vector<uint32_t>& frags, const std::vector<uint32_t>& partner, UnitigList& out,
                   cdbg_stitch_stats& st) {
    const uint64_t nc = frags.size();
    if (partner.size() != 2 * nc || rec.size() != nc)
        throw Error("stitch: cycle partners or records do not match the cycle fragments");
    Stitcher S{buf, n, k, 1, {}, {}, {}, {}};
    S.off.resize(nc);
    S.partner.resize(2 * nc);
    auto local = [&](uint32_t f) -> uint64_t {
        const auto it = std::lower_bound(frags.begin(), frags.end(), f);
        if (it == frags.end() || *it != f) throw Error("stitch: an unvisited fragment is not on a cycle");
        return uint64_t(it - frags.begin());
    };
    for (uint64_t l = 0; l < nc; ++l) {
        if ((l && frags[l] <= frags[l - 1]) || rec[l] >= n || record_size(buf, n, rec[l]) == 0)
            throw Error("stitch: cycle fragments out of order or their records out of range");
        S.off[l] = rec[l];
        for (uint32_t s = 0; s < 2; ++s) {
            const uint32_t p = partner[2 * l + s];
            S.partner[2 * l + s] = p == ~0u ? NONE : (local(p >> 1) << 1) | (p & 1u);
        }
    }
    S.used.assign(nc, 0);
    Tig tig;
    for (uint64_t l = 0; l < nc; ++l) {
        if (S.used[l]) continue;
        if (!S.walk(l, 0, tig, S.used)) throw Error("stitch: an unvisited fragment is not on a cycle");
        ++st.cycles;
        cut_cycle(tig, k, out, st);
    }
    std::stable_sort(out.items.begin(), out.items.end(),
                     [](const UnitigList::Item& a, const UnitigList::Item& b) { return a.cid < b.cid; });
}

uint64_t Unitigs::size() const {
    uint64_t total = 0;
    for (auto& l : lists) total += l.size();
    return total;
}

void Unitigs::build_order() {
    if (single != -2) return;
    int nonempty = 0, which = -1;
    for (uint32_t l = 0; l < lists.size(); ++l)
        if (lists[l].size()) {
            ++nonempty;
            which = int(l);
        }
    if (nonempty <= 1) {
        single = which;
        return;
    }
    single = -1;
    order.reserve(size());
    // stable merge by cid: the smallest head, ties to the earlier list
    std::vector<uint64_t> pos(lists.size(), 0);
    using H = std::pair<uint64_t, uint32_t>;  // (cid, list)
    std::vector<H> heap;
    auto cmp = [](const H& a, const H& b) { return a.first > b.first || (a.first == b.first && a.second > b.second); };
    for (uint32_t l = 0; l < lists.size(); ++l)
        if (lists[l].size()) heap.push_back(H{lists[l].item(0).cid, l});
    std::make_heap(heap.begin(), heap.end(), cmp);
    while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        const uint32_t l = heap.back().second;
        heap.pop_back();
        const UnitigList& L = lists[l];
        const uint64_t cid = L.item(pos[l]).cid;
        // take this list's whole run of equal cids
        while (pos[l] < L.size() && L.item(pos[l]).cid == cid) order.push_back({l, pos[l]++});
        if (pos[l] < L.size()) {
            heap.push_back(H{L.item(pos[l]).cid, l});
            std::push_heap(heap.begin(), heap.end(), cmp);
        }
    }
}

Unitigs* stitch_fragments(const uint8_t* buf, uint64_t n, uint32_t k, uint32_t threads) {
    if (k < 2 || k > 63) throw Error("k must satisfy 2 <= k <= 63");
    if (threads == 0) threads = 1;
    Stitcher S{buf, n, k, threads, {}, {}, {}, {}};
    S.index();
    S.match();
    const uint64_t nf = S.off.size();
    // chains: started from the end with the smaller (fragment, side)
    std::vector<std::vector<uint64_t>> part(threads);
    parallel_for(nf, threads, [&](uint64_t b, uint64_t e, uint32_t t) {
        for (uint64_t i = b; i < e; ++i)
            for (uint32_t side = 0; side < 2; ++side) {
                if (S.partner[(i << 1) | side] != NONE) continue;
                const uint64_t me = (i << 1) | side, other = S.far_end(i, side);
                if (me <= other) part[t].push_back(me);
                break;  // a fragment starts at most one chain: its left side first
            }
    });
    std::vector<uint64_t> starts;
    for (auto& p : part) {
        starts.insert(starts.end(), p.begin(), p.end());
        std::vector<uint64_t>().swap(p);
    }
    auto* U = new Unitigs();
    U->k = k;
    U->threads = threads;
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads, std::max<uint64_t>(starts.size() / 4096, 1)));
    U->lists.resize(T + 1);  // the last list holds the cycles
    S.used.assign(nf, 0);
    std::vector<cdbg_stitch_stats> tst(T + 1, cdbg_stitch_stats{});
    try {
        parallel_for(starts.size(), T, [&](uint64_t b, uint64_t e, uint32_t t) {
            Tig tig;
            auto& out = U->lists[t];
            for (uint64_t j = b; j < e; ++j) {
                if (S.walk(starts[j] >> 1, uint32_t(starts[j] & 1), tig, S.used))
                    throw Error("stitch: a chain from a free end closed on itself");
                cut_linear(tig, k, out, tst[t]);
            }
        });
        // what is left are cycles of joined fragments
        Tig tig;
        for (uint64_t i = 0; i < nf; ++i) {
            if (S.used[i]) continue;
            if (!S.walk(i, 0, tig, S.used)) throw Error("stitch: an unvisited fragment is not on a cycle");
            ++tst[T].cycles;
            cut_cycle(tig, k, U->lists[T], tst[T]);
        }
        for (auto& l : U->lists)
            std::stable_sort(l.items.begin(), l.items.end(),
                             [](const UnitigList::Item& a, const UnitigList::Item& b) { return a.cid < b.cid; });
    } catch (...) {
        delete U;
        throw;
    }
    U->stats = S.st;
    for (auto& s : tst) {
        U->stats.cycles += s.cycles;
        U->stats.unitig_kmers += s.unitig_kmers;
        U->stats.foreign_dropped += s.foreign_dropped;
    }
    for (auto& l : U->lists) U->stats.num_unitigs += l.size();
    return U;
}

}  // namespace cdbg_amd
