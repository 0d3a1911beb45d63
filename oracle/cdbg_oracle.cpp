// ORACLE — test infrastructure only.
//
// A plain CPU restatement of cdbg-builder's phase 1 (bucket-write) and phase 2
// (bucket-process), used as the checker for the HIP kernels and as the
// "port" CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library; the product never links it.
//
// Every function cites the reference file:line it follows.  The restatement
// keeps the reference's semantics exactly, including the ones that make its
// output order-dependent:
//   * k-mer nodes are kept in first-insertion order (ankerl::unordered_dense
//     iterates in insertion order), which fixes the walk's seed order and the
//     local colour-class order            (bucket_walker.hpp:160,387,563-566);
//   * buckets are processed 0..B-1 and merged into the global dictionary in
//     that order, i.e. the reference at -t 1 (bucket_walker.hpp:645,713-757).
//
// Parity status per function (details in the test files):
//   ntHash / window minimum / reverse complement / canonical / varint / 2-bit
//     packing / compute_best_m   : pinned against the reference's own headers
//                                   compiled into oracle/_ref, plus brute force.
//   super-k-mer split + flags     : pinned by brute force from the definitions
//                                   and by invariants (tests/test_oracle_phase1.py).
//   per-bucket load/resolve/walk, : pinned by the ground-truth ccdBG definition of
//   colour-set dictionary,          test_data/verify.py:103-220 through a test-side
//   fragment records                stitch + monochromatic split on small inputs, and
//                                   by the salmonella_10 known answers (86,630
//                                   unitigs, 171 colour sets; algorithm.md:36-39)
//                                   (tests/test_oracle_phase2.py).
//   hybrid colour-set code        : decoded by a reader written from the reference's
//                                   own decoder (stream.hpp:196-237); the Elias-delta
//                                   bit order inside it is parity unpinned (jermp/bits
//                                   is absent; the layout follows its write_delta).
//   fragment decomposition order  : (which seed starts which fragment) follows
//                                   bucket_walker.hpp:563-566 at -t 1; pinned only
//                                   through the final unitig set above.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;

// ---------------------------------------------------------------------------
// Primitives (kmer.hpp, minimizer.hpp, super_kmer.hpp)
// ---------------------------------------------------------------------------

// ntHash per-base seeds: the published ntHash table (minimizer.hpp:21-26).
static const u64 NT_SEED[4] = {0x3c8bfbb395c60474ULL, 0x3193c18562a02b4cULL,
                               0x20323ed082572324ULL, 0x295549f54be24456ULL};

static inline u64 rol(u64 x, unsigned r) {
    r &= 63u;
    return r ? (x << r) | (x >> (64 - r)) : x;
}
static inline u64 ror(u64 x, unsigned r) {
    r &= 63u;
    return r ? (x >> r) | (x << (64 - r)) : x;
}

// A/C/G/T in either case -> 0..3, anything else -> 0xff (kmer.hpp:35-52).
static inline u8 code_of(u8 c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return 0xff;
    }
}

// m chosen from k when -m is absent (minimizer.hpp:65-74).
static u32 best_m(u32 k) {
    if (k <= 13) return (k / 2 > k - 4) ? (k / 2) : (k > 4 ? (k - 4) : 1);
    if (k <= 15) return 9;
    if (k <= 21) return 10;
    if (k <= 30) return 11;
    if (k <= 37) return 12;
    if (k <= 42) return 13;
    if (k <= 64) return 14;
    return (k + 2) / 4;
}

// Canonical ntHash of one m-mer, from scratch: fwd = xor_i rol(S[b_i], m-1-i),
// rc = xor_i rol(S[b_i ^ 3], i), h = min(fwd, rc)  (minimizer.hpp:40-55,62).
static inline u64 mmer_hash(const u8* b, u32 m) {
    u64 f = 0, r = 0;
    for (u32 i = 0; i < m; ++i) {
        f = rol(f, 1) ^ NT_SEED[b[i]];
        r = rol(r, 1) ^ NT_SEED[b[m - 1 - i] ^ 3u];
    }
    return f <= r ? f : r;
}

// k-mer words: first base in the highest used bits (kmer.hpp:11-18).
template <class T>
static inline T kmask(u32 k) {
    if (2 * k >= sizeof(T) * 8) return ~T(0);
    return (T(1) << (2 * k)) - 1;
}
// Reverse complement by complementing and reversing the 2-bit groups; written
// as a plain loop here (the reference uses a SWAR form, kmer.hpp:79-106).
template <class T>
static inline T revcomp(T x, u32 k) {
    T r = 0;
    for (u32 i = 0; i < k; ++i) {
        r = (r << 2) | ((x & 3) ^ 3);
        x >>= 2;
    }
    return r;
}
template <class T>
static inline T canon(T x, u32 k) {
    T r = revcomp<T>(x, k);
    return x <= r ? x : r;
}
template <class T>
static inline T shift_append(T x, u32 nt, u32 k) {
    return ((x << 2) | T(nt & 3)) & kmask<T>(k);
}

// LEB128, 7-bit groups, least significant first (super_kmer.hpp:61-82).
static inline void put_varint(u64 v, std::vector<u8>& out) {
    while (v >= 0x80) {
        out.push_back(u8((v & 0x7f) | 0x80));
        v >>= 7;
    }
    out.push_back(u8(v));
}
static inline u64 get_varint(const u8* buf, size_t len, size_t& pos) {
    u64 v = 0;
    unsigned sh = 0;
    while (pos < len) {
        u8 b = buf[pos++];
        v |= u64(b & 0x7f) << sh;
        if (!(b & 0x80)) return v;
        sh += 7;
        if (sh >= 64) break;
    }
    return v;
}

// Super-k-mer flag bits (super_kmer.hpp:85-97).
enum : u8 { SK_BEGIN = 1, SK_END = 2, SK_OWNS_FIRST = 4, SK_OWNS_LAST = 8 };
// Per-k-mer boundary bits and fragment open bits (bucket_walker.hpp:80-81,132-133).
enum : u8 { BOUND_LEFT = 1, BOUND_RIGHT = 2 };
enum : u8 { OPEN_LEFT = 1, OPEN_RIGHT = 2 };
static const u64 FOREIGN = ~u64(0);  // COLOR_RUN_FOREIGN (bucket_walker.hpp:96)

struct Super {
    u64 start;  // absolute base offset of the first base
    u32 len;    // bases (>= k)
    u8 flags;   // SK_*
    u32 bucket;
    u32 color;
};

// ---------------------------------------------------------------------------
// Phase 1: ACGT runs -> minimizers -> super-k-mers
// ---------------------------------------------------------------------------

// One ACGT run of length L >= k (codes in b[0..L)), absolute offset `abs0`.
// Follows emit_super_kmers (bucket_ingester.hpp:36-127): minimizer of
// (k-1)-mer j is the minimum canonical m-mer hash over the W = k-m m-mers it
// contains (minimizer.hpp:94-157, bucket_ingester.hpp:52); a super-k-mer is a
// maximal stretch of consecutive (k-1)-mers with equal minimizer VALUE
// (:106-109) and spans k-mers [a-1 .. b], clamped to the run (:110,121).
// Bucket map of the build being run: null for (h >> 1) % B
// (bucket_ingester.hpp:67), else B - 1 ascending cuts of the low 32 bits of
// h >> 1 (a bucket map of this repository's library, for skewed-bucket
// workloads; the reference has none).
static const u64* g_cuts = nullptr;
static u32 bucket_of(u64 h, u32 B) {
    if (!g_cuts) return u32((h >> 1) % B);
    const u64 x = (h >> 1) & 0xFFFFFFFFull;
    return u32(std::upper_bound(g_cuts, g_cuts + (B - 1), x) - g_cuts);
}

static void supers_of_run(const u8* b, u32 L, u32 k, u32 m, u32 B, u32 color, u64 abs0,
                          std::vector<Super>& out) {
    const u32 K = L - k + 1;  // k-mers 0..K-1, (k-1)-mers 0..K
    const u32 W = k - m;
    const u32 nm = L - m + 1;
    std::vector<u64> h(nm);
    for (u32 p = 0; p < nm; ++p) h[p] = mmer_hash(b + p, m);
    std::vector<u64> mn(K + 1);
    for (u32 j = 0; j <= K; ++j) {
        u64 best = ~u64(0);
        for (u32 p = j; p < j + W; ++p) best = std::min(best, h[p]);
        mn[j] = best;
    }
    auto emit = [&](u32 a, u32 e, bool begin, bool end, bool own_first, bool own_last) {
        u32 first_kmer = (a == 0) ? 0 : a - 1;
        u32 last_kmer = e;  // e <= K-1 here
        Super s;
        s.start = abs0 + first_kmer;
        s.len = (last_kmer - first_kmer) + k;
        s.flags = u8((begin ? SK_BEGIN : 0) | (end ? SK_END : 0) | (own_first ? SK_OWNS_FIRST : 0) |
                     (own_last ? SK_OWNS_LAST : 0));
        s.bucket = bucket_of(mn[a], B);  // bucket_ingester.hpp:67
        s.color = color;
        out.push_back(s);
    };
    u32 run_a = 0;
    bool first = true, own_first = true;
    for (u32 j = 1; j <= K; ++j) {
        if (mn[j] != mn[j - 1]) {
            bool own_last = mn[j - 1] < mn[j];
            emit(run_a, j - 1, first, false, own_first, own_last);
            first = false;
            own_first = mn[j] < mn[j - 1];
            run_a = j;
        }
    }
    emit(run_a, K - 1, first, true, own_first, true);
}

// Scan one colour's sequence bytes: maximal runs of ACGT codes; runs shorter
// than k are dropped (bucket_ingester.hpp:152-172).
static void supers_of_seq(const u8* seq, u64 len, u64 abs_base, u32 k, u32 m, u32 B, u32 color,
                          std::vector<Super>& out) {
    std::vector<u8> codes;
    u64 pos = 0;
    while (pos < len) {
        while (pos < len && code_of(seq[pos]) == 0xff) ++pos;
        u64 s = pos;
        codes.clear();
        while (pos < len) {
            u8 c = code_of(seq[pos]);
            if (c == 0xff) break;
            codes.push_back(c);
            ++pos;
        }
        u64 L = pos - s;
        if (L >= k) supers_of_run(codes.data(), u32(L), k, m, B, color, abs_base + s, out);
    }
}

// ---------------------------------------------------------------------------
// Hybrid colour-set codec (hybrid_color_sets.hpp:41-96) over an LSB-first bit
// stream with jermp/bits-style gamma/delta codes.
// ---------------------------------------------------------------------------
struct BitStream {
    std::vector<u64> words;
    u64 nbits = 0;
    void append(u64 v, u32 len) {  // low `len` bits of v, bit 0 first
        if (len == 0) return;
        if (len < 64) v &= (u64(1) << len) - 1;
        u64 w = nbits >> 6;
        u32 o = u32(nbits & 63);
        if (w >= words.size()) words.push_back(0);
        words[w] |= v << o;
        if (o && o + len > 64) {
            words.push_back(v >> (64 - o));
        }
        nbits += len;
    }
};
static inline u32 msb64(u64 x) { return 63u - u32(__builtin_clzll(x)); }
static void write_gamma(BitStream& bs, u64 x) {
    u64 xx = x + 1;
    u32 b = msb64(xx);
    bs.append(0, b);
    u64 low = xx & ((u64(1) << b) - 1);
    bs.append((low << 1) | 1, b + 1);
}
static void write_delta(BitStream& bs, u64 x) {
    u64 xx = x + 1;
    u32 b = msb64(xx);
    write_gamma(bs, b);
    bs.append(xx & ((u64(1) << b) - 1), b);
}
static void encode_set(BitStream& bs, const u32* c, u64 n, u64 num_colors, u64 sparse, u64 dense) {
    write_delta(bs, n);
    if (n < sparse) {
        u32 prev = c[0];
        write_delta(bs, prev);
        for (u64 i = 1; i < n; ++i) {
            write_delta(bs, u64(c[i] - (prev + 1)));
            prev = c[i];
        }
    } else if (n < dense) {
        std::vector<u64> bm((num_colors + 63) / 64, 0);
        for (u64 i = 0; i < n; ++i) bm[c[i] >> 6] |= u64(1) << (c[i] & 63);
        for (u64 i = 0, left = num_colors; left; ++i) {
            u32 take = left < 64 ? u32(left) : 64u;
            bs.append(bm[i], take);
            left -= take;
        }
    } else {
        // delta gaps over the absent colours (hybrid_color_sets.hpp:67-94)
        bool first = true;
        u32 val = 0, prev = u32(-1);
        for (u64 i = 0; i < n; ++i) {
            u32 x = c[i];
            while (val < x) {
                if (first) {
                    write_delta(bs, val);
                    first = false;
                } else {
                    write_delta(bs, u64(val - (prev + 1)));
                }
                prev = val;
                ++val;
            }
            ++val;
        }
        while (val < num_colors) {
            if (first) {
                write_delta(bs, val);
                first = false;
            } else {
                write_delta(bs, u64(val - (prev + 1)));
            }
            prev = val;
            ++val;
        }
    }
}

// ---------------------------------------------------------------------------
// Phase 2 (bucket_walker.hpp)
// ---------------------------------------------------------------------------

// One bucket-stream record: sorted unique colours, flags, 2-bit codes.
struct Rec {
    std::vector<u32> colors;
    u8 flags;
    std::vector<u8> bases;
};

template <class T>
struct KHash {
    static inline u64 mix(u64 z) {
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    static inline u64 of(T x) {
        if constexpr (sizeof(T) > 8)
            return mix(u64(x)) ^ (mix(u64(x >> 64) + 0x9e3779b97f4a7c15ULL) << 1);
        else
            return mix(u64(x));
    }
};

// Insertion-ordered k-mer -> node index (stands in for the reference's
// ankerl map, whose iteration order is insertion order).
template <class T>
struct NodeIndex {
    std::vector<T> keys;
    std::vector<u32> slots;  // node index + 1, 0 = empty
    u64 mask = 0;
    NodeIndex() { rehash(1024); }
    void rehash(u64 cap) {
        slots.assign(cap, 0);
        mask = cap - 1;
        for (u32 i = 0; i < keys.size(); ++i) place(keys[i], i);
    }
    void place(T key, u32 idx) {
        u64 h = KHash<T>::of(key) & mask;
        while (slots[h]) h = (h + 1) & mask;
        slots[h] = idx + 1;
    }
    int64_t find(T key) const {
        u64 h = KHash<T>::of(key) & mask;
        while (slots[h]) {
            if (keys[slots[h] - 1] == key) return slots[h] - 1;
            h = (h + 1) & mask;
        }
        return -1;
    }
    u32 get_or_add(T key, bool& added) {
        int64_t f = find(key);
        if (f >= 0) {
            added = false;
            return u32(f);
        }
        added = true;
        if ((keys.size() + 1) * 2 > slots.size()) rehash(slots.size() * 2);
        keys.push_back(key);
        place(key, u32(keys.size() - 1));
        return u32(keys.size() - 1);
    }
    size_t size() const { return keys.size(); }
};

struct Fragment {
    std::string seq;
    std::vector<std::pair<u64, u32>> runs;  // (cid, num_kmers)
    u8 open = 0;
};

// Local colour-set dictionary with exact-content dedup.  The reference dedups
// by a 128-bit content hash without a byte compare
// (compact_color_set_dict.hpp:158-172); exact dedup gives the same ids unless
// that hash collides.
struct LocalDict {
    std::map<std::vector<u32>, u32> index;
    std::vector<const std::vector<u32>*> sets;
    u32 intern(const std::vector<u32>& v) {
        auto it = index.find(v);
        if (it != index.end()) return it->second;
        u32 id = u32(sets.size());
        auto ins = index.emplace(v, id);
        sets.push_back(&ins.first->first);
        return id;
    }
};

template <class T>
struct BucketResult {
    std::vector<Fragment> frags;
    LocalDict local;
    u64 n_primary = 0;
    u64 n_nodes = 0;
    u64 n_occ = 0;
};

// process_bucket (bucket_walker.hpp:338-570) over one bucket's record stream.
template <class T>
static void process_bucket(const std::vector<Rec>& recs, u32 k, BucketResult<T>& res) {
    struct Info {
        u8 flags = 0;
        bool primary = false;
        std::vector<u32> rsids;  // record ids of owned occurrences
        u32 cid = UINT32_MAX;
    };
    NodeIndex<T> idx;
    std::vector<Info> info;
    const T mask = kmask<T>(k);
    const u32 sh = 2 * (k - 1);
    // load_bucket (bucket_walker.hpp:166-255).  Records are interned by their
    // colour list (record_sets.intern :187); here the record index itself is
    // kept and the colour lists are unioned in resolve, which yields the same
    // sets and, through first-use order, the same local class order.
    for (u32 ri = 0; ri < recs.size(); ++ri) {
        const Rec& r = recs[ri];
        if (r.bases.size() < k) continue;
        res.n_occ += r.bases.size() - (k - 1);
        T fwd = 0, rc = 0;
        for (u32 i = 0; i + 1 < k; ++i) {
            u8 v = r.bases[i];
            fwd = ((fwd << 2) | T(v)) & mask;
            rc = (rc >> 2) | (T(v ^ 3) << sh);
        }
        const bool begin_incl = r.flags & SK_BEGIN, end_incl = r.flags & SK_END;
        const bool own_first = r.flags & SK_OWNS_FIRST, own_last = r.flags & SK_OWNS_LAST;
        const u32 nk = u32(r.bases.size() - (k - 1));
        for (u32 i = k - 1, ix = 0; i < r.bases.size(); ++i, ++ix) {
            u8 v = r.bases[i];
            fwd = ((fwd << 2) | T(v)) & mask;
            rc = (rc >> 2) | (T(v ^ 3) << sh);
            bool is_fwd = fwd <= rc;
            T can = is_fwd ? fwd : rc;
            u8 contrib = 0;
            if (!begin_incl && ix == 0) contrib |= is_fwd ? BOUND_LEFT : BOUND_RIGHT;
            if (!end_incl && ix == nk - 1) contrib |= is_fwd ? BOUND_RIGHT : BOUND_LEFT;
            bool added;
            u32 n = idx.get_or_add(can, added);
            if (added) info.emplace_back();
            Info& in = info[n];
            in.flags |= contrib;
            bool owned = !((ix == 0 && !own_first) || (ix == nk - 1 && !own_last));
            if (owned) {
                in.rsids.push_back(ri);
                in.primary = true;
            }
        }
    }
    res.n_nodes = info.size();
    // resolve (bucket_walker.hpp:355-429): primary nodes in insertion order;
    // colour set = union of the owning records' colours; first intern order
    // defines the local class ids.
    std::vector<u32> merged;
    for (u32 n = 0; n < info.size(); ++n) {
        Info& in = info[n];
        if (!in.primary) continue;
        ++res.n_primary;
        std::sort(in.rsids.begin(), in.rsids.end());
        in.rsids.erase(std::unique(in.rsids.begin(), in.rsids.end()), in.rsids.end());
        merged.clear();
        for (u32 ri : in.rsids) merged.insert(merged.end(), recs[ri].colors.begin(), recs[ri].colors.end());
        std::sort(merged.begin(), merged.end());
        merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
        in.cid = res.local.intern(merged);
        std::vector<u32>().swap(in.rsids);
    }
    // walk (bucket_walker.hpp:257-336,431-566)
    auto ext_mask = [&](T can) -> u8 {
        u8 out = 0;
        T rv = revcomp<T>(can, k);
        for (u32 nt = 0; nt < 4; ++nt) {
            if (idx.find(canon<T>(shift_append<T>(can, nt, k), k)) >= 0) out |= u8(1u << nt);
            if (idx.find(canon<T>(shift_append<T>(rv, nt, k), k)) >= 0) out |= u8(1u << (4 + nt));
        }
        return out;
    };
    auto fwd_nib = [](u8 m, bool rc) -> u8 { return u8((rc ? (m >> 4) : m) & 0xf); };
    auto back_nib = [](u8 m, bool rc) -> u8 { return u8((rc ? m : (m >> 4)) & 0xf); };
    auto step = [&](T can, bool rc, u32 nt, T& ncan, bool& nrc) {
        T cur = rc ? revcomp<T>(can, k) : can;
        T nf = shift_append<T>(cur, nt, k);
        T nr = revcomp<T>(nf, k);
        if (nf <= nr) {
            ncan = nf;
            nrc = false;
        } else {
            ncan = nr;
            nrc = true;
        }
    };
    auto walk_step = [&](T can, bool rc, bool forward, T& ncan, bool& nrc) -> bool {
        u8 m = ext_mask(can);
        u8 nib = forward ? fwd_nib(m, rc) : back_nib(m, rc);
        if (__builtin_popcount(nib) != 1) return false;
        u32 nt = u32(__builtin_ctz(nib));
        T c2;
        bool r2;
        if (forward) {
            step(can, rc, nt, c2, r2);
        } else {
            step(can, !rc, nt, c2, r2);
            r2 = !r2;
        }
        u8 nm = ext_mask(c2);
        u8 facing = forward ? back_nib(nm, r2) : fwd_nib(nm, r2);
        if (__builtin_popcount(facing) != 1) return false;
        ncan = c2;
        nrc = r2;
        return true;
    };
    auto is_break = [](u8 f) { return f == BOUND_LEFT || f == BOUND_RIGHT; };
    std::vector<u8> visited(info.size(), 0);
    std::vector<std::pair<u32, bool>> bw, fw;
    for (u32 seed = 0; seed < info.size(); ++seed) {
        if (visited[seed]) continue;
        visited[seed] = 1;
        const T scan = idx.keys[seed];
        const u8 sflags = info[seed].flags;
        bw.clear();
        fw.clear();
        bool open_l = false, open_r = false;
        if (sflags == BOUND_LEFT) {
            open_l = __builtin_popcount(back_nib(ext_mask(scan), false)) <= 1;
        } else {
            T can = scan;
            bool rc = false;
            for (;;) {
                T nc;
                bool nr;
                if (!walk_step(can, rc, false, nc, nr)) break;
                u32 n = u32(idx.find(nc));
                if (visited[n]) break;
                visited[n] = 1;
                bw.push_back({n, nr});
                can = nc;
                rc = nr;
                if (is_break(info[n].flags)) {
                    open_l = __builtin_popcount(back_nib(ext_mask(can), rc)) <= 1;
                    break;
                }
            }
        }
        if (sflags == BOUND_RIGHT) {
            open_r = __builtin_popcount(fwd_nib(ext_mask(scan), false)) <= 1;
        } else {
            T can = scan;
            bool rc = false;
            for (;;) {
                T nc;
                bool nr;
                if (!walk_step(can, rc, true, nc, nr)) break;
                u32 n = u32(idx.find(nc));
                if (visited[n]) break;
                visited[n] = 1;
                fw.push_back({n, nr});
                can = nc;
                rc = nr;
                if (is_break(info[n].flags)) {
                    open_r = __builtin_popcount(fwd_nib(ext_mask(can), rc)) <= 1;
                    break;
                }
            }
        }
        Fragment f;
        bool firstk = true;
        auto append = [&](u32 n, bool rc) {
            T km = rc ? revcomp<T>(idx.keys[n], k) : idx.keys[n];
            if (firstk) {
                f.seq.assign(k, 'N');
                T x = km;
                for (u32 i = 0; i < k; ++i) {
                    f.seq[k - 1 - i] = "ACGT"[u32(x & 3)];
                    x >>= 2;
                }
                firstk = false;
            } else {
                f.seq.push_back("ACGT"[u32(km & 3)]);
            }
            u64 cid = info[n].primary ? u64(info[n].cid) : FOREIGN;
            if (!f.runs.empty() && f.runs.back().first == cid)
                f.runs.back().second += 1;
            else
                f.runs.push_back({cid, 1});
        };
        for (auto it = bw.rbegin(); it != bw.rend(); ++it) append(it->first, it->second);
        append(seed, false);
        for (auto& p : fw) append(p.first, p.second);
        if (open_l) f.open |= OPEN_LEFT;
        if (open_r) f.open |= OPEN_RIGHT;
        res.frags.push_back(std::move(f));
    }
}

// Global dictionary (streaming_color_set_dict.hpp:86-392) at -t 1: classes
// get dense ids in first-intern order; each new class is hybrid-encoded and
// appended to one bit stream, its bit offset recorded (:148-182,350-365).
struct GlobalDict {
    u32 n, sparse, dense;
    std::map<std::vector<u32>, u64> index;
    BitStream bits;
    std::vector<u64> offsets;
    u64 total_integers = 0;
    explicit GlobalDict(u32 ncol)
        : n(ncol), sparse(u32(u64(ncol) / 4)), dense(u32(u64(ncol) * 3 / 4)) {}
    u64 intern(const std::vector<u32>& v) {
        auto it = index.find(v);
        if (it != index.end()) return it->second;
        u64 id = offsets.size();
        offsets.push_back(bits.nbits);
        encode_set(bits, v.data(), v.size(), n, sparse, dense);
        total_integers += v.size();
        index.emplace(v, id);
        return id;
    }
};

// frag_unitig_writer::write_batch record layout (unitig_spill.hpp:491-522):
// [u8 open][u32 nruns]{[u64 cid][u32 n]}x nruns [u32 seq_len][seq], LE.
static void serialize_frag(const Fragment& f, std::vector<u8>& out) {
    auto put = [&](const void* p, size_t n) {
        const u8* q = (const u8*)p;
        out.insert(out.end(), q, q + n);
    };
    u8 fl = f.open;
    u32 nr = u32(f.runs.size()), sl = u32(f.seq.size());
    put(&fl, 1);
    put(&nr, 4);
    for (auto& r : f.runs) {
        put(&r.first, 8);
        put(&r.second, 4);
    }
    put(&sl, 4);
    put(f.seq.data(), sl);
}

// ---------------------------------------------------------------------------
// Whole-job driver: phases 1+2 over the per-occurrence record stream.
// ---------------------------------------------------------------------------
struct Result {
    std::vector<Super> supers;         // phase 1, position order
    std::vector<u64> bucket_offsets;   // supers grouped by bucket (B+1)
    std::vector<Super> bucketed;       // phase 1, bucket-major, position order inside
    std::vector<u8> frag_bytes;        // phase 2 frag_unitigs.bin bytes
    std::vector<u64> cs_words;         // all words of the colour-set bit stream
    std::vector<u64> cs_offsets;       // bit offset of each class
    u64 cs_bits = 0;
    u64 num_kmers = 0;                 // distinct k-mers (sum of primary nodes)
    u64 num_nodes = 0;                 // nodes incl. foreign copies
    u64 num_occ = 0;
    u64 num_frags = 0;
    u64 num_sets = 0;
    u64 total_integers = 0;
    double t_phase1 = 0, t_phase2 = 0;
};

static double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

template <class T>
static void phase2(const std::vector<std::vector<Rec>>& bucket_recs, u32 k, u32 num_colors,
                   u32 threads, Result& R) {
    const u32 B = u32(bucket_recs.size());
    std::vector<BucketResult<T>> per(B);
    std::atomic<u32> next{0};
    auto work = [&]() {
        for (;;) {
            u32 b = next.fetch_add(1);
            if (b >= B) break;
            process_bucket<T>(bucket_recs[b], k, per[b]);
        }
    };
    if (threads <= 1) {
        work();
    } else {
        std::vector<std::thread> th;
        for (u32 t = 0; t < threads; ++t) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    // Merge in bucket order, local class order (bucket_walker.hpp:713-761).
    GlobalDict gd(num_colors);
    for (u32 b = 0; b < B; ++b) {
        BucketResult<T>& br = per[b];
        std::vector<u64> l2g(br.local.sets.size());
        for (u32 lc = 0; lc < l2g.size(); ++lc) l2g[lc] = gd.intern(*br.local.sets[lc]);
        for (auto& f : br.frags) {
            for (auto& r : f.runs)
                if (r.first != FOREIGN) r.first = l2g[r.first];
            serialize_frag(f, R.frag_bytes);
        }
        R.num_frags += br.frags.size();
        R.num_kmers += br.n_primary;
        R.num_nodes += br.n_nodes;
        R.num_occ += br.n_occ;
        br = BucketResult<T>();
    }
    R.cs_words = gd.bits.words;
    R.cs_bits = gd.bits.nbits;
    R.cs_offsets = gd.offsets;
    R.num_sets = gd.offsets.size();
    R.total_integers = gd.total_integers;
}

static void phase1(const u8* seq, const u64* color_off, u32 num_colors, u32 k, u32 m, u32 B,
                   u32 threads, Result& R) {
    std::vector<std::vector<Super>> per(num_colors);
    std::atomic<u32> next{0};
    auto work = [&]() {
        for (;;) {
            u32 c = next.fetch_add(1);
            if (c >= num_colors) break;
            supers_of_seq(seq + color_off[c], color_off[c + 1] - color_off[c], color_off[c], k, m, B,
                          c, per[c]);
        }
    };
    if (threads <= 1) {
        work();
    } else {
        std::vector<std::thread> th;
        for (u32 t = 0; t < threads; ++t) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    for (auto& v : per) R.supers.insert(R.supers.end(), v.begin(), v.end());
    R.bucket_offsets.assign(B + 1, 0);
    for (auto& s : R.supers) R.bucket_offsets[s.bucket + 1]++;
    for (u32 b = 0; b < B; ++b) R.bucket_offsets[b + 1] += R.bucket_offsets[b];
    R.bucketed.resize(R.supers.size());
    std::vector<u64> fill(R.bucket_offsets.begin(), R.bucket_offsets.end() - 1);
    for (auto& s : R.supers) R.bucketed[fill[s.bucket]++] = s;
}

static void build(const u8* seq, const u64* color_off, u32 num_colors, u32 k, u32 m, u32 B,
                  u32 threads, Result& R) {
    double t0 = now_s();
    phase1(seq, color_off, num_colors, k, m, B, threads, R);
    double t1 = now_s();
    // The bucket stream: one record per super-k-mer occurrence, colour-major
    // position order inside each bucket.
    std::vector<std::vector<Rec>> recs(B);
    for (u32 b = 0; b < B; ++b) {
        for (u64 i = R.bucket_offsets[b]; i < R.bucket_offsets[b + 1]; ++i) {
            const Super& s = R.bucketed[i];
            Rec r;
            r.colors = {s.color};
            r.flags = s.flags;
            r.bases.resize(s.len);
            for (u32 j = 0; j < s.len; ++j) r.bases[j] = code_of(seq[s.start + j]);
            recs[b].push_back(std::move(r));
        }
    }
    if (k <= 31)
        phase2<u64>(recs, k, num_colors, threads, R);
    else
        phase2<u128>(recs, k, num_colors, threads, R);
    double t2 = now_s();
    R.t_phase1 = t1 - t0;
    R.t_phase2 = t2 - t1;
}

// ---------------------------------------------------------------------------
// Bucket files (bucket_io.hpp:19-40,201-212,255-276,651-720)
// ---------------------------------------------------------------------------
static std::vector<u8> slurp(const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<u8> buf;
    u8 tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    return buf;
}

// LZ4 block decoder after the public LZ4 block format specification.
static bool lz4_decode(const u8* src, size_t n, u8* dst, size_t cap, size_t& out_len) {
    size_t ip = 0, op = 0;
    while (ip < n) {
        u8 tok = src[ip++];
        size_t lit = tok >> 4;
        if (lit == 15) {
            u8 b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (ip + lit > n || op + lit > cap) return false;
        memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip >= n) break;  // last sequence has literals only
        if (ip + 2 > n) return false;
        size_t off = size_t(src[ip]) | (size_t(src[ip + 1]) << 8);
        ip += 2;
        if (off == 0 || off > op) return false;
        size_t ml = tok & 15;
        if (ml == 15) {
            u8 b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += 4;
        if (op + ml > cap) return false;
        for (size_t i = 0; i < ml; ++i) dst[op + i] = dst[op - off + i];
        op += ml;
    }
    out_len = op;
    return true;
}

// bucket_reader (bucket_io.hpp:665-705) + read_super_kmer (super_kmer.hpp:121-143).
static std::vector<Rec> read_bucket_file(const std::string& path) {
    std::vector<u8> f = slurp(path);
    std::vector<u8> dec;
    size_t p = 0;
    while (p + 4 <= f.size()) {
        u32 u, c;
        memcpy(&u, &f[p], 4);
        p += 4;
        if (u == 0) break;
        if (p + 4 > f.size()) throw std::runtime_error("short frame header in " + path);
        memcpy(&c, &f[p], 4);
        p += 4;
        if (p + c > f.size()) throw std::runtime_error("short frame in " + path);
        size_t off = dec.size(), got = 0;
        dec.resize(off + u);
        if (!lz4_decode(&f[p], c, dec.data() + off, u, got) || got != u)
            throw std::runtime_error("LZ4 decode failed on " + path);
        p += c;
    }
    std::vector<Rec> recs;
    size_t q = 0;
    while (q < dec.size()) {
        Rec r;
        u64 nc = get_varint(dec.data(), dec.size(), q);
        u32 prev = 0;
        for (u64 i = 0; i < nc; ++i) {
            prev += u32(get_varint(dec.data(), dec.size(), q));
            r.colors.push_back(prev);
        }
        if (q >= dec.size()) break;
        u64 lf = get_varint(dec.data(), dec.size(), q);
        u64 len = lf >> 4;
        r.flags = u8(lf & 15);
        size_t nb = (len + 3) / 4;
        if (q + nb > dec.size()) break;
        r.bases.resize(len);
        for (u64 i = 0; i < len; ++i) r.bases[i] = (dec[q + (i >> 2)] >> (2 * (i & 3))) & 3;
        q += nb;
        recs.push_back(std::move(r));
    }
    return recs;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes) and the CPU-baseline leg of bench.py.
// ---------------------------------------------------------------------------
using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

uint32_t orc_best_m(uint32_t k) { return best_m(k); }

// Canonical m-mer hash at every position of a 2-bit code array.
void orc_mmer_hashes(const uint8_t* codes, uint64_t n, uint32_t m, uint64_t* out) {
    for (uint64_t p = 0; p + m <= n; ++p) out[p] = mmer_hash(codes + p, m);
}

// Window minima over W consecutive m-mer hashes for every (k-1)-mer.
void orc_window_min(const uint8_t* codes, uint64_t n, uint32_t k, uint32_t m, uint64_t* out) {
    uint32_t W = k - m;
    std::vector<uint64_t> h(n - m + 1);
    for (uint64_t p = 0; p + m <= n; ++p) h[p] = mmer_hash(codes + p, m);
    for (uint64_t j = 0; j + k - 1 <= n; ++j) {
        uint64_t b = ~0ULL;
        for (uint64_t p = j; p < j + W; ++p) b = std::min(b, h[p]);
        out[j] = b;
    }
}

uint64_t orc_revcomp64(uint64_t x, uint32_t k) { return revcomp<uint64_t>(x, k); }
void orc_revcomp128(uint64_t lo, uint64_t hi, uint32_t k, uint64_t* out) {
    u128 x = (u128(hi) << 64) | lo;
    u128 r = revcomp<u128>(x, k);
    out[0] = uint64_t(r);
    out[1] = uint64_t(r >> 64);
}

// Encode one colour set with the hybrid codec; returns the bit count and
// writes ceil(bits/64) words to `out` (capacity `cap` words).
uint64_t orc_encode_set(const uint32_t* c, uint64_t n, uint32_t num_colors, uint64_t* out,
                        uint64_t cap) {
    BitStream bs;
    encode_set(bs, c, n, num_colors, u64(num_colors) / 4, u64(num_colors) * 3 / 4);
    for (uint64_t i = 0; i < bs.words.size() && i < cap; ++i) out[i] = bs.words[i];
    return bs.nbits;
}

void orc_write_delta(uint64_t x, uint64_t* out_words, uint64_t* out_bits) {
    BitStream bs;
    write_delta(bs, x);
    out_words[0] = bs.words.empty() ? 0 : bs.words[0];
    out_words[1] = bs.words.size() > 1 ? bs.words[1] : 0;
    *out_bits = bs.nbits;
}

// Phase-1 super-k-mers of one colour's bytes (absolute base offset `abs0`).
// Returns the count; fills up to `cap` entries of each output array.
uint64_t orc_supers(const uint8_t* seq, uint64_t len, uint64_t abs0, uint32_t k, uint32_t m,
                    uint32_t B, uint32_t color, uint64_t cap, uint64_t* start, uint32_t* slen,
                    uint8_t* flags, uint32_t* bucket) {
    std::vector<Super> v;
    supers_of_seq(seq, len, abs0, k, m, B, color, v);
    for (uint64_t i = 0; i < v.size() && i < cap; ++i) {
        start[i] = v[i].start;
        slen[i] = v[i].len;
        flags[i] = v[i].flags;
        bucket[i] = v[i].bucket;
    }
    return v.size();
}

// Whole job (phases 1+2).  Returns an opaque handle or NULL (orc_last_error).
void* orc_build(const uint8_t* seq, const uint64_t* color_off, uint32_t num_colors, uint32_t k,
                uint32_t m, uint32_t B, uint32_t threads) {
    try {
        if (k < 2 || k > 63 || m < 2 || m >= k || B == 0) throw std::runtime_error("bad k/m/B");
        Result* R = new Result();
        build(seq, color_off, num_colors, k, m, B, threads, *R);
        return R;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// The same under a bucket map (n_cuts = B - 1 ascending cuts).
void* orc_build_mapped(const uint8_t* seq, const uint64_t* color_off, uint32_t num_colors, uint32_t k,
                       uint32_t m, uint32_t B, uint32_t threads, const uint64_t* cuts, uint32_t n_cuts) {
    if (!cuts || n_cuts + 1 != B) {
        g_err = "bucket map: n_cuts must be B - 1";
        return nullptr;
    }
    g_cuts = cuts;
    void* r = orc_build(seq, color_off, num_colors, k, m, B, threads);
    g_cuts = nullptr;
    return r;
}

// Phase 2 over bucket files bucket_<b>.bin in `dir` (the drop-in input).
void* orc_process_bucket_files(const char* dir, uint32_t B, uint32_t k, uint32_t num_colors,
                               uint32_t threads) {
    try {
        std::vector<std::vector<Rec>> recs(B);
        for (uint32_t b = 0; b < B; ++b)
            recs[b] = read_bucket_file(std::string(dir) + "/bucket_" + std::to_string(b) + ".bin");
        Result* R = new Result();
        double t0 = now_s();
        if (k <= 31)
            phase2<u64>(recs, k, num_colors, threads, *R);
        else
            phase2<u128>(recs, k, num_colors, threads, *R);
        R->t_phase2 = now_s() - t0;
        return R;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_free(void* h) { delete (Result*)h; }

// Scalar statistics: 0 num_kmers, 1 num_nodes, 2 num_occ, 3 num_frags,
// 4 num_sets, 5 cs_bits, 6 total_integers, 7 num_supers, 8 frag bytes.
uint64_t orc_stat(void* h, int which) {
    Result* R = (Result*)h;
    switch (which) {
        case 0: return R->num_kmers;
        case 1: return R->num_nodes;
        case 2: return R->num_occ;
        case 3: return R->num_frags;
        case 4: return R->num_sets;
        case 5: return R->cs_bits;
        case 6: return R->total_integers;
        case 7: return R->supers.size();
        case 8: return R->frag_bytes.size();
        default: return 0;
    }
}
double orc_time(void* h, int which) {
    Result* R = (Result*)h;
    return which == 0 ? R->t_phase1 : R->t_phase2;
}

void orc_copy_frags(void* h, uint8_t* dst) {
    Result* R = (Result*)h;
    memcpy(dst, R->frag_bytes.data(), R->frag_bytes.size());
}
void orc_copy_cs(void* h, uint64_t* words, uint64_t* offsets) {
    Result* R = (Result*)h;
    u64 nw = (R->cs_bits + 63) / 64;
    for (u64 i = 0; i < nw; ++i) words[i] = i < R->cs_words.size() ? R->cs_words[i] : 0;
    memcpy(offsets, R->cs_offsets.data(), R->cs_offsets.size() * 8);
}
// which = 0: supers in position order; 1: bucket-major order.
void orc_copy_supers(void* h, int which, uint64_t* start, uint32_t* len, uint8_t* flags,
                     uint32_t* bucket, uint32_t* color) {
    Result* R = (Result*)h;
    auto& v = which ? R->bucketed : R->supers;
    for (size_t i = 0; i < v.size(); ++i) {
        start[i] = v[i].start;
        len[i] = v[i].len;
        flags[i] = v[i].flags;
        bucket[i] = v[i].bucket;
        color[i] = v[i].color;
    }
}

int orc_lz4_decode(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap, uint64_t* out_len) {
    size_t o = 0;
    bool ok = lz4_decode(src, n, dst, cap, o);
    *out_len = o;
    return ok ? 0 : -1;
}

}  // extern "C"
