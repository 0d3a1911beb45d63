// Phase 3 (stitch) on the device: the unitigs of the host stitch
// (stitch.cpp), in the same order, from sorts, scans and list ranking.
//
// Semantics (stitch_extmem.hpp:112-355, the same as the host stitch): open
// fragment sides are keyed on their canonical boundary k-mer; two sides with
// one key and compatible orientation join (ext_pair_compatible :296-300);
// the shared k-mer of a join is kept once with the colour of the side that
// owns it (ext_concat_runs :170-226); tigs are cut at colour-set changes and
// foreign placeholder runs at their ends are dropped (ext_split_monochromatic
// :112-151).
//
// Method.  The stream is on the device (a file goes through pinned staging
// buffers); its record starts are found there (see k_st_parse_guess).  A
// fragment side s = fragment << 1 | side; entering fragment i at
// side s and leaving it by s ^ 1 continues into partner[s ^ 1], so
// next(s) = partner[s ^ 1] is injective and its graph is paths and cycles.
//   1. one thread per fragment checks the record (length, run coverage),
//      keeps its first/last colour ids and emits the canonical key of every
//      open side;
//   2. the keys are radix sorted; equal neighbours pair up;
//   3. pointer jumping over next() gives every side the last side of its
//      path and its distance to it, until a round resolves no side (what is
//      left lies on cycles);
//   4. a fragment is walked in the direction whose free start side is the
//      smaller of the path's two free sides (the host stitch's rule); its
//      position is the reverse direction's distance, chains are ordered by
//      their start side, and a scan of chain lengths places every fragment
//      at its slot in tig order;
//   5. per slot: the k-mers, bases and colour runs it adds to its tig (the
//      first k-mer of a joined fragment is the previous one's last; that
//      shared k-mer takes the first side's colour unless it is foreign);
//      scans give the offsets;
//   6. runs of k-mers with one colour set across a tig are the unitigs;
//      foreign ones at tig ends are dropped, inside a tig they are an error;
//   7. a stable radix sort by colour-set id gives the emit order, and two
//      copies (fragments -> tig sequences -> unitigs) give the bases.
// Fragments on cycles go back to the host stitch (stitch_cycles) with their
// records, and are walked in fragment order exactly as it would have.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <fcntl.h>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "cdbg_gpu.h"
#include "common.h"
#include "job.h"
#include "kmer_ops.h"
#include "prims.h"
#include "stitch.h"

namespace cdbg_amd {

namespace {

constexpr int ST_TPB = 256;
constexpr uint32_t NONE32 = ~0u;
constexpr uint32_t DONE = 1u << 31;  // jump state: .x is the last side of the path
constexpr uint32_t PIECE = 4096;     // bytes one wave copies per work item
constexpr uint32_t SMALL_COPY = 128; // copies up to this size take one thread, larger ones waves

// error bits, in the order the host stitch meets them
enum : uint32_t {
    E_SHORT = 1,
    E_COVER = 2,
    E_BASE = 4,
    E_GROUP = 8,
    E_REENTER = 16,
    E_FOREIGN = 32,
    E_INTERNAL = 64,
};

inline uint32_t st_blocks(uint64_t n, uint32_t t = ST_TPB) {
    const uint64_t b = (n + t - 1) / t;
    if (b >= (1ull << 31)) throw Error("stitch: launch too large");
    return uint32_t(b ? b : 1);
}

__device__ __forceinline__ uint32_t ld32(const uint8_t* p) {
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ __forceinline__ uint64_t ld64(const uint8_t* p) {
    uint64_t v;
    __builtin_memcpy(&v, p, 8);
    return v;
}

// A counter is N_STRIPE words: a block adds its total to the word of its
// own stripe, so no address takes more than a few atomics at a time (one
// word taking an atomic from every wave serialises the launch); the host
// adds the words up.
constexpr uint32_t N_STRIPE = 256;

// Sum over the block of each thread's v (every thread of the block calls it).
__device__ __forceinline__ uint64_t block_sum(uint64_t v) {
    __shared__ unsigned long long s_part[ST_TPB / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int w = 0; w < ST_TPB / 64; ++w) t += s_part[w];
    __syncthreads();  // s_part is free for the next call
    return t;
}

__device__ __forceinline__ uint64_t block_max(uint64_t v) {
    __shared__ unsigned long long s_part[ST_TPB / 64];
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(v, o, 64);
        v = y > v ? y : v;
    }
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int w = 0; w < ST_TPB / 64; ++w) t = s_part[w] > t ? s_part[w] : t;
    __syncthreads();
    return t;
}

__device__ __forceinline__ void stripe_add(unsigned long long* c, uint64_t v) {
    const uint64_t t = block_sum(v);
    if (threadIdx.x == 0 && t) atomicAdd(&c[blockIdx.x % N_STRIPE], (unsigned long long)t);
}

__device__ __forceinline__ uint8_t comp_of(uint8_t c) {
    return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
}

template <class K>
__device__ __forceinline__ K push_base(const K& x, uint32_t c);
template <>
__device__ __forceinline__ uint64_t push_base<uint64_t>(const uint64_t& x, uint32_t c) {
    return (x << 2) | c;
}
template <>
__device__ __forceinline__ Key128 push_base<Key128>(const Key128& x, uint32_t c) {
    return Key128{(x.lo << 2) | c, (x.hi << 2) | (x.lo >> 62)};
}
// four bases (8 bits of codes) at once
template <class K>
__device__ __forceinline__ K push_4(const K& x, uint32_t c8);
template <>
__device__ __forceinline__ uint64_t push_4<uint64_t>(const uint64_t& x, uint32_t c8) {
    return (x << 8) | c8;
}
template <>
__device__ __forceinline__ Key128 push_4<Key128>(const Key128& x, uint32_t c8) {
    return Key128{(x.lo << 8) | c8, (x.hi << 8) | (x.lo >> 56)};
}

// 2-bit codes of the four ASCII bytes of w (first byte lowest), first base
// in the highest bits; *bad |= 1 when a byte is not A, C, G or T.
__device__ __forceinline__ uint32_t codes4(uint32_t w, uint32_t& bad) {
    uint32_t c8 = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t b = (w >> (8 * i)) & 0xFFu;
        bad |= uint32_t(b != 'A' && b != 'C' && b != 'G' && b != 'T');
        c8 = (c8 << 2) | (((b >> 1) ^ (b >> 2)) & 3u);  // A 0, C 1, G 2, T 3
    }
    return c8;
}

__device__ __forceinline__ bool key_le(uint64_t a, uint64_t b) { return a <= b; }
__device__ __forceinline__ bool key_le(const Key128& a, const Key128& b) { return !(b < a); }

// The canonical k-mer of s[0, k) and whether the forward word is it: the
// bytes are read once, four per load where whole words lie inside the
// k-mer; the reverse complement comes from the forward word.
template <class K>
__device__ __forceinline__ K canonical_at(const uint8_t* __restrict__ s, uint32_t k, bool& fwd, uint32_t& bad) {
    K f{};
    uint32_t j = 0;
    for (; j + 4 <= k; j += 4) f = push_4<K>(f, codes4(ld32(s + j), bad));
    for (; j < k; ++j) {
        const uint32_t b = s[j];
        bad |= uint32_t(b != 'A' && b != 'C' && b != 'G' && b != 'T');
        f = push_base<K>(f, ((b >> 1) ^ (b >> 2)) & 3u);
    }
    const K r = KOps<K>::rc(f, k);
    fwd = key_le(f, r);
    return fwd ? f : r;
}

// 1. Open sides per block of fragments (their ends are laid out block by
// block, in fragment order).
__global__ __launch_bounds__(ST_TPB) void k_st_count_ends(uint64_t nf, const uint8_t* __restrict__ buf,
                                                          const uint64_t* __restrict__ off,
                                                          uint64_t* __restrict__ block_cnt) {
    const uint64_t i = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    const uint32_t c = i < nf ? uint32_t(__popc(buf[off[i]] & 3u)) : 0u;
    const uint64_t t = block_sum(c);
    if (threadIdx.x == 0) block_cnt[blockIdx.x] = t;
}

// Per fragment: checks, colour ids, keys of the open sides (at the block's
// base plus the open sides before it in the block).
template <class K>
__global__ __launch_bounds__(ST_TPB) void k_st_frags(uint64_t nf, const uint8_t* __restrict__ buf,
                                                     const uint64_t* __restrict__ off, uint32_t k,
                                                     const uint64_t* __restrict__ block_base,
                                                     uint32_t* __restrict__ f_len, uint32_t* __restrict__ f_nr,
                                                     uint64_t* __restrict__ f_cid, uint2* __restrict__ f_pos,
                                                     K* __restrict__ ekey,
                                                     uint32_t* __restrict__ eid, uint32_t* __restrict__ err) {
    __shared__ uint64_t s_tmp[ST_TPB / 64 + 1];
    const uint64_t i = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    const bool live = i < nf;
    uint8_t open = 0;
    uint32_t len = 0, nr = 0;
    bool ok = false;
    const uint8_t* rec = nullptr;
    if (live) {
        rec = buf + off[i];
        open = rec[0] & 3u;
        nr = ld32(rec + 1);
        len = ld32(rec + 5 + 12 * uint64_t(nr));
        f_len[i] = len;
        f_nr[i] = nr;
        uint32_t bad = 0;
        if (len < k) {
            bad = E_SHORT;
        } else {
            // colour ids: first and last run, first and last run with k-mers
            uint64_t sum = 0, c0 = 0, cl = 0, p0 = COLOR_FOREIGN, pl = COLOR_FOREIGN;
            uint32_t i0 = 0, il = 0;  // runs of the first and last colour with k-mers
            bool any = false;
            for (uint32_t r = 0; r < nr; ++r) {
                const uint64_t cid = ld64(rec + 5 + 12 * uint64_t(r));
                const uint32_t n = ld32(rec + 13 + 12 * uint64_t(r));
                if (r == 0) c0 = cid;
                cl = cid;
                if (n) {
                    if (!any) {
                        p0 = cid;
                        i0 = r;
                    }
                    pl = cid;
                    il = r;
                    any = true;
                }
                sum += n;
            }
            if (sum != uint64_t(len) - k + 1) bad = E_COVER;
            f_cid[4 * i + 0] = c0;
            f_cid[4 * i + 1] = cl;
            f_cid[4 * i + 2] = p0;
            f_cid[4 * i + 3] = pl;
            f_pos[i] = make_uint2(i0, il);
        }
        if (bad) atomicOr(err, bad);
        ok = !bad;
    }
    uint64_t tot;
    uint64_t at = block_base[blockIdx.x] + block_exclusive_scan<uint64_t>(uint64_t(__popc(open)), s_tmp, tot);
    // open sides: canonical boundary k-mer, and whether the forward word is it
    // (a record that failed its checks gets placeholder keys: the stitch stops)
    for (uint32_t side = 0; side < 2; ++side) {
        if (!(open & (side ? 2u : 1u))) continue;
        K key{};
        bool fwd = false;
        if (ok) {
            uint32_t bad = 0;
            key = canonical_at<K>(rec + 9 + 12 * uint64_t(nr) + (side ? len - k : 0), k, fwd, bad);
            if (bad) atomicOr(err, E_BASE);
        }
        ekey[at] = key;
        eid[at] = uint32_t(i << 2) | (side << 1) | uint32_t(fwd);
        ++at;
    }
}

template <class K>
__device__ __forceinline__ bool key_eq(const K& a, const K& b) {
    return a == b;
}

// 2. Equal keys pair up (counters: 1 joins, 2 unmatched ends).
template <class K>
__global__ __launch_bounds__(ST_TPB) void k_st_pair(uint64_t ne, const K* __restrict__ key,
                                                    const uint32_t* __restrict__ id, uint32_t* __restrict__ partner,
                                                    unsigned long long* __restrict__ ctr, uint32_t* __restrict__ err) {
    const uint64_t j = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    uint64_t joins = 0, unmatched = 0;
    if (j < ne && (j == 0 || !key_eq(key[j], key[j - 1]))) {
        const K x = key[j];
        const bool two = j + 1 < ne && key_eq(key[j + 1], x);
        if (two && j + 2 < ne && key_eq(key[j + 2], x)) atomicOr(err, E_GROUP);
        bool paired = false;
        if (two) {
            const uint32_t a = id[j], b = id[j + 1];
            const uint32_t sa = (a >> 1) & 1u, sb = (b >> 1) & 1u, fa = a & 1u, fb = b & 1u;
            const bool ok = sa != sb ? fa == fb : fa != fb;
            if (ok && (a >> 1) != (b >> 1)) {
                partner[a >> 1] = b >> 1;
                partner[b >> 1] = a >> 1;
                paired = true;
                joins = 1;
            }
        }
        if (!paired) unmatched = two ? 2 : 1;
    }
    stripe_add(ctr + 1 * N_STRIPE, joins);
    stripe_add(ctr + 2 * N_STRIPE, unmatched);
}

// 3. Pointer jumping over next(s) = partner[s ^ 1].  A side's state is
// (target side, distance to it), packed in 64 bits; with DONE set in the
// target, the target is the last side of the path.
__device__ __forceinline__ uint64_t st_pack(uint32_t x, uint32_t y) { return uint64_t(x) | (uint64_t(y) << 32); }
__device__ __forceinline__ uint32_t st_x(uint64_t v) { return uint32_t(v); }
__device__ __forceinline__ uint32_t st_y(uint64_t v) { return uint32_t(v >> 32); }

__global__ __launch_bounds__(ST_TPB) void k_st_jump_init(uint64_t ns, const uint32_t* __restrict__ partner,
                                                         uint64_t* __restrict__ S) {
    const uint64_t s = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (s >= ns) return;
    const uint32_t nx = partner[s ^ 1];
    S[s] = nx == NONE32 ? st_pack(uint32_t(s) | DONE, 0u) : st_pack(nx, 1u);
}

// One round, in place: every open side takes its target's state.  A state
// is read and written whole (one 64-bit access), so it is always some
// (side, distance to it) pair of the path, and a round moves every open
// side at least twice as far as the round before (the target's state is at
// least as advanced as at the round's start).  Counts the sides still open.
__global__ __launch_bounds__(ST_TPB) void k_st_jump(uint64_t ns, uint64_t* S, unsigned long long* __restrict__ open_ctr) {
    const uint64_t s = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    bool open = false;
    if (s < ns) {
        uint64_t v = __hip_atomic_load(&S[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(st_x(v) & DONE)) {
            const uint64_t w = __hip_atomic_load(&S[st_x(v)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = st_pack(st_x(w), st_y(v) + st_y(w));
            __hip_atomic_store(&S[s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            open = !(st_x(w) & DONE);
        }
    }
    stripe_add(open_ctr, open ? 1u : 0u);
}

// 4. The walked direction of every fragment and its place.
__global__ __launch_bounds__(ST_TPB) void k_st_choose(uint64_t nf, const uint64_t* __restrict__ S,
                                                      uint32_t* __restrict__ ch_side, uint32_t* __restrict__ ch_start,
                                                      uint32_t* __restrict__ ch_pos, uint32_t* __restrict__ chain_len,
                                                      uint8_t* __restrict__ on_cycle, uint32_t* __restrict__ err) {
    const uint64_t i = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (i >= nf) return;
    const uint64_t v0 = S[2 * i], v1 = S[2 * i + 1];
    const uint2 s0 = make_uint2(st_x(v0), st_y(v0)), s1 = make_uint2(st_x(v1), st_y(v1));
    if (!(s0.x & DONE) || !(s1.x & DONE)) {
        // a side on a cycle: so is the other (the reverse direction of a cycle is one)
        if ((s0.x & DONE) || (s1.x & DONE)) atomicOr(err, E_INTERNAL);
        on_cycle[i] = 1;
        ch_side[i] = NONE32;
        return;
    }
    on_cycle[i] = 0;
    const uint32_t end0 = (s0.x & ~DONE) ^ 1u, end1 = (s1.x & ~DONE) ^ 1u;  // free exit side of each direction
    if (end0 == end1) {
        atomicOr(err, E_REENTER);  // both directions of one path
        ch_side[i] = NONE32;
        return;
    }
    // entering at side c, the path starts at the other direction's free end
    const bool take0 = end1 < end0;
    const uint32_t c = uint32_t(2 * i) + (take0 ? 0u : 1u);
    const uint2 sc = take0 ? s0 : s1, so = take0 ? s1 : s0;
    const uint32_t start = take0 ? end1 : end0;
    ch_side[i] = c;
    ch_start[i] = start;
    ch_pos[i] = so.y;
    if (so.y == 0) {
        if (start != c) atomicOr(err, E_INTERNAL);
        chain_len[c] = sc.y + 1;
    }
}

__global__ __launch_bounds__(ST_TPB) void k_st_place(uint64_t nf, const uint32_t* __restrict__ ch_side,
                                                     const uint32_t* __restrict__ ch_start,
                                                     const uint32_t* __restrict__ ch_pos,
                                                     const uint32_t* __restrict__ chain_off, uint64_t n_lin,
                                                     uint32_t* __restrict__ slot_side, uint32_t* __restrict__ err) {
    const uint64_t i = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (i >= nf || ch_side[i] == NONE32) return;
    const uint64_t slot = uint64_t(chain_off[ch_start[i]]) + ch_pos[i];
    if (slot >= n_lin) {
        atomicOr(err, E_INTERNAL);
        return;
    }
    slot_side[slot] = ch_side[i];
}

// 5. What each slot adds to its tig: k-mers, bases, run entries (nr + 1),
// and the tig it belongs to (flag of a tig's first slot, scanned).
__global__ __launch_bounds__(ST_TPB) void k_st_slot_sizes(uint64_t n_lin, const uint32_t* __restrict__ slot_side,
                                                          const uint32_t* __restrict__ ch_pos,
                                                          const uint64_t* __restrict__ S, const uint32_t* __restrict__ f_len,
                                                          const uint32_t* __restrict__ f_nr, uint32_t k,
                                                          uint64_t* __restrict__ s_kmers, uint64_t* __restrict__ s_bases,
                                                          uint64_t* __restrict__ s_runs, uint32_t* __restrict__ s_first,
                                                          uint32_t* __restrict__ err) {
    const uint64_t j = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (j >= n_lin) return;
    const uint32_t s = slot_side[j];
    if (s == NONE32) {
        atomicOr(err, E_INTERNAL);
        s_kmers[j] = s_bases[j] = s_runs[j] = 0;
        s_first[j] = 0;
        return;
    }
    const uint32_t i = s >> 1;
    const bool joined = ch_pos[i] > 0, last = st_y(S[s]) == 0;
    const uint64_t len = f_len[i], nk = len - k + 1;
    const uint64_t m = nk - (joined ? 1 : 0);
    // a joined fragment that is not the last of its tig adds a k-mer of its own
    if (joined && !last && m == 0) atomicOr(err, E_INTERNAL);
    s_kmers[j] = m;
    s_bases[j] = joined ? len - k : len;
    s_runs[j] = uint64_t(f_nr[i]) + 1;
    s_first[j] = joined ? 0u : 1u;
}

// Run entries of a slot in walk order: the fragment's runs (reversed when it
// is entered at its right side), less its first k-mer when joined (taken
// from its first run with k-mers), and its last k-mer recoloured by the next
// join when that differs (one k-mer less in its last run with k-mers, and the
// slot's extra entry holds the new colour); the extra entry is empty
// otherwise.  One thread per entry, its slot found among the block's slots
// in LDS (every slot has at least one entry), so consecutive entries are
// written by consecutive lanes.
__global__ __launch_bounds__(ST_TPB) void k_st_runs(uint64_t n_ent, uint64_t n_lin, const uint64_t* __restrict__ r_off,
                                                    const uint32_t* __restrict__ slot_side,
                                                    const uint32_t* __restrict__ ch_pos, const uint64_t* __restrict__ S,
                                                    const uint8_t* __restrict__ buf, const uint64_t* __restrict__ off,
                                                    const uint32_t* __restrict__ f_nr, const uint64_t* __restrict__ f_cid,
                                                    const uint2* __restrict__ f_pos, const uint32_t* __restrict__ s_tig,
                                                    uint64_t* __restrict__ e_cid, uint32_t* __restrict__ e_n,
                                                    uint32_t* __restrict__ e_tig, uint32_t* __restrict__ err) {
    __shared__ uint64_t s_off[ST_TPB + 2];
    __shared__ uint64_t s_j0;
    const uint64_t t0 = uint64_t(blockIdx.x) * ST_TPB;
    if (threadIdx.x == 0) {  // the slot holding entry t0: the last with r_off <= t0
        uint64_t lo = 0, hi = n_lin;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (r_off[mid] <= t0)
                lo = mid;
            else
                hi = mid;
        }
        s_j0 = lo;
    }
    __syncthreads();
    const uint64_t j0 = s_j0;
    for (uint32_t q = threadIdx.x; q < ST_TPB + 2; q += ST_TPB) s_off[q] = j0 + q < n_lin ? r_off[j0 + q] : n_ent;
    __syncthreads();
    const uint64_t t = t0 + threadIdx.x;
    if (t >= n_ent) return;
    uint32_t lo = 0, hi = ST_TPB + 1;  // s_off[lo] <= t < s_off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    const uint64_t j = j0 + lo;
    if (j >= n_lin || s_off[lo] > t) {
        atomicOr(err, E_INTERNAL);
        return;
    }
    const uint32_t s = slot_side[j];
    const uint32_t i = s >> 1;
    const bool rev = s & 1u, joined = ch_pos[i] > 0, last = st_y(S[s]) == 0;
    const uint32_t nr = f_nr[i];
    const uint64_t x = t - s_off[lo];  // entry of the slot, walk order; nr is the extra one
    if (x > nr) {
        atomicOr(err, E_INTERNAL);
        return;
    }
    const uint32_t tig = joined ? s_tig[j] - 1u : s_tig[j];
    // the shared k-mer with the next fragment: this side's colour unless
    // foreign, else the next fragment's first (walk order) run's
    bool recolour = false;
    uint64_t real = 0;
    if (!last) {
        const uint32_t nx = slot_side[j + 1];
        const uint64_t mine = f_cid[4 * uint64_t(i) + (rev ? 2 : 3)];
        const uint64_t theirs = f_cid[4 * uint64_t(nx >> 1) + ((nx & 1u) ? 1 : 0)];
        real = mine != COLOR_FOREIGN ? mine : theirs;
        recolour = real != mine;
    }
    if (x == nr) {
        e_cid[t] = recolour ? real : 0;
        e_n[t] = recolour ? 1u : 0u;
        e_tig[t] = tig;
        return;
    }
    const uint2 pos = f_pos[i];
    const uint32_t first_pos = rev ? pos.y : pos.x, last_pos = rev ? pos.x : pos.y;  // stored run indices
    const uint32_t r = rev ? nr - 1 - uint32_t(x) : uint32_t(x);
    const uint8_t* run = buf + off[i] + 5 + 12 * uint64_t(r);
    uint32_t take = ld32(run + 8);
    if (joined && r == first_pos) --take;   // the shared first k-mer (the run has k-mers)
    if (recolour && r == last_pos) --take;  // the recoloured last k-mer
    e_cid[t] = ld64(run);
    e_n[t] = take;
    e_tig[t] = tig;
}

// 6. Unitig heads among the entries with k-mers (compacted, in tig order):
// the first entry, a new tig or a new colour id.
__global__ __launch_bounds__(ST_TPB) void k_st_heads(uint64_t nk_e, const uint32_t* __restrict__ list,
                                                     const uint64_t* __restrict__ e_cid,
                                                     const uint32_t* __restrict__ e_tig, uint8_t* __restrict__ head) {
    const uint64_t q = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (q >= nk_e) return;
    const uint32_t a = list[q];
    bool h = q == 0;
    if (!h) {
        const uint32_t b = list[q - 1];
        h = e_tig[a] != e_tig[b] || e_cid[a] != e_cid[b];
    }
    head[q] = h ? 1 : 0;
}

// Unitig candidates (one per head): colour id, first k-mer, k-mers, tig;
// foreign ones at a tig's ends are dropped, inside a tig they are an error.
__global__ __launch_bounds__(ST_TPB) void k_st_unitigs(uint64_t n_cand, uint64_t nk_e, const uint32_t* __restrict__ hlist,
                                                       const uint32_t* __restrict__ list,
                                                       const uint64_t* __restrict__ e_cid,
                                                       const uint32_t* __restrict__ e_tig,
                                                       const uint64_t* __restrict__ e_x, uint64_t total_kmers,
                                                       uint64_t* __restrict__ u_cid, uint64_t* __restrict__ u_x,
                                                       uint64_t* __restrict__ u_n, uint32_t* __restrict__ u_tig,
                                                       uint8_t* __restrict__ keep,
                                                       unsigned long long* __restrict__ ctr,
                                                       uint32_t* __restrict__ err) {
    const uint64_t u = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    uint64_t dropped = 0, kept_kmers = 0;
    if (u < n_cand) {
        const uint64_t q = hlist[u];
        const uint32_t a = list[q];
        const uint64_t x = e_x[q];
        const uint64_t x1 = u + 1 < n_cand ? e_x[hlist[u + 1]] : total_kmers;
        const uint64_t cid = e_cid[a];
        const uint32_t tig = e_tig[a];
        const bool first = u == 0 || e_tig[list[hlist[u - 1]]] != tig;
        const bool lastu = u + 1 == n_cand || e_tig[list[hlist[u + 1]]] != tig;
        bool k_ = true;
        if (cid == COLOR_FOREIGN) {
            k_ = false;
            if (first || lastu)
                dropped = x1 - x;
            else
                atomicOr(err, E_FOREIGN);
        } else {
            kept_kmers = x1 - x;
        }
        u_cid[u] = cid;
        u_x[u] = x;
        u_n[u] = x1 - x;
        u_tig[u] = tig;
        keep[u] = k_ ? 1 : 0;
    }
    stripe_add(ctr + 3 * N_STRIPE, dropped);
    stripe_add(ctr + 4 * N_STRIPE, kept_kmers);
}

// 7. Emit order: kept unitigs (tig order) with their colour ids as sort keys.
__global__ __launch_bounds__(ST_TPB) void k_st_keys(uint64_t nu, const uint32_t* __restrict__ klist,
                                                    const uint64_t* __restrict__ u_cid, uint64_t* __restrict__ key,
                                                    uint32_t* __restrict__ val, unsigned long long* __restrict__ cmax) {
    const uint64_t t = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    uint64_t c = 0;
    if (t < nu) {
        c = u_cid[klist[t]];
        key[t] = c;
        val[t] = klist[t];
    }
    c = block_max(c);
    if (threadIdx.x == 0) atomicMax(&cmax[blockIdx.x % N_STRIPE], (unsigned long long)c);
}

__global__ __launch_bounds__(ST_TPB) void k_st_out_len(uint64_t nu, const uint32_t* __restrict__ ord,
                                                       const uint64_t* __restrict__ u_n, uint32_t k,
                                                       uint64_t* __restrict__ olen) {
    const uint64_t t = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (t < nu) olen[t] = u_n[ord[t]] + k - 1;
}

// The host's unitig items {cid, offset, length} in emit order.
__global__ __launch_bounds__(ST_TPB) void k_st_items(uint64_t nu, const uint32_t* __restrict__ ord,
                                                     const uint64_t* __restrict__ u_cid, const uint64_t* __restrict__ u_n,
                                                     const uint64_t* __restrict__ ooff, uint32_t k,
                                                     uint64_t* __restrict__ items) {
    const uint64_t t = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (t >= nu) return;
    const uint32_t u = ord[t];
    items[3 * t + 0] = u_cid[u];
    items[3 * t + 1] = ooff[t];
    items[3 * t + 2] = u_n[u] + k - 1;
}

// Work items of PIECE bytes: item_of[piece] = the element it copies from.
__global__ __launch_bounds__(ST_TPB) void k_st_pieces(uint64_t n, const uint64_t* __restrict__ bytes_scan,
                                                      uint64_t total, const uint64_t* __restrict__ piece_off,
                                                      uint32_t* __restrict__ item_of) {
    const uint64_t e = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (e >= n) return;
    const uint64_t b = (e + 1 < n ? bytes_scan[e + 1] : total) - bytes_scan[e];
    const uint64_t p0 = piece_off[e], np = b > SMALL_COPY ? (b + PIECE - 1) / PIECE : 0;
    for (uint64_t p = 0; p < np; ++p) item_of[p0 + p] = uint32_t(e);
}

__global__ __launch_bounds__(ST_TPB) void k_st_piece_count(uint64_t n, const uint64_t* __restrict__ bytes_scan,
                                                           uint64_t total, uint64_t* __restrict__ np) {
    const uint64_t e = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (e >= n) return;
    const uint64_t b = (e + 1 < n ? bytes_scan[e + 1] : total) - bytes_scan[e];
    np[e] = b > SMALL_COPY ? (b + PIECE - 1) / PIECE : 0;
}

// Tig sequences: slot j's bases (reverse complement when entered at the
// right side; without the shared k bases when joined) at s_boff[j].
__global__ __launch_bounds__(ST_TPB) void k_st_fill_tigs(uint64_t n_pieces, const uint32_t* __restrict__ item_of,
                                                         const uint64_t* __restrict__ piece_off,
                                                         const uint32_t* __restrict__ slot_side,
                                                         const uint32_t* __restrict__ ch_pos,
                                                         const uint8_t* __restrict__ buf,
                                                         const uint64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ f_nr,
                                                         const uint32_t* __restrict__ f_len,
                                                         const uint64_t* __restrict__ s_boff, uint32_t k,
                                                         uint64_t total_tig, uint8_t* __restrict__ tig,
                                                         uint32_t* __restrict__ err) {
    const uint64_t w = (uint64_t(blockIdx.x) * ST_TPB + threadIdx.x) >> 6;
    if (w >= n_pieces) return;
    const int lane = threadIdx.x & 63;
    const uint32_t j = item_of[w];
    const uint32_t s = slot_side[j], i = s >> 1;
    const bool rev = s & 1u;
    const uint32_t len = f_len[i];
    const uint32_t from = ch_pos[i] > 0 ? k : 0u;
    const uint8_t* seq = buf + off[i] + 9 + 12 * uint64_t(f_nr[i]);
    const uint64_t q0 = (w - piece_off[j]) * PIECE;
    const uint64_t q1 = uint64_t(len - from) < q0 + PIECE ? uint64_t(len - from) : q0 + PIECE;
    if (len < from || s_boff[j] + (len - from) > total_tig) {
        if (lane == 0) atomicOr(err, E_INTERNAL);
        return;
    }
    uint8_t* dst = tig + s_boff[j];
    for (uint64_t q = q0 + lane; q < q1; q += 64) {
        const uint64_t t = from + q;
        dst[q] = rev ? comp_of(seq[len - 1 - t]) : seq[t];
    }
}

// Unitig bases, in emit order, from the tig sequences.
__global__ __launch_bounds__(ST_TPB) void k_st_copy_out(uint64_t n_pieces, const uint32_t* __restrict__ item_of,
                                                        const uint64_t* __restrict__ piece_off, uint64_t nu,
                                                        const uint32_t* __restrict__ ord,
                                                        const uint64_t* __restrict__ u_x,
                                                        const uint32_t* __restrict__ u_tig,
                                                        const uint64_t* __restrict__ ooff, uint64_t total_out,
                                                        uint32_t k, const uint8_t* __restrict__ tig,
                                                        uint64_t total_tig, uint8_t* __restrict__ out,
                                                        uint32_t* __restrict__ err) {
    const uint64_t w = (uint64_t(blockIdx.x) * ST_TPB + threadIdx.x) >> 6;
    if (w >= n_pieces) return;
    const int lane = threadIdx.x & 63;
    const uint32_t t = item_of[w];
    const uint32_t u = ord[t];
    const uint64_t from = u_x[u] + uint64_t(k - 1) * u_tig[u];  // the tig base of the unitig's first k-mer
    const uint64_t o0 = ooff[t];
    const uint64_t len = (t + 1 < nu ? ooff[t + 1] : total_out) - o0;
    if (from > total_tig || len > total_tig - from || o0 + len > total_out) {
        if (lane == 0) atomicOr(err, E_INTERNAL);
        return;
    }
    const uint8_t* src = tig + from;
    const uint64_t q0 = (w - piece_off[t]) * PIECE;
    const uint64_t q1 = len < q0 + PIECE ? len : q0 + PIECE;
    for (uint64_t q = q0 + lane; q < q1; q += 64) out[o0 + q] = src[q];
}

// The slots (tig sequences) and unitigs (output) of SMALL_COPY bytes or
// fewer, one thread each.
__global__ __launch_bounds__(ST_TPB) void k_st_fill_tigs_small(uint64_t n_lin, const uint32_t* __restrict__ slot_side,
                                                               const uint32_t* __restrict__ ch_pos,
                                                               const uint8_t* __restrict__ buf,
                                                               const uint64_t* __restrict__ off,
                                                               const uint32_t* __restrict__ f_nr,
                                                               const uint32_t* __restrict__ f_len,
                                                               const uint64_t* __restrict__ s_boff, uint32_t k,
                                                               uint64_t total_tig, uint8_t* __restrict__ tig,
                                                               uint32_t* __restrict__ err) {
    const uint64_t j = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (j >= n_lin) return;
    const uint32_t s = slot_side[j], i = s >> 1;
    const uint32_t len = f_len[i];
    const uint32_t from = ch_pos[i] > 0 ? k : 0u;
    if (len < from || s_boff[j] + (len - from) > total_tig) {
        atomicOr(err, E_INTERNAL);
        return;
    }
    const uint32_t b = len - from;
    if (b > SMALL_COPY) return;
    const uint8_t* seq = buf + off[i] + 9 + 12 * uint64_t(f_nr[i]);
    uint8_t* dst = tig + s_boff[j];
    if (s & 1u) {
        for (uint32_t q = 0; q < b; ++q) dst[q] = comp_of(seq[len - 1 - from - q]);
    } else {
        for (uint32_t q = 0; q < b; ++q) dst[q] = seq[from + q];
    }
}

__global__ __launch_bounds__(ST_TPB) void k_st_copy_out_small(uint64_t nu, const uint32_t* __restrict__ ord,
                                                              const uint64_t* __restrict__ u_x,
                                                              const uint32_t* __restrict__ u_tig,
                                                              const uint64_t* __restrict__ ooff, uint64_t total_out,
                                                              uint32_t k, const uint8_t* __restrict__ tig,
                                                              uint64_t total_tig, uint8_t* __restrict__ out,
                                                              uint32_t* __restrict__ err) {
    const uint64_t t = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (t >= nu) return;
    const uint32_t u = ord[t];
    const uint64_t from = u_x[u] + uint64_t(k - 1) * u_tig[u];
    const uint64_t o0 = ooff[t];
    const uint64_t len = (t + 1 < nu ? ooff[t + 1] : total_out) - o0;
    if (from > total_tig || len > total_tig - from || o0 + len > total_out) {
        atomicOr(err, E_INTERNAL);
        return;
    }
    if (len > SMALL_COPY) return;
    for (uint32_t q = 0; q < uint32_t(len); ++q) out[o0 + q] = tig[from + q];
}

__global__ void k_st_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(ST_TPB) void k_st_gather_n(uint64_t n, const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ e_n, uint64_t* __restrict__ g) {
    const uint64_t q = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (q < n) g[q] = e_n[list[q]];
}

// The partners of both sides of every fragment on a cycle.
__global__ __launch_bounds__(ST_TPB) void k_st_cycle_partners(uint64_t nc, const uint32_t* __restrict__ list,
                                                              const uint32_t* __restrict__ partner,
                                                              uint32_t* __restrict__ out) {
    const uint64_t q = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (q >= nc) return;
    const uint64_t i = list[q];
    out[2 * q] = partner[2 * i];
    out[2 * q + 1] = partner[2 * i + 1];
}

template <class T>
T st_d2h(const void* d, hipStream_t st) {
    T v;
    CDBG_HIP(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    return v;
}

void throw_on(uint32_t e) {
    if (e & E_SHORT) throw Error("stitch: a fragment is shorter than k");
    if (e & E_COVER) throw Error("stitch: colour runs do not cover a fragment's k-mers");
    if (e & E_BASE) throw Error("stitch: non-ACGT base in a fragment sequence");
    if (e & E_GROUP) throw Error("stitch: more than two open fragment ends share a boundary k-mer");
    if (e & E_REENTER) throw Error("stitch: a chain re-enters a fragment");
    if (e & E_FOREIGN) throw Error("stitch: a foreign k-mer inside a tig (its owner never joined)");
    if (e) throw Error("stitch: inconsistent fragment chains on the device");
}

int st_bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

// Work items of PIECE bytes over elements whose byte offsets are `scan`
// (exclusive, `total` in all): item_of[p] = element of piece p, piece_off[e]
// = first piece of element e.  Returns the number of pieces.
uint64_t make_pieces(uint64_t n, const uint64_t* scan, uint64_t total, DevBuf& np, DevBuf& poff, DevBuf& items,
                     DevBuf& tot, DevBuf& tmp, hipStream_t st) {
    uint64_t* d_np = np.ensure<uint64_t>(n);
    uint64_t* d_po = poff.ensure<uint64_t>(n);
    uint64_t* d_tot = tot.ensure<uint64_t>(1);
    { KSCOPE("k_st_piece_count", st); k_st_piece_count<<<st_blocks(n), ST_TPB, 0, st>>>(n, scan, total, d_np); }
    exclusive_scan<uint64_t, uint64_t>(d_np, d_po, n, d_tot, tmp, st);
    const uint64_t P = st_d2h<uint64_t>(d_tot, st);
    if (P >= (1ull << 32) || P * 64 / ST_TPB >= (1ull << 31)) throw Error("stitch: too many copy pieces");
    uint32_t* d_items = items.ensure<uint32_t>(P + 1);
    { KSCOPE("k_st_pieces", st); k_st_pieces<<<st_blocks(n), ST_TPB, 0, st>>>(n, scan, total, d_po, d_items); }
    return P;
}

// Record starts on the device.  The stream is cut into PARSE_CHUNK-byte
// chunks; a thread per chunk guesses the chunk's first record (an open-flag
// byte after a base, opening four records that fit with ACGT at both ends
// of their sequences; chunk 0 starts at 0) and follows the chain of records
// past the chunk's end.  The host then follows the true chain chunk by
// chunk: a chunk whose guess is where the chain enters it keeps its count,
// any other is walked again from the entry (one thread), and a second pass
// writes every chunk's starts at its place.  Same starts and the same error
// as one sequential parse.
constexpr uint64_t PARSE_CHUNK = uint64_t(64) << 10;
constexpr uint64_t NO_START = ~0ull;

__device__ __forceinline__ bool d_is_base(uint8_t c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

// size of the record at p, 0 when it does not fit in [p, n)
__device__ __forceinline__ uint64_t d_record_size(const uint8_t* __restrict__ buf, uint64_t n, uint64_t p) {
    if (p >= n || n - p < 9) return 0;
    const uint64_t nr = ld32(buf + p + 1);
    if (n - p < 9 + 12 * nr) return 0;
    const uint64_t len = ld32(buf + p + 5 + 12 * nr);
    if (n - p - 9 - 12 * nr < len) return 0;
    return 9 + 12 * nr + len;
}

__device__ bool d_plausible(const uint8_t* __restrict__ buf, uint64_t n, uint64_t p) {
    if (p == 0 || p >= n || buf[p] > 3 || !d_is_base(buf[p - 1])) return false;
    for (int r = 0; r < 4 && p < n; ++r) {
        if (buf[p] > 3) return false;
        const uint64_t s = d_record_size(buf, n, p);
        if (!s) return false;
        const uint64_t seq = p + 9 + 12 * uint64_t(ld32(buf + p + 1));
        if (seq < p + s && (!d_is_base(buf[seq]) || !d_is_base(buf[p + s - 1]))) return false;
        p += s;
    }
    return true;
}

// chain from `from` while it is below hi: records, where it leaves, whether
// a record did not fit
__device__ void d_walk(const uint8_t* __restrict__ buf, uint64_t n, uint64_t from, uint64_t hi, uint64_t& cnt,
                       uint64_t& exitp, uint8_t& bad) {
    uint64_t p = from, m = 0;
    bad = 0;
    while (p < hi) {
        const uint64_t s = d_record_size(buf, n, p);
        if (!s) {
            bad = 1;
            break;
        }
        ++m;
        p += s;
    }
    cnt = m;
    exitp = p;
}

__global__ __launch_bounds__(ST_TPB) void k_st_parse_guess(uint64_t nc, const uint8_t* __restrict__ buf, uint64_t n,
                                                           uint64_t* __restrict__ guess, uint64_t* __restrict__ cnt,
                                                           uint64_t* __restrict__ exitp, uint8_t* __restrict__ bad) {
    const uint64_t c = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (c >= nc) return;
    const uint64_t lo = c * PARSE_CHUNK, hi = n - lo < PARSE_CHUNK ? n : lo + PARSE_CHUNK;
    uint64_t p = lo;
    if (c > 0)
        while (p < hi && !d_plausible(buf, n, p)) ++p;
    if (p >= hi) {
        guess[c] = NO_START;
        cnt[c] = 0;
        exitp[c] = 0;
        bad[c] = 0;
        return;
    }
    guess[c] = p;
    uint64_t m, e;
    uint8_t b;
    d_walk(buf, n, p, hi, m, e, b);
    cnt[c] = m;
    exitp[c] = e;
    bad[c] = b;
}

// one chunk walked again from where the true chain enters it
__global__ void k_st_parse_walk(const uint8_t* __restrict__ buf, uint64_t n, uint64_t from, uint64_t hi,
                                uint64_t* __restrict__ out) {
    if (threadIdx.x || blockIdx.x) return;
    uint64_t m, e;
    uint8_t b;
    d_walk(buf, n, from, hi, m, e, b);
    out[0] = m;
    out[1] = e;
    out[2] = b;
}

__global__ __launch_bounds__(ST_TPB) void k_st_parse_write(uint64_t nc, const uint8_t* __restrict__ buf, uint64_t n,
                                                           const uint64_t* __restrict__ entry,
                                                           const uint64_t* __restrict__ base, uint64_t nf,
                                                           uint64_t* __restrict__ off, uint32_t* __restrict__ err) {
    const uint64_t c = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (c >= nc || entry[c] == NO_START) return;
    const uint64_t lo = c * PARSE_CHUNK, hi = n - lo < PARSE_CHUNK ? n : lo + PARSE_CHUNK;
    uint64_t p = entry[c], j = base[c];
    while (p < hi) {
        const uint64_t s = d_record_size(buf, n, p);
        if (!s || j >= nf) {
            atomicOr(err, E_INTERNAL);
            return;
        }
        off[j++] = p;
        p += s;
    }
}

// Records of the fragments on cycles, gathered into one block (a wave each).
__global__ __launch_bounds__(ST_TPB) void k_st_gather_records(uint64_t nc, const uint32_t* __restrict__ list,
                                                              const uint64_t* __restrict__ off, uint64_t nf, uint64_t n,
                                                              const uint64_t* __restrict__ dst_off,
                                                              const uint8_t* __restrict__ buf,
                                                              uint8_t* __restrict__ dst) {
    const uint64_t w = (uint64_t(blockIdx.x) * ST_TPB + threadIdx.x) >> 6;
    if (w >= nc) return;
    const uint32_t f = list[w];
    const uint64_t a = off[f], b = f + 1 < nf ? off[f + 1] : n;
    for (uint64_t q = (threadIdx.x & 63); q < b - a; q += 64) dst[dst_off[w] + q] = buf[a + q];
}

__global__ __launch_bounds__(ST_TPB) void k_st_record_sizes(uint64_t nc, const uint32_t* __restrict__ list,
                                                            const uint64_t* __restrict__ off, uint64_t nf, uint64_t n,
                                                            uint64_t* __restrict__ size) {
    const uint64_t l = uint64_t(blockIdx.x) * ST_TPB + threadIdx.x;
    if (l >= nc) return;
    const uint32_t f = list[l];
    size[l] = (f + 1 < nf ? off[f + 1] : n) - off[f];
}


// Record starts of the device stream buf[0, n) into b_off; returns their
// number.
uint64_t device_offsets(const uint8_t* buf, uint64_t n, DevBuf& b_off, uint32_t* err, hipStream_t st) {
    if (n == 0) return 0;
    const uint64_t nc = (n + PARSE_CHUNK - 1) / PARSE_CHUNK;
    DevBuf b_g, b_c, b_e, b_b, b_w, b_entry, b_base;
    uint64_t* g = b_g.ensure<uint64_t>(nc);
    uint64_t* c = b_c.ensure<uint64_t>(nc);
    uint64_t* e = b_e.ensure<uint64_t>(nc);
    uint8_t* bad = b_b.ensure<uint8_t>(nc);
    { KSCOPE("k_st_parse_guess", st); k_st_parse_guess<<<st_blocks(nc), ST_TPB, 0, st>>>(nc, buf, n, g, c, e, bad); }
    std::vector<uint64_t> hg(nc), hc(nc), he(nc), entry(nc, NO_START), base(nc, 0);
    std::vector<uint8_t> hb(nc);
    CDBG_HIP(hipMemcpyAsync(hg.data(), g, nc * 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipMemcpyAsync(hc.data(), c, nc * 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipMemcpyAsync(he.data(), e, nc * 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipMemcpyAsync(hb.data(), bad, nc, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    // the true chain, chunk by chunk
    uint64_t at = 0, nf = 0;
    uint64_t* w = b_w.ensure<uint64_t>(3);
    for (uint64_t ch = 0; ch < nc; ++ch) {
        const uint64_t lo = ch * PARSE_CHUNK, hi = std::min(n, lo + PARSE_CHUNK);
        base[ch] = nf;
        if (at >= hi) continue;  // a record spans the whole chunk
        uint64_t m, ex, bd;
        if (hg[ch] == at && !hb[ch]) {
            m = hc[ch];
            ex = he[ch];
            bd = 0;
        } else {
            { KSCOPE("k_st_parse_walk", st); k_st_parse_walk<<<1, 64, 0, st>>>(buf, n, at, hi, w); }
            uint64_t r[3];
            CDBG_HIP(hipMemcpyAsync(r, w, 24, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipStreamSynchronize(st));
            m = r[0];
            ex = r[1];
            bd = r[2];
        }
        if (bd || ex < hi) throw Error("stitch: truncated fragment record");
        entry[ch] = at;
        nf += m;
        at = ex;
    }
    if (at != n) throw Error("stitch: truncated fragment record");
    if (nf >= (1ull << 30)) throw Error("stitch: 2^30 or more fragments for the device stitch");
    uint64_t* d_entry = b_entry.ensure<uint64_t>(nc);
    uint64_t* d_base = b_base.ensure<uint64_t>(nc);
    CDBG_HIP(hipMemcpyAsync(d_entry, entry.data(), nc * 8, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipMemcpyAsync(d_base, base.data(), nc * 8, hipMemcpyHostToDevice, st));
    uint64_t* off = b_off.ensure<uint64_t>(nf);
    { KSCOPE("k_st_parse_write", st); k_st_parse_write<<<st_blocks(nc), ST_TPB, 0, st>>>(nc, buf, n, d_entry, d_base, nf, off, err); }
    CDBG_HIP(hipStreamSynchronize(st));  // the host tables are local
    return nf;
}

// The stitch of the fragment stream buf[0, n) already on the device.
template <class K>
Unitigs* stitch_device(hipStream_t st, const uint8_t* buf, uint64_t n, uint32_t k, uint32_t threads) {
    std::unique_ptr<Unitigs> U(new Unitigs());
    U->k = k;
    U->threads = threads;
    U->lists.resize(2);  // the linear tigs' unitigs, then the cycles'
    cdbg_stitch_stats& S = U->stats;
    DevBuf b_off, b_len, b_nr, b_cid, b_ekey, b_eid, b_ekey_t, b_eid_t, b_ctr, b_err, b_partner, b_A;
    DevBuf b_chs, b_chst, b_chp, b_clen, b_coff, b_cyc, b_tot, b_tmp, b_cnt, b_slot;
    RadixScratch rs;
    uint32_t* err = b_err.ensure<uint32_t>(1);
    CDBG_HIP(hipMemsetAsync(err, 0, 4, st));
    const uint64_t nf = device_offsets(buf, n, b_off, err, st);
    throw_on(st_d2h<uint32_t>(err, st));
    S.num_fragments = nf;
    if (nf == 0) return U.release();
    const uint64_t ns = 2 * nf;
    const uint64_t* d_off = b_off.as<uint64_t>();
    // counters (N_STRIPE words each): 1 joins, 2 unmatched ends, 3 foreign
    // k-mers dropped, 4 unitig k-mers, 5 open sides of a jump round, 6 largest id
    unsigned long long* ctr = b_ctr.ensure<unsigned long long>(8 * N_STRIPE);
    CDBG_HIP(hipMemsetAsync(ctr, 0, 8 * N_STRIPE * sizeof(unsigned long long), st));
    auto stripes = [&](int c, bool max) {
        std::vector<unsigned long long> h(N_STRIPE);
        CDBG_HIP(hipMemcpyAsync(h.data(), ctr + uint64_t(c) * N_STRIPE, N_STRIPE * 8, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
        uint64_t v = 0;
        for (unsigned long long x : h) v = max ? std::max<uint64_t>(v, x) : v + x;
        return v;
    };
    auto check = [&]() { throw_on(st_d2h<uint32_t>(err, st)); };

    // 1-2. keys of the open sides, sorted; equal keys pair up
    uint32_t* f_len = b_len.ensure<uint32_t>(nf);
    uint32_t* f_nr = b_nr.ensure<uint32_t>(nf);
    uint64_t* f_cid = b_cid.ensure<uint64_t>(4 * nf);
    DevBuf b_fpos;
    uint2* f_pos = b_fpos.ensure<uint2>(nf);
    K* ekey = b_ekey.ensure<K>(ns);
    uint32_t* eid = b_eid.ensure<uint32_t>(ns);
    const uint64_t nbf = st_blocks(nf);
    DevBuf b_bc, b_bb, b_bt;
    uint64_t* block_cnt = b_bc.ensure<uint64_t>(nbf);
    uint64_t* block_base = b_bb.ensure<uint64_t>(nbf);
    uint64_t* d_ne = b_bt.ensure<uint64_t>(1);
    { KSCOPE("k_st_count_ends", st); k_st_count_ends<<<uint32_t(nbf), ST_TPB, 0, st>>>(nf, buf, d_off, block_cnt); }
    exclusive_scan<uint64_t, uint64_t>(block_cnt, block_base, nbf, d_ne, b_tmp, st);
    { KSCOPE("k_st_frags", st); k_st_frags<K><<<uint32_t(nbf), ST_TPB, 0, st>>>(nf, buf, d_off, k, block_base, f_len, f_nr, f_cid, f_pos, ekey, eid, err); }
    check();
    const uint64_t ne = st_d2h<uint64_t>(d_ne, st);
    if (ne > ns) throw_on(E_INTERNAL);
    S.open_ends = ne;
    uint32_t* partner = b_partner.ensure<uint32_t>(ns);
    { KSCOPE("k_st_fill32", st); k_st_fill32<<<st_blocks(ns), ST_TPB, 0, st>>>(partner, ns, NONE32); }
    if (ne) {
        SortPlan plan;
        plan.build({0, ne});
        segmented_radix_sort<K, uint32_t>(ekey, eid, b_ekey_t.ensure<K>(ne), b_eid_t.ensure<uint32_t>(ne), plan, 0,
                                          int(2 * k), 8, rs, st, "stitch_ends");
        { KSCOPE("k_st_pair", st); k_st_pair<K><<<st_blocks(ne), ST_TPB, 0, st>>>(ne, ekey, eid, partner, ctr, err); }
        check();
    }
    b_ekey_t.release();
    b_eid_t.release();

    // 3. list ranking over next(s) = partner[s ^ 1]
    uint64_t* A = b_A.ensure<uint64_t>(ns);
    { KSCOPE("k_st_jump_init", st); k_st_jump_init<<<st_blocks(ns), ST_TPB, 0, st>>>(ns, partner, A); }
    uint64_t prev = ~0ull;
    for (int r = 0; r < 64; ++r) {
        CDBG_HIP(hipMemsetAsync(ctr + 5 * N_STRIPE, 0, N_STRIPE * sizeof(unsigned long long), st));
        { KSCOPE("k_st_jump", st); k_st_jump<<<st_blocks(ns), ST_TPB, 0, st>>>(ns, A, ctr + 5 * N_STRIPE); }
        const uint64_t open = stripes(5, false);
        // a round that resolves no side leaves only sides on cycles
        if (open == 0 || open == prev) break;
        prev = open;
    }

    // 4. walked direction, chain starts and lengths, slots in tig order
    uint32_t* ch_side = b_chs.ensure<uint32_t>(nf);
    uint32_t* ch_start = b_chst.ensure<uint32_t>(nf);
    uint32_t* ch_pos = b_chp.ensure<uint32_t>(nf);
    uint32_t* chain_len = b_clen.ensure<uint32_t>(ns);
    uint8_t* on_cycle = b_cyc.ensure<uint8_t>(nf);
    { KSCOPE("k_st_fill32", st); k_st_fill32<<<st_blocks(ns), ST_TPB, 0, st>>>(chain_len, ns, 0u); }
    { KSCOPE("k_st_choose", st); k_st_choose<<<st_blocks(nf), ST_TPB, 0, st>>>(nf, A, ch_side, ch_start, ch_pos, chain_len, on_cycle, err); }
    check();
    uint32_t* chain_off = b_coff.ensure<uint32_t>(ns);
    uint32_t* d_tot32 = b_tot.ensure<uint32_t>(2);
    exclusive_scan<uint32_t, uint32_t>(chain_len, chain_off, ns, d_tot32, b_tmp, st);
    const uint64_t n_lin = st_d2h<uint32_t>(d_tot32, st);
    if (n_lin > nf) throw_on(E_INTERNAL);
    b_clen.release();

    // host: the fragments on cycles (numbers, partners, their records)
    std::vector<uint32_t> cyc_frags, cyc_partner;
    std::vector<uint64_t> cyc_rec;
    std::vector<uint8_t> cyc_bytes;
    if (n_lin < nf) {
        DevBuf b_list, b_pp;
        uint32_t* list = b_list.ensure<uint32_t>(nf);
        compact_flags<uint8_t>(on_cycle, nf, list, d_tot32, b_cnt, b_tmp, st);
        const uint64_t nc = st_d2h<uint32_t>(d_tot32, st);
        if (nc != nf - n_lin) throw_on(E_INTERNAL);
        uint32_t* pp = b_pp.ensure<uint32_t>(2 * nc);
        { KSCOPE("k_st_cycle_partners", st); k_st_cycle_partners<<<st_blocks(nc), ST_TPB, 0, st>>>(nc, list, partner, pp); }
        cyc_frags.resize(nc);
        cyc_partner.resize(2 * nc);
        CDBG_HIP(hipMemcpyAsync(cyc_frags.data(), list, nc * 4, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipMemcpyAsync(cyc_partner.data(), pp, 2 * nc * 4, hipMemcpyDeviceToHost, st));
        // their records, gathered into one block for the host walk
        DevBuf b_sz, b_do, b_recs;
        uint64_t* sz = b_sz.ensure<uint64_t>(nc);
        { KSCOPE("k_st_record_sizes", st); k_st_record_sizes<<<st_blocks(nc), ST_TPB, 0, st>>>(nc, list, d_off, nf, n, sz); }
        std::vector<uint64_t> hsz(nc);
        CDBG_HIP(hipMemcpyAsync(hsz.data(), sz, nc * 8, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
        cyc_rec.resize(nc);
        uint64_t total = 0;
        for (uint64_t l = 0; l < nc; ++l) {
            cyc_rec[l] = total;
            total += hsz[l];
        }
        uint64_t* dof = b_do.ensure<uint64_t>(nc);
        CDBG_HIP(hipMemcpyAsync(dof, cyc_rec.data(), nc * 8, hipMemcpyHostToDevice, st));
        uint8_t* recs = b_recs.ensure<uint8_t>(total + 16);
        { KSCOPE("k_st_gather_records", st); k_st_gather_records<<<st_blocks(nc * 64), ST_TPB, 0, st>>>(nc, list, d_off, nf, n, dof, buf, recs); }
        cyc_bytes.resize(total);
        CDBG_HIP(hipMemcpyAsync(cyc_bytes.data(), recs, total, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
    }
    b_partner.release();

    if (n_lin) {
        uint32_t* slot_side = b_slot.ensure<uint32_t>(n_lin);
        { KSCOPE("k_st_fill32", st); k_st_fill32<<<st_blocks(n_lin), ST_TPB, 0, st>>>(slot_side, n_lin, NONE32); }
        { KSCOPE("k_st_place", st); k_st_place<<<st_blocks(nf), ST_TPB, 0, st>>>(nf, ch_side, ch_start, ch_pos, chain_off, n_lin, slot_side, err); }

        // 5. what each slot adds; offsets
        DevBuf b_sk, b_sb, b_sr, b_sf, b_sko, b_sbo, b_sro, b_sfo, b_t64;
        uint64_t* s_kmers = b_sk.ensure<uint64_t>(n_lin);
        uint64_t* s_bases = b_sb.ensure<uint64_t>(n_lin);
        uint64_t* s_runs = b_sr.ensure<uint64_t>(n_lin);
        uint32_t* s_first = b_sf.ensure<uint32_t>(n_lin);
        { KSCOPE("k_st_slot_sizes", st); k_st_slot_sizes<<<st_blocks(n_lin), ST_TPB, 0, st>>>(n_lin, slot_side, ch_pos, A, f_len, f_nr, k, s_kmers, s_bases, s_runs, s_first, err); }
        check();
        uint64_t* t64 = b_t64.ensure<uint64_t>(4);
        uint64_t* s_boff = b_sbo.ensure<uint64_t>(n_lin);
        uint64_t* r_off = b_sro.ensure<uint64_t>(n_lin);
        uint32_t* s_tig = b_sfo.ensure<uint32_t>(n_lin);
        exclusive_scan<uint64_t, uint64_t>(s_bases, s_boff, n_lin, t64 + 0, b_tmp, st);
        exclusive_scan<uint64_t, uint64_t>(s_runs, r_off, n_lin, t64 + 1, b_tmp, st);
        exclusive_scan<uint32_t, uint32_t>(s_first, s_tig, n_lin, d_tot32, b_tmp, st);
        uint64_t h64[2];
        CDBG_HIP(hipMemcpyAsync(h64, t64, 16, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
        const uint64_t total_tig = h64[0], n_ent = h64[1];
        if (n_ent >= (1ull << 32)) throw Error("stitch: 2^32 or more colour runs for the device stitch");
        b_sk.release();
        b_sr.release();

        // run entries in tig order; those with k-mers, their first k-mers
        DevBuf b_ec, b_en, b_et, b_list, b_g, b_x, b_head, b_hlist;
        uint64_t* e_cid = b_ec.ensure<uint64_t>(n_ent);
        uint32_t* e_n = b_en.ensure<uint32_t>(n_ent);
        uint32_t* e_tig = b_et.ensure<uint32_t>(n_ent);
        { KSCOPE("k_st_runs", st); k_st_runs<<<st_blocks(n_ent), ST_TPB, 0, st>>>(n_ent, n_lin, r_off, slot_side, ch_pos, A, buf, d_off, f_nr, f_cid, f_pos, s_tig, e_cid, e_n, e_tig, err); }
        check();
        uint32_t* list = b_list.ensure<uint32_t>(n_ent);
        compact_flags<uint32_t>(e_n, n_ent, list, d_tot32, b_cnt, b_tmp, st);
        const uint64_t nk_e = st_d2h<uint32_t>(d_tot32, st);
        if (nk_e == 0) throw_on(E_INTERNAL);
        uint64_t* g = b_g.ensure<uint64_t>(nk_e);
        uint64_t* e_x = b_x.ensure<uint64_t>(nk_e);
        { KSCOPE("k_st_gather_n", st); k_st_gather_n<<<st_blocks(nk_e), ST_TPB, 0, st>>>(nk_e, list, e_n, g); }
        exclusive_scan<uint64_t, uint64_t>(g, e_x, nk_e, t64 + 2, b_tmp, st);
        const uint64_t total_kmers = st_d2h<uint64_t>(t64 + 2, st);
        b_g.release();
        uint8_t* head = b_head.ensure<uint8_t>(nk_e);
        { KSCOPE("k_st_heads", st); k_st_heads<<<st_blocks(nk_e), ST_TPB, 0, st>>>(nk_e, list, e_cid, e_tig, head); }
        uint32_t* hlist = b_hlist.ensure<uint32_t>(nk_e);
        compact_flags<uint8_t>(head, nk_e, hlist, d_tot32, b_cnt, b_tmp, st);
        const uint64_t n_cand = st_d2h<uint32_t>(d_tot32, st);

        // 6. unitigs of the linear tigs
        DevBuf b_uc, b_ux, b_un, b_ut, b_keep, b_klist;
        uint64_t* u_cid = b_uc.ensure<uint64_t>(n_cand);
        uint64_t* u_x = b_ux.ensure<uint64_t>(n_cand);
        uint64_t* u_n = b_un.ensure<uint64_t>(n_cand);
        uint32_t* u_tig = b_ut.ensure<uint32_t>(n_cand);
        uint8_t* keep = b_keep.ensure<uint8_t>(n_cand);
        { KSCOPE("k_st_unitigs", st); k_st_unitigs<<<st_blocks(n_cand), ST_TPB, 0, st>>>(n_cand, nk_e, hlist, list, e_cid, e_tig, e_x, total_kmers, u_cid, u_x, u_n, u_tig, keep, ctr, err); }
        check();
        b_ec.release();
        b_en.release();
        b_et.release();
        b_list.release();
        b_x.release();
        b_head.release();
        b_hlist.release();
        uint32_t* klist = b_klist.ensure<uint32_t>(n_cand);
        compact_flags<uint8_t>(keep, n_cand, klist, d_tot32, b_cnt, b_tmp, st);
        const uint64_t nu = st_d2h<uint32_t>(d_tot32, st);

        // 7. emit order (stable by colour id), items, bases
        DevBuf b_key, b_val, b_key_t, b_val_t, b_olen, b_ooff, b_items, b_tig, b_out, b_np, b_po, b_it;
        UnitigList& L = U->lists[0];
        if (nu) {
            uint64_t* key = b_key.ensure<uint64_t>(nu);
            uint32_t* ord = b_val.ensure<uint32_t>(nu);
            { KSCOPE("k_st_keys", st); k_st_keys<<<st_blocks(nu), ST_TPB, 0, st>>>(nu, klist, u_cid, key, ord, ctr + 6 * N_STRIPE); }
            const int cbits = st_bits_for(stripes(6, true));
            if (cbits) {
                SortPlan plan;
                plan.build({0, nu});
                segmented_radix_sort<uint64_t, uint32_t>(key, ord, b_key_t.ensure<uint64_t>(nu), b_val_t.ensure<uint32_t>(nu),
                                                         plan, 0, cbits, 8, rs, st, "stitch_cids");
            }
            uint64_t* olen = b_olen.ensure<uint64_t>(nu);
            uint64_t* ooff = b_ooff.ensure<uint64_t>(nu);
            { KSCOPE("k_st_out_len", st); k_st_out_len<<<st_blocks(nu), ST_TPB, 0, st>>>(nu, ord, u_n, k, olen); }
            exclusive_scan<uint64_t, uint64_t>(olen, ooff, nu, t64 + 3, b_tmp, st);
            const uint64_t total_out = st_d2h<uint64_t>(t64 + 3, st);
            uint64_t* items = b_items.ensure<uint64_t>(3 * nu);
            { KSCOPE("k_st_items", st); k_st_items<<<st_blocks(nu), ST_TPB, 0, st>>>(nu, ord, u_cid, u_n, ooff, k, items); }

            // fragments -> tig sequences
            uint8_t* tig = b_tig.ensure<uint8_t>(total_tig + 16);
            DevBuf b_ptot;
            const uint64_t P1 = make_pieces(n_lin, s_boff, total_tig, b_np, b_po, b_it, b_ptot, b_tmp, st);
            { KSCOPE("k_st_fill_tigs_small", st); k_st_fill_tigs_small<<<st_blocks(n_lin), ST_TPB, 0, st>>>(n_lin, slot_side, ch_pos, buf, d_off, f_nr, f_len, s_boff, k, total_tig, tig, err); }
            if (P1) { KSCOPE("k_st_fill_tigs", st); k_st_fill_tigs<<<st_blocks(P1 * 64), ST_TPB, 0, st>>>(P1, b_it.as<uint32_t>(), b_po.as<uint64_t>(), slot_side, ch_pos, buf, d_off, f_nr, f_len, s_boff, k, total_tig, tig, err); }
            // tig sequences -> unitigs in emit order
            uint8_t* out = b_out.ensure<uint8_t>(total_out + 16);
            DevBuf b_np2, b_po2, b_it2;
            const uint64_t P2 = make_pieces(nu, ooff, total_out, b_np2, b_po2, b_it2, b_ptot, b_tmp, st);
            { KSCOPE("k_st_copy_out_small", st); k_st_copy_out_small<<<st_blocks(nu), ST_TPB, 0, st>>>(nu, ord, u_x, u_tig, ooff, total_out, k, tig, total_tig, out, err); }
            if (P2) { KSCOPE("k_st_copy_out", st); k_st_copy_out<<<st_blocks(P2 * 64), ST_TPB, 0, st>>>(P2, b_it2.as<uint32_t>(), b_po2.as<uint64_t>(), nu, ord, u_x, u_tig, ooff, total_out, k, tig, total_tig, out, err); }
            check();
            static_assert(sizeof(UnitigList::Item) == 24, "items are three u64");
            // items, then bases, in one pinned block the list keeps
            void* hp = nullptr;
            CDBG_HIP(hipHostMalloc(&hp, nu * 24 + total_out + 16, hipHostMallocDefault));
            L.keep = std::shared_ptr<void>(hp, [](void* q) { (void)hipHostFree(q); });
            L.ext_items = (const UnitigList::Item*)hp;
            L.ext_n = nu;
            L.ext_bases = (const char*)hp + nu * 24;
            CDBG_HIP(hipMemcpyAsync(hp, items, nu * 24, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipMemcpyAsync((uint8_t*)hp + nu * 24, out, total_out, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipStreamSynchronize(st));  // the block is complete before anything can free it
        }
        S.foreign_dropped = stripes(3, false);
        S.unitig_kmers = stripes(4, false);
        S.num_unitigs = nu;
    }
    S.joins = stripes(1, false);
    S.unmatched_ends = stripes(2, false);
    if (!cyc_frags.empty()) {
        stitch_cycles(cyc_bytes.data(), cyc_bytes.size(), cyc_rec, k, cyc_frags, cyc_partner, U->lists[1], S);
        S.num_unitigs += U->lists[1].size();
    }
    return U.release();
}

}  // namespace

namespace {

// A file goes to the device through two pinned buffers kept for the process
// (pinned pages are mapped for the host on first touch, so fresh ones would
// fault page by page): host threads read a chunk into one while the other's
// copy runs.
constexpr uint64_t FILE_STAGE = uint64_t(256) << 20;

struct FileStagePool {
    std::mutex mu;
    std::vector<uint8_t*> free;
    uint8_t* take() {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (!free.empty()) {
                uint8_t* p = free.back();
                free.pop_back();
                return p;
            }
        }
        void* p = nullptr;
        CDBG_HIP(hipHostMalloc(&p, FILE_STAGE, hipHostMallocDefault));
        return (uint8_t*)p;
    }
    void give(uint8_t* p) {
        std::lock_guard<std::mutex> lk(mu);
        free.push_back(p);
    }
};
FileStagePool g_file_stage;

void upload_file(const char* path, uint64_t n, uint8_t* d, hipStream_t st, uint32_t threads) {
    struct Stage {
        uint8_t* b[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        int fd = -1;
        hipStream_t st = nullptr;
        ~Stage() {
            if (st) (void)hipStreamSynchronize(st);  // no copy still reads a buffer given back
            for (int i = 0; i < 2; ++i) {
                if (ev[i]) (void)hipEventDestroy(ev[i]);
                if (b[i]) g_file_stage.give(b[i]);
            }
            if (fd >= 0) ::close(fd);
        }
    } S;
    S.st = st;
    S.fd = ::open(path, O_RDONLY);
    if (S.fd < 0) throw Error(std::string("cannot open ") + path);
    for (int i = 0; i < 2; ++i) {
        S.b[i] = g_file_stage.take();
        CDBG_HIP(hipEventCreateWithFlags(&S.ev[i], hipEventDisableTiming));
    }
    const uint64_t nchunks = (n + FILE_STAGE - 1) / FILE_STAGE;
    for (uint64_t c = 0; c < nchunks; ++c) {
        const int b = int(c & 1);
        if (c >= 2) CDBG_HIP(hipEventSynchronize(S.ev[b]));  // this buffer's previous copy is done
        const uint64_t at = c * FILE_STAGE, len = std::min<uint64_t>(FILE_STAGE, n - at);
        const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads, len >> 22));
        std::atomic<bool> short_read{false};
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                uint64_t p = len * t / T;
                const uint64_t e = len * (t + 1) / T;
                while (p < e) {
                    const ssize_t got = ::pread(S.fd, S.b[b] + p, size_t(e - p), off_t(at + p));
                    if (got <= 0) {
                        short_read = true;
                        return;
                    }
                    p += uint64_t(got);
                }
            });
        for (auto& x : th) x.join();
        if (short_read) throw Error(std::string("short read of ") + path);
        CDBG_HIP(hipMemcpyAsync(d + at, S.b[b], len, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipEventRecord(S.ev[b], st));
    }
    CDBG_HIP(hipStreamSynchronize(st));
}

// Device, stream and allocation scope of one stitch; `load` puts the stream
// in the device buffer it is given.
template <class F>
Unitigs* run_stitch(int device, uint64_t n, uint32_t k, uint32_t threads, F&& load) {
    if (k < 2 || k > 63) throw Error("k must satisfy 2 <= k <= 63");
    if (threads == 0) threads = 1;
    CDBG_HIP(hipSetDevice(device));
    StreamOwner own;
    own.s = acquire_stream(device);
    own.device = device;
    AllocStream on_stream(own.s);
    DevBuf b_buf;  // released (to the block cache, on own.s) before the stream is given back
    uint8_t* buf = b_buf.ensure<uint8_t>(n + 16);
    load(buf, own.s);
    if (k <= 31) return stitch_device<uint64_t>(own.s, buf, n, k, threads);
    return stitch_device<Key128>(own.s, buf, n, k, threads);
}

}  // namespace

Unitigs* stitch_fragments_gpu(int device, const uint8_t* hbuf, uint64_t n, uint32_t k, uint32_t threads) {
    return run_stitch(device, n, k, threads, [&](uint8_t* buf, hipStream_t st) {
        if (n) CDBG_HIP(hipMemcpyAsync(buf, hbuf, n, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipStreamSynchronize(st));
    });
}

Unitigs* stitch_job_gpu(Job& J, uint32_t threads) {
    if (threads == 0) threads = 1;
    bool on_device = false;
    uint64_t n = 0;
    const uint8_t* src = phase2_frags_where(J, on_device, n);
    if (!on_device)  // phase 2 kept the stream in host memory: it goes up like a caller's buffer
        return stitch_fragments_gpu(J.ctx->device, src, n, J.k, threads);
    // in HBM: stitched where it lies, on the job's stream
    CDBG_HIP(hipSetDevice(J.ctx->device));
    AllocStream on_stream(J.st);
    if (J.k <= 31) return stitch_device<uint64_t>(J.st, src, n, J.k, threads);
    return stitch_device<Key128>(J.st, src, n, J.k, threads);
}

Unitigs* stitch_file_gpu(int device, const char* path, uint32_t k, uint32_t threads) {
    const uint64_t n = file_size_of(path);
    return run_stitch(device, n, k, threads, [&](uint8_t* buf, hipStream_t st) { upload_file(path, n, buf, st, threads); });
}

}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

CDBG_API void* cdbg_gpu_stitch(cdbg_gpu_ctx* ctx, const uint8_t* frags, uint64_t n_bytes, uint32_t k,
                               uint32_t threads) {
    try {
        if (!ctx || (!frags && n_bytes)) throw Error("null argument");
        return stitch_fragments_gpu(ctx->device, frags, n_bytes, k, threads);
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

CDBG_API void* cdbg_gpu_job_stitch(cdbg_gpu_job* job, uint32_t threads) {
    try {
        if (!job) throw Error("null argument");
        return stitch_job_gpu(job->J, threads);
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

CDBG_API void* cdbg_gpu_stitch_file(cdbg_gpu_ctx* ctx, const char* frag_path, uint32_t k, uint32_t threads) {
    try {
        if (!ctx || !frag_path) throw Error("null argument");
        return stitch_file_gpu(ctx->device, frag_path, k, threads);
    } catch (std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

}  // extern "C"
