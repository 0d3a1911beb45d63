// Phase 2 (bucket-process) on the device, over the phase-1 record stream.
//
// Semantics are those of process_buckets/process_bucket at -t 1
// (bucket_walker.hpp:166-566,607-779), reproduced exactly:
//   * a node is every canonical k-mer of a bucket's records; flags = OR of
//     the per-occurrence boundary bits, primary = any owned occurrence
//     (bucket_walker.hpp:212-250);
//   * a node's colour set is the union of the colours of its owned
//     occurrences; its rank is its first occurrence in the bucket stream,
//     which is the reference's ankerl insertion order (:160,387,563);
//   * global colour-set ids are assigned in order of first appearance over
//     (bucket, rank) -- buckets merged 0..B-1, local classes in first-intern
//     order (:713-757, streaming_color_set_dict.hpp:148-182);
//   * fragments are the walks of emit_from_seed with seeds taken in rank
//     order (:431-566), emitted in (bucket, seed rank) order.
//
// Buckets are processed in chunks of consecutive buckets whose k-mer
// occurrences fit the chunk budget, so HBM use is bounded by the chunk, not
// the job.  Because colour-set keys are (bucket << 32 | rank), every set first
// seen in a chunk sorts after all sets of earlier chunks: its global id and
// its place in the colour-set bit stream are final when the chunk ends, as in
// the reference's streaming dictionary.  The singleton table, the multi-set
// hash table (128-bit content hashes, no colour lists), the colour-set stream
// and the fragment stream persist across chunks.
//
// Kernels per chunk (tuple = one k-mer occurrence):
//   k_expand<K>               thread per tuple: canonical k-mer straight from
//                             the 2-bit window, rank|colour|owned|contrib
//   tuple sort                global radix passes on each bucket's top bits,
//                             then LDS cell sorts, which flag the node heads
//   k_tile_heads / k_nodes    node per run of equal keys, numbered per tile:
//                             flags, primary, min rank, owned colours
//   k_set_*                   colour-set interning: singleton table keyed by
//                             colour, content-hash table for larger sets keyed
//                             by 64 bits and checked on the other 64 of a
//                             128-bit content hash (the reference's rule:
//                             dedup on the hash alone); new sets sorted by key,
//                             ids assigned, hybrid encoded
//                             (hybrid_color_sets.hpp:41-96)
//   k_dir_build<K>            per-bucket directory over the sorted node keys
//   k_masks<K> / k_links<K>   8-neighbour ext masks (local_ext_mask :260-271)
//                             and clean links (walk_step :317-336), kept with
//                             the k-mer, rank and colour-set id in one record
//                             per node (one gather per walked position)
//   k_walk_ends/record/frags  path ends walk their paths, paths are recorded
//   (+ k_walk_other/cycles)   end to end, the rank-ordered seed loop runs over
//                             the recorded positions (general loop for hairpin
//                             paths and cycles)
//   fragment sort + k_frag_*  frag_unitigs.bin records (unitig_spill.hpp:491-522)
// Lists of flagged items (path ends, fragments, long runs, new sets) come
// from stream compaction (compact_flags).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <type_traits>
#include <cstdlib>
#include <cstring>

#include "hbm_model.h"
#include "job.h"
#include "kmer_ops.h"

namespace cdbg_amd {

namespace {

constexpr uint32_t NONE32 = 0xFFFFFFFFu;
constexpr uint32_t PORT_L = 0, PORT_R = 1;
constexpr uint8_t BOUND_LEFT = 1, BOUND_RIGHT = 2;
constexpr int TPB = 256;

inline uint32_t blocks_for(uint64_t n, uint32_t t = TPB) {
    uint64_t b = (n + t - 1) / t;
    return uint32_t(b ? b : 1);
}

// Tuple value: rank (32) | colour (29) | owned (1) | contrib (2).
__device__ __forceinline__ uint32_t v_rank(uint64_t v) { return uint32_t(v >> 32); }
__device__ __forceinline__ uint32_t v_color(uint64_t v) { return uint32_t(v >> 3) & 0x1FFFFFFFu; }
__device__ __forceinline__ bool v_owned(uint64_t v) { return (v >> 2) & 1u; }
__device__ __forceinline__ uint32_t v_contrib(uint64_t v) { return uint32_t(v) & 3u; }

// last b in [0, B) with off[b] <= i (off has B+1 entries, off[B] > i)
__device__ __forceinline__ uint32_t bucket_of(const uint64_t* off, uint32_t B, uint64_t i) {
    uint32_t lo = 0, hi = B;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <class T>
__global__ void k_fill(T* p, uint64_t n, T v) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// p[2i] = a, p[2i+1] = b for i < n
template <class T>
__global__ void k_fill2(T* p, uint64_t n, T a, T b) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        p[2 * i] = a;
        p[2 * i + 1] = b;
    }
}

// block-wide sum added atomically to *dst
__device__ __forceinline__ void block_add(unsigned long long* dst, uint64_t v) {
    __shared__ uint64_t s_part[TPB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < int(blockDim.x >> 6); ++w) t += s_part[w];
        if (t) atomicAdd(dst, (unsigned long long)t);
    }
    __syncthreads();  // s_part is reused by the next call
}

// ---------------------------------------------------------------------------
// expand
// ---------------------------------------------------------------------------
__global__ void k_count_kmers(uint64_t n, const uint32_t* __restrict__ len, uint32_t k, uint32_t* __restrict__ cnt) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = len[i] - k + 1;
}

// v[i] += add for i < n
__global__ void k_add_u64(uint64_t* __restrict__ v, uint64_t n, uint64_t add) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] += add;
}

__global__ void k_bucket_tuple_off(uint32_t B, const uint64_t* __restrict__ boff, const uint64_t* __restrict__ tup_off,
                                   uint64_t* __restrict__ btup) {
    const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b <= B) btup[b] = tup_off[boff[b]];  // tup_off has n_supers+1 entries
}

// bases p..p+31 of the 2-bit array, base p in bits 0-1
__device__ __forceinline__ uint64_t window64(const uint32_t* __restrict__ packed, uint64_t p) {
    const uint64_t w = p >> 4;
    const uint32_t sh = 2 * uint32_t(p & 15);
    const uint64_t lo = uint64_t(packed[w]) | (uint64_t(packed[w + 1]) << 32);
    uint64_t v = lo >> sh;
    if (sh) v |= uint64_t(packed[w + 2]) << (64 - sh);
    return v;
}

// forward and reverse-complement words of the k-mer starting at base p
__device__ __forceinline__ void kmer_at(const uint32_t* packed, uint64_t p, uint32_t k, uint64_t& f, uint64_t& r) {
    const uint64_t w = window64(packed, p);
    f = rc_word64(~w) >> (64 - 2 * k);        // bases in reading order
    r = ~w & ((uint64_t(1) << (2 * k)) - 1);  // complemented, reversed order
}
__device__ __forceinline__ void kmer_at(const uint32_t* packed, uint64_t p, uint32_t k, Key128& f, Key128& r) {
    const uint64_t w0 = window64(packed, p), w1 = window64(packed, p + 32);
    f = KOps<Key128>::shr(Key128{rc_word64(~w1), rc_word64(~w0)}, 128 - 2 * k);
    const Key128 m = KOps<Key128>::mask(k);
    r = Key128{~w0 & m.lo, ~w1 & m.hi};
}

struct ExpArgs {
    uint64_t n;       // tuples in the chunk
    uint64_t t0;      // global index of the chunk's first tuple
    uint64_t s0, s1;  // super range of the chunk
    const uint64_t* tup_off;
    const uint64_t* btup;
    const uint32_t* bucket;
    const uint64_t* start;
    const uint32_t* len;
    const uint8_t* flags;
    const uint32_t* color;
    const uint32_t* packed;
    uint32_t k;
    const uint32_t* rank0;  // per record: stream rank of its first k-mer (null: records are the stream)
};

// Expand tiles: EXP_PER tuples per thread, strided by the block size so that
// the stores stay coalesced and each thread has EXP_PER independent lookups
// in flight.
constexpr uint32_t EXP_PER = 4, EXP_TILE = TPB * EXP_PER;

// First super-k-mer of every expand tile: tile b starts at chunk tuple
// b * EXP_TILE; block_super[n_blocks] = s1 - 1 closes the table.
__global__ void k_block_super(uint64_t s0, uint64_t s1, uint64_t t0, uint64_t n_blocks,
                              const uint64_t* __restrict__ tup_off, uint32_t* __restrict__ block_super) {
    const uint64_t s = s0 + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= s1) return;
    const uint64_t lo = tup_off[s] - t0, hi = tup_off[s + 1] - t0;  // local tuples [lo, hi)
    for (uint64_t b = (lo + EXP_TILE - 1) / EXP_TILE; b * EXP_TILE < hi && b < n_blocks; ++b)
        block_super[b] = uint32_t(s - s0);
    if (s == s1 - 1) block_super[n_blocks] = uint32_t(s - s0);
}

template <class K>
__global__ __launch_bounds__(TPB) void k_expand(ExpArgs a, const uint32_t* __restrict__ block_super,
                                                K* __restrict__ keys, uint64_t* __restrict__ vals) {
    using O = KOps<K>;
    __shared__ uint64_t s_off[EXP_TILE + 2];
    const uint64_t T0 = a.t0 + uint64_t(blockIdx.x) * EXP_TILE;
    const uint64_t TE = min(T0 + EXP_TILE, a.t0 + a.n);  // tiles end
    const uint64_t sa = a.s0 + block_super[blockIdx.x];
    const uint32_t ns = block_super[blockIdx.x + 1] - block_super[blockIdx.x] + 1;  // <= EXP_TILE + 1
    // the record of every tuple of the tile: each record marks its first
    // tuple, and an inclusive max-scan carries the mark to the tuples after it
    __shared__ uint32_t s_rec[EXP_TILE];
    __shared__ uint32_t s_wmax[TPB / 64];
    for (uint32_t i = threadIdx.x; i <= ns; i += TPB) s_off[i] = a.tup_off[sa + i];
#pragma unroll
    for (uint32_t j = 0; j < EXP_PER; ++j) s_rec[threadIdx.x * EXP_PER + j] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ns; i += TPB) {
        const uint64_t f = s_off[i] > T0 ? s_off[i] - T0 : 0;  // only the first record starts before the tile
        if (f < EXP_TILE) s_rec[f] = i;
    }
    __syncthreads();
    {
        uint32_t v[EXP_PER];
        uint32_t m = 0;
#pragma unroll
        for (uint32_t j = 0; j < EXP_PER; ++j) {
            m = max(m, s_rec[threadIdx.x * EXP_PER + j]);
            v[j] = m;
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        uint32_t x = m;  // inclusive max over the lanes before, then the waves before
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x = max(x, y);
        }
        if (lane == 63) s_wmax[wave] = x;
        uint32_t before = __shfl_up(x, 1, 64);
        if (lane == 0) before = 0;
        __syncthreads();
        for (int w = 0; w < wave; ++w) before = max(before, s_wmax[w]);
#pragma unroll
        for (uint32_t j = 0; j < EXP_PER; ++j) s_rec[threadIdx.x * EXP_PER + j] = max(v[j], before);
    }
    __syncthreads();
    uint64_t sup[EXP_PER];
    uint32_t ix[EXP_PER];
#pragma unroll
    for (uint32_t j = 0; j < EXP_PER; ++j) {
        const uint64_t t = T0 + threadIdx.x + j * TPB;
        const uint32_t r = s_rec[threadIdx.x + j * TPB];  // s_off[r] <= t < s_off[r + 1]
        sup[j] = sa + r;
        ix[j] = uint32_t(t - s_off[r]);
    }
    uint32_t len[EXP_PER], fl[EXP_PER], col[EXP_PER];
    uint64_t st[EXP_PER], rk[EXP_PER];
#pragma unroll
    for (uint32_t j = 0; j < EXP_PER; ++j) {
        if (T0 + threadIdx.x + j * TPB >= TE) continue;
        const uint64_t s = sup[j];
        len[j] = a.len[s];
        fl[j] = a.flags[s];
        col[j] = a.color[s];
        st[j] = a.start[s];
        rk[j] = a.rank0 ? uint64_t(a.rank0[s]) : a.btup[a.bucket[s]];
    }
#pragma unroll
    for (uint32_t j = 0; j < EXP_PER; ++j) {
        const uint64_t t = T0 + threadIdx.x + j * TPB;
        if (t >= TE) continue;
        const uint32_t nk = len[j] - a.k + 1;
        K f, r;
        kmer_at(a.packed, st[j] + ix[j], a.k, f, r);
        const bool fwd = O::le(f, r);
        uint32_t contrib = 0;
        if (!(fl[j] & 1u) && ix[j] == 0) contrib |= fwd ? BOUND_LEFT : BOUND_RIGHT;
        if (!(fl[j] & 2u) && ix[j] == nk - 1) contrib |= fwd ? BOUND_RIGHT : BOUND_LEFT;
        const bool owned = !((ix[j] == 0 && !(fl[j] & 4u)) || (ix[j] == nk - 1 && !(fl[j] & 8u)));
        const uint32_t rank = a.rank0 ? uint32_t(rk[j]) + ix[j] : uint32_t(t - rk[j]);
        keys[t - a.t0] = O::select(fwd, f, r);
        vals[t - a.t0] = (uint64_t(rank) << 32) | (uint64_t(col[j]) << 3) | (uint64_t(owned) << 2) | contrib;
    }
}

// ---------------------------------------------------------------------------
// Cell sort: the last stage of a stable sort whose top bits are already
// sorted.  A cell (a run of keys equal above end_bit) is sorted by one
// workgroup in LDS: packed words merge-sorted (below), or, where key bits
// and position do not fit one word, LSD passes of 8-bit digits, each ranking
// the items with the per-wave scheme of k_radix_scatter and placing them in
// LDS, keys and their original positions held in registers between passes;
// the values follow once, at the end.  Size classes: a launch of
// THREADS threads and capacity CAP takes the listed cells (or all cells)
// with more than `lo` items; cells above CAP go to the next class's list.
// ---------------------------------------------------------------------------
// PACK (keys whose unsorted bits and a position of log2(CAP) <= 13 bits fit
// one key word): each item is the word (key bits below end_bit) << bits |
// position, distinct within the cell, so sorting the words is the stable
// sort of the keys; they are merge-sorted in LDS instead (the bits above
// end_bit are the cell's own).
constexpr int PACK_IB = 13;
template <class K, class V, uint32_t CAP, int THREADS, bool PACK = false>
inline size_t cell_sort_lds() {
    if (PACK) return size_t(CAP + THREADS) * sizeof(K);  // the words, spread (mslot)
    return size_t(CAP) * (sizeof(K) + sizeof(uint16_t)) + ((THREADS / 64) * 256 + THREADS / 64 + 1) * sizeof(uint32_t);
}
// position bits of a packed item in a class of capacity CAP (at most 13)
template <uint32_t CAP>
constexpr int pos_bits() {
    int b = 0;
    while ((1u << b) < CAP) ++b;
    return b;
}
// whether a class of capacity CAP sorting the bits below end_bit packs its items
template <class K, uint32_t CAP>
inline bool cell_packed(int end_bit) {
    return end_bit + pos_bits<CAP>() <= int(8 * sizeof(K));
}

// Packed words of 64- and 128-bit keys: order by <, selection word by word.
__device__ __forceinline__ bool w_lt(uint64_t a, uint64_t b) { return a < b; }
__device__ __forceinline__ bool w_lt(const Key128& a, const Key128& b) { return a < b; }
__device__ __forceinline__ uint64_t w_sel(bool c, uint64_t a, uint64_t b) { return c ? a : b; }
__device__ __forceinline__ Key128 w_sel(bool c, const Key128& a, const Key128& b) {
    return KOps<Key128>::select(c, a, b);
}
template <class K>
__device__ __forceinline__ K w_max();
template <>
__device__ __forceinline__ uint64_t w_max<uint64_t>() { return ~0ull; }
template <>
__device__ __forceinline__ Key128 w_max<Key128>() { return Key128{~0ull, ~0ull}; }
// bits [0, b) set, b < 128
__device__ __forceinline__ uint64_t w_low(uint64_t, int b) { return b >= 64 ? ~0ull : (uint64_t(1) << b) - 1; }
__device__ __forceinline__ Key128 w_low(const Key128&, int b) {
    return b >= 64 ? Key128{~0ull, (uint64_t(1) << (b - 64)) - 1} : Key128{(uint64_t(1) << b) - 1, 0};
}
template <int PB>
__device__ __forceinline__ uint64_t w_pack(uint64_t key, uint64_t low, uint32_t i) { return ((key & low) << PB) | i; }
template <int PB>
__device__ __forceinline__ Key128 w_pack(const Key128& key, const Key128& low, uint32_t i) {
    Key128 w = KOps<Key128>::shl(Key128{key.lo & low.lo, key.hi & low.hi}, PB);
    w.lo |= i;
    return w;
}
template <int PB>
__device__ __forceinline__ uint64_t w_key_bits(uint64_t w) { return w >> PB; }
template <int PB>
__device__ __forceinline__ Key128 w_key_bits(const Key128& w) { return KOps<Key128>::shr(w, PB); }
template <int PB>
__device__ __forceinline__ uint32_t w_pos(uint64_t w) { return uint32_t(w) & ((1u << PB) - 1); }
template <int PB>
__device__ __forceinline__ uint32_t w_pos(const Key128& w) { return uint32_t(w.lo) & ((1u << PB) - 1); }
__device__ __forceinline__ uint64_t w_or(uint64_t a, uint64_t b) { return a | b; }
__device__ __forceinline__ Key128 w_or(const Key128& a, const Key128& b) { return Key128{a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ uint64_t w_andn(uint64_t a, uint64_t m) { return a & ~m; }
__device__ __forceinline__ Key128 w_andn(const Key128& a, const Key128& m) { return Key128{a.lo & ~m.lo, a.hi & ~m.hi}; }

// LDS slot of word i of a merge-sorted cell: one spare word after every
// `IPT` words, so that a wave whose lanes each touch their own IPT
// consecutive words spreads them over the banks instead of landing on one.
template <uint32_t IPT>
__device__ __forceinline__ uint32_t mslot(uint32_t i) {
    return i + i / IPT;
}

// Sorts the distinct words s[mslot(0 .. count - 1)] in LDS ascending
// (count <= CAP): each thread sorts IPT = CAP / THREADS consecutive words in
// registers by a bitonic network, then runs are merged pairwise, each thread
// producing IPT outputs of a merge after finding where its outputs start by
// a binary search along the merge path.  The tail up to a multiple of IPT is
// padded with ~0 (equal words are only ~0s, and which copy of a value goes
// first does not change the values).  Ends with a barrier.
template <class W, uint32_t CAP, int THREADS>
__device__ __forceinline__ void merge_sort_words(W* s, uint32_t count) {
    constexpr uint32_t IPT = CAP / THREADS;
    static_assert(IPT >= 2 && (IPT & (IPT - 1)) == 0 && IPT * THREADS == CAP, "items per thread");
    const W top = w_max<W>();
    const uint32_t n_pad = (count + IPT - 1) / IPT * IPT;
    for (uint32_t i = count + threadIdx.x; i < n_pad; i += THREADS) s[mslot<IPT>(i)] = top;
    __syncthreads();
    const uint32_t out0 = threadIdx.x * IPT;
    const bool act = out0 < n_pad;
    W* const mine = s + mslot<IPT>(out0);  // this thread's IPT words, contiguous
    W r[IPT];
    if (act) {
#pragma unroll
        for (uint32_t j = 0; j < IPT; ++j) r[j] = mine[j];
#pragma unroll
        for (uint32_t k = 2; k <= IPT; k <<= 1)
#pragma unroll
            for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                for (uint32_t i = 0; i < IPT; ++i) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const W x = r[i], y = r[l];
                        const bool sw = ((i & k) == 0) ? w_lt(y, x) : w_lt(x, y);
                        r[i] = w_sel(sw, y, x);
                        r[l] = w_sel(sw, x, y);
                    }
                }
#pragma unroll
        for (uint32_t j = 0; j < IPT; ++j) mine[j] = r[j];  // this thread's own words only
    }
    __syncthreads();
    for (uint32_t L = IPT; L < n_pad; L <<= 1) {
        if (act) {
            // run A = words [p0, p0 + la), run B = words [p0 + L, p0 + L + lb);
            // this thread writes outputs d .. d + IPT - 1 of their merge
            const uint32_t p0 = out0 & ~(2 * L - 1), d = out0 - p0;
            const uint32_t la = min(L, n_pad - p0);
            const uint32_t lb = n_pad - p0 > L ? min(L, n_pad - p0 - L) : 0u;
            const uint32_t pb = p0 + L;
            uint32_t lo = d > lb ? d - lb : 0u, hi = min(d, la);
            while (lo < hi) {  // the number of A's words among the first d outputs
                const uint32_t mid = (lo + hi) >> 1;
                if (w_lt(s[mslot<IPT>(p0 + mid)], s[mslot<IPT>(pb + d - 1 - mid)]))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            uint32_t ia = lo, ib = d - lo;
            W a = w_sel(ia < la, s[mslot<IPT>(ia < la ? p0 + ia : 0u)], top);
            W b = w_sel(ib < lb, s[mslot<IPT>(ib < lb ? pb + ib : 0u)], top);
#pragma unroll
            for (uint32_t j = 0; j < IPT; ++j) {
                const bool ta = ib >= lb || (ia < la && !w_lt(b, a));
                r[j] = w_sel(ta, a, b);
                ia += ta ? 1u : 0u;
                ib += ta ? 0u : 1u;
                const bool ok = ta ? ia < la : ib < lb;
                const W v = w_sel(ok, s[mslot<IPT>(ok ? (ta ? p0 + ia : pb + ib) : 0u)], top);
                a = w_sel(ta, v, a);
                b = w_sel(ta, b, v);
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (uint32_t j = 0; j < IPT; ++j) mine[j] = r[j];
        }
        __syncthreads();
    }
}

struct CellSortArgs {
    const uint32_t* cell_start;
    const uint32_t* list;  // cells of this class (null: every cell)
    uint32_t lo;           // cells of at most lo keys were sorted by a smaller class
    int end_bit;
    uint32_t* next;        // cells above CAP, for the next class (null: none)
    unsigned long long* n_next;
    uint8_t* heads;
};

// Sorts the item-th cell of the class's list (every thread of the
// workgroup calls it; it returns uniformly).
template <class K, class V, uint32_t CAP, int THREADS, bool PACK>
__device__ __forceinline__ void cell_sort_one(K* __restrict__ keys, V* __restrict__ vals, const CellSortArgs& a,
                                              uint32_t item, uint64_t* sm64) {
    constexpr int WAVES = THREADS / 64;
    constexpr int RMAX = CAP / THREADS;  // rounds per wave at most
    constexpr uint32_t IPT = CAP / THREADS;  // packed: words per thread of the merge sort
    static_assert(CAP <= 65536 && CAP % 2 == 0, "item indices are 16-bit");
    constexpr int PB = pos_bits<CAP>();  // packed: position bits
    static_assert(!PACK || (PB <= PACK_IB && CAP == (1u << PB)), "packed items");
    // keys (packed: words) and their positions in the cell move through LDS;
    // each value is fetched once at the end from its original position
    using E = K;  // what moves per item
    E* s_key = (E*)sm64;
    uint16_t* s_idx = (uint16_t*)(s_key + CAP);
    uint32_t* wcnt = (uint32_t*)(PACK ? (void*)(s_key + CAP) : (void*)(s_idx + CAP));  // [WAVES][256]
    uint32_t* s_tmp = wcnt + WAVES * 256;
    const uint32_t* cell_start = a.cell_start;
    const uint32_t lo = a.lo;
    const int end_bit = a.end_bit;
    uint8_t* heads = a.heads;
    const uint32_t cell = a.list ? a.list[item] : item;
    const uint32_t start = cell_start[cell], count = cell_start[cell + 1] - start;
    if (count == 1 && lo == 0 && threadIdx.x == 0) heads[start] = 1;  // a one-key cell is one node
    if (count <= 1 || count <= lo) return;
    if (count > CAP) {
        if (a.next && threadIdx.x == 0) a.next[atomicAdd(a.n_next, 1ull)] = cell;
        return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t rounds = (count + THREADS - 1) / THREADS;
    const uint32_t wbase = uint32_t(wave) * rounds * 64;  // wave w owns items [wbase, wbase + 64 rounds)
    E key[RMAX];
    uint32_t idx[RMAX];
    uint32_t dr[RMAX];
    K hi{};  // packed: the cell's bits above end_bit
    if constexpr (PACK) {
        // distinct words (the position breaks ties): merge-sorted in LDS,
        // then read in the layout the epilogue uses
        const K low = w_low(K{}, end_bit);  // end_bit + 13 fits the word
        hi = w_andn(keys[start], low);
        for (uint32_t i = threadIdx.x; i < count; i += THREADS) s_key[mslot<IPT>(i)] = w_pack<PB>(keys[start + i], low, i);
        if (end_bit > 0)
            merge_sort_words<K, CAP, THREADS>(s_key, count);
        else
            __syncthreads();
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
            if (uint32_t(r) < rounds && i < count) key[r] = s_key[mslot<IPT>(i)];
        }
    } else {
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
            const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
            if (uint32_t(r) < rounds && i < count) {
                key[r] = keys[start + i];
                idx[r] = i;
            }
        }
        uint32_t* my = wcnt + wave * 256;
        for (int sh = 0; sh < end_bit; sh += 8) {
            const int bits = (end_bit - sh) < 8 ? (end_bit - sh) : 8;
            const uint32_t mask = (1u << bits) - 1;
            for (uint32_t b = threadIdx.x; b < WAVES * 256; b += THREADS) wcnt[b] = 0;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                if (uint32_t(r) >= rounds) break;
                const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
                const bool active = i < count;
                const uint32_t dig = active ? digit_of(key[r], sh, mask) : 0u;
                uint64_t peers = __ballot(active);
                for (int b = 0; b < bits; ++b) {
                    const bool bit = (dig >> b) & 1u;
                    const uint64_t m = __ballot(bit);
                    peers &= bit ? m : ~m;
                }
                const uint32_t rk = __popcll(peers & lt);
                uint32_t before = 0;
                if (active) before = my[dig];
                if (active && rk == 0) my[dig] = before + __popcll(peers);
                dr[r] = (dig << 16) | (before + rk);
            }
            __syncthreads();
            {
                // digit-major, wave-minor bases; thread t owns digits DPT t .. DPT t + DPT - 1
                constexpr uint32_t DPT = THREADS >= 256 ? 1u : 256u / THREADS;
                uint32_t tot = 0;
#pragma unroll
                for (uint32_t j = 0; j < DPT; ++j) {
                    const uint32_t d = threadIdx.x * DPT + j;
                    if (d < 256)
                        for (int w = 0; w < WAVES; ++w) tot += wcnt[w * 256 + d];
                }
                uint32_t all;
                uint32_t run = block_exclusive_scan<uint32_t>(tot, s_tmp, all);
#pragma unroll
                for (uint32_t j = 0; j < DPT; ++j) {
                    const uint32_t d = threadIdx.x * DPT + j;
                    if (d < 256)
                        for (int w = 0; w < WAVES; ++w) {
                            const uint32_t c = wcnt[w * 256 + d];
                            wcnt[w * 256 + d] = run;
                            run += c;
                        }
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                if (uint32_t(r) >= rounds) break;
                const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
                if (i < count) {
                    const uint32_t pos = my[dr[r] >> 16] + (dr[r] & 0xFFFFu);
                    s_key[pos] = key[r];
                    s_idx[pos] = uint16_t(idx[r]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                if (uint32_t(r) >= rounds) break;
                const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
                if (i < count) {
                    key[r] = s_key[i];
                    idx[r] = s_idx[i];
                }
            }
        }
    }
    V val[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
        const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
        if (uint32_t(r) < rounds && i < count) {
            if constexpr (PACK) idx[r] = min(w_pos<PB>(key[r]), count - 1);  // in the cell whatever LDS held
            val[r] = vals[start + idx[r]];
        }
    }
    __syncthreads();  // every value is read before any is overwritten
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
        if (uint32_t(r) >= rounds) break;
        const uint32_t i = wbase + uint32_t(r) * 64 + uint32_t(lane);
        if (i < count) {
            if constexpr (PACK)
                keys[start + i] = w_or(hi, w_key_bits<PB>(key[r]));
            else
                keys[start + i] = key[r];
            vals[start + i] = val[r];
        }
    }
    // node heads of the cell: its first key, and every key unlike the one
    // before it (the last pass left the sorted keys in LDS; a cell is one
    // bucket, so a bucket's first tuple is a head)
    const bool sorted_in_lds = end_bit > 0;
    for (uint32_t i = threadIdx.x; i < count; i += THREADS) {
        bool differs;
        if constexpr (PACK)
            differs = !KOps<K>::eq(w_key_bits<PB>(s_key[mslot<IPT>(i)]), w_key_bits<PB>(s_key[mslot<IPT>(i - (i > 0))]));
        else
            differs = !KOps<K>::eq(s_key[i], s_key[i - (i > 0)]);
        heads[start + i] = (i == 0 || (sorted_in_lds && differs)) ? 1 : 0;
    }
}

// One workgroup per listed cell.
template <class K, class V, uint32_t CAP, int THREADS, bool PACK = false>
__global__ __launch_bounds__(THREADS) void k_cell_sort(K* __restrict__ keys, V* __restrict__ vals, CellSortArgs a,
                                                       uint32_t n_items) {
    extern __shared__ uint64_t sm64[];
    if (blockIdx.x >= n_items) return;
    cell_sort_one<K, V, CAP, THREADS, PACK>(keys, vals, a, blockIdx.x, sm64);
}

// Persistent workgroups: each takes CELL_Q cells at a time from a device
// queue (one atomic per batch) until the class's cells are gone.  The cell
// count is read from device memory (*n32 or *n64, else n_items), so a class
// whose list a previous kernel built needs no host read-back, and the stage
// can be captured in a graph.
constexpr uint32_t CELL_Q = 4;
template <class K, class V, uint32_t CAP, int THREADS, bool PACK = false>
__global__ __launch_bounds__(THREADS) void k_cell_sort_q(K* __restrict__ keys, V* __restrict__ vals, CellSortArgs a,
                                                         uint32_t n_items, const uint32_t* __restrict__ n32,
                                                         const unsigned long long* __restrict__ n64,
                                                         unsigned long long* __restrict__ queue) {
    extern __shared__ uint64_t sm64[];
    __shared__ uint32_t s_base;
    const uint64_t n = n32 ? uint64_t(*n32) : n64 ? uint64_t(*n64) : uint64_t(n_items);
    for (;;) {
        if (threadIdx.x == 0) s_base = uint32_t(atomicAdd(queue, (unsigned long long)CELL_Q));
        __syncthreads();
        const uint32_t base = s_base;
        __syncthreads();  // every thread has read s_base before it is written again
        if (base >= n) break;
        for (uint32_t j = 0; j < CELL_Q && base + j < n; ++j) {
            cell_sort_one<K, V, CAP, THREADS, PACK>(keys, vals, a, base + j, sm64);
            __syncthreads();  // the cell's last LDS reads are done before the next cell's writes
        }
    }
}

// Lets the k_cell_sort kernels of a class use their LDS (once per process).
template <class K, class V, uint32_t CAP, int THREADS, bool PACK = false>
inline void cell_sort_prepare() {
    static const bool done = [] {
        CDBG_HIP(hipFuncSetAttribute((const void*)&k_cell_sort<K, V, CAP, THREADS, PACK>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(cell_sort_lds<K, V, CAP, THREADS, PACK>())));
        CDBG_HIP(hipFuncSetAttribute((const void*)&k_cell_sort_q<K, V, CAP, THREADS, PACK>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(cell_sort_lds<K, V, CAP, THREADS, PACK>())));
        return true;
    }();
    (void)done;
}

// Workgroups of a persistent class that fit the device at once.
template <class K, class V, uint32_t CAP, int THREADS, bool PACK>
inline uint32_t cell_sort_resident() {
    static const uint32_t n = [] {
        int dev = 0, per_cu = 0;
        CDBG_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop{};
        CDBG_HIP(hipGetDeviceProperties(&prop, dev));
        CDBG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)&k_cell_sort_q<K, V, CAP, THREADS, PACK>,
                                                              THREADS, cell_sort_lds<K, V, CAP, THREADS, PACK>()));
        return uint32_t(std::max(1, per_cu) * std::max(1, prop.multiProcessorCount));
    }();
    return n;
}

// the two LDS classes: (capacity, threads) per key width
template <class K>
struct CellClasses {
    static constexpr uint32_t TINY = 1024, SMALL = 2048, LARGE = 8192;
};
template <>
struct CellClasses<Key128> {
    static constexpr uint32_t TINY = 512, SMALL = 1024, LARGE = 4096;
};

// One size class of the cell sort over `grid` cells; packed items when the
// key is 64-bit and its unsorted bits leave room for the position.
// Sets up a class's kernels outside any graph capture.
template <class K, uint32_t CAP, int THREADS>
static void cell_sort_warm(int end_bit) {
    if (cell_packed<K, CAP>(end_bit)) {
        cell_sort_prepare<K, uint64_t, CAP, THREADS, true>();
        (void)cell_sort_resident<K, uint64_t, CAP, THREADS, true>();
        return;
    }
    cell_sort_prepare<K, uint64_t, CAP, THREADS, false>();
    (void)cell_sort_resident<K, uint64_t, CAP, THREADS, false>();
}

// With `queue` the class runs as persistent workgroups over a device cell
// count (n32 / n64, else n_items), otherwise one workgroup per cell.
template <class K, uint32_t CAP, int THREADS, bool PACK>
static void launch_cell_class(hipStream_t st, K* keys, uint64_t* vals, const CellSortArgs& a, uint32_t n_items,
                              const uint32_t* n32, const unsigned long long* n64, unsigned long long* queue) {
    cell_sort_prepare<K, uint64_t, CAP, THREADS, PACK>();
    const size_t lds = cell_sort_lds<K, uint64_t, CAP, THREADS, PACK>();
    if (queue) {
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(cell_sort_resident<K, uint64_t, CAP, THREADS, PACK>(),
                                                                       (n_items + CELL_Q - 1) / CELL_Q));
        k_cell_sort_q<K, uint64_t, CAP, THREADS, PACK><<<grid, THREADS, lds, st>>>(keys, vals, a, n_items, n32, n64, queue);
    } else if (n_items) {
        k_cell_sort<K, uint64_t, CAP, THREADS, PACK><<<n_items, THREADS, lds, st>>>(keys, vals, a, n_items);
    }
}

template <class K, uint32_t CAP, int THREADS>
static void launch_cell_sort(hipStream_t st, K* keys, uint64_t* vals, const CellSortArgs& a, uint32_t n_items,
                             const uint32_t* n32 = nullptr, const unsigned long long* n64 = nullptr,
                             unsigned long long* queue = nullptr) {
    if (cell_packed<K, CAP>(a.end_bit)) {
        launch_cell_class<K, CAP, THREADS, true>(st, keys, vals, a, n_items, n32, n64, queue);
        return;
    }
    launch_cell_class<K, CAP, THREADS, false>(st, keys, vals, a, n_items, n32, n64, queue);
}

// flag[c] = cell c holds more than cap keys
__global__ void k_cell_over(uint32_t n_cells, const uint32_t* __restrict__ cell_start, uint32_t cap,
                            uint32_t* __restrict__ flag) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < n_cells) flag[c] = cell_start[c + 1] - cell_start[c] > cap ? 1u : 0u;
}

// Cells are dense: local bucket lb owns cells cell_base[lb] .. cell_base[lb+1]
// - 1, one per value of its top D_lb key bits (2^D_lb of them), and
// cell_start[c] is the first tuple of the bucket whose top D_lb bits are >=
// the cell's value (binary search over the bucket's sorted keys); empty cells
// have equal starts.  cell_start[n_cells] = n_tuples.
template <class K>
__global__ void k_cell_starts(uint32_t n_cells, uint32_t nb, const uint32_t* __restrict__ cell_base,
                              const uint8_t* __restrict__ dbits, const uint64_t* __restrict__ lbtup,
                              uint64_t n_tuples, const K* __restrict__ keys, int kb,
                              uint32_t* __restrict__ cell_start) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > n_cells) return;
    if (c == n_cells) {
        cell_start[c] = uint32_t(n_tuples);
        return;
    }
    uint32_t l = 0, h = nb;  // cell_base[l] <= c < cell_base[h]
    while (h - l > 1) {
        const uint32_t mid = (l + h) >> 1;
        if (cell_base[mid] <= c)
            l = mid;
        else
            h = mid;
    }
    const uint32_t v = uint32_t(c) - cell_base[l], D = dbits[l];
    uint64_t lo = lbtup[l], hi = lbtup[l + 1];
    if (v == 0) hi = lo;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (digit_of(keys[mid], kb - int(D), 0xFFFFFFFFu) < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    cell_start[c] = uint32_t(lo);
}

// Node heads of listed cells from their sorted compact copy: compact index j
// of cell c maps to start[c] + j - pre[c].
template <class K>
__global__ void k_cells_heads(const uint32_t* __restrict__ start, const uint64_t* __restrict__ pre, uint32_t n_list,
                              const K* __restrict__ ck, uint8_t* __restrict__ heads) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= pre[n_list]) return;
    uint32_t lo = 0, hi = n_list;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= j)
            lo = mid;
        else
            hi = mid;
    }
    heads[start[lo] + (j - pre[lo])] = (j == pre[lo] || !KOps<K>::eq(ck[j], ck[j - 1])) ? 1 : 0;
}

// Moves the elements of listed cells between the tuple arrays and a compact
// buffer (to_compact) or back: compact index j lies in cell c with
// pre[c] <= j < pre[c + 1] and maps to start[c] + j - pre[c].
template <class K, class V>
__global__ void k_cells_move(bool to_compact, const uint32_t* __restrict__ start, const uint64_t* __restrict__ pre,
                             uint32_t n_list, K* __restrict__ keys, V* __restrict__ vals, K* __restrict__ ck,
                             V* __restrict__ cv) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= pre[n_list]) return;
    uint32_t lo = 0, hi = n_list;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= j)
            lo = mid;
        else
            hi = mid;
    }
    const uint64_t i = start[lo] + (j - pre[lo]);
    if (to_compact) {
        ck[j] = keys[i];
        cv[j] = vals[i];
    } else {
        keys[i] = ck[j];
        vals[i] = cv[j];
    }
}

// ---------------------------------------------------------------------------
// nodes
// ---------------------------------------------------------------------------
// Heads (node starts) of every NT_TILE-tuple tile of k_nodes.
constexpr uint32_t NT_TILE = 1024;
__global__ __launch_bounds__(TPB) void k_tile_heads(uint64_t nt, const uint8_t* __restrict__ heads,
                                                    uint32_t* __restrict__ tile_cnt) {
    static_assert(NT_TILE == 4 * TPB, "one 4-byte load per thread");
    __shared__ uint32_t s_tmp[TPB / 64 + 1];
    const uint64_t t0 = uint64_t(blockIdx.x) * NT_TILE + 4ull * threadIdx.x;
    uint32_t c = 0;
    if (t0 + 4 <= nt)
        c = nonzero_bytes(*reinterpret_cast<const uint32_t*>(heads + t0));
    else
        for (uint64_t t = t0; t < nt; ++t) c += heads[t] != 0;
    uint32_t tot;
    (void)block_exclusive_scan<uint32_t>(c, s_tmp, tot);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// First node of every bucket: the nodes that start before its first tuple
// (a bucket's first tuple is a head), by binary search over the node starts.
__global__ void k_node_off(uint32_t nb, const uint64_t* __restrict__ lbtup, const uint32_t* __restrict__ node_start,
                           uint32_t n_nodes, uint64_t* __restrict__ node_off) {
    const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b > nb) return;
    const uint64_t t = lbtup[b];
    uint32_t lo = 0, hi = n_nodes;  // first node starting at or after t
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (node_start[mid] < t)
            lo = mid + 1;
        else
            hi = mid;
    }
    node_off[b] = lo;
}

// Nodes with more than BIG_RUN occurrences are reduced by a whole wave.
constexpr uint32_t BIG_RUN = 32;
// counters: 0 primary nodes, 1 multi-colour primary nodes, 2 total integers,
// 3 nodes placed by the walk, 4 fragments, 5 walk errors, 6 colour-set
// collisions, 7 colours of multi-colour nodes, 8 big nodes, 10 paths outside
// the end walk, 12 fragment ranks out of range, 13 new multi-colour sets
constexpr int N_COUNTERS = 16;

// Colour-set hash: a sum of mixed terms over the sorted unique list, so a
// wave can add its lanes' terms in any order and get the thread's value.
__device__ __forceinline__ uint64_t set_term(uint32_t c) { return mix64(uint64_t(c) + 0x9E3779B97F4A7C15ull); }
// a second, independent term: the two sums together are the set's 128-bit
// content hash (the reference dedups on a 128-bit content hash alone,
// streaming_color_set_dict.hpp:329-341)
__device__ __forceinline__ uint64_t set_term2(uint32_t c) {
    return mix64((uint64_t(c) * 0xD6E8FEB86659FD93ull) ^ 0x5851F42D4C957F2Dull);
}
__device__ __forceinline__ uint64_t set_hash_fin(uint64_t sum, uint32_t n) {
    return mix64(sum ^ (uint64_t(n) * 0xC2B2AE3D27D4EB4Full)) | 1ull;  // 0 marks an empty slot
}

// One thread per node (a run of equal keys [t0, t1) in the sorted tuples):
// flags OR, seed rank = min, primary = any owned occurrence, the sorted unique
// owned colours (written in place at colors[t0..]) and their set hash
// (bucket_walker.hpp:228-249 merge of map entries).  A block takes the nodes
// whose first occurrence lies in its tile of NT_TILE tuples and stages the
// tile (plus the BIG_RUN tuples after it) in LDS; runs longer than BIG_RUN
// are queued for k_nodes_big.  A node's index is its tile's first node
// (scanned tile head counts) plus its rank among the tile's heads; the node
// writes its own start.
__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

template <class K>
__global__ __launch_bounds__(TPB) void k_nodes(uint64_t nt, const uint8_t* __restrict__ heads,
                                               const uint32_t* __restrict__ tile_base,
                                               const uint64_t* __restrict__ lbtup, uint32_t nb,
                                               uint32_t* __restrict__ node_bucket,
                                               uint32_t* __restrict__ node_start, const K* __restrict__ keys,
                                               const uint64_t* __restrict__ vals, K* __restrict__ node_key,
                                               uint8_t* __restrict__ node_meta, uint32_t* __restrict__ node_rank,
                                               uint32_t* __restrict__ node_ncol, uint64_t* __restrict__ node_hash,
                                               uint64_t* __restrict__ node_hash2, uint32_t* __restrict__ colors,
                                               uint8_t* __restrict__ big_flag) {
    __shared__ uint64_t s_val[NT_TILE + BIG_RUN];
    __shared__ uint32_t s_list[NT_TILE];
    __shared__ uint32_t s_wn[TPB / 64 + 1];
    __shared__ uint32_t s_next;  // first head in [T1, TE); nt when the tuples end first; NONE32: none
    __shared__ uint32_t s_blo, s_bhi;  // local buckets of the tile's first and last tuples
    const uint64_t T0 = uint64_t(blockIdx.x) * NT_TILE;
    const uint64_t T1 = min(T0 + NT_TILE, nt), TE = min(T1 + BIG_RUN, nt);
    if (threadIdx.x == 64) s_blo = bucket_of(lbtup, nb, T0);
    if (threadIdx.x == 128) s_bhi = bucket_of(lbtup, nb, T1 - 1);
    for (uint32_t i = threadIdx.x; i < TE - T0; i += TPB) s_val[i] = vals[T0 + i];
    static_assert(BIG_RUN <= 64, "one wave looks past the tile");
    if (threadIdx.x < 64) {
        const uint64_t t = T1 + threadIdx.x;
        const uint64_t m = __ballot(threadIdx.x < BIG_RUN && t < TE && heads[t]);
        if (threadIdx.x == 0) s_next = m ? uint32_t(T1 + __ffsll((unsigned long long)m) - 1) : (TE == nt ? uint32_t(nt) : NONE32);
    }
    // heads of the tile, in order: wave w lists the heads of its 4 x 64 tuples
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lanes_below(lane);
    uint32_t my = 0;
    bool h[NT_TILE / TPB];
#pragma unroll
    for (int r = 0; r < int(NT_TILE / TPB); ++r) {
        const uint32_t i = uint32_t(wave) * (NT_TILE / (TPB / 64)) + uint32_t(r) * 64 + uint32_t(lane);
        h[r] = T0 + i < T1 && heads[T0 + i];
        my += __popcll(__ballot(h[r]));
    }
    if (lane == 0) s_wn[wave] = my;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < TPB / 64; ++w) {
            const uint32_t c = s_wn[w];
            s_wn[w] = run;
            run += c;
        }
        s_wn[TPB / 64] = run;
    }
    __syncthreads();
    uint32_t at = s_wn[wave];
#pragma unroll
    for (int r = 0; r < int(NT_TILE / TPB); ++r) {
        const uint32_t i = uint32_t(wave) * (NT_TILE / (TPB / 64)) + uint32_t(r) * 64 + uint32_t(lane);
        const uint64_t m = __ballot(h[r]);
        if (h[r]) s_list[at + __popcll(m & lt)] = i;
        at += __popcll(m);
    }
    __syncthreads();
    const uint32_t n_heads = s_wn[TPB / 64];
    const uint32_t base = tile_base[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < n_heads; j += TPB) {
        const uint32_t hd = s_list[j];
        const uint32_t t0 = uint32_t(T0 + hd);
        const uint32_t n = base + j;
        const uint32_t t1 = j + 1 < n_heads ? uint32_t(T0 + s_list[j + 1]) : s_next;
        node_start[n] = t0;
        node_key[n] = keys[t0];
        uint32_t bk = s_blo;  // a tile seldom spans two buckets
        if (s_bhi != s_blo) {
            uint32_t hi = s_bhi + 1;
            while (hi - bk > 1) {
                const uint32_t mid = (bk + hi) >> 1;
                if (lbtup[mid] <= t0)
                    bk = mid;
                else
                    hi = mid;
            }
        }
        node_bucket[n] = bk;
        const bool is_big = t1 == NONE32 || t1 - t0 > BIG_RUN;
        big_flag[n] = is_big;
        if (is_big) continue;
        uint32_t fl = 0, rank = 0xFFFFFFFFu, ncol = 0, last = 0xFFFFFFFFu;
        uint64_t sum = 0, sum2 = 0;
        bool primary = false;
        for (uint32_t t = t0; t < t1; ++t) {
            const uint64_t v = s_val[t - T0];
            fl |= v_contrib(v);
            const uint32_t r = v_rank(v);
            rank = r < rank ? r : rank;
            if (v_owned(v)) {
                primary = true;
                const uint32_t c = v_color(v);
                if (c != last) {
                    colors[t0 + ncol] = c;
                    sum += set_term(c);
                    sum2 += set_term2(c);
                    ++ncol;
                    last = c;
                }
            }
        }
        node_meta[n] = uint8_t(fl | (primary ? 4u : 0u));
        node_rank[n] = rank;
        node_ncol[n] = ncol;
        node_hash[n] = set_hash_fin(sum, ncol);
        node_hash2[n] = sum2;
    }
}

// One wave per queued node (the grid holds one wave per queued node): the
// same reduction with the run read 256 occurrences at a time.  A lane's
// colour is new when it differs from the colour of the nearest owned
// occurrence before it (runs are colour-sorted).
constexpr int BIG_UNROLL = 4;
__global__ __launch_bounds__(TPB) void k_nodes_big(const uint32_t* __restrict__ big, uint64_t nbig,
                                                   const uint32_t* __restrict__ node_start,
                                                   const uint64_t* __restrict__ vals, uint8_t* __restrict__ node_meta,
                                                   uint32_t* __restrict__ node_rank, uint32_t* __restrict__ node_ncol,
                                                   uint64_t* __restrict__ node_hash, uint64_t* __restrict__ node_hash2,
                                                   uint32_t* __restrict__ colors) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below(lane);
    const uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (i < nbig) {
        const uint32_t n = big[i];
        const uint32_t t0 = node_start[n], t1 = node_start[n + 1];
        uint32_t fl = 0, rank = 0xFFFFFFFFu, ncol = 0, last = 0xFFFFFFFFu;
        uint64_t sum = 0, sum2 = 0;
        bool primary = false;
        for (uint32_t base = t0; base < t1; base += 64 * BIG_UNROLL) {
            uint64_t v[BIG_UNROLL];
#pragma unroll
            for (int u = 0; u < BIG_UNROLL; ++u) {
                const uint32_t t = base + uint32_t(64 * u + lane);
                v[u] = t < t1 ? vals[t] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < BIG_UNROLL; ++u) {
                const bool active = base + uint32_t(64 * u + lane) < t1;
                const bool owned = active && v_owned(v[u]);
                const uint32_t c = v_color(v[u]);
                if (active) {
                    fl |= v_contrib(v[u]);
                    const uint32_t r = v_rank(v[u]);
                    rank = r < rank ? r : rank;
                }
                const uint64_t om = __ballot(owned);
                const uint64_t below = om & lt;
                const int src = below ? 63 - __clzll(below) : lane;
                const uint32_t pc = __shfl(c, src, 64);
                const uint32_t prev = below ? pc : last;
                const bool fresh = owned && c != prev;
                const uint64_t fm = __ballot(fresh);
                if (fresh) {
                    colors[t0 + ncol + __popcll(fm & lt)] = c;
                    sum += set_term(c);
                    sum2 += set_term2(c);
                }
                ncol += __popcll(fm);
                if (om) {
                    last = __shfl(c, 63 - __clzll(om), 64);
                    primary = true;
                }
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            fl |= __shfl_xor(fl, o, 64);
            const uint32_t r = __shfl_xor(rank, o, 64);
            rank = r < rank ? r : rank;
            sum += __shfl_xor(sum, o, 64);
            sum2 += __shfl_xor(sum2, o, 64);
        }
        if (lane == 0) {
            node_meta[n] = uint8_t(fl | (primary ? 4u : 0u));
            node_rank[n] = rank;
            node_ncol[n] = ncol;
            node_hash[n] = set_hash_fin(sum, ncol);
            node_hash2[n] = sum2;
        }
    }
}

// Indices of flagged items: list[scan[i]] = i.
template <class F>
__global__ void k_compact(uint64_t n, const F* __restrict__ flag, const uint32_t* __restrict__ scan,
                          uint32_t* __restrict__ list) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) list[scan[i]] = uint32_t(i);
}

// counters[0] += primary nodes, counters[1] += primary nodes of >= 2 colours
__global__ void k_count_nodes(uint32_t n_nodes, const uint8_t* __restrict__ meta, const uint32_t* __restrict__ ncol,
                              unsigned long long* __restrict__ counters) {
    uint64_t prim = 0, multi = 0;
    for (uint64_t n = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; n < n_nodes; n += uint64_t(gridDim.x) * blockDim.x) {
        const bool p = meta[n] & 4u;
        prim += p;
        multi += p && ncol[n] >= 2;
    }
    block_add(&counters[0], prim);
    block_add(&counters[1], multi);
}



// ---------------------------------------------------------------------------
// colour sets
// ---------------------------------------------------------------------------
struct SetTables {
    unsigned long long* single_min;  // [num_colors] min key of the set {c}
    uint32_t* single_cid;            // [num_colors]
    unsigned long long* mt_hk;       // [2 cap] per slot: hash (0 = empty), min key; one line per probe
    uint32_t* mt_cid;                // [cap]
    uint64_t* mt_h2;                 // [cap] the second half of the set's 128-bit content hash
    uint64_t* mt_loff;               // [cap] colour list of a set created in this chunk, in cur_colors
    uint32_t* mt_llen;               // [cap]
    uint64_t cap;
    const uint32_t* cur_colors;      // node colour lists of the chunk being processed
};

// The colour list of a set created in the chunk being processed (the list
// of the node that created it); sets of earlier chunks keep no list.
__device__ __forceinline__ const uint32_t* slot_list(const SetTables& T, uint64_t s) {
    return T.cur_colors + T.mt_loff[s];
}

// Colour lists live at colors[coff[n]] (coff = the node's first occurrence).
__global__ void k_set_insert(uint32_t n_nodes, const uint8_t* __restrict__ meta, const uint32_t* __restrict__ ncol,
                             const uint32_t* __restrict__ coff, const uint32_t* __restrict__ colors,
                             const uint64_t* __restrict__ node_hash, const uint32_t* __restrict__ rank,
                             const uint32_t* __restrict__ node_bucket, uint32_t b0, SetTables T,
                             const uint64_t* __restrict__ node_hash2, uint32_t* __restrict__ node_slot,
                             uint8_t* __restrict__ created) {
    const uint64_t n = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    if (!(meta[n] & 4u)) {
        created[n] = 0;
        return;
    }
    const uint32_t b = b0 + node_bucket[n];
    const unsigned long long key = (uint64_t(b) << 32) | rank[n];
    const uint32_t nc = ncol[n];
    const uint32_t* c = colors + coff[n];
    if (nc == 1) {
        // keys only decrease: a plain read that is already <= key skips the
        // atomic; the one thread that finds the slot empty creates the set
        bool made = false;
        if (T.single_min[c[0]] > key) made = atomicMin(&T.single_min[c[0]], key) == ~0ull;
        created[n] = made ? 1 : 0;
        return;
    }
    bool made = false;
    const unsigned long long h = node_hash[n];
    uint64_t s = h & (T.cap - 1);
    for (uint64_t probe = 0; probe < T.cap; ++probe) {
        unsigned long long cur = T.mt_hk[2 * s];
        if (cur == 0ull) {
            cur = atomicCAS(&T.mt_hk[2 * s], 0ull, h);
            if (cur == 0ull) {  // this thread created the slot: its list stands for the set
                T.mt_loff[s] = coff[n];
                T.mt_llen[s] = nc;
                T.mt_h2[s] = node_hash2[n];
                cur = h;
                made = true;
            }
        }
        if (cur == h) {
            if (T.mt_hk[2 * s + 1] > key) atomicMin(&T.mt_hk[2 * s + 1], key);
            node_slot[n] = uint32_t(s);
            created[n] = made ? 1 : 0;
            return;
        }
        s = (s + 1) & (T.cap - 1);
    }
    created[n] = 0;  // table full (cannot happen at the load kept below 1/2)
}

// Every multi-colour primary node has the set of its hash slot: the second
// half of the 128-bit content hash and the size agree (a mismatch is a 64-bit
// hash collision between two sets, reported as an error).
__global__ void k_set_verify(uint32_t n_nodes, const uint8_t* __restrict__ meta, const uint32_t* __restrict__ ncol,
                             const uint64_t* __restrict__ node_hash2, const uint32_t* __restrict__ node_slot,
                             SetTables T, unsigned long long* __restrict__ err) {
    const uint64_t n = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= n_nodes || !(meta[n] & 4u) || ncol[n] < 2) return;
    const uint32_t s = node_slot[n];
    if (T.mt_h2[s] != node_hash2[n] || T.mt_llen[s] != ncol[n]) atomicAdd(err, 1ull);
}

__global__ void k_mt_rehash(uint64_t old_cap, const unsigned long long* __restrict__ ohk,
                            const uint32_t* __restrict__ oc, const uint64_t* __restrict__ oh2,
                            const uint32_t* __restrict__ ol, SetTables T) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= old_cap || ohk[2 * i] == 0ull) return;
    const unsigned long long h = ohk[2 * i];
    uint64_t s = h & (T.cap - 1);
    for (uint64_t probe = 0; probe < T.cap; ++probe) {
        if (atomicCAS(&T.mt_hk[2 * s], 0ull, h) == 0ull) {
            T.mt_hk[2 * s + 1] = ohk[2 * i + 1];
            T.mt_cid[s] = oc[i];
            T.mt_h2[s] = oh2[i];
            T.mt_llen[s] = ol[i];
            return;
        }
        s = (s + 1) & (T.cap - 1);
    }
}

// The sets created in this chunk, one per listed creating node: the
// singleton {c} or the node's hash slot, with the set's min key.
__global__ void k_newset_emit(uint64_t n_new, const uint32_t* __restrict__ list, const uint32_t* __restrict__ ncol,
                              const uint32_t* __restrict__ coff, const uint32_t* __restrict__ colors,
                              const uint32_t* __restrict__ node_slot, SetTables T, uint64_t* __restrict__ set_key,
                              uint64_t* __restrict__ set_desc) {
    const uint64_t d = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (d >= n_new) return;
    const uint32_t n = list[d];
    if (ncol[n] == 1) {
        const uint32_t c = colors[coff[n]];
        set_key[d] = T.single_min[c];
        set_desc[d] = c;
    } else {
        const uint32_t s = node_slot[n];
        set_key[d] = T.mt_hk[2 * uint64_t(s) + 1];
        set_desc[d] = (uint64_t(1) << 63) | s;
    }
}

__device__ __forceinline__ uint32_t msb64d(uint64_t x) { return 63u - uint32_t(__clzll(x)); }
__device__ __forceinline__ uint64_t delta_len(uint64_t x) {
    const uint64_t b = msb64d(x + 1);
    return 2ull * msb64d(b + 1) + 1 + b;
}

// The colours of a set: a list of a hash slot, or one colour held by value
// (a pointer to a local would put it in scratch memory).
struct SetColors {
    const uint32_t* list;  // null for a singleton
    uint32_t n, single;
    __device__ __forceinline__ uint32_t operator[](uint32_t j) const { return list ? list[j] : single; }
};

__device__ __forceinline__ SetColors set_list(uint64_t desc, const SetTables& T) {
    if (desc >> 63) {
        const uint64_t s = desc & 0xFFFFFFFFFFFFull;
        return SetColors{slot_list(T, s), T.mt_llen[s], 0u};
    }
    return SetColors{nullptr, 1u, uint32_t(desc)};
}

__global__ void k_set_cid_len(uint64_t n_new, uint64_t cid_base, const uint64_t* __restrict__ set_desc, SetTables T,
                              uint32_t num_colors, uint64_t* __restrict__ bits,
                              unsigned long long* __restrict__ counters) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t ints = 0, multi = 0;
    if (i < n_new) {
        const uint64_t d = set_desc[i];
        if (d >> 63)
            T.mt_cid[d & 0xFFFFFFFFFFFFull] = uint32_t(cid_base + i);
        else
            T.single_cid[uint32_t(d)] = uint32_t(cid_base + i);
        const SetColors c = set_list(d, T);
        const uint32_t n = c.n;
        const uint32_t sparse = num_colors / 4, dense = uint32_t(uint64_t(num_colors) * 3 / 4);
        uint64_t L = delta_len(n);
        if (n < sparse) {
            L += delta_len(c[0]);
            for (uint32_t j = 1; j < n; ++j) L += delta_len(uint64_t(c[j] - (c[j - 1] + 1)));
        } else if (n < dense) {
            L += num_colors;
        } else {
            // delta gaps over the absent colours (hybrid_color_sets.hpp:67-94)
            uint32_t j = 0, prev = 0xFFFFFFFFu;
            for (uint32_t v = 0; v < num_colors; ++v) {
                if (j < n && c[j] == v) {
                    ++j;
                    continue;
                }
                L += delta_len(uint64_t(v - (prev + 1)));
                prev = v;
            }
        }
        bits[i] = L;
        ints = n;
        multi = d >> 63;
    }
    block_add(&counters[2], ints);
    block_add(&counters[13], multi);
}

// ORs the low `len` (<= 64) bits of v in at bit `pos` (words zeroed beforehand).
__device__ __forceinline__ void or_bits(unsigned long long* words, uint64_t pos, uint64_t v, uint32_t len) {
    if (len == 0) return;
    const uint64_t w = pos >> 6;
    const uint32_t o = uint32_t(pos & 63);
    atomicOr(&words[w], (unsigned long long)(v << o));
    if (o && o + len > 64) atomicOr(&words[w + 1], (unsigned long long)(v >> (64 - o)));
}

// Elias delta code of x (x + 1 <= 2^32) as one bit string, LSB first: gamma(b)
// = b' zero bits, then b' + 1 bits holding (b + 1) with its leading 1 moved to
// the front, then the b low bits of x + 1 (b = floor(log2(x + 1)),
// b' = floor(log2(b + 1))); returns its length (<= 43).
__device__ __forceinline__ uint32_t delta_code(uint64_t x, uint64_t& code) {
    const uint64_t xx = x + 1;
    const uint32_t b = msb64d(xx);
    const uint64_t g = uint64_t(b) + 1;
    const uint32_t bb = msb64d(g);
    const uint64_t gbits = ((g & ((uint64_t(1) << bb) - 1)) << 1) | 1u;
    code = (gbits << bb) | ((xx & ((uint64_t(1) << b) - 1)) << (2 * bb + 1));
    return 2 * bb + 1 + b;
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}

// Colours of c in the round [lo0, lo0 + 4096) as 64 words in the wave's LDS
// row; returns lane q's word (colours lo0 + 64 q + bit).  `j`, the same in
// every lane, is the first list element >= lo0 and moves past the round: the
// lanes take 64 list elements at a time (coalesced) and OR their bits in.
constexpr uint32_t ROUND_COLORS = 64 * 64;
__device__ __forceinline__ uint64_t round_bits(const SetColors& c, uint64_t lo0, uint32_t& j,
                                               unsigned long long* row) {
    const int lane = threadIdx.x & 63;
    row[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (;;) {
        const uint32_t jj = j + uint32_t(lane);
        const uint64_t v = jj < c.n ? uint64_t(c[jj]) : ~0ull;
        const bool in = v < lo0 + ROUND_COLORS;  // the list is sorted: a prefix of the lanes
        if (in) atomicOr(&row[(v - lo0) >> 6], 1ull << (v & 63));
        const uint32_t cnt = uint32_t(__popcll(__ballot(in)));
        j += cnt;
        if (cnt < 64) break;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    return row[lane];
}

// Absent colours of lane's range [lo, lo + 64) given the present bits.
__device__ __forceinline__ uint64_t absent_bits(uint64_t present, uint64_t lo, uint32_t num_colors) {
    if (lo >= num_colors) return 0;
    const uint64_t w = num_colors - lo;
    return ~present & (w >= 64 ? ~0ull : (uint64_t(1) << w) - 1);
}

// Absent colour before lane's range: the last absent colour of the lower
// lanes' ranges, else `prev` (the last one of earlier rounds, -1: none).
// `top` receives the last absent colour up to this round's end (or prev).
__device__ __forceinline__ int64_t absent_before(uint64_t absent, uint64_t lo, int64_t prev, int64_t& top) {
    const int lane = threadIdx.x & 63;
    uint64_t mx = absent ? lo + msb64d(absent) + 1 : 0;  // last absent + 1 (0: none)
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(mx, o, 64);
        if (lane >= o && y > mx) mx = y;
    }
    const uint64_t below = __shfl_up(mx, 1, 64);
    const uint64_t all = __shfl(mx, 63, 64);
    top = all ? int64_t(all) - 1 : prev;
    return lane > 0 && below ? int64_t(below) - 1 : prev;
}

// Hybrid colour-set codes (hybrid_color_sets.hpp:41-96), one wave per set:
// delta(n), then below n_colors / 4 colours the delta gaps of the list, below
// 3/4 the n_colors-bit bitmap, else the delta gaps of the absent colours.  A
// lane encodes one list element, or one 64-colour range of the bitmap / of
// the absent colours; wave scans place the codes.
__global__ __launch_bounds__(TPB) void k_set_encode(uint64_t n_new, uint64_t cid_base, uint64_t bit_base,
                                                    const uint64_t* __restrict__ set_desc,
                                                    const uint64_t* __restrict__ boff, SetTables T, uint32_t num_colors,
                                                    unsigned long long* __restrict__ words,
                                                    uint64_t* __restrict__ set_off) {
    __shared__ unsigned long long s_row[TPB / 64][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (i >= n_new) return;  // the same for the whole wave
    const SetColors c = set_list(set_desc[i], T);
    const uint32_t n = c.n;
    const uint32_t sparse = num_colors / 4, dense = uint32_t(uint64_t(num_colors) * 3 / 4);
    const uint64_t at = bit_base + boff[i];
    uint64_t code;
    const uint32_t hl = delta_code(n, code);
    if (lane == 0) {
        set_off[cid_base + i] = at;
        or_bits(words, at, code, hl);
    }
    uint64_t pos = at + hl;
    if (n < sparse) {
        for (uint32_t j0 = 0; j0 < n; j0 += 64) {
            const uint32_t j = j0 + uint32_t(lane);
            uint32_t len = 0;
            if (j < n) {
                const uint32_t v = c[j];
                len = delta_code(j == 0 ? uint64_t(v) : uint64_t(v - (c[j - 1] + 1)), code);
            }
            const uint32_t incl = wave_incl_sum(len);
            if (j < n) or_bits(words, pos + incl - len, code, len);
            pos += __shfl(incl, 63, 64);
        }
    } else if (n < dense) {
        uint32_t j = 0;
        for (uint64_t lo0 = 0; lo0 < num_colors; lo0 += ROUND_COLORS) {
            const uint64_t present = round_bits(c, lo0, j, s_row[wave]);
            const uint64_t lo = lo0 + uint64_t(lane) * 64;
            if (lo < num_colors) {
                const uint32_t w = uint32_t(num_colors - lo < 64 ? num_colors - lo : 64);
                or_bits(words, pos + lo, present, w);
            }
        }
    } else {
        int64_t prev = -1;  // last absent colour before this round of ranges
        uint32_t j = 0;
        for (uint64_t lo0 = 0; lo0 < num_colors; lo0 += ROUND_COLORS) {
            const uint64_t lo = lo0 + uint64_t(lane) * 64;
            const uint64_t absent = absent_bits(round_bits(c, lo0, j, s_row[wave]), lo, num_colors);
            int64_t top;
            const int64_t p0 = absent_before(absent, lo, prev, top);
            uint32_t len = 0;
            int64_t p = p0;
            for (uint64_t m = absent; m; m &= m - 1) {
                const int64_t v = int64_t(lo) + __ffsll((unsigned long long)m) - 1;
                len += uint32_t(delta_len(uint64_t(v - (p + 1))));
                p = v;
            }
            const uint32_t incl = wave_incl_sum(len);
            uint64_t q = pos + incl - len;
            p = p0;
            for (uint64_t m = absent; m; m &= m - 1) {
                const int64_t v = int64_t(lo) + __ffsll((unsigned long long)m) - 1;
                const uint32_t l = delta_code(uint64_t(v - (p + 1)), code);
                or_bits(words, q, code, l);
                q += l;
                p = v;
            }
            pos += __shfl(incl, 63, 64);
            prev = top;
        }
    }
}

// Everything the walk reads of a node, in one record: a path position
// gathers one cache line (32 B for 64-bit keys, 64 B for 128-bit keys).
template <class K>
struct alignas(sizeof(K) == 8 ? 32 : 64) NodeRec {
    K key;
    uint32_t rank;
    uint32_t cid;
    uint32_t link[2];  // per port: (neighbour << 1) | entry port, NONE32 when not a unitig link
    uint16_t mm;       // meta | mask << 8
};

// ---------------------------------------------------------------------------
// per-bucket directory over the sorted node keys, masks, links
// ---------------------------------------------------------------------------
// The nodes of a bucket are its distinct k-mers in increasing order.  A key's
// cell d in [0, n_b) is a monotone function of the key, flattened by the
// distribution of canonical k-mers (the smaller of a random word and its
// reverse complement: CDF 1 - (1 - x)^2) and computed in integers so that it
// is exactly monotone.  dir[d] = the first node of the bucket whose cell is
// >= d, so the nodes of cell d are [dir[d], dir[d + 1]): a lookup reads two
// directory words and about one key, and the node index is the position.
// Bucket b's directory has n_b + 1 words and starts at node_off[b] + b.
//
// A presence filter in front of it: the bitmap of occupied fine cells,
// 2^FINE_SHIFT per node (fine cell f lies in cell f >> FINE_SHIFT), so an
// absent key passes one bit with probability about 2^-FINE_SHIFT, from a
// bitmap that stays in L2 while an XCD works through a bucket.  Bucket b's
// bitmap has ceil(n_b / NODES_PER_WORD) 32-bit words and starts at word
// node_off[b] / NODES_PER_WORD + b.
constexpr uint32_t FINE_SHIFT = 3;  // 16 per node measured no faster (fewer searches, more bitmap lines)
constexpr uint32_t NODES_PER_WORD = 32u >> FINE_SHIFT;

template <class K>
__device__ __forceinline__ uint32_t top32(const K& key, uint32_t k);
template <>
__device__ __forceinline__ uint32_t top32<uint64_t>(const uint64_t& key, uint32_t k) {
    const uint32_t b = 2 * k;
    return b >= 32 ? uint32_t(key >> (b - 32)) : uint32_t(key << (32 - b));
}
template <>
__device__ __forceinline__ uint32_t top32<Key128>(const Key128& key, uint32_t k) {
    const uint32_t b = 2 * k;  // 66 .. 126 for k > 32
    if (b <= 64) return top32<uint64_t>(key.lo, k);
    const uint32_t sh = b - 32;  // >= 34
    return sh >= 64 ? uint32_t(key.hi >> (sh - 64)) : uint32_t((key.lo >> sh) | (key.hi << (64 - sh)));
}

template <class K>
__device__ __forceinline__ uint64_t cell_of(const K& key, uint32_t k, uint64_t size) {
    const uint64_t u = (uint64_t(1) << 32) - top32<K>(key, k);  // 1 .. 2^32
    const uint64_t f = uint64_t(0) - u * u;                    // 2^64 - u^2 (0 when u = 2^32)
    return ((f >> 32) * size) >> 32;
}

template <class K>
__device__ __forceinline__ bool key_lt(const K& a, const K& b) {
    return !KOps<K>::le(b, a);
}

// Node n writes the directory words of the cells after its predecessor's
// cell up to its own; the bucket's last node also writes the words after it.
// Filter words: each lane's bit is ORed over the lanes of the wave that share
// its word (words are non-decreasing along the nodes), and the first node of
// the word stores it, after adding the nodes past the wave's end that still
// fall in it; it also clears the empty words before it (the bucket's last
// node: the words after it).
template <class K>
__global__ __launch_bounds__(TPB) void k_dir_build(uint32_t n_nodes, const K* __restrict__ node_key,
                                                   const uint32_t* __restrict__ node_bucket,
                                                   const uint64_t* __restrict__ node_off, uint32_t k,
                                                   uint32_t* __restrict__ dir, uint32_t* __restrict__ filt) {
    const uint64_t n = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const bool live = n < n_nodes;
    uint64_t first = 0, last = 0, size = 0, f = 0, fp = 0, gw = ~0ull;
    uint32_t b = 0, bits = 0;
    if (live) {
        b = node_bucket[n];
        first = node_off[b], last = node_off[b + 1], size = last - first;
        const uint64_t fsize = size << FINE_SHIFT;
        f = cell_of<K>(node_key[n], k, fsize);
        fp = n == first ? 0 : cell_of<K>(node_key[n - 1], k, fsize);
        uint32_t* d = dir + first + b;
        const uint64_t c = f >> FINE_SHIFT, c0 = n == first ? 0 : (fp >> FINE_SHIFT) + 1;
        for (uint64_t j = c0; j <= c; ++j) d[j] = uint32_t(n);
        if (n + 1 == last)
            for (uint64_t j = c + 1; j <= size; ++j) d[j] = uint32_t(last);
        gw = first / NODES_PER_WORD + b + (f >> 5);
        bits = 1u << (f & 31);
    }
#pragma unroll
    for (uint32_t sft = 1; sft < 64; sft <<= 1) {
        const uint32_t ob = __shfl_down(bits, sft, 64);
        const uint64_t ow = __shfl_down(gw, sft, 64);
        if (lane + sft < 64 && ow == gw) bits |= ob;
    }
    const uint64_t gw_end = __shfl(gw, 63, 64);
    if (!live) return;
    uint32_t* fw = filt + first / NODES_PER_WORD + b;
    const uint64_t w = f >> 5, nw = (size + NODES_PER_WORD - 1) / NODES_PER_WORD;
    if (n == first || (fp >> 5) < w) {
        const uint64_t w0 = n == first ? 0 : (fp >> 5) + 1;
        for (uint64_t j = w0; j < w; ++j) fw[j] = 0u;
        if (gw_end == gw) {  // the word continues past this wave
            const uint64_t fsize = size << FINE_SHIFT;
            for (uint64_t m = n - lane + 64; m < last; ++m) {
                const uint64_t g = cell_of<K>(node_key[m], k, fsize);
                if ((g >> 5) != w) break;
                bits |= 1u << (g & 31);
            }
        }
        fw[w] = bits;
    }
    if (n + 1 == last)
        for (uint64_t j = w + 1; j < nw; ++j) fw[j] = 0u;
}

// mask bits 0..3: successors of the canonical k-mer, 4..7: successors of its
// reverse complement (local_ext_mask, bucket_walker.hpp:260-271).  For a side
// with exactly one present extension, nb[2n+port] = (neighbour << 1) | entry
// port: the neighbour is entered at its L side iff the extension is canonical.
// The four lookups of a side are independent: their filter words, then their
// directory words, then the first key of each cell are loaded together;
// cells holding more keys continue by binary search.  The two sides run one after
// the other (register pressure: eight in flight halve the waves per SIMD).
template <class K>
__device__ __forceinline__ uint32_t side_lookup(const K& x, uint32_t k, const K* __restrict__ node_key,
                                                const uint32_t* __restrict__ d, const uint32_t* __restrict__ fwords,
                                                uint64_t size, uint32_t& found, uint32_t& port) {
    using O = KOps<K>;
    K q[4];
    bool fw[4];
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) {
        const K e = O::shift_append(x, nt, k);
        const K er = O::rc(e, k);
        fw[nt] = O::le(e, er);
        q[nt] = fw[nt] ? e : er;
    }
    uint64_t c[4];
    uint32_t word[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        c[i] = cell_of<K>(q[i], k, size << FINE_SHIFT);
        word[i] = fwords[c[i] >> 5];
    }
    uint32_t lo[4], hi[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        lo[i] = hi[i] = 0;
        if ((word[i] >> (c[i] & 31)) & 1u) {  // the key's fine cell is occupied
            lo[i] = d[c[i] >> FINE_SHIFT];
            hi[i] = d[(c[i] >> FINE_SHIFT) + 1];
        }
    }
    K v0[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i)
        if (hi[i] > lo[i]) v0[i] = node_key[lo[i]];
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        uint32_t r = NONE32;
        if (hi[i] > lo[i]) {
            if (O::eq(v0[i], q[i])) {
                r = lo[i];
            } else if (hi[i] > lo[i] + 1 && key_lt<K>(v0[i], q[i])) {
                uint32_t a = lo[i] + 1, z = hi[i];  // first key >= q in [a, z)
                while (a < z) {
                    const uint32_t mid = (a + z) >> 1;
                    if (key_lt<K>(node_key[mid], q[i]))
                        a = mid + 1;
                    else
                        z = mid;
                }
                if (a < hi[i] && O::eq(node_key[a], q[i])) r = a;
            }
        }
        if (r != NONE32) {
            mask |= 1u << i;
            found = r;
            port = fw[i] ? PORT_L : PORT_R;
        }
    }
    return mask;
}

template <class K>
__global__ __attribute__((amdgpu_waves_per_eu(sizeof(K) == 8 ? 8 : 6))) void k_masks(uint32_t n_nodes, const K* __restrict__ node_key, const uint32_t* __restrict__ node_bucket,
                        const uint64_t* __restrict__ node_off, const uint32_t* __restrict__ dir,
                        const uint32_t* __restrict__ filt, uint32_t k, uint8_t* __restrict__ mask_out,
                        uint32_t* __restrict__ nb) {
    const uint32_t nblk = (n_nodes + TPB - 1) / TPB, blk = tile_of(blockIdx.x, nblk);
    const uint64_t n = uint64_t(blk) * TPB + threadIdx.x;
    if (blk >= nblk || n >= n_nodes) return;
    const uint32_t b = node_bucket[n];
    const uint64_t first = node_off[b], size = node_off[b + 1] - first;
    const uint32_t* d = dir + first + b;
    const uint32_t* fwords = filt + first / NODES_PER_WORD + b;
    const K can = node_key[n];
    uint32_t sR = NONE32, pR = 0, sL = NONE32, pL = 0, mR = 0, mL = 0;
#pragma unroll 1
    for (uint32_t side = 0; side < 2; ++side) {
        uint32_t s = NONE32, p = 0;
        const uint32_t m = side_lookup<K>(side ? KOps<K>::rc(can, k) : can, k, node_key, d, fwords, size, s, p);
        if (side) {
            mL = m, sL = s, pL = p;
        } else {
            mR = m, sR = s, pR = p;
        }
    }
    const uint32_t mask = mR | (mL << 4);
    mask_out[n] = uint8_t(mask);
    nb[2 * n + PORT_R] = __popc(mR) == 1 ? (sR << 1) | pR : NONE32;
    nb[2 * n + PORT_L] = __popc(mL) == 1 ? (sL << 1) | pL : NONE32;
}

__device__ __forceinline__ uint32_t nib(uint32_t m, uint32_t port) { return port == PORT_R ? (m & 0xfu) : (m >> 4); }

// Keeps a neighbour as a link when the neighbour's entry side has exactly one
// extension too, and assembles the node's walk record (k-mer, seed rank,
// colour-set id, links, meta and mask); the links are also stored alone, 8 B
// per node, for the walks that read nothing else (a bucket's links stay in
// cache).  Also flags the path ends: nodes with at most one link.
template <class K>
__global__ void k_links(uint32_t n_nodes, const uint8_t* __restrict__ mask, const uint32_t* __restrict__ nb,
                        const uint8_t* __restrict__ meta, const uint32_t* __restrict__ ncol,
                        const uint32_t* __restrict__ coff, const uint32_t* __restrict__ colors,
                        const uint32_t* __restrict__ node_slot, SetTables T, const K* __restrict__ node_key,
                        const uint32_t* __restrict__ rank, NodeRec<K>* __restrict__ rec,
                        uint32_t* __restrict__ link, uint8_t* __restrict__ end_flag) {
    const uint64_t n = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    NodeRec<K> r;
    bool end = false;
#pragma unroll
    for (uint32_t p = 0; p < 2; ++p) {
        uint32_t l = nb[2 * n + p];
        if (l != NONE32 && __popc(nib(mask[l >> 1], l & 1u)) != 1) l = NONE32;
        r.link[p] = l;
        link[2 * n + p] = l;
        end |= l == NONE32;
    }
    end_flag[n] = end ? 1 : 0;
    const uint8_t mt = meta[n];
    uint32_t cid = NONE32;
    if (mt & 4u) cid = ncol[n] == 1 ? T.single_cid[colors[coff[n]]] : T.mt_cid[node_slot[n]];
    r.key = node_key[n];
    r.rank = rank[n];
    r.cid = cid;
    r.mm = uint16_t(mt | (uint32_t(mask[n]) << 8));
    rec[n] = r;
}

// ---------------------------------------------------------------------------
// walk: one thread per component
// ---------------------------------------------------------------------------
struct WalkArgs {
    uint32_t n_nodes;
    const uint32_t* link;          // [2 n + port]: (neighbour << 1) | entry port, or NONE32
    const uint8_t* meta;
    const uint8_t* mask;
    const uint32_t* rank;
    const uint32_t* node_bucket;
    uint32_t b0;
    uint8_t* visited;
    unsigned long long* counters;  // [3] positions placed by the general loop, [5] errors, [9] paths, [10] other
    // fragment descriptors, indexed by the fragment's first position
    uint32_t* frag_n;
    uint8_t* frag_open;
    uint64_t* frag_key;
    uint8_t* frag_dir;             // 1: the fragment is its position range read backwards
    uint8_t* fflag;                // 1 at the first position of a fragment
    // simple paths, recorded end to end
    const uint32_t* comp_x;        // owning end (the end list: a path is stored at its owner's index)
    uint8_t* comp_ex;              // port the path leaves it by
    uint32_t* comp_len;
    uint32_t* comp_nodes;          // [PATH_BUF] per path: node | orientation << 31 (short paths)
    const uint32_t* comp_off;      // first position
    uint32_t* other_x;             // paths that pass a node twice (hairpins)
    uint8_t* other_ex;
    // per position
    uint8_t* prc;                  // bit 0: orientation, bit 1: placed
    uint16_t* pinfo;               // meta | mask << 8
    uint32_t* prank;
    void* pkey;                    // the node's k-mer
    uint32_t* pcid;                // the node's colour-set id
    const void* node_rec;          // NodeRec<K> per node
    bool mark_visited;             // nodes outside simple paths exist: keep `visited` exact
    uint32_t walk_cap;             // links an end follows before its path is left to list ranking
    uint8_t* long_bucket;          // per chunk bucket: 1 when an end of it stopped at walk_cap
};

// comp_ex bits beside the exit port
constexpr uint8_t COMP_RANKED = 2;  // long path: positions come from list ranking
constexpr uint8_t COMP_LONG = 4;    // the end stopped at walk_cap


constexpr uint32_t PATH_BUF = 8;

template <class K>
__device__ __forceinline__ void put_position(const WalkArgs& a, uint64_t pos, uint32_t node, uint8_t o) {
    const NodeRec<K> r = ((const NodeRec<K>*)a.node_rec)[node];
    ((K*)a.pkey)[pos] = r.key;
    a.pcid[pos] = r.cid;
    a.prank[pos] = r.rank;
    a.pinfo[pos] = r.mm;
    a.prc[pos] = o;
    a.fflag[pos] = 0;
}

__device__ __forceinline__ bool is_break(uint8_t meta) {
    const uint32_t f = meta & 3u;
    return f == BOUND_LEFT || f == BOUND_RIGHT;
}

// Index of this lane's item among the wave's requesting lanes, added to a
// counter once per wave.  Every lane of the wave that is still running must
// call it.
__device__ __forceinline__ uint64_t wave_alloc(unsigned long long* ctr, bool want) {
    const int lane = threadIdx.x & 63;
    const uint64_t m = __ballot(want);
    if (!m) return 0;
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + __popcll(m & lanes_below(lane));
}

// emit_from_seed (bucket_walker.hpp:451-542) over the link graph; returns the
// number of nodes placed in the fragment.
template <class K>
__device__ uint32_t emit_fragment(const WalkArgs& a, uint32_t s, uint32_t b) {
    a.visited[s] = 1;
    const uint32_t sflags = a.meta[s] & 3u;
    const uint32_t ms = a.mask[s];
    const uint32_t limit = 2 * a.n_nodes + 2;
    uint32_t nbk = 0, nfw = 0;
    bool open_l = false, open_r = false;
    if (sflags == BOUND_LEFT) {
        open_l = __popc(ms >> 4) <= 1;
    } else {
        uint32_t cur = s, ex = PORT_L;
        for (uint32_t h = 0; h < limit; ++h) {
            const uint32_t l = a.link[2 * cur + ex];
            if (l == NONE32) break;
            const uint32_t y = l >> 1, q = l & 1u;
            if (a.visited[y]) break;
            a.visited[y] = 1;
            ++nbk;
            cur = y;
            ex = 1u - q;
            if (is_break(a.meta[y])) {
                open_l = __popc(nib(a.mask[y], ex)) <= 1;
                break;
            }
        }
    }
    if (sflags == BOUND_RIGHT) {
        open_r = __popc(ms & 0xfu) <= 1;
    } else {
        uint32_t cur = s, ex = PORT_R;
        for (uint32_t h = 0; h < limit; ++h) {
            const uint32_t l = a.link[2 * cur + ex];
            if (l == NONE32) break;
            const uint32_t y = l >> 1, q = l & 1u;
            if (a.visited[y]) break;
            a.visited[y] = 1;
            ++nfw;
            cur = y;
            ex = 1u - q;
            if (is_break(a.meta[y])) {
                open_r = __popc(nib(a.mask[y], ex)) <= 1;
                break;
            }
        }
    }
    const uint32_t total = nbk + 1 + nfw;
    const uint64_t off = atomicAdd(&a.counters[3], (unsigned long long)total);
    if (off + total > a.n_nodes) {
        atomicAdd(&a.counters[5], 1ull);
        return total;
    }
    uint32_t cur = s, ex = PORT_L;
    for (uint32_t i = 0; i < nbk; ++i) {  // walking backward: entered at L <=> reverse orientation
        const uint32_t l = a.link[2 * cur + ex];
        const uint32_t y = l >> 1, q = l & 1u;
        put_position<K>(a, off + nbk - 1 - i, y, q == PORT_L);
        cur = y;
        ex = 1u - q;
    }
    put_position<K>(a, off + nbk, s, 0);
    cur = s;
    ex = PORT_R;
    for (uint32_t i = 0; i < nfw; ++i) {  // walking forward: entered at R <=> reverse orientation
        const uint32_t l = a.link[2 * cur + ex];
        const uint32_t y = l >> 1, q = l & 1u;
        put_position<K>(a, off + nbk + 1 + i, y, q == PORT_R);
        cur = y;
        ex = 1u - q;
    }
    a.fflag[off] = 1;
    a.frag_n[off] = total;
    a.frag_open[off] = uint8_t((open_l ? 1u : 0u) | (open_r ? 2u : 0u));
    a.frag_key[off] = (uint64_t(b) << 32) | a.rank[s];
    a.frag_dir[off] = 0;
    return total;
}

// Seed loop of one component (path from end x, or cycle through x): seeds are
// the unvisited nodes in rank order.  `len` is the component size when the
// first traversal saw every node once (0 when it bounced off a node linked to
// itself, where nodes are passed twice), `seed` the first seed.
//
// A cycle is traversed from x leaving through R until it re-enters x through
// L; a re-entry through R is a bounce (a hairpin node sends the walk back) and
// the traversal continues through x's other side.
template <class K>
__device__ void simulate_component(const WalkArgs& a, uint32_t x, uint32_t ex0, bool cycle, uint32_t len,
                                   uint32_t seed) {
    const uint32_t b = a.b0 + a.node_bucket[x];
    const uint32_t limit = 2 * a.n_nodes + 2;
    uint32_t placed = 0;
    for (uint32_t guard = 0; guard < limit; ++guard) {
        placed += emit_fragment<K>(a, seed, b);
        if (len && placed >= len) return;
        uint32_t best = NONE32, best_rank = 0xFFFFFFFFu;
        uint32_t cur = x, ex = ex0;
        for (uint32_t h = 0; h < limit; ++h) {
            if (!a.visited[cur] && a.rank[cur] < best_rank) {
                best = cur;
                best_rank = a.rank[cur];
            }
            const uint32_t l = a.link[2 * cur + ex];
            if (l == NONE32) break;
            const uint32_t y = l >> 1, q = l & 1u;
            if (cycle && y == x && q == PORT_L) break;
            cur = y;
            ex = 1u - q;
        }
        if (best == NONE32) return;  // every node of the component is placed
        seed = best;
    }
    atomicAdd(&a.counters[5], 1ull);
}

// Path ends walk to the other end; the end with the smaller index owns the
// path.  Paths that never pass a node twice are queued for recording (with
// their nodes when they have at most PATH_BUF), the others (a node linked to
// itself turns the walk back) for the general seed loop.  An end that has
// followed walk_cap links stops: its path is long and is ranked by pointer
// jumping instead (k_long_*), so no lane chases a long path.  One thread
// per end of the dense end list; XCD x takes the x-th eighth of it, which
// keeps a bucket's links in one L2.
__global__ __launch_bounds__(TPB) void k_walk_ends(WalkArgs a, const uint32_t* __restrict__ ends, uint32_t n_ends) {
    const uint32_t nblk = (n_ends + TPB - 1) / TPB;
    const uint32_t blk = tile_of(blockIdx.x, nblk);
    const uint64_t e = uint64_t(blk) * TPB + threadIdx.x;
    bool owner = false, simple = true, lng = false;
    uint32_t x = 0, ex0 = 0, len = 0;
    uint32_t buf[PATH_BUF];
    if (blk < nblk && e < n_ends) {
        x = ends[e];
        const uint32_t lR = a.link[2 * x + PORT_R];
        ex0 = lR == NONE32 ? PORT_L : PORT_R;
        uint32_t cur = x, ex = ex0;
        buf[0] = cur | (ex0 == PORT_L ? 0x80000000u : 0u);
        len = 1;
        const uint32_t limit = 2 * a.n_nodes + 2;
        const uint32_t cap = a.walk_cap < limit ? a.walk_cap : limit;
        bool reached = false;
        for (uint32_t h = 0; h < cap; ++h) {
            const uint32_t l = a.link[2 * cur + ex];
            if (l == NONE32) {
                reached = true;
                break;
            }
            const uint32_t y = l >> 1, q = l & 1u;
            if (y == cur) simple = false;  // hairpin: the walk comes back
            cur = y;
            ex = 1u - q;
#pragma unroll
            for (uint32_t j = 1; j < PATH_BUF; ++j)
                if (j == len) buf[j] = cur | (q == PORT_R ? 0x80000000u : 0u);
            ++len;
        }
        lng = !reached;
        owner = reached && cur >= x;
    }
    const uint64_t o = wave_alloc(&a.counters[10], owner && !simple);
    (void)wave_alloc(&a.counters[11], lng);
    if (blk < nblk && e < n_ends) {
        // a path is stored at the index of its owning end (length 0: none)
        a.comp_len[e] = owner && simple ? len : 0u;
        a.comp_ex[e] = uint8_t(ex0 | (lng ? COMP_LONG : 0u));
        if (lng) a.long_bucket[a.node_bucket[x]] = 1;
        if (owner && simple && len <= PATH_BUF)
#pragma unroll
            for (uint32_t j = 0; j < PATH_BUF; ++j)
                if (j < len) a.comp_nodes[e * PATH_BUF + j] = buf[j];
    }
    if (owner && !simple) {
        a.other_x[o] = x;
        a.other_ex[o] = uint8_t(ex0);
    }
}

// Writes path c end to end from position comp_off[c]: per position the
// node's data and its orientation along the path (1 when it is passed
// entering by R).  The paths of a wave's ends occupy one contiguous range of
// positions; when it holds at most rec_cap<K>() positions it is assembled in
// LDS and written out with coalesced stores.  The stage size was measured:
// for 64-bit keys 320 positions (24 KB per block, six blocks per CU: the
// gathers want waves in flight) beat 384, 512 and 256 (too many waves
// overflow the stage); 128-bit keys did best with 512.
template <class K>
constexpr uint32_t rec_cap() { return sizeof(K) == 8 ? 320 : 512; }

template <class K>
__global__ __launch_bounds__(TPB) void k_walk_record(WalkArgs a, uint64_t n_comp) {
    constexpr uint32_t REC_CAP = rec_cap<K>();
    __shared__ K s_key[TPB / 64][REC_CAP];
    __shared__ uint32_t s_cid[TPB / 64][REC_CAP];
    __shared__ uint32_t s_rank[TPB / 64][REC_CAP];
    __shared__ uint16_t s_info[TPB / 64][REC_CAP];
    __shared__ uint8_t s_rc[TPB / 64][REC_CAP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = c < n_comp;
    const uint32_t L = valid ? a.comp_len[c] : 0u;
    const uint64_t off = valid ? a.comp_off[c] : 0ull;
    // lane 0 holds the wave's first position (offsets grow with the lane)
    const uint64_t base = __shfl(off, 0, 64);
    uint64_t last = valid ? off + L : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(last, o, 64);
        last = y > last ? y : last;
    }
    const bool staged = last - base <= REC_CAP;  // the same for the whole wave
    const NodeRec<K>* rec = (const NodeRec<K>*)a.node_rec;
    // one record gather per position: the next node of a long path comes
    // from the record of the current one
    auto put = [&](uint32_t i, const NodeRec<K>& r, uint32_t node, uint8_t o) {
        if (staged) {
            const uint32_t j = uint32_t(off - base) + i;
            s_key[wave][j] = r.key;
            s_cid[wave][j] = r.cid;
            s_rank[wave][j] = r.rank;
            s_info[wave][j] = r.mm;
            s_rc[wave][j] = o;
        } else {
            const uint64_t pos = off + i;
            ((K*)a.pkey)[pos] = r.key;
            a.pcid[pos] = r.cid;
            a.prank[pos] = r.rank;
            a.pinfo[pos] = r.mm;
            a.prc[pos] = o;
            a.fflag[pos] = 0;
        }
        if (a.mark_visited) a.visited[node] = 1;
    };
    // A ranked path's positions are written by k_long_place, which runs
    // after this kernel (a staged wave also stores the slots of such a
    // path).  k_walk_ends buffers the nodes of the short paths it walked to
    // the end only, so a ranked path of at most PATH_BUF nodes has none.
    const uint32_t cex = L > 0 ? a.comp_ex[c] : 0u;
    if (L > 0 && L <= PATH_BUF && !(cex & COMP_RANKED)) {
        for (uint32_t i = 0; i < L; ++i) {
            const uint32_t v = a.comp_nodes[c * PATH_BUF + i], node = v & 0x7FFFFFFFu;
            put(i, rec[node], node, uint8_t(v >> 31));
        }
    } else if (L > PATH_BUF && !(cex & COMP_RANKED)) {
        uint32_t cur = a.comp_x[c], ex = cex & 1u;
        uint8_t o = ex == PORT_L ? 1 : 0;
        for (uint32_t i = 0;; ++i) {
            const NodeRec<K> r = rec[cur];
            put(i, r, cur, o);
            if (i + 1 == L) break;
            const uint32_t l = ex ? r.link[1] : r.link[0];  // a constant index keeps r in registers
            const uint32_t q = l & 1u;
            cur = l >> 1;
            ex = 1u - q;
            o = q == PORT_R ? 1 : 0;
        }
    }
    if (!staged) return;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t total = uint32_t(last - base);
    K* pkey = (K*)a.pkey;
    for (uint32_t j = uint32_t(lane); j < total; j += 64) {
        pkey[base + j] = s_key[wave][j];
        a.pcid[base + j] = s_cid[wave][j];
        a.prank[base + j] = s_rank[wave][j];
        a.pinfo[base + j] = s_info[wave][j];
        a.prc[base + j] = s_rc[wave][j];
        a.fflag[base + j] = 0;
    }
}

// The seed loop of bucket_walker.hpp:451-570 on a recorded path: seeds are
// the unplaced positions in rank order; from a seed the fragment grows
// backward then forward along the path (positions -1 / +1 when the seed lies
// forward on the path, +1 / -1 otherwise), stopping before a placed position
// or after a contig-break node.
// The paths of a wave's components occupy one contiguous range of positions;
// when it holds at most FR_CAP positions their orientation / info / rank are
// staged in LDS (coalesced loads) and the seed loops run there, else on the
// global arrays.  Placement marks (bit 1 of the orientation byte) live where
// the loop runs; only the fragment descriptors are written out.
constexpr uint32_t FR_CAP = 512;
__global__ __launch_bounds__(TPB) void k_walk_frags(WalkArgs a, uint64_t n_comp) {
    __shared__ uint8_t s_rc[TPB / 64][FR_CAP];
    __shared__ uint16_t s_info[TPB / 64][FR_CAP];
    __shared__ uint32_t s_rk[TPB / 64][FR_CAP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = c < n_comp;
    const uint32_t L = valid ? a.comp_len[c] : 0u;
    const uint64_t off = valid ? a.comp_off[c] : 0ull;
    const uint64_t base = __shfl(off, 0, 64);  // offsets grow with the lane
    uint64_t last = valid ? off + L : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(last, o, 64);
        last = y > last ? y : last;
    }
    const bool staged = last - base <= FR_CAP;  // the same for the whole wave
    if (staged) {
        const uint32_t total = uint32_t(last - base);
        for (uint32_t j = uint32_t(lane); j < total; j += 64) {
            s_rc[wave][j] = a.prc[base + j];
            s_info[wave][j] = a.pinfo[base + j];
            s_rk[wave][j] = a.prank[base + j];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    if (L == 0 || (a.comp_ex[c] & COMP_RANKED)) return;  // ranked paths: k_walk_frags_long
    uint8_t* rc = staged ? &s_rc[wave][off - base] : a.prc + off;
    const uint16_t* info = staged ? &s_info[wave][off - base] : a.pinfo + off;
    const uint32_t* rk = staged ? &s_rk[wave][off - base] : a.prank + off;
    const uint32_t b = a.b0 + a.node_bucket[a.comp_x[c]];
    uint32_t placed = 0;
    while (placed < L) {
        uint32_t p = NONE32, best = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < L; ++i)
            if (!(rc[i] & 2u) && rk[i] < best) {
                best = rk[i];
                p = i;
            }
        const uint32_t o = rc[p] & 1u;
        const int fwd = o ? -1 : 1;
        const uint32_t sflags = info[p] & 3u, ms = info[p] >> 8;
        rc[p] |= 2u;
        int lo = int(p), hi = int(p);
        bool open_l = false, open_r = false;
        // backward, then forward; the port a node is left by in walking
        // direction d is its path exit port (R when forward on the path)
        // for d = +1 and its path entry port for d = -1
        for (int side = 0; side < 2; ++side) {
            const int d = side == 0 ? -fwd : fwd;
            bool& open = side == 0 ? open_l : open_r;
            if (sflags == (side == 0 ? BOUND_LEFT : BOUND_RIGHT)) {
                open = __popc(side == 0 ? (ms >> 4) : (ms & 0xfu)) <= 1;
                continue;
            }
            for (int i = int(p) + d; i >= 0 && i < int(L); i += d) {
                if (rc[i] & 2u) break;
                rc[i] |= 2u;
                if (d < 0)
                    lo = i;
                else
                    hi = i;
                const uint32_t mi = info[i];
                if (is_break(uint8_t(mi))) {
                    const uint32_t oi = rc[i] & 1u;
                    const uint32_t ex = (d > 0) == (oi == 0) ? PORT_R : PORT_L;
                    open = __popc(nib(mi >> 8, ex)) <= 1;
                    break;
                }
            }
        }
        const uint32_t n = uint32_t(hi - lo + 1);
        placed += n;
        const uint64_t f = off + uint32_t(lo);
        a.fflag[f] = 1;
        a.frag_n[f] = n;
        a.frag_dir[f] = uint8_t(o);
        a.frag_open[f] = uint8_t((open_l ? 1u : 0u) | (open_r ? 2u : 0u));
        a.frag_key[f] = (uint64_t(b) << 32) | best;
    }
}

// General seed loop for the queued paths that pass a node twice.
template <class K>
__global__ void k_walk_other(WalkArgs a, uint64_t n_other) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_other) return;
    const uint32_t x = a.other_x[c], ex0 = a.other_ex[c];
    uint32_t cur = x, ex = ex0, seed = x, seed_rank = a.rank[x];
    const uint32_t limit = 2 * a.n_nodes + 2;
    for (uint32_t h = 0; h < limit; ++h) {
        const uint32_t l = a.link[2 * cur + ex];
        if (l == NONE32) break;
        cur = l >> 1;
        ex = 1u - (l & 1u);
        const uint32_t r = a.rank[cur];
        if (r < seed_rank) {
            seed = cur;
            seed_rank = r;
        }
    }
    simulate_component<K>(a, x, ex0, false, 0, seed);
}

__global__ void k_set_u64(unsigned long long* p, uint64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

template <class K>
__global__ void k_walk_cycles(WalkArgs a) {
    const uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (x >= a.n_nodes || a.visited[x]) return;
    uint32_t cur = uint32_t(x), ex = PORT_R, len = 1, seed = uint32_t(x), seed_rank = a.rank[x];
    bool simple = true;
    const uint32_t limit = 2 * a.n_nodes + 2;
    for (uint32_t h = 0; h < limit; ++h) {
        const uint32_t l = a.link[2 * cur + ex];
        if (l == NONE32) {  // not a cycle: should have been a path
            atomicAdd(&a.counters[5], 1ull);
            return;
        }
        const uint32_t y = l >> 1, q = l & 1u;
        if (y == x && q == PORT_L) break;
        if (y == cur || y == x) simple = false;  // hairpin or bounce through x
        cur = y;
        ex = 1u - q;
        if (cur < x) return;
        ++len;
        const uint32_t r = a.rank[cur];
        if (r < seed_rank) {
            seed = cur;
            seed_rank = r;
        }
    }
    simulate_component<K>(a, uint32_t(x), PORT_R, true, simple ? len : 0, seed);
}

// ---------------------------------------------------------------------------
// List ranking of long paths (emit_from_seed's walk, bucket_walker.hpp:451-566,
// for paths longer than walk_cap links).
//
// The nodes U of the buckets that hold a long path get two half-edge states
// each, state 2u + p = "leaving U[u] by port p".  A linked port's successor
// is the neighbour's state leaving by its other port (distance 1); a port
// without a link is terminal.  Links are symmetric, so every state has at
// most one predecessor: the states form simple chains, one per path and
// direction, and rings (cycles, handled by the cycle loop).  Pointer jumping
// (next = next[next], dist += dist[next]) gives every chain state its
// terminal and its distance in ceil(log2(2|U|)) rounds, so the owning end
// learns its path's length and every node its position and orientation on
// its path, each in one parallel pass.  A hairpin link (a node linked to
// itself) is or-ed along the chain: such a path goes to the general loop.
// ---------------------------------------------------------------------------
constexpr uint32_t ST_TERM = 0x80000000u, ST_HAIR = 0x40000000u, ST_DIST = 0x3FFFFFFFu;

__global__ void k_long_nodes(uint32_t n_nodes, const uint32_t* __restrict__ node_bucket,
                             const uint8_t* __restrict__ long_bucket, uint8_t* __restrict__ nflag) {
    const uint64_t n = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n < n_nodes) nflag[n] = long_bucket[node_bucket[n]];
}

__global__ void k_long_index(uint32_t nu, const uint32_t* __restrict__ U, uint32_t* __restrict__ uidx,
                             uint32_t* __restrict__ own_e) {
    const uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    uidx[U[u]] = uint32_t(u);
    own_e[u] = NONE32;
}

__global__ void k_long_states(uint32_t nu, const uint32_t* __restrict__ U, const uint32_t* __restrict__ uidx,
                              const uint32_t* __restrict__ link, uint2* __restrict__ st,
                              unsigned long long* __restrict__ err) {
    const uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t v = U[u];
#pragma unroll
    for (uint32_t p = 0; p < 2; ++p) {
        const uint32_t l = link[2 * v + p], s = uint32_t(2 * u + p);
        const uint32_t y = l >> 1, q = l & 1u;
        // a neighbour lies in the node's bucket, so it is in U
        const uint32_t uy = l == NONE32 ? NONE32 : uidx[y];
        const bool in_u = uy < nu && U[uy] == y;
        if (l != NONE32 && !in_u) atomicAdd(err, 1ull);
        if (l == NONE32 || !in_u)
            st[s] = make_uint2(s, ST_TERM);
        else
            st[s] = make_uint2(2 * uy + (1u - q), 1u | (y == v ? ST_HAIR : 0u));
    }
}

__global__ void k_long_jump(uint64_t ns, const uint2* __restrict__ A, uint2* __restrict__ Bn) {
    const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const uint2 a = A[s];
    if (a.y & ST_TERM) {
        Bn[s] = a;
        return;
    }
    const uint2 b = A[a.x];
    uint32_t d = (a.y & ST_DIST) + (b.y & ST_DIST);
    if (d > ST_DIST) d = ST_DIST;  // rings only: their distances are not used
    Bn[s] = make_uint2(b.x, d | ((a.y | b.y) & ST_HAIR));
}

// Long ends: the owning end (the smaller index, as in k_walk_ends) takes the
// path's length; hairpin paths are queued for the general loop.  Every lane
// of a wave calls wave_alloc.
__global__ __launch_bounds__(TPB) void k_long_ends(WalkArgs a, const uint32_t* __restrict__ ends, uint32_t n_ends,
                                                   uint32_t nu, const uint32_t* __restrict__ U,
                                                   const uint32_t* __restrict__ uidx,
                                                   const uint2* __restrict__ st, uint32_t* __restrict__ own_e,
                                                   uint32_t* __restrict__ long_list) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool ranked = false, other = false;
    uint32_t x = 0, ex = 0, u = 0;
    if (e < n_ends) {
        ex = a.comp_ex[e];
        if (ex & COMP_LONG) {
            x = ends[e];
            u = uidx[x];  // the end's bucket is flagged, so the end is in U
            const bool in_u = u < nu && U[u] == x;
            const uint2 s = in_u ? st[2 * u + (ex & 1u)] : make_uint2(0u, 0u);
            if (!in_u || !(st[s.x].y & ST_TERM)) {
                atomicAdd(&a.counters[5], 1ull);  // a path end cannot lie on a ring
            } else if (U[s.x >> 1] >= x) {
                if (s.y & ST_HAIR) {
                    other = true;
                } else {
                    ranked = true;
                    a.comp_len[e] = (s.y & ST_DIST) + 1;
                    own_e[u] = uint32_t(e);
                }
            }
            a.comp_ex[e] = uint8_t((ex & 1u) | (ranked ? COMP_RANKED : 0u));
        }
    }
    const uint64_t o = wave_alloc(&a.counters[10], other);
    const uint64_t r = wave_alloc(&a.counters[14], ranked);
    if (other) {
        a.other_x[o] = x;
        a.other_ex[o] = uint8_t(ex & 1u);
    }
    if (ranked) long_list[r] = uint32_t(e);
}

// Every node of a ranked path writes its own position: its distance from
// the owning end, entered by L (orientation 0) when the owner lies behind
// its L port, by R (orientation 1) otherwise.
template <class K>
__global__ void k_long_place(WalkArgs a, uint32_t nu, const uint32_t* __restrict__ U, const uint2* __restrict__ st,
                             const uint32_t* __restrict__ own_e) {
    const uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint2 sl = st[2 * u + PORT_L], sr = st[2 * u + PORT_R];
    if (!(st[sl.x].y & ST_TERM) || !(st[sr.x].y & ST_TERM)) return;  // ring
    const uint32_t zl = U[sl.x >> 1], zr = U[sr.x >> 1];
    if (zl == zr) return;  // one-node path, or a path turned back by a hairpin
    const bool via_l = zl < zr;
    const uint32_t e = own_e[via_l ? (sl.x >> 1) : (sr.x >> 1)];
    if (e == NONE32) return;  // a short path of the bucket: already recorded
    const uint32_t i = via_l ? (sl.y & ST_DIST) : (sr.y & ST_DIST);
    const uint32_t v = U[u];
    const uint64_t pos = uint64_t(a.comp_off[e]) + i;
    if (i >= a.comp_len[e] || pos >= a.n_nodes) {
        atomicAdd(&a.counters[5], 1ull);
        return;
    }
    put_position<K>(a, pos, v, via_l ? 0 : 1);
    if (a.mark_visited) a.visited[v] = 1;
}

// The seed loop of k_walk_frags on a ranked path, one wave per path: the
// smallest unplaced rank by a wave reduction, each side's growth 64
// positions per step (the first placed position or path end stops before,
// the first contig break stops after).  Placement marks live in prc; the
// fences make one lane's marks visible to the others before they are read.
__global__ __launch_bounds__(TPB) void k_walk_frags_long(WalkArgs a, const uint32_t* __restrict__ list,
                                                         uint32_t n_list) {
    const int lane = threadIdx.x & 63;
    const uint64_t w = uint64_t(blockIdx.x) * (TPB / 64) + (threadIdx.x >> 6);
    if (w >= n_list) return;  // the whole wave
    const uint32_t c = list[w];
    const uint32_t L = a.comp_len[c];
    const uint64_t off = a.comp_off[c];
    uint8_t* rc = a.prc + off;
    const uint16_t* info = a.pinfo + off;
    const uint32_t* rk = a.prank + off;
    const uint32_t b = a.b0 + a.node_bucket[a.comp_x[c]];
    uint32_t placed = 0;
    while (placed < L) {
        uint64_t best = ~0ull;  // rank << 32 | position: the smallest rank, then the first position
        for (uint32_t i = uint32_t(lane); i < L; i += 64)
            if (!(rc[i] & 2u)) {
                const uint64_t key = (uint64_t(rk[i]) << 32) | i;
                best = key < best ? key : best;
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint64_t y = __shfl_xor(best, o, 64);
            best = y < best ? y : best;
        }
        if (best == ~0ull) {  // cannot happen while positions are unplaced
            if (lane == 0) atomicAdd(&a.counters[5], 1ull);
            return;
        }
        const int64_t p = int64_t(uint32_t(best));
        const uint32_t o = rc[p] & 1u;
        const int64_t fwd = o ? -1 : 1;
        const uint32_t sflags = info[p] & 3u, ms = info[p] >> 8;
        int64_t lo = p, hi = p;
        bool open_l = false, open_r = false;
        for (int side = 0; side < 2; ++side) {
            const int64_t d = side == 0 ? -fwd : fwd;
            bool open = false;
            if (sflags == (side == 0 ? BOUND_LEFT : BOUND_RIGHT)) {
                open = __popc(side == 0 ? (ms >> 4) : (ms & 0xfu)) <= 1;
            } else {
            for (int64_t step = 1;; step += 64) {
                const int64_t i = p + d * (step + lane);
                const bool inb = i >= 0 && i < int64_t(L);
                const uint32_t r = inb ? rc[i] : 2u;
                const bool before = !inb || (r & 2u);
                uint32_t mi = 0;
                if (!before) mi = info[i];
                const bool brk = !before && is_break(uint8_t(mi));
                const uint64_t mb = __ballot(before), mk = __ballot(brk);
                const int jb = mb ? __ffsll((unsigned long long)mb) - 1 : 64;
                const int jk = mk ? __ffsll((unsigned long long)mk) - 1 : 64;
                const int last = jk < jb ? jk : jb - 1;  // last lane whose position joins the fragment
                if (lane <= last) rc[i] = uint8_t(r | 2u);
                if (last >= 0) {
                    const int64_t idx = p + d * (step + last);
                    if (d < 0)
                        lo = idx;
                    else
                        hi = idx;
                }
                if (jk < jb) {
                    uint32_t ov = 0;
                    if (lane == jk) {
                        const uint32_t ex = (d > 0) == ((r & 1u) == 0) ? PORT_R : PORT_L;
                        ov = __popc(nib(mi >> 8, ex)) <= 1 ? 1u : 0u;
                    }
                    open = __shfl(ov, jk, 64) != 0;
                    break;
                }
                if (jb < 64) break;
            }
            }
            if (side == 0)
                open_l = open;
            else
                open_r = open;
        }
        if (lane == 0) {
            rc[p] = uint8_t(rc[p] | 2u);
            const uint64_t f = off + uint64_t(lo);
            a.fflag[f] = 1;
            a.frag_n[f] = uint32_t(hi - lo + 1);
            a.frag_dir[f] = uint8_t(o);
            a.frag_open[f] = uint8_t((open_l ? 1u : 0u) | (open_r ? 2u : 0u));
            a.frag_key[f] = (uint64_t(b) << 32) | uint32_t(best >> 32);
        }
        placed += uint32_t(hi - lo + 1);
        __threadfence();
        __builtin_amdgcn_wave_barrier();
    }
}

// Sort keys of the fragments: forder[d] = the d-th fragment's first position
// (ascending), fkey[d] = its (bucket, seed rank) key.
__global__ void k_frag_keys(uint64_t nf, const uint32_t* __restrict__ forder, const uint64_t* __restrict__ frag_key,
                            uint64_t* __restrict__ fkey) {
    const uint64_t d = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (d < nf) fkey[d] = frag_key[forder[d]];
}

// The same with 32-bit sort keys (local bucket << rb) | seed rank, for a chunk
// whose buckets and ranks fit them (a rank is below its bucket's tuple count,
// which is below 2^rb); a rank that does not fit is counted in *overflow.
__global__ void k_frag_keys32(uint64_t nf, const uint32_t* __restrict__ forder,
                              const uint64_t* __restrict__ frag_key, uint32_t b0, int rb,
                              uint32_t* __restrict__ fkey, unsigned long long* __restrict__ overflow) {
    const uint64_t d = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (d >= nf) return;
    const uint64_t key = frag_key[forder[d]];
    const uint32_t r = uint32_t(key);
    if (rb < 32 && (r >> rb)) atomicAdd(overflow, 1ull);
    fkey[d] = uint32_t((uint64_t(uint32_t(key >> 32) - b0) << rb) | r);
}

// ---------------------------------------------------------------------------
// fragment records
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put_bytes(uint8_t* p, uint64_t v, int n) {
    for (int i = 0; i < n; ++i) p[i] = uint8_t(v >> (8 * i));
}

// `[u8 open][u32 nruns]{[u64 cid][u32 n]}[u32 seq_len][seq]`, the node range
// [o, o + n) read forwards or (rev) backwards with orientations flipped.
template <class K>
__device__ void write_record(uint8_t* p, uint8_t open, uint64_t o, uint32_t n, uint32_t rev,
                             const uint8_t* __restrict__ frc, const K* __restrict__ pkey,
                             const uint32_t* __restrict__ pcid, uint32_t k) {
    using O = KOps<K>;
    p[0] = open;
    uint8_t* pr = p + 5;
    uint32_t runs = 0, last = 0, cnt = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t c = pcid[o + (rev ? n - 1 - j : j)];
        if (j > 0 && c != last) {
            put_bytes(pr, last == NONE32 ? ~0ull : uint64_t(last), 8);
            put_bytes(pr + 8, cnt, 4);
            pr += 12;
            ++runs;
            cnt = 0;
        }
        last = c;
        ++cnt;
    }
    put_bytes(pr, last == NONE32 ? ~0ull : uint64_t(last), 8);
    put_bytes(pr + 8, cnt, 4);
    pr += 12;
    ++runs;
    put_bytes(p + 1, runs, 4);
    const uint32_t seq_len = k + n - 1;
    put_bytes(pr, seq_len, 4);
    uint8_t* ps = pr + 4;
    const char ACGT[4] = {'A', 'C', 'G', 'T'};
    for (uint32_t j = 0; j < n; ++j) {
        const uint64_t at = o + (rev ? n - 1 - j : j);
        const K key = pkey[at];
        const K km = ((frc[at] & 1u) ^ rev) ? O::rc(key, k) : key;
        if (j == 0) {
            for (uint32_t t = 0; t < k; ++t) ps[t] = ACGT[O::base(km, t, k)];
        } else {
            ps[k + j - 1] = ACGT[O::base(km, k - 1, k)];
        }
    }
}

// The same record from positions staged in LDS: per position its colour-set
// id and pb = orientation | first base << 1 | last base << 3 of its k-mer;
// only the fragment's first k-mer in output order (`first`, at position
// o + (rev ? n - 1 : 0)) is read whole.  The last base of the k-mer read
// reverse-complemented is the complement (3 - b) of its first base.
template <class K>
__device__ __forceinline__ void write_record_lds(uint8_t* p, uint8_t open, uint32_t o, uint32_t n, uint32_t rev,
                                                 const uint32_t* s_cid, const uint8_t* s_pb, const K& first,
                                                 uint32_t k) {
    using O = KOps<K>;
    p[0] = open;
    uint8_t* pr = p + 5;
    uint32_t runs = 0, last = 0, cnt = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t c = s_cid[o + (rev ? n - 1 - j : j)];
        if (j > 0 && c != last) {
            put_bytes(pr, last == NONE32 ? ~0ull : uint64_t(last), 8);
            put_bytes(pr + 8, cnt, 4);
            pr += 12;
            ++runs;
            cnt = 0;
        }
        last = c;
        ++cnt;
    }
    put_bytes(pr, last == NONE32 ? ~0ull : uint64_t(last), 8);
    put_bytes(pr + 8, cnt, 4);
    pr += 12;
    ++runs;
    put_bytes(p + 1, runs, 4);
    const uint32_t seq_len = k + n - 1;
    put_bytes(pr, seq_len, 4);
    uint8_t* ps = pr + 4;
    const char ACGT[4] = {'A', 'C', 'G', 'T'};
    const uint32_t o0 = o + (rev ? n - 1 : 0);
    const K km = ((s_pb[o0] & 1u) ^ rev) ? O::rc(first, k) : first;
    for (uint32_t t = 0; t < k; ++t) ps[t] = ACGT[O::base(km, t, k)];
    for (uint32_t j = 1; j < n; ++j) {
        const uint32_t pb = s_pb[o + (rev ? n - 1 - j : j)];
        const uint32_t lb = ((pb & 1u) ^ rev) ? 3u - ((pb >> 1) & 3u) : (pb >> 3) & 3u;
        ps[k + j - 1] = ACGT[lb];
    }
}

// Record bytes (unitig_spill.hpp:491-522) are staged in LDS up to fw_lds<K>
// per block, the positions the block's records read up to fw_pos<K>.  For
// 64-bit keys 16 KB + 1024 positions (~25 KB per block, six blocks per CU)
// measured fastest: 24 KB + 2048 positions was 20% slower (four blocks per
// CU), a 12 KB record stage twice as slow (most blocks no longer fit and write
// their records byte by byte).  Records of 128-bit keys' k-mers are longer:
// their stage keeps 32 KB.
template <class K>
constexpr uint32_t fw_lds() { return sizeof(K) == 8 ? 16384 : 32768; }
template <class K>
constexpr uint32_t fw_pos() { return sizeof(K) == 8 ? 1024 : 512; }

// Fragments in position order (d = the d-th fragment start, ascending): the
// record size of each (read from consecutive positions) and d itself, the
// value the output sort carries.
// Fragments of more than LONG_FRAG k-mers (a unitig that spans a bucket's
// whole graph, with few buckets) are listed instead and sized and written by
// a block each (k_frag_psize_long, k_frag_write_long).
constexpr uint32_t LONG_FRAG = 4096;

__global__ void k_frag_psize(uint64_t nf, const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ frag_n,
                             const uint32_t* __restrict__ pcid, uint32_t k, uint32_t* __restrict__ psize,
                             uint32_t* __restrict__ dval, uint32_t* __restrict__ long_list,
                             unsigned long long* __restrict__ n_long) {
    const uint64_t d = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = d < nf;
    const uint64_t o = live ? fpos[d] : 0ull;
    const uint32_t n = live ? frag_n[o] : 0u;
    const bool lng = n > LONG_FRAG;
    const uint64_t li = wave_alloc(n_long, lng);  // every lane of the wave calls it
    if (!live) return;
    dval[d] = uint32_t(d);
    if (lng) {
        long_list[li] = uint32_t(d);
        return;
    }
    uint32_t runs = 1;  // run count is the same read either way
    for (uint32_t j = 1; j < n; ++j) runs += pcid[o + j] != pcid[o + j - 1];
    psize[d] = 1 + 4 + 12 * runs + 4 + (k + n - 1);
}

__global__ __launch_bounds__(TPB) void k_frag_psize_long(const uint32_t* __restrict__ long_list,
                                                         const unsigned long long* __restrict__ n_long,
                                                         const uint32_t* __restrict__ fpos,
                                                         const uint32_t* __restrict__ frag_n,
                                                         const uint32_t* __restrict__ pcid, uint32_t k,
                                                         uint32_t* __restrict__ psize) {
    __shared__ uint32_t s_tmp[TPB / 64 + 1];
    const uint64_t nl = *n_long;
    for (uint64_t li = blockIdx.x; li < nl; li += gridDim.x) {  // the same trip count for the whole block
        const uint32_t d = long_list[li];
        const uint64_t o = fpos[d];
        const uint32_t n = frag_n[o];
        uint32_t c = 0;
        for (uint32_t j = 1 + threadIdx.x; j < n; j += TPB) c += pcid[o + j] != pcid[o + j - 1];
        uint32_t changes;
        (void)block_exclusive_scan<uint32_t>(c, s_tmp, changes);
        if (threadIdx.x == 0) psize[d] = 1 + 4 + 12 * (changes + 1) + 4 + (k + n - 1);
    }
}

// Block-wide exclusive max over threads (identity -1); total = the block's max.
__device__ inline int64_t block_exclusive_max(int64_t v, int64_t* s_tmp, int64_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x = y > x ? y : x;
    }
    int64_t ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = -1;
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = -1;
        for (int w = 0; w < nw; ++w) {
            const int64_t t = s_tmp[w];
            s_tmp[w] = run;
            run = t > run ? t : run;
        }
        s_tmp[nw] = run;
    }
    __syncthreads();
    const int64_t before = s_tmp[wave];
    total = s_tmp[nw];
    __syncthreads();
    return before > ex ? before : ex;
}

// One block per long fragment: the sequence (first k-mer whole, then the
// last base of each following k-mer) by all threads, the colour runs in
// tiles of TPB * LF_ITEMS k-mers: a thread's run index and the start of its
// current run come from a block sum of run starts and a block max of start
// positions, carried from tile to tile.
constexpr uint32_t LF_ITEMS = 8;
template <class K>
__device__ void write_long_one(uint32_t d, const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ frag_n,
                               const uint8_t* __restrict__ frag_open, const uint8_t* __restrict__ frag_dir,
                               const uint8_t* __restrict__ frc, const K* __restrict__ pkey,
                               const uint32_t* __restrict__ pcid, uint32_t k, const uint32_t* __restrict__ psize,
                               const uint64_t* __restrict__ pboff, uint8_t* __restrict__ out, uint32_t* s_sum,
                               int64_t* s_max);
template <class K>
__global__ __launch_bounds__(TPB) void k_frag_write_long(const uint32_t* __restrict__ long_list,
                                                         const unsigned long long* __restrict__ n_long,
                                                         const uint32_t* __restrict__ fpos,
                                                         const uint32_t* __restrict__ frag_n,
                                                         const uint8_t* __restrict__ frag_open,
                                                         const uint8_t* __restrict__ frag_dir,
                                                         const uint8_t* __restrict__ frc, const K* __restrict__ pkey,
                                                         const uint32_t* __restrict__ pcid, uint32_t k,
                                                         const uint32_t* __restrict__ psize,
                                                         const uint64_t* __restrict__ pboff,
                                                         uint8_t* __restrict__ out) {
    __shared__ uint32_t s_sum[TPB / 64 + 1];
    __shared__ int64_t s_max[TPB / 64 + 1];
    const uint64_t nl = *n_long;
    for (uint64_t li = blockIdx.x; li < nl; li += gridDim.x)  // the same trip count for the whole block
        write_long_one<K>(long_list[li], fpos, frag_n, frag_open, frag_dir, frc, pkey, pcid, k, psize, pboff, out,
                          s_sum, s_max);
}

template <class K>
__device__ void write_long_one(uint32_t d, const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ frag_n,
                               const uint8_t* __restrict__ frag_open, const uint8_t* __restrict__ frag_dir,
                               const uint8_t* __restrict__ frc, const K* __restrict__ pkey,
                               const uint32_t* __restrict__ pcid, uint32_t k, const uint32_t* __restrict__ psize,
                               const uint64_t* __restrict__ pboff, uint8_t* __restrict__ out, uint32_t* s_sum,
                               int64_t* s_max) {
    using O = KOps<K>;
    const uint64_t o = fpos[d];
    const uint32_t n = frag_n[o], rev = frag_dir[o];
    const uint32_t seq_len = k + n - 1;
    const uint32_t runs = (psize[d] - (1 + 4 + 4 + seq_len)) / 12;
    uint8_t* p = out + pboff[d];
    uint8_t* pr = p + 5;
    uint8_t* ps = pr + 12 * uint64_t(runs) + 4;
    const char ACGT[4] = {'A', 'C', 'G', 'T'};
    auto at = [&](uint32_t j) { return o + (rev ? n - 1 - j : j); };
    if (threadIdx.x == 0) {
        p[0] = frag_open[o];
        put_bytes(p + 1, runs, 4);
        put_bytes(ps - 4, seq_len, 4);
    }
    if (threadIdx.x < k) {
        const uint64_t a0 = at(0);
        const K key = pkey[a0];
        const K km = ((frc[a0] & 1u) ^ rev) ? O::rc(key, k) : key;
        ps[threadIdx.x] = ACGT[O::base(km, threadIdx.x, k)];
    }
    for (uint32_t j = 1 + threadIdx.x; j < n; j += TPB) {
        const uint64_t a = at(j);
        const K key = pkey[a];
        const K km = ((frc[a] & 1u) ^ rev) ? O::rc(key, k) : key;
        ps[k + j - 1] = ACGT[O::base(km, k - 1, k)];
    }
    uint32_t carry_runs = 0;    // runs started before the tile
    int64_t carry_start = -1;   // the last run start before the tile
    for (uint32_t t0 = 0; t0 < n; t0 += TPB * LF_ITEMS) {
        const uint32_t j0 = t0 + threadIdx.x * LF_ITEMS;
        uint32_t c[LF_ITEMS + 2];  // cids of j0 - 1 .. j0 + LF_ITEMS
#pragma unroll
        for (uint32_t i = 0; i < LF_ITEMS + 2; ++i) {
            const int64_t j = int64_t(j0) + int64_t(i) - 1;
            c[i] = (j >= 0 && j < int64_t(n)) ? pcid[at(uint32_t(j))] : NONE32;
        }
        uint32_t nstart = 0;
        int64_t last = -1;
#pragma unroll
        for (uint32_t i = 0; i < LF_ITEMS; ++i) {
            const uint32_t j = j0 + i;
            if (j < n && (j == 0 || c[i + 1] != c[i])) {
                ++nstart;
                last = j;
            }
        }
        uint32_t tile_runs;
        const uint32_t ex = block_exclusive_scan<uint32_t>(nstart, s_sum, tile_runs) + carry_runs;
        int64_t tile_max;
        const int64_t mx = block_exclusive_max(last, s_max, tile_max);
        int64_t r = int64_t(ex) - 1;  // index of the run the thread's first k-mer continues
        int64_t cur = mx > carry_start ? mx : carry_start;
#pragma unroll
        for (uint32_t i = 0; i < LF_ITEMS; ++i) {
            const uint32_t j = j0 + i;
            if (j >= n) break;
            const uint32_t cid = c[i + 1];
            if (j == 0 || cid != c[i]) {
                ++r;
                cur = j;
                put_bytes(pr + 12 * r, cid == NONE32 ? ~0ull : uint64_t(cid), 8);
            }
            if (j + 1 == n || c[i + 2] != cid) put_bytes(pr + 12 * r + 8, uint32_t(int64_t(j) - cur + 1), 4);
        }
        carry_runs += tile_runs;
        carry_start = tile_max > carry_start ? tile_max : carry_start;
    }
}

// osize[i] = size of the i-th fragment in output order
__global__ void k_frag_osize(uint64_t nf, const uint32_t* __restrict__ dval, const uint32_t* __restrict__ psize,
                             uint64_t* __restrict__ osize) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nf) osize[i] = psize[dval[i]];
}

// pboff[d] = output offset of the d-th fragment in position order
__global__ void k_frag_pboff(uint64_t nf, const uint32_t* __restrict__ dval, const uint64_t* __restrict__ boff,
                             uint64_t* __restrict__ pboff) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nf) pboff[dval[i]] = boff[i];
}

// A block builds the records of TPB consecutive fragments in position order
// in LDS, then each wave copies whole records, one at a time with its 64
// lanes, to their output offsets.  The fragments tile one contiguous range of
// positions; when it holds at most FW_POS positions, their colour-set ids and
// orientation / end bases are first loaded into LDS with coalesced loads.
// Blocks whose records exceed the LDS stage write them directly.
template <class K>
__global__ __launch_bounds__(TPB) void k_frag_write_pos(uint64_t nf, const uint32_t* __restrict__ fpos,
                                                        const uint32_t* __restrict__ frag_n,
                                                        const uint8_t* __restrict__ frag_open,
                                                        const uint8_t* __restrict__ frag_dir,
                                                        const uint8_t* __restrict__ frc, const K* __restrict__ pkey,
                                                        const uint32_t* __restrict__ pcid, uint32_t k,
                                                        const uint32_t* __restrict__ psize,
                                                        const uint64_t* __restrict__ pboff, uint8_t* __restrict__ out,
                                                        uint32_t pos_cap) {
    using O = KOps<K>;
    constexpr uint32_t FW_LDS = fw_lds<K>(), FW_POS = fw_pos<K>();
    __shared__ uint8_t s_out[FW_LDS];
    __shared__ uint32_t s_cid[FW_POS];
    __shared__ uint8_t s_pb[FW_POS];
    __shared__ uint32_t s_loff[TPB], s_size[TPB];
    __shared__ uint64_t s_dst[TPB];
    __shared__ uint32_t s_tmp[TPB / 64 + 1];
    const uint64_t d0 = uint64_t(blockIdx.x) * TPB, d = d0 + threadIdx.x;
    const bool live = d < nf;
    // the block's positions: from its first fragment's start to its last one's end
    const uint64_t dl = (nf < d0 + TPB ? nf : d0 + TPB) - 1;
    const uint32_t p0 = fpos[d0], pl = fpos[dl];
    const uint32_t np = pl + frag_n[pl] - p0;
    const bool pstaged = np <= FW_POS && np <= pos_cap;
    if (pstaged)
        for (uint32_t i = threadIdx.x; i < np; i += TPB) {
            const K key = pkey[p0 + i];
            s_cid[i] = pcid[p0 + i];
            s_pb[i] = uint8_t((frc[p0 + i] & 1u) | (O::base(key, 0, k) << 1) | (O::base(key, k - 1, k) << 3));
        }
    const uint32_t sz = live ? psize[d] : 0u;
    uint32_t total;
    const uint32_t loff = block_exclusive_scan<uint32_t>(sz, s_tmp, total);  // its barriers publish the stage
    const bool staged = total <= FW_LDS;
    if (live) {
        const uint32_t f = fpos[d];
        const uint64_t dst = pboff[d];
        uint8_t* p = staged ? s_out + loff : out + dst;
        const uint32_t n = frag_n[f], rev = frag_dir[f];
        if (pstaged)
            write_record_lds<K>(p, frag_open[f], f - p0, n, rev, s_cid, s_pb, pkey[f + (rev ? n - 1 : 0)], k);
        else if (n <= LONG_FRAG)  // (a block with a long fragment is never staged)
            write_record<K>(p, frag_open[f], f, n, rev, frc, pkey, pcid, k);
        s_loff[threadIdx.x] = loff;
        s_size[threadIdx.x] = n <= LONG_FRAG ? sz : 0u;  // a long one is written by k_frag_write_long
        s_dst[threadIdx.x] = dst;
    } else {
        s_size[threadIdx.x] = 0;
    }
    if (!staged) return;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int r = wave; r < TPB; r += TPB / 64) {
        const uint32_t n = s_size[r];
        if (n == 0) continue;
        const uint8_t* src = s_out + s_loff[r];
        uint8_t* dst = out + s_dst[r];
        for (uint32_t j = uint32_t(lane); j < n; j += 64) dst[j] = src[j];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
// An instantiated graph of one chunk's stage and the arguments it was
// captured with.
struct StageGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    ~StageGraph() { reset(); }
};

struct Phase2Bufs {
    std::map<uint32_t, std::unique_ptr<StageGraph>> graphs;  // cell-sort stage per chunk (first bucket)
    StageGraph& cell_graph(uint32_t b0) {
        auto& g = graphs[b0];
        if (!g) g.reset(new StageGraph());
        return *g;
    }
    // job-wide
    DevBuf kcnt, tup_off, btup, total, lbtup;
    DevBuf single_min, single_cid, mt_hk, mt_cid, mt_h2, mt_loff, mt_llen;
    DevBuf set_off, cs_words, frag_bytes;
    uint64_t mt_cap = 0, mt_count = 0;
    // per chunk (reused)
    DevBuf keys, vals, keys_t, vals_t, block_super, ckcnt, ctup_off;
    DevBuf cell_start, cell_base, cell_dbits, cell_mid, cell_mid2, cell_over, cell_novf, cell_oflag, cell_oscan, fb_start, fb_pre, fb_k, fb_v, fb_kt, fb_vt;
    DevBuf heads, tile_heads, node_start, node_off;
    DevBuf node_key, node_meta, node_rank, node_ncol, node_hash, node_hash2, big, big_flag, big_cnt, colors, counters;
    DevBuf node_slot, set_fl, set_pos, set_key, set_desc, set_key_t, set_desc_t, set_bits, set_boff;
    DevBuf node_bucket, dir, filt, node_mask, nb, link, visited;
    DevBuf frag_n, frag_open, frag_key, frag_dir, fflag, fcnt, fkey;
    DevBuf comp_ex, comp_len, comp_nodes, comp_off, other_x, other_ex, prc, pinfo, prank, pkey, pcid;
    DevBuf node_rec, eflag, ecnt, ends;
    DevBuf lr_bucket, lr_u, lr_uidx, lr_own, lr_st0, lr_st1, lr_list, lr_cnt;  // list ranking of long paths
    DevBuf flong;                                                            // fragments of > LONG_FRAG k-mers
    DevBuf fkey_t, forder, forder_t, fdval, fpsize, fpboff, fsize, fboff;
    RadixScratch rs2;
    // bytes of the per-chunk buffers (what a new chunk reuses)
    uint64_t chunk_bytes() const {
        uint64_t b = 0;
        for (const DevBuf* x : {&keys, &vals, &keys_t, &vals_t, &block_super, &ckcnt, &ctup_off, &cell_start, &cell_mid,
                                &cell_mid2, &cell_over, &cell_oflag, &cell_oscan, &fb_k, &fb_v, &fb_kt, &fb_vt, &heads,
                                &tile_heads, &node_start, &node_key, &node_meta, &node_rank, &node_ncol, &node_hash,
                                &node_hash2, &big, &big_flag, &colors, &node_slot, &set_fl, &set_pos, &set_key,
                                &set_desc, &set_key_t, &set_desc_t, &set_bits, &set_boff, &node_bucket, &dir, &filt,
                                &node_mask, &nb, &link, &visited, &frag_n, &frag_open, &frag_key, &frag_dir, &fflag,
                                &fkey, &comp_ex, &comp_len, &comp_nodes, &comp_off, &other_x, &other_ex, &prc, &pinfo,
                                &prank, &pkey, &pcid, &node_rec, &eflag, &ends, &fkey_t, &forder, &forder_t, &fdval,
                                &fpsize, &fpboff, &fsize, &fboff, &lr_bucket, &lr_u, &lr_uidx, &lr_own, &lr_st0,
                                &lr_st1, &lr_list, &flong})
            b += x->cap;
        return b;
    }
    // fragment stream kept in host memory, chunk by chunk, when the job's
    // projected stream does not fit beside its working set in HBM
    bool spill = false;
    PinnedBuf h_frags;
    // results
    uint64_t n_tuples = 0, n_nodes = 0, n_sets = 0, n_frags = 0, frag_total = 0, cs_bits = 0;
    uint64_t n_primary = 0, total_integers = 0, words_zeroed = 0;
};

template <class T>
static T d2h(const void* d, hipStream_t st) {
    T v;
    CDBG_HIP(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    return v;
}

static int bits_for(uint64_t n) {
    int b = 0;
    while ((uint64_t(1) << b) < n) ++b;
    return b;
}

static SetTables set_tables(Phase2Bufs& P) {
    SetTables T;
    T.single_min = P.single_min.as<unsigned long long>();
    T.single_cid = P.single_cid.as<uint32_t>();
    T.mt_hk = P.mt_hk.as<unsigned long long>();
    T.mt_cid = P.mt_cid.as<uint32_t>();
    T.mt_h2 = P.mt_h2.as<uint64_t>();
    T.mt_loff = P.mt_loff.as<uint64_t>();
    T.mt_llen = P.mt_llen.as<uint32_t>();
    T.cap = P.mt_cap;
    T.cur_colors = nullptr;
    return T;
}

// (Re)allocates the colour-set hash table with `cap` slots, keeping entries.
// The list offsets belong to the sets of the chunk about to run, which has
// not inserted any yet: they are not carried over (the sizes are).
static void mt_resize(Phase2Bufs& P, uint64_t cap, hipStream_t st) {
    DevBuf ohk, oc, oh2, ol;
    const uint64_t old_cap = P.mt_cap;
    ohk.swap(P.mt_hk);
    oc.swap(P.mt_cid);
    oh2.swap(P.mt_h2);
    ol.swap(P.mt_llen);
    P.mt_cap = cap;
    P.mt_hk.ensure<unsigned long long>(2 * cap);
    P.mt_cid.ensure<uint32_t>(cap);
    P.mt_h2.ensure<uint64_t>(cap);
    P.mt_loff.ensure<uint64_t>(cap);
    P.mt_llen.ensure<uint32_t>(cap);
    { KSCOPE("k_fill2", st); k_fill2<unsigned long long><<<blocks_for(cap), TPB, 0, st>>>(P.mt_hk.as<unsigned long long>(), cap, 0ull, ~0ull); }
    { KSCOPE("k_fill", st); k_fill<uint32_t><<<blocks_for(cap), TPB, 0, st>>>(P.mt_cid.as<uint32_t>(), cap, NONE32); }
    if (old_cap) {
        KSCOPE("k_mt_rehash", st);
        k_mt_rehash<<<blocks_for(old_cap), TPB, 0, st>>>(old_cap, ohk.as<unsigned long long>(), oc.as<uint32_t>(),
                                                          oh2.as<uint64_t>(), ol.as<uint32_t>(), set_tables(P));
    }
    CDBG_HIP(hipStreamSynchronize(st));
}

template <class K>
static void run_chunk(Job& J, Phase2Bufs& P, uint32_t b0, uint32_t b1, const std::vector<uint64_t>& h_btup) {
    hipStream_t st = J.st;
    Phase1Bufs& Q = J.p1;
    const uint32_t k = J.k, nb = b1 - b0;
    const uint64_t t0 = h_btup[b0], nt = h_btup[b1] - t0;
    const uint64_t s0 = J.h_bucket_off[b0], s1 = J.h_bucket_off[b1];
    if (nt == 0) return;
    if (nt >= (1ull << 32)) throw Error("phase 2: a chunk holds more than 2^32 k-mer occurrences");
    const int bbits = bits_for(J.B);
    // size of a job-wide stream at the job's end, extrapolated from its size
    // after this chunk by the share of occurrences processed (+5%); 0 after
    // the last chunk
    const uint64_t t_all = h_btup.back(), t_done = h_btup[b1];
    auto grown = [&](uint64_t n_now) -> size_t {
        if (t_done >= t_all || t_done == 0) return 0;
        return size_t(double(n_now) * double(t_all) / double(t_done) * 1.05);
    };

    // local bucket tuple offsets
    std::vector<uint64_t> lb(nb + 1);
    for (uint32_t b = 0; b <= nb; ++b) lb[b] = h_btup[b0 + b] - t0;
    uint64_t* lbtup = P.lbtup.ensure<uint64_t>(nb + 1);
    CDBG_HIP(hipMemcpyAsync(lbtup, lb.data(), (nb + 1) * 8, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipStreamSynchronize(st));

    // tuple offsets of the chunk's records: the job's array, or (jobs of
    // several chunks, which drop it) the chunk's own, indexed by record
    const uint64_t* tup_off = P.tup_off.as<uint64_t>();
    if (!tup_off) {
        const uint64_t nc = s1 - s0;
        uint32_t* ck = P.ckcnt.ensure<uint32_t>(nc);
        uint64_t* co = P.ctup_off.ensure<uint64_t>(nc + 1);
        { KSCOPE("k_count_kmers", st); k_count_kmers<<<blocks_for(nc), TPB, 0, st>>>(nc, Q.b_len.as<uint32_t>() + s0, k, ck); }
        exclusive_scan<uint32_t, uint64_t>(ck, co, nc, co + nc, J.rs.scan_tmp, st);
        { KSCOPE("k_add_u64", st); k_add_u64<<<blocks_for(nc + 1), TPB, 0, st>>>(co, nc + 1, t0); }
        tup_off = co - s0;
    }

    // --- expand
    K* keys = P.keys.ensure<K>(nt);
    uint64_t* vals = P.vals.ensure<uint64_t>(nt);
    K* keys_t = P.keys_t.ensure<K>(nt);
    uint64_t* vals_t = P.vals_t.ensure<uint64_t>(nt);
    ExpArgs ea{nt, t0, s0, s1, tup_off, P.btup.as<uint64_t>(), Q.key.as<uint32_t>(),
               Q.b_start.as<uint64_t>(), Q.b_len.as<uint32_t>(), Q.b_flags.as<uint8_t>(), Q.b_color.as<uint32_t>(),
               Q.packed.as<uint32_t>(), k, J.from_files ? Q.rank0.as<uint32_t>() : nullptr};
    const uint64_t n_eblocks = (nt + EXP_TILE - 1) / EXP_TILE;
    uint32_t* block_super = P.block_super.ensure<uint32_t>(n_eblocks + 1);
    { KSCOPE("k_block_super", st); k_block_super<<<blocks_for(s1 - s0), TPB, 0, st>>>(s0, s1, t0, n_eblocks, tup_off, block_super); }
    { KSCOPE("k_expand", st); k_expand<K><<<uint32_t(n_eblocks), TPB, 0, st>>>(ea, block_super, keys, vals); }

    // --- sort tuples by k-mer inside every bucket (stable): the top D_max
    // bits by one or two global passes, which cut every bucket b into 2^D_b
    // cells of 500 - 1000 keys with 64-bit keys, 350 - 700 with 128-bit
    // ones (measured best on MI355X against 350-700 and
    // 700-1400: larger cells amortise the per-cell passes, smaller ones fill
    // the CUs; an 11-bit single global pass scatters too slowly; D_b follows
    // the bucket's size, so skewed buckets get cells of the same size), then
    // the rest of each cell in LDS.  Cells sort the
    // bits below the shallowest bucket's prefix; the deeper buckets' extra
    // bits are constant in a cell or already in order, so sorting them again
    // changes nothing.  D_b >= D_max - 8 bounds the extra passes to one and
    // the cell count of small buckets; with 64-bit keys D_b >= 2k - 51 when
    // the largest bucket needs it, so the packed cell sort applies to all.
    const int kb = 2 * int(k);
    // mean keys per cell (the job's tuning; by default 1000 for 64-bit keys,
    // 700 for 128-bit ones, whose 1024-key LDS class larger cells overflow)
    const double cell_target = J.tune.cell_target > 0.0 ? J.tune.cell_target : (sizeof(K) == 8 ? 1000.0 : 700.0);
    std::vector<uint8_t> h_dbits(nb);
    int D_max = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        int d = 0;
        while (d < 20 && double(lb[b + 1] - lb[b]) / double(1ull << d) > cell_target) ++d;
        h_dbits[b] = uint8_t(std::min(d, kb));
        D_max = std::max(D_max, int(h_dbits[b]));
    }
    int D_floor = std::max(D_max - 8, 0);
    // the cells sort the bits below the shallowest bucket's cut: raise the
    // shallow buckets (up to the global passes' depth, at no extra pass) so
    // that the most size classes can pack their items, largest class first
    {
        const int wbits = int(8 * sizeof(K));
        const int pb[3] = {pos_bits<CellClasses<K>::LARGE>(), pos_bits<CellClasses<K>::SMALL>(),
                           pos_bits<CellClasses<K>::TINY>()};
        for (int p : pb) {
            const int need = kb - (wbits - p);
            if (D_max >= need) {
                D_floor = std::max(D_floor, need);
                break;
            }
        }
    }
    int D_min = kb;
    std::vector<uint32_t> h_cbase(nb + 1, 0);
    for (uint32_t b = 0; b < nb; ++b) {
        h_dbits[b] = uint8_t(std::max<int>(h_dbits[b], D_floor));
        D_min = std::min(D_min, int(h_dbits[b]));
        const uint64_t nc = uint64_t(h_cbase[b]) + (1ull << h_dbits[b]);
        if (nc >= (1ull << 32) - 1) throw Error("phase 2: too many sort cells in one chunk");
        h_cbase[b + 1] = uint32_t(nc);
    }
    SortPlan plan;
    plan.build(lb);
    if (D_max > 0) {
        // the fewest passes of at most digit_bits bits each, bits spread
        // evenly; a pass's tile and counters must fit the LDS (128-bit keys
        // allow 10 bits, 64-bit keys 11)
        int dmax = J.tune.digit_bits ? int(J.tune.digit_bits) : 8;
        while (dmax > 1 && radix_scatter_lds<K, uint64_t>(dmax) > 160 * 1024) --dmax;
        const int passes = (D_max + dmax - 1) / dmax;
        segmented_radix_sort<K, uint64_t>(keys, vals, keys_t, vals_t, plan, kb - D_max, kb,
                                          (D_max + passes - 1) / passes, P.rs2, st, "tuples");
    }
    const int D = D_min;  // the cells sort bits [0, kb - D_min)
    uint8_t* heads = P.heads.ensure<uint8_t>(nt);  // 1 at the first tuple of every node
    {
        const uint32_t n_cells = h_cbase[nb];
        uint32_t* cbase = P.cell_base.ensure<uint32_t>(uint64_t(nb) + 1);
        uint8_t* dbits = P.cell_dbits.ensure<uint8_t>(nb);
        CDBG_HIP(hipMemcpyAsync(cbase, h_cbase.data(), (uint64_t(nb) + 1) * 4, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipMemcpyAsync(dbits, h_dbits.data(), nb, hipMemcpyHostToDevice, st));
        uint32_t* cell_start = P.cell_start.ensure<uint32_t>(uint64_t(n_cells) + 1);
        { KSCOPE("k_cell_starts", st); k_cell_starts<K><<<blocks_for(uint64_t(n_cells) + 1), TPB, 0, st>>>(n_cells, nb, cbase, dbits, lbtup, nt, keys, kb, cell_start); }
        CDBG_HIP(hipStreamSynchronize(st));  // h_cbase / h_dbits are pageable
        // the tiny class over every cell, then each larger class over the
        // cells the previous one passed on; what is left goes to the global
        // passes
        constexpr uint32_t C0 = CellClasses<K>::TINY, C1 = CellClasses<K>::SMALL, C2 = CellClasses<K>::LARGE;
        uint32_t* mid = P.cell_mid.ensure<uint32_t>(n_cells);
        uint32_t* mid2 = P.cell_mid2.ensure<uint32_t>(n_cells);
        uint32_t* over = P.cell_over.ensure<uint32_t>(n_cells);
        auto* novf = P.cell_novf.ensure<unsigned long long>(6);  // listed cells of classes 0-2, queue heads 0-2
        uint32_t* oflag = P.cell_oflag.ensure<uint32_t>(n_cells);
        uint32_t* oscan = P.cell_oscan.ensure<uint32_t>(uint64_t(n_cells) + 1);
        const CellSortArgs a0{cell_start, nullptr, 0, kb - D, nullptr, novf, heads};
        const CellSortArgs a1{cell_start, mid, C0, kb - D, mid2, novf + 1, heads};
        const CellSortArgs a2{cell_start, mid2, C1, kb - D, over, novf + 2, heads};
        uint64_t n_over = 0;
        if (!J.tune.persistent) {
            { KSCOPE("k_fill", st); k_fill<unsigned long long><<<1, 64, 0, st>>>(novf, 3, 0ull); }
            { KSCOPE("k_cell_sort", st); launch_cell_sort<K, C0, 128>(st, keys, vals, a0, n_cells); }
            // cells above the tiny class, listed by a scan (many: no shared counter)
            { KSCOPE("k_cell_over", st); k_cell_over<<<blocks_for(n_cells), TPB, 0, st>>>(n_cells, cell_start, C0, oflag); }
            exclusive_scan<uint32_t, uint32_t>(oflag, oscan, n_cells, oscan + n_cells, J.rs.scan_tmp, st);
            const uint64_t n_mid = d2h<uint32_t>(oscan + n_cells, st);
            if (n_mid) { KSCOPE("k_compact", st); k_compact<<<blocks_for(n_cells), TPB, 0, st>>>(n_cells, oflag, oscan, mid); }
            if (n_mid) { KSCOPE("k_cell_sort", st); launch_cell_sort<K, C1, 256>(st, keys, vals, a1, uint32_t(n_mid)); }
            const uint64_t n_mid2 = n_mid ? d2h<unsigned long long>(novf + 1, st) : 0;
            if (n_mid2 > n_mid) throw Error("phase 2: cell sort listed more cells than exist");
            if (n_mid2) { KSCOPE("k_cell_sort", st); launch_cell_sort<K, C2, 512>(st, keys, vals, a2, uint32_t(n_mid2)); }
            n_over = n_mid2 ? d2h<unsigned long long>(novf + 2, st) : 0;
            if (n_over > n_mid2) throw Error("phase 2: cell sort listed more cells than exist");
        } else {
            // persistent classes over device cell counts: no read-back until
            // the cells too large for LDS are counted; with graphs on, the
            // stage is one graph launch, kept per chunk and replayed while
            // its arguments are unchanged (a repeated run of the job)
            J.rs.scan_tmp.ensure<uint32_t>((uint64_t(n_cells) + SCAN_CHUNK - 1) / SCAN_CHUNK + 2);
            cell_sort_warm<K, C0, 128>(kb - D);
            cell_sort_warm<K, C1, 256>(kb - D);
            cell_sort_warm<K, C2, 512>(kb - D);
            auto stage = [&]() {
                { KSCOPE("k_fill", st); k_fill<unsigned long long><<<1, 64, 0, st>>>(novf, 6, 0ull); }
                { KSCOPE("k_cell_sort", st); launch_cell_sort<K, C0, 128>(st, keys, vals, a0, n_cells, nullptr, nullptr, novf + 3); }
                { KSCOPE("k_cell_over", st); k_cell_over<<<blocks_for(n_cells), TPB, 0, st>>>(n_cells, cell_start, C0, oflag); }
                exclusive_scan<uint32_t, uint32_t>(oflag, oscan, n_cells, oscan + n_cells, J.rs.scan_tmp, st);
                { KSCOPE("k_compact", st); k_compact<<<blocks_for(n_cells), TPB, 0, st>>>(n_cells, oflag, oscan, mid); }
                { KSCOPE("k_cell_sort", st); launch_cell_sort<K, C1, 256>(st, keys, vals, a1, n_cells, oscan + n_cells, nullptr, novf + 4); }
                { KSCOPE("k_cell_sort", st); launch_cell_sort<K, C2, 512>(st, keys, vals, a2, n_cells, nullptr, novf + 1, novf + 5); }
            };
            if (J.tune.graph) {
                const std::vector<uint64_t> key{uint64_t(b0), n_cells, uint64_t(kb - D), uint64_t(uintptr_t(keys)),
                                                uint64_t(uintptr_t(vals)), uint64_t(uintptr_t(cell_start)),
                                                uint64_t(uintptr_t(mid)), uint64_t(uintptr_t(mid2)),
                                                uint64_t(uintptr_t(over)), uint64_t(uintptr_t(novf)),
                                                uint64_t(uintptr_t(heads)), uint64_t(uintptr_t(oflag)),
                                                uint64_t(uintptr_t(oscan)), uint64_t(uintptr_t(J.rs.scan_tmp.p))};
                StageGraph& g = P.cell_graph(b0);
                if (!g.exec || g.key != key) {
                    g.reset();
                    Profiler* saved = g_prof;  // no per-kernel events inside a capture
                    g_prof = nullptr;
                    hipGraph_t graph = nullptr;
                    CDBG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    try {
                        stage();
                    } catch (...) {
                        (void)hipStreamEndCapture(st, &graph);
                        if (graph) (void)hipGraphDestroy(graph);
                        g_prof = saved;
                        throw;
                    }
                    CDBG_HIP(hipStreamEndCapture(st, &graph));
                    g_prof = saved;
                    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    if (e != hipSuccess) {
                        g.exec = nullptr;
                        (void)hipGetLastError();
                        throw Error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                    }
                    g.key = key;
                }
                KSCOPE("k_cell_sort:graph", st);
                CDBG_HIP(hipGraphLaunch(g.exec, st));
            } else {
                stage();
            }
            n_over = d2h<unsigned long long>(novf + 2, st);
            if (n_over > n_cells) throw Error("phase 2: cell sort listed more cells than exist");
        }
        if (n_over) {
            // cells too large for LDS: gathered, sorted by the global passes, put back
            std::vector<uint32_t> ids(n_over), cs(n_cells + 1);
            CDBG_HIP(hipMemcpyAsync(ids.data(), over, n_over * 4, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipMemcpyAsync(cs.data(), cell_start, (uint64_t(n_cells) + 1) * 4, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipStreamSynchronize(st));
            std::sort(ids.begin(), ids.end());
            std::vector<uint32_t> st_l(n_over);
            std::vector<uint64_t> pre(n_over + 1, 0);
            for (uint64_t i = 0; i < n_over; ++i) {
                st_l[i] = cs[ids[i]];
                pre[i + 1] = pre[i] + (cs[ids[i] + 1] - cs[ids[i]]);
            }
            const uint64_t tot = pre[n_over];
            uint32_t* d_st = P.fb_start.ensure<uint32_t>(n_over);
            uint64_t* d_pre = P.fb_pre.ensure<uint64_t>(n_over + 1);
            CDBG_HIP(hipMemcpyAsync(d_st, st_l.data(), n_over * 4, hipMemcpyHostToDevice, st));
            CDBG_HIP(hipMemcpyAsync(d_pre, pre.data(), (n_over + 1) * 8, hipMemcpyHostToDevice, st));
            K* ck = P.fb_k.ensure<K>(tot);
            uint64_t* cv = P.fb_v.ensure<uint64_t>(tot);
            K* ckt = P.fb_kt.ensure<K>(tot);
            uint64_t* cvt = P.fb_vt.ensure<uint64_t>(tot);
            { KSCOPE("k_cells_move", st); k_cells_move<K, uint64_t><<<blocks_for(tot), TPB, 0, st>>>(true, d_st, d_pre, uint32_t(n_over), keys, vals, ck, cv); }
            SortPlan fplan;
            fplan.build(pre);
            segmented_radix_sort<K, uint64_t>(ck, cv, ckt, cvt, fplan, 0, kb - D, 8, P.rs2, st, "cells");
            { KSCOPE("k_cells_move", st); k_cells_move<K, uint64_t><<<blocks_for(tot), TPB, 0, st>>>(false, d_st, d_pre, uint32_t(n_over), keys, vals, ck, cv); }
            { KSCOPE("k_cells_heads", st); k_cells_heads<K><<<blocks_for(tot), TPB, 0, st>>>(d_st, d_pre, uint32_t(n_over), ck, heads); }
            CDBG_HIP(hipStreamSynchronize(st));  // st_l / pre are pageable
        }
    }

    // --- nodes
    // node numbers: heads counted per k_nodes tile and scanned over the tiles
    const uint64_t n_ntiles = (nt + NT_TILE - 1) / NT_TILE;
    uint32_t* tile_cnt = P.tile_heads.ensure<uint32_t>(2 * (n_ntiles + 1));
    uint32_t* tile_base = tile_cnt + n_ntiles + 1;
    uint64_t* dtot = P.total.ensure<uint64_t>(8);
    uint32_t* dtot32 = (uint32_t*)(dtot + 1);
    { KSCOPE("k_tile_heads", st); k_tile_heads<<<uint32_t(n_ntiles), TPB, 0, st>>>(nt, heads, tile_cnt); }
    exclusive_scan<uint32_t, uint32_t>(tile_cnt, tile_base, n_ntiles, dtot32, J.rs.scan_tmp, st);
    const uint32_t nn = d2h<uint32_t>(dtot32, st);
    if (nn >= (1u << 31)) throw Error("phase 2: more than 2^31 nodes in one chunk");
    uint32_t* node_start = P.node_start.ensure<uint32_t>(uint64_t(nn) + 1);
    uint64_t* node_off = P.node_off.ensure<uint64_t>(nb + 1);
    uint32_t* node_bucket = P.node_bucket.ensure<uint32_t>(nn);
    K* node_key = P.node_key.ensure<K>(nn);
    uint8_t* meta = P.node_meta.ensure<uint8_t>(nn);
    uint32_t* rank = P.node_rank.ensure<uint32_t>(nn);
    uint32_t* ncol = P.node_ncol.ensure<uint32_t>(nn);
    uint64_t* nhash = P.node_hash.ensure<uint64_t>(nn);
    uint64_t* nhash2 = P.node_hash2.ensure<uint64_t>(nn);
    uint32_t* big = P.big.ensure<uint32_t>(nn);
    uint32_t* colors = P.colors.ensure<uint32_t>(nt);  // node lists in place at their first occurrence
    const uint32_t* coff = node_start;
    unsigned long long* cnt = P.counters.ensure<unsigned long long>(N_COUNTERS);
    { KSCOPE("k_fill", st); k_fill<unsigned long long><<<1, 64, 0, st>>>(cnt, N_COUNTERS, 0ull); }
    uint8_t* big_flag = P.big_flag.ensure<uint8_t>(nn);
    { KSCOPE("k_nodes", st); k_nodes<K><<<uint32_t(n_ntiles), TPB, 0, st>>>(nt, heads, tile_base, lbtup, nb, node_bucket, node_start, keys, vals, node_key, meta, rank, ncol, nhash, nhash2, colors, big_flag); }
    { KSCOPE("k_set_u32", st); k_set_u32<<<1, 64, 0, st>>>(node_start + nn, uint32_t(nt)); }
    { KSCOPE("k_node_off", st); k_node_off<<<blocks_for(nb + 1), TPB, 0, st>>>(nb, lbtup, node_start, nn, node_off); }
    uint32_t* d_nbig = (uint32_t*)(dtot + 2);
    compact_flags<uint8_t>(big_flag, nn, big, d_nbig, P.big_cnt, J.rs.scan_tmp, st);
    const uint64_t nbig = d2h<uint32_t>(d_nbig, st);
    unsigned long long h_cnt[N_COUNTERS];
    if (nbig) { KSCOPE("k_nodes_big", st); k_nodes_big<<<blocks_for(nbig, TPB / 64), TPB, 0, st>>>(big, nbig, node_start, vals, meta, rank, ncol, nhash, nhash2, colors); }
    { KSCOPE("k_count_nodes", st); k_count_nodes<<<4096, TPB, 0, st>>>(nn, meta, ncol, cnt); }
    CDBG_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof h_cnt, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    const uint64_t n_primary = h_cnt[0], n_multi = h_cnt[1];

    // --- colour sets: intern into the job-wide tables
    uint64_t cap = P.mt_cap;
    const double slack = J.tune.set_slack > 0.0 ? J.tune.set_slack : 2.0;  // slots per multi-colour set
    while (double(cap) < slack * double(P.mt_count + n_multi)) cap <<= 1;
    if (cap != P.mt_cap) mt_resize(P, cap, st);
    SetTables T = set_tables(P);
    T.cur_colors = colors;
    uint32_t* node_slot = P.node_slot.ensure<uint32_t>(nn);
    uint8_t* created = P.set_fl.ensure<uint8_t>(nn);  // 1: the node created its set in this chunk
    { KSCOPE("k_set_insert", st); k_set_insert<<<blocks_for(nn), TPB, 0, st>>>(nn, meta, ncol, coff, colors, nhash, rank, node_bucket, b0, T, nhash2, node_slot, created); }
    { KSCOPE("k_set_verify", st); k_set_verify<<<blocks_for(nn), TPB, 0, st>>>(nn, meta, ncol, nhash2, node_slot, T, cnt + 6); }
    const uint32_t NC = J.num_colors;
    uint32_t* set_list = P.set_pos.ensure<uint32_t>(nn);
    uint32_t* d_new = (uint32_t*)(dtot + 3);
    compact_flags<uint8_t>(created, nn, set_list, d_new, P.set_bits, J.rs.scan_tmp, st);
    const uint64_t n_new = d2h<uint32_t>(d_new, st);
    uint64_t* set_key = P.set_key.ensure<uint64_t>(n_new);
    uint64_t* set_desc = P.set_desc.ensure<uint64_t>(n_new);
    uint64_t* set_key_t = P.set_key_t.ensure<uint64_t>(n_new);
    uint64_t* set_desc_t = P.set_desc_t.ensure<uint64_t>(n_new);
    if (n_new) { KSCOPE("k_newset_emit", st); k_newset_emit<<<blocks_for(n_new), TPB, 0, st>>>(n_new, set_list, ncol, coff, colors, node_slot, T, set_key, set_desc); }
    SortPlan splan;
    splan.build({0, n_new});
    segmented_radix_sort<uint64_t, uint64_t>(set_key, set_desc, set_key_t, set_desc_t, splan, 0, 32 + bbits, 8, P.rs2,
                                             st, "sets");
    const uint64_t cid_base = P.n_sets;
    uint64_t* set_bits = P.set_bits.ensure<uint64_t>(n_new);
    uint64_t* set_boff = P.set_boff.ensure<uint64_t>(n_new + 1);
    if (n_new) { KSCOPE("k_set_cid_len", st); k_set_cid_len<<<blocks_for(n_new), TPB, 0, st>>>(n_new, cid_base, set_desc, T, NC, set_bits, cnt); }
    exclusive_scan<uint64_t, uint64_t>(set_bits, set_boff, n_new, set_boff + n_new, J.rs.scan_tmp, st);
    const uint64_t new_bits = d2h<uint64_t>(set_boff + n_new, st);
    const uint64_t bit_base = P.cs_bits;
    const uint64_t words_needed = (bit_base + new_bits + 63) / 64 + 1;
    auto* words = P.cs_words.ensure_keep<unsigned long long>(words_needed, P.words_zeroed * 8, st, grown(words_needed));
    if (words_needed > P.words_zeroed) {
        KSCOPE("k_fill", st);
        k_fill<unsigned long long><<<blocks_for(words_needed - P.words_zeroed), TPB, 0, st>>>(
            words + P.words_zeroed, words_needed - P.words_zeroed, 0ull);
        P.words_zeroed = words_needed;
    }
    uint64_t* set_off = P.set_off.ensure_keep<uint64_t>(cid_base + n_new, cid_base * 8, st, grown(cid_base + n_new));
    if (n_new) { KSCOPE("k_set_encode", st); k_set_encode<<<blocks_for(n_new, TPB / 64), TPB, 0, st>>>(n_new, cid_base, bit_base, set_desc, set_boff, T, NC, words, set_off); }
    P.n_sets += n_new;
    P.cs_bits += new_bits;
    // P.mt_count (the multi-colour sets in the table) grows at the chunk's end

    // --- per-bucket directory over the sorted keys, masks, links
    uint32_t* dir = P.dir.ensure<uint32_t>(uint64_t(nn) + nb);
    uint32_t* filt = P.filt.ensure<uint32_t>(uint64_t(nn) / NODES_PER_WORD + nb + 1);
    { KSCOPE("k_dir_build", st); k_dir_build<K><<<blocks_for(nn), TPB, 0, st>>>(nn, node_key, node_bucket, node_off, k, dir, filt); }
    uint8_t* node_mask = P.node_mask.ensure<uint8_t>(nn);
    uint32_t* nbr = P.nb.ensure<uint32_t>(2ull * nn);
    { KSCOPE("k_masks", st); k_masks<K><<<xcd_grid(blocks_for(nn)), TPB, 0, st>>>(nn, node_key, node_bucket, node_off, dir, filt, k, node_mask, nbr); }
    uint8_t* eflag = P.eflag.ensure<uint8_t>(nn);
    NodeRec<K>* node_rec = P.node_rec.ensure<NodeRec<K>>(nn);
    uint32_t* link = P.link.ensure<uint32_t>(2ull * nn);
    { KSCOPE("k_links", st); k_links<K><<<blocks_for(nn), TPB, 0, st>>>(nn, node_mask, nbr, meta, ncol, coff, colors, node_slot, T, node_key, rank, node_rec, link, eflag); }

    // --- walk
    uint8_t* visited = P.visited.ensure<uint8_t>(nn);
    WalkArgs wa{};
    wa.n_nodes = nn;
    wa.link = link;
    wa.meta = meta;
    wa.mask = node_mask;
    wa.rank = rank;
    wa.node_bucket = node_bucket;
    wa.b0 = b0;
    wa.visited = visited;
    wa.counters = cnt;
    wa.frag_n = P.frag_n.ensure<uint32_t>(nn);
    wa.frag_open = P.frag_open.ensure<uint8_t>(nn);
    wa.frag_key = P.frag_key.ensure<uint64_t>(nn);
    wa.frag_dir = P.frag_dir.ensure<uint8_t>(nn);
    uint8_t* fflag = P.fflag.ensure<uint8_t>(nn);
    wa.fflag = fflag;
    wa.other_x = P.other_x.ensure<uint32_t>(nn);
    wa.other_ex = P.other_ex.ensure<uint8_t>(nn);
    uint8_t* prc = P.prc.ensure<uint8_t>(nn);
    wa.prc = prc;
    wa.pinfo = P.pinfo.ensure<uint16_t>(nn);
    wa.prank = P.prank.ensure<uint32_t>(nn);
    K* pkey = P.pkey.ensure<K>(nn);
    uint32_t* pcid = P.pcid.ensure<uint32_t>(nn);
    wa.pkey = pkey;
    wa.pcid = pcid;
    wa.node_rec = node_rec;
    const char* wcap = getenv("CDBG_WALK_CAP");  // testing: rank shorter paths by pointer jumping too
    wa.walk_cap = wcap && *wcap ? uint32_t(std::max(1l, atol(wcap))) : 1024u;
    uint8_t* long_bucket = P.lr_bucket.ensure<uint8_t>(nb);
    { KSCOPE("k_fill", st); k_fill<uint8_t><<<blocks_for(nb), TPB, 0, st>>>(long_bucket, nb, uint8_t(0)); }
    wa.long_bucket = long_bucket;
    uint32_t* ends = P.ends.ensure<uint32_t>(nn);
    uint32_t* d_count = (uint32_t*)(P.total.ensure<uint64_t>(8) + 2);
    compact_flags<uint8_t>(eflag, nn, ends, d_count, P.ecnt, J.rs.scan_tmp, st);
    const uint32_t n_ends = d2h<uint32_t>(d_count, st);
    wa.comp_x = ends;
    wa.comp_ex = P.comp_ex.ensure<uint8_t>(n_ends);
    wa.comp_len = P.comp_len.ensure<uint32_t>(n_ends);
    wa.comp_nodes = P.comp_nodes.ensure<uint32_t>(uint64_t(n_ends) * PATH_BUF);
    uint32_t* comp_off = P.comp_off.ensure<uint32_t>(uint64_t(n_ends) + 1);
    wa.comp_off = comp_off;
    if (n_ends) { KSCOPE("k_walk_ends", st); k_walk_ends<<<xcd_grid(blocks_for(n_ends)), TPB, 0, st>>>(wa, ends, n_ends); }
    CDBG_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof h_cnt, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    const uint64_t n_comp = n_ends;
    uint64_t n_other = h_cnt[10], n_ranked = 0;
    uint32_t nu = 0;
    uint32_t *lr_u = nullptr, *lr_own = nullptr, *long_list = nullptr;
    uint2* lr_st = nullptr;
    if (h_cnt[11]) {  // ends stopped at walk_cap: list ranking over their buckets' nodes
        uint8_t* nflag = P.big_flag.ensure<uint8_t>(nn);  // free after the node reduction
        { KSCOPE("k_long_nodes", st); k_long_nodes<<<blocks_for(nn), TPB, 0, st>>>(nn, node_bucket, long_bucket, nflag); }
        lr_u = P.lr_u.ensure<uint32_t>(nn);
        uint32_t* d_nu = (uint32_t*)(dtot + 4);
        compact_flags<uint8_t>(nflag, nn, lr_u, d_nu, P.lr_cnt, J.rs.scan_tmp, st);
        nu = d2h<uint32_t>(d_nu, st);
        if (2ull * nu >= ST_DIST) throw Error("phase 2: too many nodes in buckets with long paths for list ranking");
        uint32_t* uidx = P.lr_uidx.ensure<uint32_t>(nn);
        lr_own = P.lr_own.ensure<uint32_t>(nu);
        { KSCOPE("k_long_index", st); k_long_index<<<blocks_for(nu), TPB, 0, st>>>(nu, lr_u, uidx, lr_own); }
        uint2* sa = P.lr_st0.ensure<uint2>(2ull * nu);
        uint2* sb = P.lr_st1.ensure<uint2>(2ull * nu);
        { KSCOPE("k_long_states", st); k_long_states<<<blocks_for(nu), TPB, 0, st>>>(nu, lr_u, uidx, link, sa, cnt + 5); }
        const int rounds = bits_for(2ull * nu) + 1;  // 2^rounds > the longest chain of states
        for (int r = 0; r < rounds; ++r) {
            { KSCOPE("k_long_jump", st); k_long_jump<<<blocks_for(2ull * nu), TPB, 0, st>>>(2ull * nu, sa, sb); }
            std::swap(sa, sb);
        }
        lr_st = sa;
        long_list = P.lr_list.ensure<uint32_t>(h_cnt[11]);
        { KSCOPE("k_long_ends", st); k_long_ends<<<blocks_for(n_ends), TPB, 0, st>>>(wa, ends, n_ends, nu, lr_u, uidx, lr_st, lr_own, long_list); }
        CDBG_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof h_cnt, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
        if (h_cnt[5]) throw Error("phase 2: list ranking left a path end unresolved");
        n_other = h_cnt[10];
        n_ranked = h_cnt[14];
    }
    exclusive_scan<uint32_t, uint32_t>(wa.comp_len, comp_off, n_comp, comp_off + n_comp, J.rs.scan_tmp, st);
    const uint32_t path_nodes = n_comp ? d2h<uint32_t>(comp_off + n_comp, st) : 0u;
    if (path_nodes > nn) throw Error("phase 2: paths hold more positions than the chunk has nodes");
    // nodes outside simple paths (hairpin paths, cycles) need the visited marks
    wa.mark_visited = path_nodes < nn;
    if (wa.mark_visited) { KSCOPE("k_fill", st); k_fill<uint8_t><<<blocks_for(nn), TPB, 0, st>>>(visited, nn, uint8_t(0)); }
    // the general seed loop places its nodes after the recorded paths
    { KSCOPE("k_set_u64", st); k_set_u64<<<1, 64, 0, st>>>(cnt + 3, path_nodes); }
    if (n_comp) { KSCOPE("k_walk_record", st); k_walk_record<K><<<blocks_for(n_comp), TPB, 0, st>>>(wa, n_comp); }
    if (n_ranked) { KSCOPE("k_long_place", st); k_long_place<K><<<blocks_for(nu), TPB, 0, st>>>(wa, nu, lr_u, lr_st, lr_own); }
    if (n_comp) { KSCOPE("k_walk_frags", st); k_walk_frags<<<blocks_for(n_comp), TPB, 0, st>>>(wa, n_comp); }
    if (n_ranked) { KSCOPE("k_walk_frags_long", st); k_walk_frags_long<<<blocks_for(n_ranked, TPB / 64), TPB, 0, st>>>(wa, long_list, uint32_t(n_ranked)); }
    if (n_other) { KSCOPE("k_walk_other", st); k_walk_other<K><<<blocks_for(n_other), TPB, 0, st>>>(wa, n_other); }
    if (wa.mark_visited) { KSCOPE("k_walk_cycles", st); k_walk_cycles<K><<<blocks_for(nn), TPB, 0, st>>>(wa); }
    CDBG_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof h_cnt, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    if (h_cnt[6]) throw Error("phase 2: colour-set content hash collision (" + std::to_string(h_cnt[6]) + " nodes)");
    if (h_cnt[5] || h_cnt[3] != nn)
        throw Error("phase 2: walk covered " + std::to_string(h_cnt[3]) + " of " + std::to_string(nn) +
                    " nodes (" + std::to_string(h_cnt[5]) + " errors)");
    uint32_t* forder = P.forder.ensure<uint32_t>(nn);  // first positions of the fragments, ascending
    compact_flags<uint8_t>(fflag, nn, forder, d_count, P.fcnt, J.rs.scan_tmp, st);
    const uint64_t nf = d2h<uint32_t>(d_count, st);

    // --- fragments in (bucket, seed rank) order, appended to the job stream.
    // Keys are 32-bit when the chunk's local bucket and rank bits fit (4
    // passes of 4+4 bytes instead of 6 of 8+4).
    uint64_t* fkey = P.fkey.ensure<uint64_t>(nf);
    uint64_t* fkey_t = P.fkey_t.ensure<uint64_t>(nf);
    uint32_t* dval = P.fdval.ensure<uint32_t>(nf);  // sorted: output rank -> position-order index
    uint32_t* dval_t = P.forder_t.ensure<uint32_t>(nf);
    uint32_t* psize = P.fpsize.ensure<uint32_t>(nf);
    // fragments of more than LONG_FRAG k-mers: a block each, the blocks
    // striding over the device-side list (no host read of its length)
    const uint32_t max_long = uint32_t(uint64_t(nn) / (LONG_FRAG + 1) + 1);
    const uint32_t long_grid = std::min<uint32_t>(max_long, 1024);
    uint32_t* long_frags = P.flong.ensure<uint32_t>(max_long);
    if (nf) { KSCOPE("k_frag_psize", st); k_frag_psize<<<blocks_for(nf), TPB, 0, st>>>(nf, forder, wa.frag_n, pcid, k, psize, dval, long_frags, cnt + 15); }
    if (nf) { KSCOPE("k_frag_psize_long", st); k_frag_psize_long<<<long_grid, TPB, 0, st>>>(long_frags, cnt + 15, forder, wa.frag_n, pcid, k, psize); }
    uint64_t max_btup = 0;
    for (uint32_t b = b0; b < b1; ++b) max_btup = std::max<uint64_t>(max_btup, h_btup[b + 1] - h_btup[b]);
    const int rbits = bits_for(max_btup), lbits = bits_for(b1 - b0);
    SortPlan fplan;
    fplan.build({0, nf});
    const char* force64 = getenv("CDBG_FRAG_KEY64");  // testing: the 64-bit key path
    const char* nostage = getenv("CDBG_FRAG_NOSTAGE");  // testing: records read positions from HBM
    const uint32_t pos_cap = nostage && *nostage == '1' ? 0u : ~0u;
    if (lbits + rbits <= 32 && !(force64 && *force64 == '1')) {
        uint32_t* fk32 = (uint32_t*)fkey;
        uint32_t* fk32_t = (uint32_t*)fkey_t;
        if (nf) { KSCOPE("k_frag_keys", st); k_frag_keys32<<<blocks_for(nf), TPB, 0, st>>>(nf, forder, wa.frag_key, b0, rbits, fk32, cnt + 12); }
        bool in_temp = false;
        segmented_radix_sort<uint32_t, uint32_t>(fk32, dval, fk32_t, dval_t, fplan, 0, lbits + rbits, 8, P.rs2, st,
                                                 "frags", &in_temp);
        if (in_temp) dval = dval_t;
    } else {
        if (nf) { KSCOPE("k_frag_keys", st); k_frag_keys<<<blocks_for(nf), TPB, 0, st>>>(nf, forder, wa.frag_key, fkey); }
        segmented_radix_sort<uint64_t, uint32_t>(fkey, dval, fkey_t, dval_t, fplan, 0, 32 + bbits, 8, P.rs2, st,
                                                 "frags");
    }
    // record sizes in output order, their offsets, and the offsets in
    // position order; the records are then built in position order
    uint64_t* fsize = P.fsize.ensure<uint64_t>(nf);
    uint64_t* fboff = P.fboff.ensure<uint64_t>(nf + 1);
    uint64_t* pboff = P.fpboff.ensure<uint64_t>(nf);
    if (nf) { KSCOPE("k_frag_osize", st); k_frag_osize<<<blocks_for(nf), TPB, 0, st>>>(nf, dval, psize, fsize); }
    exclusive_scan<uint64_t, uint64_t>(fsize, fboff, nf, fboff + nf, J.rs.scan_tmp, st);
    if (nf) { KSCOPE("k_frag_pboff", st); k_frag_pboff<<<blocks_for(nf), TPB, 0, st>>>(nf, dval, fboff, pboff); }
    const uint64_t fbytes = d2h<uint64_t>(fboff + nf, st);
    if (P.frag_total == 0 && t_done < t_all && !P.spill) {
        // the first chunk decides where the job's fragment stream lives: in
        // HBM only when its size, extrapolated from this chunk, leaves room
        // for the chunks' working set; else in host memory, chunk by chunk
        size_t free_b = 0, total_b = 0;
        CDBG_HIP(hipMemGetInfo(&free_b, &total_b));
        const double projected = double(fbytes) * double(t_all) / double(t_done);
        const double room = double(free_b) + double(g_block_cache.cached.load()) + double(P.frag_bytes.cap);
        const char* force = getenv("CDBG_FRAG_SPILL");  // testing: keep the stream in host memory
        P.spill = projected > 0.35 * room || (force && *force == '1');
    }
    // the chunk's records go after the stream's (at its start when the
    // stream is kept in host memory)
    uint8_t* fout = P.spill ? P.frag_bytes.ensure<uint8_t>(fbytes)
                            : P.frag_bytes.ensure_keep<uint8_t>(P.frag_total + fbytes, P.frag_total, st,
                                                                grown(P.frag_total + fbytes)) + P.frag_total;
    if (nf) { KSCOPE("k_frag_write", st); k_frag_write_pos<K><<<blocks_for(nf), TPB, 0, st>>>(nf, forder, wa.frag_n, wa.frag_open, wa.frag_dir, prc, pkey, pcid, k, psize, pboff, fout, pos_cap); }
    if (nf) { KSCOPE("k_frag_write_long", st); k_frag_write_long<K><<<long_grid, TPB, 0, st>>>(long_frags, cnt + 15, forder, wa.frag_n, wa.frag_open, wa.frag_dir, prc, pkey, pcid, k, psize, pboff, fout); }
    CDBG_HIP(hipGetLastError());
    CDBG_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof h_cnt, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    if (h_cnt[12]) throw Error("phase 2: fragment seed rank outside its bucket's tuples");
    if (P.spill && fbytes) {
        uint8_t* h = P.h_frags.ensure_keep(P.frag_total + fbytes, P.frag_total, grown(P.frag_total + fbytes));
        CDBG_HIP(hipMemcpyAsync(h + P.frag_total, fout, fbytes, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
    }
    P.frag_total += fbytes;
    P.n_frags += nf;
    P.n_nodes += nn;
    P.n_tuples += nt;
    P.n_primary += n_primary;
    P.total_integers += h_cnt[2];
    P.mt_count += h_cnt[13];  // multi-colour sets in the table: sizes it
}

template <class K>
static void phase2_impl(Job& J) {
    if (!J.p2) J.p2 = new Phase2Bufs();
    Phase2Bufs& P = *J.p2;
    Phase1Bufs& Q = J.p1;
    hipStream_t st = J.st;
    const uint32_t k = J.k, B = J.B;
    const uint64_t ns = J.n_supers;
    if (J.num_colors >= (1u << 29)) throw Error("more than 2^29 colours");

    hipEvent_t e0, e1;
    CDBG_HIP(hipEventCreate(&e0));
    CDBG_HIP(hipEventCreate(&e1));
    CDBG_HIP(hipEventRecord(e0, st));

    // job-wide state
    P.n_tuples = P.n_nodes = P.n_sets = P.n_frags = P.frag_total = P.cs_bits = 0;
    P.n_primary = P.total_integers = P.words_zeroed = 0;
    P.mt_count = 0;
    const uint32_t NC = J.num_colors;
    auto* single_min = P.single_min.ensure<unsigned long long>(NC);
    auto* single_cid = P.single_cid.ensure<uint32_t>(NC);
    { KSCOPE("k_fill", st); k_fill<unsigned long long><<<blocks_for(NC), TPB, 0, st>>>(single_min, NC, ~0ull); }
    { KSCOPE("k_fill", st); k_fill<uint32_t><<<blocks_for(NC), TPB, 0, st>>>(single_cid, NC, NONE32); }
    if (P.mt_cap == 0) {
        mt_resize(P, 1024, st);
    } else {  // clear, keeping the capacity of the previous run
        { KSCOPE("k_fill2", st); k_fill2<unsigned long long><<<blocks_for(P.mt_cap), TPB, 0, st>>>(P.mt_hk.as<unsigned long long>(), P.mt_cap, 0ull, ~0ull); }
        { KSCOPE("k_fill", st); k_fill<uint32_t><<<blocks_for(P.mt_cap), TPB, 0, st>>>(P.mt_cid.as<uint32_t>(), P.mt_cap, NONE32); }
    }

    // k-mer occurrences of every record, per bucket
    uint32_t* kcnt = P.kcnt.ensure<uint32_t>(ns);
    uint64_t* tup_off = P.tup_off.ensure<uint64_t>(ns + 1);
    if (ns) { KSCOPE("k_count_kmers", st); k_count_kmers<<<blocks_for(ns), TPB, 0, st>>>(ns, Q.b_len.as<uint32_t>(), k, kcnt); }
    exclusive_scan<uint32_t, uint64_t>(kcnt, tup_off, ns, tup_off + ns, J.rs.scan_tmp, st);
    uint64_t* btup = P.btup.ensure<uint64_t>(B + 1);
    { KSCOPE("k_bucket_tuple_off", st); k_bucket_tuple_off<<<blocks_for(B + 1), TPB, 0, st>>>(B, Q.bucket_off.as<uint64_t>(), tup_off, btup); }
    std::vector<uint64_t> h_btup(B + 1);
    CDBG_HIP(hipMemcpyAsync(h_btup.data(), btup, (B + 1) * 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));

    // chunk budget: k-mer occurrences whose working set fits free HBM
    uint64_t budget = J.chunk_tuples ? J.chunk_tuples : J.auto_chunk;
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        CDBG_HIP(hipMemGetInfo(&free_b, &total_b));
        free_b += size_t(g_block_cache.cached.load());  // cached blocks are free for this job
        // the job's per-occurrence model (tuples, sort buffers, nodes,
        // table, walk and fragment arrays; wider for 128-bit keys); buffers
        // this phase already holds from an earlier run count as available
        const uint64_t held = g_device_bytes.load();
        const uint64_t usable = uint64_t(free_b) + std::min<uint64_t>(held, uint64_t(free_b)) / 4;
        budget = chunk_budget(usable, k);
        J.auto_chunk = budget;  // later runs of this job plan the same chunks
    }
    uint32_t nchunks = 0;
    P.spill = false;  // the host stream keeps its capacity for the next run
    {
        // a job of several chunks computes its tuple offsets chunk by chunk
        uint32_t b1 = 1;
        while (b1 < B && h_btup[b1 + 1] <= budget) ++b1;
        if (b1 < B) {
            P.kcnt.release();
            P.tup_off.release();
        }
    }
    const bool adaptive = J.chunk_tuples == 0;
    for (uint32_t b0 = 0; b0 < B;) {
        // the chunk also fits what the job-wide structures (colour-set table
        // and streams) leave free now: they grow from chunk to chunk
        uint64_t cap = budget;
        if (adaptive && b0 > 0) {
            size_t free_b = 0, total_b = 0;
            CDBG_HIP(hipMemGetInfo(&free_b, &total_b));
            cap = std::min(cap, chunk_budget(uint64_t(free_b) + g_block_cache.cached.load() + P.chunk_bytes(), k));
        }
        uint32_t b1 = b0 + 1;
        while (b1 < B && h_btup[b1 + 1] - h_btup[b0] <= cap) ++b1;
        try {
            run_chunk<K>(J, P, b0, b1, h_btup);
        } catch (Error& e) {
            throw Error(std::string(e.what()) + " [phase-2 chunk " + std::to_string(nchunks + 1) + ": buckets " +
                        std::to_string(b0) + ".." + std::to_string(b1 - 1) + ", " +
                        std::to_string(h_btup[b1] - h_btup[b0]) + " of " + std::to_string(h_btup[B]) +
                        " occurrences, budget " + std::to_string(budget) + "]");
        }
        ++nchunks;
        if (J.progress) __atomic_fetch_add(J.progress, uint64_t(b1 - b0), __ATOMIC_RELAXED);
        b0 = b1;
        if (J.p2_max_chunks && nchunks >= J.p2_max_chunks) break;  // a trial over the first chunks only
    }
    J.frags_in_host = P.spill;
    J.n_chunks = nchunks;

    CDBG_HIP(hipEventRecord(e1, st));
    CDBG_HIP(hipStreamSynchronize(st));
    float ms = 0;
    CDBG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    J.stats.num_occ = P.n_tuples;
    J.stats.num_tuples = P.n_tuples;
    J.stats.num_nodes = P.n_nodes;
    J.stats.num_kmers = P.n_primary;
    J.stats.num_color_sets = P.n_sets;
    J.stats.num_frags = P.n_frags;
    J.stats.frag_bytes = P.frag_total;
    J.stats.cs_bits = P.cs_bits;
    J.stats.total_integers = P.total_integers;
    J.stats.ms_phase2 = ms;
    J.phase2_done = true;
}

void run_phase2(Job& J) {
    AllocStream alloc_on_job_stream(J.st);
    if (J.k <= 31)
        phase2_impl<uint64_t>(J);
    else
        phase2_impl<Key128>(J);
}

void free_phase2(Job& J) {
    delete J.p2;
    J.p2 = nullptr;
}

void copy_phase2_results_device(Job& J, void* frags, void* words, void* offsets, hipStream_t st) {
    if (!J.phase2_done) throw Error("phase 2 has not run");
    Phase2Bufs& P = *J.p2;
    CDBG_HIP(hipStreamSynchronize(J.st));
    if (frags && P.frag_total)
        CDBG_HIP(hipMemcpyAsync(frags, P.spill ? (const void*)P.h_frags.p : P.frag_bytes.p, P.frag_total,
                                P.spill ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, st));
    const uint64_t nw = (P.cs_bits + 63) / 64;
    if (words && nw) CDBG_HIP(hipMemcpyAsync(words, P.cs_words.p, nw * 8, hipMemcpyDeviceToDevice, st));
    if (offsets && P.n_sets) CDBG_HIP(hipMemcpyAsync(offsets, P.set_off.p, P.n_sets * 8, hipMemcpyDeviceToDevice, st));
    CDBG_HIP(hipStreamSynchronize(st));
}

void copy_phase2_results(Job& J, uint8_t* frags, uint64_t* words, uint64_t* offsets) {
    AllocStream alloc_on_job_stream(J.st);
    if (!J.phase2_done) throw Error("phase 2 has not run");
    Phase2Bufs& P = *J.p2;
    if (frags && P.frag_total) {
        if (P.spill)
            std::memcpy(frags, P.h_frags.p, P.frag_total);
        else
            CDBG_HIP(hipMemcpy(frags, P.frag_bytes.p, P.frag_total, hipMemcpyDeviceToHost));
    }
    const uint64_t nw = (P.cs_bits + 63) / 64;
    if (words && nw) CDBG_HIP(hipMemcpy(words, P.cs_words.p, nw * 8, hipMemcpyDeviceToHost));
    if (offsets && P.n_sets) CDBG_HIP(hipMemcpy(offsets, P.set_off.p, P.n_sets * 8, hipMemcpyDeviceToHost));
}

const uint8_t* phase2_frags_where(Job& J, bool& on_device, uint64_t& n) {
    if (!J.phase2_done) throw Error("phase 2 has not run");
    Phase2Bufs& P = *J.p2;
    CDBG_HIP(hipStreamSynchronize(J.st));
    n = P.frag_total;
    on_device = !P.spill;
    return P.spill ? P.h_frags.p : (const uint8_t*)P.frag_bytes.p;
}

}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

CDBG_API int cdbg_gpu_job_frags(cdbg_gpu_job* job, uint8_t* dst) {
    try {
        copy_phase2_results(job->J, dst, nullptr, nullptr);
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API int cdbg_gpu_job_color_sets(cdbg_gpu_job* job, uint64_t* words, uint64_t* offsets) {
    try {
        copy_phase2_results(job->J, nullptr, words, offsets);
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

CDBG_API int cdbg_gpu_job_set_chunk_tuples(cdbg_gpu_job* job, uint64_t max_tuples) {
    if (!job) return set_error("null job");
    job->J.chunk_tuples = max_tuples;
    return 0;
}

CDBG_API int64_t cdbg_gpu_job_num_chunks(cdbg_gpu_job* job) {
    if (!job) return set_error("null job");
    return job->J.n_chunks;
}

}  // extern "C"
