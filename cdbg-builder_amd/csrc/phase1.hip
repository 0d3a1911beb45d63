// Phase 1 (bucket-write) on the device.
//
//   k_pack2bit      ASCII -> 2-bit codes (16 per u32, base i at bits 2(i%16),
//                   i.e. the byte packing of super_kmer.hpp:43-49 read as LE
//                   words) + an ACGT bitmap; nuc_to_2bit semantics
//                   (kmer.hpp:35-52).
//   k_superkmers    per tile of 4096 (k-1)-mer positions: canonical ntHash of
//                   every m-mer (minimizer.hpp:40-62), window minimum over the
//                   W = k-m m-mers of every (k-1)-mer (minimizer.hpp:94-157,
//                   bucket_ingester.hpp:52), and the super-k-mer starts/ends
//                   with their BEGIN/END/OWNS flags (bucket_ingester.hpp:80-126).
//                   One pass: each tile stores its starts (4 bytes: offset in
//                   the tile, flags, bucket) and ends (2 bytes) in its own
//                   slot range and its two counts.
//   k_zip_supers    after a scan of the tile counts: pairs the i-th start
//                   with the i-th end (they alternate along every ACGT run),
//                   assigns the colour.
//   bucket scatter  stable radix sort of super indices by bucket
//                   (h >> 1) % B (bucket_ingester.hpp:67) + gather.
//
// A super-k-mer is described by [start, start+len) in the job's base
// coordinates, its 4 flag bits, its bucket and its colour.  Inside a bucket
// the supers stay in position order, which is colour-major: this is the
// record stream phase 2 consumes.
#include <hip/hip_runtime.h>

#include "hbm_model.h"
#include "job.h"

namespace cdbg_amd {

namespace {

constexpr int SK_THREADS = 256;
constexpr int SK_PER_THREAD = 16;
constexpr int SK_TILE = SK_THREADS * SK_PER_THREAD;  // (k-1)-mer positions per tile
constexpr int SK_NPK = SK_TILE / 16 + 16;            // packed words staged per tile
constexpr int SK_NVD = SK_TILE / 32 + 16;            // validity words staged per tile

__host__ __device__ constexpr int sk_hidx(int p) { return p + (p >> 4); }

// ntHash per-base seeds: the published ntHash table (minimizer.hpp:21-26).
constexpr uint64_t NT_SEED0 = 0x3c8bfbb395c60474ULL;
constexpr uint64_t NT_SEED1 = 0x3193c18562a02b4cULL;
constexpr uint64_t NT_SEED2 = 0x20323ed082572324ULL;
constexpr uint64_t NT_SEED3 = 0x295549f54be24456ULL;

__device__ __forceinline__ uint64_t rotl(uint64_t x, uint32_t r) {
    r &= 63u;
    return r ? (x << r) | (x >> (64u - r)) : x;
}

struct HashTabs {
    uint64_t s[4];    // S[b]
    uint64_t sm[4];   // rol(S[b], m)          (leaving base, forward strand)
    uint64_t rc1[4];  // ror(S[b^3], 1)        (leaving base, reverse strand)
    uint64_t rcm[4];  // rol(S[b^3], m-1)      (entering base, reverse strand)
};

__global__ __launch_bounds__(256) void k_pack2bit(const uint8_t* __restrict__ bytes, uint64_t ngroups,
                                                  uint32_t* __restrict__ packed, uint32_t* __restrict__ valid) {
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const uint4* src = reinterpret_cast<const uint4*>(bytes + g * 32);
    const uint4 a = src[0], b = src[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t p0 = 0, p1 = 0, v = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        const uint32_t c = ((w[i >> 2] >> (8 * (i & 3))) & 0xffu) | 0x20u;  // fold case
        const bool ok = (c == 'a') | (c == 'c') | (c == 'g') | (c == 't');
        // a->0, c->1, g->2, t->3 from the ASCII bits of the lower-case letter
        const uint32_t code = ok ? (((c >> 1) & 3u) ^ ((c >> 2) & 1u)) : 0u;
        if (i < 16)
            p0 |= code << (2 * i);
        else
            p1 |= code << (2 * (i - 16));
        v |= uint32_t(ok) << i;
    }
    packed[2 * g] = p0;
    packed[2 * g + 1] = p1;
    valid[g] = v;
}

struct SkArgs {
    const uint32_t* packed;
    const uint32_t* valid;
    uint64_t n_bases;
    uint64_t pos_lo, pos_hi;  // super-k-mers starting and ending in [pos_lo, pos_hi): this launch's share
    uint64_t tile0;           // first tile of the launch (block b works on tile tile0 + b)
    uint64_t n_pk_words, n_vd_words;
    uint32_t k, m, W, B;
    const uint64_t* cuts;  // null: bucket (h >> 1) % B; else B - 1 ascending cuts of the low 32 bits of h >> 1
    HashTabs T;
    uint32_t* cnt_s;  // per-tile start counts
    uint32_t* cnt_e;  // per-tile end counts
    uint32_t* st_s;   // per tile SK_TILE slots: bucket << 16 | flags << 12 | offset in tile
    uint16_t* st_e;   // per tile SK_TILE slots: flags << 12 | offset in tile
};

// Bucket of a minimizer: (h >> 1) % B (bucket_ingester.hpp:67), or under a
// bucket map the number of cuts <= the low 32 bits of h >> 1 (uniform, where
// the high bits of a window minimum are not; every minimizer still has one
// bucket, so all super-k-mers of a minimizer stay together).
__device__ __forceinline__ uint32_t bucket_of(uint64_t h, const SkArgs& a) {
    if (!a.cuts) {
        // B from -b is a power of two: a mask instead of a 64-bit division
        if ((a.B & (a.B - 1)) == 0) return uint32_t(h >> 1) & (a.B - 1);
        return uint32_t((h >> 1) % a.B);
    }
    const uint64_t x = (h >> 1) & 0xFFFFFFFFull;
    uint32_t lo = 0, hi = a.B - 1;  // first cut > x
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.cuts[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Minima of the 18 windows of W hashes starting at s_h positions q0 .. q0+17
// (the exact sliding-window minimum of windowed_min, minimizer.hpp:94-157;
// only the value is used).  Windows are taken in groups of G (G <= W): the
// windows starting at t0 .. t0+G-1 all contain the core [t0+G-1, t0+W-1],
// so window t0+i is the minimum of the core, the suffix [t0+i, t0+G-2] of
// the group's first G-1 hashes and the prefix [t0+W, t0+W+i-1] of the G-1
// hashes after the core: W + G - 1 reads per group instead of G W.  G = 1 is
// the direct minimum.
template <int G>
__device__ __forceinline__ void window_minima(const uint64_t* s_h, int q0, uint32_t W,
                                              uint64_t (&mn)[SK_PER_THREAD + 2]) {
    static_assert((SK_PER_THREAD + 2) % G == 0, "whole groups");
#pragma unroll
    for (int g = 0; g < (SK_PER_THREAD + 2) / G; ++g) {
        const int t0 = q0 + g * G;
        uint64_t core = ~0ull;
        for (uint32_t i = G - 1; i < W; ++i) {
            const uint64_t h = s_h[sk_hidx(t0 + int(i))];
            core = h < core ? h : core;
        }
        uint64_t suf[G], pre[G];
        suf[G - 1] = ~0ull;
        pre[0] = ~0ull;
#pragma unroll
        for (int i = G - 2; i >= 0; --i) {
            const uint64_t h = s_h[sk_hidx(t0 + i)];
            suf[i] = h < suf[i + 1] ? h : suf[i + 1];
        }
#pragma unroll
        for (int i = 1; i < G; ++i) {
            const uint64_t h = s_h[sk_hidx(t0 + int(W) + i - 1)];
            pre[i] = h < pre[i - 1] ? h : pre[i - 1];
        }
#pragma unroll
        for (int i = 0; i < G; ++i) {
            uint64_t v = core;
            v = suf[i] < v ? suf[i] : v;
            v = pre[i] < v ? pre[i] : v;
            mn[g * G + i] = v;
        }
    }
}

__global__ __launch_bounds__(SK_THREADS) void k_superkmers(SkArgs a) {
    // one pad slot per 16 hashes: threads own 16 consecutive hashes, so an
    // unpadded stride of 128 bytes would put a wave's accesses in 2 banks
    __shared__ uint64_t s_h[sk_hidx(SK_TILE + 72)];
    __shared__ uint32_t s_pk[SK_NPK];
    __shared__ uint32_t s_vd[SK_NVD];
    __shared__ uint64_t s_tmp[SK_THREADS / 64 + 1];
    __shared__ uint64_t s_tab[16];  // the four per-base tables S, sm, rc1, rcm: a lookup is one LDS read

    const int64_t p0 = int64_t(a.tile0 + blockIdx.x) * SK_TILE;
    if (threadIdx.x < 16) {
        const uint32_t t = threadIdx.x >> 2, b = threadIdx.x & 3u;
        s_tab[threadIdx.x] = t == 0 ? a.T.s[b] : t == 1 ? a.T.sm[b] : t == 2 ? a.T.rc1[b] : a.T.rcm[b];
    }
    const int64_t pk_w0 = p0 / 16 - 1;  // first staged packed word
    const int64_t vd_w0 = p0 / 32 - 1;  // first staged validity word
    for (int i = threadIdx.x; i < SK_NPK; i += SK_THREADS) {
        const int64_t w = pk_w0 + i;
        s_pk[i] = (w >= 0 && uint64_t(w) < a.n_pk_words) ? a.packed[w] : 0u;
    }
    for (int i = threadIdx.x; i < SK_NVD; i += SK_THREADS) {
        const int64_t w = vd_w0 + i;
        s_vd[i] = (w >= 0 && uint64_t(w) < a.n_vd_words) ? a.valid[w] : 0u;
    }
    __syncthreads();

    const uint32_t m = a.m, W = a.W;
    auto base_at = [&](int64_t p) -> uint32_t {  // p >= pk_w0*16
        const int64_t lp = p - pk_w0 * 16;
        return (s_pk[lp >> 4] >> (2 * (lp & 15))) & 3u;
    };

    // 1. canonical m-mer hashes for positions p0-1 .. p0-1+SK_TILE+W: one run
    //    per thread, rolled from one hash computed whole (runs of
    //    ceil(nq / SK_THREADS) positions, so the W + 1 positions past the
    //    tile leave no thread a second run)
    const int nq = SK_TILE + int(W) + 1;
    const int rlen = (nq + SK_THREADS - 1) / SK_THREADS;
    if (int(threadIdx.x) * rlen < nq) {  // the last threads may have no run
        const int q0 = int(threadIdx.x) * rlen;
        const int cnt = min(rlen, nq - q0);
        const int64_t pos = p0 - 1 + q0;
        uint64_t f = 0, r = 0;
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t b = base_at(pos + i);
            f = rotl(f, 1) ^ s_tab[b];
            r ^= rotl(s_tab[b ^ 3u], i);
        }
        s_h[sk_hidx(q0)] = f <= r ? f : r;
        for (int i = 1; i < cnt; ++i) {
            const uint32_t bo = base_at(pos + i - 1);
            const uint32_t bi = base_at(pos + i + m - 1);
            f = rotl(f, 1) ^ s_tab[4 + bo] ^ s_tab[bi];
            r = rotl(r, 63) ^ s_tab[8 + bo] ^ s_tab[12 + bi];
            s_h[sk_hidx(q0 + i)] = f <= r ? f : r;
        }
    }
    __syncthreads();
    // 2. window minima of this thread's (k-1)-mers and one neighbour on each
    //    side: mn[t] = minimizer of (k-1)-mer j0 - 1 + t, t = 0 .. 17, in
    //    groups of G consecutive windows that share their middle hashes
    const int q0 = int(threadIdx.x) * SK_PER_THREAD;  // s_h index of (k-1)-mer j0 - 1
    uint64_t mn[SK_PER_THREAD + 2];
    if (W >= 18)
        window_minima<18>(s_h, q0, W, mn);
    else if (W >= 9)
        window_minima<9>(s_h, q0, W, mn);
    else if (W >= 6)
        window_minima<6>(s_h, q0, W, mn);
    else
        window_minima<1>(s_h, q0, W, mn);

    // 3. (k-1)-mer validity and super-k-mer boundaries
    const uint64_t kmask = (uint64_t(1) << (a.k - 1)) - 1;
    auto valid_k1 = [&](int64_t j) -> bool {
        if (j < 0) return false;
        const int64_t lb = j - vd_w0 * 32;
        const int w = int(lb >> 5), o = int(lb & 31);
        const uint64_t lo = uint64_t(s_vd[w]) | (uint64_t(s_vd[w + 1]) << 32);
        uint64_t win = lo >> o;
        if (o) win |= uint64_t(s_vd[w + 2]) << (64 - o);
        return (win & kmask) == kmask;
    };
    const int64_t j0 = p0 + int64_t(threadIdx.x) * SK_PER_THREAD;
    uint32_t starts = 0, ends = 0;  // bit i: position j0+i starts / ends a super
    bool vprev = valid_k1(j0 - 1);
    bool vcur = valid_k1(j0);
#pragma unroll
    for (int i = 0; i < SK_PER_THREAD; ++i) {
        const int64_t j = j0 + i;
        const bool vnext = valid_k1(j + 1);
        if (vcur && (vprev || vnext) && uint64_t(j) >= a.pos_lo && uint64_t(j) < a.pos_hi) {
            if (!vprev || mn[i + 1] != mn[i]) starts |= 1u << i;
            if (!vnext || mn[i + 1] != mn[i + 2]) ends |= 1u << i;
        }
        vprev = vcur;
        vcur = vnext;
    }
    const uint32_t ns = __popc(starts), ne = __popc(ends);
    uint64_t ts, te;
    uint32_t os = uint32_t(block_exclusive_scan<uint64_t>(uint64_t(ns), s_tmp, ts));
    uint32_t oe = uint32_t(block_exclusive_scan<uint64_t>(uint64_t(ne), s_tmp, te));
    if (threadIdx.x == 0) {
        a.cnt_s[blockIdx.x] = uint32_t(ts);
        a.cnt_e[blockIdx.x] = uint32_t(te);
    }
    uint32_t* st_s = a.st_s + uint64_t(blockIdx.x) * SK_TILE;
    uint16_t* st_e = a.st_e + uint64_t(blockIdx.x) * SK_TILE;
#pragma unroll
    for (int i = 0; i < SK_PER_THREAD; ++i) {
        const int64_t j = j0 + i;
        const uint32_t lo = uint32_t(threadIdx.x) * SK_PER_THREAD + uint32_t(i);  // j - p0
        if (starts & (1u << i)) {
            const bool begin = !valid_k1(j - 1);
            const bool own_first = begin || mn[i + 1] < mn[i];
            const uint32_t fl = (begin ? 1u : 0u) | (own_first ? 4u : 0u);
            st_s[os++] = (bucket_of(mn[i + 1], a) << 16) | (fl << 12) | lo;
        }
        if (ends & (1u << i)) {
            const bool end = !valid_k1(j + 1);
            const bool own_last = end || mn[i + 1] < mn[i + 2];
            const uint32_t fl = (end ? 2u : 0u) | (own_last ? 8u : 0u);
            st_e[oe++] = uint16_t((fl << 12) | lo);
        }
    }
}

// One super-k-mer record as the bucket sort moves it.
struct SuperRec {
    uint64_t start;
    uint32_t lenfl;  // length << 4 | flags
    uint32_t color;
};

// Pairs the i-th start with the i-th end: record, and its bucket as sort key.
// One block per tile; its starts are global supers off_s[t] .. off_s[t] +
// cnt_s[t] - 1, and the i-th end lies in the first tile whose ends reach i.
__global__ __launch_bounds__(256) void k_zip_supers(uint64_t tile0, uint64_t ntiles, const uint32_t* __restrict__ cnt_s,
                                                    const uint32_t* __restrict__ cnt_e,
                                                    const uint64_t* __restrict__ off_s,
                                                    const uint64_t* __restrict__ off_e,
                                                    const uint32_t* __restrict__ st_s,
                                                    const uint16_t* __restrict__ st_e, uint32_t k,
                                                    const uint64_t* __restrict__ color_off, uint32_t num_colors,
                                                    SuperRec* __restrict__ rec, uint32_t* __restrict__ key,
                                                    unsigned long long* __restrict__ too_long) {
    const uint64_t t = blockIdx.x;
    const uint32_t ns = cnt_s[t];
    for (uint32_t j = threadIdx.x; j < ns; j += blockDim.x) {
        const uint64_t i = off_s[t] + j;
        const uint32_t se = st_s[t * SK_TILE + j];
        const uint64_t sj = (tile0 + t) * SK_TILE + (se & 0xFFFu);
        const uint32_t sfl = (se >> 12) & 15u;
        const uint64_t s = (sfl & 1u) ? sj : sj - 1;
        uint64_t te = t;
        while (te > 0 && off_e[te] > i) --te;
        while (te + 1 < ntiles && off_e[te] + cnt_e[te] <= i) ++te;
        const uint32_t ee = st_e[te * SK_TILE + (i - off_e[te])];
        const uint64_t ej = (tile0 + te) * SK_TILE + (ee & 0xFFFu);
        const uint32_t efl = (ee >> 12) & 15u;
        const uint64_t e = ((efl & 2u) ? ej - 1 : ej) + k;
        // colour = last c with color_off[c] <= s
        uint32_t lo = 0, hi = num_colors;  // invariant: color_off[lo] <= s < color_off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (color_off[mid] <= s)
                lo = mid;
            else
                hi = mid;
        }
        const uint64_t len = e - s;
        if (len >= (uint64_t(1) << 28)) atomicAdd(too_long, 1ull);
        rec[i] = SuperRec{s, uint32_t(len << 4) | sfl | efl, lo};
        key[i] = se >> 16;
    }
}

// The bucket-ordered records as separate arrays.
__global__ void k_split_supers(uint64_t n, const SuperRec* __restrict__ rec, uint64_t* __restrict__ b_start,
                               uint32_t* __restrict__ b_len, uint8_t* __restrict__ b_flags,
                               uint32_t* __restrict__ b_color) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SuperRec r = rec[i];
    b_start[i] = r.start;
    b_len[i] = r.lenfl >> 4;
    b_flags[i] = uint8_t(r.lenfl & 15u);
    b_color[i] = r.color;
}

// off[b] = first index i with key[i] >= b, for b in [0, B].
__global__ void k_bucket_offsets(uint64_t n, const uint32_t* __restrict__ key, uint32_t B,
                                 uint64_t* __restrict__ off) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (n == 0) {
        if (i <= B) off[i] = 0;
        return;
    }
    if (i >= n) return;
    const int64_t prev = i == 0 ? -1 : int64_t(key[i - 1]);
    const int64_t cur = key[i];
    for (int64_t b = prev + 1; b <= cur; ++b) off[b] = i;
    if (i == n - 1)
        for (int64_t b = cur + 1; b <= int64_t(B); ++b) off[b] = n;
}

// Chunked phase 1: the bucket-sorted records of one colour range (n_c of
// them, bucket b at [off_c[b], off_c[b+1])) placed into the job's
// bucket-major arrays, bucket b of this range starting at base_c[b] (after
// the same bucket of every earlier range: the records stay in position order
// inside a bucket).
__global__ void k_merge_supers(uint64_t n_c, uint32_t B, const SuperRec* __restrict__ rec,
                               const uint64_t* __restrict__ off_c, const uint64_t* __restrict__ base_c,
                               uint64_t* __restrict__ b_start, uint32_t* __restrict__ b_len,
                               uint8_t* __restrict__ b_flags, uint32_t* __restrict__ b_color,
                               uint32_t* __restrict__ bucket) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_c) return;
    uint32_t lo = 0, hi = B;  // last b with off_c[b] <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off_c[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const uint64_t d = base_c[lo] + (i - off_c[lo]);
    const SuperRec r = rec[i];
    b_start[d] = r.start;
    b_len[d] = r.lenfl >> 4;
    b_flags[d] = uint8_t(r.lenfl & 15u);
    b_color[d] = r.color;
    bucket[d] = lo;
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline uint64_t d2h_u64(const uint64_t* d, hipStream_t st) {
    uint64_t v = 0;
    CDBG_HIP(hipMemcpyAsync(&v, d, sizeof v, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    return v;
}

}  // namespace

namespace {

// Super-k-mers that start and end in [lo, hi) (a union of whole colours): the
// tiles covering the range are processed, their records (position order) and
// buckets written to rec_b / key_b; returns the count.
uint64_t supers_of_range(Job& J, SkArgs a, uint64_t lo, uint64_t hi, DevBuf& rec_b, DevBuf& key_b,
                         unsigned long long* too_long) {
    hipStream_t st = J.st;
    Phase1Bufs& P = J.p1;
    const uint64_t t0 = lo / SK_TILE, t1 = (hi + SK_TILE - 1) / SK_TILE, ntiles = t1 - t0;
    if (ntiles == 0) return 0;
    if (ntiles >= (1ull << 31)) throw Error("input too large for one job");
    a.pos_lo = lo;
    a.pos_hi = hi;
    a.tile0 = t0;
    uint32_t* cnt = P.tile_cnt.ensure<uint32_t>(2 * ntiles);
    uint64_t* off = P.tile_off.ensure<uint64_t>(2 * ntiles);
    uint64_t* tot = P.totals.as<uint64_t>();
    a.cnt_s = cnt;
    a.cnt_e = cnt + ntiles;
    a.st_s = P.st_s.ensure<uint32_t>(ntiles * SK_TILE);
    a.st_e = P.st_e.ensure<uint16_t>(ntiles * SK_TILE);
    { KSCOPE("k_superkmers", st); k_superkmers<<<uint32_t(ntiles), SK_THREADS, 0, st>>>(a); }
    exclusive_scan<uint32_t, uint64_t>(cnt, off, ntiles, tot, J.rs.scan_tmp, st);
    exclusive_scan<uint32_t, uint64_t>(cnt + ntiles, off + ntiles, ntiles, tot + 1, J.rs.scan_tmp, st);
    uint64_t h_tot[2];
    CDBG_HIP(hipMemcpyAsync(h_tot, tot, sizeof h_tot, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    if (h_tot[0] != h_tot[1]) throw Error("phase 1: super-k-mer starts and ends do not pair up");
    const uint64_t n = h_tot[0];
    if (n >= (1ull << 32)) throw Error("phase 1: more than 2^32 super-k-mers in one job");
    SuperRec* rec = (SuperRec*)rec_b.ensure<SuperRec>(n);
    uint32_t* key = key_b.ensure<uint32_t>(n);
    if (n) {
        KSCOPE("k_zip_supers", st);
        k_zip_supers<<<uint32_t(ntiles), 256, 0, st>>>(t0, ntiles, cnt, cnt + ntiles, off, off + ntiles, a.st_s,
                                                      a.st_e, J.k, J.color_off.as<uint64_t>(), J.num_colors, rec,
                                                      key, too_long);
    }
    return n;
}

// Colour ranges of phase 1: one range unless the job's phase-1 working set
// (staging 6 B/base, records and their sort buffers 40 B per super-k-mer,
// the 2-bit bases and the bucket-major records) would not fit 70% of the
// free device memory; then ranges of whole colours whose staging and sort
// buffers fit what the job-wide arrays leave (the plan of the first run is
// kept for the job).
std::vector<std::pair<uint64_t, uint64_t>> phase1_ranges(Job& J) {
    uint64_t cap = J.p1_chunk_bases;
    if (cap == 0) {
        if (J.p1_auto_bases == 0) {
            size_t free_b = 0, total_b = 0;
            CDBG_HIP(hipMemGetInfo(&free_b, &total_b));
            const double usable = double(free_b) + double(g_block_cache.cached.load());
            const double spb = supers_per_base(J.k, J.m), nb = double(J.n_bytes);
            const double per_base = 6.0 + 40.0 * spb, global = 0.375 + 37.0 * spb;
            if (nb * (per_base + global) <= 0.7 * usable) {
                J.p1_auto_bases = ~0ull;
            } else {
                const double left = 0.6 * usable - nb * global;
                J.p1_auto_bases = uint64_t(std::max(double(64ull << 20), left / per_base));
            }
        }
        cap = J.p1_auto_bases;
    }
    std::vector<std::pair<uint64_t, uint64_t>> r;
    const auto& co = J.h_color_off;
    uint64_t lo = 0;
    if (cap < J.n_bytes && co.size() == size_t(J.num_colors) + 1) {
        for (uint32_t c = 1; c < J.num_colors; ++c)
            if (co[c + 1] - lo > cap && co[c] > lo) {
                r.push_back({lo, co[c]});
                lo = co[c];
            }
    }
    r.push_back({lo, J.n_bytes});
    return r;
}

}  // namespace

// Phase 1 over J's staged input on J's stream: fills J.p1's bucket-major
// records and bucket offsets (device and host) and J.n_supers.  Every device
// buffer comes from DevBuf::ensure, so the caller's allocation scope (block
// cache or arena) decides where they live.  A job too large for one pass runs
// in colour ranges: each range's records are sorted by bucket and kept, and
// the ranges are then merged bucket by bucket in range order, which is the
// position order of one pass.
void phase1_core(Job& J) {
    hipStream_t st = J.st;
    Phase1Bufs& P = J.p1;

    const uint64_t ngroups = J.n_pad / 32;
    uint32_t* packed = P.packed.ensure<uint32_t>(ngroups * 2);
    uint32_t* valid = P.valid.ensure<uint32_t>(ngroups);
    { KSCOPE("k_pack2bit", st); k_pack2bit<<<blocks_for(ngroups, 256), 256, 0, st>>>(J.bytes.as<uint8_t>(), ngroups, packed, valid); }

    SkArgs a{};
    a.packed = packed;
    a.valid = valid;
    a.n_bases = J.n_bytes;
    a.n_pk_words = ngroups * 2;
    a.n_vd_words = ngroups;
    a.k = J.k;
    a.m = J.m;
    a.W = J.k - J.m;
    a.B = J.B;
    if (!J.h_cuts.empty() && J.h_cuts.size() != J.B - 1) throw Error("phase 1: the bucket map has the wrong size");
    a.cuts = nullptr;
    if (!J.h_cuts.empty()) {
        uint64_t* d = P.cuts.ensure<uint64_t>(J.h_cuts.size());
        CDBG_HIP(hipMemcpyAsync(d, J.h_cuts.data(), J.h_cuts.size() * 8, hipMemcpyHostToDevice, st));
        a.cuts = d;
    }
    const uint64_t S[4] = {NT_SEED0, NT_SEED1, NT_SEED2, NT_SEED3};
    auto hrot = [](uint64_t x, uint32_t r) { r &= 63; return r ? (x << r) | (x >> (64 - r)) : x; };
    for (int b = 0; b < 4; ++b) {
        a.T.s[b] = S[b];
        a.T.sm[b] = hrot(S[b], J.m);
        a.T.rc1[b] = hrot(S[b ^ 3], 63);
        a.T.rcm[b] = hrot(S[b ^ 3], J.m - 1);
    }
    if (J.B > 65536) throw Error("phase 1: more than 2^16 buckets");
    auto* too_long = (unsigned long long*)(P.totals.ensure<uint64_t>(4) + 2);
    { KSCOPE("k_fill", st); k_fill_u64<<<1, 64, 0, st>>>((uint64_t*)too_long, 1, 0ull); }
    int bbits = 0;
    while ((uint64_t(1) << bbits) < J.B) ++bbits;
    const int dbits = bbits > 10 ? (bbits + 1) / 2 : bbits;
    const auto ranges = phase1_ranges(J);
    J.p1_ranges = uint32_t(ranges.size());
    // a job too large for one pass gives its phase-2 buffers (of an earlier
    // run) back before phase 1 takes its room: the results are being redone
    if (ranges.size() > 1) free_phase2(J);
    uint64_t n = 0;
    if (ranges.size() == 1) {
        // Bucket scatter: stable sort of the records by bucket.
        n = supers_of_range(J, a, 0, J.n_bytes, P.sk_rec, P.key, too_long);
        SuperRec* rec = (SuperRec*)P.sk_rec.ensure<SuperRec>(n);
        SuperRec* rec_t = (SuperRec*)P.sk_rec_t.ensure<SuperRec>(n);
        uint32_t* key = P.key.ensure<uint32_t>(n);
        uint32_t* key_t = P.key_t.ensure<uint32_t>(n);
        SortPlan plan;
        plan.build({0, n});
        bool in_temp = false;
        segmented_radix_sort<uint32_t, SuperRec>(key, rec, key_t, rec_t, plan, 0, bbits, dbits, J.rs, st, "buckets",
                                                 &in_temp);
        if (in_temp) {  // the sorted records are in the second buffers: use them as they are
            P.sk_rec.swap(P.sk_rec_t);
            P.key.swap(P.key_t);
            rec = (SuperRec*)P.sk_rec.p;
            key = (uint32_t*)P.key.p;
        }
        uint64_t* b_start = P.b_start.ensure<uint64_t>(n);
        uint32_t* b_len = P.b_len.ensure<uint32_t>(n);
        uint8_t* b_flags = P.b_flags.ensure<uint8_t>(n);
        uint32_t* b_color = P.b_color.ensure<uint32_t>(n);
        uint64_t* boff = P.bucket_off.ensure<uint64_t>(J.B + 1);
        if (n) { KSCOPE("k_split_supers", st); k_split_supers<<<blocks_for(n, 256), 256, 0, st>>>(n, rec, b_start, b_len, b_flags, b_color); }
        { KSCOPE("k_bucket_offsets", st); k_bucket_offsets<<<blocks_for(std::max<uint64_t>(n, uint64_t(J.B) + 1), 256), 256, 0, st>>>(n, key, J.B, boff); }
        CDBG_HIP(hipGetLastError());
        J.h_bucket_off.resize(J.B + 1);
        CDBG_HIP(hipMemcpyAsync(J.h_bucket_off.data(), boff, (J.B + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    } else {
        // every range sorted by bucket on its own and kept in a buffer of its
        // own (blocks of a range's size are what the block cache reuses for
        // phase-2 chunks and the next run), with its bucket offsets on the host
        const uint64_t B = J.B;
        P.range_rec.clear();
        for (size_t r = 0; r < ranges.size(); ++r) P.range_rec.emplace_back(new DevBuf());
        std::vector<uint64_t> roff{0};
        std::vector<std::vector<uint64_t>> h_off;
        for (size_t r = 0; r < ranges.size(); ++r) {
            const uint64_t nr = supers_of_range(J, a, ranges[r].first, ranges[r].second, P.sk_rec, P.key_t, too_long);
            SuperRec* rec = (SuperRec*)P.sk_rec.ensure<SuperRec>(nr);
            SuperRec* rec_t = (SuperRec*)P.sk_rec_t.ensure<SuperRec>(nr);
            uint32_t* key = P.key_t.ensure<uint32_t>(nr);
            uint32_t* key_t = P.rkey_t.ensure<uint32_t>(nr);
            SortPlan plan;
            plan.build({0, nr});
            bool in_temp = false;
            segmented_radix_sort<uint32_t, SuperRec>(key, rec, key_t, rec_t, plan, 0, bbits, dbits, J.rs, st,
                                                     "buckets", &in_temp);
            if (in_temp) {
                rec = rec_t;
                key = key_t;
            }
            if (nr) {
                SuperRec* kept = (SuperRec*)P.range_rec[r]->ensure<SuperRec>(nr);
                CDBG_HIP(hipMemcpyAsync(kept, rec, nr * sizeof(SuperRec), hipMemcpyDeviceToDevice, st));
            }
            uint64_t* boff = P.bucket_off.ensure<uint64_t>(B + 1);
            { KSCOPE("k_bucket_offsets", st); k_bucket_offsets<<<blocks_for(std::max<uint64_t>(nr, B + 1), 256), 256, 0, st>>>(nr, key, J.B, boff); }
            h_off.emplace_back(B + 1);
            CDBG_HIP(hipMemcpyAsync(h_off.back().data(), boff, (B + 1) * 8, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipStreamSynchronize(st));
            roff.push_back(roff.back() + nr);
        }
        n = roff.back();
        if (n >= (1ull << 32)) throw Error("phase 1: more than 2^32 super-k-mers in one job");
        // bucket b of range r starts after bucket b of every earlier range
        J.h_bucket_off.assign(B + 1, 0);
        for (uint64_t b = 0; b < B; ++b) {
            uint64_t c = 0;
            for (auto& o : h_off) c += o[b + 1] - o[b];
            J.h_bucket_off[b + 1] = J.h_bucket_off[b] + c;
        }
        uint64_t* b_start = P.b_start.ensure<uint64_t>(n);
        uint32_t* b_len = P.b_len.ensure<uint32_t>(n);
        uint8_t* b_flags = P.b_flags.ensure<uint8_t>(n);
        uint32_t* b_color = P.b_color.ensure<uint32_t>(n);
        uint32_t* bucket = P.key.ensure<uint32_t>(n);
        uint64_t* tab = P.range_tab.ensure<uint64_t>(2 * B + 1);
        std::vector<uint64_t> base(J.h_bucket_off.begin(), J.h_bucket_off.end() - 1);
        for (size_t r = 0; r < ranges.size(); ++r) {
            const uint64_t nr = roff[r + 1] - roff[r];
            if (nr) {
                // tab = [off_r (B + 1) | base_r (B)]; uploaded from pageable memory: waited for below
                CDBG_HIP(hipMemcpyAsync(tab, h_off[r].data(), (B + 1) * 8, hipMemcpyHostToDevice, st));
                CDBG_HIP(hipMemcpyAsync(tab + B + 1, base.data(), B * 8, hipMemcpyHostToDevice, st));
                KSCOPE("k_merge_supers", st);
                k_merge_supers<<<blocks_for(nr, 256), 256, 0, st>>>(nr, J.B, P.range_rec[r]->as<SuperRec>(), tab,
                                                                    tab + B + 1, b_start, b_len, b_flags, b_color,
                                                                    bucket);
                CDBG_HIP(hipStreamSynchronize(st));
            }
            for (uint64_t b = 0; b < B; ++b) base[b] += h_off[r][b + 1] - h_off[r][b];
        }
        uint64_t* boff = P.bucket_off.ensure<uint64_t>(B + 1);
        CDBG_HIP(hipMemcpyAsync(boff, J.h_bucket_off.data(), (B + 1) * 8, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipGetLastError());
        CDBG_HIP(hipStreamSynchronize(st));
        // a job this large needs the room in phase 2: the per-range buffers
        // and the kept ranges go back to the cache
        for (DevBuf* b : {&P.st_s, &P.st_e, &P.sk_rec, &P.sk_rec_t, &P.key_t, &P.rkey_t}) b->release();
        P.range_rec.clear();
    }
    J.n_supers = n;
    if (d2h_u64((const uint64_t*)too_long, st)) throw Error("phase 1: a super-k-mer is longer than 2^28 bases");
    CDBG_HIP(hipStreamSynchronize(st));
}

void run_phase1(Job& J) {
    AllocStream alloc_on_job_stream(J.st);
    hipStream_t st = J.st;
    hipEvent_t e0, e1;
    CDBG_HIP(hipEventCreate(&e0));
    CDBG_HIP(hipEventCreate(&e1));
    CDBG_HIP(hipEventRecord(e0, st));
    phase1_core(J);
    CDBG_HIP(hipEventRecord(e1, st));
    CDBG_HIP(hipStreamSynchronize(st));
    float ms = 0;
    CDBG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    J.stats.ms_phase1 = ms;
    J.stats.num_bases = J.n_bytes;
    J.stats.num_supers = J.n_supers;
    J.phase1_done = true;
}

}  // namespace cdbg_amd
