// On-disk artifacts of phases 1 and 2 (the drop-in boundary).
//
// Bucket files (bucket_io.hpp:19-40, super_kmer.hpp:43-68,105-117):
//   frames [u32 uncompressed][u32 compressed][LZ4 block] ... [u32 0];
//   decompressed stream = records
//     varint(n_colors) varint(c0) varint(c1-c0)... varint((len<<4)|flags)
//     ceil(len/4) bytes, base i in bits 2(i%4) of byte i/4.
//
// Writing (the compacting writer, bucket_io.hpp:127-277), all on the device:
//   * records with equal bases, length and ownership bits in one bucket merge
//     into one record carrying the sorted unique colours of all of them and
//     the AND of their run-begin/run-end bits (:145-160, 221-240); records
//     keep the order of their first occurrence (the insertion order of the
//     reference's map).  The reference merges inside a spill window and on
//     the bases alone; merging over the whole bucket only merges more
//     (dedup "is purely a memory bound, never a correctness one",
//     algorithm.md:264-266), and keying on the ownership bits as well keeps
//     the phase-2 result the same as over unmerged records (a one-k-mer
//     record met at a run start and at a run end of one bucket carries
//     different ownership bits, and AND-ing them would drop a colour).
//     Grouping: a 64-bit hash of (bucket, length, ownership, bases) in an
//     open-addressing table whose slot keeps the group's first record;
//     every record is compared base by base with that record, and one that
//     differs (a hash collision) stays a record of its own;
//   * the stream is cut into frames of at most 64 KiB inside each bucket
//     (the reference's spill size, builder.hpp:525), LZ4-compressed one lane
//     per frame, packed with their headers into the files' image, and the
//     host only writes each file's bytes.
//
// Reading (bucket_reader, bucket_io.hpp:651-720): host threads read the
// files; frames are LZ4-decoded on the device into one stream; one thread
// per bucket walks its records (stopping at the first malformed one, like
// read_super_kmer, super_kmer.hpp:121-143) and emits one (record, colour)
// pair per colour of every record with len >= k (load_bucket skips shorter
// ones, bucket_walker.hpp:179); pairs are stably sorted by colour inside
// each bucket.  Each pair carries the stream rank of its record's first
// k-mer, so first-occurrence order is that of the file.  The decoded stream
// itself holds the 2-bit bases phase 2 reads.
//
// Phase-2 files: frag_unitigs.bin (unitig_spill.hpp:491-522), the colour-set
// stream <out>.color_sets as the reference leaves it at the end of phase 2 (a
// 28-byte zero header, then every complete 64-bit word;
// streaming_color_set_dict.hpp:93-97,350-365) and <out>.color_sets.tmp_offsets
// (one u64 bit offset per class, :170-175).  The trailing partial word and
// the totals are returned for the host's finalize step.
#include <hip/hip_runtime.h>
#include <fcntl.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <mutex>
#include <thread>
#include <vector>

#include "job.h"
#include "stitch.h"
#include "kmer_ops.h"
#include "lz4_device.h"

namespace cdbg_amd {

namespace {

constexpr int TPB = 256;

inline uint32_t blocks_for(uint64_t n, int tpb = TPB) {
    return uint32_t(std::max<uint64_t>(1, (n + tpb - 1) / tpb));
}

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

// LEB128 read bounded by `end` (varint_read, super_kmer.hpp:71-82): at most
// ten bytes, a truncated value is what was read so far.
__device__ __forceinline__ uint64_t get_varint(const uint8_t* __restrict__ d, uint64_t end, uint64_t& q) {
    uint64_t v = 0;
    uint32_t sh = 0;
    while (q < end) {
        const uint32_t b = d[q++];
        v |= uint64_t(b & 0x7f) << sh;
        if (!(b & 0x80)) return v;
        sh += 7;
        if (sh >= 64) break;
    }
    return v;
}

// 32 bases from base p of a 2-bit stream (16 bases per u32, first base low).
__device__ __forceinline__ uint64_t window64(const uint32_t* __restrict__ packed, uint64_t p) {
    const uint64_t w = p >> 4;
    const uint32_t sh = 2 * uint32_t(p & 15);
    const uint64_t lo = uint64_t(packed[w]) | (uint64_t(packed[w + 1]) << 32);
    uint64_t v = lo >> sh;
    if (sh) v |= uint64_t(packed[w + 2]) << (64 - sh);
    return v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Largest b with off[b] <= i (off ascending, off[0] = 0, B + 1 entries).
__device__ __forceinline__ uint32_t segment_of(const uint64_t* __restrict__ off, uint32_t B, uint64_t i) {
    uint32_t lo = 0, hi = B;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr uint8_t OWN_BITS = 0xC;  // SK_FLAG_OWNS_FIRST | SK_FLAG_OWNS_LAST (super_kmer.hpp:96-97)

// ---- writer: merge identical records --------------------------------------

struct RecArgs {
    uint64_t n;
    uint32_t B;
    const uint64_t* boff;  // records of bucket b: [boff[b], boff[b+1])
    const uint64_t* start;
    const uint32_t* len;
    const uint8_t* flags;
    const uint32_t* color;
    const uint32_t* packed;
};

__global__ void k_dedup_insert(RecArgs a, uint64_t* __restrict__ tab_key, uint32_t* __restrict__ tab_first,
                               uint64_t mask, uint32_t* __restrict__ rslot, uint32_t* __restrict__ fl_and) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t b = segment_of(a.boff, a.B, i);
    const uint32_t L = a.len[i];
    const uint64_t s = a.start[i];
    uint64_t h = mix64((uint64_t(b) << 36) ^ (uint64_t(L) << 4) ^ (a.flags[i] & OWN_BITS) ^ 0x9e3779b97f4a7c15ull);
    for (uint32_t j = 0; j < L; j += 32) {
        uint64_t w = window64(a.packed, s + j);
        if (L - j < 32) w &= (uint64_t(1) << (2 * (L - j))) - 1;
        h = mix64(h + w * 0xff51afd7ed558ccdull + j);
    }
    if (h == 0) h = 1;  // 0 marks an empty slot
    uint64_t slot = h & mask;
    for (;;) {
        const uint64_t cur = tab_key[slot];
        if (cur == h) break;
        if (cur == 0) {
            const uint64_t prev = atomicCAS((unsigned long long*)&tab_key[slot], 0ull, (unsigned long long)h);
            if (prev == 0 || prev == h) break;
        }
        slot = (slot + 1) & mask;
    }
    atomicMin(&tab_first[slot], uint32_t(i));
    rslot[i] = uint32_t(slot);
    fl_and[i] = 15;
}

// grp[i] = the first record of i's group (i itself when it leads, or when
// the record in its slot differs from it); lkey[i] = grp[i] local to the
// bucket, the sort key that gathers each group's colours.
__global__ void k_dedup_group(RecArgs a, const uint32_t* __restrict__ tab_first, const uint32_t* __restrict__ rslot,
                              uint32_t* __restrict__ grp, uint32_t* __restrict__ lkey, uint32_t* __restrict__ fl_and) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t b = segment_of(a.boff, a.B, i);
    uint64_t g = tab_first[rslot[i]];
    if (g != i) {
        bool same = g >= a.boff[b] && a.len[g] == a.len[i] && ((a.flags[g] ^ a.flags[i]) & OWN_BITS) == 0;
        const uint32_t L = a.len[i];
        for (uint32_t j = 0; same && j < L; j += 32) {
            uint64_t x = window64(a.packed, a.start[g] + j) ^ window64(a.packed, a.start[i] + j);
            if (L - j < 32) x &= (uint64_t(1) << (2 * (L - j))) - 1;
            same = x == 0;
        }
        if (!same) g = i;
    }
    grp[i] = uint32_t(g);
    lkey[i] = uint32_t(g - a.boff[b]);
    atomicAnd(&fl_and[g], uint32_t(a.flags[i]));
}

// After the stable sort by lkey: each group's members are a run of sorted
// positions, in record order.  gspan[g] = (first, end) of group g's run.
__global__ void k_group_span(uint64_t n, uint32_t B, const uint64_t* __restrict__ boff,
                             const uint32_t* __restrict__ sk, uint32_t* __restrict__ gbeg,
                             uint32_t* __restrict__ gend) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t b = segment_of(boff, B, p);
    const uint64_t g = boff[b] + sk[p];
    if (p == boff[b] || sk[p - 1] != sk[p]) gbeg[g] = uint32_t(p);
    if (p + 1 == boff[b + 1] || sk[p + 1] != sk[p]) gend[g] = uint32_t(p + 1);
}

// Record bytes of every group leader (0 for other records).  Colours of a
// group arrive in record order, which is colour order inside a bucket;
// *bad is set when they do not.
__global__ void k_dedup_size(RecArgs a, const uint32_t* __restrict__ grp, const uint32_t* __restrict__ gbeg,
                             const uint32_t* __restrict__ gend, const uint32_t* __restrict__ scol,
                             const uint32_t* __restrict__ fl_and, uint64_t* __restrict__ size,
                             uint32_t* __restrict__ bad) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    if (grp[i] != i) {
        size[i] = 0;
        return;
    }
    uint64_t nc = 0, bytes = 0;
    uint32_t prev = 0;
    for (uint32_t p = gbeg[i]; p < gend[i]; ++p) {
        const uint32_t c = scol[p];
        if (nc && c < prev) *bad = 1;
        if (nc == 0 || c != prev) {
            bytes += varint_len(nc ? c - prev : c);
            ++nc;
            prev = c;
        }
    }
    const uint64_t lf = (uint64_t(a.len[i]) << 4) | (fl_and[i] & 15u);
    size[i] = varint_len(nc) + bytes + varint_len(lf) + (uint64_t(a.len[i]) + 3) / 4;
}

__global__ void k_dedup_write(RecArgs a, const uint32_t* __restrict__ grp, const uint32_t* __restrict__ gbeg,
                              const uint32_t* __restrict__ gend, const uint32_t* __restrict__ scol,
                              const uint32_t* __restrict__ fl_and, const uint64_t* __restrict__ off,
                              uint8_t* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n || grp[i] != i) return;
    const uint32_t g0 = gbeg[i], g1 = gend[i];
    uint64_t nc = 0;
    for (uint32_t p = g0; p < g1; ++p)
        nc += (p == g0 || scol[p] != scol[p - 1]);
    uint8_t* p8 = put_varint(out + off[i], nc);
    for (uint32_t p = g0; p < g1; ++p)
        if (p == g0) p8 = put_varint(p8, scol[p]);
        else if (scol[p] != scol[p - 1]) p8 = put_varint(p8, scol[p] - scol[p - 1]);
    const uint32_t L = a.len[i];
    p8 = put_varint(p8, (uint64_t(L) << 4) | (fl_and[i] & 15u));
    const uint64_t s = a.start[i];
    const uint32_t nb = (L + 3) / 4;
    for (uint32_t j = 0; j < nb; ++j) {
        uint32_t v = uint32_t(window64(a.packed, s + 4ull * j)) & 0xffu;
        const uint32_t rem = L - 4 * j;
        if (rem < 4) v &= (1u << (2 * rem)) - 1;  // unused base slots stay zero
        p8[j] = uint8_t(v);
    }
}

__global__ void k_gather_u64(const uint64_t* __restrict__ idx, uint64_t n, const uint64_t* __restrict__ src,
                             uint64_t* __restrict__ dst) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// ---- reader: walk each bucket's records -----------------------------------
//
// One wavefront per bucket.  The bucket's decoded bytes pass through an LDS
// window; the walk reads 64 bytes at a time (one byte per lane) and finds
// varint ends with a ballot of the bytes whose high bit is clear, so a
// record header (count, colour deltas, length) costs a few ballots however
// many colours it lists.  Every varint of more than ten bytes (only in a
// malformed stream) is read byte by byte with the ten-byte rule of
// varint_read.  Kept records (len >= k) are described in a slot range per
// bucket sized by the smallest kept record; a second kernel compacts the
// descriptors, a third decodes each record's colours into its pairs.

constexpr uint32_t WALK_WIN = 8192;  // LDS window of one walker, bytes

enum : uint32_t { WALK_OK = 0, WALK_RANK = 1, WALK_SLOTS = 2, WALK_BIG = 3 };

struct WalkCounts {
    uint64_t read;   // records parsed
    uint64_t kept;   // records with len >= k
    uint64_t pairs;  // colours of kept records
    uint32_t err, pad;
};

struct RecDesc {  // kept records in bucket order
    uint64_t* cpos;  // colour deltas: byte offset in the decoded stream
    uint64_t* bpos;  // packed bases: byte offset in the decoded stream
    uint32_t* nc;
    uint32_t* len;
    uint8_t* flags;
    uint32_t* rank0;  // stream rank of the record's first k-mer in its bucket
    uint64_t* pair0;  // first pair
};

// Kept-record descriptor as written by a walker: 8 dwords per slot, stored
// by lanes 0-7 in one instruction.
enum : uint32_t { DS_CPOS_LO, DS_CPOS_HI, DS_BDELTA, DS_NC, DS_LEN, DS_RANK0, DS_PAIR0, DS_FLAGS, DS_WORDS };

__global__ __launch_bounds__(64) void k_walk(const uint8_t* __restrict__ d, const uint64_t* __restrict__ dbo,
                                             uint32_t k, const uint64_t* __restrict__ region,
                                             uint32_t* __restrict__ desc, WalkCounts* __restrict__ wc) {
    __shared__ uint32_t win32[WALK_WIN / 4];
    const uint8_t* win = (const uint8_t*)win32;
    const uint32_t b = blockIdx.x, lane = threadIdx.x;
    const uint64_t end = dbo[b + 1];
    uint64_t wbase = 0, wend = 0;  // the window holds bytes [wbase, wend)
    // after need(q): bytes [q, min(q + 64, end)) are in the window
    auto need = [&](uint64_t q) {
        if (q >= wbase && (q + 64 <= wend || wend == end)) return;
        __syncthreads();  // every lane is done with the old window
        wbase = q & ~uint64_t(3);
        wend = min(wbase + WALK_WIN, end);
        const uint32_t nd = uint32_t((wend - wbase + 3) / 4);  // ends <= 3 bytes past `end`: inside the padding
        const uint32_t* src = (const uint32_t*)(d + wbase);
        for (uint32_t i0 = 0; i0 < nd; i0 += 16 * 64) {  // 16 loads in flight per lane
            uint32_t t[16];
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const uint32_t i = i0 + u * 64 + lane;
                t[u] = i < nd ? src[i] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const uint32_t i = i0 + u * 64 + lane;
                if (i < nd) win32[i] = t[u];
            }
        }
        __syncthreads();
    };
    // byte q + lane (0x80, a continuation byte, past the bucket's end) and
    // the mask of varint-terminating bytes among the 64
    auto block = [&](uint64_t q, uint64_t& term) -> uint32_t {
        need(q);
        const uint32_t v = q + lane < end ? uint32_t(win[q + lane - wbase]) : 0x80u;
        term = __ballot(!(v & 0x80u));
        return v;
    };
    // varint_read at q < end (super_kmer.hpp:71-82): up to the terminator,
    // at most ten bytes, at most to the end
    auto varint = [&](uint64_t& q) -> uint64_t {
        uint64_t term;
        const uint32_t v = block(q, term);
        const uint64_t avail = end - q;
        uint32_t nb = term ? uint32_t(__builtin_ctzll(term)) + 1 : 64;
        if (nb > 10) nb = 10;
        if (nb > avail) nb = uint32_t(avail);
        uint64_t x = 0;
        for (uint32_t j = 0; j < nb; ++j) x |= uint64_t(__builtin_amdgcn_readlane(int(v), int(j)) & 0x7f) << (7 * j);
        q += nb;
        return x;
    };
    // skips n varints; false when the stream ends first
    auto skip = [&](uint64_t& q, uint64_t n) -> bool {
        while (n) {
            if (q >= end) return false;
            uint64_t term;
            (void)block(q, term);
            const uint64_t avail = end - q;
            const uint64_t valid = avail >= 64 ? ~0ull : ((1ull << avail) - 1);
            const uint64_t z = ~term & valid;
            uint64_t run = z;  // bit a: bytes a..a+9 all continue a varint
            for (int s = 1; s < 10; ++s) run &= z >> s;
            const uint32_t cnt = uint32_t(__builtin_popcountll(term));
            if (cnt == 0) {  // no end in sight: an overlong varint, or the stream's end
                (void)varint(q);
                --n;
                continue;
            }
            const uint64_t take = n < cnt ? n : cnt;
            uint64_t t = term;
            for (uint64_t i = 1; i < take; ++i) t &= t - 1;
            const uint32_t pos = uint32_t(__builtin_ctzll(t));  // the take-th end
            const uint64_t below = pos >= 63 ? ~0ull : ((1ull << (pos + 1)) - 1);
            if (run & below) {  // an overlong varint among them: one at a time
                (void)varint(q);
                --n;
                continue;
            }
            q += pos + 1;
            n -= take;
        }
        return true;
    };
    WalkCounts c{0, 0, 0, WALK_OK, 0};
    uint64_t slot = region[b], rank = 0;
    const uint64_t slot_end = region[b + 1];
    // one parsed record; false stops the walk
    auto keep = [&](uint64_t cpos, uint64_t nc, uint64_t len, uint32_t fl, uint64_t bpos) -> bool {
        ++c.read;
        if (len < k) return true;
        if (slot >= slot_end) {
            c.err = WALK_SLOTS;
            return false;
        }
        if (nc >= (uint64_t(1) << 32) || len >= (uint64_t(1) << 32) || bpos - cpos >= (uint64_t(1) << 32)) {
            c.err = WALK_BIG;
            return false;
        }
        if (lane < DS_WORDS) {
            const uint32_t x = lane == DS_CPOS_LO ? uint32_t(cpos)
                             : lane == DS_CPOS_HI ? uint32_t(cpos >> 32)
                             : lane == DS_BDELTA  ? uint32_t(bpos - cpos)
                             : lane == DS_NC      ? uint32_t(nc)
                             : lane == DS_LEN     ? uint32_t(len)
                             : lane == DS_RANK0   ? uint32_t(rank)
                             : lane == DS_PAIR0   ? uint32_t(c.pairs)
                                                  : fl;
            desc[slot * DS_WORDS + lane] = x;
        }
        ++slot;
        ++c.kept;
        c.pairs += nc;
        rank += len - k + 1;
        if (rank >= (uint64_t(1) << 32)) {
            c.err = WALK_RANK;
            return false;
        }
        return true;
    };
    // read_super_kmer (super_kmer.hpp:121-143) at q, one block at a time;
    // 0: parsed, 1: malformed (the walk stops there)
    auto general = [&](uint64_t& q) -> int {
        const uint64_t nc = varint(q);
        const uint64_t cpos = q;
        if (!skip(q, nc) || q >= end) return 1;
        const uint64_t lf = varint(q);
        const uint64_t len = lf >> 4;
        const uint64_t nbytes = (len + 3) / 4;
        if (q + nbytes > end) return 1;
        if (!keep(cpos, nc, len, uint32_t(lf & 15u), q)) return 1;
        q += nbytes;
        return 0;
    };
    uint64_t q = dbo[b];
    bool stop = false;
    while (!stop && q < end) {
        // every record whose header (count, colours, length) ends inside the
        // 64 bytes at q is parsed from this one ballot
        uint64_t T;
        const uint32_t v = block(q, T);
        const uint64_t avail = end - q;
        const uint64_t valid = avail >= 64 ? ~0ull : ((1ull << avail) - 1);
        const uint64_t Z = ~T & valid;
        uint64_t R = Z;  // bit a: bytes a..a+9 all continue a varint
        for (int s = 1; s < 10; ++s) R &= Z >> s;
        uint32_t o = 0;
        bool slow = false;
        while (o < 64 && o < avail) {
            const uint64_t T0 = T >> o;
            if (!T0) {
                slow = true;
                break;
            }
            const uint32_t t0 = o + uint32_t(__builtin_ctzll(T0));  // end of the colour count
            if (t0 - o >= 10) {
                slow = true;
                break;
            }
            uint64_t nc = 0;
            for (uint32_t j = o; j <= t0; ++j) nc |= uint64_t(__builtin_amdgcn_readlane(int(v), int(j)) & 0x7f) << (7 * (j - o));
            uint64_t rest = t0 >= 63 ? 0 : (T & (~0ull << (t0 + 1)));
            if (uint64_t(__builtin_popcountll(rest)) < nc + 1) {
                slow = true;
                break;
            }
            uint32_t t1 = t0;  // end of the last colour
            for (uint64_t i = 0; i < nc; ++i) {
                t1 = uint32_t(__builtin_ctzll(rest));
                rest &= rest - 1;
            }
            const uint32_t t2 = uint32_t(__builtin_ctzll(rest));  // end of the length
            const uint32_t span = t2 - o + 1;
            if (t2 - t1 > 10 || ((R >> o) & (span >= 64 ? ~0ull : ((1ull << span) - 1)))) {
                slow = true;  // a varint of more than ten bytes
                break;
            }
            uint64_t lf = 0;
            for (uint32_t j = t1 + 1; j <= t2; ++j)
                lf |= uint64_t(__builtin_amdgcn_readlane(int(v), int(j)) & 0x7f) << (7 * (j - t1 - 1));
            const uint64_t len = lf >> 4;
            const uint64_t nbytes = (len + 3) / 4;
            const uint64_t bpos = q + t2 + 1;
            if (bpos + nbytes > end || !keep(q + t0 + 1, nc, len, uint32_t(lf & 15u), bpos)) {
                stop = true;
                break;
            }
            const uint64_t next = uint64_t(t2) + 1 + nbytes;
            if (next >= 64) {
                o = 64;
                q += next;
                break;
            }
            o = uint32_t(next);
        }
        if (stop) break;
        if (o < 64) q += o;
        if (slow && q < end) stop = general(q) != 0;
    }
    if (lane == 0) wc[b] = c;
}

// Descriptor slots -> records in bucket order (r = rec_base[b] + local).
__global__ void k_walk_compact(uint64_t n_slots, uint32_t B, const uint64_t* __restrict__ region,
                               const uint64_t* __restrict__ rec_base, const uint64_t* __restrict__ pair_base,
                               const uint32_t* __restrict__ desc, RecDesc r, uint32_t* __restrict__ rec_bucket) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_slots) return;
    const uint32_t b = segment_of(region, B, i);
    const uint64_t local = i - region[b];
    if (local >= rec_base[b + 1] - rec_base[b]) return;
    const uint64_t j = rec_base[b] + local;
    const uint32_t* x = desc + i * DS_WORDS;
    const uint64_t cpos = uint64_t(x[DS_CPOS_LO]) | (uint64_t(x[DS_CPOS_HI]) << 32);
    r.cpos[j] = cpos;
    r.bpos[j] = cpos + x[DS_BDELTA];
    r.nc[j] = x[DS_NC];
    r.len[j] = x[DS_LEN];
    r.flags[j] = uint8_t(x[DS_FLAGS]);
    r.rank0[j] = x[DS_RANK0];
    r.pair0[j] = pair_base[b] + x[DS_PAIR0];
    rec_bucket[j] = b;
}

// Colours of record j into its pairs; *bad = min (bucket << 32 | colour)
// over colours >= num_colors.
__global__ void k_walk_pairs(uint64_t nr, const uint8_t* __restrict__ d, const uint64_t* __restrict__ dbo,
                             const RecDesc r, const uint32_t* __restrict__ rec_bucket, uint32_t num_colors,
                             uint32_t* __restrict__ pair_color, uint32_t* __restrict__ pair_rec,
                             unsigned long long* __restrict__ bad) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nr) return;
    const uint32_t b = rec_bucket[j];
    const uint64_t end = dbo[b + 1];
    uint64_t q = r.cpos[j], p = r.pair0[j];
    uint32_t col = 0;
    for (uint32_t i = 0; i < r.nc[j]; ++i, ++p) {
        col += uint32_t(get_varint(d, end, q));
        if (col >= num_colors) atomicMin(bad, (unsigned long long)((uint64_t(b) << 32) | col));
        pair_color[p] = col;
        pair_rec[p] = uint32_t(j);
    }
}

__global__ void k_pair_gather(uint64_t n, const uint32_t* __restrict__ pair_rec, const RecDesc r,
                              const uint32_t* __restrict__ rec_bucket, uint64_t* __restrict__ start,
                              uint32_t* __restrict__ len, uint8_t* __restrict__ flags, uint32_t* __restrict__ rank0,
                              uint32_t* __restrict__ bucket) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t j = pair_rec[p];
    start[p] = 4 * r.bpos[j];
    len[p] = r.len[j];
    flags[p] = r.flags[j];
    rank0[p] = r.rank0[j];
    bucket[p] = rec_bucket[j];
}

uint32_t bits_for(uint64_t max_value) {  // bits to hold every value in [0, max_value]
    uint32_t b = 1;
    while (b < 64 && (max_value >> b)) ++b;
    return b;
}

std::string bucket_path(const std::string& dir, uint32_t b) { return dir + "/bucket_" + std::to_string(b) + ".bin"; }

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

// Sorts (keys, vals) stably by the low `bits` key bits inside each segment.
void sort_in_segments(uint32_t* keys, uint32_t* vals, uint64_t n, const std::vector<uint64_t>& seg_off,
                      uint32_t bits, RadixScratch& S, hipStream_t st, const char* tag) {
    if (n == 0) return;
    SortPlan plan;
    plan.build(seg_off);
    DevBuf kt, vt;
    const int digit = bits <= 8 ? int(bits) : int((bits + 1) / 2) <= 11 ? int((bits + 1) / 2) : 11;
    segmented_radix_sort<uint32_t, uint32_t>(keys, vals, kt.ensure<uint32_t>(n), vt.ensure<uint32_t>(n), plan, 0,
                                             int(bits), digit, S, st, tag);
}

}  // namespace

void write_bucket_files(Job& J, const std::string& dir, uint32_t threads, uint64_t* unc_bytes) {
    AllocStream alloc_on_job_stream(J.st);
    if (!J.phase1_done || !J.bytes.p) throw Error("bucket files need phase 1 of a sequence job");
    hipStream_t st = J.st;
    Phase1Bufs& Q = J.p1;
    const uint32_t B = J.B;
    const uint64_t n = J.n_supers;
    if (n >= (uint64_t(1) << 32)) throw Error("bucket files: more than 2^32 super-k-mer records in one job");
    const std::vector<uint64_t>& bo = J.h_bucket_off;
    uint64_t max_bucket = 1;
    for (uint32_t b = 0; b < B; ++b) max_bucket = std::max(max_bucket, bo[b + 1] - bo[b]);
    DevBuf d_boff;
    uint64_t* boff = d_boff.ensure<uint64_t>(B + 1);
    CDBG_HIP(hipMemcpyAsync(boff, bo.data(), (B + 1) * 8, hipMemcpyHostToDevice, st));
    RecArgs a{n, B, boff, Q.b_start.as<uint64_t>(), Q.b_len.as<uint32_t>(), Q.b_flags.as<uint8_t>(),
              Q.b_color.as<uint32_t>(), Q.packed.as<uint32_t>()};
    // 1. groups of identical records
    uint64_t slots = 1024;
    while (slots < 2 * n) slots <<= 1;
    DevBuf d_tkey, d_tfirst, d_rslot, d_grp, d_lkey, d_fl, d_col;
    uint64_t* tkey = d_tkey.ensure<uint64_t>(slots);
    uint32_t* tfirst = d_tfirst.ensure<uint32_t>(slots);
    uint32_t* rslot = d_rslot.ensure<uint32_t>(n);
    uint32_t* grp = d_grp.ensure<uint32_t>(n);
    uint32_t* lkey = d_lkey.ensure<uint32_t>(n);
    uint32_t* fl = d_fl.ensure<uint32_t>(n);
    uint32_t* col = d_col.ensure<uint32_t>(n);
    CDBG_HIP(hipMemsetAsync(tkey, 0, slots * 8, st));
    CDBG_HIP(hipMemsetAsync(tfirst, 0xff, slots * 4, st));
    if (n) {
        { KSCOPE("k_dedup_insert", st); k_dedup_insert<<<blocks_for(n), TPB, 0, st>>>(a, tkey, tfirst, slots - 1, rslot, fl); }
        { KSCOPE("k_dedup_group", st); k_dedup_group<<<blocks_for(n), TPB, 0, st>>>(a, tfirst, rslot, grp, lkey, fl); }
        CDBG_HIP(hipGetLastError());
        CDBG_HIP(hipMemcpyAsync(col, a.color, n * 4, hipMemcpyDeviceToDevice, st));
    }
    d_tkey.release();
    d_tfirst.release();
    d_rslot.release();
    // 2. each group's colours, in record order, as one run
    sort_in_segments(lkey, col, n, bo, bits_for(max_bucket - 1), J.rs, st, "dedup");
    DevBuf d_gbeg, d_gend, d_size, d_off, d_bad, scan_tmp;
    uint32_t* gbeg = d_gbeg.ensure<uint32_t>(n);
    uint32_t* gend = d_gend.ensure<uint32_t>(n);
    uint64_t* size = d_size.ensure<uint64_t>(n);
    uint64_t* off = d_off.ensure<uint64_t>(n + 1);
    uint32_t* bad = d_bad.ensure<uint32_t>(1);
    CDBG_HIP(hipMemsetAsync(bad, 0, 4, st));
    if (n) {
        { KSCOPE("k_group_span", st); k_group_span<<<blocks_for(n), TPB, 0, st>>>(n, B, boff, lkey, gbeg, gend); }
        { KSCOPE("k_dedup_size", st); k_dedup_size<<<blocks_for(n), TPB, 0, st>>>(a, grp, gbeg, gend, col, fl, size, bad); }
        CDBG_HIP(hipGetLastError());
    }
    exclusive_scan<uint64_t, uint64_t>(size, off, n, off + n, scan_tmp, st);
    // byte range of every bucket in the merged stream
    std::vector<uint64_t> bb(B + 1);
    DevBuf d_bb;
    uint64_t* dbb = d_bb.ensure<uint64_t>(B + 1);
    k_gather_u64<<<blocks_for(B + 1), TPB, 0, st>>>(boff, B + 1, off, dbb);
    CDBG_HIP(hipGetLastError());
    uint32_t h_bad = 0;
    CDBG_HIP(hipMemcpyAsync(bb.data(), dbb, (B + 1) * 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    if (h_bad) throw Error("bucket files: the records of a bucket are not in colour order");
    const uint64_t total = bb[B];
    DevBuf d_out;
    uint8_t* out = d_out.ensure<uint8_t>(total + 16);
    if (n) {
        KSCOPE("k_dedup_write", st);
        k_dedup_write<<<blocks_for(n), TPB, 0, st>>>(a, grp, gbeg, gend, col, fl, off, out);
    }
    CDBG_HIP(hipGetLastError());
    for (DevBuf* d : {&d_grp, &d_lkey, &d_fl, &d_col, &d_gbeg, &d_gend, &d_size, &d_off}) d->release();
    // 3. frames of at most LZ_FRAME bytes inside each bucket, compressed
    std::vector<uint64_t> foff{0}, slot{0};
    std::vector<uint32_t> bframe(B + 1, 0);
    for (uint32_t b = 0; b < B; ++b) {
        bframe[b] = uint32_t(foff.size() - 1);
        for (uint64_t o = bb[b]; o < bb[b + 1]; o += LZ_FRAME) {
            const uint64_t e = std::min(o + LZ_FRAME, bb[b + 1]);
            foff.push_back(e);
            slot.push_back(slot.back() + lz4_slot(e - o));
        }
    }
    const uint32_t nf = uint32_t(foff.size() - 1);
    bframe[B] = nf;
    DevBuf d_ftab, d_slots, d_cs;
    uint64_t* ftab = d_ftab.ensure<uint64_t>(3 * (uint64_t(nf) + 1));
    uint64_t* d_foff = ftab;
    uint64_t* d_slot = ftab + nf + 1;
    uint64_t* d_ooff = ftab + 2 * (uint64_t(nf) + 1);
    uint8_t* slots_buf = d_slots.ensure<uint8_t>(slot[nf] + 16);
    uint32_t* cs = d_cs.ensure<uint32_t>(nf + 1);
    CDBG_HIP(hipMemcpyAsync(d_foff, foff.data(), (nf + 1) * 8, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipMemcpyAsync(d_slot, slot.data(), (nf + 1) * 8, hipMemcpyHostToDevice, st));
    launch_lz4_compress(out, d_foff, nf, slots_buf, d_slot, cs, st);
    std::vector<uint32_t> h_cs(nf);
    if (nf) CDBG_HIP(hipMemcpyAsync(h_cs.data(), cs, nf * 4ull, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    // 4. file images: frames with headers, a 4-byte end marker per bucket
    std::vector<uint64_t> ooff(nf + 1), fbase(B + 1);
    uint64_t pos = 0;
    for (uint32_t b = 0; b < B; ++b) {
        fbase[b] = pos;
        for (uint32_t f = bframe[b]; f < bframe[b + 1]; ++f) {
            if (h_cs[f] > slot[f + 1] - slot[f]) throw Error("lz4: a compressed frame overran its slot");
            ooff[f] = pos;
            pos += 8 + h_cs[f];
        }
        pos += 4;
    }
    fbase[B] = pos;
    DevBuf d_img;
    uint8_t* img = d_img.ensure<uint8_t>(pos + 16);
    CDBG_HIP(hipMemcpyAsync(d_ooff, ooff.data(), nf * 8ull, hipMemcpyHostToDevice, st));
    launch_lz4_pack(slots_buf, d_slot, cs, d_foff, nf, img, d_ooff, st);
    PinnedBuf h_img;
    h_img.ensure_keep(pos + 16, 0);
    if (pos) CDBG_HIP(hipMemcpyAsync(h_img.p, img, pos, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    std::filesystem::create_directories(dir);
    parallel_for(B, threads, [&](uint32_t b) {
        std::memset(h_img.p + fbase[b + 1] - 4, 0, 4);  // [u32 0]: end of stream
        const std::string path = bucket_path(dir, b);
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) throw Error("cannot open bucket file: " + path);
        const size_t m = fbase[b + 1] - fbase[b];
        const bool ok = std::fwrite(h_img.p + fbase[b], 1, m, f) == m;
        if ((std::fclose(f) != 0) || !ok) throw Error("short write to " + path);
        if (unc_bytes) unc_bytes[b] = bb[b + 1] - bb[b];
    });
    if (J.prof.on) J.prof.collect();  // the stream is idle: every launch above has finished
}

// With delete_consumed the bucket files are removed once the whole load has
// succeeded (process_buckets' delete_consumed_buckets,
// bucket_walker.hpp:684-691: scratch does not keep what phase 2 has read).
void load_bucket_files(Job& J, const std::string& dir, uint32_t threads, bool delete_consumed) {
    AllocStream alloc_on_job_stream(J.st);
    hipStream_t st = J.st;
    const uint32_t B = J.B, k = J.k;
    // 1. every file's bytes into one pinned image
    std::vector<uint64_t> fbase(B + 1, 0);
    for (uint32_t b = 0; b < B; ++b) {
        std::error_code ec;
        const uint64_t sz = std::filesystem::file_size(bucket_path(dir, b), ec);
        if (ec) throw Error("cannot open " + bucket_path(dir, b));
        fbase[b + 1] = fbase[b] + sz;
    }
    PinnedBuf img;
    img.ensure_keep(fbase[B] + 16, 0);
    // frames of each bucket: (block position in the image, block size, decoded size)
    struct Frame {
        uint64_t pos;
        uint32_t c, u;
    };
    std::vector<std::vector<Frame>> frames(B);
    parallel_for(B, threads, [&](uint32_t b) {
        const std::string path = bucket_path(dir, b);
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw Error("cannot open " + path);
        const uint64_t sz = fbase[b + 1] - fbase[b];
        const bool ok = std::fread(img.p + fbase[b], 1, sz, f) == sz;
        std::fclose(f);
        if (!ok) throw Error("short read of " + path);
        const uint8_t* p = img.p + fbase[b];
        uint64_t q = 0;
        for (;;) {
            uint32_t u = 0, c = 0;
            if (q + 4 > sz) break;  // a missing end marker is tolerated (bucket_io.hpp:675-679)
            std::memcpy(&u, p + q, 4);
            if (u == 0) break;
            if (q + 8 > sz) throw Error("short read of compressed_size on " + path);
            std::memcpy(&c, p + q + 4, 4);
            if (q + 8 + c > sz) throw Error("short read of compressed payload on " + path);
            frames[b].push_back({fbase[b] + q + 8, c, u});
            q += 8 + uint64_t(c);
        }
    });
    // 2. decode every frame on the device into one stream
    std::vector<uint64_t> soff, doff{0}, dbo(B + 1, 0);
    std::vector<uint32_t> csz, fbucket;
    for (uint32_t b = 0; b < B; ++b) {
        for (const Frame& fr : frames[b]) {
            soff.push_back(fr.pos);
            csz.push_back(fr.c);
            doff.push_back(doff.back() + fr.u);
            fbucket.push_back(b);
        }
        dbo[b + 1] = doff.back();
        std::vector<Frame>().swap(frames[b]);
    }
    const uint32_t nf = uint32_t(csz.size());
    const uint64_t D = doff.back();
    Phase1Bufs& Q = J.p1;
    DevBuf d_img, d_ftab;
    uint8_t* dimg = d_img.ensure<uint8_t>(fbase[B] + 16);
    if (fbase[B]) CDBG_HIP(hipMemcpyAsync(dimg, img.p, fbase[B], hipMemcpyHostToDevice, st));
    uint64_t* ftab = d_ftab.ensure<uint64_t>(2 * uint64_t(nf) + 1 + (B + 1) + (nf + 7) / 8 + nf / 2 + 2);
    uint64_t* d_soff = ftab;
    uint64_t* d_doff = d_soff + nf;
    uint64_t* d_dbo = d_doff + nf + 1;
    uint32_t* d_cs = (uint32_t*)(d_dbo + B + 1);
    uint8_t* d_bad = (uint8_t*)(d_cs + nf);
    if (nf) {
        CDBG_HIP(hipMemcpyAsync(d_soff, soff.data(), nf * 8ull, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipMemcpyAsync(d_cs, csz.data(), nf * 4ull, hipMemcpyHostToDevice, st));
    }
    CDBG_HIP(hipMemcpyAsync(d_doff, doff.data(), (nf + 1) * 8ull, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipMemcpyAsync(d_dbo, dbo.data(), (B + 1) * 8ull, hipMemcpyHostToDevice, st));
    uint8_t* dec = Q.packed.ensure<uint8_t>(D + 64);
    CDBG_HIP(hipMemsetAsync(dec + D, 0, 64, st));  // k-mer windows read past the last base
    launch_lz4_decompress(dimg, d_soff, d_cs, nf, dec, d_doff, d_bad, st);
    std::vector<uint8_t> bad(nf);
    if (nf) CDBG_HIP(hipMemcpyAsync(bad.data(), d_bad, nf, hipMemcpyDeviceToHost, st));
    // 3. walk each bucket: kept records into slot ranges sized by the
    //    smallest kept record (count byte, length byte, ceil(k/4) bases)
    const uint64_t min_rec = 2 + (k + 3) / 4;
    std::vector<uint64_t> region(B + 1, 0);
    for (uint32_t b = 0; b < B; ++b) region[b + 1] = region[b] + (dbo[b + 1] - dbo[b]) / min_rec + 1;
    const uint64_t ns = region[B];
    DevBuf d_region, d_wc, d_desc;
    uint64_t* dreg = d_region.ensure<uint64_t>(3 * (uint64_t(B) + 1));
    CDBG_HIP(hipMemcpyAsync(dreg, region.data(), (B + 1) * 8ull, hipMemcpyHostToDevice, st));
    WalkCounts* wc = d_wc.ensure<WalkCounts>(B);
    uint32_t* desc = d_desc.ensure<uint32_t>(ns * DS_WORDS);
    if (B) {
        KSCOPE("k_walk", st);
        k_walk<<<B, 64, 0, st>>>(dec, d_dbo, k, dreg, desc, wc);
    }
    CDBG_HIP(hipGetLastError());
    std::vector<WalkCounts> h_wc(B);
    CDBG_HIP(hipMemcpyAsync(h_wc.data(), wc, B * sizeof(WalkCounts), hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    for (uint32_t f = 0; f < nf; ++f)
        if (bad[f]) throw Error("LZ4 decode failed on " + bucket_path(dir, fbucket[f]));
    std::vector<uint64_t> rec_base(B + 1, 0), pair_base(B + 1, 0);
    for (uint32_t b = 0; b < B; ++b) {
        const WalkCounts& c = h_wc[b];
        if (c.err == WALK_RANK)
            throw Error("bucket holds more than 2^32 k-mer occurrences: " + bucket_path(dir, b));
        if (c.err == WALK_BIG) throw Error("a record of 4 GiB or more in " + bucket_path(dir, b));
        if (c.err != WALK_OK) throw Error("bucket walk overran its record slots: " + bucket_path(dir, b));
        rec_base[b + 1] = rec_base[b] + c.kept;
        pair_base[b + 1] = pair_base[b] + c.pairs;
        J.stats.num_supers += c.read;
    }
    const uint64_t nr = rec_base[B], np = pair_base[B];
    if (np >= (1ull << 32)) throw Error("more than 2^32 (record, colour) pairs in one job");
    // 4. records in bucket order, then one pair per colour
    CDBG_HIP(hipMemcpyAsync(dreg + B + 1, rec_base.data(), (B + 1) * 8ull, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipMemcpyAsync(dreg + 2 * (B + 1), pair_base.data(), (B + 1) * 8ull, hipMemcpyHostToDevice, st));
    DevBuf r_cpos, r_bpos, r_nc, r_len, r_fl, r_rank, r_pair, r_bkt, d_pcol, d_prec, d_bad2;
    RecDesc rd{r_cpos.ensure<uint64_t>(nr), r_bpos.ensure<uint64_t>(nr), r_nc.ensure<uint32_t>(nr),
               r_len.ensure<uint32_t>(nr), r_fl.ensure<uint8_t>(nr), r_rank.ensure<uint32_t>(nr),
               r_pair.ensure<uint64_t>(nr)};
    uint32_t* rbkt = r_bkt.ensure<uint32_t>(nr);
    uint32_t* pcol = d_pcol.ensure<uint32_t>(np);
    uint32_t* prec = d_prec.ensure<uint32_t>(np);
    auto* bad2 = d_bad2.ensure<unsigned long long>(1);
    CDBG_HIP(hipMemsetAsync(bad2, 0xff, 8, st));
    if (ns) {
        KSCOPE("k_walk_compact", st);
        k_walk_compact<<<blocks_for(ns), TPB, 0, st>>>(ns, B, dreg, dreg + B + 1, dreg + 2 * (B + 1), desc, rd, rbkt);
    }
    CDBG_HIP(hipGetLastError());
    d_desc.release();
    d_img.release();
    if (nr) {
        KSCOPE("k_walk_pairs", st);
        k_walk_pairs<<<blocks_for(nr), TPB, 0, st>>>(nr, dec, d_dbo, rd, rbkt, J.num_colors, pcol, prec, bad2);
    }
    CDBG_HIP(hipGetLastError());
    unsigned long long h_bad2 = 0;
    CDBG_HIP(hipMemcpyAsync(&h_bad2, bad2, 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    if (h_bad2 != ~0ull)
        throw Error("bucket " + std::to_string(h_bad2 >> 32) + " names colour " +
                    std::to_string(h_bad2 & 0xffffffffu) + " >= num_colors");
    // 5. colour-major inside each bucket, record order kept among equal colours
    sort_in_segments(pcol, prec, np, pair_base, bits_for(std::max(1u, J.num_colors) - 1), J.rs, st, "pairs");
    uint64_t* start = Q.b_start.ensure<uint64_t>(np);
    uint32_t* len = Q.b_len.ensure<uint32_t>(np);
    uint8_t* flags = Q.b_flags.ensure<uint8_t>(np);
    uint32_t* rank0 = Q.rank0.ensure<uint32_t>(np);
    uint32_t* bucket = Q.key.ensure<uint32_t>(np);
    uint32_t* color = Q.b_color.ensure<uint32_t>(np);
    if (np) {
        { KSCOPE("k_pair_gather", st); k_pair_gather<<<blocks_for(np), TPB, 0, st>>>(np, prec, rd, rbkt, start, len, flags, rank0, bucket); }
        CDBG_HIP(hipGetLastError());
        CDBG_HIP(hipMemcpyAsync(color, pcol, np * 4, hipMemcpyDeviceToDevice, st));
    }
    uint64_t* bucket_off = Q.bucket_off.ensure<uint64_t>(B + 1);
    CDBG_HIP(hipMemcpyAsync(bucket_off, pair_base.data(), (B + 1) * 8ull, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipStreamSynchronize(st));  // host tables are pageable and local
    J.h_bucket_off = pair_base;
    J.n_supers = np;
    J.phase1_done = true;
    J.from_files = true;
    if (delete_consumed)
        for (uint32_t b = 0; b < B; ++b) {
            std::error_code ec;
            std::filesystem::remove(bucket_path(dir, b), ec);
        }
}

namespace {

// Pinned memory is mapped for the host on first touch, a fault per page:
// a fresh 3.5 GB block the device filled took seconds to write out, its
// pages faulting one by one under the file's write lock.  Device bytes are
// written through two 64 MiB pinned buffers kept for the process (the copy
// of one chunk overlaps the write of the other); bytes already in pinned
// host memory have their pages touched by `threads` threads first.
constexpr uint64_t STAGE = uint64_t(64) << 20;

struct StagePool {
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
        CDBG_HIP(hipHostMalloc(&p, STAGE, hipHostMallocDefault));
        return (uint8_t*)p;
    }
    void give(uint8_t* p) {
        std::lock_guard<std::mutex> lk(mu);
        free.push_back(p);
    }
};
StagePool g_stage;

void write_device_stream(const std::string& path, const uint8_t* d, uint64_t n, hipStream_t st) {
    struct Bufs {
        uint8_t* b[2];
        hipEvent_t ev[2] = {nullptr, nullptr};
        hipStream_t st = nullptr;
        ~Bufs() {
            if (st) (void)hipStreamSynchronize(st);  // no copy still writes a buffer given back (error paths)
            for (int i = 0; i < 2; ++i) {
                if (ev[i]) (void)hipEventDestroy(ev[i]);
                if (b[i]) g_stage.give(b[i]);
            }
        }
    } B{{nullptr, nullptr}};
    B.st = st;
    for (int i = 0; i < 2; ++i) {
        B.b[i] = g_stage.take();
        CDBG_HIP(hipEventCreateWithFlags(&B.ev[i], hipEventDisableTiming));
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error("cannot open " + path);
    if (n) (void)posix_fallocate(fileno(f), 0, off_t(n));  // the extent up front (best effort)
    bool ok = true;
    try {
        const uint64_t nc = (n + STAGE - 1) / STAGE;
        auto issue = [&](uint64_t c) {
            const uint64_t len = std::min<uint64_t>(STAGE, n - c * STAGE);
            CDBG_HIP(hipMemcpyAsync(B.b[c & 1], d + c * STAGE, len, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipEventRecord(B.ev[c & 1], st));
        };
        if (nc) issue(0);
        for (uint64_t c = 0; c < nc && ok; ++c) {
            if (c + 1 < nc) issue(c + 1);  // its buffer was written out last round
            CDBG_HIP(hipEventSynchronize(B.ev[c & 1]));
            const uint64_t len = std::min<uint64_t>(STAGE, n - c * STAGE);
            ok = std::fwrite(B.b[c & 1], 1, len, f) == len;
        }
        CDBG_HIP(hipStreamSynchronize(st));
    } catch (...) {
        std::fclose(f);
        throw;
    }
    if (std::fclose(f) != 0 || !ok) throw Error("short write to " + path);
}

void write_host_stream(const std::string& path, const uint8_t* h, uint64_t n, uint32_t threads) {
    const uint64_t pages = (n + 4095) / 4096;
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(threads, pages / 1024 + 1));
    std::vector<std::thread> th;
    std::atomic<uint64_t> sink{0};
    for (uint32_t t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            uint64_t acc = 0;
            for (uint64_t p = pages * t / T; p < pages * (t + 1) / T; ++p) acc += ((const volatile uint8_t*)h)[p * 4096];
            sink += acc;
        });
    for (auto& x : th) x.join();
    write_file_parallel(path.c_str(), h, n, 1);
}

}  // namespace

void write_phase2_files(Job& J, const std::string& frag_path, const std::string& cs_path,
                        cdbg_gpu_color_state* state) {
    AllocStream alloc_on_job_stream(J.st);
    if (!J.phase2_done) throw Error("phase 2 has not run");
    const cdbg_gpu_stats& s = J.stats;
    const uint64_t nw = (s.cs_bits + 63) / 64;
    std::vector<uint64_t> words(nw), offs(s.num_color_sets);
    copy_phase2_results(J, nullptr, words.data(), offs.data());
    auto write_all = [](const std::string& path, const void* p, size_t n, const char* mode) {
        FILE* f = std::fopen(path.c_str(), mode);
        if (!f) throw Error("cannot open " + path);
        const bool ok = (n == 0 || std::fwrite(p, 1, n, f) == n);
        if (std::fclose(f) != 0 || !ok) throw Error("short write to " + path);
    };
    if (!frag_path.empty()) {
        bool on_device = false;
        uint64_t n = 0;
        const uint8_t* src = phase2_frags_where(J, on_device, n);
        if (on_device)
            write_device_stream(frag_path, src, n, J.st);
        else
            write_host_stream(frag_path, src, n, J.io_threads);
    }
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
