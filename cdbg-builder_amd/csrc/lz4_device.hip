// LZ4 block codec on the device, for the frames of bucket files.
//
// Format: the public LZ4 block format (sequences of a token, literal-length
// bytes, literals, a 16-bit little-endian match offset and match-length
// bytes; matches of at least 4 bytes; the last 5 bytes of a block are
// literals and the last match starts at least 12 bytes before its end).  A
// bucket file frames each block as [u32 uncompressed][u32 compressed][block]
// (bucket_io.hpp:19-40, written with LZ4_compress_default :255-265 and read
// with LZ4_decompress_safe :695-696); any valid block is accepted.
//
// Parallel layout:
//   k_lz4_compress   one wavefront per frame of at most LZ_FRAME (64 KiB, so
//                    every position and match offset fits 16 bits); the frame
//                    and a 2048-entry hash table live in LDS; matches are
//                    found 64 positions at a time and taken greedily.
//   k_lz4_decompress one lane per frame, any frame size; every length,
//                    offset and capacity is checked, and a malformed frame is
//                    flagged, never read or written outside its ranges.
//   k_lz4_pack       one workgroup per frame: header + block into the
//                    bucket-file image.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "job.h"
#include "lz4_device.h"

namespace cdbg_amd {

namespace {

constexpr int LZ_HASH_LOG = 11;  // 8 KiB table + 64 KiB frame: two workgroups per CU
constexpr int LZ_HASH = 1 << LZ_HASH_LOG;
constexpr int LZ_LANES = 64;
constexpr uint32_t LZ_MIN_MATCH = 4, LZ_LAST_LITERALS = 5, LZ_MF_LIMIT = 12;
constexpr uint32_t LZ_CAP = 64;  // match bytes each lane measures; longer matches are extended by the wave

// Frame bytes staged in LDS as dwords; frame byte p sits at LDS byte p + sh.
__device__ __forceinline__ uint32_t rd32(const uint32_t* fr, uint32_t a) {
    return __builtin_amdgcn_alignbyte(fr[(a >> 2) + 1], fr[a >> 2], a & 3);
}
__device__ __forceinline__ uint32_t rd8(const uint32_t* fr, uint32_t a) { return (fr[a >> 2] >> (8 * (a & 3))) & 0xffu; }

// Length bytes past a token's 15: r / 255 bytes of 255, then r % 255.
__device__ __forceinline__ void put_run(uint8_t* o, uint32_t& op, uint32_t r, uint32_t lane) {
    const uint32_t full = r / 255, nb = full + 1;
    for (uint32_t i = lane; i < nb; i += LZ_LANES) o[op + i] = i < full ? uint8_t(255) : uint8_t(r - 255 * full);
    op += nb;
}

// One sequence, written by the whole wave (op is the same in every lane):
// the literals at frame bytes [a, a + nlit), then (mlen > 0) a match of mlen
// bytes at distance off.
__device__ __forceinline__ void emit_seq(uint8_t* o, uint32_t& op, const uint32_t* fr, uint32_t sh, uint32_t a,
                                         uint32_t nlit, uint32_t off, uint32_t mlen, uint32_t lane) {
    const uint32_t ml = mlen ? mlen - LZ_MIN_MATCH : 0;
    if (lane == 0) o[op] = uint8_t(((nlit >= 15 ? 15u : nlit) << 4) | (mlen ? (ml >= 15 ? 15u : ml) : 0u));
    op += 1;
    if (nlit >= 15) put_run(o, op, nlit - 15, lane);
    for (uint32_t i = lane; i < nlit; i += LZ_LANES) o[op + i] = uint8_t(rd8(fr, sh + a + i));
    op += nlit;
    if (!mlen) return;
    if (lane < 2) o[op + lane] = uint8_t(lane ? off >> 8 : off & 0xffu);
    op += 2;
    if (ml >= 15) put_run(o, op, ml - 15, lane);
}

// One wavefront per frame.  Positions are taken 64 at a time: every lane
// hashes the 4 bytes at its position, reads the last earlier position with
// that hash from the table (positions of earlier rounds), and measures its
// match up to LZ_CAP bytes; the wave then takes the matches greedily in
// position order (a match shadows the positions it covers), extending any
// that reached LZ_CAP 64 bytes at a time, and writes each sequence.  Every
// position enters the table, inside matches too.
__global__ __launch_bounds__(LZ_LANES) void k_lz4_compress(const uint8_t* __restrict__ src,
                                                           const uint64_t* __restrict__ foff, uint32_t nf,
                                                           uint8_t* __restrict__ slots,
                                                           const uint64_t* __restrict__ slot_off,
                                                           uint32_t* __restrict__ csize) {
    __shared__ uint32_t fr[LZ_FRAME / 4 + 4];
    __shared__ uint32_t tab[LZ_HASH];  // last position + 1 per hash, 0 when none
    const uint32_t f = blockIdx.x, lane = threadIdx.x;
    const uint8_t* s = src + foff[f];
    const uint32_t n = uint32_t(foff[f + 1] - foff[f]);
    const uint32_t sh = uint32_t(uintptr_t(s) & 3);
    const uint32_t* w = (const uint32_t*)(s - sh);
    const uint32_t nw = (n + sh + 3) / 4;  // reads end at most 3 bytes past the frame
    for (uint32_t i0 = 0; i0 < nw + 2; i0 += 16 * LZ_LANES) {  // 16 loads in flight per lane
        uint32_t t[16];
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u) {
            const uint32_t i = i0 + u * LZ_LANES + lane;
            t[u] = i < nw ? w[i] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u) {
            const uint32_t i = i0 + u * LZ_LANES + lane;
            if (i < nw + 2) fr[i] = t[u];
        }
    }
    for (uint32_t i = lane; i < uint32_t(LZ_HASH); i += LZ_LANES) tab[i] = 0;
    __syncthreads();
    uint8_t* o = slots + slot_off[f];
    uint32_t op = 0, anchor = 0, cur = 0;
    if (n >= LZ_MF_LIMIT + 1) {
        const uint32_t match_limit = n - LZ_LAST_LITERALS;  // matches end at or before this
        const uint32_t search_end = n - LZ_MF_LIMIT;        // and start before this
        for (uint32_t p0 = 0; p0 < search_end; p0 += LZ_LANES) {
            const uint32_t p = p0 + lane;
            const bool act = p < search_end;
            uint32_t seq = 0, h = 0, c = 0;
            if (act) {
                seq = rd32(fr, sh + p);
                h = (seq * 2654435761u) >> (32 - LZ_HASH_LOG);
                c = tab[h];
            }
            __syncthreads();  // this round's reads see only earlier rounds
            if (act) atomicMax(&tab[h], p + 1);
            uint32_t L = 0;
            if (act && c != 0 && rd32(fr, sh + c - 1) == seq) {
                const uint32_t q = c - 1;
                bool exact = false;
                L = LZ_MIN_MATCH;
                while (L < LZ_CAP && p + L + 4 <= match_limit) {
                    const uint32_t x = rd32(fr, sh + q + L) ^ rd32(fr, sh + p + L);
                    if (x) {
                        L += uint32_t(__builtin_ctz(x)) >> 3;
                        exact = true;
                        break;
                    }
                    L += 4;
                }
                if (!exact)
                    while (L < LZ_CAP && p + L < match_limit && rd8(fr, sh + q + L) == rd8(fr, sh + p + L)) ++L;
            }
            const uint64_t M = __ballot(L >= LZ_MIN_MATCH);
            while (cur < p0 + LZ_LANES) {
                const uint32_t lo = cur > p0 ? cur - p0 : 0;
                const uint64_t E = M & (~0ull << lo);
                if (!E) break;
                const uint32_t j = uint32_t(__builtin_ctzll(E));
                const uint32_t mp = p0 + j;
                uint32_t mL = uint32_t(__builtin_amdgcn_readlane(int(L), int(j)));
                const uint32_t mq = uint32_t(__builtin_amdgcn_readlane(int(c), int(j))) - 1;
                if (mL >= LZ_CAP) {
                    for (;;) {
                        const uint32_t i = mL + lane;
                        const bool eq = mp + i < match_limit && rd8(fr, sh + mq + i) == rd8(fr, sh + mp + i);
                        const uint64_t ne = __ballot(!eq);
                        if (ne) {
                            mL += uint32_t(__builtin_ctzll(ne));
                            break;
                        }
                        mL += LZ_LANES;
                    }
                }
                emit_seq(o, op, fr, sh, anchor, mp - anchor, mp - mq, mL, lane);
                cur = mp + mL;
                anchor = cur;
            }
        }
    }
    emit_seq(o, op, fr, sh, anchor, n - anchor, 0, 0, lane);  // the last literals
    if (lane == 0) csize[f] = op;
}

__global__ __launch_bounds__(LZ_LANES) void k_lz4_decompress(const uint8_t* __restrict__ src,
                                                             const uint64_t* __restrict__ soff,
                                                             const uint32_t* __restrict__ csize, uint32_t nf,
                                                             uint8_t* __restrict__ dst,
                                                             const uint64_t* __restrict__ doff,
                                                             uint8_t* __restrict__ bad) {
    const uint64_t f = uint64_t(blockIdx.x) * LZ_LANES + threadIdx.x;
    if (f >= nf) return;
    const uint8_t* s = src + soff[f];
    const uint64_t n = csize[f];
    uint8_t* d = dst + doff[f];
    const uint64_t cap = doff[f + 1] - doff[f];
    uint64_t ip = 0, op = 0;
    bool ok = true;
    while (ip < n) {
        const uint32_t tok = s[ip++];
        uint64_t lit = tok >> 4;
        if (lit == 15) {
            uint32_t b = 255;
            while (b == 255 && ip < n) {
                b = s[ip++];
                lit += b;
            }
            if (b == 255) {
                ok = false;
                break;
            }
        }
        if (ip + lit > n || op + lit > cap) {
            ok = false;
            break;
        }
        for (uint64_t i = 0; i < lit; ++i) d[op + i] = s[ip + i];
        ip += lit;
        op += lit;
        if (ip >= n) break;  // the last sequence carries literals only
        if (ip + 2 > n) {
            ok = false;
            break;
        }
        const uint64_t off = uint64_t(s[ip]) | (uint64_t(s[ip + 1]) << 8);
        ip += 2;
        uint64_t ml = tok & 15u;
        if (ml == 15) {
            uint32_t b = 255;
            while (b == 255 && ip < n) {
                b = s[ip++];
                ml += b;
            }
            if (b == 255) {
                ok = false;
                break;
            }
        }
        ml += LZ_MIN_MATCH;
        if (off == 0 || off > op || op + ml > cap) {
            ok = false;
            break;
        }
        for (uint64_t i = 0; i < ml; ++i) d[op + i] = d[op - off + i];  // overlapping copies repeat the pattern
        op += ml;
    }
    bad[f] = (ok && op == cap) ? 0 : 1;
}

__global__ __launch_bounds__(256) void k_lz4_pack(const uint8_t* __restrict__ slots,
                                                  const uint64_t* __restrict__ slot_off,
                                                  const uint32_t* __restrict__ csize,
                                                  const uint64_t* __restrict__ foff, uint8_t* __restrict__ out,
                                                  const uint64_t* __restrict__ out_off) {
    const uint32_t f = blockIdx.x;
    const uint32_t c = csize[f];
    const uint32_t u = uint32_t(foff[f + 1] - foff[f]);
    uint8_t* o = out + out_off[f];
    const uint8_t* s = slots + slot_off[f];
    for (uint32_t i = threadIdx.x; i < 8 + c; i += blockDim.x)
        o[i] = i < 4 ? uint8_t(u >> (8 * i)) : i < 8 ? uint8_t(c >> (8 * (i - 4))) : s[i - 8];
}

inline uint32_t lz_blocks(uint64_t nf) { return uint32_t((nf + LZ_LANES - 1) / LZ_LANES); }

}  // namespace

void launch_lz4_compress(const uint8_t* src, const uint64_t* foff, uint32_t nf, uint8_t* slots,
                         const uint64_t* slot_off, uint32_t* csize, hipStream_t st) {
    if (!nf) return;
    { KSCOPE("k_lz4_compress", st); k_lz4_compress<<<nf, LZ_LANES, 0, st>>>(src, foff, nf, slots, slot_off, csize); }
    CDBG_HIP(hipGetLastError());
}

void launch_lz4_decompress(const uint8_t* src, const uint64_t* soff, const uint32_t* csize, uint32_t nf,
                           uint8_t* dst, const uint64_t* doff, uint8_t* bad, hipStream_t st) {
    if (!nf) return;
    { KSCOPE("k_lz4_decompress", st); k_lz4_decompress<<<lz_blocks(nf), LZ_LANES, 0, st>>>(src, soff, csize, nf, dst, doff, bad); }
    CDBG_HIP(hipGetLastError());
}

void launch_lz4_pack(const uint8_t* slots, const uint64_t* slot_off, const uint32_t* csize, const uint64_t* foff,
                     uint32_t nf, uint8_t* out, const uint64_t* out_off, hipStream_t st) {
    if (!nf) return;
    { KSCOPE("k_lz4_pack", st); k_lz4_pack<<<nf, 256, 0, st>>>(slots, slot_off, csize, foff, out, out_off); }
    CDBG_HIP(hipGetLastError());
}

}  // namespace cdbg_amd

using namespace cdbg_amd;

namespace {

// Frame table of a C-ABI call re-based to its first entry; it must ascend.
std::vector<uint64_t> rebase(const uint64_t* off, uint32_t nf) {
    std::vector<uint64_t> v(off, off + nf + 1);
    for (uint32_t f = 0; f < nf; ++f)
        if (v[f + 1] < v[f]) throw Error("frame offsets must ascend");
    const uint64_t b = v[0];
    for (auto& x : v) x -= b;
    return v;
}

}  // namespace

extern "C" {

CDBG_API int cdbg_gpu_lz4_compress_frames(cdbg_gpu_ctx* ctx, const uint8_t* src, const uint64_t* frame_off,
                                          uint32_t n_frames, uint8_t* dst, uint64_t dst_cap, uint64_t* comp_off) {
    CDBG_TRY({
        if (!ctx || !frame_off || !comp_off) throw Error("null argument");
        const std::vector<uint64_t> fo = rebase(frame_off, n_frames);
        const uint64_t n = fo[n_frames];
        if (n && (!src || !dst)) throw Error("null argument");
        std::vector<uint64_t> slot(n_frames + 1, 0);
        for (uint32_t f = 0; f < n_frames; ++f) {
            if (fo[f + 1] - fo[f] > LZ_FRAME) throw Error("lz4: a frame holds more than 64 KiB");
            slot[f + 1] = slot[f] + lz4_slot(fo[f + 1] - fo[f]);
        }
        CDBG_HIP(hipSetDevice(ctx->device));
        StreamOwner own;
        own.s = acquire_stream(ctx->device);
        own.device = ctx->device;
        hipStream_t st = own.s;
        AllocStream on_stream(st);
        DevBuf d_src, d_slots, d_tab;
        uint8_t* s = d_src.ensure<uint8_t>(n + 8);
        if (n) CDBG_HIP(hipMemcpyAsync(s, src + frame_off[0], n, hipMemcpyHostToDevice, st));
        uint8_t* out = d_slots.ensure<uint8_t>(slot[n_frames] + 16);
        uint64_t* t = d_tab.ensure<uint64_t>(2 * (n_frames + 1) + n_frames / 2 + 1);
        uint64_t* d_fo = t;
        uint64_t* d_slot = t + n_frames + 1;
        uint32_t* d_cs = (uint32_t*)(t + 2 * (n_frames + 1));
        CDBG_HIP(hipMemcpyAsync(d_fo, fo.data(), (n_frames + 1) * 8, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipMemcpyAsync(d_slot, slot.data(), (n_frames + 1) * 8, hipMemcpyHostToDevice, st));
        launch_lz4_compress(s, d_fo, n_frames, out, d_slot, d_cs, st);
        std::vector<uint32_t> cs(n_frames);
        std::vector<uint8_t> h(slot[n_frames]);
        if (n_frames) CDBG_HIP(hipMemcpyAsync(cs.data(), d_cs, n_frames * 4, hipMemcpyDeviceToHost, st));
        if (!h.empty()) CDBG_HIP(hipMemcpyAsync(h.data(), out, h.size(), hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
        comp_off[0] = 0;
        for (uint32_t f = 0; f < n_frames; ++f) {
            if (cs[f] > slot[f + 1] - slot[f]) throw Error("lz4: a compressed frame overran its slot");
            comp_off[f + 1] = comp_off[f] + cs[f];
        }
        if (comp_off[n_frames] > dst_cap) throw Error("lz4: output capacity too small");
        for (uint32_t f = 0; f < n_frames; ++f) std::memcpy(dst + comp_off[f], h.data() + slot[f], cs[f]);
    })
}

CDBG_API int cdbg_gpu_lz4_decompress_frames(cdbg_gpu_ctx* ctx, const uint8_t* comp, const uint64_t* comp_off,
                                            uint32_t n_frames, uint8_t* dst, const uint64_t* dst_off) {
    CDBG_TRY({
        if (!ctx || !comp_off || !dst_off) throw Error("null argument");
        const std::vector<uint64_t> co = rebase(comp_off, n_frames), dof = rebase(dst_off, n_frames);
        const uint64_t nc = co[n_frames], nd = dof[n_frames];
        if ((nc && !comp) || (nd && !dst)) throw Error("null argument");
        std::vector<uint32_t> cs(n_frames);
        for (uint32_t f = 0; f < n_frames; ++f) {
            if (co[f + 1] - co[f] >= (1ull << 32)) throw Error("lz4: a block holds 4 GiB or more");
            cs[f] = uint32_t(co[f + 1] - co[f]);
        }
        CDBG_HIP(hipSetDevice(ctx->device));
        StreamOwner own;
        own.s = acquire_stream(ctx->device);
        own.device = ctx->device;
        hipStream_t st = own.s;
        AllocStream on_stream(st);
        DevBuf d_src, d_dst, d_tab;
        uint8_t* s = d_src.ensure<uint8_t>(nc + 8);
        uint8_t* d = d_dst.ensure<uint8_t>(nd + 8);
        uint64_t* t = d_tab.ensure<uint64_t>(2 * (n_frames + 1) + n_frames);
        uint64_t* d_so = t;
        uint64_t* d_do = t + n_frames + 1;
        uint32_t* d_cs = (uint32_t*)(t + 2 * (n_frames + 1));
        uint8_t* d_bad = (uint8_t*)(d_cs + n_frames);
        if (nc) CDBG_HIP(hipMemcpyAsync(s, comp + comp_off[0], nc, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipMemcpyAsync(d_so, co.data(), (n_frames + 1) * 8, hipMemcpyHostToDevice, st));
        CDBG_HIP(hipMemcpyAsync(d_do, dof.data(), (n_frames + 1) * 8, hipMemcpyHostToDevice, st));
        if (n_frames) CDBG_HIP(hipMemcpyAsync(d_cs, cs.data(), n_frames * 4, hipMemcpyHostToDevice, st));
        launch_lz4_decompress(s, d_so, d_cs, n_frames, d, d_do, d_bad, st);
        std::vector<uint8_t> bad(n_frames);
        if (n_frames) CDBG_HIP(hipMemcpyAsync(bad.data(), d_bad, n_frames, hipMemcpyDeviceToHost, st));
        if (nd) CDBG_HIP(hipMemcpyAsync(dst + dst_off[0], d, nd, hipMemcpyDeviceToHost, st));
        CDBG_HIP(hipStreamSynchronize(st));
        for (uint32_t f = 0; f < n_frames; ++f)
            if (bad[f]) throw Error("lz4: frame " + std::to_string(f) + " is malformed or has the wrong size");
    })
}

}  // extern "C"
