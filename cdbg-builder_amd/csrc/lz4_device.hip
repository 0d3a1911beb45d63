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
// Parallel layout: one lane per frame, thousands of frames per launch.
//   k_lz4_compress   frames of at most LZ_FRAME (64 KiB), so every position
//                    and match offset fits 16 bits; a 512-entry table of
//                    positions per lane in LDS (64 lanes x 1 KiB per
//                    workgroup); greedy parse; the input is read as aligned
//                    dwords (a two-dword window slides along the scan) and
//                    the output is gathered into dwords before it is stored.
//   k_lz4_decompress any frame size; every length, offset and capacity is
//                    checked, and a malformed frame is flagged, never read or
//                    written outside its ranges.
//   k_lz4_pack       one workgroup per frame: header + block into the
//                    bucket-file image.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "job.h"
#include "lz4_device.h"

namespace cdbg_amd {

namespace {

constexpr int LZ_HASH_LOG = 9;
constexpr int LZ_HASH = 1 << LZ_HASH_LOG;
constexpr int LZ_LANES = 64;
constexpr uint32_t LZ_MIN_MATCH = 4, LZ_LAST_LITERALS = 5, LZ_MF_LIMIT = 12;
constexpr uint16_t LZ_EMPTY = 0xFFFF;

// 4 bytes at any address, from the two aligned dwords that cover them.
__device__ __forceinline__ uint32_t ld32(const uint8_t* p) {
    const uintptr_t a = uintptr_t(p);
    const uint32_t* w = (const uint32_t*)(a & ~uintptr_t(3));
    const uint32_t sh = uint32_t(a & 3) * 8;
    const uint32_t lo = w[0];
    return sh ? (lo >> sh) | (w[1] << (32 - sh)) : lo;
}

// Output gathered into dwords; the slot starts 4-byte aligned and has room
// for the last partial dword.
struct Out {
    uint32_t* o;
    uint32_t op, acc;
    __device__ __forceinline__ void put(uint32_t b) {
        acc |= b << (8 * (op & 3));
        if ((op & 3) == 3) {
            o[op >> 2] = acc;
            acc = 0;
        }
        ++op;
    }
    __device__ __forceinline__ void len(uint32_t r) {  // length bytes past a token's 15
        while (r >= 255) {
            put(255);
            r -= 255;
        }
        put(r);
    }
    __device__ __forceinline__ void flush() {
        if (op & 3) o[op >> 2] = acc;
    }
};

// One sequence: the literals s[a, a + nlit), then (mlen > 0) a match of mlen
// bytes at distance off.
__device__ __forceinline__ void emit_seq(Out& w, const uint8_t* s, uint32_t a, uint32_t nlit, uint32_t off,
                                         uint32_t mlen) {
    const uint32_t ml = mlen ? mlen - LZ_MIN_MATCH : 0;
    w.put(((nlit >= 15 ? 15u : nlit) << 4) | (mlen ? (ml >= 15 ? 15u : ml) : 0u));
    if (nlit >= 15) w.len(nlit - 15);
    uint32_t i = 0;
    while (i < nlit) {  // literals a dword at a time
        const uint32_t v = ld32(s + a + i);
        const uint32_t c = nlit - i < 4 ? nlit - i : 4;
        for (uint32_t j = 0; j < c; ++j) w.put((v >> (8 * j)) & 0xffu);
        i += c;
    }
    if (!mlen) return;
    w.put(off & 0xffu);
    w.put(off >> 8);
    if (ml >= 15) w.len(ml - 15);
}

__global__ __launch_bounds__(LZ_LANES) void k_lz4_compress(const uint8_t* __restrict__ src,
                                                           const uint64_t* __restrict__ foff, uint32_t nf,
                                                           uint8_t* __restrict__ slots,
                                                           const uint64_t* __restrict__ slot_off,
                                                           uint32_t* __restrict__ csize) {
    __shared__ uint16_t tab[LZ_HASH][LZ_LANES];  // lane-interleaved: a lane's probes spread over the banks
    const uint32_t lane = threadIdx.x;
    for (int i = 0; i < LZ_HASH; ++i) tab[i][lane] = LZ_EMPTY;
    const uint64_t f = uint64_t(blockIdx.x) * LZ_LANES + lane;
    if (f >= nf) return;
    const uint8_t* s = src + foff[f];
    const uint32_t n = uint32_t(foff[f + 1] - foff[f]);
    Out w{(uint32_t*)(slots + slot_off[f]), 0, 0};
    uint32_t anchor = 0, ip = 0;
    if (n >= LZ_MF_LIMIT + 1) {
        const uint32_t match_limit = n - LZ_LAST_LITERALS;  // matches end at or before this
        const uint32_t search_end = n - LZ_MF_LIMIT;        // and start before this
        const uintptr_t a0 = uintptr_t(s);
        // window: the aligned dwords at wp[0], wp[1], wp = (a0 + ip) & ~3
        const uint32_t* wp = (const uint32_t*)(a0 & ~uintptr_t(3));
        uint32_t lo = wp[0], hi = wp[1];
        while (ip < search_end) {
            const uint32_t sh = uint32_t((a0 + ip) & 3) * 8;
            const uint32_t seq = sh ? (lo >> sh) | (hi << (32 - sh)) : lo;
            const uint32_t h = (seq * 2654435761u) >> (32 - LZ_HASH_LOG);
            const uint32_t c = tab[h][lane];
            tab[h][lane] = uint16_t(ip);
            if (c != LZ_EMPTY && ld32(s + c) == seq) {
                uint32_t m = LZ_MIN_MATCH;
                while (ip + m + 4 <= match_limit) {  // extend a dword at a time
                    const uint32_t x = ld32(s + c + m) ^ ld32(s + ip + m);
                    if (x) break;
                    m += 4;
                }
                while (ip + m < match_limit && s[c + m] == s[ip + m]) ++m;  // the last few bytes
                emit_seq(w, s, anchor, ip - anchor, ip - c, m);
                ip += m;
                anchor = ip;
                if (ip < search_end) {
                    wp = (const uint32_t*)((a0 + ip) & ~uintptr_t(3));
                    lo = wp[0];
                    hi = wp[1];
                }
            } else {
                ++ip;
                if (((a0 + ip) & 3) == 0 && ip < search_end) {
                    ++wp;
                    lo = hi;
                    hi = wp[1];
                }
            }
        }
    }
    emit_seq(w, s, anchor, n - anchor, 0, 0);  // the last literals
    w.flush();
    csize[f] = w.op;
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
    { KSCOPE("k_lz4_compress", st); k_lz4_compress<<<lz_blocks(nf), LZ_LANES, 0, st>>>(src, foff, nf, slots, slot_off, csize); }
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
