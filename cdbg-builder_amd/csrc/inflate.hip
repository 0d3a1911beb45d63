// Input decode on the device: gzip inflate and FASTA record parsing, from
// the files' bytes to a job's input (every colour's records, each followed by
// '\n').
//
//   k_gunzip       one wavefront per file, running the decoder of
//                  inflate_core.h: the compressed bytes pass through an LDS
//                  window, the output through a 64 KiB LDS ring (the 32 KiB
//                  history every match may reach, plus what is not yet
//                  written out), match copies and table fills spread over
//                  the lanes; files without the gzip magic are taken as
//                  they are (copied, not decoded)
//   k_crc_members  one workgroup per file: CRC-32 of every member by 256
//                  threads over contiguous parts, combined in GF(2)
//   k_nl_*         newline positions (every decoded file is closed by one)
//   k_first_header the first '>' or '@' of each file (where records begin)
//   k_line_size /  one item per line plus one per file: a sequence line
//   k_line_write   keeps its bytes but the '\n', a header line after the
//                  first becomes the '\n' closing the previous record, the
//                  file's last item closes its last record
// The record rules are those of kseq_read as seq_reader uses it
// (seq_reader.hpp:44-77, external/kseq.h:161-201; the host reader in
// fasta.cpp states them): records start at '>' or '@', the header line is
// skipped, sequence lines run to the next line starting with '>', '@' or '+',
// '\n' is dropped and every other byte kept.  A line starting with '+' (a
// FASTQ quality header) is not handled here: such inputs go to the host
// reader.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <thread>
#include <vector>

#include <zlib.h>

#include "inflate_core.h"
#include "job.h"

namespace cdbg_amd {

namespace {

constexpr uint32_t IN_WIN = 4096;    // LDS input window, bytes (with the ring and tables: two files per CU)
constexpr uint32_t RING = 65536;     // LDS output ring, bytes
constexpr uint32_t RMASK = RING - 1;
constexpr uint32_t FLUSH = 16384;    // written out once this much is pending
constexpr uint64_t MIN_MEMBER = 18;  // header, an empty final block and the trailer take at least this
constexpr int TPB = 256;

struct Member {
    uint64_t start, len;
    uint32_t crc, pad;
};

struct FileOut {
    uint64_t produced;  // output bytes (may exceed the capacity: then nothing past it was written)
    int32_t status;     // infl::OK or an error code
    uint32_t members;
};

inline uint32_t blocks_for(uint64_t n, int tpb = TPB) { return uint32_t(std::max<uint64_t>(1, (n + tpb - 1) / tpb)); }

struct DevIO {
    const uint8_t* src;
    uint64_t n;
    uint8_t* dst;
    uint64_t cap;
    uint32_t* win32;
    uint8_t* ring;
    infl::Tables* T;
    Member* mem;  // room for in_len / MIN_MEMBER + 1 members
    uint64_t wbase, wend, op, fpos;
    uint32_t ln, nmem;

    __device__ __forceinline__ uint64_t in_len() { return n; }
    __device__ __forceinline__ void fill(uint64_t pos) {  // window from pos (4-aligned down)
        __syncthreads();  // every lane is done with the old window
        wbase = pos & ~uint64_t(3);
        wend = min(wbase + IN_WIN, n);
        const uint32_t nd = uint32_t((wend - wbase + 3) / 4);  // ends <= 3 bytes past the file: its padding
        const uint32_t* s32 = (const uint32_t*)(src + wbase);
        for (uint32_t i0 = 0; i0 < nd + 4; i0 += 16 * 64) {
            uint32_t t[16];
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const uint32_t i = i0 + u * 64 + ln;
                t[u] = i < nd ? s32[i] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const uint32_t i = i0 + u * 64 + ln;
                if (i < nd + 4) win32[i] = t[u];
            }
        }
        __syncthreads();
    }
    __device__ __forceinline__ uint64_t load8(uint64_t pos) {
        if (pos >= n) return 0;
        if (pos < wbase || (pos + 8 > wend && wend < n)) fill(pos);
        const uint32_t a = uint32_t(pos - wbase);
        const uint32_t w0 = win32[a >> 2], w1 = win32[(a >> 2) + 1], w2 = win32[(a >> 2) + 2];
        const uint32_t sh = a & 3;
        const uint64_t v = uint64_t(__builtin_amdgcn_alignbyte(w1, w0, sh)) |
                           (uint64_t(__builtin_amdgcn_alignbyte(w2, w1, sh)) << 32);
        const uint64_t left = n - pos;
        return left >= 8 ? v : v & ((uint64_t(1) << (8 * left)) - 1);
    }
    __device__ __forceinline__ void flush() {
        __syncthreads();
        for (uint64_t p = fpos + ln; p < op; p += 64)
            if (p < cap) dst[p] = ring[p & RMASK];
        __syncthreads();
        fpos = op;
    }
    __device__ __forceinline__ void put(uint32_t b) {
        if (ln == 0) ring[op & RMASK] = uint8_t(b);
        ++op;
        if (op - fpos >= FLUSH) flush();
    }
    // The decoder's ring traffic is one wavefront's: its LDS operations are
    // performed in the order issued, so a copy needs no barrier to see lane
    // 0's literals, only the compiler kept from reordering across it.
    __device__ __forceinline__ void wave_order() {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __device__ __forceinline__ bool copy(uint32_t dist, uint32_t len) {
        if (dist == 0 || dist > op) return false;
        wave_order();
        if (dist >= len) {
            for (uint32_t i = ln; i < len; i += 64) ring[(op + i) & RMASK] = ring[(op - dist + i) & RMASK];
        } else {  // the match repeats its last `dist` bytes: every source still precedes op
            for (uint32_t i = ln; i < len; i += 64) ring[(op + i) & RMASK] = ring[(op - dist + (i % dist)) & RMASK];
        }
        wave_order();
        op += len;
        if (op - fpos >= FLUSH) flush();
        return true;
    }
    __device__ __forceinline__ void copy_in(uint64_t pos, uint32_t len) {
        while (len) {
            const uint32_t c = min(len, FLUSH);
            __syncthreads();
            for (uint32_t i = ln; i < c; i += 64) ring[(op + i) & RMASK] = src[pos + i];
            __syncthreads();
            op += c;
            pos += c;
            len -= c;
            if (op - fpos >= FLUSH) flush();
        }
    }
    __device__ __forceinline__ uint64_t produced() { return op; }
    __device__ __forceinline__ uint32_t lane() { return ln; }
    __device__ __forceinline__ uint32_t nlanes() { return 64; }
    __device__ __forceinline__ void sync() { __syncthreads(); }
    __device__ __forceinline__ infl::Tables* tables() { return T; }
    __device__ __forceinline__ bool member(uint64_t s, uint64_t l, uint32_t crc) {
        if (nmem > n / MIN_MEMBER) return false;
        if (ln == 0) mem[nmem] = Member{s, l, crc, 0};
        ++nmem;
        return true;
    }
};

// File j: bytes img[in_off[j], + in_len[j]), output room out[out_off[j], +
// out_cap[j]), members from mems + mem_off[j].
__global__ __launch_bounds__(64) void k_gunzip(const uint8_t* __restrict__ img, const uint64_t* __restrict__ in_off,
                                               const uint64_t* __restrict__ in_len, uint8_t* __restrict__ out,
                                               const uint64_t* __restrict__ out_off,
                                               const uint64_t* __restrict__ out_cap,
                                               const uint64_t* __restrict__ mem_off, Member* __restrict__ mems,
                                               FileOut* __restrict__ res) {
    __shared__ uint32_t win32[IN_WIN / 4 + 8];
    __shared__ uint8_t ring[RING];
    __shared__ infl::Tables T;
    const uint32_t f = blockIdx.x;
    DevIO io{img + in_off[f], in_len[f], out + out_off[f], out_cap[f], win32, ring, &T, mems + mem_off[f],
             0, 0, 0, 0, threadIdx.x, 0};
    infl::Inflater<DevIO> inf(io);
    const int st = inf.gunzip();
    io.flush();
    if (threadIdx.x == 0) {
        res[f].produced = io.op;
        res[f].status = st;
        res[f].members = io.nmem;
    }
}

// CRC-32 (bitwise definition, 0xEDB88320 reflected) of every member, by a
// table in LDS, 256 contiguous parts per member combined in order.
__global__ __launch_bounds__(TPB) void k_crc_members(const uint8_t* __restrict__ out,
                                                     const uint64_t* __restrict__ out_off,
                                                     const uint64_t* __restrict__ out_cap,
                                                     const uint64_t* __restrict__ mem_off,
                                                     const Member* __restrict__ mems, const FileOut* __restrict__ res,
                                                     uint32_t* __restrict__ bad) {
    __shared__ uint32_t tab[256];
    __shared__ uint32_t part[TPB];
    const uint32_t f = blockIdx.x;
    {
        uint32_t c = threadIdx.x;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        tab[threadIdx.x] = c;
    }
    __syncthreads();
    const FileOut& r = res[f];
    if (r.status != infl::OK || r.produced > out_cap[f]) return;
    const uint8_t* base = out + out_off[f];
    const Member* M = mems + mem_off[f];
    for (uint32_t m = 0; m < r.members; ++m) {
        const uint64_t s = M[m].start, len = M[m].len;
        const uint64_t chunk = (len + TPB - 1) / TPB;
        const uint64_t a = min(len, chunk * threadIdx.x), b = min(len, a + chunk);
        uint32_t c = 0xFFFFFFFFu;
        for (uint64_t i = a; i < b; ++i) c = tab[(c ^ base[s + i]) & 0xFFu] ^ (c >> 8);
        part[threadIdx.x] = ~c;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t pw = chunk ? infl::x_pow_8n(chunk) : 0x80000000u;
            uint32_t crc = 0;  // of the empty prefix
            for (uint32_t t = 0; t < TPB; ++t) {
                const uint64_t ta = min(len, chunk * t), tb = min(len, ta + chunk);
                if (tb == ta) continue;
                crc = (tb - ta == chunk) ? infl::gf2_mulmod(pw, crc) ^ part[t] : infl::crc32_combine(crc, part[t], tb - ta);
            }
            if (crc != M[m].crc) bad[f] = 1;
        }
        __syncthreads();
    }
}

// Every decoded file ends with '\n' (added when its last byte is another).
__global__ void k_close_files(uint32_t nf, uint8_t* __restrict__ dec, const uint64_t* __restrict__ dec_off,
                              uint64_t* __restrict__ len) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const uint64_t n = len[f];
    if (n && dec[dec_off[f] + n - 1] != '\n') {
        dec[dec_off[f] + n] = '\n';
        len[f] = n + 1;
    }
}

// Largest f with off[f] <= i.
__device__ __forceinline__ uint32_t file_of(const uint64_t* __restrict__ off, uint32_t nf, uint64_t i) {
    uint32_t lo = 0, hi = nf;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Newlines: 16-byte groups, four per thread; a byte counts when it is '\n'
// and inside its file's decoded length.
constexpr uint32_t NL_CHUNK = 64;
__device__ __forceinline__ uint32_t nl_mask(const uint8_t* __restrict__ dec, uint64_t pos, uint64_t end) {
    const uint4 w = *reinterpret_cast<const uint4*>(dec + pos);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (((ws[q] >> (8 * j)) & 0xFFu) == '\n') m |= 1u << (4 * q + j);
    const uint64_t room = end > pos ? end - pos : 0;
    return room >= 16 ? m : (m & ((1u << room) - 1));
}
__global__ void k_nl_count(uint64_t n_chunks, const uint8_t* __restrict__ dec, const uint64_t* __restrict__ dec_off,
                           const uint64_t* __restrict__ len, uint32_t nf, uint32_t* __restrict__ cnt) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_chunks) return;
    const uint64_t pos = c * NL_CHUNK;
    const uint32_t f = file_of(dec_off, nf, pos);  // files start 16-byte aligned: a group is in one file
    uint32_t n = 0;
#pragma unroll
    for (uint32_t g = 0; g < NL_CHUNK; g += 16) {
        const uint32_t fg = file_of(dec_off, nf, pos + g);
        n += __popc(nl_mask(dec, pos + g, dec_off[fg] + len[fg]));
    }
    (void)f;
    cnt[c] = n;
}
__global__ void k_nl_write(uint64_t n_chunks, const uint8_t* __restrict__ dec, const uint64_t* __restrict__ dec_off,
                           const uint64_t* __restrict__ len, uint32_t nf, const uint64_t* __restrict__ at,
                           uint64_t* __restrict__ nl) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_chunks) return;
    const uint64_t pos = c * NL_CHUNK;
    uint64_t o = at[c];
#pragma unroll
    for (uint32_t g = 0; g < NL_CHUNK; g += 16) {
        const uint32_t fg = file_of(dec_off, nf, pos + g);
        uint32_t m = nl_mask(dec, pos + g, dec_off[fg] + len[fg]);
        while (m) {
            nl[o++] = pos + g + uint32_t(__ffs(m) - 1);
            m &= m - 1;
        }
    }
}

// First '>' or '@' of each file (one workgroup per file, 4 KiB a round).
__global__ __launch_bounds__(TPB) void k_first_header(const uint8_t* __restrict__ dec,
                                                      const uint64_t* __restrict__ dec_off,
                                                      const uint64_t* __restrict__ len, uint64_t* __restrict__ p0) {
    __shared__ unsigned long long best;
    const uint32_t f = blockIdx.x;
    const uint64_t s = dec_off[f], e = s + len[f];
    if (threadIdx.x == 0) best = ~0ull;
    __syncthreads();
    for (uint64_t b = s; b < e; b += TPB * 16) {
        for (uint32_t j = 0; j < 16; ++j) {
            const uint64_t i = b + uint64_t(j) * TPB + threadIdx.x;
            if (i < e && (dec[i] == '>' || dec[i] == '@')) atomicMin(&best, (unsigned long long)i);
        }
        __syncthreads();
        if (best != ~0ull) break;
        __syncthreads();
    }
    if (threadIdx.x == 0) p0[f] = best;  // ~0: no record
}

// Lines of file f: its newlines nl[lb[f] .. lb[f + 1]); item of line k is
// k + f, the file's closing item lb[f + 1] + f.
__global__ void k_file_lines(uint32_t nf, const uint64_t* __restrict__ nl, uint64_t n_nl,
                             const uint64_t* __restrict__ dec_off, uint64_t* __restrict__ lb) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nf) return;
    const uint64_t x = dec_off[f];  // first newline at or after the file's start
    uint64_t lo = 0, hi = n_nl;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (nl[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    lb[f] = lo;
}

struct LineArgs {
    const uint8_t* dec;
    const uint64_t* dec_off;
    const uint64_t* nl;
    const uint64_t* lb;
    const uint64_t* p0;
    uint32_t nf;
    uint64_t n_items;
};

// Item i: (file, line or closing item) and its output bytes; *plus = 1 when a
// line after the first header starts with '+'.
__device__ __forceinline__ uint64_t item_size(const LineArgs& a, uint64_t i, uint32_t& f, uint64_t& s, uint64_t& e,
                                              int& kind) {
    // the item's file: largest f with lb[f] + f <= i
    uint32_t lo = 0, hi = a.nf;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.lb[mid] + mid <= i) lo = mid;
        else hi = mid;
    }
    f = lo;
    const uint64_t k = i - f;  // newline index, or lb[f + 1] for the closing item
    const uint64_t p0 = a.p0[f];
    if (k == a.lb[f + 1]) {  // closing item
        kind = 3;
        return p0 != ~0ull ? 1 : 0;
    }
    e = a.nl[k];
    s = k == a.lb[f] ? a.dec_off[f] : a.nl[k - 1] + 1;
    if (p0 == ~0ull || e < p0) {  // before the first record
        kind = 0;
        return 0;
    }
    if (s <= p0) {  // the first header line
        kind = 0;
        return 0;
    }
    if (s == e) {  // empty line
        kind = 0;
        return 0;
    }
    const uint8_t c = a.dec[s];
    if (c == '>' || c == '@') {
        kind = 2;
        return 1;
    }
    if (c == '+') {
        kind = 4;
        return 0;
    }
    kind = 1;
    return e - s;
}

__global__ void k_line_size(LineArgs a, uint64_t* __restrict__ size, uint32_t* __restrict__ plus) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n_items) return;
    uint32_t f;
    uint64_t s = 0, e = 0;
    int kind;
    size[i] = item_size(a, i, f, s, e, kind);
    if (kind == 4) atomicOr(plus, 1u);
}

__global__ void k_line_write(LineArgs a, const uint64_t* __restrict__ off, uint8_t* __restrict__ outb) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n_items) return;
    uint32_t f;
    uint64_t s = 0, e = 0;
    int kind;
    const uint64_t n = item_size(a, i, f, s, e, kind);
    if (!n) return;
    uint8_t* o = outb + off[i];
    if (kind == 1) {
        for (uint64_t j = 0; j < n; ++j) o[j] = a.dec[s + j];
    } else {
        o[0] = '\n';
    }
}

// Colour offsets: the first item of each file (and the total).
__global__ void k_color_offsets(uint32_t nf, const uint64_t* __restrict__ lb, const uint64_t* __restrict__ off,
                                uint64_t* __restrict__ coff) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f <= nf) coff[f] = off[lb[f] + f];
}

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

// A whole gzip file through zlib (gzread: members in turn, trailing bytes
// ignored) into out[0, cap); returns the decoded size (> cap: more data).
uint64_t gunzip_host(const std::string& path, uint8_t* out, uint64_t cap) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) throw Error("could not open input file: " + path);
    gzbuffer(f, 256 * 1024);
    uint64_t n = 0;
    for (;;) {
        uint8_t spill[65536];
        uint8_t* dst = n < cap ? out + n : spill;
        const uint64_t room = n < cap ? std::min<uint64_t>(cap - n, 1u << 30) : sizeof spill;
        const int r = gzread(f, dst, unsigned(room));
        if (r < 0) {
            int err = 0;
            const std::string msg = gzerror(f, &err);
            gzclose(f);
            throw Error("gzip: " + msg + " in " + path);
        }
        if (r == 0) break;
        n += uint64_t(r);
    }
    gzclose(f);
    return n;
}

// Where gzip files inflate fastest, by the rates measured on MI355X: one
// file per wavefront decodes ~6 MB/s of output with two files per CU at a
// time, a host thread of zlib ~350 MB/s.  The device wins only with few
// host threads (the FASTA parse runs on the device either way).
bool gzip_on_device_faster(const std::vector<uint64_t>& out_sizes, uint32_t threads, int cus) {
    if (out_sizes.empty()) return false;
    const double dev_rate = 6e6, host_rate = 350e6;
    uint64_t total = 0, largest = 0;
    for (uint64_t x : out_sizes) {
        total += x;
        largest = std::max(largest, x);
    }
    const double slots = 2.0 * std::max(1, cus);
    const double rounds = std::ceil(double(out_sizes.size()) / slots);
    const double t_dev = rounds * double(largest) / dev_rate;
    const double t_host = double(total) / (host_rate * std::max(1u, threads));
    return t_dev < t_host;
}

const char* status_text(int s) {
    switch (s) {
        case infl::E_TRUNCATED: return "truncated";
        case infl::E_HEADER: return "bad gzip header";
        case infl::E_BTYPE: return "invalid block type";
        case infl::E_STORED: return "invalid stored block lengths";
        case infl::E_CODES: return "invalid code lengths";
        case infl::E_SYMBOL: return "invalid code";
        case infl::E_DISTANCE: return "invalid distance";
        case infl::E_REPEAT: return "invalid length repeat";
        case infl::E_SIZE: return "size field differs from the data";
        case infl::E_MEMBERS: return "gzip members smaller than possible";
        default: return "decode error";
    }
}

}  // namespace

// Reads `files` (colour = position) into J's input on the device.  Returns
// false, leaving J's input unset, when an input needs the host reader (a
// line starting with '+').
bool load_inputs_device(Job& J, const std::vector<std::string>& files, uint32_t threads) {
    AllocStream alloc_on_job_stream(J.st);
    hipStream_t st = J.st;
    const uint32_t nf = uint32_t(files.size());
    if (nf == 0) throw Error("no input files");
    // 1. the files' bytes, each at a 16-byte aligned offset with 16 bytes of slack
    std::vector<uint64_t> fsize(nf), ioff(nf + 1, 0);
    for (uint32_t f = 0; f < nf; ++f) {
        std::error_code ec;
        fsize[f] = std::filesystem::file_size(files[f], ec);
        if (ec) throw Error("could not open input file: " + files[f]);
        ioff[f + 1] = ioff[f] + ((fsize[f] + 31) & ~uint64_t(15));
    }
    PinnedBuf img;
    img.ensure_keep(ioff[nf] + 16, 0);
    std::vector<uint8_t> gz(nf);
    std::vector<uint64_t> cap(nf);
    parallel_for(nf, threads, [&](uint32_t f) {
        FILE* fp = std::fopen(files[f].c_str(), "rb");
        if (!fp) throw Error("could not open input file: " + files[f]);
        const bool ok = std::fread(img.p + ioff[f], 1, fsize[f], fp) == fsize[f];
        std::fclose(fp);
        if (!ok) throw Error("short read of " + files[f]);
        std::memset(img.p + ioff[f] + fsize[f], 0, ioff[f + 1] - ioff[f] - fsize[f]);
        const uint8_t* p = img.p + ioff[f];
        gz[f] = fsize[f] >= 18 && p[0] == 0x1f && p[1] == 0x8b;
        // output room: the last member's size field (a single member's
        // exact size), or the file itself; +1 for a closing '\n'
        cap[f] = gz[f] ? (uint64_t(p[fsize[f] - 4]) | uint64_t(p[fsize[f] - 3]) << 8 |
                          uint64_t(p[fsize[f] - 2]) << 16 | uint64_t(p[fsize[f] - 1]) << 24)
                       : fsize[f];
    });
    std::vector<uint32_t> glist;
    for (uint32_t f = 0; f < nf; ++f)
        if (gz[f]) glist.push_back(f);
    bool on_device = J.tune.gzip_device == 1;
    if (J.tune.gzip_device == 0 && !glist.empty()) {
        std::vector<uint64_t> sizes;
        for (uint32_t f : glist) sizes.push_back(cap[f]);
        hipDeviceProp_t prop{};
        CDBG_HIP(hipGetDeviceProperties(&prop, J.ctx->device));
        on_device = gzip_on_device_faster(sizes, threads, prop.multiProcessorCount);
    }
    // on host threads instead: zlib inflates into pinned memory, and the
    // device takes the decoded bytes as it takes plain files
    PinnedBuf dec_h;
    std::vector<const uint8_t*> host_dec(nf, nullptr);
    if (!on_device && !glist.empty()) {
        std::vector<uint64_t> hoff(glist.size() + 1, 0), got(glist.size());
        for (int pass = 0; pass < 2; ++pass) {
            for (size_t j = 0; j < glist.size(); ++j) hoff[j + 1] = hoff[j] + ((cap[glist[j]] + 15) & ~uint64_t(15));
            dec_h.release();
            dec_h.ensure_keep(hoff.back() + 16, 0);
            parallel_for(uint32_t(glist.size()), threads, [&](uint32_t j) {
                got[j] = gunzip_host(files[glist[j]], dec_h.p + hoff[j], cap[glist[j]]);
            });
            bool short_room = false;
            for (size_t j = 0; j < glist.size(); ++j) {
                short_room |= got[j] > cap[glist[j]];
                cap[glist[j]] = got[j];  // exact now; again only if it did not fit
            }
            if (!short_room) break;  // members beyond the last one's size field: again with the sizes found
        }
        for (size_t j = 0; j < glist.size(); ++j) host_dec[glist[j]] = dec_h.p + hoff[j];
        glist.clear();
    }
    const uint32_t ng = uint32_t(glist.size());
    DevBuf d_img, d_tab, d_dec, d_res, d_bad, d_mem;
    uint8_t* dimg = d_img.ensure<uint8_t>(ioff[nf] + 16);
    CDBG_HIP(hipMemcpyAsync(dimg, img.p, ioff[nf], hipMemcpyHostToDevice, st));
    std::vector<uint64_t> doff(nf + 1, 0), len(nf, 0);
    uint8_t* dec = nullptr;
    uint64_t D = 0;
    // 2. inflate the gzip files, plain ones copied.  A file of several
    //    members may decode to more than its last member's size field: the
    //    decoder then counts without writing, and the pass is repeated once
    //    with the sizes it counted.
    for (int pass = 0;; ++pass) {
        for (uint32_t f = 0; f < nf; ++f) doff[f + 1] = doff[f] + ((cap[f] + 1 + 15) & ~uint64_t(15));
        D = doff[nf];
        dec = d_dec.ensure<uint8_t>(D + 128);  // newline groups of 64 bytes read up to 48 bytes past D
        for (uint32_t f = 0; f < nf; ++f)
            if (host_dec[f]) {
                if (cap[f]) CDBG_HIP(hipMemcpyAsync(dec + doff[f], host_dec[f], cap[f], hipMemcpyHostToDevice, st));
                len[f] = cap[f];
            } else if (!gz[f]) {
                if (fsize[f]) CDBG_HIP(hipMemcpyAsync(dec + doff[f], dimg + ioff[f], fsize[f], hipMemcpyDeviceToDevice, st));
                len[f] = fsize[f];
            }
        bool again = false;
        if (ng) {
            std::vector<uint64_t> tab(5 * uint64_t(ng));  // in_off, in_len, out_off, out_cap, mem_off per gzip file
            uint64_t nm = 0;
            for (uint32_t j = 0; j < ng; ++j) {
                const uint32_t f = glist[j];
                tab[j] = ioff[f];
                tab[ng + j] = fsize[f];
                tab[2 * ng + j] = doff[f];
                tab[3 * ng + j] = cap[f];
                tab[4 * ng + j] = nm;
                nm += fsize[f] / MIN_MEMBER + 1;
            }
            uint64_t* t = d_tab.ensure<uint64_t>(tab.size());
            Member* mems = d_mem.ensure<Member>(nm);
            FileOut* res = d_res.ensure<FileOut>(ng);
            uint32_t* bad = d_bad.ensure<uint32_t>(ng);
            CDBG_HIP(hipMemsetAsync(bad, 0, ng * 4ull, st));
            CDBG_HIP(hipMemcpyAsync(t, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, st));
            { KSCOPE("k_gunzip", st); k_gunzip<<<ng, 64, 0, st>>>(dimg, t, t + ng, dec, t + 2 * ng, t + 3 * ng, t + 4 * ng, mems, res); }
            CDBG_HIP(hipGetLastError());
            { KSCOPE("k_crc_members", st); k_crc_members<<<ng, TPB, 0, st>>>(dec, t + 2 * ng, t + 3 * ng, t + 4 * ng, mems, res, bad); }
            CDBG_HIP(hipGetLastError());
            std::vector<FileOut> h_res(ng);
            std::vector<uint32_t> h_bad(ng);
            CDBG_HIP(hipMemcpyAsync(h_res.data(), res, ng * sizeof(FileOut), hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipMemcpyAsync(h_bad.data(), bad, ng * 4ull, hipMemcpyDeviceToHost, st));
            CDBG_HIP(hipStreamSynchronize(st));
            for (uint32_t j = 0; j < ng; ++j) {
                const uint32_t f = glist[j];
                if (h_res[j].status != infl::OK)
                    throw Error(std::string("gzip: ") + status_text(h_res[j].status) + " in " + files[f]);
                if (h_res[j].produced > cap[f]) {
                    if (pass) throw Error("gzip: " + files[f] + " decoded to another size the second time");
                    cap[f] = h_res[j].produced;
                    again = true;
                    continue;
                }
                if (h_bad[j]) throw Error("gzip: CRC-32 mismatch in " + files[f]);
                len[f] = h_res[j].produced;
            }
        }
        if (!again) break;
    }
    d_img.release();
    // 3. every file closed by '\n'; newline positions
    DevBuf d_len, d_cnt, d_at, d_nl, d_tmp, d_p0, d_lb, d_size, d_off, d_plus;
    uint64_t* dlen = d_len.ensure<uint64_t>(2 * (uint64_t(nf) + 1));
    uint64_t* ddoff = dlen + nf + 1;
    CDBG_HIP(hipMemcpyAsync(dlen, len.data(), nf * 8ull, hipMemcpyHostToDevice, st));
    CDBG_HIP(hipMemcpyAsync(ddoff, doff.data(), (nf + 1) * 8ull, hipMemcpyHostToDevice, st));
    { KSCOPE("k_close_files", st); k_close_files<<<blocks_for(nf), TPB, 0, st>>>(nf, dec, ddoff, dlen); }
    const uint64_t n_chunks = (D + NL_CHUNK - 1) / NL_CHUNK;
    uint32_t* cnt = d_cnt.ensure<uint32_t>(n_chunks);
    uint64_t* at = d_at.ensure<uint64_t>(n_chunks + 1);
    { KSCOPE("k_nl_count", st); k_nl_count<<<blocks_for(n_chunks), TPB, 0, st>>>(n_chunks, dec, ddoff, dlen, nf, cnt); }
    exclusive_scan<uint32_t, uint64_t>(cnt, at, n_chunks, at + n_chunks, d_tmp, st);
    uint64_t n_nl = 0;
    CDBG_HIP(hipMemcpyAsync(&n_nl, at + n_chunks, 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    uint64_t* nl = d_nl.ensure<uint64_t>(n_nl + 1);
    { KSCOPE("k_nl_write", st); k_nl_write<<<blocks_for(n_chunks), TPB, 0, st>>>(n_chunks, dec, ddoff, dlen, nf, at, nl); }
    d_cnt.release();
    d_at.release();
    // 4. where records start; each file's lines; item sizes
    uint64_t* p0 = d_p0.ensure<uint64_t>(nf);
    uint64_t* lb = d_lb.ensure<uint64_t>(uint64_t(nf) + 1);
    { KSCOPE("k_first_header", st); k_first_header<<<nf, TPB, 0, st>>>(dec, ddoff, dlen, p0); }
    { KSCOPE("k_file_lines", st); k_file_lines<<<blocks_for(uint64_t(nf) + 1), TPB, 0, st>>>(nf, nl, n_nl, ddoff, lb); }
    const uint64_t n_items = n_nl + nf;
    LineArgs la{dec, ddoff, nl, lb, p0, nf, n_items};
    uint64_t* size = d_size.ensure<uint64_t>(n_items);
    uint64_t* off = d_off.ensure<uint64_t>(n_items + 1);
    uint32_t* plus = d_plus.ensure<uint32_t>(1);
    CDBG_HIP(hipMemsetAsync(plus, 0, 4, st));
    { KSCOPE("k_line_size", st); k_line_size<<<blocks_for(n_items), TPB, 0, st>>>(la, size, plus); }
    exclusive_scan<uint64_t, uint64_t>(size, off, n_items, off + n_items, d_tmp, st);
    uint64_t total = 0;
    uint32_t h_plus = 0;
    CDBG_HIP(hipMemcpyAsync(&total, off + n_items, 8, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipMemcpyAsync(&h_plus, plus, 4, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    if (h_plus) return false;  // FASTQ quality lines: the host reader's rules
    // 5. the job's input: records, '\n' padding, colour offsets
    J.n_bytes = total;
    J.n_pad = ((total + 64) / 64) * 64 + 64;
    uint8_t* jb = J.bytes.ensure<uint8_t>(J.n_pad);
    CDBG_HIP(hipMemsetAsync(jb + total, '\n', J.n_pad - total, st));
    { KSCOPE("k_line_write", st); k_line_write<<<blocks_for(n_items), TPB, 0, st>>>(la, off, jb); }
    uint64_t* coff = J.color_off.ensure<uint64_t>(uint64_t(nf) + 1);
    { KSCOPE("k_color_offsets", st); k_color_offsets<<<blocks_for(uint64_t(nf) + 1), TPB, 0, st>>>(nf, lb, off, coff); }
    CDBG_HIP(hipGetLastError());
    J.h_color_off.resize(nf + 1);
    CDBG_HIP(hipMemcpyAsync(J.h_color_off.data(), coff, (nf + 1) * 8ull, hipMemcpyDeviceToHost, st));
    CDBG_HIP(hipStreamSynchronize(st));
    return true;
}

}  // namespace cdbg_amd
