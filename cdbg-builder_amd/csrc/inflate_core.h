// gzip (RFC 1952) / DEFLATE (RFC 1951) decoding, written once for the
// device inflater (one wavefront per file, every lane running the decoder in
// step, the table fills and match copies spread over the lanes) and for its
// host check.  The IO policy supplies input bytes, output, the tables'
// memory, the lane index and a barrier.
//
// Huffman decoding is canonical (RFC 1951 3.2.2): codes of one length are
// consecutive and shorter codes precede longer ones, so a symbol is found by
// walking the lengths with the running first code of each length and the
// count of codes of that length.  A 2^FAST-entry table holds the symbol and
// length of every bit pattern whose code is at most FAST bits long (each
// entry decoded by that same walk, so the table needs no code assignment);
// longer codes take the walk itself.  Everything is inlined into the kernel:
// the decoder's state then lives in registers, not in private memory.
#pragma once

#include <stdint.h>

namespace cdbg_amd {
namespace infl {

constexpr int MAXBITS = 15, FAST = 10;
constexpr uint32_t NLIT = 288, NDIST = 32, NCL = 19;

struct Huff {
    uint16_t fast[1 << FAST];  // symbol | length << 9; 0: a code longer than FAST bits
    uint16_t count[MAXBITS + 1];
    uint16_t sym[NLIT];  // symbols ordered by (code length, symbol)
};

struct Tables {
    Huff lit, dist;
    // literal runs: for each FAST-bit pattern, the literals (up to three, 8
    // bits each) whose codes lie wholly inside it, their count << 24 and
    // their total code length << 26; 0 when it does not start with a literal
    uint32_t multi[1 << FAST];
    uint16_t offs[MAXBITS + 2];  // table build: next slot per code length
    uint8_t lens[NLIT + NDIST];
    int32_t err;                 // shared verdict of a table build
};

enum : int {
    OK = 0,
    E_TRUNCATED = 1,   // the stream ends inside a block or header
    E_HEADER = 2,      // not a gzip member (bad method or flags)
    E_BTYPE = 3,       // block type 3
    E_STORED = 4,      // stored length and its complement disagree
    E_CODES = 5,       // over-subscribed or incomplete code lengths
    E_SYMBOL = 6,      // no symbol for the bits read
    E_DISTANCE = 7,    // a distance before the start of the output
    E_REPEAT = 8,      // a length repeat with nothing to repeat, or past the end
    E_SIZE = 9,        // trailer size differs from the output
    E_MEMBERS = 10,    // more members than the member table holds
};

// RFC 1951 3.2.5: length codes 257..285 (index 0..28) and distance codes
// 0..29, computed (no tables: nothing indexed lands in private memory).
// Lengths: 3..10 with no extra bits, then groups of four with 1..5 extra
// bits starting at 11, 19, 35, 67, 131; code 285 is 258.  Distances: 1..4,
// then pairs with 1..13 extra bits starting at 5, 9, 17, ...
__host__ __device__ __forceinline__ uint32_t len_extra(uint32_t i) { return i < 8 || i == 28 ? 0u : (i - 4) >> 2; }
__host__ __device__ __forceinline__ uint32_t len_base(uint32_t i) {
    return i < 8 ? 3 + i : i == 28 ? 258u : ((4 + (i & 3)) << len_extra(i)) + 3;
}
__host__ __device__ __forceinline__ uint32_t dist_extra(uint32_t i) { return i < 4 ? 0u : (i >> 1) - 1; }
__host__ __device__ __forceinline__ uint32_t dist_base(uint32_t i) { return i < 4 ? i + 1 : ((2 + (i & 1)) << dist_extra(i)) + 1; }
// Order of the code-length code lengths (RFC 1951 3.2.7): 16 17 18 0 8 7 9 6
// 10 5 11 4 12 3 13 2 14 1 15, five bits each.
__host__ __device__ __forceinline__ uint32_t cl_order(uint32_t i) {
    const uint64_t lo = 16ull | 17ull << 5 | 18ull << 10 | 0ull << 15 | 8ull << 20 | 7ull << 25 | 9ull << 30 |
                        6ull << 35 | 10ull << 40 | 5ull << 45 | 11ull << 50 | 4ull << 55;
    const uint64_t hi = 12ull | 3ull << 5 | 13ull << 10 | 2ull << 15 | 14ull << 20 | 1ull << 25 | 15ull << 30;
    return uint32_t((i < 12 ? lo >> (5 * i) : hi >> (5 * (i - 12))) & 31u);
}

// The decoder over an IO policy.  IO provides:
//   uint64_t in_len();                  input bytes
//   uint64_t load8(uint64_t pos);       bytes [pos, pos + 8), zero past the end
//   void put(uint32_t b);               one output byte (lane 0 writes)
//   bool copy(uint32_t dist, uint32_t n);  n bytes from dist back (false: dist > output)
//   void copy_in(uint64_t pos, uint32_t n);  n input bytes (a stored block)
//   uint64_t produced();                output bytes so far
//   uint32_t lane(), nlanes(); void sync();
//   Tables* tables();
//   bool member(uint64_t out_start, uint64_t out_len, uint32_t crc);  a finished member (false: table full)
template <class IO>
struct Inflater {
    IO& io;
    uint64_t bits = 0;  // bit buffer, next bit lowest
    uint32_t cnt = 0;   // bits in it
    uint64_t ipos = 0;  // input bytes taken into the buffer (may pass the end: zero bytes)

    __host__ __device__ __forceinline__ explicit Inflater(IO& i) : io(i) {}

    __host__ __device__ __forceinline__ void refill() {
        if (cnt > 56) return;
        const uint32_t take = (63 - cnt) >> 3;  // whole bytes that fit
        const uint64_t w = io.load8(ipos);
        const uint64_t m = take >= 8 ? ~0ull : ((uint64_t(1) << (8 * take)) - 1);
        bits |= (w & m) << cnt;
        cnt += 8 * take;
        ipos += take;
    }
    __host__ __device__ __forceinline__ uint32_t need(uint32_t n) {  // n <= 32 bits, consumed
        refill();
        const uint32_t v = uint32_t(bits & ((uint64_t(1) << n) - 1));
        bits >>= n;
        cnt -= n;
        return v;
    }
    __host__ __device__ __forceinline__ uint64_t consumed_bits() const { return 8 * ipos - cnt; }
    __host__ __device__ __forceinline__ bool overrun() const { return consumed_bits() > 8 * io.in_len(); }
    __host__ __device__ __forceinline__ void align() {  // to the next byte of the stream
        const uint32_t drop = cnt & 7;
        bits >>= drop;
        cnt -= drop;
    }
    __host__ __device__ __forceinline__ uint64_t byte_pos() const { return consumed_bits() >> 3; }  // after align()
    __host__ __device__ __forceinline__ void seek(uint64_t pos) {  // restart the bit buffer at byte pos
        bits = 0;
        cnt = 0;
        ipos = pos;
    }

    // Canonical walk over lengths 1..maxlen of the bits in `b` (next bit
    // lowest): symbol | length << 9, or 0.
    static __host__ __device__ __forceinline__ uint32_t walk(const Huff& h, uint64_t b, int maxlen) {
        int code = 0, first = 0, index = 0;
        for (int len = 1; len <= maxlen; ++len) {
            code |= int((b >> (len - 1)) & 1u);
            const int count = h.count[len];
            if (code - first < count) return uint32_t(h.sym[index + code - first]) | (uint32_t(len) << 9);
            index += count;
            first = (first + count) << 1;
            code <<= 1;
        }
        return 0;
    }

    // Builds h from n code lengths; incomplete codes are allowed (their
    // unused patterns decode to an error), over-subscribed ones are not.
    __host__ __device__ __forceinline__ int build(Huff& h, const uint8_t* len, uint32_t n) {
        Tables* T = io.tables();
        if (io.lane() == 0) {
            for (int l = 0; l <= MAXBITS; ++l) h.count[l] = 0;
            for (uint32_t s = 0; s < n; ++s) h.count[len[s]]++;
            h.count[0] = 0;
            int left = 1, err = OK;
            for (int l = 1; l <= MAXBITS; ++l) {
                left = (left << 1) - int(h.count[l]);
                if (left < 0) err = E_CODES;
            }
            uint16_t* offs = T->offs;
            offs[1] = 0;
            for (int l = 1; l <= MAXBITS; ++l) offs[l + 1] = uint16_t(offs[l] + h.count[l]);
            for (uint32_t s = 0; s < n; ++s)
                if (len[s]) h.sym[offs[len[s]]++] = uint16_t(s);
            T->err = err;
        }
        io.sync();
        const int err = T->err;
        if (err != OK) return err;
        for (uint32_t e = io.lane(); e < (1u << FAST); e += io.nlanes()) h.fast[e] = uint16_t(walk(h, e, FAST));
        io.sync();
        return OK;
    }

    __host__ __device__ __forceinline__ int decode(const Huff& h, uint32_t& sym) {
        refill();
        uint32_t e = h.fast[bits & ((1u << FAST) - 1)];
        if (!e) e = walk(h, bits, MAXBITS);
        if (!e) return E_SYMBOL;
        const uint32_t n = e >> 9;
        bits >>= n;
        cnt -= n;
        sym = e & 511u;
        return overrun() ? E_TRUNCATED : OK;  // past the end every bit reads 0: stop there
    }

    // multi[] from the literal/length table (its fast entries decode the
    // patterns' prefixes; a later code counts only if its bits are known).
    __host__ __device__ __forceinline__ void build_multi() {
        Tables* T = io.tables();
        for (uint32_t e = io.lane(); e < (1u << FAST); e += io.nlanes()) {
            uint32_t used = 0, n = 0, packed = 0;
            while (n < 3) {
                const uint32_t ent = T->lit.fast[e >> used];  // the unknown top bits read as 0
                const uint32_t len = ent >> 9, sym = ent & 511u;
                if (!ent || len > uint32_t(FAST) - used || sym >= 256) break;
                packed |= sym << (8 * n);
                ++n;
                used += len;
            }
            T->multi[e] = n ? packed | (n << 24) | (used << 26) : 0u;
        }
        io.sync();
    }

    __host__ __device__ __forceinline__ int fixed_tables() {
        Tables* T = io.tables();
        for (uint32_t s = io.lane(); s < NLIT + NDIST; s += io.nlanes())
            T->lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : s < NLIT ? 8 : 5;
        io.sync();
        int r = build(T->lit, T->lens, NLIT);
        if (r == OK) r = build(T->dist, T->lens + NLIT, NDIST);
        return r;
    }

    __host__ __device__ __forceinline__ int dynamic_tables() {
        Tables* T = io.tables();
        const uint32_t nlen = need(5) + 257, ndist = need(5) + 1, ncode = need(4) + 4;
        if (nlen > 286 || ndist > 30) return E_CODES;
        uint8_t* lens = T->lens;
        io.sync();
        if (io.lane() == 0)
            for (uint32_t i = 0; i < NLIT + NDIST; ++i) lens[i] = 0;
        io.sync();
        for (uint32_t i = 0; i < ncode; ++i) {
            const uint32_t v = need(3);
            if (io.lane() == 0) lens[NLIT + NDIST - NCL + cl_order(i)] = uint8_t(v);
        }
        io.sync();
        // the code-length code, in the literal table for now; its lengths sit
        // at the end of `lens`, below which the decoded lengths are written
        // (the literal/length ones from 0, the distance ones from NLIT)
        int r = build(T->lit, lens + NLIT + NDIST - NCL, NCL);
        if (r != OK) return r;
        uint32_t idx = 0, prev = 0;  // every lane decodes the same lengths; lane 0 stores them
        auto store = [&](uint32_t at, uint32_t v) {
            if (io.lane() == 0) lens[at < nlen ? at : NLIT + (at - nlen)] = uint8_t(v);
        };
        while (idx < nlen + ndist) {
            uint32_t sym;
            if ((r = decode(T->lit, sym)) != OK) return r;
            if (sym < 16) {
                store(idx++, sym);
                prev = sym;
                continue;
            }
            uint32_t rep, val = 0;
            if (sym == 16) {
                if (idx == 0) return E_REPEAT;
                val = prev;
                rep = 3 + need(2);
            } else if (sym == 17) {
                rep = 3 + need(3);
            } else {
                rep = 11 + need(7);
            }
            if (idx + rep > nlen + ndist) return E_REPEAT;
            while (rep--) store(idx++, val);
            prev = val;
        }
        io.sync();
        if (lens[256] == 0) return E_CODES;  // no end-of-block code
        // the code-length code's lengths (at NLIT + NDIST - NCL and up) may
        // be left beyond the distance lengths: clear what the distance code
        // does not use
        if (io.lane() == 0)
            for (uint32_t i = NLIT + ndist; i < NLIT + NDIST; ++i) lens[i] = 0;
        io.sync();
        r = build(T->lit, lens, nlen);
        if (r == OK) r = build(T->dist, lens + NLIT, ndist);
        return r;
    }

    // One DEFLATE stream from the current bit position.
    __host__ __device__ __forceinline__ int inflate() {
        Tables* T = io.tables();
        for (;;) {
            const uint32_t final = need(1), type = need(2);
            int r = OK;
            if (type == 0) {
                align();
                const uint32_t len = need(16), nlen = need(16);
                if ((len ^ 0xFFFFu) != nlen) return E_STORED;
                // the buffer holds whole bytes now: take them, then the input
                const uint64_t at = byte_pos();
                if (at + len > io.in_len()) return E_TRUNCATED;
                io.copy_in(at, len);
                seek(at + len);
            } else if (type == 1 || type == 2) {
                r = type == 1 ? fixed_tables() : dynamic_tables();
                if (r != OK) return r;
                build_multi();
                for (;;) {
                    refill();
                    const uint32_t mt = T->multi[bits & ((1u << FAST) - 1)];
                    if (mt) {  // one to three literals
                        const uint32_t used = mt >> 26, nl = (mt >> 24) & 3u;
                        bits >>= used;
                        cnt -= used;
                        if (overrun()) return E_TRUNCATED;
                        io.put(mt & 0xFFu);
                        if (nl > 1) io.put((mt >> 8) & 0xFFu);
                        if (nl > 2) io.put((mt >> 16) & 0xFFu);
                        continue;
                    }
                    uint32_t sym;
                    if ((r = decode(T->lit, sym)) != OK) return r;
                    if (sym < 256) {
                        io.put(sym);
                        continue;
                    }
                    if (sym == 256) break;
                    sym -= 257;
                    if (sym >= 29) return E_SYMBOL;
                    const uint32_t n = len_base(sym) + need(len_extra(sym));
                    uint32_t ds;
                    if ((r = decode(T->dist, ds)) != OK) return r;
                    if (ds >= 30) return E_SYMBOL;
                    const uint32_t dist = dist_base(ds) + need(dist_extra(ds));
                    if (!io.copy(dist, n)) return E_DISTANCE;
                }
            } else {
                return E_BTYPE;
            }
            if (overrun()) return E_TRUNCATED;
            if (final) return OK;
        }
    }

    __host__ __device__ __forceinline__ uint32_t byte_at(uint64_t pos) { return uint32_t(io.load8(pos) & 0xFFu); }

    // Every gzip member of the input (RFC 1952 2.3); bytes after a member
    // that do not start another are ignored (as gzread does).  A file that
    // does not start with the gzip magic is taken as it is.
    __host__ __device__ __forceinline__ int gunzip() {
        const uint64_t n = io.in_len();
        if (n < 2 || byte_at(0) != 0x1f || byte_at(1) != 0x8b) {
            io.copy_in(0, uint32_t(n));  // callers cap plain files below 4 GiB
            return io.member(0, n, 0) ? OK : E_MEMBERS;
        }
        uint64_t pos = 0;
        bool first = true;
        while (first || (pos + 2 <= n && byte_at(pos) == 0x1f && byte_at(pos + 1) == 0x8b)) {
            first = false;
            if (pos + 10 > n) return E_TRUNCATED;
            if (byte_at(pos + 2) != 8) return E_HEADER;
            const uint32_t flg = byte_at(pos + 3);
            if (flg & 0xE0) return E_HEADER;
            uint64_t p = pos + 10;
            if (flg & 4) {  // FEXTRA
                if (p + 2 > n) return E_TRUNCATED;
                p += 2 + (byte_at(p) | (byte_at(p + 1) << 8));
            }
            if (flg & 8) {  // FNAME
                while (p < n && byte_at(p)) ++p;
                ++p;
            }
            if (flg & 16) {  // FCOMMENT
                while (p < n && byte_at(p)) ++p;
                ++p;
            }
            if (flg & 2) p += 2;  // FHCRC
            if (p > n) return E_TRUNCATED;
            const uint64_t out0 = io.produced();
            seek(p);
            const int r = inflate();
            if (r != OK) return r;
            align();
            const uint64_t t = byte_pos();
            if (t + 8 > n) return E_TRUNCATED;
            const uint32_t crc = byte_at(t) | (byte_at(t + 1) << 8) | (byte_at(t + 2) << 16) | (byte_at(t + 3) << 24);
            const uint32_t isize =
                byte_at(t + 4) | (byte_at(t + 5) << 8) | (byte_at(t + 6) << 16) | (byte_at(t + 7) << 24);
            const uint64_t got = io.produced() - out0;
            if (uint32_t(got) != isize) return E_SIZE;
            if (!io.member(out0, got, crc)) return E_MEMBERS;
            pos = t + 8;
        }
        return OK;
    }
};

// CRC-32 (ISO 3309, reflected polynomial 0xEDB88320) over GF(2): a*b mod P
// for reflected polynomials, and x^(8n) mod P by squaring, so that the CRC of
// A followed by B is crc(A) * x^(8|B|) + crc(B).
__host__ __device__ __forceinline__ uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (uint32_t m = 0x80000000u; m; m >>= 1) {  // the reflected x^0 term is the top bit
        if (a & m) p ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;  // b * x
    }
    return p;
}
__host__ __device__ __forceinline__ uint32_t x_pow_8n(uint64_t n) {  // x^(8n) mod P, reflected
    uint32_t r = 0x80000000u;  // 1
    uint32_t sq = 0x00800000u;  // x^8
    while (n) {
        if (n & 1) r = gf2_mulmod(r, sq);
        sq = gf2_mulmod(sq, sq);
        n >>= 1;
    }
    return r;
}
__host__ __device__ __forceinline__ uint32_t crc32_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b) {
    return gf2_mulmod(x_pow_8n(len_b), crc_a) ^ crc_b;
}

}  // namespace infl
}  // namespace cdbg_amd
