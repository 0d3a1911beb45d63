// LZ4 block codec written from the public LZ4 block format specification
// (sequences of token | literal-length bytes | literals | 16-bit LE offset |
// match-length bytes; minimum match 4; the last 5 bytes of a block are
// literals and the last match starts at least 12 bytes before the end).
// Bucket files frame each block as [u32 uncompressed][u32 compressed][block]
// (bucket_io.hpp:19-40); any valid block decodes with LZ4_decompress_safe.
#include "lz4_block.h"

#include <cstring>
#include <vector>

namespace cdbg_amd {

namespace {
constexpr size_t MIN_MATCH = 4;
constexpr size_t LAST_LITERALS = 5;
constexpr size_t MF_LIMIT = 12;
constexpr int HASH_LOG = 16;

inline uint32_t read32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> (32 - HASH_LOG); }

inline void put_len(std::vector<uint8_t>& out, size_t len) {  // length beyond the token's 15
    while (len >= 255) {
        out.push_back(255);
        len -= 255;
    }
    out.push_back(uint8_t(len));
}

void emit(std::vector<uint8_t>& out, const uint8_t* lit, size_t nlit, size_t off, size_t mlen) {
    const size_t ml = mlen ? mlen - MIN_MATCH : 0;
    uint8_t tok = uint8_t((nlit >= 15 ? 15 : nlit) << 4) | uint8_t(mlen ? (ml >= 15 ? 15 : ml) : 0);
    out.push_back(tok);
    if (nlit >= 15) put_len(out, nlit - 15);
    out.insert(out.end(), lit, lit + nlit);
    if (!mlen) return;
    out.push_back(uint8_t(off & 0xff));
    out.push_back(uint8_t(off >> 8));
    if (ml >= 15) put_len(out, ml - 15);
}
}  // namespace

void lz4_compress_block(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
    out.clear();
    out.reserve(n + n / 255 + 16);
    std::vector<int64_t> table(size_t(1) << HASH_LOG, -1);
    size_t anchor = 0, ip = 0;
    if (n >= MF_LIMIT + 1) {
        const size_t match_limit = n - LAST_LITERALS;  // matches end before this
        const size_t search_end = n - MF_LIMIT;        // matches start before this
        while (ip < search_end) {
            const uint32_t seq = read32(src + ip);
            const uint32_t h = hash4(seq);
            const int64_t cand = table[h];
            table[h] = int64_t(ip);
            if (cand >= 0 && ip - size_t(cand) <= 65535 && read32(src + cand) == seq) {
                size_t m = MIN_MATCH;
                while (ip + m < match_limit && src[cand + m] == src[ip + m]) ++m;
                emit(out, src + anchor, ip - anchor, ip - size_t(cand), m);
                ip += m;
                anchor = ip;
                continue;
            }
            ++ip;
        }
    }
    emit(out, src + anchor, n - anchor, 0, 0);  // last literals
}

bool lz4_decompress_block(const uint8_t* src, size_t n, uint8_t* dst, size_t cap, size_t& out_len) {
    size_t ip = 0, op = 0;
    while (ip < n) {
        const uint8_t tok = src[ip++];
        size_t lit = tok >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (ip + lit > n || op + lit > cap) return false;
        std::memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip >= n) break;  // the last sequence carries literals only
        if (ip + 2 > n) return false;
        const size_t off = size_t(src[ip]) | (size_t(src[ip + 1]) << 8);
        ip += 2;
        if (off == 0 || off > op) return false;
        size_t ml = tok & 15;
        if (ml == 15) {
            uint8_t b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += MIN_MATCH;
        if (op + ml > cap) return false;
        for (size_t i = 0; i < ml; ++i) dst[op + i] = dst[op - off + i];
        op += ml;
    }
    out_len = op;
    return true;
}

}  // namespace cdbg_amd
