// Host check of the device inflater's decoding core: decodes a gzip file with
// the same Inflater the device kernel runs (one lane, a plain output vector),
// checks every member's CRC-32 with a bitwise CRC and its combination rule
// over two halves, and writes the output, or prints the error code.
//   inflate_host_check <in.gz> <out>      exit 0 ok, 3 decode error
#define __host__
#define __device__
#define __forceinline__ inline __attribute__((always_inline))
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../cdbg-builder_amd/csrc/inflate_core.h"

using namespace cdbg_amd::infl;

struct HostIO {
    const uint8_t* in;
    uint64_t n;
    std::vector<uint8_t> out;
    Tables T;
    struct M {
        uint64_t start, len;
        uint32_t crc;
    };
    std::vector<M> members;
    uint64_t in_len() { return n; }
    uint64_t load8(uint64_t p) {
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i)
            if (p + i < n) v |= uint64_t(in[p + i]) << (8 * i);
        return v;
    }
    void put(uint32_t b) { out.push_back(uint8_t(b)); }
    bool copy(uint32_t d, uint32_t len) {
        if (d == 0 || d > out.size()) return false;
        for (uint32_t i = 0; i < len; ++i) out.push_back(out[out.size() - d]);
        return true;
    }
    void copy_in(uint64_t p, uint32_t len) { out.insert(out.end(), in + p, in + p + len); }
    uint64_t produced() { return out.size(); }
    uint32_t lane() { return 0; }
    uint32_t nlanes() { return 1; }
    void sync() {}
    Tables* tables() { return &T; }
    bool member(uint64_t s, uint64_t l, uint32_t c) {
        members.push_back({s, l, c});
        return members.size() <= 16;
    }
};

static uint32_t crc_bitwise(const uint8_t* p, uint64_t n) {
    uint32_t c = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
    }
    return ~c;
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<uint8_t> in;
    uint8_t buf[65536];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) in.insert(in.end(), buf, buf + r);
    std::fclose(f);
    HostIO io{in.data(), in.size(), {}, {}, {}};
    Inflater<HostIO> inf(io);
    const int e = inf.gunzip();
    if (e != OK) {
        std::printf("error %d\n", e);
        return 3;
    }
    const bool plain = !(in.size() >= 2 && in[0] == 0x1f && in[1] == 0x8b);
    for (auto& m : io.members) {
        if (plain) break;
        const uint8_t* p = io.out.data() + m.start;
        const uint32_t whole = crc_bitwise(p, m.len);
        const uint64_t h = m.len / 3;  // the combination rule over two unequal parts
        const uint32_t comb = crc32_combine(crc_bitwise(p, h), crc_bitwise(p + h, m.len - h), m.len - h);
        if (whole != m.crc || comb != whole) {
            std::printf("crc mismatch\n");
            return 4;
        }
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    if (!io.out.empty()) std::fwrite(io.out.data(), 1, io.out.size(), o);
    std::fclose(o);
    std::printf("ok %zu members %zu\n", io.out.size(), io.members.size());
    return 0;
}
