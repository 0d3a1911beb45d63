// 2-bit k-mer arithmetic on 64-bit (k <= 31) and 128-bit (k <= 63) words.
// First base in the highest used bits, A=0 C=1 G=2 T=3 (kmer.hpp:11-18).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "prims.h"

namespace cdbg_amd {

__host__ __device__ __forceinline__ uint64_t rc_word64(uint64_t x) {
    // complement, then reverse the 32 two-bit groups of the word
    x = ~x;
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    x = ((x >> 8) & 0x00FF00FF00FF00FFULL) | ((x & 0x00FF00FF00FF00FFULL) << 8);
    x = ((x >> 16) & 0x0000FFFF0000FFFFULL) | ((x & 0x0000FFFF0000FFFFULL) << 16);
    return (x >> 32) | (x << 32);
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <class K>
struct KOps;

template <>
struct KOps<uint64_t> {
    using K = uint64_t;
    static constexpr int WORDS = 1;
    __device__ static K zero() { return 0; }
    __device__ static K mask(uint32_t k) { return (uint64_t(1) << (2 * k)) - 1; }
    __device__ static K rc(K x, uint32_t k) { return rc_word64(x) >> (64 - 2 * k); }
    __device__ static K shift_append(K x, uint32_t nt, uint32_t k) { return ((x << 2) | nt) & mask(k); }
    // forward / reverse rolling update with base v
    __device__ static void roll(K& f, K& r, uint32_t v, uint32_t k) {
        f = ((f << 2) | v) & mask(k);
        r = (r >> 2) | (K(v ^ 3u) << (2 * (k - 1)));
    }
    __device__ static bool le(const K& a, const K& b) { return a <= b; }
    __device__ static K select(bool c, const K& a, const K& b) { return c ? a : b; }
    __device__ static bool eq(const K& a, const K& b) { return a == b; }
    __device__ static uint32_t base(const K& x, uint32_t i, uint32_t k) {  // i-th base from the left
        return uint32_t(x >> (2 * (k - 1 - i))) & 3u;
    }
};

template <>
struct KOps<Key128> {
    using K = Key128;
    static constexpr int WORDS = 2;
    __device__ static K zero() { return K{0, 0}; }
    __device__ static K mask(uint32_t k) {
        const uint32_t b = 2 * k;
        K m;
        if (b >= 64) {
            m.lo = ~0ull;
            m.hi = (b >= 128) ? ~0ull : ((uint64_t(1) << (b - 64)) - 1);
        } else {
            m.lo = (uint64_t(1) << b) - 1;
            m.hi = 0;
        }
        return m;
    }
    __device__ static K shl(const K& x, uint32_t s) {  // 0 < s < 64
        return K{x.lo << s, (x.hi << s) | (x.lo >> (64 - s))};
    }
    __device__ static K shr(const K& x, uint32_t s) {  // 0 <= s < 128
        if (s == 0) return x;
        if (s >= 64) return K{x.hi >> (s - 64), 0};
        return K{(x.lo >> s) | (x.hi << (64 - s)), x.hi >> s};
    }
    __device__ static K rc(const K& x, uint32_t k) {
        const K r{rc_word64(x.hi), rc_word64(x.lo)};
        return shr(r, 128 - 2 * k);
    }
    __device__ static K shift_append(const K& x, uint32_t nt, uint32_t k) {
        K y = shl(x, 2);
        y.lo |= nt;
        const K m = mask(k);
        return K{y.lo & m.lo, y.hi & m.hi};
    }
    __device__ static void roll(K& f, K& r, uint32_t v, uint32_t k) {
        f = shift_append(f, v, k);
        r = shr(r, 2);
        const uint32_t sh = 2 * (k - 1);
        const uint64_t c = uint64_t(v ^ 3u);
        if (sh >= 64)
            r.hi |= c << (sh - 64);
        else
            r.lo |= c << sh;
    }
    __device__ static bool le(const K& a, const K& b) { return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo); }
    __device__ static K select(bool c, const K& a, const K& b) {  // word by word: no select of addresses
        return K{c ? a.lo : b.lo, c ? a.hi : b.hi};
    }
    __device__ static bool eq(const K& a, const K& b) { return a.lo == b.lo && a.hi == b.hi; }
    __device__ static uint32_t base(const K& x, uint32_t i, uint32_t k) {
        const uint32_t s = 2 * (k - 1 - i);
        return s >= 64 ? uint32_t(x.hi >> (s - 64)) & 3u : uint32_t(x.lo >> s) & 3u;
    }
};

}  // namespace cdbg_amd
