// Checker shim: exposes the reference's own header-only primitives through a
// C interface so the tests can compare the oracle against them.  Built only
// where the reference tree is mounted (oracle/Makefile target `ref`), output
// goes to oracle/_ref/ and is never committed or shipped.
//
// Compiled headers (read in place from the reference tree):
//   include/kmer.hpp                       reverse_complement, canonical, nuc_to_2bit
//   include/phase1_bucket_write/minimizer.hpp   nthash_init/roll, windowed_min, compute_best_m
//   include/phase1_bucket_write/super_kmer.hpp  pack_2bit, varint_write/read
#include <cstdint>
#include <vector>

#include "kmer.hpp"
#include "phase1_bucket_write/minimizer.hpp"
#include "phase1_bucket_write/super_kmer.hpp"

extern "C" {

uint32_t ref_compute_best_m(uint32_t k) { return cdbg::compute_best_m(k); }

uint8_t ref_nuc_to_2bit(char c) { return cdbg::nuc_to_2bit(c); }

// Rolled canonical ntHash at every m-mer position (init at 0, then roll).
void ref_rolled_hashes(const uint8_t* codes, uint64_t n, uint32_t m, uint64_t* out) {
    uint64_t f = 0, r = 0;
    if (!cdbg::nthash_init(codes, m, f, r)) return;
    out[0] = cdbg::canonical_mhash(f, r);
    for (uint64_t p = 1; p + m <= n; ++p) {
        cdbg::nthash_roll(codes[p - 1], codes[p + m - 1], m, f, r);
        out[p] = cdbg::canonical_mhash(f, r);
    }
}

// windowed_min over an ACGT-only run: min hash of every (k-1)-mer.
void ref_window_min(const uint8_t* codes, uint64_t n, uint32_t k, uint32_t m, uint64_t* out) {
    cdbg::windowed_min mq;
    if (!mq.init(codes, m, (int32_t)(k - m))) return;
    out[0] = mq.min_hash();
    for (uint64_t j = 1; j + k - 1 <= n; ++j) {
        mq.advance();
        out[j] = mq.min_hash();
    }
}

uint64_t ref_revcomp64(uint64_t x, uint32_t k) {
    return cdbg::reverse_complement<uint64_t>(x, k);
}
void ref_revcomp128(uint64_t lo, uint64_t hi, uint32_t k, uint64_t* out) {
    __uint128_t x = ((__uint128_t)hi << 64) | lo;
    __uint128_t r = cdbg::reverse_complement<__uint128_t>(x, k);
    out[0] = (uint64_t)r;
    out[1] = (uint64_t)(r >> 64);
}

// varint_write of v into out (returns the byte count).
uint32_t ref_varint(uint64_t v, uint8_t* out) {
    std::vector<uint8_t> b;
    cdbg::varint_write(v, b);
    for (size_t i = 0; i < b.size(); ++i) out[i] = b[i];
    return (uint32_t)b.size();
}

// write_super_kmer_packed over unpacked codes; returns the byte count.
uint32_t ref_record(uint8_t flags, const uint32_t* colors, uint32_t nc, const uint8_t* codes,
                    uint32_t len, uint8_t* out) {
    std::vector<uint8_t> packed;
    cdbg::pack_2bit(codes, len, packed);
    std::vector<uint8_t> rec;
    cdbg::write_super_kmer_packed(flags, colors, nc, packed.data(), len, rec);
    for (size_t i = 0; i < rec.size(); ++i) out[i] = rec[i];
    return (uint32_t)rec.size();
}

}  // extern "C"
