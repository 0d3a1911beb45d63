// Device-memory model of one build job, shared by the phase-2 chunk planner
// and the server's admission control.
//
// Random sequence has a minimizer every (k - m + 2) / 2 positions, so about
// 2 / (k - m + 2) super-k-mers per base, and a super-k-mer's records carry
// one k-mer more than it starts: occurrences ~ bases + super-k-mers
// (measured 1.19 per base at k=31 m=21, 1.05 at k=63 m=27; 5% margin).
// Phase 1 holds the input (1 B/base), 2-bit bases and ACGT bitmap (3/8
// B/base), the tile staging of super-k-mer starts and ends (6 B/base) and
// the records with their sort buffers (~57 B per super-k-mer); phase 2 holds
// ~160 B per k-mer occurrence of its chunk with 64-bit k-mers (~230 B with
// 128-bit ones: wider tuples, sort buffers and node keys) plus the output
// streams (~6 B per occurrence).
#pragma once

#include <algorithm>
#include <cstdint>

namespace cdbg_amd {

constexpr uint64_t HBM_FIXED_BYTES = 64ull << 20;
inline double supers_per_base(uint32_t k, uint32_t m) { return 2.0 / double(k - m + 2); }
inline double occ_per_base(uint32_t k, uint32_t m) { return 1.05 * (1.0 + supers_per_base(k, m)); }
inline double p1_per_base(uint32_t k, uint32_t m) { return 1.0 + 0.375 + 6.0 + 57.0 * supers_per_base(k, m); }
// phase-2 device bytes per k-mer occurrence of a chunk, by key width
inline double p2_per_occ(uint32_t k) { return k <= 31 ? 166.0 : 236.0; }

// Phase-2 chunk (k-mer occurrences) planned when `usable` device bytes are
// free for phase 2: 60% of them for the chunk's working set, the rest left
// to the output streams that grow across chunks; at least 2^20, below 2^31.
inline uint64_t chunk_budget(uint64_t usable, uint32_t k) {
    const uint64_t b = uint64_t(0.6 * double(usable) / p2_per_occ(k));
    return std::min<uint64_t>(std::max<uint64_t>(b, 1u << 20), (1ull << 31) - 1);
}

}  // namespace cdbg_amd
