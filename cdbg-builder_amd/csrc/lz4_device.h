// LZ4 block codec on the device (frames of bucket files).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdbg_amd {

// Largest frame the compressor takes: positions and match offsets fit 16 bits.
constexpr uint64_t LZ_FRAME = 64 * 1024;

// Output slot of a frame of n bytes: the block format's worst case
// (n + n/255 + 16) rounded up to 16 bytes, so every slot starts aligned.
__host__ __device__ inline uint64_t lz4_slot(uint64_t n) { return (n + n / 255 + 16 + 15) & ~uint64_t(15); }

// Frame f = src[foff[f], foff[f+1]) (<= LZ_FRAME bytes) is compressed into
// slots + slot_off[f] (16-byte aligned, lz4_slot bytes); its block size goes
// to csize[f].  src must be 4-byte aligned and readable 8 bytes past its
// last frame.
void launch_lz4_compress(const uint8_t* src, const uint64_t* foff, uint32_t nf, uint8_t* slots,
                         const uint64_t* slot_off, uint32_t* csize, hipStream_t st);

// Frame f's block src[soff[f], soff[f] + csize[f]) is decoded into
// dst[doff[f], doff[f+1]); bad[f] = 1 when the block is malformed or does not
// decode to exactly that many bytes (nothing is read or written outside the
// frame's ranges either way).
void launch_lz4_decompress(const uint8_t* src, const uint64_t* soff, const uint32_t* csize, uint32_t nf,
                           uint8_t* dst, const uint64_t* doff, uint8_t* bad, hipStream_t st);

// Block f's bytes slots[slot_off[f], + csize[f]) moved to out + out_off[f],
// after an 8-byte frame header [u32 uncompressed][u32 compressed]
// (uncompressed = foff[f+1] - foff[f]).
void launch_lz4_pack(const uint8_t* slots, const uint64_t* slot_off, const uint32_t* csize, const uint64_t* foff,
                     uint32_t nf, uint8_t* out, const uint64_t* out_off, hipStream_t st);

}  // namespace cdbg_amd
