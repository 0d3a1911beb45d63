// LZ4 block codec (public block format) used for bucket-file frames.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace cdbg_amd {

// Compresses src[0..n) into one LZ4 block (replaces `out`).
void lz4_compress_block(const uint8_t* src, size_t n, std::vector<uint8_t>& out);
// Decodes one block into dst (capacity cap); false on malformed input.
bool lz4_decompress_block(const uint8_t* src, size_t n, uint8_t* dst, size_t cap, size_t& out_len);

}  // namespace cdbg_amd
