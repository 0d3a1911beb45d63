// Phase 3 (stitch) and phase 4 (emit) on the host: frag_unitigs.bin ->
// monochromatic unitigs -> <out>.fa, <out>.u2c, the finished
// <out>.color_sets and <out>.metadata.txt.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cdbg_gpu.h"

namespace cdbg_amd {

constexpr uint64_t COLOR_FOREIGN = ~0ull;  // COLOR_RUN_FOREIGN (bucket_walker.hpp:96)
constexpr uint8_t FRAG_OPEN_LEFT = 1, FRAG_OPEN_RIGHT = 2;

// Unitigs of one stitch, kept as `lists` each sorted stably by colour-set id;
// the emit order is the stable merge of the lists in list order (ties: the
// earlier list first), i.e. ascending cid, then the order the stitch made
// them in.
struct UnitigList {
    std::string bases;
    struct Item {
        uint64_t cid;
        uint64_t off;  // into bases
        uint64_t len;
    };
    std::vector<Item> items;
};

struct Unitigs {
    uint32_t k = 0;
    std::vector<UnitigList> lists;
    cdbg_stitch_stats stats{};
    // emit order, computed on first use
    std::vector<std::pair<uint32_t, uint64_t>> order;  // (list, item)
    void build_order();
};

// Stitches the fragment records in buf[0, n) (the frag_unitigs.bin byte
// stream) with `threads` host threads.
Unitigs* stitch_fragments(const uint8_t* buf, uint64_t n, uint32_t k, uint32_t threads);

}  // namespace cdbg_amd
