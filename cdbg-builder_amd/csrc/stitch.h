// Phase 3 (stitch) and phase 4 (emit) on the host: frag_unitigs.bin ->
// monochromatic unitigs -> <out>.fa, <out>.u2c, the finished
// <out>.color_sets and <out>.metadata.txt.
#pragma once

#include <cstdint>
#include <memory>
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
    struct Item {
        uint64_t cid;
        uint64_t off;  // into the list's bases
        uint64_t len;
    };
    std::string bases;
    std::vector<Item> items;
    // A list the device stitch filled: bases and items live in host memory
    // that `keep` owns (pinned), and stand in for the two vectors.
    std::shared_ptr<void> keep;
    const char* ext_bases = nullptr;
    const Item* ext_items = nullptr;
    uint64_t ext_n = 0;

    uint64_t size() const { return ext_items ? ext_n : items.size(); }
    const Item& item(uint64_t i) const { return ext_items ? ext_items[i] : items[i]; }
    const char* seq() const { return ext_items ? ext_bases : bases.data(); }
};

struct Unitigs {
    uint32_t k = 0;
    uint32_t threads = 1;  // host threads of the stitch, used again by emit
    std::vector<UnitigList> lists;
    cdbg_stitch_stats stats{};
    // emit order, computed on first use (left empty when a single list holds
    // every unitig: that list's order is the emit order)
    std::vector<std::pair<uint32_t, uint64_t>> order;  // (list, item)
    int single = -2;  // the one non-empty list, -1 none, -2 not computed
    void build_order();
    uint64_t size() const;
    std::pair<uint32_t, uint64_t> at(uint64_t i) const {
        return single >= 0 ? std::pair<uint32_t, uint64_t>(uint32_t(single), i) : order[i];
    }
    const UnitigList::Item& item(uint64_t i) const {
        const auto a = at(i);
        return lists[a.first].item(a.second);
    }
    const char* seq(uint64_t i) const {
        const auto a = at(i);
        return lists[a.first].seq() + lists[a.first].item(a.second).off;
    }
};

// Stitches the fragment records in buf[0, n) (the frag_unitigs.bin byte
// stream) with `threads` host threads.
Unitigs* stitch_fragments(const uint8_t* buf, uint64_t n, uint32_t k, uint32_t threads);

// Start of every record of the frag_unitigs.bin stream buf[0, n), parsed by
// `threads` threads; throws on a record that does not fit.
std::vector<uint64_t> fragment_offsets(const uint8_t* buf, uint64_t n, uint32_t threads);

// The tigs that close on themselves, over the fragments `frags` (ascending
// fragment numbers; the record of frags[l] at buf + rec[l]) whose sides pair
// up among themselves: partner[2 l + s] is the (fragment << 1 | side) joined
// to side s of frags[l], or ~0u.  Walked in fragment order as the host stitch
// walks them; the unitigs go to `out`, sorted stably by colour-set id.
void stitch_cycles(const uint8_t* buf, uint64_t n, const std::vector<uint64_t>& rec, uint32_t k,
                   const std::vector<uint32_t>& frags, const std::vector<uint32_t>& partner, UnitigList& out,
                   cdbg_stitch_stats& st);

// The whole file at `path`, read by up to `threads` threads; n = its size.
std::unique_ptr<uint8_t[]> read_file(const char* path, uint32_t threads, uint64_t& n);
uint64_t file_size_of(const char* path);
// Reads the n bytes of the file at `path` into dst with up to `threads` threads.
void read_file_into(const char* path, uint8_t* dst, uint64_t n, uint32_t threads);
// Writes n bytes from src as the file at `path` (created or truncated) with
// up to `threads` threads, each writing its range at its offset.
void write_file_parallel(const char* path, const void* src, uint64_t n, uint32_t threads);
// memcpy split over up to `threads` threads (16 MiB or more each).
void parallel_copy(void* dst, const void* src, uint64_t n, uint32_t threads);

// Phase 3 on the device (stitch_gpu.hip): the same unitigs, in the same
// order, as stitch_fragments.
Unitigs* stitch_fragments_gpu(int device, const uint8_t* buf, uint64_t n, uint32_t k, uint32_t threads);
// The same over the file at `path`, streamed to the device.
Unitigs* stitch_file_gpu(int device, const char* path, uint32_t k, uint32_t threads);

}  // namespace cdbg_amd
