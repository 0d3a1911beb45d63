// Device primitives shared by the phase kernels: device buffers, exclusive
// scans, and a stable segmented LSD radix sort.
//
// Radix sort layout.  The input is a list of segments (buckets) laid out one
// after another.  Every pass splits each segment into tiles of RS_TILE keys
// that never cross a segment boundary; tile t of segment s writes its digit
// histogram at  base_s*R + digit*ntiles_s + local_t  (R = 2^bits), so one
// exclusive scan over the whole histogram array yields, for every
// (segment, digit, tile), the first output slot: segments stay in place and
// each is sorted on its own.  Ranks inside a tile are stable: the tile is
// consumed in rounds of 256 keys, each wave finds the lanes holding its digit
// with `bits` ballots, and per-wave LDS histograms order the waves.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "prof.h"

namespace cdbg_amd {

// ---------------------------------------------------------------------------
// Block cache.  hipFree waits for the whole device and hipMalloc maps fresh
// pages, so buffers released on a job stream are kept: a block goes back to
// the cache with an event recorded on its stream, and a later request takes
// the smallest cached block of at least its size (and at most twice it) that
// is on the same stream or whose event has completed.  When free HBM runs
// short, a block still in use on another stream is taken (the new stream
// waits for the old one's event), or one up to four times the size, before
// cached blocks are handed back to the driver (device-wide wait) and
// hipMalloc is called; when hipMalloc still
// fails the cache is emptied and it is tried again.  Process-wide and
// thread-safe.
// ---------------------------------------------------------------------------
struct BlockCache {
    struct Block {
        void* p;
        hipStream_t st;
        hipEvent_t ev;
    };
    std::mutex mu;
    std::multimap<size_t, Block> blocks;
    std::atomic<uint64_t> cached{0};
    // counters: requests served from the cache, hipMalloc calls and their
    // bytes and time, trims and the bytes they handed back
    std::atomic<uint64_t> n_hit{0}, n_malloc{0}, malloc_bytes{0}, malloc_ns{0}, n_trim{0}, trim_bytes{0}, trim_ns{0};
    static constexpr size_t RESERVE = size_t(2) << 30;

    std::vector<hipEvent_t> retired;  // events a stream was made to wait on; destroyed once complete

    // The smallest cached block of bytes .. max_bytes that is usable on
    // stream st now (released on st, or its event has completed), or with
    // `wait`, also one still in use on another stream: st then waits for
    // that stream's work up to the release (stream-ordered reuse).  Null if
    // none.
    void* take(size_t bytes, size_t max_bytes, hipStream_t st, size_t& got, bool wait = false) {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < retired.size();) {
            if (hipEventQuery(retired[i]) == hipSuccess) {
                (void)hipEventDestroy(retired[i]);
                retired[i] = retired.back();
                retired.pop_back();
            } else {
                ++i;
            }
        }
        for (auto it = blocks.lower_bound(bytes); it != blocks.end() && it->first <= max_bytes; ++it) {
            const Block& b = it->second;
            const bool ready = b.st == st || hipEventQuery(b.ev) == hipSuccess;
            if (ready || (wait && hipStreamWaitEvent(st, b.ev, 0) == hipSuccess)) {
                void* p = b.p;
                got = it->first;
                if (ready)
                    (void)hipEventDestroy(b.ev);
                else
                    retired.push_back(b.ev);
                cached -= got;
                blocks.erase(it);
                ++n_hit;
                return p;
            }
        }
        return nullptr;
    }

    // fault injection for tests: the request after this many more fails
    // as out of memory (-1: none)
    std::atomic<int64_t> inject_oom{-1};

    void* get(size_t bytes, hipStream_t st, size_t& got) {
        if (inject_oom.load() >= 0 && inject_oom.fetch_sub(1) == 0)
            throw Error("device allocation: out of memory (injected)");
        if (void* p = take(bytes, 2 * bytes, st, got)) return p;
        // Leave RESERVE bytes to the runtime: it allocates at dispatch time
        // (a queue's scratch for kernels with a private segment, kernel
        // arguments, signals), and a dispatch whose allocation fails aborts
        // its queue.  Short of that, a cached block still in use on another
        // stream (waited for), then one up to four times the size, is taken
        // before any is handed back: HBM freed to the driver is cleared when it is
        // allocated again, at tens of GB/s (~400 ms for 18 GB measured on
        // MI355X), where fresh or cached memory costs next to nothing.
        size_t free_b = 0, total_b = 0;
        if (cached.load() && hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b < bytes + RESERVE) {
            if (void* p = take(bytes, 2 * bytes, st, got, true)) return p;
            if (void* p = take(bytes, 4 * bytes, st, got, true)) return p;
            trim(bytes + RESERVE - free_b);
        }
        void* p = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim(2 * bytes);  // the largest blocks first, about what is asked and as much again
            if (hipMalloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                trim();
                if (hipMalloc(&p, bytes) != hipSuccess) {
                    (void)hipGetLastError();
                    size_t f = 0, t = 0;
                    (void)hipMemGetInfo(&f, &t);
                    throw Error("device allocation of " + std::to_string(bytes) + " bytes: out of memory (" +
                                std::to_string(g_device_bytes.load()) + " bytes held by device buffers, " +
                                std::to_string(f) + " free)");
                }
            }
        }
        ++n_malloc;
        malloc_bytes += bytes;
        malloc_ns += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count());
        got = bytes;
        return p;
    }
    void put(void* p, size_t size, hipStream_t st) {
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, st) != hipSuccess) {
            if (ev) (void)hipEventDestroy(ev);
            (void)hipStreamSynchronize(st);
            (void)hipFree(p);
            return;
        }
        std::lock_guard<std::mutex> lk(mu);
        blocks.emplace(size, Block{p, st, ev});
        cached += size;
    }
    // hands cached blocks back to the driver, the largest first, until at
    // least `want` bytes are released (all of them by default)
    void trim(uint64_t want = ~0ull) {
        std::lock_guard<std::mutex> lk(mu);
        if (blocks.empty()) return;
        const auto t0 = std::chrono::steady_clock::now();
        (void)hipDeviceSynchronize();
        for (hipEvent_t e : retired) (void)hipEventDestroy(e);
        retired.clear();
        uint64_t freed = 0;
        while (!blocks.empty() && freed < want) {
            auto it = std::prev(blocks.end());
            (void)hipEventDestroy(it->second.ev);
            (void)hipFree(it->second.p);
            freed += it->first;
            cached -= it->first;
            blocks.erase(it);
        }
        ++n_trim;
        trim_bytes += freed;
        trim_ns += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count());
    }
};
inline BlockCache g_block_cache;

// ---------------------------------------------------------------------------
// Bump allocator over a device range the caller owns.  While a thread has one
// set (ArenaScope), every DevBuf that grows takes its block from the range
// and nothing is allocated or freed; running past the end throws before any
// kernel sees the block.  Blocks are 256-byte aligned.
// ---------------------------------------------------------------------------
struct Arena {
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0;
    void* take(size_t bytes) {
        const size_t at = (used + 255) & ~size_t(255);
        if (at > cap || bytes > cap - at)
            throw Error("scratch exhausted: " + std::to_string(at + bytes) + " bytes needed, " +
                        std::to_string(cap) + " available");
        used = at + bytes;
        return base + at;
    }
};
inline thread_local Arena* t_arena = nullptr;

struct ArenaScope {
    Arena* prev;
    explicit ArenaScope(Arena* a) : prev(t_arena) { t_arena = a; }
    ~ArenaScope() { t_arena = prev; }
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;
};

// ---------------------------------------------------------------------------
// Device buffer that only grows (contents are not preserved on growth).
// With a t_alloc_stream set it comes from, and goes back to, the block
// cache; with an arena set it comes from the arena; otherwise hipMalloc /
// hipFree.  A buffer that borrows memory (an arena block, or a caller's
// buffer adopted with borrow()) frees nothing.
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipStream_t st = nullptr;  // stream the buffer is used on (null: not cached)
    bool borrowed = false;     // memory owned elsewhere
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && !borrowed) {
            if (st)
                g_block_cache.put(p, cap, st);
            else
                (void)hipFree(p);
            g_device_bytes -= cap;
        }
        p = nullptr;
        cap = 0;
        st = nullptr;
        borrowed = false;
    }
    // adopts `bytes` bytes at q that stay owned by the caller
    void borrow(void* q, size_t bytes) {
        release();
        p = q;
        cap = bytes;
        borrowed = true;
    }
    template <class T>
    T* ensure(size_t n) {
        size_t bytes = n * sizeof(T);
        if (bytes < 256) bytes = 256;
        if (bytes > cap) {
            release();
            if (t_arena) {
                p = t_arena->take(bytes);
                cap = bytes;
                borrowed = true;
                return (T*)p;
            }
            if (t_alloc_stream) {
                p = g_block_cache.get(bytes, t_alloc_stream, bytes);
            } else {
                CDBG_HIP(hipMalloc(&p, bytes));
            }
            st = t_alloc_stream;
            cap = bytes;
            g_device_bytes += cap;
        }
        return (T*)p;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(st, o.st);
        std::swap(borrowed, o.borrowed);
    }
    // Grows to hold n elements keeping the first `keep_bytes` bytes
    // (stream-ordered copy; waits for it before freeing the old block).
    template <class T>
    // Grows to at least n elements keeping the first keep_bytes; a growth
    // allocates at least 1.5x the old capacity and at least `hint` elements
    // (the caller's estimate of the final size, so a stream that grows chunk
    // by chunk is copied about once).
    T* ensure_keep(size_t n, size_t keep_bytes, hipStream_t st, size_t hint = 0) {
        size_t bytes = n * sizeof(T);
        if (bytes <= cap) return (T*)p;
        DevBuf nb;
        nb.ensure<uint8_t>(std::max(std::max(bytes, cap + cap / 2), hint * sizeof(T)));
        if (keep_bytes && p) CDBG_HIP(hipMemcpyAsync(nb.p, p, std::min(keep_bytes, cap), hipMemcpyDeviceToDevice, st));
        CDBG_HIP(hipStreamSynchronize(st));
        swap(nb);
        return (T*)p;
    }
};

// Pinned host buffer that grows keeping its contents (hipHostMalloc: device
// copies into it run at full link speed, without staging).
struct PinnedBuf {
    uint8_t* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    // at least n bytes, the first `keep` kept; a growth takes at least
    // `hint` bytes and 1.5x the old capacity
    uint8_t* ensure_keep(size_t n, size_t keep, size_t hint = 0) {
        if (n <= cap) return p;
        const size_t want = std::max(std::max(n, cap + cap / 2), hint);
        void* q = nullptr;
        CDBG_HIP(hipHostMalloc(&q, want, hipHostMallocDefault));
        if (keep && p) std::memcpy(q, p, std::min(keep, cap));
        release();
        p = (uint8_t*)q;
        cap = want;
        return p;
    }
};

struct Key128 {
    uint64_t lo, hi;
};

__host__ __device__ inline bool operator==(const Key128& a, const Key128& b) {
    return a.lo == b.lo && a.hi == b.hi;
}
__host__ __device__ inline bool operator!=(const Key128& a, const Key128& b) { return !(a == b); }
__host__ __device__ inline bool operator<(const Key128& a, const Key128& b) {
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}

__device__ __forceinline__ uint32_t digit_of(uint32_t k, int shift, uint32_t mask) {
    return (k >> shift) & mask;
}
__device__ __forceinline__ uint32_t digit_of(uint64_t k, int shift, uint32_t mask) {
    return uint32_t(k >> shift) & mask;
}
__device__ __forceinline__ uint32_t digit_of(const Key128& k, int shift, uint32_t mask) {
    if (shift >= 64) return uint32_t(k.hi >> (shift - 64)) & mask;
    uint64_t v = k.lo >> shift;
    if (shift > 0) v |= k.hi << (64 - shift);
    return uint32_t(v) & mask;
}

// ---------------------------------------------------------------------------
// Exclusive scan (reduce, scan of block sums, down-sweep).
// ---------------------------------------------------------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;

template <class T>
__device__ inline T block_exclusive_scan(T v, T* s_tmp, T& total) {
    // s_tmp holds >= blockDim/64 entries.  Wave-level inclusive scan first.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int w = 0; w < nw; ++w) {
            T t = s_tmp[w];
            s_tmp[w] = run;
            run += t;
        }
        s_tmp[nw] = run;
    }
    __syncthreads();
    T res = s_tmp[wave] + x - v;
    total = s_tmp[nw];
    __syncthreads();
    return res;
}

// Chunk sums; loads are striped across the block (coalesced).
template <class Tin, class Tout>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const Tin* __restrict__ in, uint64_t n,
                                                              Tout* __restrict__ sums) {
    __shared__ Tout s_tmp[SCAN_THREADS / 64 + 1];
    const uint64_t base = uint64_t(blockIdx.x) * SCAN_CHUNK;
    Tout acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint64_t idx = base + uint64_t(i) * SCAN_THREADS + threadIdx.x;
        if (idx < n) acc += Tout(in[idx]);
    }
    Tout total;
    (void)block_exclusive_scan<Tout>(acc, s_tmp, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <class T>
__global__ __launch_bounds__(1024) void k_scan_single(T* __restrict__ a, uint64_t n, T* __restrict__ total_out) {
    __shared__ T s_tmp[1024 / 64 + 1];
    T carry = 0;
    for (uint64_t base = 0; base < n; base += 1024) {
        uint64_t idx = base + threadIdx.x;
        T v = idx < n ? a[idx] : T(0);
        T tot;
        T ex = block_exclusive_scan<T>(v, s_tmp, tot);
        if (idx < n) a[idx] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Down-sweep: the chunk is staged in LDS with striped (coalesced) loads,
// each thread scans 16 consecutive items there, and the result leaves with
// striped stores.  One pad word per 32 keeps the per-thread runs on
// different banks.
__device__ __forceinline__ uint32_t scan_pad(uint32_t i) { return i + (i >> 5); }

template <class Tin, class Tout>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const Tin* __restrict__ in, uint64_t n,
                                                            const Tout* __restrict__ sums,
                                                            Tout* __restrict__ out) {
    __shared__ Tout s_tmp[SCAN_THREADS / 64 + 1];
    __shared__ Tout s_v[SCAN_CHUNK + SCAN_CHUNK / 32];
    const uint64_t base = uint64_t(blockIdx.x) * SCAN_CHUNK;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t j = uint32_t(i) * SCAN_THREADS + threadIdx.x;
        const uint64_t idx = base + j;
        s_v[scan_pad(j)] = idx < n ? Tout(in[idx]) : Tout(0);
    }
    __syncthreads();
    const uint32_t first = threadIdx.x * SCAN_ITEMS;
    Tout acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) acc += s_v[scan_pad(first + i)];
    Tout tot;
    Tout ex = block_exclusive_scan<Tout>(acc, s_tmp, tot) + sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t j = scan_pad(first + i);
        const Tout v = s_v[j];
        s_v[j] = ex;
        ex += v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t j = uint32_t(i) * SCAN_THREADS + threadIdx.x;
        const uint64_t idx = base + j;
        if (idx < n) out[idx] = s_v[scan_pad(j)];
    }
}

// out[i] = sum(in[0..i)); *d_total (device) = sum(in).  `tmp` is scratch.
template <class Tin, class Tout>
inline void exclusive_scan(const Tin* in, Tout* out, uint64_t n, Tout* d_total, DevBuf& tmp,
                           hipStream_t st) {
    uint64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nb == 0) nb = 1;
    Tout* sums = tmp.ensure<Tout>(nb + 1);
    { KSCOPE("k_scan_reduce", st); k_scan_reduce<Tin, Tout><<<dim3(uint32_t(nb)), SCAN_THREADS, 0, st>>>(in, n, sums); }
    { KSCOPE("k_scan_single", st); k_scan_single<Tout><<<1, 1024, 0, st>>>(sums, nb, d_total); }
    { KSCOPE("k_scan_down", st); k_scan_down<Tin, Tout><<<dim3(uint32_t(nb)), SCAN_THREADS, 0, st>>>(in, n, sums, out); }
    CDBG_HIP(hipGetLastError());
}
// ---------------------------------------------------------------------------
// Stream compaction: list[j] = index of the j-th nonzero flag, in order.
// Flags are counted per 4096-flag tile, the tile counts scanned, and each
// tile writes its indices at its base: the flags are read twice and nothing
// of size n is written (a scan of the flags would write an n-entry array).
// ---------------------------------------------------------------------------
constexpr int CF_ITEMS = SCAN_ITEMS;
constexpr int CF_TILE = SCAN_THREADS * CF_ITEMS;

// nonzero bytes of a word
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t w) {
    return uint32_t(__popc((((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u));
}

// Byte flags are read 16 per lane (one 16-byte load), element i of the tile
// at lane i / 16; wider flags one per lane per round.
template <class F>
__global__ __launch_bounds__(SCAN_THREADS) void k_flag_count(const F* __restrict__ flags, uint64_t n,
                                                             uint32_t* __restrict__ cnt) {
    __shared__ uint32_t s_tmp[SCAN_THREADS / 64 + 1];
    const uint64_t T0 = uint64_t(blockIdx.x) * CF_TILE;
    uint32_t c = 0;
    if constexpr (sizeof(F) == 1) {
        const uint64_t i0 = T0 + uint64_t(threadIdx.x) * CF_ITEMS;
        if (i0 + CF_ITEMS <= n) {
            const uint4 w = *reinterpret_cast<const uint4*>(flags + i0);
            c = nonzero_bytes(w.x) + nonzero_bytes(w.y) + nonzero_bytes(w.z) + nonzero_bytes(w.w);
        } else {
            for (uint64_t i = i0; i < n; ++i) c += flags[i] != 0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < CF_ITEMS; ++r) {
            const uint64_t i = T0 + uint64_t(r) * SCAN_THREADS + threadIdx.x;
            c += i < n && flags[i] != 0;
        }
    }
    uint32_t tot;
    (void)block_exclusive_scan<uint32_t>(c, s_tmp, tot);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// A flag's rank in its tile is the count of flags before it: byte flags by a
// block scan of the lanes' 16-flag counts and the flags before it in its
// lane; wider flags by rounds (round r = flags T0 + 256 r .. T0 + 256 r +
// 255): the rounds before it, the waves before it in its round and the lanes
// before it in its wave.
template <class F>
__global__ __launch_bounds__(SCAN_THREADS) void k_flag_compact(const F* __restrict__ flags, uint64_t n,
                                                               const uint32_t* __restrict__ base,
                                                               uint32_t* __restrict__ list) {
    constexpr int WAVES = SCAN_THREADS / 64;
    const uint64_t T0 = uint64_t(blockIdx.x) * CF_TILE;
    const uint32_t b = base[blockIdx.x];
    if constexpr (sizeof(F) == 1) {
        __shared__ uint32_t s_tmp[WAVES + 1];
        const uint64_t i0 = T0 + uint64_t(threadIdx.x) * CF_ITEMS;
        uint32_t bits = 0;  // bit j: flag i0 + j is set
        if (i0 + CF_ITEMS <= n) {
            const uint4 w = *reinterpret_cast<const uint4*>(flags + i0);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int j = 0; j < 4; ++j) bits |= uint32_t(((ws[q] >> (8 * j)) & 0xFFu) != 0) << (4 * q + j);
        } else {
            for (uint64_t i = i0; i < n; ++i) bits |= uint32_t(flags[i] != 0) << (i - i0);
        }
        uint32_t tot;
        uint32_t at = b + block_exclusive_scan<uint32_t>(uint32_t(__popc(bits)), s_tmp, tot);
        while (bits) {
            const int j = __ffs(bits) - 1;
            list[at++] = uint32_t(i0 + j);
            bits &= bits - 1;
        }
    } else {
        __shared__ uint32_t s_cnt[CF_ITEMS * WAVES];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        uint64_t m[CF_ITEMS];
#pragma unroll
        for (int r = 0; r < CF_ITEMS; ++r) {
            const uint64_t i = T0 + uint64_t(r) * SCAN_THREADS + threadIdx.x;
            m[r] = __ballot(i < n && flags[i] != 0);
            if (lane == 0) s_cnt[r * WAVES + wave] = uint32_t(__popcll(m[r]));
        }
        __syncthreads();
        if (threadIdx.x < 64) {  // exclusive scan of the CF_ITEMS * WAVES counts (<= 64) by one wave
            const uint32_t v = threadIdx.x < CF_ITEMS * WAVES ? s_cnt[threadIdx.x] : 0u;
            uint32_t x = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (threadIdx.x < CF_ITEMS * WAVES) s_cnt[threadIdx.x] = x - v;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CF_ITEMS; ++r) {
            if ((m[r] >> lane) & 1ull) {
                const uint64_t i = T0 + uint64_t(r) * SCAN_THREADS + threadIdx.x;
                list[b + s_cnt[r * WAVES + wave] + uint32_t(__popcll(m[r] & lt))] = uint32_t(i);
            }
        }
    }
}

// Compacts the nonzero flags of [0, n) into list; *d_count (device) = their
// number.  `cnt` is scratch of 2 (n / 4096 + 2) words.  Byte flags must
// start 16-byte aligned (device buffers are).
template <class F>
inline void compact_flags(const F* flags, uint64_t n, uint32_t* list, uint32_t* d_count, DevBuf& cnt, DevBuf& tmp,
                          hipStream_t st) {
    static_assert(CF_ITEMS * (SCAN_THREADS / 64) <= 64, "one wave scans a tile's round counts");
    uint64_t nb = (n + CF_TILE - 1) / CF_TILE;
    if (nb == 0) nb = 1;
    uint32_t* c = cnt.ensure<uint32_t>(2 * (nb + 1));
    uint32_t* b = c + nb + 1;
    { KSCOPE("k_flag_count", st); k_flag_count<F><<<dim3(uint32_t(nb)), SCAN_THREADS, 0, st>>>(flags, n, c); }
    exclusive_scan<uint32_t, uint32_t>(c, b, nb, d_count, tmp, st);
    { KSCOPE("k_flag_compact", st); k_flag_compact<F><<<dim3(uint32_t(nb)), SCAN_THREADS, 0, st>>>(flags, n, b, list); }
    CDBG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Segmented stable LSD radix sort.
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 512;
constexpr int RS_TILE = 4096;

struct TileDesc {
    uint64_t start;       // first element of the tile
    uint32_t count;       // elements in the tile
    uint32_t ntiles_seg;  // tiles of the tile's segment
    uint64_t hist_base;   // first tile index of the segment (times R gives the region)
    uint32_t local;       // tile index inside the segment
    uint32_t pad;
};

// Workgroups are dispatched to the 8 XCDs round-robin; tile_of() gives XCD x
// the x-th eighth of the tiles, so the tiles of one segment (whose outputs
// interleave) share one L2.  Grids are launched with 8 * ceil(nt / 8) blocks.
constexpr uint32_t N_XCD = 8;
__device__ __forceinline__ uint32_t tile_of(uint32_t block, uint32_t nt) {
    const uint32_t per = (nt + N_XCD - 1) / N_XCD;
    return (block % N_XCD) * per + block / N_XCD;
}
inline uint32_t xcd_grid(uint32_t nt) { return ((nt + N_XCD - 1) / N_XCD) * N_XCD; }

template <class K>
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const K* __restrict__ keys,
                                                           const TileDesc* __restrict__ tiles, uint32_t nt,
                                                           int shift, int bits, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t s_hist[];
    const uint32_t tile = tile_of(blockIdx.x, nt);
    if (tile >= nt) return;
    const uint32_t R = 1u << bits, mask = R - 1;
    const TileDesc d = tiles[tile];
    for (uint32_t i = threadIdx.x; i < R; i += RS_THREADS) s_hist[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < d.count; i += RS_THREADS)
        atomicAdd(&s_hist[digit_of(keys[d.start + i], shift, mask)], 1u);
    __syncthreads();
    const uint64_t region = d.hist_base * R;
    for (uint32_t b = threadIdx.x; b < R; b += RS_THREADS)
        hist[region + uint64_t(b) * d.ntiles_seg + d.local] = s_hist[b];
}

// Stable scatter of one tile.  Wave w owns keys [1024w, 1024w + 1024) of
// the tile and ranks them alone, 64 at a time: `bits` ballots give each
// lane its peers (same digit) in the round, and a per-wave LDS count per
// digit carries the rank across rounds.  With the per-wave counts known, the
// tile is placed in LDS in digit order and written out digit run by digit
// run, so consecutive lanes store consecutive addresses.  Four barriers per
// tile.
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ITEMS = RS_TILE / RS_THREADS;  // rounds per wave

template <class K, class V>
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(const K* __restrict__ keys,
                                                              const V* __restrict__ vals,
                                                              K* __restrict__ okeys, V* __restrict__ ovals,
                                                              const TileDesc* __restrict__ tiles, uint32_t nt,
                                                              int shift, int bits, const uint32_t* __restrict__ hist,
                                                              const uint32_t* __restrict__ offs) {
    extern __shared__ uint64_t sm64[];
    const uint32_t tile = tile_of(blockIdx.x, nt);
    if (tile >= nt) return;
    const uint32_t R = 1u << bits, mask = R - 1;
    K* s_key = (K*)sm64;
    V* s_val = (V*)(s_key + RS_TILE);
    uint32_t* dstart = (uint32_t*)(s_val + RS_TILE);  // [R] tile-local first slot of each digit
    uint32_t* gbase = dstart + R;                      // [R] global first slot of each digit
    uint32_t* wcnt = gbase + R;                        // [RS_WAVES][R] per-wave counts, then bases
    uint32_t* s_tmp = wcnt + RS_WAVES * R;             // scan scratch
    const TileDesc d = tiles[tile];
    const uint64_t region = d.hist_base * R;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t b = threadIdx.x; b < RS_WAVES * R; b += RS_THREADS) wcnt[b] = 0;
    __syncthreads();

    // rank: digit and wave-local rank of every item, kept in registers.
    // Values wider than 8 bytes are read when they are placed (held through
    // the ranking they would not fit the registers and spill to scratch).
    constexpr bool LATE = sizeof(V) > 8;
    K key[RS_ITEMS];
    V val[LATE ? 1 : RS_ITEMS];
    uint32_t dr[RS_ITEMS];  // digit << 16 | wave-local rank
    uint32_t* my = wcnt + wave * R;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint32_t i = uint32_t(wave) * (RS_TILE / RS_WAVES) + uint32_t(r) * 64 + uint32_t(lane);
        const bool active = i < d.count;
        uint32_t dig = 0;
        if (active) {
            key[r] = keys[d.start + i];
            if constexpr (!LATE) val[r] = vals[d.start + i];
            dig = digit_of(key[r], shift, mask);
        }
        uint64_t peers = __ballot(active);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rk = __popcll(peers & lt);
        uint32_t before = 0;
        if (active) before = my[dig];
        // leaders write after every lane has read (the write uses the read)
        if (active && rk == 0) my[dig] = before + __popcll(peers);
        dr[r] = (dig << 16) | (before + rk);
    }
    __syncthreads();

    // digit starts of the tile and of every wave inside it
    {
        const uint32_t per = R >= RS_THREADS ? R / RS_THREADS : 1u;  // <= 8 for R <= 2048
        const uint32_t b0 = threadIdx.x * per;
        uint32_t c[8];
        uint32_t acc = 0;
        for (uint32_t j = 0; j < per; ++j) {
            c[j] = (b0 + j < R) ? hist[region + uint64_t(b0 + j) * d.ntiles_seg + d.local] : 0u;
            acc += c[j];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan<uint32_t>(acc, s_tmp, tot);
        for (uint32_t j = 0; j < per; ++j) {
            const uint32_t b = b0 + j;
            if (b < R) {
                dstart[b] = ex;
                gbase[b] = offs[region + uint64_t(b) * d.ntiles_seg + d.local];
                uint32_t run = ex;
                for (int w = 0; w < RS_WAVES; ++w) {
                    const uint32_t cw = wcnt[w * R + b];
                    wcnt[w * R + b] = run;
                    run += cw;
                }
            }
            ex += c[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint32_t i = uint32_t(wave) * (RS_TILE / RS_WAVES) + uint32_t(r) * 64 + uint32_t(lane);
        if (i < d.count) {
            const uint32_t pos = my[dr[r] >> 16] + (dr[r] & 0xFFFFu);
            s_key[pos] = key[r];
            if constexpr (LATE)
                s_val[pos] = vals[d.start + i];
            else
                s_val[pos] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < d.count; j += RS_THREADS) {
        const K k = s_key[j];
        const uint32_t dig = digit_of(k, shift, mask);
        const uint32_t pos = gbase[dig] + (j - dstart[dig]);
        okeys[pos] = k;
        ovals[pos] = s_val[j];
    }
}

template <class K, class V>
inline size_t radix_scatter_lds(int bits) {
    const size_t R = size_t(1) << bits;
    return RS_TILE * (sizeof(K) + sizeof(V)) + ((2 + RS_WAVES) * R + RS_WAVES + 1) * sizeof(uint32_t);
}

// Lets k_radix_scatter<K, V> use up to the 160 KiB of LDS of a gfx950
// workgroup (set once per instantiation, thread-safe: jobs of a server run
// on several host threads); a pass of `bits`-bit digits must fit it.
template <class K, class V>
inline void radix_scatter_prepare(int bits) {
    constexpr size_t LDS_MAX = 160 * 1024;
    if (radix_scatter_lds<K, V>(bits) > LDS_MAX) throw Error("segmented_radix_sort: digit width too large for LDS");
    static const bool done = [] {
        CDBG_HIP(hipFuncSetAttribute((const void*)&k_radix_scatter<K, V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(LDS_MAX)));
        return true;
    }();
    (void)done;
}

// Host-side tile plan for a segmented sort.  seg_off has nseg+1 entries.
struct SortPlan {
    std::vector<TileDesc> tiles;
    uint64_t n = 0;
    void build(const std::vector<uint64_t>& seg_off) {
        tiles.clear();
        n = seg_off.empty() ? 0 : seg_off.back();
        uint64_t tile_base = 0;
        for (size_t s = 0; s + 1 < seg_off.size(); ++s) {
            uint64_t len = seg_off[s + 1] - seg_off[s];
            uint32_t nt = uint32_t((len + RS_TILE - 1) / RS_TILE);
            for (uint32_t t = 0; t < nt; ++t) {
                TileDesc d{};
                d.start = seg_off[s] + uint64_t(t) * RS_TILE;
                uint64_t rem = len - uint64_t(t) * RS_TILE;
                d.count = uint32_t(rem < RS_TILE ? rem : RS_TILE);
                d.ntiles_seg = nt;
                d.hist_base = tile_base;
                d.local = t;
                tiles.push_back(d);
            }
            tile_base += nt;
        }
    }
};

// Sorts (keys, vals) in place by key bits [begin_bit, end_bit) inside every
// segment of `plan`, `digit_bits` bits per pass.  kt/vt are ping-pong
// buffers of the same length; the result always ends in keys/vals.
struct RadixScratch {
    DevBuf tiles, hist, offs, scan_tmp, total;
};

template <class K, class V>
// The result is in keys/vals; with in_temp non-null an odd number of passes
// leaves it in kt/vt instead (*in_temp = true) and saves the copy back.
inline void segmented_radix_sort(K* keys, V* vals, K* kt, V* vt, const SortPlan& plan, int begin_bit,
                                 int end_bit, int digit_bits, RadixScratch& S, hipStream_t st,
                                 const char* tag = "", bool* in_temp = nullptr) {
    if (in_temp) *in_temp = false;
    if (plan.n == 0 || begin_bit >= end_bit) return;
    if (plan.n >= (uint64_t(1) << 32)) throw Error("segmented_radix_sort: more than 2^32 elements");
    if (digit_bits < 1 || digit_bits > 11) throw Error("segmented_radix_sort: digit width must be 1..11 bits");
    const uint32_t nt = uint32_t(plan.tiles.size());
    TileDesc* d_tiles = S.tiles.ensure<TileDesc>(nt);
    // The previous sort may still read d_tiles: order the upload on the stream
    // and wait for it, since plan.tiles is pageable host memory.
    CDBG_HIP(hipMemcpyAsync(d_tiles, plan.tiles.data(), nt * sizeof(TileDesc), hipMemcpyHostToDevice, st));
    CDBG_HIP(hipStreamSynchronize(st));
    K* src_k = keys;
    V* src_v = vals;
    K* dst_k = kt;
    V* dst_v = vt;
    int passes = 0;
    for (int sh = begin_bit; sh < end_bit; sh += digit_bits) {
        const int bits = (end_bit - sh) < digit_bits ? (end_bit - sh) : digit_bits;
        const uint32_t R = 1u << bits;
        radix_scatter_prepare<K, V>(bits);
        const uint64_t hl = uint64_t(nt) * R;
        uint32_t* hist = S.hist.ensure<uint32_t>(hl);
        uint32_t* offs = S.offs.ensure<uint32_t>(hl);
        { KSCOPE(std::string("k_radix_hist:") + tag, st); k_radix_hist<K><<<xcd_grid(nt), RS_THREADS, R * sizeof(uint32_t), st>>>(src_k, d_tiles, nt, sh, bits, hist); }
        exclusive_scan<uint32_t, uint32_t>(hist, offs, hl, S.total.ensure<uint32_t>(1), S.scan_tmp, st);
        { KSCOPE(std::string("k_radix_scatter:") + tag, st); k_radix_scatter<K, V><<<xcd_grid(nt), RS_THREADS, radix_scatter_lds<K, V>(bits), st>>>(
              src_k, src_v, dst_k, dst_v, d_tiles, nt, sh, bits, hist, offs); }
        CDBG_HIP(hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
        ++passes;
    }
    if ((passes & 1) && in_temp) {
        *in_temp = true;
    } else if (passes & 1) {
        CDBG_HIP(hipMemcpyAsync(keys, src_k, plan.n * sizeof(K), hipMemcpyDeviceToDevice, st));
        CDBG_HIP(hipMemcpyAsync(vals, src_v, plan.n * sizeof(V), hipMemcpyDeviceToDevice, st));
    }
}

}  // namespace cdbg_amd
