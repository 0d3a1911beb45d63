// Phase 1 into caller-owned device buffers without allocating device memory.
//
// Callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc) pass the input sequences and the
// output arrays; every intermediate buffer of phase 1 (the padded input copy,
// 2-bit bases, ACGT bitmap, super-k-mer staging, records and their sort
// buffers, scan partials) is carved out of `scratch`, a file-scope device
// array of NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the
// process.  The kernels are those of cdbg_gpu_job_phase1.
//
// In host code `scratch` names a host-side stand-in, not the device array:
// its device address comes from hipGetSymbolAddress, once per device, with
// that device current, and only that address is handed to kernels and copies.
// Calls in one process are serialised by a lock (they share the array); each
// call returns after its stream has finished the work, so nothing of one call
// is in flight when the next starts.  Nothing in the array is assumed zero:
// phase 1 clears the one counter it accumulates into.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "job.h"

namespace cdbg_amd {

int set_error(const std::string& msg);

namespace {

static __device__ __align__(256) uint8_t scratch[CDBG_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;                     // one call at a time: the calls share each device's array
std::vector<void*> g_scratch_addr;   // device address of `scratch` per device ordinal (null: not yet set up)

// Device address of this device's scratch array; the device must be current.
void* scratch_on(int device) {
    if (device < 0) throw Error("bad device ordinal");
    if (size_t(device) >= g_scratch_addr.size()) g_scratch_addr.resize(size_t(device) + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        CDBG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_addr[device] = p;
    }
    return g_scratch_addr[device];
}

__global__ void k_fill_bytes(uint8_t* p, uint64_t n, uint8_t v) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace
}  // namespace cdbg_amd

using namespace cdbg_amd;

extern "C" {

CDBG_API int cdbg_gpu_phase1_noalloc(cdbg_gpu_ctx* ctx, const uint8_t* d_bytes, const uint64_t* color_off,
                                     uint32_t num_colors, uint32_t k, uint32_t m, uint32_t num_buckets,
                                     uint64_t max_supers, uint64_t* d_start, uint32_t* d_len, uint8_t* d_flags,
                                     uint32_t* d_bucket, uint32_t* d_color, uint64_t* d_bucket_off,
                                     uint64_t* n_supers, void* stream) {
    try {
        if (!ctx || !d_bytes || !color_off || !n_supers || !d_bucket_off) throw Error("null argument");
        if (max_supers && (!d_start || !d_len || !d_flags || !d_bucket || !d_color)) throw Error("null output array");
        if (k < 2 || k > 63) throw Error("k must satisfy 2 <= k <= 63");
        if (m < 2 || m >= k) throw Error("m must satisfy 2 <= m < k");
        if (num_buckets == 0 || num_buckets > 65536) throw Error("num_buckets must be in [1, 65536]");
        if (num_colors == 0) throw Error("no colours");
        for (uint32_t c = 0; c < num_colors; ++c)
            if (color_off[c + 1] < color_off[c]) throw Error("colour offsets must be non-decreasing");
        const uint64_t n_bytes = color_off[num_colors] - color_off[0];
        if (n_bytes > CDBG_NOALLOC_MAX_BASES)
            throw Error("input of " + std::to_string(n_bytes) + " bytes exceeds the allocation-free limit of " +
                        std::to_string(CDBG_NOALLOC_MAX_BASES));
        std::lock_guard<std::mutex> lk(g_mu);
        CDBG_HIP(hipSetDevice(ctx->device));
        Arena A;
        A.base = (uint8_t*)scratch_on(ctx->device);
        A.cap = CDBG_NOALLOC_SCRATCH_BYTES;
        ArenaScope in_scratch(&A);

        Job J;  // its buffers below are arena blocks or the caller's arrays: none is allocated or freed
        J.ctx = ctx;
        J.st = (hipStream_t)stream;
        J.k = k;
        J.m = m;
        J.B = num_buckets;
        J.num_colors = num_colors;
        J.n_bytes = n_bytes;
        J.n_pad = ((n_bytes + 64) / 64) * 64 + 64;  // as cdbg_gpu_job_create pads
        J.p1_chunk_bases = ~0ull;                   // one pass over the input
        hipStream_t st = J.st;
        uint8_t* d = J.bytes.ensure<uint8_t>(J.n_pad);
        if (n_bytes) CDBG_HIP(hipMemcpyAsync(d, d_bytes + color_off[0], n_bytes, hipMemcpyDeviceToDevice, st));
        k_fill_bytes<<<uint32_t((J.n_pad - n_bytes + 255) / 256), 256, 0, st>>>(d + n_bytes, J.n_pad - n_bytes,
                                                                              uint8_t('\n'));
        std::vector<uint64_t> off(num_colors + 1);
        for (uint32_t c = 0; c <= num_colors; ++c) off[c] = color_off[c] - color_off[0];
        uint64_t* doff = J.color_off.ensure<uint64_t>(num_colors + 1);
        CDBG_HIP(hipMemcpyAsync(doff, off.data(), off.size() * 8, hipMemcpyHostToDevice, st));
        Phase1Bufs& P = J.p1;
        if (max_supers) {
            P.b_start.borrow(d_start, max_supers * 8);
            P.b_len.borrow(d_len, max_supers * 4);
            P.b_flags.borrow(d_flags, max_supers);
            P.b_color.borrow(d_color, max_supers * 4);
        }
        P.bucket_off.borrow(d_bucket_off, (uint64_t(num_buckets) + 1) * 8);
        phase1_core(J);  // synchronises the stream before it returns
        const uint64_t n = J.n_supers;
        if (n > max_supers)
            throw Error("phase 1 found " + std::to_string(n) + " super-k-mers, more than max_supers = " +
                        std::to_string(max_supers));
        // the arrays a small count moved into the arena (a buffer holds at
        // least 256 bytes), and the bucket of every record, go to the caller
        auto out = [&](void* dst, const DevBuf& b, size_t bytes) {
            if (bytes && b.p != dst) CDBG_HIP(hipMemcpyAsync(dst, b.p, bytes, hipMemcpyDeviceToDevice, st));
        };
        out(d_start, P.b_start, n * 8);
        out(d_len, P.b_len, n * 4);
        out(d_flags, P.b_flags, n);
        out(d_color, P.b_color, n * 4);
        out(d_bucket, P.key, n * 4);
        out(d_bucket_off, P.bucket_off, (uint64_t(num_buckets) + 1) * 8);
        CDBG_HIP(hipGetLastError());
        CDBG_HIP(hipStreamSynchronize(st));
        *n_supers = n;
        return 0;
    } catch (std::exception& e) {
        return set_error(e.what());
    }
}

}  // extern "C"
