// Phase 2 (bucket-process) on the device.  (Kernels follow in this file.)
#include <hip/hip_runtime.h>

#include "job.h"

namespace cdbg_amd {

struct Phase2Bufs {
    DevBuf frag_bytes;
    DevBuf cs_words, cs_offsets;
};

void run_phase2(Job& J) { throw Error("phase 2 not built yet"); }

void free_phase2(Job& J) {
    delete J.p2;
    J.p2 = nullptr;
}

}  // namespace cdbg_amd

extern "C" {
CDBG_API int cdbg_gpu_job_frags(cdbg_gpu_job* job, uint8_t* dst) {
    return cdbg_amd::set_error("phase 2 not built yet");
}
CDBG_API int cdbg_gpu_job_color_sets(cdbg_gpu_job* job, uint64_t* words, uint64_t* offsets) {
    return cdbg_amd::set_error("phase 2 not built yet");
}
}
