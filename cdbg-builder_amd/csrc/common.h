// Shared host-side helpers: thread-local error reporting and HIP checks.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cdbg_amd {

// Records `msg` as this thread's last error; returns the generic failure code.
int set_error(const std::string& msg);
const char* last_error();

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace cdbg_amd

#define CDBG_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (void)hipGetLastError(); /* reset: a later launch check must not see it */     \
            throw ::cdbg_amd::Error(std::string(#call) + ": " + hipGetErrorString(e_) + " (" + \
                                    __FILE__ + ":" + std::to_string(__LINE__) + ")");      \
        }                                                                                   \
    } while (0)

// Body of a C-ABI entry point: 0 on success, else the exception's message
// becomes the thread's last error and the generic failure code is returned.
#define CDBG_TRY(...)                              \
    try {                                          \
        __VA_ARGS__;                               \
        return 0;                                  \
    } catch (std::exception & e) {                 \
        return ::cdbg_amd::set_error(e.what());    \
    } catch (...) {                                \
        return ::cdbg_amd::set_error("unknown C++ exception"); \
    }
