// rkg_internal.hpp -- error reporting shared by the translation units of
// librakali_amd.so (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

namespace rk {

// Records the message returned by rkg_last_error() on this thread.
__attribute__((visibility("hidden"), format(printf, 1, 2))) void set_error(const char* fmt, ...);

}  // namespace rk

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            rk::set_error("%s failed: %s", #expr, hipGetErrorString(e_));      \
            return RKG_ERR_HIP;                                                \
        }                                                                      \
    } while (0)
