// Shared host-side helpers of the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace mtlib {
// Records the calling thread's last error message and returns `code`.
int set_error(int code, const std::string& what);
// Per-device initialisation hooks of the other translation units, called by mt_init.
int rate_init_device(int device);
}  // namespace mtlib

#define MT_HIP_TRY(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return mtlib::set_error(MT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
