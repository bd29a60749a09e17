// sdp_internal.h -- shared helpers of the HIP translation units (not part
// of the public C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sdp_device.h"

namespace sdp {

// Records a message for sdp_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);

// Resolves a user stream (NULL means the library's default stream).
hipStream_t stream_of(sdp_stream s);

}  // namespace sdp

#define SDP_HIP_TRY(expr)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::sdp::set_error(SDP_E_HIP, "%s failed: %s (%s:%d)", #expr,            \
                                    hipGetErrorString(e_), __FILE__, __LINE__);           \
    } while (0)

#define SDP_LAUNCH_CHECK()                                                                \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            return ::sdp::set_error(SDP_E_HIP, "kernel launch failed: %s (%s:%d)",        \
                                    hipGetErrorString(e_), __FILE__, __LINE__);           \
    } while (0)

#define SDP_REQUIRE(cond, ...)                                                            \
    do {                                                                                  \
        if (!(cond)) return ::sdp::set_error(SDP_E_ARG, __VA_ARGS__);                     \
    } while (0)

static inline unsigned int sdp_div_up(size_t a, size_t b) { return (unsigned int)((a + b - 1) / b); }
