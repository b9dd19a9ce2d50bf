// Shared plumbing for libvipe_hip: status codes, last-error text, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/vipe_hip.h"

namespace vh {

// Thread-local text of the last failure, returned by vh_last_error().
inline char* last_error_buf() {
    static thread_local char buf[512] = {0};
    return buf;
}

inline int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(last_error_buf(), 512, fmt, ap);
    va_end(ap);
    return code;
}

inline int hip_fail(hipError_t e, const char* where) {
    return fail(VH_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

#define VH_HIP(call)                                      \
    do {                                                  \
        hipError_t _e = (call);                           \
        if (_e != hipSuccess) return vh::hip_fail(_e, #call); \
    } while (0)

#define VH_LAUNCH_CHECK(name)                                 \
    do {                                                      \
        hipError_t _e = hipGetLastError();                    \
        if (_e != hipSuccess) return vh::hip_fail(_e, name);  \
    } while (0)

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b on XCD b % 8),
// and each XCD has its own L2.  xcd_logical renumbers workgroup b of nb so
// that XCD x runs the x-th contiguous range of logical blocks: neighbouring
// blocks, which share input rows, then share one L2.
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t nb) {
    const int64_t xcd = b % kXcds, local = b / kXcds;
    const int64_t per = (nb + kXcds - 1) / kXcds, r = nb % kXcds;
    if (r == 0 || xcd < r) return xcd * per + local;
    return r * per + (xcd - r) * (per - 1) + local;
}

template <typename T>
struct dtype_of;
template <>
struct dtype_of<float> {
    static constexpr int value = VH_F32;
};
template <>
struct dtype_of<double> {
    static constexpr int value = VH_F64;
};

}  // namespace vh
