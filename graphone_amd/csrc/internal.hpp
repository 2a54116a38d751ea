// Internal helpers of libgraphone_hip.so: status plumbing and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include <string>

#include "graphone_hip.h"

namespace g1 {

// Records a thread-local message for g1_last_error_message() and returns `status`.
g1_status set_error(g1_status status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline hipStream_t as_stream(g1_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

int device_cu_count();  // cached per process (device 0 of the process)

}  // namespace g1

#define G1_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t g1_e_ = (call);                                                                 \
        if (g1_e_ != hipSuccess)                                                                   \
            return ::g1::set_error(g1_e_ == hipErrorOutOfMemory ? G1_ERR_OUT_OF_MEMORY : G1_ERR_HIP, \
                                   "%s failed: %s (%s:%d)", #call, hipGetErrorString(g1_e_),       \
                                   __FILE__, __LINE__);                                            \
    } while (0)

#define G1_LAUNCHED() G1_HIP(hipGetLastError())

#define G1_CHECK(cond, status, ...)                                                                \
    do {                                                                                           \
        if (!(cond)) return ::g1::set_error((status), __VA_ARGS__);                                \
    } while (0)

#define G1_TRY(expr)                                                                               \
    do {                                                                                           \
        g1_status g1_s_ = (expr);                                                                  \
        if (g1_s_ != G1_OK) return g1_s_;                                                          \
    } while (0)
