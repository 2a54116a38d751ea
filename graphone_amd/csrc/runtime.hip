// Status codes, device memory, streams, events and fill kernels of the C-ABI.
// Replaces the reference's archi::mem wrappers (include/GraphOne/archi/mem/cuda.cuh:11-59) and
// archi::fill (include/GraphOne/archi/thrust/functions.h:28-31).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "internal.hpp"

namespace g1 {

namespace {
thread_local char t_last_error[1024] = "";
}

g1_status set_error(g1_status status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_last_error, sizeof(t_last_error), fmt, ap);
    va_end(ap);
    return status;
}

void clear_error() { t_last_error[0] = '\0'; }

int device_cu_count() {
    static int cus = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    });
    return cus;
}

template <typename T>
__global__ void __launch_bounds__(256) fill_kernel(T* __restrict__ dst, T value, size_t count) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = value;
}

template <typename T>
g1_status fill(T* dst, T value, size_t count, g1_stream_t stream) {
    if (count == 0) return G1_OK;
    G1_CHECK(dst != nullptr, G1_ERR_INVALID_ARG, "fill: null destination");
    size_t blocks = (count + 255) / 256;
    const size_t cap = static_cast<size_t>(device_cu_count()) * 16;
    if (blocks > cap) blocks = cap;
    fill_kernel<T><<<dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, as_stream(stream)>>>(dst, value, count);
    G1_LAUNCHED();
    return G1_OK;
}

}  // namespace g1

extern "C" {

const char* g1_status_string(g1_status status) {
    switch (status) {
        case G1_OK: return "ok";
        case G1_ERR_INVALID_ARG: return "invalid argument";
        case G1_ERR_HIP: return "HIP runtime error";
        case G1_ERR_OUT_OF_MEMORY: return "out of device memory";
        case G1_ERR_BUDGET: return "HBM budget exceeded";
        case G1_ERR_NOT_FOUND: return "not found";
        case G1_ERR_BUSY: return "busy";
        case G1_ERR_CANCELLED: return "cancelled";
        case G1_ERR_INTERNAL: return "internal error";
        case G1_ERR_NOT_CONVERGED: return "not converged";
        default: return "unknown status";
    }
}

const char* g1_last_error_message(void) { return g1::t_last_error; }

int32_t g1_abi_version(void) { return 1; }

g1_status g1_device_info(int32_t* device_count, int32_t* cu_count, uint64_t* hbm_bytes, char* name,
                         size_t name_len) {
    int count = 0;
    G1_HIP(hipGetDeviceCount(&count));
    if (device_count) *device_count = count;
    if (count == 0) return g1::set_error(G1_ERR_NOT_FOUND, "no HIP device");
    int dev = 0;
    G1_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    G1_HIP(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    if (name && name_len) {
        snprintf(name, name_len, "%s", prop.gcnArchName);
    }
    return G1_OK;
}

g1_status g1_malloc(void** ptr, size_t bytes) {
    G1_CHECK(ptr != nullptr, G1_ERR_INVALID_ARG, "g1_malloc: null out pointer");
    *ptr = nullptr;
    if (bytes == 0) return G1_OK;
    G1_HIP(hipMalloc(ptr, bytes));
    return G1_OK;
}

g1_status g1_free(void* ptr) {
    if (ptr) G1_HIP(hipFree(ptr));
    return G1_OK;
}

g1_status g1_memcpy_h2d(void* dst, const void* src, size_t bytes, g1_stream_t stream) {
    if (bytes == 0) return G1_OK;
    G1_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g1::as_stream(stream)));
    G1_HIP(hipStreamSynchronize(g1::as_stream(stream)));
    return G1_OK;
}

g1_status g1_memcpy_d2h(void* dst, const void* src, size_t bytes, g1_stream_t stream) {
    if (bytes == 0) return G1_OK;
    G1_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g1::as_stream(stream)));
    G1_HIP(hipStreamSynchronize(g1::as_stream(stream)));
    return G1_OK;
}

g1_status g1_memcpy_d2d(void* dst, const void* src, size_t bytes, g1_stream_t stream) {
    if (bytes == 0) return G1_OK;
    G1_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g1::as_stream(stream)));
    return G1_OK;
}

g1_status g1_fill_u32(uint32_t* dst, uint32_t value, size_t count, g1_stream_t stream) {
    return g1::fill<uint32_t>(dst, value, count, stream);
}

g1_status g1_fill_f32(float* dst, float value, size_t count, g1_stream_t stream) {
    return g1::fill<float>(dst, value, count, stream);
}

g1_status g1_stream_create(g1_stream_t* stream) {
    G1_CHECK(stream != nullptr, G1_ERR_INVALID_ARG, "g1_stream_create: null out pointer");
    hipStream_t s = nullptr;
    G1_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<g1_stream_t>(s);
    return G1_OK;
}

g1_status g1_stream_destroy(g1_stream_t stream) {
    if (stream) G1_HIP(hipStreamDestroy(g1::as_stream(stream)));
    return G1_OK;
}

g1_status g1_stream_synchronize(g1_stream_t stream) {
    G1_HIP(hipStreamSynchronize(g1::as_stream(stream)));
    return G1_OK;
}

g1_status g1_event_create(g1_event_t* event) {
    G1_CHECK(event != nullptr, G1_ERR_INVALID_ARG, "g1_event_create: null out pointer");
    hipEvent_t e = nullptr;
    G1_HIP(hipEventCreate(&e));
    *event = reinterpret_cast<g1_event_t>(e);
    return G1_OK;
}

g1_status g1_event_destroy(g1_event_t event) {
    if (event) G1_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
    return G1_OK;
}

g1_status g1_event_record(g1_event_t event, g1_stream_t stream) {
    G1_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), g1::as_stream(stream)));
    return G1_OK;
}

g1_status g1_event_synchronize(g1_event_t event) {
    G1_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
    return G1_OK;
}

g1_status g1_event_elapsed_ms(float* ms, g1_event_t start, g1_event_t stop) {
    G1_CHECK(ms != nullptr, G1_ERR_INVALID_ARG, "g1_event_elapsed_ms: null out pointer");
    G1_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    return G1_OK;
}

}  // extern "C"
