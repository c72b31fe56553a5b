// Error state and the device-memory / stream / event plumbing of the C ABI.
#include "common.hpp"

#include <algorithm>
#include <cstdio>

namespace palz {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    set_error(buf);
    return e == hipErrorOutOfMemory ? PALZ_ERR_OUT_OF_MEMORY : PALZ_ERR_HIP;
}

namespace {

// Streaming copy used to measure the achievable HBM bandwidth and to calibrate
// the FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic.
__global__ void __launch_bounds__(256) k_copy16(float4* __restrict__ dst, const float4* __restrict__ src,
                                                size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

}  // namespace palz

extern "C" {

int palz_copy_bytes(void* d_dst, const void* d_src, size_t bytes, void* stream) {
    if ((reinterpret_cast<uintptr_t>(d_dst) | reinterpret_cast<uintptr_t>(d_src) | bytes) & 15u)
        return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_copy_bytes: pointers and size must be 16-byte multiples");
    const size_t n16 = bytes / 16u;
    if (n16 == 0) return PALZ_OK;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 32);
    palz::k_copy16<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>((float4*)d_dst, (const float4*)d_src, n16);
    PALZ_LAUNCH_CHECK();
    return PALZ_OK;
}

const char* palz_last_error(void) { return palz::g_last_error.c_str(); }

int palz_version(void) { return 100; }

int palz_device_count(int* count) {
    if (!count) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_device_count: NULL");
    *count = 0;
    PALZ_HIP(hipGetDeviceCount(count));
    return PALZ_OK;
}

int palz_set_device(int device) {
    PALZ_HIP(hipSetDevice(device));
    return PALZ_OK;
}

int palz_malloc(void** d_ptr, size_t bytes) {
    if (!d_ptr) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_malloc: NULL");
    *d_ptr = nullptr;
    if (bytes == 0) return PALZ_OK;
    PALZ_HIP(hipMalloc(d_ptr, bytes));
    return PALZ_OK;
}

int palz_free(void* d_ptr) {
    if (d_ptr) PALZ_HIP(hipFree(d_ptr));
    return PALZ_OK;
}

int palz_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, void* stream) {
    if (bytes == 0) return PALZ_OK;
    PALZ_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    PALZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PALZ_OK;
}

int palz_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, void* stream) {
    if (bytes == 0) return PALZ_OK;
    PALZ_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PALZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PALZ_OK;
}

int palz_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream) {
    if (bytes == 0) return PALZ_OK;
    PALZ_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return PALZ_OK;
}

int palz_memset(void* d_dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return PALZ_OK;
    PALZ_HIP(hipMemsetAsync(d_dst, value, bytes, (hipStream_t)stream));
    return PALZ_OK;
}

int palz_stream_create(void** stream) {
    if (!stream) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_stream_create: NULL");
    hipStream_t s = nullptr;
    PALZ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return PALZ_OK;
}

int palz_stream_destroy(void* stream) {
    if (stream) PALZ_HIP(hipStreamDestroy((hipStream_t)stream));
    return PALZ_OK;
}

int palz_stream_sync(void* stream) {
    PALZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PALZ_OK;
}

int palz_device_sync(void) {
    PALZ_HIP(hipDeviceSynchronize());
    return PALZ_OK;
}

int palz_event_create(void** event) {
    if (!event) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_event_create: NULL");
    hipEvent_t e = nullptr;
    PALZ_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return PALZ_OK;
}

int palz_event_destroy(void* event) {
    if (event) PALZ_HIP(hipEventDestroy((hipEvent_t)event));
    return PALZ_OK;
}

int palz_event_record(void* event, void* stream) {
    PALZ_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return PALZ_OK;
}

int palz_event_elapsed_ms(float* ms, void* start, void* stop) {
    if (!ms) return palz::fail(PALZ_ERR_INVALID_ARGUMENT, "palz_event_elapsed_ms: NULL");
    PALZ_HIP(hipEventSynchronize((hipEvent_t)stop));
    PALZ_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return PALZ_OK;
}

}  // extern "C"
