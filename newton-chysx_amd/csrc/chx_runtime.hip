// C ABI: error state, device memory and stream utilities
// (the reference's CudaArray / check_cuda layer:
// ChysX/src/memory/cuda_array.h:64-284, cuda_array.cu:41-90).
#include <algorithm>
#include <cstring>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "chx_device.h"
#include "chx_internal.h"

namespace chx {

namespace {
thread_local std::string g_last_error;

__global__ void zero_kernel(uint32_t* __restrict__ p, size_t words) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0u;
}
__global__ void zero_bytes_kernel(uint8_t* __restrict__ p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

}  // namespace chx

using namespace chx;

extern "C" {

const char* chx_last_error_message(void) { return g_last_error.c_str(); }

int chx_device_count(int* count) {
    return guarded([&] {
        require(count != nullptr, "chx_device_count: null output");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *count = c;
    });
}

int chx_get_device(int* device) {
    return guarded([&] {
        require(device != nullptr, "chx_get_device: null output");
        hip_check(hipGetDevice(device), "hipGetDevice");
    });
}

int chx_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        require(ptr != nullptr, "chx_malloc: null output pointer");
        *ptr = nullptr;
        if (bytes == 0) return;
        hip_check(hipMalloc(ptr, bytes), "hipMalloc");
    });
}

int chx_free(void* ptr) {
    return guarded([&] {
        if (ptr) hip_check(hipFree(ptr), "hipFree");
    });
}

int chx_host_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        require(ptr != nullptr, "chx_host_malloc: null output pointer");
        *ptr = nullptr;
        if (bytes == 0) return;
        hip_check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
    });
}

int chx_host_free(void* ptr) {
    return guarded([&] {
        if (ptr) hip_check(hipHostFree(ptr), "hipHostFree");
    });
}

int chx_memcpy_h2d_async(void* dst, const void* src, size_t bytes, uintptr_t stream) {
    return guarded([&] {
        if (bytes == 0) return;
        require(dst && src, "chx_memcpy_h2d_async: null pointer");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)), "hipMemcpyAsync(H2D)");
        if (stream == 0) hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    });
}

int chx_event_create(void** event) {
    return guarded([&] {
        require(event != nullptr, "chx_event_create: null output");
        hipEvent_t e = nullptr;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
        *event = e;
    });
}

int chx_event_destroy(void* event) {
    return guarded([&] {
        if (event) hip_check(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
    });
}

int chx_event_record(void* event, uintptr_t stream) {
    return guarded([&] {
        require(event != nullptr, "chx_event_record: null event");
        hip_check(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)), "hipEventRecord");
    });
}

int chx_event_synchronize(void* event) {
    return guarded([&] {
        require(event != nullptr, "chx_event_synchronize: null event");
        hip_check(hipEventSynchronize(static_cast<hipEvent_t>(event)), "hipEventSynchronize");
    });
}

int chx_memcpy_h2d(void* dst, const void* src, size_t bytes, uintptr_t stream) {
    return guarded([&] {
        if (bytes == 0) return;
        require(dst && src, "chx_memcpy_h2d: null pointer");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)), "hipMemcpyAsync(H2D)");
        // The source may be reused as soon as this returns.
        hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    });
}

int chx_memcpy_d2h(void* dst, const void* src, size_t bytes, uintptr_t stream) {
    return guarded([&] {
        if (bytes == 0) return;
        require(dst && src, "chx_memcpy_d2h: null pointer");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)), "hipMemcpyAsync(D2H)");
        hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    });
}

int chx_memcpy_d2d(void* dst, const void* src, size_t bytes, uintptr_t stream) {
    return guarded([&] {
        if (bytes == 0) return;
        require(dst && src, "chx_memcpy_d2d: null pointer");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)), "hipMemcpyAsync(D2D)");
    });
}

int chx_zero(void* dst, size_t bytes, uintptr_t stream) {
    return guarded([&] {
        if (bytes == 0) return;
        require(dst != nullptr, "chx_zero: null pointer");
        const uintptr_t a = reinterpret_cast<uintptr_t>(dst);
        if ((a % 4u) == 0 && (bytes % 4u) == 0) {
            const size_t words = bytes / 4;
            const int blocks = static_cast<int>(std::min<size_t>((words + 255) / 256, 8192));
            zero_kernel<<<blocks, 256, 0, as_stream(stream)>>>(static_cast<uint32_t*>(dst), words);
        } else {
            zero_bytes_kernel<<<blocks_for(static_cast<long>(bytes), 256), 256, 0, as_stream(stream)>>>(
                static_cast<uint8_t*>(dst), bytes);
        }
        hip_check(hipGetLastError(), "zero kernel launch");
    });
}

int chx_profile_range_push(const char* name) {
    return guarded([&] {
        require(name != nullptr, "chx_profile_range_push: null name");
        roctxRangePushA(name);
    });
}

int chx_profile_range_pop(void) {
    return guarded([&] { roctxRangePop(); });
}

int chx_stream_synchronize(uintptr_t stream) {
    return guarded([&] { hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize"); });
}

int chx_device_synchronize(void) {
    return guarded([&] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

}  // extern "C"
