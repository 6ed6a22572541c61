// Host-side helpers of the C ABI implementation: thread-local error state,
// HIP status checks and launch geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "chysx_hip.h"

namespace chx {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Error(CHX_ERR_DEVICE, std::string("chysx::hip: ") + what + " failed: " + hipGetErrorString(e));
}

inline void require(bool cond, const char* msg) {
    if (!cond) throw Error(CHX_ERR_INVALID_ARGUMENT, msg);
}

inline hipStream_t as_stream(uintptr_t s) { return reinterpret_cast<hipStream_t>(s); }

// Runs `fn`, converting exceptions into an error code + message.
template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return CHX_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CHX_ERR_RUNTIME;
    }
}

inline int blocks_for(long n, int block) { return static_cast<int>((n + block - 1) / block); }

}  // namespace chx
