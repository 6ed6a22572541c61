// Host-side basics of the C++ API: plain vector / block types with the
// reference's memory layout (ChysX/src/math/vec.cuh, matrix.cuh:130-133),
// a non-owning device span (ChysX/src/memory/device_span.h:34-78), an owning
// host-mirrored device array (the role of CudaArray<T>,
// ChysX/src/memory/cuda_array.h:64-284) and the C-ABI error translation.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "chysx_hip.h"

namespace chysx {

namespace math {
template <class T> struct Vec2 {
    T x{}, y{};
    Vec2() = default;
    Vec2(T a, T b) : x(a), y(b) {}
};
template <class T> struct Vec3 {
    T x{}, y{}, z{};
    Vec3() = default;
    Vec3(T a, T b, T c) : x(a), y(b), z(c) {}
};
template <class T> struct Vec4 {
    T x{}, y{}, z{}, w{};
    Vec4() = default;
    Vec4(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
};
using Vec2i = Vec2<int>;
using Vec3i = Vec3<int>;
using Vec4i = Vec4<int>;
using Vec3f = Vec3<float>;
using Vec4f = Vec4<float>;
struct Mat3f {
    float data[9]{};
};
struct Mat2f {
    float data[4]{};
};
static_assert(sizeof(Vec3f) == 12 && sizeof(Vec3i) == 12 && sizeof(Vec4i) == 16 && sizeof(Vec4f) == 16 &&
                  sizeof(Mat3f) == 36,
              "layout must match the reference");
}  // namespace math

// Translate a C-ABI status into the reference's exception types:
// "chysx::<component>: <what> failed: <message>".
inline void check(int rc, const char* component, const char* what) {
    if (rc == CHX_OK) return;
    const std::string msg = std::string("chysx::") + component + ": " + what + " failed: " + chx_last_error_message();
    if (rc == CHX_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    if (rc == CHX_ERR_OUT_OF_RANGE) throw std::out_of_range(msg);
    throw std::runtime_error(msg);
}

// Non-owning view of device memory.
template <class T> class DeviceSpan {
public:
    DeviceSpan() = default;
    DeviceSpan(T* p, std::size_t n) : p_(p), n_(n) {}
    static DeviceSpan from_raw(std::uintptr_t ptr, std::size_t n) { return DeviceSpan(reinterpret_cast<T*>(ptr), n); }
    T* data() const noexcept { return p_; }
    std::size_t size() const noexcept { return n_; }
    bool empty() const noexcept { return n_ == 0; }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Owning device buffer with a host mirror.  Device memory is allocated on
// first use, so host-only set-up (topology) runs without a GPU.
template <class T> class DeviceArray {
public:
    DeviceArray() = default;
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept { *this = std::move(o); }
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        if (this != &o) {
            release();
            d_ = o.d_;
            dn_ = o.dn_;
            host_ = std::move(o.host_);
            o.d_ = nullptr;
            o.dn_ = 0;
        }
        return *this;
    }

    // Host mirror sized to n; the device buffer follows on the next device use.
    void resize(std::size_t n) { host_.resize(n); }
    std::size_t cpu_size() const noexcept { return host_.size(); }
    std::size_t gpu_size() const noexcept { return dn_; }
    T* cpu_data() noexcept { return host_.empty() ? nullptr : host_.data(); }
    const T* cpu_data() const noexcept { return host_.empty() ? nullptr : host_.data(); }
    std::vector<T>& host() noexcept { return host_; }
    const std::vector<T>& host() const noexcept { return host_; }
    T* gpu_data() const noexcept { return d_; }

    // Device buffer of exactly `n` elements (contents undefined after a
    // reallocation).  Returns true when it was (re)allocated.
    bool ensure_device(std::size_t n) {
        if (n == dn_ && (d_ != nullptr || n == 0)) return false;
        release();
        if (n > 0) {
            void* p = nullptr;
            check(chx_malloc(&p, n * sizeof(T)), "memory", "device allocation");
            d_ = static_cast<T*>(p);
        }
        dn_ = n;
        return true;
    }
    void zero_device(std::uintptr_t stream = 0) {
        if (dn_ > 0) check(chx_zero(d_, dn_ * sizeof(T), stream), "memory", "zero fill");
    }
    // Host mirror -> device (allocating to the mirror's size).
    void copy_to_device(std::uintptr_t stream = 0) {
        ensure_device(host_.size());
        if (!host_.empty()) check(chx_memcpy_h2d(d_, host_.data(), host_.size() * sizeof(T), stream), "memory", "host-to-device copy");
    }
    // Device -> host mirror.
    void copy_to_host(std::uintptr_t stream = 0) {
        host_.resize(dn_);
        if (dn_ > 0) check(chx_memcpy_d2h(host_.data(), d_, dn_ * sizeof(T), stream), "memory", "device-to-host copy");
    }
    void release() noexcept {
        if (d_ != nullptr) (void)chx_free(d_);
        d_ = nullptr;
        dn_ = 0;
    }

private:
    T* d_ = nullptr;
    std::size_t dn_ = 0;
    std::vector<T> host_;
};

// Host-side profiler range for the life of a scope (the reference's
// CHYSX_NVTX_RANGE, ChysX/src/profile/nvtx_range.h:58-100; roctx here, shown
// by rocprofv3 --marker-trace).  Never throws.
class ProfileRange {
public:
    explicit ProfileRange(const char* name) noexcept { (void)chx_profile_range_push(name); }
    ~ProfileRange() { (void)chx_profile_range_pop(); }
    ProfileRange(const ProfileRange&) = delete;
    ProfileRange& operator=(const ProfileRange&) = delete;
};

// Page-locked staging for small asynchronous host-to-device uploads (the
// reference's pinned host mirrors, cuda_array.cu:41-62): a ring of two
// buffers, each reused only after the copy that read it has run on its
// stream, so the caller's data may change as soon as upload() returns and
// the host waits only when it is two uploads ahead of the device.
class HostStager {
public:
    HostStager() = default;
    ~HostStager() { release(); }
    HostStager(const HostStager&) = delete;
    HostStager& operator=(const HostStager&) = delete;
    HostStager(HostStager&& o) noexcept { *this = std::move(o); }
    HostStager& operator=(HostStager&& o) noexcept {
        if (this != &o) {
            release();
            for (int i = 0; i < 2; ++i) {
                slot_[i] = o.slot_[i];
                o.slot_[i] = Slot{};
            }
            next_ = o.next_;
        }
        return *this;
    }
    void upload(void* dst, const void* src, std::size_t bytes, std::uintptr_t stream) {
        if (bytes == 0) return;
        Slot& s = slot_[next_];
        next_ ^= 1;
        if (s.event == nullptr) check(chx_event_create(&s.event), "memory", "staging event");
        if (s.pending) check(chx_event_synchronize(s.event), "memory", "staging buffer reuse");
        s.pending = false;
        if (s.cap < bytes) {
            if (s.buf) check(chx_host_free(s.buf), "memory", "staging buffer");
            s.buf = nullptr;
            s.cap = 0;
            check(chx_host_malloc(&s.buf, bytes), "memory", "staging buffer");
            s.cap = bytes;
        }
        std::memcpy(s.buf, src, bytes);
        check(chx_memcpy_h2d_async(dst, s.buf, bytes, stream), "memory", "host-to-device copy");
        check(chx_event_record(s.event, stream), "memory", "staging event");
        s.pending = true;
    }
    void release() noexcept {
        for (Slot& s : slot_) {
            if (s.pending && s.event) (void)chx_event_synchronize(s.event);
            if (s.buf) (void)chx_host_free(s.buf);
            if (s.event) (void)chx_event_destroy(s.event);
            s = Slot{};
        }
    }

private:
    struct Slot {
        void* buf = nullptr;
        std::size_t cap = 0;
        void* event = nullptr;
        bool pending = false;
    };
    Slot slot_[2];
    int next_ = 0;
};

}  // namespace chysx
