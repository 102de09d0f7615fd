// device.hpp -- device memory and error plumbing shared by the host C++
// components (include/prodm/components.hpp, gpu_mdr.hpp).
//
// Errors follow ProDM's convention for fatal paths (print, then exit; the
// reference has no exceptions and no status returns), while the C ABI
// underneath returns status codes.
#ifndef PRODM_DEVICE_HPP
#define PRODM_DEVICE_HPP

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../prodm_mdr.h"

namespace prodm {

[[noreturn]] inline void fatal(const std::string &what) {
    std::fprintf(stderr, "prodm: %s\n", what.c_str());
    std::exit(-1);
}

inline void check(int rc, const char *what) {
    if (rc != PRODM_OK) fatal(std::string(what) + " failed: " + prodm_last_error());
}

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fatal(std::string(what) + ": " + hipGetErrorString(e));
}

// Grow-only device allocation, reused across calls.
class DeviceArena {
  public:
    DeviceArena() = default;
    DeviceArena(const DeviceArena &) = delete;
    DeviceArena &operator=(const DeviceArena &) = delete;
    ~DeviceArena() {
        if (p_) (void)hipFree(p_);
    }
    void *get(size_t bytes) {
        if (bytes > cap_) {
            if (p_) hip_check(hipFree(p_), "hipFree");
            p_ = nullptr;
            hip_check(hipMalloc(&p_, bytes ? bytes : 256), "hipMalloc");
            cap_ = bytes;
        }
        return p_;
    }
    template <class T> T *as(size_t count) { return static_cast<T *>(get(count * sizeof(T))); }
    size_t capacity() const { return cap_; }

  private:
    void *p_ = nullptr;
    size_t cap_ = 0;
};

// The device state a component needs: a stream and a few reusable buffers.
// Components are copyable handles onto one context (ProDM composes its
// components by value, ComposedRefactor.hpp:15-19).
struct DeviceContext {
    hipStream_t stream = nullptr;  // default stream
    DeviceArena a, b, ws, aux;
};
using Context = std::shared_ptr<DeviceContext>;
inline Context make_context() { return std::make_shared<DeviceContext>(); }

template <class T> constexpr int dtype_of() { return sizeof(T) == 8 ? PRODM_F64 : PRODM_F32; }

inline uint64_t product(const std::vector<uint32_t> &v) {
    uint64_t n = 1;
    for (uint32_t d : v) n *= d;
    return n;
}

// Number of elements from the first to the last element of a box of `dims`
// laid out with `strides` (dense when strides is empty).
inline uint64_t span_of(const std::vector<uint32_t> &dims, const std::vector<uint32_t> &strides) {
    if (strides.empty()) return product(dims);
    uint64_t last = 0;
    for (size_t d = 0; d < dims.size(); d++) {
        if (dims[d] == 0) return 0;
        last += (uint64_t)(dims[d] - 1) * strides[d];
    }
    return last + 1;
}

template <class T> void to_device(T *d, const T *h, uint64_t n, hipStream_t s) {
    if (n) hip_check(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}
template <class T> void to_host(T *h, const T *d, uint64_t n, hipStream_t s) {
    if (n) hip_check(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

}  // namespace prodm
#endif
