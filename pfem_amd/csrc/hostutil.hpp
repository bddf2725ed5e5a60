// Host-side helpers: error type, HIP call checking, RAII device buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <stdexcept>
#include <string>

namespace pfem {

// Status codes of the C-ABI (mirrored in include/pfem_gpu.h).
enum Status : int {
    OK = 0,
    E_INVALID = -1,
    E_OOM = -2,
    E_HIP = -3,
    E_UNSUPPORTED = -4,
    E_STATE = -5,
    NOT_CONVERGED = 1,
    NAN_RESIDUAL = 2,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PFEM_HIP(call)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            throw ::pfem::Error(e_ == hipErrorOutOfMemory ? ::pfem::E_OOM : ::pfem::E_HIP,                  \
                                std::string(#call) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                                    std::to_string(__LINE__) + ")");                                        \
    } while (0)

#define PFEM_REQUIRE(cond, code, msg)                          \
    do {                                                       \
        if (!(cond)) throw ::pfem::Error((code), (msg));       \
    } while (0)

// Device allocation owned by one case or context; grows, never shrinks.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void reserve(size_t nb) {
        if (nb < 256) nb = 256;
        if (nb <= bytes) return;
        release();
        PFEM_HIP(hipMalloc(&p, nb));
        bytes = nb;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

inline int grid_for(long long work, int per_block, int cap = 1 << 20) {
    long long g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace pfem
