// Shared host-side helpers: error types and device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "layout.hpp"

namespace imps {

struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };
struct NotConverged : std::runtime_error { using std::runtime_error::runtime_error; };
struct Numerical : std::runtime_error { using std::runtime_error::runtime_error; };

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Growable device allocation (hipMalloc on growth only).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) { hip_check(hipFree(p), "hipFree"); p = nullptr; bytes = 0; }
        if (n == 0) return;
        hip_check(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
    }
    double* d() const { return static_cast<double*>(p); }
};

}  // namespace imps
