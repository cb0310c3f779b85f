// runtime.h -- host-side helpers shared by the entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

#include "sedi_internal.h"

namespace sedi {

int ensure_device();
bool is_device_pointer(const void* p);

// Device buffer owned by the library (RAII, no copies).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows to at least `n` bytes (contents not kept).
    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        else p = nullptr;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace sedi
