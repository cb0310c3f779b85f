// sedi_internal.h -- shared host/device definitions of libsedi_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "sedi_c.h"

namespace sedi {

// Thread-local error text behind sedi_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

#define SEDI_HIP(call)                                                           \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::sedi::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, \
                              hipGetErrorString(e_));                            \
            return SEDI_ERR_HIP;                                                 \
        }                                                                        \
    } while (0)

#define SEDI_REQUIRE(cond, code, ...)          \
    do {                                       \
        if (!(cond)) {                         \
            ::sedi::set_error(__VA_ARGS__);    \
            return (code);                     \
        }                                      \
    } while (0)

// Grid shape as the kernels see it: dense n1*n2*n3, i1 fastest, no ghosts.
struct Shape {
    int n1, n2, n3;
    long long n() const { return (long long)n1 * n2 * n3; }
};

// Scalar or vector inverse diagonal (the reference's precond vector,
// stochastic-estimator.cc:102-103, is the constant 1/a).
struct Precond {
    const double* v;
    double s;
    __device__ __forceinline__ double at(long long idx) const { return v ? v[idx] : s; }
};

}  // namespace sedi
