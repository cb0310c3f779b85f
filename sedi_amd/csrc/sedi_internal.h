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

// The kernels address a vector with 32-bit byte offsets from its base, so
// every vector they touch must be smaller than 4 GiB: `cells` values of
// `bytes_per_cell` bytes each.
inline bool offsets_fit_32(long long cells, double bytes_per_cell) {
    return (double)cells * bytes_per_cell < 4294967296.0;
}

// Inverse diagonal of the Jacobi preconditioner.  The estimator's is the
// constant 1/a (precond.all_cells.fill(1.0/stencil.a),
// stochastic-estimator.cc:102-103), which kernels instantiated with
// PrecondScalar keep in a scalar register; sedi_solve with an explicit iM
// vector uses PrecondVec.
// at() takes the cell's index.
struct PrecondScalar {
    double s;
    __device__ __forceinline__ double at(unsigned) const { return s; }
};
struct PrecondVec {
    const double* v;
    __device__ __forceinline__ double at(unsigned i) const {
        return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(v) + (unsigned)(i * 8u));
    }
};

}  // namespace sedi
