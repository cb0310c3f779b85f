// plain.h -- the one-probe-at-a-time FP64 solvers (drop-in for solver_t::solve).
#pragma once

#include <hip/hip_runtime.h>
#include <vector>

#include "march.h"
#include "runtime.h"
#include "sedi_internal.h"

namespace sedi {

// Device-resident solver scalars and control.  The reference keeps these as
// locals of solve() (Cg.hpp:136-140, PipelinedCg.hpp:130-134,
// BiCgStab.hpp:148-151); here the last workgroup of each pass updates them,
// so the host never waits inside an iteration.
struct DevState {
    double bb, tol, err, final_err;
    double ru, alpha, beta;          // CG (alpha, beta shared with the others)
    double apend[3];                 // CG, pipelined CG: alphas of the group's iterations whose x updates are pending
    double gamma, delta, err_pending;  // pipelined CG
    double rho, rho_prev, omega, ss, cbw;  // BiCGStab
    int maxiter, iter, done, code, final_iter, nhist, sexit, kind;
};

// the control block holds the state in its first 256 bytes
static_assert(sizeof(DevState) <= 256, "DevState outgrew its slot of the control block");

struct Ctl {
    DevState* st;
    unsigned* counter;
    double* partials;
    double* hist;
};

int stencil_apply_dev(const Shape& s, double a, double b, double c, const double* x, double* y,
                      hipStream_t stream);

// Where a solve's device working set lives and where it is launched: shape,
// stencil and kind, the control block (state | arrival counter at +256 |
// BiCGStab's vector table at +384 | partial sums at +512), the err history
// and the vectors.  PlainSolver owns the memory behind one; the
// allocation-free solve points one into a fixed device array.
struct PlainCore {
    Shape shape{};
    double a = 6.0, b = 1.0, c = 1.0;
    int kind = SEDI_CG;
    hipStream_t stream = nullptr;
    Ctl ctl{};
    double* const* vv = nullptr;  // device table of v[0..6] (BiCGStab)
    double* v[11] = {};
};

// Vectors a solver of `kind` keeps: CG r, p x4 (one per iteration of a group
// of solution updates); pipelined CG r, u, w x2, z, q, s, p x4; BiCGStab r x2,
// r^, p x2, v x2.
inline int plain_vectors(int kind) { return kind == SEDI_CG ? 5 : (kind == SEDI_PCG ? 11 : 7); }
static_assert(sizeof(PlainCore::v) / sizeof(double*) == 11, "PlainCore::v holds the most vectors a solver keeps");

// Bytes of the control block for `nblocks` workgroups (3 partial sums each).
inline size_t plain_ctl_bytes(int nblocks) { return 512 + sizeof(double) * 3 * (size_t)nblocks; }

// The allocation-free solve (sedi_solve_noalloc): b, x, iM device arrays,
// every temporary in this library's fixed device array.  Launches on the
// null stream and returns when the solve is done.
int plain_solve_noalloc(const Shape& s, double a, double b, double c, int kind, const double* rhs, double* x,
                        const double* iM, int maxiter, double tol, double* final_err, int* final_iter);
// Bytes of that array a solve needs, and its size.
size_t plain_noalloc_bytes(const Shape& s, int kind, int maxiter);
// Looks up this device's address of the fixed array (once per device).
int plain_noalloc_setup();

class PlainSolver {
  public:
    ~PlainSolver();
    int create(const Shape& s, double a, double b, double c, int kind, int mode);
    // All pointers are device pointers.  Synchronises `stream_` at the end.
    int solve_dev(const double* b, double* x, const double* iM, int maxiter, double tol,
                  double* final_err, int* final_iter);
    // Host-or-device front end (the C-ABI's sedi_solve).
    int solve_any(const double* b, double* x, const double* iM, int maxiter, double tol,
                  double* final_err, int* final_iter);
    int err_history(double* out, int cap) const;
    hipStream_t stream() const { return stream_; }
    const Shape& shape() const { return shape_; }

  private:
    // count iterations: direct launches, or the replay of a captured graph
    int iterate(int count, const double* iM, double* x);
    Shape shape_{};
    double a_ = 6.0, b_ = 1.0, c_ = 1.0;
    int kind_ = 2, mode_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    DevState* pinned_ = nullptr;
    DevBuf ctl_, hist_, vec_;
    DevBuf stage_b_, stage_x_, stage_m_;
    PlainCore core_;            // launch view of ctl_, hist_ and vec_
    std::vector<double> hist_host_;
    int nhist_ = 0;
    // captured chunk of iterations and what its launches were bound to
    hipGraphExec_t graph_ = nullptr;
    int graph_count_ = 0;
    const void *graph_x_ = nullptr, *graph_m_ = nullptr, *graph_hist_ = nullptr;
    bool geometry_ready_[2] = {false, false};  // scalar / vector preconditioner kernels
};

}  // namespace sedi
