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
    double apend;                    // CG: alpha of the odd iteration whose x update is pending
    double gamma, delta, err_pending;  // pipelined CG
    double rho, rho_prev, omega, ss, cbw;  // BiCGStab
    int maxiter, iter, done, code, final_iter, nhist, sexit, kind;
};

struct Ctl {
    DevState* st;
    unsigned* counter;
    double* partials;
    double* hist;
};

int stencil_apply_dev(const Shape& s, double a, double b, double c, const double* x, double* y,
                      hipStream_t stream);

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
    template <class PC> int launch_setup(PC pc, const double* b, double* x);
    template <class PC> int launch_iterations(int count, PC pc, double* x);
    // count iterations: direct launches, or the replay of a captured graph
    int iterate(int count, const double* iM, double* x);
    Shape shape_{};
    double a_ = 6.0, b_ = 1.0, c_ = 1.0;
    int kind_ = 2, mode_ = 0;
    int nblocks_max_ = 0;  // partials are sized for this many workgroups
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    DevState* pinned_ = nullptr;
    DevBuf ctl_, hist_, vec_;
    DevBuf stage_b_, stage_x_, stage_m_;
    double* v_[10] = {};        // the solver's device vectors, named in create()
    std::vector<double> hist_host_;
    int nhist_ = 0;
    // captured chunk of iterations and what its launches were bound to
    hipGraphExec_t graph_ = nullptr;
    int graph_count_ = 0;
    const void *graph_x_ = nullptr, *graph_m_ = nullptr, *graph_hist_ = nullptr;
    bool geometry_ready_[2] = {false, false};  // scalar / vector preconditioner kernels
};

}  // namespace sedi
