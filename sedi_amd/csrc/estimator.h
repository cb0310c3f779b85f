// estimator.h -- the sample loop (sedi_estimator).
#pragma once

#include <vector>

#include <memory>

#include "batched.h"
#include "plain.h"
#include "rng.h"

namespace sedi {

int fill_dev(double* x, long long n, double v, hipStream_t s);

class Estimator {
  public:
    ~Estimator();
    int create(const sedi_estimate_args& a);
    int run(double* D, int* iters, double* errs, int* codes, double* seconds);
    void set_log(sedi_log_fn fn, void* user) { log_fn_ = fn; log_user_ = user; }
    // Validation mode (stochastic-estimator.cc:162-171): every sample solves
    // with this right-hand side (n cells, host or device) instead of a probe;
    // NULL returns to the probes.
    int set_rhs(const double* v);
    void set_sample_hook(sedi_sample_fn fn, int want, void* user) {
        hook_ = fn;
        hook_want_ = want;
        hook_user_ = user;
    }

  private:
    void log(const char* line);
    void log_history(int sample, const std::vector<double>& hist, int nh);
    int run_batched(double* D, int* iters, double* errs, int* codes, double* seconds);
    // host copies for the sample hook, then the call
    int call_hook(int k, const double* x, const double* v, hipStream_t s);
    int call_hook_batched(int k, int slot, hipStream_t s);
    bool fixed_rhs_ = false;
    DevBuf rhs_, hook_buf_, tmp_xv_;
    sedi_sample_fn hook_ = nullptr;
    int hook_want_ = 0;
    void* hook_user_ = nullptr;
    std::vector<double> hx_, hv_, hd_;
    bool batched_ = false;
    std::unique_ptr<BatchSolver> batch_;
    DevBuf scratch_;
    sedi_estimate_args args_{};
    Shape shape_{};
    PlainSolver solver_;
    MtStream rng_;
    hipStream_t rng_stream_ = nullptr;
    hipEvent_t ev_v_[2] = {nullptr, nullptr};     // probe k ready
    hipEvent_t ev_free_[2] = {nullptr, nullptr};  // probe buffer no longer read
    DevBuf buf_;
    double* v_[2] = {nullptr, nullptr};
    double *x_ = nullptr, *t_ = nullptr, *q_ = nullptr;
    sedi_log_fn log_fn_ = nullptr;
    void* log_user_ = nullptr;
};

}  // namespace sedi
