// estimator.hip -- the stochastic diag(A^-1) sample loop on one GPU.
//
// stochastic-estimator.cc:158-216: for each sample, v = probe, x = 0,
// solve A x = v, t += x.*v, q += v.*v; finally D = t./q.  The probe stream
// (rng.hip) runs on its own HIP stream one sample ahead of the solver, into a
// double buffer guarded by events, so probe generation overlaps the solves.
#include "estimator.h"
#include "tuning.h"
#include "profiler.h"

#include <algorithm>
#include <chrono>
#include <cstdio>

namespace sedi {

__global__ void fill_kernel(double* x, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

// t += x.*v ; q += v.*v  (SliceOps::accumulate_product, SliceOps.hpp:96-108,
// called at stochastic-estimator.cc:195,198)
__global__ void accumulate_kernel(const double* __restrict__ x, const double* __restrict__ v,
                                  double* __restrict__ t, double* __restrict__ q, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double vv = v[i];
        t[i] = t[i] + x[i] * vv;
        q[i] = q[i] + vv * vv;
    }
}

// D = t./q (SliceOps::divide, SliceOps.hpp:68-79; stochastic-estimator.cc:216)
__global__ void divide_kernel(const double* __restrict__ t, const double* __restrict__ q,
                              double* __restrict__ D, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        D[i] = t[i] / q[i];
}

static dim3 ew_grid(long long n) {
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

int fill_dev(double* x, long long n, double v, hipStream_t s) {
    SEDI_HIP(launch("fill_kernel", fill_kernel, ew_grid(n), dim3(256), s, x, n, v));
    return SEDI_OK;
}

static const char* solver_name(int kind) {
    return kind == SEDI_CG ? "Cg" : (kind == SEDI_PCG ? "PipelinedCg" : "BiCgStab");
}

Estimator::~Estimator() {
    if (rng_stream_) (void)hipStreamSynchronize(rng_stream_);
    for (auto& e : ev_v_) if (e) (void)hipEventDestroy(e);
    for (auto& e : ev_free_) if (e) (void)hipEventDestroy(e);
    if (rng_stream_) (void)hipStreamDestroy(rng_stream_);
}

int Estimator::create(const sedi_estimate_args& a) {
    SEDI_REQUIRE(a.grid.n1 > 0 && a.grid.n2 > 0 && a.grid.n3 > 0, SEDI_ERR_ARG, "bad grid");
    SEDI_REQUIRE(a.grid.n1 < (1 << 30) && a.grid.n2 < (1 << 30) && a.grid.n3 < (1 << 30), SEDI_ERR_ARG,
                 "grid dimension too large");
    SEDI_REQUIRE(a.samples >= 0 && a.maxiter >= 0, SEDI_ERR_ARG, "samples and maxiter must be >= 0");
    SEDI_REQUIRE(a.precision == SEDI_FP64 || a.precision == SEDI_FP32_MIXED, SEDI_ERR_ARG,
                 "unknown precision %d", a.precision);
    SEDI_REQUIRE(a.rep_every >= 0, SEDI_ERR_ARG, "rep_every must be >= 0");
    SEDI_REQUIRE(a.batch_k >= 0, SEDI_ERR_ARG, "batch_k must be >= 0");
    args_ = a;
    shape_ = Shape{(int)a.grid.n1, (int)a.grid.n2, (int)a.grid.n3};
    const size_t n = (size_t)shape_.n();
    const bool fp32 = a.precision == SEDI_FP32_MIXED;
    int batch_k = a.batch_k;
    if (batch_k == 0) {
        // chosen by the autotuner for this solver, precision and grid (the
        // entry of nearest cell count); without a table: 64 in the mixed
        // form, one at a time in FP64 (where the plain solver is as fast)
        char key[32];
        std::snprintf(key, sizeof key, "k:%d:%d", a.kind, a.precision);
        batch_k = tuned_value(key, shape_, fp32 ? 64 : 1, true);
        // no more than a share of the free device memory holds
        size_t free_b = 0, total_b = 0;
        if (batch_k > 4 && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const double per = (double)BatchSolver::bytes_needed(shape_, a.kind, 4, fp32) / 4;
            const int fit = (int)(0.8 * (double)free_b / per) / 4 * 4;
            batch_k = std::max(4, std::min(batch_k, fit));
        }
    }
    batched_ = batch_k > 1 || fp32;
    if (batched_) {
        SEDI_REQUIRE(a.kind == SEDI_CG || a.kind == SEDI_PCG || a.kind == SEDI_BICGSTAB, SEDI_ERR_ARG,
                     "unknown solver kind %d", a.kind);
        // blocks of 4 probes, no more than the run needs
        int K = std::max(4, (batch_k + 3) / 4 * 4);
        K = std::min(K, std::max(4, (a.samples + 3) / 4 * 4));
        batch_.reset(new BatchSolver());
        int rc = batch_->create(shape_, a.stencil.a, a.stencil.b, a.stencil.c, a.kind, K,
                                a.precision == SEDI_FP32_MIXED, a.rep_every);
        if (rc) return rc;
        SEDI_HIP(buf_.reserve(sizeof(double) * n * 2));
        t_ = buf_.as<double>();
        q_ = t_ + n;
        return SEDI_OK;
    }
    SEDI_REQUIRE(offsets_fit_32(shape_.n(), sizeof(double)), SEDI_ERR_UNSUPPORTED,
                 "grid of %lld cells too large for the plain FP64 path (a vector must be < 4 GiB); "
                 "the batched path takes grids up to 4 GiB per block of 4 FP32 probes", shape_.n());
    int rc = solver_.create(shape_, a.stencil.a, a.stencil.b, a.stencil.c, a.kind, a.mode);
    if (rc) return rc;
    SEDI_HIP(buf_.reserve(sizeof(double) * n * 5));
    v_[0] = buf_.as<double>();
    v_[1] = v_[0] + n;
    x_ = v_[1] + n;
    t_ = x_ + n;
    q_ = t_ + n;
    SEDI_HIP(hipStreamCreateWithFlags(&rng_stream_, hipStreamNonBlocking));
    for (int k = 0; k < 2; k++) {
        SEDI_HIP(hipEventCreateWithFlags(&ev_v_[k], hipEventDisableTiming));
        SEDI_HIP(hipEventCreateWithFlags(&ev_free_[k], hipEventDisableTiming));
    }
    return rng_.init();
}

void Estimator::log(const char* line) {
    if (log_fn_) log_fn_(log_user_, line);
}

int Estimator::set_rhs(const double* v) {
    if (!v) {
        fixed_rhs_ = false;
        return SEDI_OK;
    }
    const size_t bytes = sizeof(double) * (size_t)shape_.n();
    SEDI_HIP(rhs_.reserve(bytes));
    SEDI_HIP(hipMemcpy(rhs_.p, v, bytes, is_device_pointer(v) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    fixed_rhs_ = true;
    return SEDI_OK;
}

// x, v: device vectors of the sample; D = t./q so far (the reference computes
// it per sample in estimator validation mode, stochastic-estimator.cc:200-205).
int Estimator::call_hook(int k, const double* x, const double* v, hipStream_t s) {
    const long long n = shape_.n();
    const size_t bytes = sizeof(double) * (size_t)n;
    const bool wx = hook_want_ & SEDI_HOOK_X, wv = hook_want_ & SEDI_HOOK_V, wd = hook_want_ & SEDI_HOOK_D;
    if (wd) {
        SEDI_HIP(hook_buf_.reserve(bytes));
        SEDI_HIP(launch("divide_kernel", divide_kernel, ew_grid(n), dim3(256), s, (const double*)t_,
                        (const double*)q_, hook_buf_.as<double>(), n));
    }
    SEDI_HIP(hipStreamSynchronize(s));
    if (wx) {
        hx_.resize(n);
        SEDI_HIP(hipMemcpy(hx_.data(), x, bytes, hipMemcpyDeviceToHost));
    }
    if (wv) {
        hv_.resize(n);
        SEDI_HIP(hipMemcpy(hv_.data(), v, bytes, hipMemcpyDeviceToHost));
    }
    if (wd) {
        hd_.resize(n);
        SEDI_HIP(hipMemcpy(hd_.data(), hook_buf_.p, bytes, hipMemcpyDeviceToHost));
    }
    hook_(hook_user_, k, wx ? hx_.data() : nullptr, wv ? hv_.data() : nullptr, wd ? hd_.data() : nullptr);
    return SEDI_OK;
}

int Estimator::call_hook_batched(int k, int slot, hipStream_t s) {
    const long long n = shape_.n();
    const size_t bytes = sizeof(double) * (size_t)n;
    SEDI_HIP(tmp_xv_.reserve(2 * bytes));
    double* x = tmp_xv_.as<double>();
    double* v = x + n;
    int rc = batch_->store_x(x, slot, slot + 1);
    if (!rc) rc = batch_->store_rhs(v, slot, slot + 1);
    if (rc) return rc;
    return call_hook(k, x, v, s);
}

void Estimator::log_history(int sample, const std::vector<double>& hist, int nh) {
    char line[160];
    snprintf(line, sizeof(line), "Solving for sample %d/%d...", sample + 1, args_.samples);
    log(line);
    for (int i = 0; i < nh && i < (int)hist.size(); i++) {
        snprintf(line, sizeof(line), "%s: after iter=%d, err=%.10e", solver_name(args_.kind), i + 1, hist[i]);
        log(line);
    }
}

// Probes in batches of K (the last one padded with zero right-hand sides);
// each batch is generated, solved together and accumulated in probe order, so
// t and q see the same sequence of additions as the one-at-a-time loop.
int Estimator::run_batched(double* D, int* iters, double* errs, int* codes, double* seconds) {
    const long long n = shape_.n();
    BatchSolver& bs = *batch_;
    hipStream_t s = bs.stream();
    int rc = fill_dev(t_, n, 0.0, s);
    if (!rc) rc = fill_dev(q_, n, 0.0, s);
    if (rc) return rc;
    SEDI_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    const int K = bs.K();
    int status = 0;
    char line[160];
    BatchResult res;
    for (int k0 = 0; k0 < args_.samples && !status; k0 += K) {
        const int kact = std::min(K, args_.samples - k0);
        if (fixed_rhs_) rc = bs.load_rhs(rhs_.as<double>(), kact, true);
        else rc = fill_probe_batch(args_.seed, n, k0, K, kact, bs.fp32(), bs.rhs(), s, scratch_);
        if (rc || (rc = bs.solve(args_.maxiter, args_.tol, &res))) return rc;
        int kuse = kact;
        for (int k = 0; k < kact; k++) {
            const int code = res.codes[k];
            if (log_fn_) {
                const int nh = code == 0 || code == 1 ? std::min(res.iters[k], args_.maxiter) : 0;
                log_history(k0 + k, bs.history(k), nh);
            }
            if (iters) iters[k0 + k] = res.iters[k];
            if (errs) errs[k0 + k] = res.errs[k];
            if (codes) codes[k0 + k] = code;
            if (!args_.ignore_errors && code != 0) {  // stochastic-estimator.cc:181-184
                snprintf(line, sizeof(line), "Incomplete convergence in solver, errcode=%d", code);
                log(line);
                status = 100 + code;
                kuse = k;
                break;
            }
            if (hook_) {  // one sample at a time, as the reference's loop
                if ((rc = bs.accumulate(t_, q_, k, k + 1)) || (rc = call_hook_batched(k0 + k, k, s))) return rc;
            }
        }
        if (!hook_ && (rc = bs.accumulate(t_, q_, 0, kuse))) return rc;
    }
    if (D) {
        double* dd = D;
        const bool host = !is_device_pointer(D);
        DevBuf tmp;
        if (host) {
            SEDI_HIP(tmp.reserve(sizeof(double) * n));
            dd = tmp.as<double>();
        }
        SEDI_HIP(launch("divide_kernel", divide_kernel, ew_grid(n), dim3(256), s, (const double*)t_,
                        (const double*)q_, dd, n));
        SEDI_HIP(hipStreamSynchronize(s));
        if (host) SEDI_HIP(hipMemcpy(D, dd, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    SEDI_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    profile_flush();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return status;
}

int Estimator::run(double* D, int* iters, double* errs, int* codes, double* seconds) {
    if (batched_) return run_batched(D, iters, errs, codes, seconds);
    const long long n = shape_.n();
    hipStream_t s = solver_.stream();
    // set-up outside the timed scope: t = q = 0, stream seeded, first probe
    int rc = fill_dev(t_, n, 0.0, s);
    if (!rc) rc = fill_dev(q_, n, 0.0, s);
    if (rc) return rc;
    SEDI_HIP(hipEventRecord(ev_free_[0], s));
    SEDI_HIP(hipEventRecord(ev_free_[1], s));
    if (!fixed_rhs_) {
        if ((rc = rng_.seed(args_.seed, rng_stream_))) return rc;
        if (args_.samples > 0) {
            if ((rc = rng_.fill(v_[0], n, rng_stream_))) return rc;
            SEDI_HIP(hipEventRecord(ev_v_[0], rng_stream_));
        }
    }
    SEDI_HIP(hipStreamSynchronize(s));
    SEDI_HIP(hipStreamSynchronize(rng_stream_));

    const auto t0 = std::chrono::steady_clock::now();
    int status = 0;
    char line[160];
    std::vector<double> hist;
    for (int k = 0; k < args_.samples; k++) {
        const int cur = k & 1, nxt = cur ^ 1;
        snprintf(line, sizeof(line), "Solving for sample %d/%d...", k + 1, args_.samples);
        log(line);
        const double* vk = fixed_rhs_ ? rhs_.as<double>() : v_[cur];
        if (!fixed_rhs_) {
            if (k + 1 < args_.samples) {
                SEDI_HIP(hipStreamWaitEvent(rng_stream_, ev_free_[nxt], 0));
                if ((rc = rng_.fill(v_[nxt], n, rng_stream_))) return rc;
                SEDI_HIP(hipEventRecord(ev_v_[nxt], rng_stream_));
            }
            SEDI_HIP(hipStreamWaitEvent(s, ev_v_[cur], 0));
        }
        if ((rc = fill_dev(x_, n, 0.0, s))) return rc;  // stochastic-estimator.cc:175
        double fe = 0.0;
        int fi = 0;
        const int code = solver_.solve_dev(vk, x_, nullptr, args_.maxiter, args_.tol, &fe, &fi);
        if (code < 0) return code;
        if (log_fn_) {
            const int nh = solver_.err_history(nullptr, 0);
            hist.resize(nh);
            solver_.err_history(hist.data(), nh);
            for (int i = 0; i < nh; i++) {
                snprintf(line, sizeof(line), "%s: after iter=%d, err=%.10e", solver_name(args_.kind), i + 1,
                         hist[i]);
                log(line);
            }
        }
        if (iters) iters[k] = fi;
        if (errs) errs[k] = fe;
        if (codes) codes[k] = code;
        if (!args_.ignore_errors && code != 0) {  // stochastic-estimator.cc:181-184
            snprintf(line, sizeof(line), "Incomplete convergence in solver, errcode=%d", code);
            log(line);
            status = 100 + code;
            break;
        }
        SEDI_HIP(launch("accumulate_kernel", accumulate_kernel, ew_grid(n), dim3(256), s,
                        (const double*)x_, vk, t_, q_, n));
        if (hook_ && (rc = call_hook(k, x_, vk, s))) return rc;
        SEDI_HIP(hipEventRecord(ev_free_[cur], s));
    }
    if (D) {
        double* dd = D;
        const bool host = !is_device_pointer(D);
        if (host) dd = x_;  // x_ is free once the loop is over
        SEDI_HIP(launch("divide_kernel", divide_kernel, ew_grid(n), dim3(256), s, (const double*)t_,
                        (const double*)q_, dd, n));
        SEDI_HIP(hipStreamSynchronize(s));
        if (host) SEDI_HIP(hipMemcpy(D, dd, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    SEDI_HIP(hipStreamSynchronize(s));
    SEDI_HIP(hipStreamSynchronize(rng_stream_));
    const auto t1 = std::chrono::steady_clock::now();
    profile_flush();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return status;
}

}  // namespace sedi
