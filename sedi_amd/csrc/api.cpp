// api.cpp -- the extern "C" entry points of include/sedi_c.h.
#include <hip/hip_runtime.h>

#include <exception>
#include <new>

#include <algorithm>

#include "batched.h"
#include "estimator.h"
#include "h5io.h"
#include "serve.h"
#include "mt_jump.h"
#include "profiler.h"
#include "plain.h"
#include "rng.h"
#include "runtime.h"
#include "sedi_c.h"

struct sedi_solver {
    sedi::PlainSolver impl;
};

struct sedi_estimator {
    sedi::Estimator impl;
};

struct sedi_server {
    sedi::Server impl;
};

struct sedi_batch {
    sedi::BatchSolver impl;
    sedi::BatchResult last;
};

namespace {

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        sedi::set_error("host allocation failed");
        return SEDI_ERR_NOMEM;
    } catch (const std::exception& e) {
        sedi::set_error("internal error: %s", e.what());
        return SEDI_ERR_STATE;
    } catch (...) {
        sedi::set_error("internal error");
        return SEDI_ERR_STATE;
    }
}

int check_grid(const sedi_grid* g, sedi::Shape* s) {
    SEDI_REQUIRE(g, SEDI_ERR_ARG, "grid is NULL");
    SEDI_REQUIRE(g->n1 > 0 && g->n2 > 0 && g->n3 > 0, SEDI_ERR_ARG,
                 "grid dimensions must be positive (got %lldx%lldx%lld)", (long long)g->n1,
                 (long long)g->n2, (long long)g->n3);
    SEDI_REQUIRE(g->n1 < (1LL << 30) && g->n2 < (1LL << 30) && g->n3 < (1LL << 30), SEDI_ERR_ARG,
                 "grid dimension too large");
    *s = sedi::Shape{(int)g->n1, (int)g->n2, (int)g->n3};
    return SEDI_OK;
}

// The plain FP64 kernels index each vector with 32-bit byte offsets.
int check_plain_size(const sedi::Shape& s) {
    SEDI_REQUIRE(sedi::offsets_fit_32(s.n(), sizeof(double)), SEDI_ERR_UNSUPPORTED,
                 "grid of %lld cells too large: an FP64 vector must be < 4 GiB (32-bit offsets)", s.n());
    return SEDI_OK;
}

}  // namespace

extern "C" {

int sedi_stencil_apply(const sedi_grid* grid, const sedi_stencil* st, const double* x, double* y) {
    return guarded([&]() -> int {
        sedi::Shape s;
        int rc = check_grid(grid, &s);
        if (rc) return rc;
        SEDI_REQUIRE(st && x && y, SEDI_ERR_ARG, "stencil, x and y must not be NULL");
        if ((rc = check_plain_size(s))) return rc;
        if ((rc = sedi::ensure_device())) return rc;
        const size_t bytes = sizeof(double) * (size_t)s.n();
        sedi::DevBuf xs, ys;
        const double* xd = x;
        double* yd = y;
        const bool yhost = !sedi::is_device_pointer(y);
        if (!sedi::is_device_pointer(x)) {
            SEDI_HIP(xs.reserve(bytes));
            SEDI_HIP(hipMemcpy(xs.p, x, bytes, hipMemcpyHostToDevice));
            xd = xs.as<double>();
        }
        if (yhost) {
            SEDI_HIP(ys.reserve(bytes));
            yd = ys.as<double>();
        }
        if ((rc = sedi::stencil_apply_dev(s, st->a, st->b, st->c, xd, yd, nullptr))) return rc;
        SEDI_HIP(hipDeviceSynchronize());
        sedi::profile_flush();
        if (yhost) SEDI_HIP(hipMemcpy(y, yd, bytes, hipMemcpyDeviceToHost));
        return SEDI_OK;
    });
}

int sedi_solver_create(const sedi_grid* grid, const sedi_stencil* st, int kind, int mode,
                       sedi_solver** out) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(out && st, SEDI_ERR_ARG, "stencil and out must not be NULL");
        *out = nullptr;
        sedi::Shape s;
        int rc = check_grid(grid, &s);
        if (rc) return rc;
        if ((rc = check_plain_size(s))) return rc;
        if ((rc = sedi::ensure_device())) return rc;
        sedi_solver* h = new sedi_solver();
        rc = h->impl.create(s, st->a, st->b, st->c, kind, mode);
        if (rc) {
            delete h;
            return rc;
        }
        *out = h;
        return SEDI_OK;
    });
}

int sedi_solve(sedi_solver* h, const double* b, double* x, const double* inv_diag, int maxiter,
               double tol, double* final_err, int* final_iter) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(h, SEDI_ERR_ARG, "solver is NULL");
        return h->impl.solve_any(b, x, inv_diag, maxiter, tol, final_err, final_iter);
    });
}

int sedi_solve_noalloc(const sedi_grid* grid, const sedi_stencil* st, int kind, const double* b, double* x,
                       const double* inv_diag, int maxiter, double tol, double* final_err, int* final_iter) {
    return guarded([&]() -> int {
        sedi::Shape s;
        int rc = check_grid(grid, &s);
        if (rc) return rc;
        SEDI_REQUIRE(st && b && x, SEDI_ERR_ARG, "stencil, b and x must not be NULL");
        SEDI_REQUIRE(kind == SEDI_CG || kind == SEDI_PCG || kind == SEDI_BICGSTAB, SEDI_ERR_ARG,
                     "unknown solver kind %d", kind);
        SEDI_REQUIRE(maxiter >= 0, SEDI_ERR_ARG, "maxiter must be >= 0");
        const size_t need = sedi::plain_noalloc_bytes(s, kind, maxiter);
        SEDI_REQUIRE(need <= (size_t)SEDI_NOALLOC_SCRATCH_BYTES, SEDI_ERR_ARG,
                     "solve needs %zu bytes of working memory, above the allocation-free limit of %zu "
                     "(SEDI_NOALLOC_SCRATCH_BYTES); use sedi_solver_create/sedi_solve",
                     need, (size_t)SEDI_NOALLOC_SCRATCH_BYTES);
        if ((rc = sedi::ensure_device())) return rc;
        // no staging: host arrays would need device copies
        SEDI_REQUIRE(sedi::is_device_pointer(b) && sedi::is_device_pointer(x) &&
                         (!inv_diag || sedi::is_device_pointer(inv_diag)),
                     SEDI_ERR_ARG, "sedi_solve_noalloc takes device arrays only");
        return sedi::plain_solve_noalloc(s, st->a, st->b, st->c, kind, b, x, inv_diag, maxiter, tol, final_err,
                                         final_iter);
    });
}

int sedi_solver_err_history(sedi_solver* h, double* out, int cap) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(h, SEDI_ERR_ARG, "solver is NULL");
        return h->impl.err_history(out, cap);
    });
}

void sedi_solver_destroy(sedi_solver* h) { delete h; }

int sedi_rademacher_fill(uint32_t seed, uint64_t first_draw, int64_t ncells, double* v) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(ncells >= 0, SEDI_ERR_ARG, "ncells must be >= 0");
        if (ncells == 0) return SEDI_OK;
        SEDI_REQUIRE(v, SEDI_ERR_ARG, "v is NULL");
        int rc = sedi::ensure_device();
        if (rc) return rc;
        sedi::MtStream mt;
        hipStream_t s = nullptr;
        const size_t bytes = sizeof(double) * (size_t)ncells;
        sedi::DevBuf vs;
        double* vd = v;
        const bool host = !sedi::is_device_pointer(v);
        if (host) {
            SEDI_HIP(vs.reserve(bytes));
            vd = vs.as<double>();
        }
        if ((rc = mt.seed_at(seed, first_draw, s))) return rc;
        if ((rc = mt.fill(vd, ncells, s))) return rc;
        SEDI_HIP(hipDeviceSynchronize());
        sedi::profile_flush();
        if (host) SEDI_HIP(hipMemcpy(v, vd, bytes, hipMemcpyDeviceToHost));
        return SEDI_OK;
    });
}

int sedi_mt_state_after(uint32_t seed, uint64_t draws, uint32_t* state624) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(state624, SEDI_ERR_ARG, "state624 is NULL");
        SEDI_REQUIRE(sedi::mt_charpoly_ok(), SEDI_ERR_STATE, "mt19937 characteristic polynomial unavailable");
        uint32_t s0[624];
        sedi::mt_seed_state(seed, s0);
        sedi::mt_jump_state(s0, draws, state624);
        return SEDI_OK;
    });
}

int sedi_batch_create(const sedi_grid* grid, const sedi_stencil* st, int kind, int K, int precision,
                      int rep_every, sedi_batch** out) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(out && st, SEDI_ERR_ARG, "stencil and out must not be NULL");
        *out = nullptr;
        sedi::Shape s;
        int rc = check_grid(grid, &s);
        if (rc) return rc;
        SEDI_REQUIRE(K >= 1, SEDI_ERR_ARG, "batch size must be >= 1");
        SEDI_REQUIRE(precision == SEDI_FP64 || precision == SEDI_FP32_MIXED, SEDI_ERR_ARG,
                     "unknown precision %d", precision);
        SEDI_REQUIRE(rep_every >= 0, SEDI_ERR_ARG, "rep_every must be >= 0");
        if ((rc = sedi::ensure_device())) return rc;
        sedi_batch* h = new sedi_batch();
        rc = h->impl.create(s, st->a, st->b, st->c, kind, (K + 3) / 4 * 4, precision == SEDI_FP32_MIXED,
                            rep_every);
        if (rc) {
            delete h;
            return rc;
        }
        *out = h;
        return SEDI_OK;
    });
}

int sedi_batch_solve(sedi_batch* h, const double* b, int nrhs, double* x, int maxiter, double tol, int* iters,
                     double* errs, int* codes) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(h && b && x, SEDI_ERR_ARG, "batch, b and x must not be NULL");
        SEDI_REQUIRE(nrhs >= 1 && nrhs <= h->impl.K(), SEDI_ERR_ARG, "nrhs %d outside 1..%d", nrhs,
                     h->impl.K());
        const size_t bytes = sizeof(double) * (size_t)h->impl.n() * nrhs;
        hipStream_t s = h->impl.stream();
        sedi::DevBuf bs, xs;
        const double* bd = b;
        double* xd = x;
        if (!sedi::is_device_pointer(b)) {
            SEDI_HIP(bs.reserve(bytes));
            SEDI_HIP(hipMemcpyAsync(bs.p, b, bytes, hipMemcpyHostToDevice, s));
            bd = bs.as<double>();
        }
        const bool xhost = !sedi::is_device_pointer(x);
        if (xhost) {
            SEDI_HIP(xs.reserve(bytes));
            xd = xs.as<double>();
        }
        int rc = h->impl.load_rhs(bd, nrhs);
        if (!rc) rc = h->impl.solve(maxiter, tol, &h->last);
        if (!rc) rc = h->impl.store_x(xd, 0, nrhs);
        if (rc) return rc;
        SEDI_HIP(hipStreamSynchronize(s));
        if (xhost) SEDI_HIP(hipMemcpy(x, xd, bytes, hipMemcpyDeviceToHost));
        for (int k = 0; k < nrhs; k++) {
            if (iters) iters[k] = h->last.iters[k];
            if (errs) errs[k] = h->last.errs[k];
            if (codes) codes[k] = h->last.codes[k];
        }
        return SEDI_OK;
    });
}

int sedi_batch_err_history(sedi_batch* h, int k, double* out, int cap) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(h, SEDI_ERR_ARG, "batch is NULL");
        SEDI_REQUIRE(k >= 0 && k < (int)h->last.iters.size(), SEDI_ERR_ARG, "no result for right-hand side %d", k);
        const std::vector<double> hist = h->impl.history(k);
        const int code = h->last.codes[k];
        const int n = (code == 0 || code == 1) ? std::min<int>(h->last.iters[k], (int)hist.size()) : 0;
        if (out)
            for (int i = 0; i < n && i < cap; i++) out[i] = hist[i];
        return n;
    });
}

void sedi_batch_destroy(sedi_batch* h) { delete h; }

int sedi_batch_lane_moves(sedi_batch* h, int64_t* moves) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(h && moves, SEDI_ERR_ARG, "batch and moves must not be NULL");
        *moves = h->impl.lane_moves();
        return SEDI_OK;
    });
}

int sedi_batch_compact_enable(int on) {
    sedi::set_batch_compact_enabled(on != 0);
    return SEDI_OK;
}

int sedi_probe_batch(uint32_t seed, int64_t ncells, int k0, int count, double* v) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(ncells >= 1 && k0 >= 0 && count >= 1 && v, SEDI_ERR_ARG, "bad probe batch arguments");
        SEDI_REQUIRE(sedi::mt_charpoly_ok(), SEDI_ERR_STATE, "mt19937 characteristic polynomial unavailable");
        int rc = sedi::ensure_device();
        if (rc) return rc;
        const int K = (count + 3) / 4 * 4;
        sedi::DevBuf tmp, scratch, vs;
        SEDI_HIP(tmp.reserve(sizeof(double) * (size_t)ncells * K));
        const size_t bytes = sizeof(double) * (size_t)ncells * count;
        const bool host = !sedi::is_device_pointer(v);
        double* vd = v;
        if (host) {
            SEDI_HIP(vs.reserve(bytes));
            vd = vs.as<double>();
        }
        hipStream_t s = nullptr;
        if ((rc = sedi::fill_probe_batch(seed, ncells, k0, K, count, false, tmp.p, s, scratch))) return rc;
        if ((rc = sedi::unbatch_rows(tmp.as<double>(), ncells, count, vd, s))) return rc;
        SEDI_HIP(hipStreamSynchronize(s));
        if (host) SEDI_HIP(hipMemcpy(v, vd, bytes, hipMemcpyDeviceToHost));
        return SEDI_OK;
    });
}

int sedi_estimator_set_rhs(sedi_estimator* est, const double* v) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(est, SEDI_ERR_ARG, "estimator is NULL");
        return est->impl.set_rhs(v);
    });
}

int sedi_estimator_set_sample_hook(sedi_estimator* est, sedi_sample_fn fn, int want, void* user) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(est, SEDI_ERR_ARG, "estimator is NULL");
        SEDI_REQUIRE((want & ~(SEDI_HOOK_X | SEDI_HOOK_V | SEDI_HOOK_D)) == 0, SEDI_ERR_ARG, "bad hook mask %d",
                     want);
        est->impl.set_sample_hook(fn, want, user);
        return SEDI_OK;
    });
}

int sedi_server_create(const sedi_server_config* cfg, sedi_server** out) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(cfg && out, SEDI_ERR_ARG, "cfg and out must not be NULL");
        *out = nullptr;
        int rc = sedi::ensure_device();
        if (rc) return rc;
        sedi::ServeConfig c;
        c.hbm_budget = cfg->hbm_budget_bytes;
        c.precision = cfg->precision;
        c.rep_every = cfg->rep_every > 0 ? cfg->rep_every : 50;
        c.chunk = cfg->chunk > 0 ? cfg->chunk : 8;
        c.max_slots = cfg->max_slots > 0 ? cfg->max_slots : 64;
        c.hist_cap = cfg->max_iter > 0 ? cfg->max_iter : 10000;
        sedi_server* h = new sedi_server();
        if ((rc = h->impl.start(c))) {
            delete h;
            return rc;
        }
        *out = h;
        return SEDI_OK;
    });
}

int64_t sedi_server_submit(sedi_server* srv, const sedi_job* job) {
    int64_t id = -1;
    const int rc = guarded([&]() -> int {
        SEDI_REQUIRE(srv && job, SEDI_ERR_ARG, "server and job must not be NULL");
        sedi::Shape s;
        int r = check_grid(&job->grid, &s);
        if (r) return r;
        sedi::ServeJob j;
        j.s = s;
        j.a = job->stencil.a;
        j.b = job->stencil.b;
        j.c = job->stencil.c;
        j.kind = job->kind;
        j.samples = job->samples;
        j.tol = job->tol;
        j.maxiter = job->maxiter;
        j.seed = job->seed;
        const long long v = srv->impl.submit(j);
        if (v < 0) return (int)v;
        id = v;
        return SEDI_OK;
    });
    return rc < 0 ? rc : id;
}

int sedi_server_wait(sedi_server* srv, int64_t job, double timeout_s, double* D, int* iters, double* errs,
                     int* codes, double* times4, int* status) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(srv, SEDI_ERR_ARG, "server is NULL");
        sedi::ServeResult r;
        const int rc = srv->impl.wait(job, &r, timeout_s);
        if (rc) return rc;
        if (status) *status = r.status;
        if (r.status == 0) {
            if (D) std::copy(r.D.begin(), r.D.end(), D);
            if (iters) std::copy(r.iters.begin(), r.iters.end(), iters);
            if (errs) std::copy(r.errs.begin(), r.errs.end(), errs);
            if (codes) std::copy(r.codes.begin(), r.codes.end(), codes);
        }
        if (times4) {
            times4[0] = r.t_submit;
            times4[1] = r.t_admit;
            times4[2] = r.t_start;
            times4[3] = r.t_done;
        }
        return SEDI_OK;
    });
}

int sedi_server_stats(sedi_server* srv, sedi_server_info* out) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(srv && out, SEDI_ERR_ARG, "server and out must not be NULL");
        const sedi::ServeStats s = srv->impl.stats();
        out->submitted = s.submitted;
        out->done = s.done;
        out->probes = s.probes;
        out->bytes_in_use = s.bytes_in_use;
        out->peak_bytes = s.peak_bytes;
        out->pools = s.pools;
        return SEDI_OK;
    });
}

void sedi_server_destroy(sedi_server* srv) { delete srv; }

int sedi_h5_available(void) {
    return guarded([&]() -> int { return sedi::h5_available() ? 1 : 0; });
}

int sedi_h5_read(const char* path, const char* dataset, const sedi_grid* grid, double* out) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(path && dataset && out, SEDI_ERR_ARG, "path, dataset and out must not be NULL");
        sedi::Shape s;
        int rc = check_grid(grid, &s);
        if (rc) return rc;
        return sedi::h5_read(path, dataset, s, out);
    });
}

int sedi_h5_write(const char* path, int count, const char* const* datasets, const double* const* arrays,
                  const sedi_grid* grid) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(path && count >= 0 && (count == 0 || (datasets && arrays)), SEDI_ERR_ARG,
                     "bad archive arguments");
        sedi::Shape s;
        int rc = check_grid(grid, &s);
        if (rc) return rc;
        for (int k = 0; k < count; k++)
            SEDI_REQUIRE(datasets[k] && arrays[k], SEDI_ERR_ARG, "dataset %d is NULL", k);
        return sedi::h5_write(path, count, datasets, arrays, s);
    });
}

int sedi_estimator_create(const sedi_estimate_args* a, sedi_estimator** out) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(a && out, SEDI_ERR_ARG, "args and out must not be NULL");
        *out = nullptr;
        if (a->precision == SEDI_FP64 && a->batch_k == 1) {
            // the plain path, chosen explicitly
            sedi::Shape s;
            int rc = check_grid(&a->grid, &s);
            if (rc) return rc;
            if ((rc = check_plain_size(s))) return rc;
        }
        int rc = sedi::ensure_device();
        if (rc) return rc;
        sedi_estimator* h = new sedi_estimator();
        rc = h->impl.create(*a);
        if (rc) {
            delete h;
            return rc;
        }
        *out = h;
        return SEDI_OK;
    });
}

int sedi_estimator_set_log(sedi_estimator* est, sedi_log_fn fn, void* user) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(est, SEDI_ERR_ARG, "estimator is NULL");
        est->impl.set_log(fn, user);
        return SEDI_OK;
    });
}

int sedi_estimator_run(sedi_estimator* est, double* D, int* iters, double* errs, int* codes,
                       double* seconds) {
    return guarded([&]() -> int {
        SEDI_REQUIRE(est, SEDI_ERR_ARG, "estimator is NULL");
        return est->impl.run(D, iters, errs, codes, seconds);
    });
}

void sedi_estimator_destroy(sedi_estimator* est) { delete est; }

int sedi_estimate(const sedi_estimate_args* a, double* D, int* iters, double* errs, int* codes,
                  double* seconds) {
    sedi_estimator* est = nullptr;
    int rc = sedi_estimator_create(a, &est);
    if (rc) return rc;
    rc = sedi_estimator_run(est, D, iters, errs, codes, seconds);
    sedi_estimator_destroy(est);
    return rc;
}

}  // extern "C"
