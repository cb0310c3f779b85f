// serve.hip -- the continuous-batching estimation server (see serve.h).
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>

#include "profiler.h"
#include "rng.h"

namespace sedi {

__global__ void serve_fill_kernel(double* x, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

__global__ void serve_divide_kernel(const double* __restrict__ t, const double* __restrict__ q,
                                    double* __restrict__ D, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        D[i] = t[i] / q[i];  // stochastic-estimator.cc:216
}

static dim3 serve_grid(long long n) {
    long long b = (n + 255) / 256;
    return dim3((unsigned)std::max<long long>(1, std::min<long long>(b, 8192)));
}

// A job's progress.  Blocks of 4 probes go out in probe order; `finished`
// holds blocks done but not yet accumulated (an earlier block is still
// running), so t and q see the probes in order.
struct Server::Job {
    long long id = 0;
    ServeJob spec;
    Pool* pool = nullptr;
    int next_k = 0, acc_k = 0;
    std::map<int, int> finished;  // first probe of the block -> slot
    ProbeSource probes;
    DevBuf tq;
    double bytes = 0;
    ServeResult res;
    bool admitted = false, done = false, started = false;
};

// A BatchSolver of one (grid, solver) with its slot bookkeeping.  Two chunks
// of iterations may be in flight; each posts the block done flags and the
// per-RHS results into its own pinned snapshot.
struct Server::Pool {
    Shape s{};
    double a = 6, b = 1, c = 1;
    int kind = SEDI_CG;
    std::unique_ptr<BatchSolver> bs;
    int nrb = 0;
    std::vector<Job*> slot_job;
    std::vector<int> slot_k0, slot_nact;
    std::vector<char> running;
    std::vector<long long> slot_from;  // first status post that can show the block
    int nrunning = 0;
    long long posts = 0;               // status posts issued
    long long snap_post[2] = {0, 0};   // the post number of each snapshot
    void* snap[2] = {nullptr, nullptr};  // pinned status snapshots
    hipEvent_t ev[2] = {nullptr, nullptr};
    int inflight = 0, head = 0;  // chunks posted, oldest snapshot index
    int chunk = 8;               // iterations per step
    double bytes = 0;
    DevBuf dtmp;
    ~Pool() {
        if (bs) (void)hipStreamSynchronize(bs->stream());
        for (int i = 0; i < 2; i++) {
            if (snap[i]) (void)hipHostFree(snap[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
    }
    bool same(const ServeJob& j) const {
        return j.s.n1 == s.n1 && j.s.n2 == s.n2 && j.s.n3 == s.n3 && j.kind == kind && j.a == a && j.b == b &&
               j.c == c;
    }
};

Server::Server() = default;
Server::~Server() { stop(); }

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int Server::start(const ServeConfig& cfg) {
    SEDI_REQUIRE(!started_, SEDI_ERR_STATE, "server already started");
    SEDI_REQUIRE(cfg.hbm_budget > 0 && cfg.chunk > 0 && cfg.max_slots >= 4 && cfg.hist_cap > 0, SEDI_ERR_ARG,
                 "bad server configuration");
    SEDI_REQUIRE(cfg.precision == SEDI_FP64 || cfg.precision == SEDI_FP32_MIXED, SEDI_ERR_ARG,
                 "unknown precision %d", cfg.precision);
    cfg_ = cfg;
    cfg_.max_slots = cfg.max_slots / 4 * 4;
    int dev = 0;
    SEDI_HIP(hipGetDevice(&dev));
    t0_ = std::chrono::steady_clock::now();
    started_ = true;
    thread_ = std::thread([this, dev]() {
        (void)hipSetDevice(dev);
        engine();
    });
    return SEDI_OK;
}

void Server::stop() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!started_) return;
        stop_ = true;
    }
    cv_submit_.notify_all();
    if (thread_.joinable()) thread_.join();
    pools_.clear();
    std::lock_guard<std::mutex> lk(mu_);
    started_ = false;
}

long long Server::submit(const ServeJob& j) {
    if (j.s.n1 <= 0 || j.s.n2 <= 0 || j.s.n3 <= 0 || j.samples < 1 || j.maxiter < 0 ||
        j.maxiter > cfg_.hist_cap || !(j.tol >= 0) ||
        !(j.kind == SEDI_CG || j.kind == SEDI_PCG || j.kind == SEDI_BICGSTAB)) {
        set_error("bad job (grid %d x %d x %d, kind %d, samples %d, maxiter %d)", j.s.n1, j.s.n2, j.s.n3, j.kind,
                  j.samples, j.maxiter);
        return SEDI_ERR_ARG;
    }
    if ((double)j.s.n() * 16 >= 4294967296.0) {
        set_error("grid too large for the batched path");
        return SEDI_ERR_UNSUPPORTED;
    }
    std::unique_ptr<Job> job(new Job());
    job->spec = j;
    job->res.t_submit = now();
    long long id;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!started_ || stop_) {
            set_error("server is not running");
            return SEDI_ERR_STATE;
        }
        if (engine_rc_) {
            set_error("server engine failed (%d)", engine_rc_);
            return engine_rc_;
        }
        id = next_id_++;
        job->id = id;
        incoming_.push_back(job.get());
        jobs_[id] = std::move(job);
    }
    cv_submit_.notify_all();
    return id;
}

int Server::wait(long long id, ServeResult* out, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) {
        set_error("no job %lld (or its result was already taken)", id);
        return SEDI_ERR_ARG;
    }
    Job* j = it->second.get();
    auto ready = [&]() { return j->done; };
    if (timeout_s < 0) cv_done_.wait(lk, ready);
    else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return 1;
    if (out) *out = std::move(j->res);
    const int st = j->res.status;
    jobs_.erase(it);
    if (st < 0 && !out) return st;
    return SEDI_OK;
}

ServeStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ServeStats s;
    s.submitted = next_id_;
    s.done = n_done_;
    s.probes = n_probes_;
    s.bytes_in_use = bytes_;
    s.peak_bytes = peak_bytes_;
    s.pools = npools_;
    return s;
}

// ---------------------------------------------------------------------------
// engine (one thread owns the device)
// ---------------------------------------------------------------------------
void Server::fail_all(int rc) {
    std::lock_guard<std::mutex> lk(mu_);
    engine_rc_ = rc;
    for (auto& kv : jobs_) {
        Job& j = *kv.second;
        if (!j.done) {
            j.res.status = rc;
            j.done = true;
        }
    }
    cv_done_.notify_all();
}

void Server::engine() {
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            while (!incoming_.empty()) {
                pending_.push_back(incoming_.front());
                incoming_.pop_front();
            }
            if (pending_.empty() && admitted_.empty()) {
                if (stop_) break;
                cv_submit_.wait(lk, [&]() { return stop_ || !incoming_.empty(); });
                continue;
            }
        }
        bool progress = false;
        const int rc = engine_step(&progress);
        if (rc) {
            fail_all(rc);
            break;
        }
        if (!progress) {
            // nothing to do until some pool's chunk completes (whichever is
            // first: pools of small grids step much faster than large ones),
            // a job arrives, or memory frees up
            bool busy = false;
            for (auto& p : pools_) busy = busy || p->inflight > 0;
            std::unique_lock<std::mutex> lk(mu_);
            if (busy) cv_submit_.wait_for(lk, std::chrono::microseconds(50));
            else cv_submit_.wait_for(lk, std::chrono::milliseconds(2));
        }
    }
}

int Server::engine_step(bool* progress) {
    int rc = admit();
    if (rc) return rc;
    for (auto& p : pools_) {
        const int before = p->nrunning;
        if ((rc = fill(*p))) return rc;
        if (p->nrunning != before) *progress = true;
    }
    for (auto& p : pools_) {
        // keep two chunks in flight while blocks run
        while (p->nrunning > 0 && p->inflight < 2) {
            if ((rc = p->bs->step(p->chunk))) return rc;
            const int slot = (p->head + p->inflight) & 1;
            if ((rc = p->bs->post_status(p->snap[slot], p->ev[slot]))) return rc;
            p->snap_post[slot] = p->posts++;
            p->inflight++;
            *progress = true;
        }
    }
    for (auto& p : pools_) {
        while (p->inflight > 0) {
            // not ready: poll again on the next step; any other result is
            // an error that no later query would report (two chunks are
            // already in flight, so nothing new is launched), so the engine
            // stops and fails every job
            const hipError_t q = hipEventQuery(p->ev[p->head]);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) {
                set_error("serving pool chunk failed: %s", hipGetErrorString(q));
                return SEDI_ERR_HIP;
            }
            if ((rc = harvest(*p))) return rc;
            p->head ^= 1;
            p->inflight--;
            *progress = true;
        }
    }
    return SEDI_OK;
}

// Pool for a job's (grid, solver); creating it takes K slots sized so a pool
// holds at most an eighth of the budget.
Server::Pool* Server::pool_for(const ServeJob& j, bool create, int* rc) {
    *rc = SEDI_OK;
    for (auto& p : pools_)
        if (p->same(j)) return p.get();
    if (!create) return nullptr;
    const bool fp32 = cfg_.precision == SEDI_FP32_MIXED;
    const double per4 = BatchSolver::bytes_needed(j.s, j.kind, 4, fp32) + 4.0 * cfg_.hist_cap * 8;
    // at most an eighth of the budget per pool, so pools of several grids and
    // solvers fit together
    int K = (int)std::min<double>(cfg_.max_slots, std::floor(cfg_.hbm_budget / 8 / per4) * 4);
    K = std::max(4, K / 4 * 4);
    const double need = per4 * K / 4;
    if (bytes_ + need > cfg_.hbm_budget && !evict_idle(need, nullptr)) return nullptr;
    std::unique_ptr<Pool> p(new Pool());
    p->s = j.s;
    p->a = j.a; p->b = j.b; p->c = j.c;
    p->kind = j.kind;
    p->bs.reset(new BatchSolver());
    if ((*rc = p->bs->create(j.s, j.a, j.b, j.c, j.kind, K, fp32, cfg_.rep_every))) return nullptr;
    if ((*rc = p->bs->slots_open(cfg_.hist_cap))) return nullptr;
    p->nrb = K / 4;
    // steps of similar duration across grids: more iterations per step on
    // small grids, whose iterations take microseconds
    const long long n = j.s.n();
    p->chunk = cfg_.chunk * (n >= (8LL << 20) ? 1 : n >= (1LL << 20) ? 2 : n >= (1LL << 18) ? 4 : 8);
    p->slot_job.assign(p->nrb, nullptr);
    p->slot_k0.assign(p->nrb, 0);
    p->slot_nact.assign(p->nrb, 0);
    p->running.assign(p->nrb, 0);
    p->slot_from.assign(p->nrb, 0);
    for (int i = 0; i < 2; i++) {
        if (hipHostMalloc(&p->snap[i], p->bs->status_bytes(), hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev[i], hipEventDisableTiming) != hipSuccess) {
            set_error("pool set-up failed");
            *rc = SEDI_ERR_HIP;
            return nullptr;
        }
    }
    p->bytes = need;
    {
        std::lock_guard<std::mutex> lk(mu_);
        bytes_ += need;
        peak_bytes_ = std::max(peak_bytes_, bytes_);
        npools_++;
    }
    pools_.push_back(std::move(p));
    return pools_.back().get();
}

// Frees idle pools (no running block, no admitted job) other than `keep`
// until `need` fits.
bool Server::evict_idle(double need, const Pool* keep) {
    for (size_t i = 0; i < pools_.size() && bytes_ + need > cfg_.hbm_budget;) {
        Pool* p = pools_[i].get();
        bool used = p == keep || p->nrunning > 0 || p->inflight > 0;
        for (Job* j : admitted_)
            if (j->pool == p) used = true;
        if (used) {
            i++;
            continue;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            bytes_ -= p->bytes;
            npools_--;
        }
        pools_.erase(pools_.begin() + i);
    }
    return bytes_ + need <= cfg_.hbm_budget;
}

// Admits waiting jobs in arrival order while the budget allows.
// Admits waiting jobs while the budget allows: in arrival order, except that
// a job that does not fit yet lets later jobs that do fit go ahead, until it
// has waited `kAgeLimit` seconds (then strict order, so it cannot starve).
static constexpr double kAgeLimit = 20.0;

int Server::admit() {
    bool blocked = false;  // an older job is waiting for memory
    for (size_t i = 0; i < pending_.size();) {
        Job* j = pending_[i];
        if (blocked && now() - pending_.front()->res.t_submit > kAgeLimit) return SEDI_OK;
        int rc = SEDI_OK;
        Pool* p = pool_for(j->spec, true, &rc);
        if (rc) return rc;
        const long long n = j->spec.s.n();
        const double need = 16.0 * n;  // t and q
        if (p && bytes_ + need > cfg_.hbm_budget) evict_idle(need, p);
        if (!p || bytes_ + need > cfg_.hbm_budget) {
            // it waits for memory, unless nothing else holds any: then it never fits
            bool idle = admitted_.empty();
            for (auto& q : pools_) idle = idle && q->nrunning == 0 && q->inflight == 0;
            if (idle) {
                if (p) evict_idle(1e300, p);
                if (!p || bytes_ + need > cfg_.hbm_budget) {
                    set_error("job %lld does not fit the device budget", j->id);
                    fail_job(*j, SEDI_ERR_NOMEM);
                    pending_.erase(pending_.begin() + i);
                    continue;
                }
            } else {
                blocked = true;
                i++;
                continue;
            }
        }
        hipStream_t st = p->bs->stream();
        SEDI_HIP(j->tq.reserve(sizeof(double) * 2 * n));
        SEDI_HIP(launch("serve_fill_kernel", serve_fill_kernel, serve_grid(2 * n), dim3(256), st,
                        j->tq.as<double>(), 2 * n, 0.0));
        if ((rc = j->probes.init(j->spec.seed, n, 0, j->spec.samples, st))) return rc;
        j->pool = p;
        j->bytes = need;
        j->admitted = true;
        j->res.iters.assign(j->spec.samples, 0);
        j->res.codes.assign(j->spec.samples, -1);
        j->res.errs.assign(j->spec.samples, 0.0);
        {
            std::lock_guard<std::mutex> lk(mu_);
            bytes_ += need;
            peak_bytes_ = std::max(peak_bytes_, bytes_);
            j->res.t_admit = now();
        }
        admitted_.push_back(j);
        pending_.erase(pending_.begin() + i);
    }
    return SEDI_OK;
}

// Fills free slots of a pool with the next blocks of its admitted jobs,
// oldest job first.
int Server::fill(Pool& p) {
    const bool fp32 = cfg_.precision == SEDI_FP32_MIXED;
    for (Job* j : admitted_) {
        if (j->pool != &p) continue;
        while (j->next_k < j->spec.samples) {
            int rb = -1;
            for (int i = 0; i < p.nrb; i++)
                if (!p.slot_job[i]) {
                    rb = i;
                    break;
                }
            if (rb < 0) return SEDI_OK;
            const int nact = std::min(4, j->spec.samples - j->next_k);
            hipStream_t st = p.bs->stream();
            int rc = j->probes.emit(j->next_k, nact, 4, fp32, p.bs->block_rhs(rb), st);
            if (!rc) rc = p.bs->start_block(rb, j->spec.tol, j->spec.maxiter);
            if (rc) return rc;
            p.slot_job[rb] = j;
            p.slot_k0[rb] = j->next_k;
            p.slot_nact[rb] = nact;
            p.running[rb] = 1;
            // a snapshot posted before this start still shows the slot's
            // previous block as done
            p.slot_from[rb] = p.posts;
            p.nrunning++;
            if (!j->started) {
                j->started = true;
                j->res.t_start = now();
            }
            j->next_k += nact;
        }
    }
    return SEDI_OK;
}

// Reads the oldest snapshot of a pool: blocks done in it are taken off the
// running set, and every job's finished blocks are accumulated in order.
int Server::harvest(Pool& p) {
    const void* snap = p.snap[p.head];
    const int* flags = p.bs->snap_block_done(snap);
    const int* iters = p.bs->snap_iters(snap);
    const int* codes = p.bs->snap_codes(snap);
    const double* errs = p.bs->snap_errs(snap);
    const long long post = p.snap_post[p.head];
    for (int rb = 0; rb < p.nrb; rb++) {
        if (!p.running[rb] || !flags[rb] || post < p.slot_from[rb]) continue;
        p.running[rb] = 0;
        p.nrunning--;
        Job* j = p.slot_job[rb];
        j->finished[p.slot_k0[rb]] = rb;
        // a done block's results stay fixed: take them from this snapshot
        for (int q = 0; q < p.slot_nact[rb]; q++) {
            j->res.iters[p.slot_k0[rb] + q] = iters[4 * rb + q];
            j->res.codes[p.slot_k0[rb] + q] = codes[4 * rb + q];
            j->res.errs[p.slot_k0[rb] + q] = errs[4 * rb + q];
        }
    }
    for (size_t i = 0; i < admitted_.size();) {
        Job* j = admitted_[i];
        if (j->pool != &p) {
            i++;
            continue;
        }
        int rc = SEDI_OK;
        for (auto it = j->finished.find(j->acc_k); it != j->finished.end(); it = j->finished.find(j->acc_k)) {
            const int rb = it->second, nact = p.slot_nact[rb];
            double* t = j->tq.as<double>();
            if ((rc = p.bs->accumulate(t, t + j->spec.s.n(), 4 * rb, 4 * rb + nact))) return rc;
            p.slot_job[rb] = nullptr;  // later work on the stream is ordered after the accumulation
            j->acc_k += nact;
            j->finished.erase(it);
        }
        if (j->acc_k == j->spec.samples) {
            if ((rc = complete(*j))) return rc;
            admitted_.erase(admitted_.begin() + i);
            continue;
        }
        i++;
    }
    return SEDI_OK;
}

void Server::fail_job(Job& j, int rc) {
    std::lock_guard<std::mutex> lk(mu_);
    j.res.status = rc;
    j.res.t_done = now();
    j.done = true;
    cv_done_.notify_all();
}

int Server::complete(Job& j) {
    Pool& p = *j.pool;
    const long long n = j.spec.s.n();
    hipStream_t st = p.bs->stream();
    SEDI_HIP(p.dtmp.reserve(sizeof(double) * n));
    const double* t = j.tq.as<double>();
    SEDI_HIP(launch("serve_divide_kernel", serve_divide_kernel, serve_grid(n), dim3(256), st, t, t + n,
                    p.dtmp.as<double>(), n));
    j.res.D.resize(n);
    SEDI_HIP(hipMemcpyAsync(j.res.D.data(), p.dtmp.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    SEDI_HIP(hipStreamSynchronize(st));
    j.tq.release();
    std::lock_guard<std::mutex> lk(mu_);
    bytes_ -= j.bytes;
    // every probe was accumulated, converged or not (the reference's -i);
    // the per-probe codes carry the solver outcomes
    j.res.status = 0;
    j.res.t_done = now();
    j.done = true;
    n_done_++;
    n_probes_ += j.spec.samples;
    cv_done_.notify_all();
    return SEDI_OK;
}

}  // namespace sedi
