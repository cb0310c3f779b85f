// serve.h -- continuous-batching estimation server on one GPU.
//
// Jobs (grid, solver, samples, tolerance, seed) arrive at any time.  Each
// (grid, solver) pair has a pool: a BatchSolver whose RHS blocks (4 probes)
// are independent slots with their own iteration counters, so a block of a
// newly admitted job starts while other blocks of the pool are mid-solve.
// A single engine thread owns the device: it admits jobs against the HBM
// budget, fills free slots with the next probes of the admitted jobs (first
// come, first served), advances every pool by a chunk of iterations on the
// pool's own stream (pools of different grids run concurrently), and harvests
// finished blocks.  A job's blocks are accumulated into its t, q strictly in
// probe order, so its D is bitwise the one the batched estimator computes for
// the same job, whatever else is in flight.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batched.h"
#include "runtime.h"
#include "sedi_internal.h"

namespace sedi {

struct ServeConfig {
    double hbm_budget = 200e9;  // bytes of device memory for pools and jobs
    int precision = SEDI_FP32_MIXED;
    int rep_every = 50;
    int chunk = 8;          // iterations per step of a pool
    int max_slots = 64;     // right-hand sides per pool at most
    int hist_cap = 10000;   // largest maxiter a job may ask for
};

struct ServeJob {
    Shape s{};
    double a = 6.0, b = 1.0, c = 1.0;
    int kind = SEDI_CG;
    int samples = 8;
    double tol = 1e-8;
    int maxiter = 2000;
    uint32_t seed = 0;
};

struct ServeResult {
    int status = 0;  // 0 done, < 0 failed (error code)
    std::vector<double> D;
    std::vector<int> iters, codes;
    std::vector<double> errs;
    double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;  // seconds since start()
};

struct ServeStats {
    long long submitted = 0, done = 0, probes = 0;
    double bytes_in_use = 0, peak_bytes = 0;
    int pools = 0;
};

class Server {
  public:
    Server();
    ~Server();
    int start(const ServeConfig& cfg);
    // Thread-safe.  Returns the job id (>= 0) or a negative error code.
    long long submit(const ServeJob& job);
    // Waits for a job; 0 done (result moved out), 1 timeout, < 0 error.
    int wait(long long id, ServeResult* out, double timeout_s);
    ServeStats stats();
    void stop();

  private:
    struct Pool;
    struct Job;
    void engine();
    int engine_step(bool* progress);
    int admit();
    int fill(Pool& p);
    int harvest(Pool& p);
    int complete(Job& j);
    void fail_job(Job& j, int rc);
    Pool* pool_for(const ServeJob& s, bool create, int* rc);
    bool evict_idle(double need, const Pool* keep);
    double now() const;
    void fail_all(int rc);

    ServeConfig cfg_;
    std::mutex mu_;
    std::condition_variable cv_submit_, cv_done_;
    std::deque<Job*> incoming_;
    std::map<long long, std::unique_ptr<Job>> jobs_;
    std::deque<Job*> pending_;    // engine only: submitted, not admitted
    std::vector<Job*> admitted_;  // engine only: admitted, not complete (FIFO)
    std::vector<std::unique_ptr<Pool>> pools_;
    long long next_id_ = 0;
    bool stop_ = false;
    bool started_ = false;
    int engine_rc_ = 0;
    double bytes_ = 0, peak_bytes_ = 0;
    int npools_ = 0;
    long long n_done_ = 0, n_probes_ = 0;
    std::thread thread_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace sedi
