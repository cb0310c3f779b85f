// batched.h -- multi-RHS Krylov solvers over blocks of 4 probes (FP64 or
// FP32 storage): CG, pipelined CG and BiCGStab with per-RHS scalars.
#pragma once

#include <hip/hip_runtime.h>
#include <vector>

#include "march.h"
#include "runtime.h"
#include "sedi_internal.h"

namespace sedi {

// Per-RHS solver results of a batch.
struct BatchResult {
    std::vector<int> iters, codes;
    std::vector<double> errs;
};

struct BRun;

// vector slots a batch can hold (b, x_lo and a solver's vectors)
constexpr int BATCH_MAX_SLOTS = 16;

class BatchSolver {
  public:
    ~BatchSolver();
    // kind: SEDI_CG, SEDI_PCG or SEDI_BICGSTAB.  K: right-hand sides per batch
    // (a multiple of 4); fp32: FP32 vectors with FP64 reductions and residual
    // replacement every rep_every iterations.
    int create(const Shape& s, double a, double b, double c, int kind, int K, bool fp32, int rep_every);
    // Solves the K systems A x_k = b_k from x = 0, where b is the batch layout
    // [K/4][N][4] of the storage type (device).  Right-hand sides that are
    // all zero are done at once (padding).  Synchronises the stream.
    int solve(int maxiter, double tol, BatchResult* res);
    // t += x_k v_k, q += v_k^2 for k = k0 .. k1-1 in probe order (the
    // reference's accumulation, stochastic-estimator.cc:195,198).
    int accumulate(double* t, double* q, int k0, int k1);
    // b: kact right-hand sides as [kact][N] FP64 (device), rounded to the
    // storage type; slots kact..K-1 are zeroed.  With `replicate`, b is one
    // right-hand side used for all kact slots.  Asynchronous on stream().
    int load_rhs(const double* b, int kact, bool replicate = false);
    // x: solutions k0 .. k1-1 as [k1-k0][N] FP64 (device; x_hi + x_lo in the
    // mixed form).  Asynchronous on stream().
    int store_x(double* x, int k0, int k1);
    // v: right-hand sides k0 .. k1-1 as stored (rounded to the storage type).
    int store_rhs(double* v, int k0, int k1);
    void* rhs() const { return v_[0]; }
    bool fp32() const { return fp32_; }
    int K() const { return K_; }
    int kind() const { return kind_; }
    long long n() const { return shape_.n(); }
    hipStream_t stream() const { return stream_; }
    std::vector<double> history(int k) const;
    // Device bytes of the vectors of a batch of K right-hand sides on grid s.
    static double bytes_needed(const Shape& s, int kind, int K, bool fp32);

    // ---- block-level control (continuous batching) -------------------------
    // The K/4 RHS blocks are independent slots: each keeps its own iteration
    // counter, so blocks can be started while others are mid-solve.
    // Opens slot mode: every block free; histories of up to hist_cap
    // iterations.  Synchronous.
    int slots_open(int hist_cap);
    // b of block rb (4 right-hand sides in the storage type), for the probe
    // generator.
    void* block_rhs(int rb) const;
    // Starts block rb from x = 0 with its b in place (asynchronous).
    int start_block(int rb, double tol, int maxiter);
    // Launches `iters` iterations of every running block (asynchronous).
    int step(int iters);
    // Queues a copy of the status (block done flags and per-RHS results) into
    // host memory `snap` of status_bytes() bytes (pinned) and records `ev`
    // after it; the snap_* accessors read it once ev has completed.
    size_t status_bytes() const;
    int post_status(void* snap, hipEvent_t ev);
    const int* snap_block_done(const void* snap) const;
    const int* snap_iters(const void* snap) const;
    const int* snap_codes(const void* snap) const;
    const double* snap_errs(const void* snap) const;
    int nblocks() const { return nrb_; }
    // Lanes moved by compaction in the last solve() (see solve_t).
    long long lane_moves() const { return moves_last_; }

  private:
    template <class T> int bind_t(int hist_need);
    template <class T> int iterate_t(int count);
    template <class T> int start_t(int rb0, int count, double tol, int maxiter);
    template <class T> int solve_t(int maxiter, double tol, BatchResult* res);
    template <class T> int accumulate_t(double* t, double* q, int k0, int k1);
    template <class T> int pack_t(const double* b, int kact, long long row_stride);
    template <class T> int unpack_t(double* x, int k0, int k1, bool rhs);
    template <class T> int compact_t();
    int store_rows(double* x, int k0, int k1, bool rhs);
    unsigned read_moves();
    Shape shape_{};
    double a_ = 6.0, b_c_ = 1.0, c_ = 1.0;
    int kind_ = 0, K_ = 0, nrb_ = 0, nvec_ = 0;
    bool fp32_ = true;
    int rep_every_ = 50;
    hipStream_t stream_ = nullptr;
    DevBuf vec_, ctl_, hist_;
    void* v_[BATCH_MAX_SLOTS] = {};
    double* xhi_ = nullptr;
    int* pinned_ = nullptr;
    int hist_cap_ = 0;
    std::vector<double> hist_host_;
    BRun* run_ = nullptr;
    // the captured chunk of iterations (count, executable graph)
    int graph_count_ = 0, graph_compact_ = 0;
    hipGraphExec_t graph_ = nullptr;
    bool geometry_ready_ = false;
    // lane_of_[k]: the lane that holds right-hand side k (compaction moves
    // right-hand sides between lanes during a solve; identity otherwise)
    std::vector<int> lane_of_;
    long long moves_last_ = 0;
};

// Lane compaction of batch solves (on by default); no effect on results.
bool batch_compact_enabled();
void set_batch_compact_enabled(bool on);

}  // namespace sedi
