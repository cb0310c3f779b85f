// batched_common.h -- what the batched CG, pipelined CG and BiCGStab passes
// share: the 4-RHS value types, the per-RHS control block, launch geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "batched.h"
#include "march.h"
#include "profiler.h"
#include "tuning.h"

#ifndef SEDI_PREFETCH_BATCH
#define SEDI_PREFETCH_BATCH 1
#endif
// The CG update pass's iterations that leave x alone: two planes of loads in
// flight (at 256^3, k = 64: headline 12.555 -> 12.60 probe-solves/s; with two
// for every batched pass, x steps included, the update pass ran 6% slower)
#ifndef SEDI_PREFETCH_CG_UPD
#define SEDI_PREFETCH_CG_UPD 2
#endif
#ifndef SEDI_BICG_RESET
#define SEDI_BICG_RESET 1
#endif
#ifndef SEDI_BICG_DROP
#define SEDI_BICG_DROP 0.1  // BiCGStab (mixed): replace and restart at each tenfold decrease
#endif


namespace sedi {

// ---- 4-wide values: one cell of one RHS block ------------------------------
template <class T>
struct alignas(4 * sizeof(T)) V4 {
    T v[4];
    __device__ __forceinline__ T& operator[](int i) { return v[i]; }
    __device__ __forceinline__ const T& operator[](int i) const { return v[i]; }
};
template <class T>
__device__ __forceinline__ V4<T> operator+(const V4<T>& a, const V4<T>& b) {
    return {{a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]}};
}
template <class T>
__device__ __forceinline__ V4<T> operator-(const V4<T>& a, const V4<T>& b) {
    return {{a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3]}};
}
template <class T>
__device__ __forceinline__ V4<T> operator*(T s, const V4<T>& a) {
    return {{s * a[0], s * a[1], s * a[2], s * a[3]}};
}
template <class T>
__device__ __forceinline__ V4<T> operator*(const V4<T>& a, const V4<T>& b) {
    return {{a[0] * b[0], a[1] * b[1], a[2] * b[2], a[3] * b[3]}};
}
template <class T>
__device__ __forceinline__ V4<double> to_f64(const V4<T>& a) {
    return {{(double)a[0], (double)a[1], (double)a[2], (double)a[3]}};
}
template <class T>
__device__ __forceinline__ V4<T> from_f64(const V4<double>& a) {
    return {{(T)a[0], (T)a[1], (T)a[2], (T)a[3]}};
}

// Arithmetic type of the pipelined CG and BiCGStab passes for storage type T.
// With FP32 storage they compute in FP64 registers: the FP32 stencil loses
// about eps32 * cond(A) of A p on smooth fields (cancellation in 6 p0 - sum of
// neighbours), which CG's recurrences absorb but pipelined CG's matrix-vector
// recurrences and BiCGStab's bi-orthogonality do not.  Memory traffic is
// unchanged (loads and stores stay FP32).
#ifndef SEDI_MIXED_COMPUTE_F64
#define SEDI_MIXED_COMPUTE_F64 1
#endif
template <class T>
struct ComputeOf { using type = T; };
#if SEDI_MIXED_COMPUTE_F64
template <>
struct ComputeOf<float> { using type = double; };
#endif
template <class T>
using compute_t = typename ComputeOf<T>::type;

template <class C, class T>
__device__ __forceinline__ V4<C> cvt(const V4<T>& a) {
    return {{(C)a[0], (C)a[1], (C)a[2], (C)a[3]}};
}

// Two FP64 fields marched together (the replacement passes of pipelined CG
// apply the stencil to two vectors in one sweep).
struct D2 {
    V4<double> a, b;
};
__device__ __forceinline__ D2 operator+(const D2& x, const D2& y) { return {x.a + y.a, x.b + y.b}; }
__device__ __forceinline__ D2 operator-(const D2& x, const D2& y) { return {x.a - y.a, x.b - y.b}; }
__device__ __forceinline__ D2 operator*(double s, const D2& x) { return {s * x.a, s * x.b}; }

// Store of a 4-RHS value, non-temporal by default (the vectors a pass writes
// are not read again until the next pass, and a batch's vectors exceed the
// caches: measured 1% faster on the 256^3 k=64 CG estimator, 2.7% on its
// direction pass).  SEDI_BATCH_NT=0 selects ordinary stores.
#ifndef SEDI_BATCH_NT
#define SEDI_BATCH_NT 1
#endif
template <class T>
__device__ __forceinline__ void stv(V4<T>* base, cidx_t i, const V4<T>& v) {
#if SEDI_BATCH_NT
    typedef T vec4 __attribute__((ext_vector_type(4)));
    const vec4 w = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(
        w, reinterpret_cast<vec4*>(reinterpret_cast<char*>(base) + (unsigned)(i * (unsigned)sizeof(V4<T>))));
#else
    stc(base, i, v);
#endif
}

// A wave-uniform value moved to scalar registers (the per-block coefficients
// a pass multiplies by: kept in vector registers they would take 4 per value).
__device__ __forceinline__ float uniform(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Per-lane select: a where m[k], else b.
template <class T>
__device__ __forceinline__ V4<T> sel(const int (&m)[4], const V4<T>& a, const V4<T>& b) {
    return {{m[0] ? a[0] : b[0], m[1] ? a[1] : b[1], m[2] ? a[2] : b[2], m[3] ? a[3] : b[3]}};
}

// ---- device control --------------------------------------------------------
// Per-RHS scalars live in device memory and are updated by the last
// workgroup of each pass of the RHS's block (grid_sum), like the plain
// solvers' DevState.
struct BCtl {
    // per RHS (K entries each)
    double *bb, *ru, *ru_old, *alpha, *beta, *err, *ferr, *elast, *hist;
    double *gamma, *delta, *errp, *gamma_prev, *alpha_prev;  // pipelined CG
    double *rho, *rho_prev, *omega, *ss, *cbw;               // BiCGStab
    double* apend;  // CG: coefficients of the pending p1 .. p_{CG_NPEND} not yet in x, [CG_NPEND][K]
    int *fiter, *code, *done, *rep, *fold, *sexit;
    int* plane;  // per lane: the right-hand side (index at load) it holds
    // per RHS block
    int *iter, *rb_done;
    double* tolv;  // per RHS: tolerance
    int* maxv;     // per RHS: iteration cap
    double* partials;
    unsigned* counters;
    int* swaps;       // lane compaction: [count, (lane a, lane b) x count]
    unsigned* moves;  // lane compaction: lanes moved since creation
    double drop;     // replacement when the residual dropped by this factor
    double restart;  // restart when true > restart * recursive residual
    int rep_every, mixed, K, kind;
    int bicg_reset;  // BiCGStab replacement restarts r^ = r and the directions
    long long hstride;  // history entries per RHS
    long long pstride;  // partial sums per RHS block
};

__device__ __forceinline__ void bset_done(const BCtl& c, int k, int code, int it, double e) {
    c.done[k] = 1;
    c.code[k] = code;
    c.fiter[k] = it;
    c.ferr[k] = e;
}

__device__ __forceinline__ void brb_update(const BCtl& c, int rb) {
    c.rb_done[rb] = c.done[4 * rb] & c.done[4 * rb + 1] & c.done[4 * rb + 2] & c.done[4 * rb + 3];
}

// Cg.hpp:196-206 / PipelinedCg.hpp:232-242 when the loop ran out.
__device__ __forceinline__ void bcg_exhausted(const BCtl& c, int k, double e) {
    if (e <= c.tolv[k]) bset_done(c, k, 0, c.maxv[k] + 1, e);
    else bset_done(c, k, 1, c.maxv[k], e);
}

// BiCgStab.hpp:235-263 for RHS k, `iter` = the loop variable when it ended.
__device__ __forceinline__ void bbicg_outcome(const BCtl& c, int k, int iter) {
    const double e = c.err[k];
    if (e <= c.tolv[k]) bset_done(c, k, 0, iter, e);
    else if (c.ss[k] <= c.tolv[k] * c.tolv[k]) bset_done(c, k, 0, iter, sqrt(c.ss[k] / c.bb[k]));
    else if (c.omega[k] == 0.0) bset_done(c, k, 3, iter, e);
    else if (c.rho[k] == 0.0) bset_done(c, k, 2, iter, e);
    else bset_done(c, k, 1, c.maxv[k], e);
}

// ---- host side -------------------------------------------------------------
inline int batch_cus() {
    static int v = 0;
    if (!v) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
    }
    return v;
}

// Geometry of a batched pass, per RHS block: the workgroups resident at once
// (a multiple of 8).  It depends on the grid and the kernel only, not on the
// batch size, so the order of every reduction -- and with it each RHS's
// result -- is the same whichever batch an RHS is solved in.
template <typename Kern>
KGeom bgeom(const char* name, Kern kern16, const Shape& s, int D, int ty = TY) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, int, int>, KGeom> cache;
    const auto key = std::make_tuple(reinterpret_cast<const void*>(kern16), s.n1, s.n2, s.n3);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern16, TX * ty, 0) != hipSuccess || occ <= 0) {
        (void)hipGetLastError();
        occ = 2;
    }
#ifdef SEDI_BATCH_MAXOCC
    occ = std::min(occ, SEDI_BATCH_MAXOCC);
#endif
    occ = std::min(occ, 8);
    const int per = std::max(8, occ * batch_cus() / 8 * 8);
    // the occupancy plan, with the autotuner's z chunk and schedule for this
    // kernel and grid (partials hold 8 workgroups per CU per RHS block)
    const KGeom planned = plan_geometry(s, per, D, ty);
    const int zc = tuned_value(std::string("zc:") + name, s, 0);
    const int sched = tuned_value(std::string("sched:") + name, s, 0);
    return cache.emplace(key, tuned_geometry(planned, s, zc, sched, 8 * batch_cus())).first->second;
}

// Workgroups per RHS block of the elementwise kernels (independent of the
// batch size, like bgeom).
inline unsigned ew_blocks(long long n) {
    const long long g = std::min<long long>((n + NTHREADS - 1) / NTHREADS, 2LL * batch_cus());
    return (unsigned)std::max<long long>(1, g);
}

// kernel<TT, ZC> for the ZC the geometry chose, one wave per tile row
#define SEDI_BZC_LAUNCH(name, K, TT, geo, grid, ...)                                             \
    ((geo).tg.zc == 8    ? launch(name, K<TT, 8>, grid, dim3(TX * (geo).tg.ty), __VA_ARGS__)     \
     : (geo).tg.zc == 32 ? launch(name, K<TT, 32>, grid, dim3(TX * (geo).tg.ty), __VA_ARGS__)    \
                         : launch(name, K<TT, 16>, grid, dim3(TX * (geo).tg.ty), __VA_ARGS__))

// Tile heights of the batched CG passes: the direction pass takes 8 rows in
// FP32 storage (16 measured no faster); the update pass 8 or 16 rows by grid
// (batched.hip, upd_rows).  FP64 blocks, whose registers do not fit two
// 8-wave workgroups per CU, take one (SEDI_FP64_BATCH_TY, batched.hip).
template <class T>
struct CgTile {
    static constexpr int min_waves = sizeof(T) == 4 ? 4 : 1;
};
#ifndef SEDI_CG_DIR_TY
#define SEDI_CG_DIR_TY TY_TALL
#endif
template <class T>
struct CgDirTile {
    static constexpr int ty = sizeof(T) == 4 ? SEDI_CG_DIR_TY : TY;
    static constexpr int min_waves = CgTile<T>::min_waves;
};

// Vector slots of a batch (each [K/4][N] of V4<T>).
enum { SLOT_B = 0, SLOT_XLO = 1, SLOT_K0 = 2 };
// CG: r0, p0, p1, r1 (the residual of iteration it is in r1 when it is odd,
// in r0 when even: the update pass writes its successor to the other buffer)
// and, with solution updates in groups of 4, p2, p3.  The direction of
// iteration it is in p[it % CG_XG].
//
// CG's solution updates are applied in groups of CG_XG (2 or 4) iterations:
// x += alpha p of iterations it-CG_XG+1 .. it in one read-modify-write at the
// iterations it that are multiples of CG_XG, in the order of successive
// updates (so an FP64 x is bitwise what one update per iteration gives).  In
// between, an iteration's alpha waits in c.apend[(j-1)*K + k] for the
// direction in buffer p_j, j = it % CG_XG: the pending directions of a
// right-hand side are always in buffers p1 .. p_{CG_XG-1}.
#ifndef SEDI_CG_XGROUP
#define SEDI_CG_XGROUP 4
#endif
constexpr int CG_XG = SEDI_CG_XGROUP;
constexpr int CG_NPEND = CG_XG - 1;
static_assert(CG_XG == 2 || CG_XG == 4, "CG solution updates come in groups of 2 or 4");
enum { CG_R = SLOT_K0, CG_P0, CG_P1, CG_R1, CG_P2, CG_P3 };
constexpr int CG_NVEC = CG_R1 + 1 + (CG_XG - 2);
__host__ __device__ constexpr int cg_pslot(int j) { return j == 0 ? CG_P0 : j == 1 ? CG_P1 : j == 2 ? CG_P2 : CG_P3; }

// CG's direction buffers p0 .. p_{CG_XG-1} (kernel argument)
template <class T>
struct CgDirs {
    V4<T>* p[4];
    // buffer j for a run-time j (selects: indexing the kernel argument would
    // copy it to scratch memory)
    __device__ V4<T>* at(int j) const { return j == 0 ? p[0] : j == 1 ? p[1] : j == 2 ? p[2] : p[3]; }
};

// The pending directions of CG or pipelined CG, buffers p1 .. p_{CG_NPEND} of
// the batch (all null: BiCGStab, whose x is always current)
template <class T>
struct PendDirs {
    const V4<T>* p[CG_NPEND];
};
// pipelined CG: r, u, w0, w1, z, q, s, and p0 .. p_{CG_XG-1}: its x updates
// are grouped like CG's (the direction of iteration it is in p[it % CG_XG])
enum { PCG_R = SLOT_K0, PCG_U, PCG_W0, PCG_W1, PCG_Z, PCG_Q, PCG_S, PCG_P };
constexpr int PCG_NVEC = PCG_P + CG_XG;
__host__ __device__ constexpr int pcg_pslot(int j) { return PCG_P + j; }
static_assert(PCG_NVEC <= BATCH_MAX_SLOTS && CG_NVEC <= BATCH_MAX_SLOTS, "vector slots of a batch");
// BiCGStab: r0, r1, r^, p0, p1, v0, v1
enum { BI_R0 = SLOT_K0, BI_R1, BI_RH, BI_P0, BI_P1, BI_V0, BI_V1, BI_NVEC };

// What a solver's launch functions need.
struct BRun {
    Shape shape;
    long long n;
    int nrb;
    BCtl c;
    double a, b, cc, im;
    hipStream_t stream;
    void* v[BATCH_MAX_SLOTS];
    double* xhi;  // FP64 part of x ([K][N]; mixed form only)
    bool mixed;
    int nvec;     // vector slots in use (v[0 .. nvec-1])
    int compact;  // lane compaction before every compact-th iteration (0: off)
};

// Per-solver kernel sets (batched_pcg.hip, batched_bicg.hip); each launches
// the set-up passes after the common initialisation, or one iteration
// including the mixed form's replacement and fold.
template <class T> int pcg_setup(const BRun& r, int rb0, int count);
template <class T> int pcg_iteration(const BRun& r);
template <class T> int bicg_iteration(const BRun& r);
// the FP64 replacement r = b - A x_hi with <r,u> (CG) or <r,r^> (BiCGStab),
// preceded, when `fold` (CG), by the fold of x_lo and the pending updates
// into x_hi (batched.hip).  fold_lanes: that fold for the right-hand sides
// flagged in c.rep (ahead of a replacement, which clears the flags) or in
// c.fold (pipelined CG's lanes that finished with updates pending; cleared).
template <class T> int replace_and_fold(const BRun& r, V4<T>* res, V4<T>* rhat, bool fold);
template <class T> int fold_lanes(const BRun& r);

}  // namespace sedi
