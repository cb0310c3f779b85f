// plain.hip -- one-probe-at-a-time FP64 CG, pipelined CG and BiCGStab.
//
// Each solver iteration is a short sequence of fused passes (march.h).  A pass
// streams the vectors it needs once, recomputes the stencil of its input
// field on the fly, applies the pointwise updates of the reference's
// recurrences in the reference's arithmetic order, and reduces its dot
// products deterministically; the pass's last workgroup then advances the
// solver scalars and the convergence test on the device.  Vectors that a pass
// both reads at halo cells and rewrites are double-buffered by iteration
// parity, so no workgroup reads a value another one has already overwritten.
//
// Per-iteration streamed vectors (FP64, N cells; the algorithmic bytes):
//   CG        2 passes: [r, p_old -> p] + [x, r, p -> x, r]          = 8 N w
//   PipelinedCg 1 pass: x,r,u,w,z,q,s,p read and written             = 16 N w
//   BiCGStab  3 passes: [r,p,v,r^ -> p,v] + [r,v] + [x,p,r,v,r^ -> x,r] = 15 N w
#include "plain.h"
#include "profiler.h"
#include "tuning.h"

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

// Prefetch distance (planes) of the march in each pass.
#ifndef SEDI_PREFETCH_APPLY
#define SEDI_PREFETCH_APPLY 2
#endif
#ifndef SEDI_PREFETCH_CG
#define SEDI_PREFETCH_CG 2
#endif
#ifndef SEDI_PREFETCH_PCG
#define SEDI_PREFETCH_PCG 1
#endif
// the pipelined-CG pass with the scalar Jacobi preconditioner (the
// estimator's): two planes (at 256^3 372 -> 364 us a launch); with a vector
// preconditioner two would push the pass past 128 VGPRs into scratch
#ifndef SEDI_PREFETCH_PCG_SCALAR
#define SEDI_PREFETCH_PCG_SCALAR 2
#endif
#ifndef SEDI_PREFETCH_BICG
#define SEDI_PREFETCH_BICG 2
#endif

namespace sedi {

// Store of a pass's output vector: non-temporal by default (at 256^3, three
// A/B runs: cg_update_kernel 112-115 -> 106-108 us, cg_direction_kernel
// 85-86 -> 83 us, bicgstab_half_kernel, which follows a storing pass,
// 73 -> 68 us; pipelined CG unchanged), although the next pass reads it;
// SEDI_PLAIN_NT=0 selects ordinary stores.
#ifndef SEDI_PLAIN_NT
#define SEDI_PLAIN_NT 1
#endif
__device__ __forceinline__ void stp(double* base, cidx_t i, double v) {
#if SEDI_PLAIN_NT
    stc_nt(base, i, v);
#else
    stc(base, i, v);
#endif
}

// ---------------------------------------------------------------------------
// device-side scalar updates (run by thread 0 of a pass's last workgroup)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void set_done(DevState* st, int code, int fi, double fe) {
    st->code = code;
    st->final_iter = fi;
    st->final_err = fe;
    st->done = 1;
}

// BiCgStab.hpp:235-263: the outcome checks in the reference's order, for the
// value the loop variable had when the loop ended (maxiter+1 if it ran out).
__device__ __forceinline__ void bicgstab_outcome(DevState* st, int iter) {
    const double tol2 = st->tol * st->tol;
    if (st->err <= st->tol) set_done(st, 0, iter, st->err);
    else if (st->ss <= tol2) set_done(st, 0, iter, sqrt(st->ss / st->bb));
    else if (st->omega == 0.0) set_done(st, 3, iter, st->err);
    else if (st->rho == 0.0) set_done(st, 2, iter, st->err);
    else set_done(st, 1, st->maxiter, st->err);
}

// Cg.hpp:196-206 / PipelinedCg.hpp:232-242 when the loop ran out.
__device__ __forceinline__ void cg_outcome_exhausted(DevState* st) {
    if (st->err <= st->tol) set_done(st, 0, st->maxiter + 1, st->err);
    else set_done(st, 1, st->maxiter, st->err);
}

__global__ void init_state_kernel(DevState* st, int kind, int maxiter, double tol) {
    if (threadIdx.x == 0) {
        DevState s;
        memset(&s, 0, sizeof(s));
        s.kind = kind;
        s.maxiter = maxiter;
        s.tol = tol;
        s.omega = 1.0;  // BiCgStab.hpp:148
        *st = s;
    }
}

// CG's direction buffers p0 .. p3: the direction of iteration it is in
// p[it % 4].
struct CgDirs {
    double* p[4];
    __device__ double* at(int j) const { return j == 0 ? p[0] : j == 1 ? p[1] : j == 2 ? p[2] : p[3]; }
};

// CG's last solution updates when the solve ended inside a group: x +=
// apend_j p_j for the pending ones, in the order of their iterations (the
// direction of the group's j-th iteration is in p_j).
__global__ void cg_flush_kernel(const DevState* st, long long n, double* x, CgDirs d) {
    const double a1 = st->apend[0], a2 = st->apend[1], a3 = st->apend[2];
    if (a1 == 0.0) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double v = 1.0 * x[i] + a1 * d.p[1][i];
        if (a2 != 0.0) v = 1.0 * v + a2 * d.p[2][i];
        if (a3 != 0.0) v = 1.0 * v + a3 * d.p[3][i];
        x[i] = v;
    }
}

// ---------------------------------------------------------------------------
// passes
// ---------------------------------------------------------------------------
// y = A x (Stencil.hpp:38-54).  DD planes of loads in flight (the
// standalone stencil's only tuning beyond the tile: each wave streams one
// 512-byte row per plane, so a few planes ahead keep enough bytes in flight
// per CU to cover HBM latency).
template <int DD = SEDI_PREFETCH_APPLY>
struct PassApplyT {
    static constexpr int ND = 1, D = DD;
    double a, b, c;
    const double* x;
    double* y;
    struct Raw { double x; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(x, i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ double f(const Raw& r) const { return r.x; }
    __device__ void finish(cidx_t i, const Raw&, const Aux&, double, double Af, double (&)[ND]) const {
        stc_nt(y, i, Af);
    }
};
using PassApply = PassApplyT<>;

// r = b - A x  (Cg.hpp:119-123; PipelinedCg.hpp:113-117; BiCgStab.hpp:131-135)
// with the dots <b,b>, <r,r> and <r, r.*iM>; optionally r_hat = r
// (BiCgStab.hpp:154).
template <class PC>
struct PassResidual {
    static constexpr int ND = 3, D = SEDI_PREFETCH_CG;
    double a, b, c;
    const double* x;
    const double* rhs;
    double* r;
    double* rhat;
    PC iM;
    struct Raw { double x; };
    struct Aux { double b, im; };
    __device__ Raw load(cidx_t i) const { return {ldc(x, i)}; }
    __device__ Aux load_aux(cidx_t i) const { return {ldc(rhs, i), iM.at(i)}; }
    __device__ double f(const Raw& r) const { return r.x; }
    __device__ void finish(cidx_t i, const Raw&, const Aux& q, double, double Af, double (&acc)[ND]) const {
        const double rv = 1.0 * q.b + (-1.0) * Af;  // combine2(1.0, b, -1.0, r, r)
        stp(r, i, rv);
        if (rhat) stp(rhat, i, rv);
        acc[0] += q.b * q.b;
        acc[1] += rv * rv;
        acc[2] += rv * (rv * q.im);
    }
};

// CG direction: p = u + beta*p with u = r.*iM (Cg.hpp:175,188; p = u at the
// first iteration, :144-147), then s = A p and <s,p> (:152-165).
template <class PC>
struct PassCgDirection {
    static constexpr int ND = 1, D = SEDI_PREFETCH_CG;
    double a, b, c;
    const double* r;
    const double* pold;
    double* pnew;
    PC iM;
    double beta;
    bool first;
    struct Raw { double r, p, im; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(pold, i), iM.at(i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ double f(const Raw& q) const {
        const double u = q.r * q.im;
        return first ? u : 1.0 * u + beta * q.p;
    }
    __device__ void finish(cidx_t i, const Raw&, const Aux&, double p, double s, double (&acc)[ND]) const {
        stp(pnew, i, p);
        acc[0] += s * p;
    }
};

// CG update: s = A p recomputed; x += alpha p; r -= alpha s; u = r.*iM;
// <r,u>, <r,r> (Cg.hpp:168-182).
//
// x is write-only in CG's recurrences, so its updates go in groups of four
// iterations: iterations 1-3 of a group leave x alone (their alphas wait in
// DevState::apend), the 4th computes x + alpha_1 p_1 + alpha_2 p_2 +
// alpha_3 p_3 + alpha p as four successive combine2 updates (SliceOps.hpp:
// 38-40) in the reference's order, so x is bitwise the same, reading the
// earlier directions from their own buffers, which still hold them.  A solve
// that ends inside a group gets its last updates from cg_flush_kernel.  The
// pass streams 3 vectors at iterations 1-3 and 8 at the 4th (4.25 a
// iteration) instead of 5 every iteration.  XS (the 4th) is a compile-time
// choice, so the other iterations carry no loads or registers for x.
template <class PC, bool XS>
struct PassCgUpdate {
    static constexpr int ND = 2, D = SEDI_PREFETCH_CG;
    double a, b, c;
    const double* p;
    const double* pend[3];
    double* x;
    double* r;
    PC iM;
    double alpha, apend[3];
    struct Raw { double p; };
    struct Aux { double x, pp[3], r, im; };
    __device__ Raw load(cidx_t i) const { return {ldc(p, i)}; }
    __device__ Aux load_aux(cidx_t i) const {
        Aux o;
        if (XS) {
            o.x = ldc(x, i);
#pragma unroll
            for (int j = 0; j < 3; j++) o.pp[j] = ldc(pend[j], i);
        }
        o.r = ldc(r, i);
        o.im = iM.at(i);
        return o;
    }
    __device__ double f(const Raw& q) const { return q.p; }
    __device__ void finish(cidx_t i, const Raw&, const Aux& q, double pv, double s, double (&acc)[ND]) const {
        if (XS) {
            double xm = q.x;
#pragma unroll
            for (int j = 0; j < 3; j++) xm = 1.0 * xm + apend[j] * q.pp[j];
            stp(x, i, 1.0 * xm + alpha * pv);
        }
        const double rv = 1.0 * q.r + (-alpha) * s;
        stp(r, i, rv);
        const double u = rv * q.im;
        acc[0] += rv * u;
        acc[1] += rv * rv;
    }
};

// Pipelined CG set-up: u = r.*iM, w = A u (PipelinedCg.hpp:138-143), z=q=s=p=0
// (:146-149), and the first iteration's <r,u>, <w,u>, <r,r> (:155,170).
template <class PC>
struct PassPcgInit {
    static constexpr int ND = 3, D = SEDI_PREFETCH_CG;
    double a, b, c;
    const double* r;
    double* u;
    double* w;
    double *z, *q, *s, *p;
    PC iM;
    struct Raw { double r, im; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), iM.at(i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ double f(const Raw& c) const { return c.r * c.im; }
    __device__ void finish(cidx_t i, const Raw& c, const Aux&, double uv, double wv, double (&acc)[ND]) const {
        stp(u, i, uv);
        stp(w, i, wv);
        stp(z, i, 0.0); stp(q, i, 0.0); stp(s, i, 0.0); stp(p, i, 0.0);
        acc[0] += c.r * uv;
        acc[1] += wv * uv;
        acc[2] += c.r * c.r;
    }
};

// Pipelined CG iteration (PipelinedCg.hpp:157-223): m = w.*iM, n = A m, the
// eight recurrences in the reference's order, and the next <r,u>, <w,u>, <r,r>.
template <class PC, bool XS>
struct PassPcgIter {
    static constexpr int ND = 3;
    static constexpr int D = std::is_same<PC, PrecondScalar>::value ? SEDI_PREFETCH_PCG_SCALAR : SEDI_PREFETCH_PCG;
    double a, b, c;
    const double* wold;
    double* wnew;
    double *x, *r, *u, *z, *q, *s;
    const double* pold;
    double* pnew;
    const double* pend[3];
    PC iM;
    double alpha, beta, apend[3];
    struct Raw { double w, im; };
    struct Aux { double q, s, u, p, z, r, x, pp[3]; };
    __device__ Raw load(cidx_t i) const { return {ldc(wold, i), iM.at(i)}; }
    __device__ Aux load_aux(cidx_t i) const {
        Aux o;
        o.q = ldc(q, i); o.s = ldc(s, i); o.u = ldc(u, i); o.p = ldc(pold, i); o.z = ldc(z, i); o.r = ldc(r, i);
        if (XS) {
            o.x = ldc(x, i);
#pragma unroll
            for (int j = 0; j < 3; j++) o.pp[j] = ldc(pend[j], i);
        }
        return o;
    }
    __device__ double f(const Raw& c) const { return c.w * c.im; }
    __device__ void finish(cidx_t i, const Raw& c, const Aux& o, double m, double n, double (&acc)[ND]) const {
        const double qv = 1.0 * m + beta * o.q;
        const double sv = 1.0 * c.w + beta * o.s;
        const double pv = 1.0 * o.u + beta * o.p;
        const double zv = 1.0 * n + beta * o.z;
        const double rv = 1.0 * o.r + (-alpha) * sv;
        const double uv = 1.0 * o.u + (-alpha) * qv;
        const double wv = 1.0 * c.w + (-alpha) * zv;
        stp(q, i, qv); stp(s, i, sv); stp(pnew, i, pv); stp(z, i, zv);
        stp(r, i, rv); stp(u, i, uv); stp(wnew, i, wv);
        if (XS) {
            // the group's pending x += alpha_j p_j, then this iteration's
            double xm = o.x;
#pragma unroll
            for (int j = 0; j < 3; j++) xm = 1.0 * xm + apend[j] * o.pp[j];
            stp(x, i, 1.0 * xm + alpha * pv);
        }
        acc[0] += rv * uv;
        acc[1] += wv * uv;
        acc[2] += rv * rv;
    }
};

// BiCGStab direction (BiCgStab.hpp:165-187): p = r + beta*p - beta*omega*v
// (p = r at the first iteration), p^ = p.*iM, v = A p^, <v, r^>.
template <class PC>
struct PassBicgDirection {
    static constexpr int ND = 1, D = SEDI_PREFETCH_BICG;
    double a, b, c;
    const double *r, *pold, *vold, *rhat;
    double *pnew, *vnew;
    PC iM;
    double beta, cbw;
    bool first;
    struct Raw { double r, p, v, im; };
    struct Aux { double rhat; };
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(pold, i), ldc(vold, i), iM.at(i)}; }
    __device__ Aux load_aux(cidx_t i) const { return {ldc(rhat, i)}; }
    __device__ double pdir(const Raw& q) const {
        return first ? q.r : 1.0 * q.r + beta * q.p + cbw * q.v;  // combine3 (:169)
    }
    __device__ double f(const Raw& q) const { return pdir(q) * q.im; }
    __device__ void finish(cidx_t i, const Raw& q, const Aux& o, double, double v, double (&acc)[ND]) const {
        stp(pnew, i, pdir(q));
        stp(vnew, i, v);
        acc[0] += v * o.rhat;
    }
};

// BiCGStab half step (BiCgStab.hpp:190-213): s = r - alpha v, s^ = s.*iM,
// t = A s^, <s,s>, <t,s>, <t,t>; s and t are not stored.
template <class PC>
struct PassBicgHalf {
    static constexpr int ND = 3, D = SEDI_PREFETCH_BICG;
    double a, b, c;
    const double *r, *v;
    PC iM;
    double alpha;
    struct Raw { double r, v, im; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(v, i), iM.at(i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ double sval(const Raw& q) const { return 1.0 * q.r + (-alpha) * q.v; }
    __device__ double f(const Raw& q) const { return sval(q) * q.im; }
    __device__ void finish(cidx_t, const Raw& q, const Aux&, double, double t, double (&acc)[ND]) const {
        const double sv = sval(q);
        acc[0] += sv * sv;
        acc[1] += t * sv;
        acc[2] += t * t;
    }
};

// BiCGStab update (BiCgStab.hpp:217-224 and :160 of the next iteration):
// s, s^, t recomputed; x = x + alpha p^ + omega s^; r = s - omega t;
// <r,r> and the next rho = <r, r^>.
template <class PC>
struct PassBicgUpdate {
    static constexpr int ND = 2, D = SEDI_PREFETCH_BICG;
    double a, b, c;
    const double *r, *v, *p, *rhat;
    double *x, *rnew;
    PC iM;
    double alpha, omega;
    struct Raw { double r, v, im; };
    struct Aux { double p, x, rhat; };
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(v, i), iM.at(i)}; }
    __device__ Aux load_aux(cidx_t i) const { return {ldc(p, i), ldc(x, i), ldc(rhat, i)}; }
    __device__ double sval(const Raw& q) const { return 1.0 * q.r + (-alpha) * q.v; }
    __device__ double f(const Raw& q) const { return sval(q) * q.im; }
    __device__ void finish(cidx_t i, const Raw& q, const Aux& o, double sh, double t, double (&acc)[ND]) const {
        const double sv = sval(q);
        const double ph = o.p * q.im;
        const double xv = 1.0 * o.x + alpha * ph + omega * sh;  // combine3 (:217)
        const double rv = 1.0 * sv + (-omega) * t;
        stp(x, i, xv);
        stp(rnew, i, rv);
        acc[0] += rv * rv;
        acc[1] += rv * o.rhat;
    }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// The plain passes use tall tiles (8 rows, 8 waves): their registers fit two
// such workgroups per CU, and the taller tile halves the y-halo re-reads.
constexpr int PTY = TY_TALL;
constexpr int PTHREADS = TX * PTY;

#define PASS_SHARED                                  \
    __shared__ double lds[2][PTY + 2][TX + 2];       \
    __shared__ double red[4 * PTY];                  \
    __shared__ int lastflag


// The standalone stencil streams only x in and y out per plane, so it takes
// the tallest tile (16 rows, 16 waves): at 256^3 and 512^3 this is 5-8%
// faster than 8 rows, with y written non-temporally.
constexpr int STY = 16;

template <int ZC, int PF = SEDI_PREFETCH_APPLY>
__global__ __launch_bounds__(TX * STY) void stencil_apply_kernel(TileGrid g, double a, double b,
                                                                 double c, const double* x,
                                                                 double* y) {
    __shared__ double lds[2][STY + 2][TX + 2];
    PassApplyT<PF> P{a, b, c, x, y};
    double acc[1];
    run_tiles<decltype(P), ZC, STY>(P, g, lds, acc);
}

// the stencil with 4 and 8 planes of loads in flight, by chunk (for the
// launch macro, which takes a template over the chunk alone)
template <int ZC>
__global__ __launch_bounds__(TX * STY) void stencil_apply_kernel_pf4(TileGrid g, double a, double b, double c,
                                                                     const double* x, double* y) {
    __shared__ double lds[2][STY + 2][TX + 2];
    PassApplyT<4> P{a, b, c, x, y};
    double acc[1];
    run_tiles<decltype(P), ZC, STY>(P, g, lds, acc);
}
template <int ZC>
__global__ __launch_bounds__(TX * STY) void stencil_apply_kernel_pf8(TileGrid g, double a, double b, double c,
                                                                     const double* x, double* y) {
    __shared__ double lds[2][STY + 2][TX + 2];
    PassApplyT<8> P{a, b, c, x, y};
    double acc[1];
    run_tiles<decltype(P), ZC, STY>(P, g, lds, acc);
}

// y = A x with two cells per lane (march_tile_x2): one 16-byte load of x and
// one 16-byte store of y per pair and plane, and one load per plane for all
// but three waves.  Measured cold at 512^3: 5-8% faster than the one-cell
// stencil; needs n1 even and 16-byte aligned x and y.
template <int DD>
struct PassApplyX2 {
    static constexpr int ND = 1, D = DD;
    double a, b, c;
    const double* x;
    double* y;
    using Raw = double2;
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return ldc2(x, i); }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ double2 f(const Raw& r) const { return r; }
    __device__ void finish(cidx_t i, const Raw&, const Aux&, const double2&, const double2& Af,
                           double (&)[ND]) const {
        stc2_nt(y, i, Af);
    }
};

template <int ZC>
__global__ __launch_bounds__(TX * STY) void stencil_apply_kernel_x2(TileGrid g, double a, double b, double c,
                                                                    const double* x, double* y) {
    __shared__ double lds[2][STY + 2][LX2];
    PassApplyX2<2> P{a, b, c, x, y};
    double acc[1];
    run_tiles_x2<decltype(P), ZC, STY>(P, g, lds, acc);
}
template <int ZC>
__global__ __launch_bounds__(TX * STY) void stencil_apply_kernel_x2_pf4(TileGrid g, double a, double b, double c,
                                                                        const double* x, double* y) {
    __shared__ double lds[2][STY + 2][LX2];
    PassApplyX2<4> P{a, b, c, x, y};
    double acc[1];
    run_tiles_x2<decltype(P), ZC, STY>(P, g, lds, acc);
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void residual_kernel(TileGrid g, Ctl ctl, double a, double b,
                                                            double c, const double* x,
                                                            const double* rhs, double* r,
                                                            double* rhat, PC iM) {
    PASS_SHARED;
    PassResidual<PC> P{a, b, c, x, rhs, r, rhat, iM};
    double acc[3];
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, false);
    if (!grid_sum<3, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    double bb = acc[0];
    if (bb == 0.0) bb = 1.0;  // Cg.hpp:115
    st->bb = bb;
    const double err = sqrt(acc[1] / bb);
    st->err = err;
    if (err < st->tol) {  // Cg.hpp:130-134 (strict <, unlike the loop's <=)
        set_done(st, 0, 0, err);
        return;
    }
    if (st->kind == SEDI_CG) {
        st->ru = acc[2];
        st->iter = 1;
    } else if (st->kind == SEDI_BICGSTAB) {
        // rho = <r, r^> with r^ = r (BiCgStab.hpp:154,160)
        st->rho = acc[1];
        st->rho_prev = 0.0;
        st->alpha = 0.0;
        st->omega = 1.0;
        st->ss = 0.0;
        st->iter = 1;
        if (st->maxiter < 1) bicgstab_outcome(st, st->maxiter + 1);
        else if (st->rho == 0.0) bicgstab_outcome(st, 1);  // :163
        return;
    }
    if (st->maxiter < 1) cg_outcome_exhausted(st);
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void cg_direction_kernel(TileGrid g, Ctl ctl, double a,
                                                                double b, double c, const double* r,
                                                                CgDirs d, PC iM) {
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    const int it = cst->iter;
    PassCgDirection<PC> P{a, b, c, r, d.at((it + 3) & 3), d.at(it & 3), iM, cst->beta, it == 1};
    double acc[1];
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, false);
    if (!grid_sum<1, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    st->alpha = st->ru / acc[0];  // Cg.hpp:166
}

template <class PC, int ZC, bool XS>
__device__ __forceinline__ void run_cg_update(const TileGrid& g, const DevState* cst, double a, double b, double c,
                                              const CgDirs& d, int it, double* x, double* r, PC iM,
                                              double (&lds)[2][PTY + 2][TX + 2], double (&acc)[2]) {
    PassCgUpdate<PC, XS> P{a, b, c, d.at(it & 3), {d.p[1], d.p[2], d.p[3]}, x, r, iM, cst->alpha,
                           {cst->apend[0], cst->apend[1], cst->apend[2]}};
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, true);
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void cg_update_kernel(TileGrid g, Ctl ctl, double a, double b,
                                                             double c, CgDirs d, double* x, double* r,
                                                             PC iM) {
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    const int it = cst->iter;
    const int jp = it & 3;
    double acc[2];
    if (jp == 0) run_cg_update<PC, ZC, true>(g, cst, a, b, c, d, it, x, r, iM, lds, acc);
    else run_cg_update<PC, ZC, false>(g, cst, a, b, c, d, it, x, r, iM, lds, acc);
    if (!grid_sum<2, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    // this alpha waits for the group's last iteration, which applies them all
    if (jp == 0) st->apend[0] = st->apend[1] = st->apend[2] = 0.0;
    else st->apend[jp - 1] = st->alpha;
    const double ru_new = acc[0];
    st->beta = ru_new / st->ru;  // Cg.hpp:185
    st->ru = ru_new;
    const double err = sqrt(acc[1] / st->bb);  // :191
    st->err = err;
    ctl.hist[it - 1] = err;
    st->nhist = it;
    if (err <= st->tol) set_done(st, 0, it, err);  // :193,197-201
    else if (it >= st->maxiter) cg_outcome_exhausted(st);
    else st->iter = it + 1;
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void pcg_init_kernel(TileGrid g, Ctl ctl, double a, double b,
                                                            double c, const double* r, double* u,
                                                            double* w, double* z, double* q,
                                                            double* s, double* p, PC iM) {
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    PassPcgInit<PC> P{a, b, c, r, u, w, z, q, s, p, iM};
    double acc[3];
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, true);
    if (!grid_sum<3, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    st->gamma = acc[0];  // PipelinedCg.hpp:171-173 at iter 1
    st->delta = acc[1];
    st->beta = 0.0;      // :179-180
    st->alpha = acc[0] / acc[1];
    st->err_pending = sqrt(acc[2] / st->bb);
    st->iter = 1;
}

template <class PC, int ZC, bool XS>
__device__ __forceinline__ void run_pcg_iter(const TileGrid& g, const DevState* cst, double a, double b, double c,
                                             double* w0, double* w1, double* x, double* r, double* u, double* z,
                                             double* q, double* s, const CgDirs& d, int it, PC iM,
                                             double (&lds)[2][PTY + 2][TX + 2], double (&acc)[3]) {
    // iteration 1 reads w from w0 (written by pcg_init_kernel) and p from p0
    PassPcgIter<PC, XS> P{a, b, c, (it & 1) ? w0 : w1, (it & 1) ? w1 : w0, x, r, u, z, q, s,
                          d.at((it + 3) & 3), d.at(it & 3), {d.p[1], d.p[2], d.p[3]}, iM, cst->alpha, cst->beta,
                          {cst->apend[0], cst->apend[1], cst->apend[2]}};
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, (it & 1) == 0);
}

// Pipelined CG's solution updates come in groups of four iterations, as
// plain CG's (the x update is the last of the eight recurrences and nothing
// else reads x): the direction of iteration it is written to p[it % 4], the
// 4th iteration of a group adds the three pending updates and its own.
template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void pcg_iter_kernel(TileGrid g, Ctl ctl, double a, double b,
                                                            double c, double* w0, double* w1,
                                                            double* x, double* r, double* u,
                                                            double* z, double* q, double* s,
                                                            CgDirs d, PC iM) {
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    const int it = cst->iter;
    const int jp = it & 3;
    double acc[3];
    if (jp == 0) run_pcg_iter<PC, ZC, true>(g, cst, a, b, c, w0, w1, x, r, u, z, q, s, d, it, iM, lds, acc);
    else run_pcg_iter<PC, ZC, false>(g, cst, a, b, c, w0, w1, x, r, u, z, q, s, d, it, iM, lds, acc);
    if (!grid_sum<3, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    {
        // the alpha this pass used waits for the group's last iteration
        DevState* st = ctl.st;
        if (jp == 0) st->apend[0] = st->apend[1] = st->apend[2] = 0.0;
        else st->apend[jp - 1] = st->alpha;
    }
    DevState* st = ctl.st;
    const double err = st->err_pending;  // rr before this iteration's update (:227)
    st->err = err;
    ctl.hist[it - 1] = err;
    st->nhist = it;
    if (err <= st->tol) { set_done(st, 0, it, err); return; }
    if (it >= st->maxiter) { cg_outcome_exhausted(st); return; }
    // scalars of iteration it+1 (PipelinedCg.hpp:167-178)
    const double gamma_old = st->gamma;
    const double gamma = acc[0], delta = acc[1];
    const double beta = gamma / gamma_old;
    const double alpha = 1.0 / (delta / gamma - beta / st->alpha);
    st->gamma = gamma;
    st->delta = delta;
    st->beta = beta;
    st->alpha = alpha;
    st->err_pending = sqrt(acc[2] / st->bb);
    st->iter = it + 1;
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void bicgstab_direction_kernel(
    TileGrid g, Ctl ctl, double a, double b, double c, double* const* vv, PC iM) {
    // vv: r0, r1, rhat, p0, p1, v0, v1
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    const int it = cst->iter;
    const int cur = it & 1, prv = cur ^ 1;
    PassBicgDirection<PC> P{a, b, c, vv[0 + cur], vv[3 + prv], vv[5 + prv], vv[2],
                        vv[3 + cur], vv[5 + cur], iM, cst->beta, cst->cbw, it == 1};
    double acc[1];
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, (it & 1) == 1);
    if (!grid_sum<1, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    st->alpha = st->rho / acc[0];  // BiCgStab.hpp:187
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void bicgstab_half_kernel(TileGrid g, Ctl ctl, double a,
                                                                 double b, double c,
                                                                 double* const* vv, PC iM) {
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    const int cur = cst->iter & 1;
    PassBicgHalf<PC> P{a, b, c, vv[0 + cur], vv[5 + cur], iM, cst->alpha};
    double acc[3];
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, (cst->iter & 1) == 0);
    if (!grid_sum<3, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    st->ss = acc[0];                               // :193-194
    if (acc[0] < st->tol * st->tol) st->sexit = 1;  // :195 (absolute test)
    else st->omega = acc[1] / acc[2];              // :213
}

template <class PC, int ZC>
__global__ __launch_bounds__(PTHREADS, 4) void bicgstab_update_kernel(TileGrid g, Ctl ctl, double a,
                                                                   double b, double c,
                                                                   double* const* vv, double* x,
                                                                   PC iM) {
    const DevState* cst = ctl.st;
    if (cst->done) return;
    PASS_SHARED;
    const int it = cst->iter;
    const int cur = it & 1;
    double acc[2];
    if (cst->sexit) {
        // x = x + alpha p^ and stop (BiCgStab.hpp:195-199)
        const double alpha = cst->alpha;
        const double* p = vv[3 + cur];
        const long long n = (long long)g.n1 * g.n2 * g.n3;
        for (long long i = (long long)blockIdx.x * PTHREADS + threadIdx.x; i < n;
             i += (long long)gridDim.x * PTHREADS) {
            const cidx_t c = (cidx_t)i;
            const double ph = ldc(p, c) * iM.at(c);
            stc(x, c, 1.0 * ldc(x, c) + alpha * ph);
        }
        acc[0] = acc[1] = 0.0;
        if (!grid_sum<2, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
        bicgstab_outcome(ctl.st, it);
        return;
    }
    PassBicgUpdate<PC> P{a, b, c, vv[0 + cur], vv[5 + cur], vv[3 + cur], vv[2],
                     x, vv[0 + (cur ^ 1)], iM, cst->alpha, cst->omega};
    run_tiles<decltype(P), ZC, PTY>(P, g, lds, acc, (it & 1) == 1);
    if (!grid_sum<2, PTY>(acc, ctl.partials, ctl.counter, red, &lastflag) || threadIdx.x != 0) return;
    DevState* st = ctl.st;
    const double err = sqrt(acc[0] / st->bb);  // :223-227
    st->err = err;
    ctl.hist[st->nhist] = err;
    st->nhist += 1;
    if (err <= st->tol || st->omega == 0.0) { bicgstab_outcome(st, it); return; }  // :229-230
    if (it >= st->maxiter) { bicgstab_outcome(st, st->maxiter + 1); return; }
    st->rho_prev = st->rho;  // :232
    st->rho = acc[1];        // next iteration's :160
    if (st->rho == 0.0) { bicgstab_outcome(st, it + 1); return; }  // :163
    const double beta = (st->rho / st->rho_prev) * (st->alpha / st->omega);  // :167
    st->beta = beta;
    st->cbw = -beta * st->omega;  // the combine3 coefficient of :169
    st->iter = it + 1;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int num_cus() {
    static int cached = 0;
    if (!cached) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached = cus;
    }
    return cached;
}

// Workgroups of 256 lanes that can be resident per CU (4 waves each; at most
// 8 with the 32 waves a CU holds).
constexpr int MAX_BLOCKS_PER_CU = 8;

// Geometry of a pass on `s`, from the occupancy of its ZC=16 instantiation
// (the ZC variants differ only in unrolling); cached per kernel and shape.
template <typename K>
static const KGeom& geom_for(const char* name, K kern16, const Shape& s, int D, int ty = PTY,
                             bool one_per_tile = false, int tx = TX) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, int, int>, KGeom> cache;
    const auto key = std::make_tuple(reinterpret_cast<const void*>(kern16), s.n1, s.n2, s.n3);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern16, TX * ty, 0) != hipSuccess || occ <= 0) {
        (void)hipGetLastError();
        occ = 4;
    }
    occ = std::min(occ, MAX_BLOCKS_PER_CU);
    // "wpc:<kernel>" > 0 launches at most that many workgroups per CU, which
    // loop over the tiles (default 1 for the pair march: one 16-wave workgroup
    // per CU measured faster than two); 0, the default of the other kernels,
    // queues up to MAX_BLOCKS_PER_CU per CU in the launch-order schedules.
    // One workgroup per tile (schedule 2) ignores it.
    const int wpc = tuned_value(std::string("wpc:") + name, s, tx == TX2 ? 1 : 0);
    if (wpc > 0) occ = std::min(occ, wpc);
    // the occupancy plan, with the autotuner's z chunk and schedule for this
    // kernel and grid (the partial sums hold MAX_BLOCKS_PER_CU per CU)
    const KGeom planned = plan_geometry(s, occ * num_cus(), D, ty, tx);
    const int zc = tuned_value(std::string("zc:") + name, s, 0);
    const int sched = tuned_value(std::string("sched:") + name, s, 0);
    const int max_blocks = (wpc > 0 ? occ : MAX_BLOCKS_PER_CU) * num_cus();
    return cache.emplace(key, tuned_geometry(planned, s, zc, sched, max_blocks, one_per_tile)).first->second;
}

// kernel<..., ZC> for the ZC the geometry chose
#define SEDI_ZC_LAUNCH(name, K, geo, ...)                                                          \
    ((geo).tg.zc == 8    ? launch(name, K<8>, dim3((geo).nblocks), dim3(TX * (geo).tg.ty), __VA_ARGS__)    \
     : (geo).tg.zc == 32 ? launch(name, K<32>, dim3((geo).nblocks), dim3(TX * (geo).tg.ty), __VA_ARGS__)   \
                         : launch(name, K<16>, dim3((geo).nblocks), dim3(TX * (geo).tg.ty), __VA_ARGS__))
#define SEDI_ZC_LAUNCH_PC(name, K, PCT, geo, ...)                                                   \
    ((geo).tg.zc == 8    ? launch(name, K<PCT, 8>, dim3((geo).nblocks), dim3(TX * (geo).tg.ty), __VA_ARGS__) \
     : (geo).tg.zc == 32 ? launch(name, K<PCT, 32>, dim3((geo).nblocks), dim3(TX * (geo).tg.ty), __VA_ARGS__)\
                         : launch(name, K<PCT, 16>, dim3((geo).nblocks), dim3(TX * (geo).tg.ty), __VA_ARGS__))

int stencil_apply_dev(const Shape& s, double a, double b, double c, const double* x, double* y,
                      hipStream_t stream) {
    SEDI_REQUIRE(offsets_fit_32(s.n(), sizeof(double)), SEDI_ERR_UNSUPPORTED,
                 "grid of %lld cells too large: an FP64 vector must be < 4 GiB (32-bit offsets)", s.n());
    // planes of loads in flight: the autotuner's "pf:stencil_apply_kernel" (2, 4 or 8;
    // the pair march has 2 and 4); "x2:stencil_apply_kernel" 0 selects the
    // one-cell march where the pair march could run (default: the pair march)
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, std::pair<int, int>> plan_cache;
    int pf, x2;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = plan_cache.find(std::make_tuple(s.n1, s.n2, s.n3));
        if (it == plan_cache.end())
            it = plan_cache.emplace(std::make_tuple(s.n1, s.n2, s.n3),
                                    std::make_pair(tuned_value("pf:stencil_apply_kernel", s, SEDI_PREFETCH_APPLY),
                                                   tuned_value("x2:stencil_apply_kernel", s, 1))).first;
        pf = it->second.first;
        x2 = it->second.second;
    }
    const bool pairs_ok = s.n1 % 2 == 0 && (reinterpret_cast<uintptr_t>(x) % 16) == 0 &&
                          (reinterpret_cast<uintptr_t>(y) % 16) == 0;
    if (x2 && pairs_ok) {
        const KGeom& k = geom_for("stencil_apply_kernel", stencil_apply_kernel_x2<16>, s, 2, STY, true, TX2);
        if (pf >= 4) {
            SEDI_HIP(SEDI_ZC_LAUNCH("stencil_apply_kernel", stencil_apply_kernel_x2_pf4, k, stream, k.tg, a, b, c,
                                    x, y));
        } else {
            SEDI_HIP(SEDI_ZC_LAUNCH("stencil_apply_kernel", stencil_apply_kernel_x2, k, stream, k.tg, a, b, c, x, y));
        }
        return SEDI_OK;
    }
    // no partial sums: one workgroup per tile is a schedule choice too
    const KGeom& k = geom_for("stencil_apply_kernel", stencil_apply_kernel<16>, s, PassApply::D, STY, true);
    if (pf == 8) {
        SEDI_HIP(SEDI_ZC_LAUNCH("stencil_apply_kernel", stencil_apply_kernel_pf8, k, stream, k.tg, a, b, c, x, y));
    } else if (pf == 4) {
        SEDI_HIP(SEDI_ZC_LAUNCH("stencil_apply_kernel", stencil_apply_kernel_pf4, k, stream, k.tg, a, b, c, x, y));
    } else {
        SEDI_HIP(SEDI_ZC_LAUNCH("stencil_apply_kernel", stencil_apply_kernel, k, stream, k.tg, a, b, c, x, y));
    }
    return SEDI_OK;
}

// ---- launch sequences over a PlainCore -------------------------------------
// Set-up of a solve: the residual r = b - A x with <b,b>, <r,r>, <r,u> (and
// r^ = r for BiCGStab); pipelined CG's first u, w and dots.
template <class PC>
static int core_setup(const PlainCore& k, PC pc, const double* b, double* x) {
    const KGeom& kr = geom_for("residual_kernel", residual_kernel<PC, 16>, k.shape, PassResidual<PC>::D);
    if (k.kind == SEDI_BICGSTAB) {
        // r_1 lives in r[1] (buffers alternate with the iteration's parity)
        SEDI_HIP(SEDI_ZC_LAUNCH_PC("residual_kernel", residual_kernel, PC, kr, k.stream, kr.tg, k.ctl,
                        k.a, k.b, k.c, (const double*)x, b, k.v[1], k.v[2], pc));
    } else {
        SEDI_HIP(SEDI_ZC_LAUNCH_PC("residual_kernel", residual_kernel, PC, kr, k.stream, kr.tg, k.ctl,
                        k.a, k.b, k.c, (const double*)x, b, k.v[0], (double*)nullptr, pc));
    }
    if (k.kind == SEDI_PCG) {
        const KGeom& kp = geom_for("pcg_init_kernel", pcg_init_kernel<PC, 16>, k.shape, PassPcgInit<PC>::D);
        SEDI_HIP(SEDI_ZC_LAUNCH_PC("pcg_init_kernel", pcg_init_kernel, PC, kp, k.stream, kp.tg, k.ctl,
                        k.a, k.b, k.c, (const double*)k.v[0], k.v[1], k.v[2], k.v[4], k.v[5], k.v[6], k.v[7], pc));  // p0
    }
    return SEDI_OK;
}

// `count` iterations of the solver, launched directly.
template <class PC>
static int core_iterations(const PlainCore& k, int count, PC pc, double* x) {
    for (int it = 0; it < count; it++) {
        if (k.kind == SEDI_CG) {
            const KGeom& g1 = geom_for("cg_direction_kernel", cg_direction_kernel<PC, 16>, k.shape, PassCgDirection<PC>::D);
            const KGeom& g2 = geom_for("cg_update_kernel", cg_update_kernel<PC, 16>, k.shape, PassCgUpdate<PC, false>::D);
            const CgDirs d{{k.v[1], k.v[2], k.v[3], k.v[4]}};
            SEDI_HIP(SEDI_ZC_LAUNCH_PC("cg_direction_kernel", cg_direction_kernel, PC, g1, k.stream, g1.tg, k.ctl, k.a, k.b, k.c, (const double*)k.v[0], d, pc));
            SEDI_HIP(SEDI_ZC_LAUNCH_PC("cg_update_kernel", cg_update_kernel, PC, g2, k.stream, g2.tg,
                            k.ctl, k.a, k.b, k.c, d, x, k.v[0], pc));
        } else if (k.kind == SEDI_PCG) {
            const KGeom& g1 = geom_for("pcg_iter_kernel", pcg_iter_kernel<PC, 16>, k.shape, PassPcgIter<PC, false>::D);
            SEDI_HIP(SEDI_ZC_LAUNCH_PC("pcg_iter_kernel", pcg_iter_kernel, PC, g1, k.stream, g1.tg,
                            k.ctl, k.a, k.b, k.c, k.v[2], k.v[3], x, k.v[0], k.v[1], k.v[4], k.v[5], k.v[6],
                            CgDirs{{k.v[7], k.v[8], k.v[9], k.v[10]}}, pc));
        } else {
            const KGeom& g1 = geom_for("bicgstab_direction_kernel", bicgstab_direction_kernel<PC, 16>, k.shape, PassBicgDirection<PC>::D);
            const KGeom& g2 = geom_for("bicgstab_half_kernel", bicgstab_half_kernel<PC, 16>, k.shape, PassBicgHalf<PC>::D);
            const KGeom& g3 = geom_for("bicgstab_update_kernel", bicgstab_update_kernel<PC, 16>, k.shape, PassBicgUpdate<PC>::D);
            SEDI_HIP(SEDI_ZC_LAUNCH_PC("bicgstab_direction_kernel", bicgstab_direction_kernel, PC, g1, k.stream, g1.tg, k.ctl, k.a, k.b, k.c, k.vv, pc));
            SEDI_HIP(SEDI_ZC_LAUNCH_PC("bicgstab_half_kernel", bicgstab_half_kernel, PC, g2, k.stream, g2.tg, k.ctl, k.a, k.b, k.c, k.vv, pc));
            SEDI_HIP(SEDI_ZC_LAUNCH_PC("bicgstab_update_kernel", bicgstab_update_kernel, PC, g3, k.stream, g3.tg, k.ctl, k.a, k.b, k.c, k.vv, x, pc));
        }
    }
    return SEDI_OK;
}

// CG's and pipelined CG's pending solution updates of a solve that ended
// inside a group (their direction buffers: CG v[1..4], pipelined CG v[7..10]).
static int core_flush(const PlainCore& k, double* x) {
    if (k.kind != SEDI_CG && k.kind != SEDI_PCG) return SEDI_OK;
    const long long n = k.shape.n();
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4LL * num_cus());
    const CgDirs d = k.kind == SEDI_CG ? CgDirs{{k.v[1], k.v[2], k.v[3], k.v[4]}}
                                       : CgDirs{{k.v[7], k.v[8], k.v[9], k.v[10]}};
    SEDI_HIP(launch("cg_flush_kernel", cg_flush_kernel, dim3(std::max(1u, blocks)), dim3(256), k.stream,
                    (const DevState*)k.ctl.st, n, x, d));
    return SEDI_OK;
}

static int core_iterate_any(const PlainCore& k, int count, const double* iM, double* x) {
    return iM ? core_iterations(k, count, PrecondVec{iM}, x) : core_iterations(k, count, PrecondScalar{1.0 / k.a}, x);
}

// Points a PlainCore's control block and vectors into `base` (device memory
// laid out as plain_ctl_bytes(nblocks) | vectors).
static void core_bind(PlainCore& k, char* base, int nblocks, double* hist) {
    k.ctl = Ctl{reinterpret_cast<DevState*>(base), reinterpret_cast<unsigned*>(base + 256),
                reinterpret_cast<double*>(base + 512), hist};
    k.vv = reinterpret_cast<double* const*>(base + 384);
    (void)nblocks;
}

// ---- PlainSolver -----------------------------------------------------------
PlainSolver::~PlainSolver() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (pinned_) (void)hipHostFree(pinned_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int PlainSolver::create(const Shape& s, double a, double b, double c, int kind, int mode) {
    SEDI_REQUIRE(s.n1 > 0 && s.n2 > 0 && s.n3 > 0, SEDI_ERR_ARG, "grid dimensions must be positive");
    SEDI_REQUIRE(kind == SEDI_CG || kind == SEDI_PCG || kind == SEDI_BICGSTAB, SEDI_ERR_ARG,
                 "unknown solver kind %d", kind);
    SEDI_REQUIRE(offsets_fit_32(s.n(), sizeof(double)), SEDI_ERR_UNSUPPORTED,
                 "grid of %lld cells too large: an FP64 vector must be < 4 GiB (32-bit offsets)", s.n());
    shape_ = s;
    a_ = a; b_ = b; c_ = c;
    kind_ = kind;
    mode_ = mode;
    const int nblocks_max = MAX_BLOCKS_PER_CU * num_cus();  // partials are sized for this many workgroups
    SEDI_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SEDI_HIP(hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming));
    SEDI_HIP(hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming));
    SEDI_HIP(hipHostMalloc((void**)&pinned_, 2 * sizeof(DevState), hipHostMallocDefault));
    const size_t ctl_bytes = plain_ctl_bytes(nblocks_max);
    SEDI_HIP(ctl_.reserve(ctl_bytes));
    SEDI_HIP(hipMemset(ctl_.p, 0, ctl_bytes));
    const int nv = plain_vectors(kind);
    const size_t n = (size_t)s.n();
    SEDI_HIP(vec_.reserve(nv * n * sizeof(double)));
    SEDI_HIP(hipMemset(vec_.p, 0, nv * n * sizeof(double)));
    core_ = PlainCore{};
    core_.shape = s;
    core_.a = a; core_.b = b; core_.c = c;
    core_.kind = kind;
    core_.stream = stream_;
    core_bind(core_, static_cast<char*>(ctl_.p), nblocks_max, nullptr);
    for (int k = 0; k < nv; k++) core_.v[k] = vec_.as<double>() + k * n;
    if (kind == SEDI_BICGSTAB) {
        // BiCGStab's vector table (r0 r1 r^ p0 p1 v0 v1) sits in the control
        // block after the counter; its kernels pick buffers by parity.
        double* table[7] = {core_.v[0], core_.v[1], core_.v[2], core_.v[3], core_.v[4], core_.v[5], core_.v[6]};
        SEDI_HIP(hipMemcpy(static_cast<char*>(ctl_.p) + 384, table, sizeof(table), hipMemcpyHostToDevice));
    }
    return SEDI_OK;
}

// The first chunk of a solver (per preconditioner form) is launched directly,
// which settles the passes' geometry; later chunks replay one captured graph.
// Every per-iteration value lives in device memory, so the graph only depends
// on the pointers its launches were given, which are checked here.
int PlainSolver::iterate(int count, const double* iM, double* x) {
    auto direct = [&]() { return core_iterate_any(core_, count, iM, x); };
    bool& ready = geometry_ready_[iM ? 1 : 0];
    if (!graphs_enabled() || profile_any() || !ready) {
        const int rc = direct();
        ready = true;
        return rc;
    }
    if (!graph_ || graph_count_ != count || graph_x_ != x || graph_m_ != iM || graph_hist_ != hist_.p) {
        if (graph_) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
        SEDI_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        const int rc = direct();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream_, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        SEDI_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SEDI_HIP(ei);
        graph_count_ = count;
        graph_x_ = x;
        graph_m_ = iM;
        graph_hist_ = hist_.p;
    }
    SEDI_HIP(hipGraphLaunch(graph_, stream_));
    return SEDI_OK;
}

int PlainSolver::solve_dev(const double* b, double* x, const double* iM, int maxiter, double tol,
                           double* final_err, int* final_iter) {
    SEDI_REQUIRE(maxiter >= 0, SEDI_ERR_ARG, "maxiter must be >= 0");
    SEDI_HIP(hist_.reserve(sizeof(double) * (size_t)std::max(maxiter, 1)));
    core_.ctl.hist = hist_.as<double>();
    DevState* st = core_.ctl.st;

    SEDI_HIP(launch("init_state_kernel", init_state_kernel, dim3(1), dim3(64), stream_, st, kind_,
                    maxiter, tol));
    int rc = iM ? core_setup(core_, PrecondVec{iM}, b, x) : core_setup(core_, PrecondScalar{1.0 / a_}, b, x);
    if (rc) return rc;

    // Iterations go out in chunks; the host looks at the device's done flag
    // of the previous chunk while the next one runs, so the GPU never idles
    // on the host.  Kernels launched after convergence return at once.
    const int chunk = 16;
    int launched = 0, slot = 0;
    bool pending = false;
    while (launched < maxiter) {
        const int cnt = std::min(chunk, maxiter - launched);
        if ((rc = iterate(cnt, iM, x))) return rc;
        launched += cnt;
        SEDI_HIP(hipMemcpyAsync(&pinned_[slot], st, sizeof(DevState), hipMemcpyDeviceToHost, stream_));
        SEDI_HIP(hipEventRecord(ev_[slot], stream_));
        if (pending) {
            SEDI_HIP(hipEventSynchronize(ev_[slot ^ 1]));
            if (pinned_[slot ^ 1].done) break;
        }
        pending = true;
        slot ^= 1;
    }
    if ((rc = core_flush(core_, x))) return rc;
    SEDI_HIP(hipStreamSynchronize(stream_));
    profile_flush();
    DevState hs;
    SEDI_HIP(hipMemcpy(&hs, st, sizeof(DevState), hipMemcpyDeviceToHost));
    SEDI_REQUIRE(hs.done, SEDI_ERR_STATE, "solver stopped without a result (iter %d)", hs.iter);
    nhist_ = hs.nhist;
    hist_host_.resize(std::max(nhist_, 0));
    if (nhist_ > 0)
        SEDI_HIP(hipMemcpy(hist_host_.data(), hist_.p, sizeof(double) * nhist_, hipMemcpyDeviceToHost));
    if (final_err) *final_err = hs.final_err;
    if (final_iter) *final_iter = hs.final_iter;
    return hs.code;
}

int PlainSolver::solve_any(const double* b, double* x, const double* iM, int maxiter, double tol,
                           double* final_err, int* final_iter) {
    SEDI_REQUIRE(b && x, SEDI_ERR_ARG, "b and x must not be NULL");
    const size_t bytes = sizeof(double) * (size_t)shape_.n();
    const double* bd = b;
    double* xd = x;
    const double* md = iM;
    const bool xhost = !is_device_pointer(x);
    if (!is_device_pointer(b)) {
        SEDI_HIP(stage_b_.reserve(bytes));
        SEDI_HIP(hipMemcpy(stage_b_.p, b, bytes, hipMemcpyHostToDevice));
        bd = stage_b_.as<double>();
    }
    if (xhost) {
        SEDI_HIP(stage_x_.reserve(bytes));
        SEDI_HIP(hipMemcpy(stage_x_.p, x, bytes, hipMemcpyHostToDevice));
        xd = stage_x_.as<double>();
    }
    if (iM && !is_device_pointer(iM)) {
        SEDI_HIP(stage_m_.reserve(bytes));
        SEDI_HIP(hipMemcpy(stage_m_.p, iM, bytes, hipMemcpyHostToDevice));
        md = stage_m_.as<double>();
    }
    const int rc = solve_dev(bd, xd, md, maxiter, tol, final_err, final_iter);
    if (rc < 0) return rc;
    if (xhost) SEDI_HIP(hipMemcpy(x, xd, bytes, hipMemcpyDeviceToHost));
    return rc;
}

int PlainSolver::err_history(double* out, int cap) const {
    if (out)
        for (int k = 0; k < std::min(cap, nhist_); k++) out[k] = hist_host_[k];
    return nhist_;
}

// ---------------------------------------------------------------------------
// the allocation-free solve
// ---------------------------------------------------------------------------
// Its working set -- control block, err history and vectors -- lives in this
// fixed device array, which the code object brings along: each device that
// loads the library holds its own copy until the process ends.  Host code
// never names the array directly (its name is a host-side stand-in, not a
// device address); it uses the address hipGetSymbolAddress returns for the
// current device, looked up once per device.  Calls on one device share the
// array, so calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[SEDI_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static char* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int plain_noalloc_setup() {
    int dev = 0;
    SEDI_HIP(hipGetDevice(&dev));
    SEDI_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, SEDI_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[dev]) return SEDI_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
        return SEDI_ERR_HIP;
    }
    g_noalloc_ptr[dev] = static_cast<char*>(p);
    return SEDI_OK;
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

size_t plain_noalloc_bytes(const Shape& s, int kind, int maxiter) {
    const int nblocks_max = MAX_BLOCKS_PER_CU * num_cus();
    return round256(plain_ctl_bytes(nblocks_max)) + round256(sizeof(double) * (size_t)std::max(maxiter, 1)) +
           (size_t)plain_vectors(kind) * round256(sizeof(double) * (size_t)s.n());
}

// Zeroes the arrival counter (a call that stopped part-way may have left it
// armed) and writes BiCGStab's vector table, in place of the memset and copy
// of PlainSolver::create.
struct VTable { double* v[7]; };
__global__ void noalloc_ctl_kernel(unsigned* counter, double** table, VTable t) {
    const int i = threadIdx.x;
    if (i == 0) *counter = 0u;
    if (i < 7) table[i] = t.v[i];
}

int plain_solve_noalloc(const Shape& s, double a, double b, double c, int kind, const double* rhs, double* x,
                        const double* iM, int maxiter, double tol, double* final_err, int* final_iter) {
    SEDI_REQUIRE(maxiter >= 0, SEDI_ERR_ARG, "maxiter must be >= 0");
    const size_t need = plain_noalloc_bytes(s, kind, maxiter);
    SEDI_REQUIRE(need <= (size_t)SEDI_NOALLOC_SCRATCH_BYTES, SEDI_ERR_ARG,
                 "solve needs %zu bytes of working memory, above the allocation-free limit of %zu "
                 "(SEDI_NOALLOC_SCRATCH_BYTES); use sedi_solver_create/sedi_solve",
                 need, (size_t)SEDI_NOALLOC_SCRATCH_BYTES);
    int rc = plain_noalloc_setup();
    if (rc) return rc;
    int dev = 0;
    SEDI_HIP(hipGetDevice(&dev));
    char* base = g_noalloc_ptr[dev];
    ProfileSuppress no_events;

    // layout: control block | err history | vectors
    const int nblocks_max = MAX_BLOCKS_PER_CU * num_cus();
    PlainCore k;
    k.shape = s;
    k.a = a; k.b = b; k.c = c;
    k.kind = kind;
    k.stream = nullptr;  // the null stream: creating one would allocate
    char* p = base + round256(plain_ctl_bytes(nblocks_max));
    core_bind(k, base, nblocks_max, reinterpret_cast<double*>(p));
    p += round256(sizeof(double) * (size_t)std::max(maxiter, 1));
    const size_t vbytes = round256(sizeof(double) * (size_t)s.n());
    for (int q = 0; q < plain_vectors(kind); q++, p += vbytes) k.v[q] = reinterpret_cast<double*>(p);

    VTable t{};
    for (int q = 0; q < 7; q++) t.v[q] = k.v[q];
    SEDI_HIP(launch("noalloc_ctl_kernel", noalloc_ctl_kernel, dim3(1), dim3(64), k.stream, k.ctl.counter,
                    reinterpret_cast<double**>(base + 384), t));
    SEDI_HIP(launch("init_state_kernel", init_state_kernel, dim3(1), dim3(64), k.stream, k.ctl.st, kind, maxiter,
                    tol));
    rc = iM ? core_setup(k, PrecondVec{iM}, rhs, x) : core_setup(k, PrecondScalar{1.0 / a}, rhs, x);
    if (rc) return rc;
    // chunks of iterations; after each the host reads the done flag with a
    // synchronous copy into its own stack (no pinned buffer to allocate)
    const int chunk = 32;
    for (int launched = 0; launched < maxiter;) {
        const int cnt = std::min(chunk, maxiter - launched);
        if ((rc = core_iterate_any(k, cnt, iM, x))) return rc;
        launched += cnt;
        int done = 0;
        SEDI_HIP(hipMemcpy(&done, &k.ctl.st->done, sizeof(int), hipMemcpyDeviceToHost));
        if (done) break;
    }
    if ((rc = core_flush(k, x))) return rc;
    DevState hs;
    SEDI_HIP(hipMemcpy(&hs, k.ctl.st, sizeof(DevState), hipMemcpyDeviceToHost));
    SEDI_REQUIRE(hs.done, SEDI_ERR_STATE, "solver stopped without a result (iter %d)", hs.iter);
    if (final_err) *final_err = hs.final_err;
    if (final_iter) *final_iter = hs.final_iter;
    return hs.code;
}

}  // namespace sedi
