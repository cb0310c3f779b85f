// batched.hip -- multi-RHS Krylov solvers: one stencil sweep serves a block of
// 4 probes.  This file: the layout, the common kernels (set-up, replacement,
// fold, accumulation, packing), the CG passes and the host driver; the
// pipelined CG and BiCGStab passes are in batched_pcg.hip / batched_bicg.hip.
//
// Layout: a batch of K right-hand sides is stored as K/4 RHS blocks, each a
// dense grid of cells holding the 4 probes' values side by side ([K/4][N][4]).
// A lane of the march owns one cell of one block and moves 4 values per
// access (16 B in FP32: a wave load covers 1 KiB, as in the plain path), so
// the stencil's neighbour reuse, the LDS plane and the halo are shared by 4
// probes.  Workgroups are laid out as (tile, RHS block); every block keeps
// its own scalars and reductions.
//
// Arithmetic, mixed form: with FP32 storage the stencil and the vector updates
// run in FP32 with the FP64 scalars rounded to FP32, while the dot products
// accumulate the exact FP64 products of the FP32 values, and the scalars
// (alpha, beta, the norms) stay FP64.  An RHS gets an FP64 residual
// replacement r = b - A (x_hi + x_lo) (x_hi += x_lo, x_lo = 0 first; computed
// in FP64, stored as FP32) every rep_every iterations, when its recursive
// residual has dropped tenfold since the last replacement, when it reaches the
// tolerance, and at maxiter (CG; the pipelined CG and BiCGStab policies are
// in bind_t).  The tenfold rule is the reliable-update
// criterion (Sleijpen & van der Vorst, "Reliable updated residuals in hybrid
// Bi-CG methods", Computing 56, 1996; delta = 0.1 as in Clark et al., "Solving
// lattice QCD systems of equations using mixed precision solvers on GPUs",
// Comput. Phys. Commun. 181, 2010): the FP32 recursion drifts from the true
// residual by about eps32 * cond(A) * |r at the last replacement|, so
// replacing at every decade keeps the drift at a percent of |r| for
// cond(A) up to ~1e5 (512^3).  After a replacement beta is recomputed from the
// new residual; if the true residual exceeds twice the recursive one the
// recursion had drifted and the directions restart (beta = 0).  An RHS is done
// only when its true residual meets the tolerance.  With FP64 storage the
// recurrences are the plain solvers'.
//
// Solution updates of CG come in groups of CG_XG (4; 2 when built with
// -DSEDI_CG_XGROUP=2): an iteration that is not a multiple of CG_XG leaves x
// alone and records its alpha as pending (apend); the group's last iteration
// adds the pending apend_j * p_j and its own alpha * p in one read-modify-write
// of x, in the order of the iterations, reading the pending directions from
// the other direction buffers, which still hold them.  Per iteration that is
// 6 + (2 + CG_XG - 1) / CG_XG vector streams instead of 8: 7.25 with groups
// of 4, 7.5 with pairs.  Whatever reads x while updates are
// pending adds them: the fold ahead of the FP64 replacement, the
// accumulation and the unpacking of x, and -- for a right-hand side that
// finished inside a group while its block goes on -- the next direction pass.
// The pending directions are always in buffers p1 .. p_{CG_XG-1}.
//
// Per-RHS state makes each probe's trajectory independent of the others in
// its batch: an RHS that is done keeps x and r (alpha = beta = 0),
// replacements are applied per RHS, and each RHS's sums have a fixed order
// for a given grid.
#include "batched_common.h"

#include <atomic>
#include <cmath>
#include <cstring>

namespace sedi {

static std::atomic<bool> g_compact{true};
bool batch_compact_enabled() { return g_compact.load(); }
void set_batch_compact_enabled(bool on) { g_compact.store(on); }

// control block layout: per-RHS doubles [CD_NDOUBLE][K], then per-RHS ints
// [CI_NINT][K] (c.bb and c.fiter are the two bases)
enum {
    CD_BB, CD_RU, CD_RUOLD, CD_ALPHA, CD_BETA, CD_ERR, CD_FERR, CD_ELAST,
    CD_GAMMA, CD_DELTA, CD_ERRP, CD_GAMMAP, CD_ALPHAP, CD_RHO, CD_RHOP, CD_OMEGA, CD_SS, CD_CBW, CD_TOL,
    CD_APEND, CD_APEND2, CD_APEND3, CD_NDOUBLE
};
enum { CI_FITER, CI_CODE, CI_DONE, CI_REP, CI_FOLD, CI_SEXIT, CI_MAXIT, CI_PLANE, CI_NINT };

// ---- set-up: r = b (x0 = 0), x = 0, <b,b>, <r, r.*iM> -----------------------
// Cg.hpp:111-147, PipelinedCg.hpp:113-136 and BiCgStab.hpp:117-163 for x = 0:
// A x = 0, so r = b exactly; BiCGStab also takes r^ = r (:154).
template <class T>
__global__ __launch_bounds__(NTHREADS) void batch_init_kernel(long long n, BCtl c, int rb0, double tol,
                                                              int maxit, const V4<T>* b, V4<T>* r,
                                                              V4<T>* rhat, V4<T>* xlo, double* xhi, T im) {
    __shared__ double red[8 * NWAVES];
    __shared__ int lastflag;
    const int rb = rb0 + blockIdx.y;
    const V4<T>* bq = b + (long long)rb * n;
    V4<T>* rq = r + (long long)rb * n;
    V4<T>* xq = xlo + (long long)rb * n;
    double acc[8];
#pragma unroll
    for (int d = 0; d < 8; d++) acc[d] = 0.0;
    for (long long i = (long long)blockIdx.x * NTHREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * NTHREADS) {
        const V4<T> v = bq[i];
        rq[i] = v;
        if (rhat) rhat[(long long)rb * n + i] = v;
        xq[i] = V4<T>{};
        if (xhi)
#pragma unroll
            for (int q = 0; q < 4; q++) xhi[(long long)(4 * rb + q) * n + i] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double bv = (double)v[q];
            const T u = v[q] * im;
            acc[q] += bv * bv;
            acc[4 + q] += bv * (double)u;
        }
    }
    if (!grid_sum<8>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        double bb = acc[q];
        if (bb == 0.0) bb = 1.0;  // Cg.hpp:115
        const double e = sqrt(acc[q] / bb);
        c.tolv[k] = tol;
        c.maxv[k] = maxit;
        c.bb[k] = bb;
        c.err[k] = e;
        c.elast[k] = e;
        c.ru[k] = acc[4 + q];
        c.ru_old[k] = acc[4 + q];
        c.alpha[k] = 0.0;
        c.beta[k] = 0.0;
        for (int j = 0; j < CG_NPEND; j++) c.apend[j * c.K + k] = 0.0;
        c.rep[k] = 0;
        c.fold[k] = 0;
        c.sexit[k] = 0;
        c.done[k] = 0;
        c.plane[k] = k;
        if (c.kind == SEDI_BICGSTAB) {
            c.rho[k] = acc[q];  // <r, r^> with r^ = r (:160)
            c.rho_prev[k] = 0.0;
            c.omega[k] = 1.0;
            c.ss[k] = 0.0;
            c.cbw[k] = 0.0;
        }
        if (e < c.tolv[k]) {  // Cg.hpp:130-134 (strict <)
            bset_done(c, k, 0, 0, e);
        } else if (c.kind == SEDI_BICGSTAB) {
            if (c.maxv[k] < 1) bbicg_outcome(c, k, c.maxv[k] + 1);
            else if (c.rho[k] == 0.0) bbicg_outcome(c, k, 1);  // :163
        } else if (c.maxv[k] < 1) {
            bcg_exhausted(c, k, e);
        }
    }
    c.iter[rb] = 1;
    brb_update(c, rb);
}

// ---- CG direction: p = u + beta p, s = A p, <s, p> (Cg.hpp:152-166,175,188) --
// A lane that finished while its block goes on still has its last updates
// pending: this pass adds them to x (flush, the lanes' pending coefficients,
// reading p1 .. p_{CG_NPEND} at the cell; only when some lane needs it).
template <class T>
struct PassBDir {
    using V = V4<T>;
    static constexpr int ND = 4, D = SEDI_PREFETCH_BATCH;
    T a, b, c;
    const V* r;
    const V* pold;
    V* pnew;
    V* x;
    T im;
    V beta;
    V flush[CG_NPEND];  // pending coefficients of lanes that finished (else 0)
    const V* pend[CG_NPEND];
    bool first, any_flush;
    struct Raw { V r, p; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(pold, i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ V f(const Raw& q) const {
        const V u = im * q.r;
        return first ? u : u + beta * q.p;
    }
    __device__ void finish(cidx_t i, const Raw& q, const Aux&, const V& p, const V& s,
                           double (&acc)[ND]) const {
        if (any_flush) {
            V xv = ldc((const V*)x, i);
#pragma unroll
            for (int j = 0; j < CG_NPEND; j++) xv = xv + flush[j] * ldc(pend[j], i);
            stv(x, i, xv);
        }
        stv(pnew, i, p);
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += (double)s[q] * (double)p[q];
    }
};

template <class T, int ZC, int TYK>
__global__ __launch_bounds__(TX * TYK, CgDirTile<T>::min_waves) void batch_dir_kernel(
    TileGrid g, BCtl c, T a, T b, T cc, long long n, const V4<T>* r0, const V4<T>* r1, CgDirs<T> P4,
    V4<T>* x, T im) {
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<T> lds[2][TYK + 2][TX + 2];
    __shared__ double red[4 * TYK];
    __shared__ int lastflag;
    const int it = c.iter[rb];
    const long long off = (long long)rb * n;
    V4<T> beta, flush[CG_NPEND];
    bool any_flush = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool d = c.done[4 * rb + q];
        beta[q] = d ? (T)0 : (T)c.beta[4 * rb + q];
        // a lane that finished with updates pending
#pragma unroll
        for (int j = 0; j < CG_NPEND; j++) {
            flush[j][q] = d ? (T)c.apend[j * c.K + 4 * rb + q] : (T)0;
            any_flush |= flush[j][q] != (T)0;
        }
    }
    PassBDir<T> P{a, b, cc, ((it & 1) ? r1 : r0) + off, P4.at((it + CG_XG - 1) % CG_XG) + off,
                  P4.at(it % CG_XG) + off, x + off, im, beta, {}, {}, it == 1, any_flush};
#pragma unroll
    for (int j = 0; j < CG_NPEND; j++) {
        P.flush[j] = flush[j];
        P.pend[j] = P4.p[j + 1] + off;
    }
    double acc[4];
    run_tiles<PassBDir<T>, ZC, TYK>(P, g, lds, acc, false);
    if (!grid_sum<4, TYK>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        c.alpha[k] = c.done[k] ? 0.0 : c.ru[k] / acc[q];  // Cg.hpp:166
        if (c.done[k])
            for (int j = 0; j < CG_NPEND; j++) c.apend[j * c.K + k] = 0.0;
    }
}

// ---- CG update: x += alpha p, r -= alpha s, <r,u>, <r,r> (Cg.hpp:168-193) ---
// XS: the group's last iteration, which updates x (a compile-time choice: the
// other iterations keep no registers for x and the pending directions)
template <class T, bool XS>
struct PassBUpd {
    using V = V4<T>;
    static constexpr int ND = 8, D = XS ? SEDI_PREFETCH_BATCH : SEDI_PREFETCH_CG_UPD;
    T a, b, c;
    const V* p;
    const V* pend[CG_NPEND];  // the pending directions p1 .. (x steps)
    V* xlo;
    const V* r;
    V* rnext;  // the next iteration's residual (the other buffer: stores stream
               // to fresh lines instead of back into the lines being read)
    T im;
    V alpha, apend[CG_NPEND];
    static constexpr bool xstep = XS;  // it % CG_XG == 0: x += apend_1 p_1 + ... + alpha p
    struct Raw { V p; };
    struct Aux { V x, pp[CG_NPEND], r; };
    __device__ Raw load(cidx_t i) const { return {ldc(p, i)}; }
    __device__ Aux load_aux(cidx_t i) const {
        Aux o;
        if (xstep) {
            o.x = ldc((const V*)xlo, i);
#pragma unroll
            for (int j = 0; j < CG_NPEND; j++) o.pp[j] = ldc(pend[j], i);
        }
        o.r = ldc(r, i);
        return o;
    }
    __device__ V f(const Raw& q) const { return q.p; }
    __device__ void finish(cidx_t i, const Raw&, const Aux& o, const V& pv, const V& s,
                           double (&acc)[ND]) const {
        if (xstep) {
            V xv = o.x;
#pragma unroll
            for (int j = 0; j < CG_NPEND; j++) xv = xv + apend[j] * o.pp[j];
            stv(xlo, i, xv + alpha * pv);
        }
        const V rn = o.r - alpha * s;
        stv(rnext, i, rn);
        const V u = im * rn;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            acc[q] += (double)rn[q] * (double)u[q];
            acc[4 + q] += (double)rn[q] * (double)rn[q];
        }
    }
};

template <class T, int ZC, int TYK, bool XS>
__device__ __forceinline__ void run_upd(const TileGrid& g, const BCtl& c, T a, T b, T cc, long long off, int it,
                                        const CgDirs<T>& P4, V4<T>* xlo, V4<T>* r0, V4<T>* r1, T im,
                                        V4<T> (&lds)[2][TYK + 2][TX + 2], double (&acc)[8]) {
    const int rb = blockIdx.y;
    PassBUpd<T, XS> P;
    P.a = a; P.b = b; P.c = cc;
    P.p = P4.at(it % CG_XG) + off;
#pragma unroll
    for (int q = 0; q < 4; q++) P.alpha[q] = uniform((T)c.alpha[4 * rb + q]);
#pragma unroll
    for (int j = 0; j < CG_NPEND; j++) {
        P.pend[j] = P4.p[j + 1] + off;
#pragma unroll
        for (int q = 0; q < 4; q++) P.apend[j][q] = XS ? uniform((T)c.apend[j * c.K + 4 * rb + q]) : (T)0;
    }
    P.xlo = xlo + off;
    P.r = ((it & 1) ? r1 : r0) + off;
    P.rnext = ((it & 1) ? r0 : r1) + off;
    P.im = im;
    run_tiles<PassBUpd<T, XS>, ZC, TYK>(P, g, lds, acc, true);
}

template <class T, int ZC, int TYK>
__global__ __launch_bounds__(TX * TYK, CgTile<T>::min_waves) void batch_upd_kernel(
    TileGrid g, BCtl c, T a, T b, T cc, long long n, CgDirs<T> P4, V4<T>* xlo, V4<T>* r0, V4<T>* r1, T im) {
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<T> lds[2][TYK + 2][TX + 2];
    __shared__ double red[8 * TYK];
    __shared__ int lastflag;
    const int it = c.iter[rb];
    const long long off = (long long)rb * n;
    const int jp = it % CG_XG;
    double acc[8];
    if (jp == 0) run_upd<T, ZC, TYK, true>(g, c, a, b, cc, off, it, P4, xlo, r0, r1, im, lds, acc);
    else run_upd<T, ZC, TYK, false>(g, c, a, b, cc, off, it, P4, xlo, r0, r1, im, lds, acc);
    if (!grid_sum<8, TYK>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        // this alpha (0 for a finished lane) waits for the group's last
        // iteration; at that one every update of the group is in x
        if (jp == 0)
            for (int j = 0; j < CG_NPEND; j++) c.apend[j * c.K + k] = 0.0;
        else
            c.apend[(jp - 1) * c.K + k] = c.alpha[k];
        if (c.done[k]) continue;
        const double ru_new = acc[q];
        c.ru_old[k] = c.ru[k];
        c.beta[k] = ru_new / c.ru[k];  // Cg.hpp:185
        c.ru[k] = ru_new;
        const double e = sqrt(acc[4 + q] / c.bb[k]);  // :191
        c.err[k] = e;
        c.hist[(long long)k * c.hstride + it - 1] = e;
        if (c.mixed) {
            // converged on the recursive residual, a replacement due, or the
            // last iteration: the FP64 pass decides.  elast is the largest
            // residual since the last replacement (reliable updates compare
            // against it, Clark et al. 2010).
            c.elast[k] = fmax(c.elast[k], e);
            c.rep[k] = (e <= c.tolv[k]) || (it % c.rep_every == 0) || (it >= c.maxv[k]) ||
                       (e < c.drop * c.elast[k]);
        } else if (e <= c.tolv[k]) {
            bset_done(c, k, 0, it, e);
        } else if (it >= c.maxv[k]) {
            bset_done(c, k, 1, c.maxv[k], e);
        }
    }
    c.iter[rb] = it + 1;
    brb_update(c, rb);
}

// ---- FP64 replacement: r = b - A x_hi ----------------------------------------
// The new residual is computed in FP64 from x_hi, which holds all of x for
// the right-hand sides being replaced: CG folds x_lo and its pending updates
// into x_hi first (batch_fold_kernel, a pointwise pass, so no tile of this
// pass can see a half-folded neighbour), and BiCGStab keeps x in FP64 (x_lo
// stays zero).  The other lanes of a block take part in the stencil but
// their results are discarded.  r is stored rounded to the vector type.  The
// second dot is <r, u> (CG) or <r, r^> (BiCGStab, RHAT).
template <class T, bool RHAT>
struct PassBRep {
    using V = V4<double>;
    static constexpr int ND = 8, D = SEDI_PREFETCH_BATCH;
    double a, b, c;
    const double* xh[4];  // per-RHS FP64 parts of x of this block ([K][N] layout)
    V4<T>* r;
    const V4<T>* rhs;
    V4<T>* rhat;
    T im;
    int rep[4];
    bool reset;  // BiCGStab: r^ restarts from the replaced residual
    struct Raw { V xh; };
    struct Aux { V4<T> b, r, rh; };
    __device__ Raw load(cidx_t i) const {
        Raw q;
#pragma unroll
        for (int k = 0; k < 4; k++) q.xh[k] = ldc(xh[k], i);
        return q;
    }
    __device__ Aux load_aux(cidx_t i) const {
        Aux o;
        o.b = ldc(rhs, i);
        o.r = ldc((const V4<T>*)r, i);
        if (RHAT) o.rh = ldc((const V4<T>*)rhat, i);
        return o;
    }
    __device__ V f(const Raw& q) const { return q.xh; }
    __device__ void finish(cidx_t i, const Raw&, const Aux& o, const V&, const V& Ax,
                           double (&acc)[ND]) const {
        V4<T> rn = o.r;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (rep[k]) rn[k] = (T)((double)o.b[k] - Ax[k]);  // Cg.hpp:123, in FP64
        stv(r, i, rn);
        V4<T> u = RHAT ? o.rh : im * rn;
        if (RHAT && reset) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (rep[k]) u[k] = rn[k];
            stv(rhat, i, u);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] += (double)rn[k] * (double)u[k];
            acc[4 + k] += (double)rn[k] * (double)rn[k];
        }
    }
};

template <class T, bool RHAT, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_rep_kernel(TileGrid g, BCtl c, double a, double b,
                                                             double cc, long long n, const double* xhi,
                                                             V4<T>* r0, V4<T>* r1, const V4<T>* rhs, V4<T>* rhat,
                                                             T im) {
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    const int* rp = c.rep + 4 * rb;
    if (!(rp[0] | rp[1] | rp[2] | rp[3])) return;
    __shared__ V4<double> lds[2][TY + 2][TX + 2];
    __shared__ double red[8 * NWAVES];
    __shared__ int lastflag;
    const long long off = (long long)rb * n;
    // the residual the next iteration reads: it alternates with the parity of
    // the iteration (r of iteration it+1 = iter[rb]: r1 when odd)
    V4<T>* r = (r1 && (c.iter[rb] & 1)) ? r1 : r0;
    PassBRep<T, RHAT> P;
    P.a = a; P.b = b; P.c = cc;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        P.xh[k] = xhi + (long long)(4 * rb + k) * n;
        P.rep[k] = rp[k];
    }
    P.r = r + off;
    P.rhs = rhs + off;
    P.rhat = RHAT ? rhat + off : nullptr;
    P.im = im;
    P.reset = c.bicg_reset != 0;
    double acc[8];
    run_tiles<PassBRep<T, RHAT>, ZC>(P, g, lds, acc, false);
    if (!grid_sum<8>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    const int it = c.iter[rb] - 1;  // the iteration that requested the update
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        if (!c.rep[k]) continue;
        c.rep[k] = 0;
        const double e = sqrt(acc[4 + q] / c.bb[k]);
        const bool drifted = e > c.restart * c.err[k];
        c.err[k] = e;
        c.elast[k] = e;
        c.hist[(long long)k * c.hstride + it - 1] = e;
        if (!RHAT) {
            c.ru[k] = acc[q];
            // beta from the replaced residual, or a restart if the recursion drifted
            c.beta[k] = drifted ? 0.0 : acc[q] / c.ru_old[k];
            if (e <= c.tolv[k]) bset_done(c, k, 0, it, e);
            else if (it >= c.maxv[k]) bset_done(c, k, 1, c.maxv[k], e);
        } else {
            // BiCgStab.hpp:229-232 and :160-169 of the next iteration on the
            // replaced residual
            c.rho[k] = acc[q];  // <r, r^> (with a restarted r^ = r: <r, r>)
            if (e <= c.tolv[k]) { bbicg_outcome(c, k, it); continue; }
            if (it >= c.maxv[k]) { bbicg_outcome(c, k, c.maxv[k] + 1); continue; }
            if (c.rho[k] == 0.0) { bbicg_outcome(c, k, it + 1); continue; }
            const double beta = (drifted || c.bicg_reset) ? 0.0
                                                          : (c.rho[k] / c.rho_prev[k]) * (c.alpha[k] / c.omega[k]);
            c.beta[k] = beta;
            c.cbw[k] = -beta * c.omega[k];
        }
    }
    brb_update(c, rb);
}

// ---- x_hi += x_lo (+ the pending updates), x_lo = 0 -------------------------
// Runs ahead of a replacement, for the right-hand sides that asked for one
// (c.rep; the replacement clears it) and for pipelined CG's right-hand sides
// that finished with updates pending (c.fold, cleared here): their x is then
// final, and the same whether their block goes on (it would otherwise apply
// them at its next group's end, rounded to FP32) or not (the read-out adds
// them in FP64).  Runs for a block whether or not it is done.  x_hi + x_lo is
// unchanged up to the one FP64 rounding of the sum, which the final
// x = x_hi + x_lo would make anyway.
template <class T>
__global__ __launch_bounds__(NTHREADS) void batch_fold_kernel(long long n, BCtl c, double* xhi,
                                                              V4<T>* xlo, PendDirs<T> pend) {
    const int rb = blockIdx.y;
    const int* rp = c.rep + 4 * rb;
    int* fp = c.fold + 4 * rb;
    const int fl[4] = {rp[0] | fp[0], rp[1] | fp[1], rp[2] | fp[2], rp[3] | fp[3]};
    if (!(fl[0] | fl[1] | fl[2] | fl[3])) return;
    __shared__ double red[NWAVES];
    __shared__ int lastflag;
    double ap[CG_NPEND][4];
    const V4<T>* pq[CG_NPEND];
#pragma unroll
    for (int j = 0; j < CG_NPEND; j++) {
        bool pending = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            ap[j][k] = (pend.p[j] && fl[k]) ? c.apend[j * c.K + 4 * rb + k] : 0.0;
            pending |= ap[j][k] != 0.0;
        }
        pq[j] = pending ? pend.p[j] + (long long)rb * n : nullptr;
    }
    V4<T>* xq = xlo + (long long)rb * n;
    for (long long i = (long long)blockIdx.x * NTHREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * NTHREADS) {
        V4<T> xl = xq[i];
        V4<T> pp[CG_NPEND];
#pragma unroll
        for (int j = 0; j < CG_NPEND; j++) pp[j] = pq[j] ? pq[j][i] : V4<T>{};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (fl[k]) {
                double* h = xhi + (long long)(4 * rb + k) * n + i;
                // x_hi + x_lo (+ CG's pending apend_j * p_j, in the order of the updates)
                double v = *h + (double)xl[k];
#pragma unroll
                for (int j = 0; j < CG_NPEND; j++)
                    if (pq[j] && ap[j][k] != 0.0) v = v + ap[j][k] * (double)pp[j][k];
                *h = v;
                xl[k] = (T)0;
            }
        }
        xq[i] = xl;
    }
    double acc[1] = {0.0};
    if (!grid_sum<1>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int k = 0; k < 4; k++)
        if (fl[k] && pend.p[0])
            for (int j = 0; j < CG_NPEND; j++) c.apend[j * c.K + 4 * rb + k] = 0.0;
    fp[0] = fp[1] = fp[2] = fp[3] = 0;
}

// ---- lane compaction ---------------------------------------------------------
// A block runs while any of its 4 lanes does, so a right-hand side that
// finished early still costs its share of every pass until the slowest lane
// of its block converges.  Between iterations batch_compact_kernel plans
// moves among the running blocks that share an iteration count (in a solve,
// all of them): while the live lanes of the emptiest block fit into finished
// lanes of the others, each trades places with a finished lane there, and the
// emptied block retires.  batch_swap_kernel then exchanges the two lanes'
// values in every vector, the FP64 rows of x and the error histories; the
// planning kernel has already exchanged their per-RHS scalars.  A lane's
// values go through the same operations in any lane and block (each lane's
// sums have a fixed order per grid), and the two blocks are at the same
// iteration, so a moved right-hand side continues exactly as if it had
// stayed; plane[] records which right-hand side each lane holds.
constexpr int COMPACT_MAX_RB = 256;

__global__ __launch_bounds__(256) void batch_compact_kernel(BCtl c, int nrb) {
    __shared__ int live[COMPACT_MAX_RB];                // live lanes per block (0: not running)
    __shared__ int itr[COMPACT_MAX_RB];                 // its iteration
    __shared__ unsigned char recv[COMPACT_MAX_RB];      // received a lane in this plan
    __shared__ unsigned char occ[4 * COMPACT_MAX_RB];   // lane holds a live RHS after the plan
    __shared__ int pairs[2 * 4 * COMPACT_MAX_RB];
    __shared__ int npairs;
    const int t = threadIdx.x;
    const int K = c.K;
    for (int rb = t; rb < nrb; rb += blockDim.x) {
        int l = 0;
        for (int q = 0; q < 4; q++) {
            const int on = !c.done[4 * rb + q];
            occ[4 * rb + q] = (unsigned char)on;
            l += on;
        }
        live[rb] = c.rb_done[rb] ? 0 : l;
        itr[rb] = c.iter[rb];
        recv[rb] = 0;
    }
    __syncthreads();
    if (t == 0) {
        int np = 0;
        for (;;) {
            // donor: the running block with the fewest live lanes that has
            // received none (if its lanes do not fit elsewhere, no block's do)
            int d = -1;
            for (int rb = 0; rb < nrb; rb++)
                if (live[rb] > 0 && !recv[rb] && (d < 0 || live[rb] < live[d])) d = rb;
            if (d < 0) break;
            int room = 0;
            for (int rb = 0; rb < nrb; rb++)
                if (rb != d && live[rb] > 0 && itr[rb] == itr[d]) room += 4 - live[rb];
            if (room < live[d]) break;
            for (int qd = 0; qd < 4; qd++) {
                const int kd = 4 * d + qd;
                if (!occ[kd]) continue;
                int r = -1;  // receiver: the fullest block with room
                for (int rb = 0; rb < nrb; rb++)
                    if (rb != d && live[rb] > 0 && live[rb] < 4 && itr[rb] == itr[d] && (r < 0 || live[rb] > live[r]))
                        r = rb;
                int ks = 4 * r;
                while (occ[ks]) ks++;
                pairs[2 * np] = kd;
                pairs[2 * np + 1] = ks;
                np++;
                occ[ks] = 1;
                occ[kd] = 0;
                live[r]++;
                recv[r] = 1;
            }
            live[d] = -1;  // retires
        }
        npairs = np;
        c.swaps[0] = np;
        for (int j = 0; j < 2 * np; j++) c.swaps[1 + j] = pairs[j];
    }
    __syncthreads();
    const int np = npairs;
    if (np == 0) return;
    double* dbase = c.bb;  // CD_BB = 0
    int* ibase = c.fiter;  // CI_FITER = 0
    for (int p = t; p < np; p += blockDim.x) {
        const int a = pairs[2 * p], b = pairs[2 * p + 1];
        for (int f = 0; f < CD_NDOUBLE; f++) {
            const double v = dbase[f * K + a];
            dbase[f * K + a] = dbase[f * K + b];
            dbase[f * K + b] = v;
        }
        for (int f = 0; f < CI_NINT; f++) {
            const int v = ibase[f * K + a];
            ibase[f * K + a] = ibase[f * K + b];
            ibase[f * K + b] = v;
        }
    }
    if (t == 0) {
        for (int rb = 0; rb < nrb; rb++)
            if (live[rb] < 0) c.rb_done[rb] = 1;  // every lane of it now holds a finished RHS
        *c.moves += (unsigned)np;
    }
}

struct VSlots {
    void* p[BATCH_MAX_SLOTS];
};

// Grid: x over cells, y over the vector slots to exchange plus one for the
// FP64 rows of x and the histories.  The two lanes of a pair are distinct
// from every other pair's (each lane moves at most once per plan), so all
// loads of a group of pairs can be in flight before their stores.
template <class T>
__global__ __launch_bounds__(NTHREADS) void batch_swap_kernel(BCtl c, long long n, VSlots v, int nslot,
                                                              double* xhi) {
    __shared__ int pairs[2 * 4 * COMPACT_MAX_RB];
    const int np = c.swaps[0];
    if (np == 0) return;
    for (int j = threadIdx.x; j < 2 * np; j += blockDim.x) pairs[j] = c.swaps[1 + j];
    __syncthreads();
    const long long g0 = (long long)blockIdx.x * NTHREADS + threadIdx.x;
    const long long stride = (long long)gridDim.x * NTHREADS;
    constexpr int G = 4;
    if ((int)blockIdx.y < nslot) {
        T* base = static_cast<T*>(v.p[blockIdx.y]);
        for (long long i = g0; i < n; i += stride) {
            for (int p0 = 0; p0 < np; p0 += G) {
                long long ia[G], ib[G];
                T va[G], vb[G];
#pragma unroll
                for (int j = 0; j < G; j++) {
                    if (p0 + j < np) {
                        const int a = pairs[2 * (p0 + j)], b = pairs[2 * (p0 + j) + 1];
                        ia[j] = ((long long)(a >> 2) * n + i) * 4 + (a & 3);
                        ib[j] = ((long long)(b >> 2) * n + i) * 4 + (b & 3);
                        va[j] = base[ia[j]];
                        vb[j] = base[ib[j]];
                    }
                }
#pragma unroll
                for (int j = 0; j < G; j++) {
                    if (p0 + j < np) {
                        base[ia[j]] = vb[j];
                        base[ib[j]] = va[j];
                    }
                }
            }
        }
        return;
    }
    if (xhi) {
        for (long long i = g0; i < n; i += stride) {
            for (int p = 0; p < np; p++) {
                const long long a = pairs[2 * p], b = pairs[2 * p + 1];
                const double xa = xhi[a * n + i];
                xhi[a * n + i] = xhi[b * n + i];
                xhi[b * n + i] = xa;
            }
        }
    }
    for (long long j = g0; j < c.hstride; j += stride) {
        for (int p = 0; p < np; p++) {
            const long long a = pairs[2 * p], b = pairs[2 * p + 1];
            const double h = c.hist[a * c.hstride + j];
            c.hist[a * c.hstride + j] = c.hist[b * c.hstride + j];
            c.hist[b * c.hstride + j] = h;
        }
    }
}

// ---- t += x.*v, q += v.*v over the batch's probes in order ------------------
// x = x_hi + x_lo (+ CG's pending apend_j * p_j: pend.p[j] non-null); the probes are
// the lanes of `lanes`, in its order
struct LaneList {
    int n;
    int lane[256];
};

template <class T>
__global__ void batch_accum_kernel(long long n, LaneList lanes, const V4<T>* xlo, const double* xhi,
                                   const V4<T>* b, PendDirs<T> pend, const double* apend, int K, double* t,
                                   double* q) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double tv = t[i], qv = q[i];
        for (int j = 0; j < lanes.n; j++) {
            const int k = lanes.lane[j];
            const long long cell = (long long)(k >> 2) * n + i;
            double x = (xhi ? xhi[(long long)k * n + i] : 0.0) + (double)xlo[cell][k & 3];
#pragma unroll
            for (int jp = 0; jp < CG_NPEND; jp++)
                if (pend.p[jp] && apend[jp * K + k] != 0.0)
                    x = x + apend[jp * K + k] * (double)pend.p[jp][cell][k & 3];
            const double v = (double)b[cell][k & 3];
            tv = tv + x * v;  // stochastic-estimator.cc:195
            qv = qv + v * v;  // :198
        }
        t[i] = tv;
        q[i] = qv;
    }
}

// ---- [K][N] FP64 <-> batch layout ------------------------------------------
template <class T>
__global__ void batch_pack_kernel(long long n, int kact, long long row_stride, const double* __restrict__ src,
                                  V4<T>* dst) {
    const int rb = blockIdx.y;
    V4<T>* d = dst + (long long)rb * n;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        V4<T> v;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = 4 * rb + q;
            v[q] = k < kact ? (T)src[(long long)k * row_stride + i] : (T)0;
        }
        d[i] = v;
    }
}

// Rows k0 .. k1-1 of the batch (x = x_hi + x_lo, or the right-hand sides when
// xhi is null and `src` is b) into [k1-k0][n] FP64.  Grid y = RHS blocks.
// (plus CG's pending apend_j * p_j for x: pend.p[j] non-null, offset like src;
// apend indexed like xhi, with stride K between the pending directions)
template <class T>
__global__ void batch_unpack_kernel(long long n, int k0, int k1, const V4<T>* __restrict__ src,
                                    const double* __restrict__ xhi, PendDirs<T> pend,
                                    const double* __restrict__ apend, int K, double* __restrict__ x) {
    const int rb = blockIdx.y;
    const V4<T>* s = src + (long long)rb * n;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const V4<T> v = s[i];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const long long k = 4 * rb + q;
            if (k >= k0 && k < k1) {
                double xv = (xhi ? xhi[k * n + i] : 0.0) + (double)v[q];
#pragma unroll
                for (int jp = 0; jp < CG_NPEND; jp++)
                    if (pend.p[jp] && apend[jp * K + k] != 0.0)
                        xv = xv + apend[jp * K + k] * (double)pend.p[jp][(long long)rb * n + i][q];
                x[(k - k0) * n + i] = xv;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// The direction buffers that hold CG's or pipelined CG's pending updates
// (p1 .. p_{CG_NPEND}); null for BiCGStab, whose x is always current.
template <class T>
static PendDirs<T> pending_dirs(const BRun& r) {
    PendDirs<T> d;
    for (int j = 0; j < CG_NPEND; j++)
        d.p[j] = r.c.kind == SEDI_CG    ? static_cast<const V4<T>*>(r.v[cg_pslot(j + 1)])
                 : r.c.kind == SEDI_PCG ? static_cast<const V4<T>*>(r.v[pcg_pslot(j + 1)])
                                        : nullptr;
    return d;
}

template <class T>
static CgDirs<T> cg_dirs(const BRun& r) {
    CgDirs<T> d;
    for (int j = 0; j < 4; j++) d.p[j] = j < CG_XG ? static_cast<V4<T>*>(r.v[cg_pslot(j)]) : nullptr;
    return d;
}

template <class T, bool RHAT>
static hipError_t launch_rep(const BRun& r, V4<T>* r0, V4<T>* r1, V4<T>* rhat) {
    const KGeom g = bgeom("batch_rep_kernel", batch_rep_kernel<T, RHAT, 16>, r.shape, SEDI_PREFETCH_BATCH);
    const dim3 grid(g.nblocks, r.nrb);
    const T im = (T)r.im;
    const double* xhi = r.xhi;
    const V4<T>* b = static_cast<const V4<T>*>(r.v[SLOT_B]);
    if (g.tg.zc == 8)
        return launch("batch_rep_kernel", batch_rep_kernel<T, RHAT, 8>, grid, dim3(NTHREADS), r.stream, g.tg, r.c,
                      r.a, r.b, r.cc, r.n, xhi, r0, r1, b, rhat, im);
    if (g.tg.zc == 32)
        return launch("batch_rep_kernel", batch_rep_kernel<T, RHAT, 32>, grid, dim3(NTHREADS), r.stream, g.tg,
                      r.c, r.a, r.b, r.cc, r.n, xhi, r0, r1, b, rhat, im);
    return launch("batch_rep_kernel", batch_rep_kernel<T, RHAT, 16>, grid, dim3(NTHREADS), r.stream, g.tg, r.c,
                  r.a, r.b, r.cc, r.n, xhi, r0, r1, b, rhat, im);
}

template <class T>
static hipError_t launch_fold(const BRun& r) {
    return launch("batch_fold_kernel", batch_fold_kernel<T>, dim3(ew_blocks(r.n), r.nrb), dim3(NTHREADS), r.stream,
                  r.n, r.c, r.xhi, static_cast<V4<T>*>(r.v[SLOT_XLO]), pending_dirs<T>(r));
}

template <class T>
int fold_lanes(const BRun& r) {
    SEDI_HIP(launch_fold<T>(r));
    return SEDI_OK;
}

// CG: res is r0 and rhat is null; x_lo and the pending updates of the
// right-hand sides to replace are folded into x_hi first (fold).  BiCGStab:
// res is r0 and rhat r^; its x is in x_hi already.  The kernel picks r0 / r1
// by the parity of the next iteration.
template <class T>
int replace_and_fold(const BRun& r, V4<T>* res, V4<T>* rhat, bool fold) {
    if (fold) SEDI_HIP(launch_fold<T>(r));
    hipError_t e;
    if (rhat) e = launch_rep<T, true>(r, res, static_cast<V4<T>*>(r.v[BI_R1]), rhat);
    else e = launch_rep<T, false>(r, res, static_cast<V4<T>*>(r.v[CG_R1]), nullptr);
    if (e != hipSuccess) {
        set_error("replacement launch failed: %s", hipGetErrorString(e));
        return SEDI_ERR_HIP;
    }
    return SEDI_OK;
}

// The update pass with TYK-row tiles: its geometry and a launch of the
// kernel for the geometry's z chunk.
template <class T, int TYK>
static hipError_t launch_upd(const BRun& r, T a, T b, T cc, T im, CgDirs<T> P4, V4<T>* XL, V4<T>* R0, V4<T>* R1) {
    const KGeom g = bgeom("batch_upd_kernel", batch_upd_kernel<T, 16, TYK>, r.shape, SEDI_PREFETCH_BATCH, TYK);
    const dim3 grid(g.nblocks, r.nrb), block(TX * TYK);
    auto kern = g.tg.zc == 8 ? batch_upd_kernel<T, 8, TYK>
                             : (g.tg.zc == 32 ? batch_upd_kernel<T, 32, TYK> : batch_upd_kernel<T, 16, TYK>);
    return launch("batch_upd_kernel", kern, grid, block, r.stream, g.tg, r.c, a, b, cc, r.n, P4, XL, R0, R1, im);
}

// The direction pass with TYK-row tiles (8: the default in FP32 storage; 16
// halves the y-halo re-reads of r and p, chosen by the tuning table's
// "ty:batch_dir_kernel"; FP64 blocks take SEDI_FP64_BATCH_TY rows).
template <class T, int TYK>
static hipError_t launch_dir(const BRun& r, T a, T b, T cc, T im, const V4<T>* R0, const V4<T>* R1, CgDirs<T> P4,
                             V4<T>* XL) {
    const KGeom g = bgeom("batch_dir_kernel", batch_dir_kernel<T, 16, TYK>, r.shape, SEDI_PREFETCH_BATCH, TYK);
    const dim3 grid(g.nblocks, r.nrb), block(TX * TYK);
    auto kern = g.tg.zc == 8 ? batch_dir_kernel<T, 8, TYK>
                             : (g.tg.zc == 32 ? batch_dir_kernel<T, 32, TYK> : batch_dir_kernel<T, 16, TYK>);
    return launch("batch_dir_kernel", kern, grid, block, r.stream, g.tg, r.c, a, b, cc, r.n, R0, R1, P4, XL, im);
}

// Rows per tile of the FP64 batched CG passes: 8.  Their registers (about 180
// VGPRs) allow two waves per SIMD, i.e. one 8-wave workgroup per CU or two
// 4-wave ones; the taller tile halves the y-halo re-reads (update pass 7.47 ->
// 6.83 ms at 256^3, k = 64; the direction pass unchanged).
#ifndef SEDI_FP64_BATCH_TY
#define SEDI_FP64_BATCH_TY TY_TALL
#endif

static int dir_rows(const Shape& s) {
    return tuned_value("ty:batch_dir_kernel", s, 0) == 2 * TY_TALL ? 2 * TY_TALL : TY_TALL;
}

// Rows per tile of the update pass in FP32 storage: 16 (one 16-wave
// workgroup per CU) halve the y-halo re-reads of p against 8 and measured
// 1-3.5% faster on 256^3 and 512^3 grids, but 6-10% slower on 64^3 and 128^3,
// where 16-row tiles leave too few workgroups to fill the GPU; so 16 from 256
// rows up.  The tuning table's "ty:batch_upd_kernel" (8 or 16) overrides.
static int upd_rows(const Shape& s) {
    const int t = tuned_value("ty:batch_upd_kernel", s, 0);
    if (t == 8 || t == 16) return t;
    return s.n2 >= 256 ? 2 * TY_TALL : TY_TALL;
}

template <class T>
static int cg_iteration(const BRun& r) {
    const T a = (T)r.a, b = (T)r.b, cc = (T)r.cc, im = (T)r.im;
    V4<T>* R0 = static_cast<V4<T>*>(r.v[CG_R]);
    V4<T>* R1 = static_cast<V4<T>*>(r.v[CG_R1]);
    const CgDirs<T> P4 = cg_dirs<T>(r);
    V4<T>* XL = static_cast<V4<T>*>(r.v[SLOT_XLO]);
    if constexpr (sizeof(T) == 4) {
        if (dir_rows(r.shape) == 2 * TY_TALL)
            SEDI_HIP((launch_dir<T, 2 * TY_TALL>(r, a, b, cc, im, R0, R1, P4, XL)));
        else
            SEDI_HIP((launch_dir<T, TY_TALL>(r, a, b, cc, im, R0, R1, P4, XL)));
    } else {
        SEDI_HIP((launch_dir<T, SEDI_FP64_BATCH_TY>(r, a, b, cc, im, R0, R1, P4, XL)));
    }
    if constexpr (sizeof(T) == 4) {
        if (upd_rows(r.shape) == 2 * TY_TALL)
            SEDI_HIP((launch_upd<T, 2 * TY_TALL>(r, a, b, cc, im, P4, XL, R0, R1)));
        else
            SEDI_HIP((launch_upd<T, TY_TALL>(r, a, b, cc, im, P4, XL, R0, R1)));
    } else {
        SEDI_HIP((launch_upd<T, SEDI_FP64_BATCH_TY>(r, a, b, cc, im, P4, XL, R0, R1)));
    }
    if (r.mixed) return replace_and_fold<T>(r, R0, nullptr, true);
    return SEDI_OK;
}


static int nvec_of(int kind) {
    return kind == SEDI_CG ? CG_NVEC : (kind == SEDI_PCG ? PCG_NVEC : BI_NVEC);
}

double BatchSolver::bytes_needed(const Shape& s, int kind, int K, bool fp32) {
    const double n = (double)s.n();
    return n * K * ((fp32 ? 4.0 : 8.0) * nvec_of(kind) + (fp32 ? 8.0 : 0.0));
}

BatchSolver::~BatchSolver() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    if (pinned_) (void)hipHostFree(pinned_);
    if (stream_) (void)hipStreamDestroy(stream_);
    delete run_;
}

int BatchSolver::create(const Shape& s, double a, double b, double c, int kind, int K, bool fp32,
                        int rep_every) {
    SEDI_REQUIRE(kind == SEDI_CG || kind == SEDI_PCG || kind == SEDI_BICGSTAB, SEDI_ERR_ARG,
                 "unknown solver kind %d", kind);
    SEDI_REQUIRE(K >= 4 && K % 4 == 0, SEDI_ERR_ARG, "batch size %d must be a positive multiple of 4", K);
    const size_t vb = (fp32 ? 4 : 8) * 4;  // bytes per cell per RHS block
    SEDI_REQUIRE((double)s.n() * vb < 4294967296.0, SEDI_ERR_UNSUPPORTED,
                 "grid too large for the batched path's 32-bit cell offsets");
    shape_ = s;
    a_ = a; b_c_ = b; c_ = c;
    kind_ = kind;
    K_ = K;
    nrb_ = K / 4;
    fp32_ = fp32;
    rep_every_ = rep_every > 0 ? rep_every : 50;
    nvec_ = nvec_of(kind);
    run_ = new BRun();
    SEDI_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SEDI_HIP(hipHostMalloc((void**)&pinned_, sizeof(int) * 2 * nrb_, hipHostMallocDefault));
    const size_t n = (size_t)s.n();
    const size_t vec = n * vb * nrb_;
    const size_t hib = fp32 ? sizeof(double) * n * K : 0;
    SEDI_HIP(vec_.reserve(vec * nvec_ + hib));
    char* base = static_cast<char*>(vec_.p);
    for (int k = 0; k < nvec_; k++) v_[k] = base + k * vec;
    xhi_ = fp32 ? reinterpret_cast<double*>(base + nvec_ * vec) : nullptr;
    SEDI_HIP(hipMemset(vec_.p, 0, vec * nvec_ + hib));
    // control: doubles [CD_NDOUBLE][K], ints [CI_NINT][K], iter[nrb], rb_done[nrb],
    // counters[nrb], swaps[1 + 2K], moves, partials[nrb][12][8*CUs]
    const size_t pst = 12 * (size_t)(8 * batch_cus());
    const size_t bytes = sizeof(double) * CD_NDOUBLE * K + sizeof(int) * (CI_NINT * K + 2 * nrb_) +
                         sizeof(unsigned) * nrb_ + sizeof(int) * (1 + 2 * (size_t)K) + sizeof(unsigned) + 256 +
                         sizeof(double) * pst * nrb_;
    SEDI_HIP(ctl_.reserve(bytes));
    SEDI_HIP(hipMemset(ctl_.p, 0, bytes));
    lane_of_.resize(K);
    for (int k = 0; k < K; k++) lane_of_[k] = k;
    return SEDI_OK;
}

// Fills run_ (the control block's pointers and the pass parameters) for a
// history capacity of at least hist_need iterations per RHS.
template <class T>
int BatchSolver::bind_t(int hist_need) {
    const int K = K_;
    if (hist_cap_ < std::max(hist_need, 1)) {
        SEDI_HIP(hipStreamSynchronize(stream_));
        SEDI_HIP(hist_.reserve(sizeof(double) * (size_t)K * std::max(hist_need, 1)));
        hist_cap_ = std::max(hist_need, 1);
        if (graph_) {  // the graph holds the old history pointer
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
            graph_count_ = 0;
        }
    }
    BRun& run = *run_;
    BCtl& c = run.c;
    double* d = ctl_.as<double>();
    c.bb = d + CD_BB * K; c.ru = d + CD_RU * K; c.ru_old = d + CD_RUOLD * K;
    c.alpha = d + CD_ALPHA * K; c.beta = d + CD_BETA * K; c.err = d + CD_ERR * K; c.ferr = d + CD_FERR * K;
    c.elast = d + CD_ELAST * K;
    c.gamma = d + CD_GAMMA * K; c.delta = d + CD_DELTA * K; c.errp = d + CD_ERRP * K;
    c.gamma_prev = d + CD_GAMMAP * K; c.alpha_prev = d + CD_ALPHAP * K;
    c.rho = d + CD_RHO * K; c.rho_prev = d + CD_RHOP * K; c.omega = d + CD_OMEGA * K; c.ss = d + CD_SS * K;
    c.cbw = d + CD_CBW * K;
    c.apend = d + CD_APEND * K;
    c.tolv = d + CD_TOL * K;
    int* ip = reinterpret_cast<int*>(d + CD_NDOUBLE * K);
    c.fiter = ip + CI_FITER * K; c.code = ip + CI_CODE * K; c.done = ip + CI_DONE * K; c.rep = ip + CI_REP * K;
    c.fold = ip + CI_FOLD * K; c.sexit = ip + CI_SEXIT * K; c.maxv = ip + CI_MAXIT * K;
    c.plane = ip + CI_PLANE * K;
    c.iter = ip + CI_NINT * K;
    c.rb_done = c.iter + nrb_;
    c.counters = reinterpret_cast<unsigned*>(c.rb_done + nrb_);
    c.swaps = reinterpret_cast<int*>(c.counters + nrb_);
    c.moves = reinterpret_cast<unsigned*>(c.swaps + 1 + 2 * K);
    char* pp = reinterpret_cast<char*>(c.moves + 1);
    pp += (256 - (reinterpret_cast<uintptr_t>(pp) & 255)) & 255;
    c.partials = reinterpret_cast<double*>(pp);
    c.pstride = 12LL * 8 * batch_cus();
    c.hist = hist_.as<double>();
    c.hstride = hist_cap_;
    // Replacement policy per solver (mixed form), from measurements at 128^3
    // to 512^3 against the FP64 solvers on the same probes:
    //  * CG: every rep_every iterations and at each tenfold decrease;
    //  * pipelined CG: its matrix-vector recurrences (w = A u, s = A p,
    //    z = A q) drift much faster, so at least every 10 iterations (at 50
    //    it stalls near 1e-6 at 512^3; at 10 it matches FP64's counts);
    //  * BiCGStab: at each tenfold decrease (no periodic replacement), with
    //    r^ and the directions restarted from the replaced residual.  With
    //    FP32 vectors its bi-orthogonality decays within a long cycle: at
    //    512^3, cycles of hundreds of iterations (a drop of sqrt(eps32) ~
    //    2.4e-4 between replacements, the criterion of van der Vorst & Ye,
    //    SIAM J. Sci. Comput. 22, 2000) end in stalls and near-breakdowns
    //    (|s| hundreds of times |r|) and 1.7x FP64's iterations; replacing
    //    without the restart diverges; restarting at each tenfold decrease
    //    keeps the cycles short and converged all probes measured in ~1.14x
    //    FP64's iterations.
    c.rep_every = kind_ == SEDI_PCG ? std::min(rep_every_, 10)
                  : kind_ == SEDI_BICGSTAB ? 1 << 30 : rep_every_;
    c.drop = kind_ == SEDI_BICGSTAB ? SEDI_BICG_DROP : 0.1;
    c.restart = 2.0;
    c.bicg_reset = SEDI_BICG_RESET;
    c.mixed = fp32_ ? 1 : 0;
    c.K = K;
    c.kind = kind_;
    run.shape = shape_;
    run.n = shape_.n();
    run.nrb = nrb_;
    run.a = a_; run.b = b_c_; run.cc = c_; run.im = 1.0 / a_;
    run.stream = stream_;
    for (int k = 0; k < BATCH_MAX_SLOTS; k++) run.v[k] = v_[k];
    run.xhi = xhi_;
    run.mixed = fp32_;
    run.nvec = nvec_;
    return SEDI_OK;
}

// Blocks rb0 .. rb0+count-1: r = b, x = 0, their scalars, tolerance and cap
// (and the pipelined set-up pass).
template <class T>
int BatchSolver::start_t(int rb0, int count, double tol, int maxiter) {
    const BRun& run = *run_;
    const long long n = shape_.n();
    const T im = (T)(1.0 / a_);
    V4<T>* B = static_cast<V4<T>*>(v_[SLOT_B]);
    V4<T>* XL = static_cast<V4<T>*>(v_[SLOT_XLO]);
    // the first residual, read by iteration 1: r1 (CG, BiCGStab) or r (pipelined CG)
    V4<T>* R = static_cast<V4<T>*>(kind_ == SEDI_BICGSTAB ? v_[BI_R1] : (kind_ == SEDI_CG ? v_[CG_R1] : v_[PCG_R]));
    V4<T>* RH = kind_ == SEDI_BICGSTAB ? static_cast<V4<T>*>(v_[BI_RH]) : nullptr;
    SEDI_HIP(launch("batch_init_kernel", batch_init_kernel<T>, dim3(ew_blocks(n), count), dim3(NTHREADS),
                    stream_, n, run.c, rb0, tol, maxiter, (const V4<T>*)B, R, RH, XL, xhi_, im));
    if (kind_ == SEDI_PCG) return pcg_setup<T>(run, rb0, count);
    return SEDI_OK;
}

// `count` iterations of every running block.  The first call on a solver
// launches them directly (which also settles every pass's geometry); later
// calls replay one captured graph of the chunk, whose kernels read the
// iteration state from device memory.
template <class T>
int BatchSolver::iterate_t(int count) {
    const BRun& run = *run_;
    auto direct = [&]() -> int {
        int rc = SEDI_OK;
        for (int k = 0; k < count && rc == SEDI_OK; k++) {
            if (run.compact > 0 && k % run.compact == 0 && (rc = compact_t<T>())) break;
            if (kind_ == SEDI_CG) rc = cg_iteration<T>(run);
            else if (kind_ == SEDI_PCG) rc = pcg_iteration<T>(run);
            else rc = bicg_iteration<T>(run);
        }
        return rc;
    };
    if (!graphs_enabled() || profile_any() || !geometry_ready_) {
        const int rc = direct();
        geometry_ready_ = true;
        return rc;
    }
    if (!graph_ || graph_count_ != count || graph_compact_ != run.compact) {
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
        graph_compact_ = run.compact;
    }
    SEDI_HIP(hipGraphLaunch(graph_, stream_));
    return SEDI_OK;
}

// The planning kernel and the lane exchange (which returns at once when
// nothing moves).
template <class T>
int BatchSolver::compact_t() {
    // the slots whose lanes carry state: all but the FP32 part of x in mixed
    // BiCGStab, whose x is wholly in the FP64 rows (x_lo stays zero)
    VSlots vs;
    int ns = 0;
    for (int k = 0; k < nvec_; k++)
        if (!(k == SLOT_XLO && fp32_ && kind_ == SEDI_BICGSTAB)) vs.p[ns++] = v_[k];
    SEDI_HIP(launch("batch_compact_kernel", batch_compact_kernel, dim3(1), dim3(256), stream_, run_->c, nrb_));
    SEDI_HIP(launch("batch_swap_kernel", batch_swap_kernel<T>, dim3(ew_blocks(shape_.n()), ns + 1), dim3(NTHREADS),
                    stream_, run_->c, shape_.n(), vs, ns, xhi_));
    return SEDI_OK;
}

unsigned BatchSolver::read_moves() {
    unsigned m = 0;
    if (run_ && run_->c.moves && hipMemcpy(&m, run_->c.moves, sizeof(m), hipMemcpyDeviceToHost) != hipSuccess)
        m = 0;
    return m;
}

// Compaction period: every iteration on grids of 2^20 cells and more, where
// an iteration of a block takes far longer than the two small launches;
// once per chunk of 16 on smaller grids.
static int compact_period(const Shape& s, int nrb) {
    if (!batch_compact_enabled() || nrb < 2 || nrb > COMPACT_MAX_RB) return 0;
    return s.n() >= (1LL << 20) ? 1 : 16;
}

template <class T>
int BatchSolver::solve_t(int maxiter, double tol, BatchResult* res) {
    SEDI_REQUIRE(maxiter >= 0, SEDI_ERR_ARG, "maxiter must be >= 0");
    const int K = K_;
    int rc = bind_t<T>(maxiter);
    if (rc) return rc;
    SEDI_HIP(hipStreamSynchronize(stream_));
    const unsigned moves0 = read_moves();
    run_->compact = compact_period(shape_, nrb_);
    for (int k = 0; k < K; k++) lane_of_[k] = k;
    if ((rc = start_t<T>(0, nrb_, tol, maxiter))) return rc;
    const BCtl& c = run_->c;
    const int chunk = 16;
    int launched = 0, slot = 0;
    bool pending = false;
    hipEvent_t ev[2];
    SEDI_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    SEDI_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    while (launched < maxiter && rc == SEDI_OK) {
        const int cnt = std::min(chunk, maxiter - launched);
        if ((rc = iterate_t<T>(cnt))) break;
        launched += cnt;
        if (hipMemcpyAsync(pinned_ + slot * nrb_, c.rb_done, sizeof(int) * nrb_, hipMemcpyDeviceToHost,
                           stream_) != hipSuccess ||
            hipEventRecord(ev[slot], stream_) != hipSuccess) {
            set_error("batched status copy failed");
            rc = SEDI_ERR_HIP;
            break;
        }
        if (pending) {
            const hipError_t we = hipEventSynchronize(ev[slot ^ 1]);
            if (we != hipSuccess) {
                set_error("batched status wait failed: %s", hipGetErrorString(we));
                rc = SEDI_ERR_HIP;
                break;
            }
            bool all = true;
            for (int q = 0; q < nrb_; q++) all = all && pinned_[(slot ^ 1) * nrb_ + q];
            if (all) break;
        }
        pending = true;
        slot ^= 1;
    }
    const hipError_t se = hipStreamSynchronize(stream_);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    profile_flush();
    if (rc) return rc;
    if (se != hipSuccess) {
        set_error("batched solve failed: %s", hipGetErrorString(se));
        return SEDI_ERR_HIP;
    }
    // which lane holds which right-hand side after the moves
    std::vector<int> plane(K);
    SEDI_HIP(hipMemcpy(plane.data(), c.plane, sizeof(int) * K, hipMemcpyDeviceToHost));
    for (int l = 0; l < K; l++) {
        SEDI_REQUIRE(plane[l] >= 0 && plane[l] < K, SEDI_ERR_STATE, "lane %d holds no right-hand side", l);
        lane_of_[plane[l]] = l;
    }
    moves_last_ = (long long)(read_moves() - moves0);
    if (res) {
        std::vector<int> it(K), code(K), done(K);
        std::vector<double> err(K);
        SEDI_HIP(hipMemcpy(it.data(), c.fiter, sizeof(int) * K, hipMemcpyDeviceToHost));
        SEDI_HIP(hipMemcpy(code.data(), c.code, sizeof(int) * K, hipMemcpyDeviceToHost));
        SEDI_HIP(hipMemcpy(err.data(), c.ferr, sizeof(double) * K, hipMemcpyDeviceToHost));
        SEDI_HIP(hipMemcpy(done.data(), c.done, sizeof(int) * K, hipMemcpyDeviceToHost));
        res->iters.assign(K, 0);
        res->codes.assign(K, 0);
        res->errs.assign(K, 0.0);
        for (int k = 0; k < K; k++) {
            const int l = lane_of_[k];
            SEDI_REQUIRE(done[l], SEDI_ERR_STATE, "batched solve stopped without a result for RHS %d", k);
            res->iters[k] = it[l];
            res->codes[k] = code[l];
            res->errs[k] = err[l];
        }
    }
    hist_host_.resize((size_t)K * hist_cap_);
    SEDI_HIP(hipMemcpy(hist_host_.data(), hist_.p, sizeof(double) * hist_host_.size(), hipMemcpyDeviceToHost));
    return SEDI_OK;
}

// ---- block-level control ---------------------------------------------------
__global__ void batch_free_all_kernel(BCtl c, int nrb) {
    for (int i = threadIdx.x; i < nrb; i += blockDim.x) c.rb_done[i] = 1;
    for (int i = threadIdx.x; i < 4 * nrb; i += blockDim.x) c.done[i] = 1;
}

int BatchSolver::slots_open(int hist_cap) {
    int rc = fp32_ ? bind_t<float>(hist_cap) : bind_t<double>(hist_cap);
    if (rc) return rc;
    // blocks start at different times here: each keeps its lanes
    run_->compact = 0;
    for (int k = 0; k < K_; k++) lane_of_[k] = k;
    SEDI_HIP(launch("batch_free_all_kernel", batch_free_all_kernel, dim3(1), dim3(256), stream_, run_->c, nrb_));
    SEDI_HIP(hipStreamSynchronize(stream_));
    return SEDI_OK;
}

void* BatchSolver::block_rhs(int rb) const {
    const size_t bytes = (size_t)shape_.n() * 4 * (fp32_ ? 4 : 8);
    return static_cast<char*>(v_[SLOT_B]) + bytes * rb;
}

int BatchSolver::start_block(int rb, double tol, int maxiter) {
    SEDI_REQUIRE(rb >= 0 && rb < nrb_, SEDI_ERR_ARG, "block %d of %d", rb, nrb_);
    SEDI_REQUIRE(maxiter >= 0 && maxiter <= hist_cap_, SEDI_ERR_ARG, "maxiter %d beyond the history capacity %d",
                 maxiter, hist_cap_);
    return fp32_ ? start_t<float>(rb, 1, tol, maxiter) : start_t<double>(rb, 1, tol, maxiter);
}

int BatchSolver::step(int iters) {
    if (iters <= 0) return SEDI_OK;
    return fp32_ ? iterate_t<float>(iters) : iterate_t<double>(iters);
}

// Status snapshot: [ferr K doubles][the control block's ints: per-RHS
// (CI_NINT x K), then iter and rb_done per block].
size_t BatchSolver::status_bytes() const {
    return sizeof(double) * K_ + sizeof(int) * ((size_t)CI_NINT * K_ + 2 * nrb_);
}

int BatchSolver::post_status(void* snap, hipEvent_t ev) {
    const BCtl& c = run_->c;
    char* h = static_cast<char*>(snap);
    SEDI_HIP(hipMemcpyAsync(h, c.ferr, sizeof(double) * K_, hipMemcpyDeviceToHost, stream_));
    SEDI_HIP(hipMemcpyAsync(h + sizeof(double) * K_, c.fiter, sizeof(int) * ((size_t)CI_NINT * K_ + 2 * nrb_),
                            hipMemcpyDeviceToHost, stream_));
    SEDI_HIP(hipEventRecord(ev, stream_));
    return SEDI_OK;
}

static const int* snap_ints(const void* snap, int K) {
    return reinterpret_cast<const int*>(static_cast<const char*>(snap) + sizeof(double) * K);
}
const int* BatchSolver::snap_block_done(const void* snap) const {
    return snap_ints(snap, K_) + CI_NINT * K_ + nrb_;
}
const int* BatchSolver::snap_iters(const void* snap) const { return snap_ints(snap, K_) + CI_FITER * K_; }
const int* BatchSolver::snap_codes(const void* snap) const { return snap_ints(snap, K_) + CI_CODE * K_; }
const double* BatchSolver::snap_errs(const void* snap) const { return static_cast<const double*>(snap); }

int BatchSolver::solve(int maxiter, double tol, BatchResult* res) {
    return fp32_ ? solve_t<float>(maxiter, tol, res) : solve_t<double>(maxiter, tol, res);
}

template <class T>
int BatchSolver::accumulate_t(double* t, double* q, int k0, int k1) {
    const long long n = shape_.n();
    long long g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    for (int k = k0; k < k1; k += 256) {
        LaneList L;
        L.n = std::min(256, k1 - k);
        for (int j = 0; j < L.n; j++) L.lane[j] = lane_of_[k + j];
        SEDI_HIP(launch("batch_accum_kernel", batch_accum_kernel<T>, dim3((unsigned)g), dim3(256), stream_, n, L,
                        (const V4<T>*)v_[SLOT_XLO], (const double*)xhi_, (const V4<T>*)v_[SLOT_B],
                        pending_dirs<T>(*run_), (const double*)run_->c.apend, K_, t, q));
    }
    return SEDI_OK;
}

int BatchSolver::accumulate(double* t, double* q, int k0, int k1) {
    SEDI_REQUIRE(0 <= k0 && k0 <= k1 && k1 <= K_, SEDI_ERR_ARG, "probe range %d..%d of a batch of %d", k0, k1, K_);
    if (k0 == k1) return SEDI_OK;
    return fp32_ ? accumulate_t<float>(t, q, k0, k1) : accumulate_t<double>(t, q, k0, k1);
}

template <class T>
int BatchSolver::pack_t(const double* b, int kact, long long row_stride) {
    const long long n = shape_.n();
    SEDI_HIP(launch("batch_pack_kernel", batch_pack_kernel<T>, dim3(ew_blocks(n), nrb_), dim3(NTHREADS),
                    stream_, n, kact, row_stride, b, static_cast<V4<T>*>(v_[SLOT_B])));
    return SEDI_OK;
}

template <class T>
int BatchSolver::unpack_t(double* x, int k0, int k1, bool rhs) {
    const long long n = shape_.n();
    const int rb0 = k0 / 4, rb1 = (k1 + 3) / 4;
    const V4<T>* src = static_cast<const V4<T>*>(rhs ? v_[SLOT_B] : v_[SLOT_XLO]) + (long long)rb0 * n;
    const double* xhi = rhs ? nullptr : xhi_;
    PendDirs<T> pend = pending_dirs<T>(*run_);
    for (int j = 0; j < CG_NPEND; j++) pend.p[j] = (rhs || !pend.p[j]) ? nullptr : pend.p[j] + (long long)rb0 * n;
    // the kernel indexes RHS blocks from rb0: shift the row range to match
    SEDI_HIP(launch("batch_unpack_kernel", batch_unpack_kernel<T>, dim3(ew_blocks(n), rb1 - rb0),
                    dim3(NTHREADS), stream_, n, k0 - 4 * rb0, k1 - 4 * rb0, src,
                    xhi ? xhi + (long long)4 * rb0 * n : nullptr, pend, run_->c.apend + 4 * rb0, K_, x));
    return SEDI_OK;
}

int BatchSolver::load_rhs(const double* b, int kact, bool replicate) {
    SEDI_REQUIRE(kact >= 0 && kact <= K_, SEDI_ERR_ARG, "%d right-hand sides for a batch of %d", kact, K_);
    const long long stride = replicate ? 0 : shape_.n();
    return fp32_ ? pack_t<float>(b, kact, stride) : pack_t<double>(b, kact, stride);
}

// Rows k0 .. k1-1 in right-hand-side order: one unpacking when they sit in
// their own lanes, else one per row.
int BatchSolver::store_rows(double* x, int k0, int k1, bool rhs) {
    bool identity = true;
    for (int k = k0; k < k1; k++) identity = identity && lane_of_[k] == k;
    if (identity) return fp32_ ? unpack_t<float>(x, k0, k1, rhs) : unpack_t<double>(x, k0, k1, rhs);
    for (int k = k0; k < k1; k++) {
        const int l = lane_of_[k];
        double* row = x + (long long)(k - k0) * shape_.n();
        const int rc = fp32_ ? unpack_t<float>(row, l, l + 1, rhs) : unpack_t<double>(row, l, l + 1, rhs);
        if (rc) return rc;
    }
    return SEDI_OK;
}

int BatchSolver::store_x(double* x, int k0, int k1) {
    SEDI_REQUIRE(0 <= k0 && k0 < k1 && k1 <= K_, SEDI_ERR_ARG, "solutions %d..%d of a batch of %d", k0, k1, K_);
    return store_rows(x, k0, k1, false);
}

int BatchSolver::store_rhs(double* v, int k0, int k1) {
    SEDI_REQUIRE(0 <= k0 && k0 < k1 && k1 <= K_, SEDI_ERR_ARG, "right-hand sides %d..%d of a batch of %d", k0,
                 k1, K_);
    return store_rows(v, k0, k1, true);
}

std::vector<double> BatchSolver::history(int k) const {
    std::vector<double> h;
    if (k < 0 || k >= K_ || hist_host_.empty()) return h;
    const size_t l = (size_t)lane_of_[k];
    h.assign(hist_host_.begin() + l * hist_cap_, hist_host_.begin() + (l + 1) * hist_cap_);
    return h;
}

template int replace_and_fold<float>(const BRun&, V4<float>*, V4<float>*, bool);
template int replace_and_fold<double>(const BRun&, V4<double>*, V4<double>*, bool);
template int fold_lanes<float>(const BRun&);
template int fold_lanes<double>(const BRun&);

}  // namespace sedi
