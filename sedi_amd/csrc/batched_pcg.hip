// batched_pcg.hip -- batched pipelined CG (Ghysels & Vanroose), per RHS the
// recurrences of PipelinedCg.hpp:138-243 in one fused pass per iteration,
// with the reference's one-iteration lag of the stop test.
//
// Mixed form: the replacement step of pipelined CG must refresh every
// vector that a recurrence carries, not only r: with u = M^-1 r, w = A u,
// s = A p, q = M^-1 s, z = A q, it sets r = b - A x, u, w, s, q, z from x and
// p in FP64 (the residual-replacement variant of Cools, Yetkin, Agullo,
// Giraud & Vanroose, "Analyzing the effect of local rounding error
// propagation on the maximal attainable accuracy of the pipelined conjugate
// gradient method", SIAM J. Matrix Anal. Appl. 39(1), 2018).  That takes two
// sweeps, each applying the stencil to two fields at once: (x, p) -> r, s
// (and u, q pointwise), then (u, q) -> w, z; the second one also forms the
// next iteration's <r,u>, <w,u>, <r,r> on the replaced vectors, from which
// alpha and beta are recomputed (or restarted, as in the CG form).
#include "batched_common.h"

#include <cmath>

namespace sedi {

// ---- set-up: u = r.*iM, w = A u, z = q = s = p = 0 and the first dots ------
// (PipelinedCg.hpp:138-149, 155, 170-181 at iteration 1)
template <class T>
struct PassBPcgInit {
    using C = compute_t<T>;
    using V = V4<C>;
    using S = V4<T>;
    static constexpr int ND = 12, D = SEDI_PREFETCH_BATCH;
    C a, b, c;
    const S* r;
    S *u, *w, *z, *q, *s, *p;
    C im;
    struct Raw { S r; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(r, i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    __device__ V f(const Raw& o) const { return im * cvt<C>(o.r); }
    __device__ void finish(cidx_t i, const Raw& o, const Aux&, const V& uv, const V& wv,
                           double (&acc)[ND]) const {
        const S zero{};
        const S us = cvt<T>(uv), ws = cvt<T>(wv);
        stv(u, i, us);
        stv(w, i, ws);
        stv(z, i, zero); stv(q, i, zero); stv(s, i, zero); stv(p, i, zero);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] += (double)o.r[k] * (double)us[k];
            acc[4 + k] += (double)ws[k] * (double)us[k];
            acc[8 + k] += (double)o.r[k] * (double)o.r[k];
        }
    }
};

template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_pcg_init_kernel(TileGrid g, BCtl c, int rb0, compute_t<T> a,
                                                                  compute_t<T> b, compute_t<T> cc, long long n,
                                                                  const V4<T>* r, V4<T>* u, V4<T>* w, V4<T>* z,
                                                                  V4<T>* q, V4<T>* s, V4<T>* p,
                                                                  compute_t<T> im) {
    const int rb = rb0 + blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<compute_t<T>> lds[2][TY + 2][TX + 2];
    __shared__ double red[12 * NWAVES];
    __shared__ int lastflag;
    const long long off = (long long)rb * n;
    PassBPcgInit<T> P{a, b, cc, r + off, u + off, w + off, z + off, q + off, s + off, p + off, im};
    double acc[12];
    run_tiles<PassBPcgInit<T>, ZC>(P, g, lds, acc, true);
    if (!grid_sum<12>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q4 = 0; q4 < 4; q4++) {
        const int k = 4 * rb + q4;
        if (c.done[k]) {
            c.alpha[k] = 0.0;
            c.beta[k] = 0.0;
            continue;
        }
        c.gamma[k] = acc[q4];  // PipelinedCg.hpp:171-173 at iter 1
        c.delta[k] = acc[4 + q4];
        c.beta[k] = 0.0;  // :179-180
        c.alpha[k] = acc[q4] / acc[4 + q4];
        c.errp[k] = sqrt(acc[8 + q4] / c.bb[k]);
    }
}

// ---- iteration (PipelinedCg.hpp:157-223): m = w.*iM, n = A m, the eight ----
// recurrences in the reference's order, the next <r,u>, <w,u>, <r,r>.
// x is write-only (the last recurrence), so its updates come in groups of
// CG_XG iterations, as CG's: p is written to p[it % CG_XG], iterations that
// are not a multiple of CG_XG leave x alone (their alphas wait in c.apend),
// the others add the pending apend_j p_j and their own alpha p in one
// read-modify-write.  XS (that last iteration) is a compile-time choice.
template <class T, bool XS>
struct PassBPcgIter {
    using C = compute_t<T>;
    using V = V4<C>;
    using S = V4<T>;
    static constexpr int ND = 12, D = SEDI_PREFETCH_BATCH;
    C a, b, c;
    const S* wold;
    S* wnew;
    S *x, *r, *u, *z, *q, *s;
    const S* pold;
    S* pnew;
    const S* pend[CG_NPEND];
    C im;
    V alpha, beta, apend[CG_NPEND];
    struct Raw { S w; };
    struct Aux { S q, s, u, p, z, r, x, pp[CG_NPEND]; };
    __device__ Raw load(cidx_t i) const { return {ldc(wold, i)}; }
    __device__ Aux load_aux(cidx_t i) const {
        Aux o;
        o.q = ldc(q, i); o.s = ldc(s, i); o.u = ldc(u, i); o.p = ldc(pold, i); o.z = ldc(z, i); o.r = ldc(r, i);
        if (XS) {
            o.x = ldc(x, i);
#pragma unroll
            for (int j = 0; j < CG_NPEND; j++) o.pp[j] = ldc(pend[j], i);
        }
        return o;
    }
    __device__ V f(const Raw& o) const { return im * cvt<C>(o.w); }
    __device__ void finish(cidx_t i, const Raw& c, const Aux& o, const V& m, const V& nv,
                           double (&acc)[ND]) const {
        const V w = cvt<C>(c.w), uo = cvt<C>(o.u);
        const V qv = m + beta * cvt<C>(o.q);
        const V sv = w + beta * cvt<C>(o.s);
        const V pv = uo + beta * cvt<C>(o.p);
        const V zv = nv + beta * cvt<C>(o.z);
        const S rs = cvt<T>(cvt<C>(o.r) - alpha * sv);
        const S us = cvt<T>(uo - alpha * qv);
        const S ws = cvt<T>(w - alpha * zv);
        stv(q, i, cvt<T>(qv)); stv(s, i, cvt<T>(sv)); stv(pnew, i, cvt<T>(pv)); stv(z, i, cvt<T>(zv));
        stv(r, i, rs); stv(u, i, us); stv(wnew, i, ws);
        if (XS) {
            V xv = cvt<C>(o.x);
#pragma unroll
            for (int j = 0; j < CG_NPEND; j++) xv = xv + apend[j] * cvt<C>(o.pp[j]);
            stv(x, i, cvt<T>(xv + alpha * pv));
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] += (double)rs[k] * (double)us[k];
            acc[4 + k] += (double)ws[k] * (double)us[k];
            acc[8 + k] += (double)rs[k] * (double)rs[k];
        }
    }
};

template <class T, int ZC, bool XS>
__device__ __forceinline__ void run_pcg_iter(const TileGrid& g, const BCtl& c, compute_t<T> a, compute_t<T> b,
                                             compute_t<T> cc, long long off, int it, V4<T>* w0, V4<T>* w1,
                                             V4<T>* x, V4<T>* r, V4<T>* u, V4<T>* z, V4<T>* q, V4<T>* s,
                                             const CgDirs<T>& P4, compute_t<T> im,
                                             const V4<compute_t<T>>& alpha, const V4<compute_t<T>>& beta,
                                             V4<compute_t<T>> (&lds)[2][TY + 2][TX + 2], double (&acc)[12]) {
    using C = compute_t<T>;
    const int rb = blockIdx.y;
    // iteration 1 reads w from w0 (written by the set-up pass) and p from p0
    PassBPcgIter<T, XS> P;
    P.a = a; P.b = b; P.c = cc;
    P.wold = ((it & 1) ? w0 : w1) + off;
    P.wnew = ((it & 1) ? w1 : w0) + off;
    P.x = x + off; P.r = r + off; P.u = u + off; P.z = z + off; P.q = q + off; P.s = s + off;
    P.pold = P4.at((it + CG_XG - 1) % CG_XG) + off;
    P.pnew = P4.at(it % CG_XG) + off;
    P.im = im;
    P.alpha = alpha;
    P.beta = beta;
#pragma unroll
    for (int j = 0; j < CG_NPEND; j++) {
        P.pend[j] = P4.p[j + 1] + off;
#pragma unroll
        for (int k = 0; k < 4; k++) P.apend[j][k] = XS ? uniform((C)c.apend[j * c.K + 4 * rb + k]) : (C)0;
    }
    run_tiles<PassBPcgIter<T, XS>, ZC>(P, g, lds, acc, (it & 1) == 0);
}

template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_pcg_iter_kernel(TileGrid g, BCtl c, compute_t<T> a,
                                                                  compute_t<T> b, compute_t<T> cc, long long n,
                                                                  V4<T>* w0, V4<T>* w1, V4<T>* x, V4<T>* r,
                                                                  V4<T>* u, V4<T>* z, V4<T>* q, V4<T>* s,
                                                                  CgDirs<T> P4, compute_t<T> im) {
    using C = compute_t<T>;
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<C> lds[2][TY + 2][TX + 2];
    __shared__ double red[12 * NWAVES];
    __shared__ int lastflag;
    const int it = c.iter[rb];
    const int jp = it % CG_XG;
    const long long off = (long long)rb * n;
    V4<C> alpha, beta;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool d = c.done[4 * rb + k];
        alpha[k] = d ? (C)0 : uniform((C)c.alpha[4 * rb + k]);
        beta[k] = d ? (C)0 : uniform((C)c.beta[4 * rb + k]);
    }
    double acc[12];
    if (jp == 0)
        run_pcg_iter<T, ZC, true>(g, c, a, b, cc, off, it, w0, w1, x, r, u, z, q, s, P4, im, alpha, beta, lds, acc);
    else
        run_pcg_iter<T, ZC, false>(g, c, a, b, cc, off, it, w0, w1, x, r, u, z, q, s, P4, im, alpha, beta, lds, acc);
    if (!grid_sum<12>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    // the alpha each right-hand side used in this pass (0 for a finished one)
    // waits for the group's last iteration, which applies them all
    for (int q4 = 0; q4 < 4; q4++) {
        const int k = 4 * rb + q4;
        const double used = c.done[k] ? 0.0 : c.alpha[k];
        if (jp == 0)
            for (int j = 0; j < CG_NPEND; j++) c.apend[j * c.K + k] = 0.0;
        else
            c.apend[(jp - 1) * c.K + k] = used;
    }
    for (int q4 = 0; q4 < 4; q4++) {
        const int k = 4 * rb + q4;
        if (c.done[k]) continue;
        const double err = c.errp[k];  // rr before this iteration's update (:227)
        c.err[k] = err;
        c.hist[(long long)k * c.hstride + it - 1] = err;
        if (err <= c.tolv[k] || it >= c.maxv[k]) {
            if (err <= c.tolv[k]) bset_done(c, k, 0, it, err);
            else bcg_exhausted(c, k, err);
            // mixed: updates still pending go into x_hi now (batch_fold_kernel)
            if (c.mixed && jp != 0) c.fold[k] = 1;
            continue;
        }
        // scalars of iteration it+1 (PipelinedCg.hpp:167-178)
        const double gamma_old = c.gamma[k];
        const double gamma = acc[q4], delta = acc[4 + q4];
        const double beta_n = gamma / gamma_old;
        const double alpha_n = 1.0 / (delta / gamma - beta_n / c.alpha[k]);
        c.gamma_prev[k] = gamma_old;
        c.alpha_prev[k] = c.alpha[k];
        c.gamma[k] = gamma;
        c.delta[k] = delta;
        c.beta[k] = beta_n;
        c.alpha[k] = alpha_n;
        const double e = sqrt(acc[8 + q4] / c.bb[k]);
        c.errp[k] = e;
        if (c.mixed) c.elast[k] = fmax(c.elast[k], e);
        if (c.mixed)
            c.rep[k] = (e <= c.tolv[k]) || (it % c.rep_every == 0) || (it + 1 >= c.maxv[k]) ||
                       (e < c.drop * c.elast[k]);
    }
    c.iter[rb] = it + 1;
    brb_update(c, rb);
}

// ---- replacement A: r = b - A x, s = A p, u = r.*iM, q = s.*iM (FP64) --------
template <class T>
struct PassBPcgRepA {
    using V = D2;
    static constexpr int ND = 1, D = SEDI_PREFETCH_BATCH;
    double a, b, c;
    const double* xh[4];
    const V4<T>* xlo;
    const V4<T>* p;
    const V4<T>* rhs;
    V4<T> *r, *s, *u, *q;
    double im;
    int rep[4];
    struct Raw { V4<double> xh; V4<T> xl, p; };
    struct Aux { V4<T> b, r, s, u, q; };
    __device__ Raw load(cidx_t i) const {
        Raw o;
#pragma unroll
        for (int k = 0; k < 4; k++) o.xh[k] = ldc(xh[k], i);
        o.xl = ldc(xlo, i);
        o.p = ldc(p, i);
        return o;
    }
    __device__ Aux load_aux(cidx_t i) const {
        return {ldc(rhs, i), ldc((const V4<T>*)r, i), ldc((const V4<T>*)s, i), ldc((const V4<T>*)u, i),
                ldc((const V4<T>*)q, i)};
    }
    __device__ V f(const Raw& o) const {
        V v;
#pragma unroll
        for (int k = 0; k < 4; k++) v.a[k] = o.xh[k] + (double)o.xl[k];
        v.b = to_f64(o.p);
        return v;
    }
    __device__ void finish(cidx_t i, const Raw&, const Aux& o, const V&, const V& Av, double (&)[ND]) const {
        V4<T> rn, sn, un, qn;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double rd = (double)o.b[k] - Av.a[k];  // PipelinedCg.hpp:125, in FP64
            const double sd = Av.b[k];
            rn[k] = rep[k] ? (T)rd : o.r[k];
            sn[k] = rep[k] ? (T)sd : o.s[k];
            un[k] = rep[k] ? (T)(rd * im) : o.u[k];
            qn[k] = rep[k] ? (T)(sd * im) : o.q[k];
        }
        stv(r, i, rn); stv(s, i, sn); stv(u, i, un); stv(q, i, qn);
    }
};

template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_pcg_repa_kernel(TileGrid g, BCtl c, double a, double b,
                                                                  double cc, long long n, const double* xhi,
                                                                  const V4<T>* xlo, CgDirs<T> P4,
                                                                  const V4<T>* rhs, V4<T>* r, V4<T>* s,
                                                                  V4<T>* u, V4<T>* q, double im) {
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    const int* rp = c.rep + 4 * rb;
    if (!(rp[0] | rp[1] | rp[2] | rp[3])) return;
    __shared__ D2 lds[2][TY + 2][TX + 2];
    const long long off = (long long)rb * n;
    PassBPcgRepA<T> P;
    P.a = a; P.b = b; P.c = cc;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        P.xh[k] = xhi + (long long)(4 * rb + k) * n;
        P.rep[k] = rp[k];
    }
    // the direction of the iteration that asked for the replacement
    P.xlo = xlo + off; P.p = P4.at((c.iter[rb] - 1) % CG_XG) + off; P.rhs = rhs + off;
    P.r = r + off; P.s = s + off; P.u = u + off; P.q = q + off;
    P.im = im;
    double acc[1];
    run_tiles<PassBPcgRepA<T>, ZC>(P, g, lds, acc, false);
}

// ---- replacement B: w = A u, z = A q, and <r,u>, <w,u>, <r,r> ---------------
template <class T>
struct PassBPcgRepB {
    using V = D2;
    static constexpr int ND = 12, D = SEDI_PREFETCH_BATCH;
    double a, b, c;
    const V4<T>* u;
    const V4<T>* q;
    const V4<T>* r;
    V4<T> *w, *z;
    int rep[4];
    struct Raw { V4<T> u, q; };
    struct Aux { V4<T> r, w, z; };
    __device__ Raw load(cidx_t i) const { return {ldc(u, i), ldc(q, i)}; }
    __device__ Aux load_aux(cidx_t i) const { return {ldc(r, i), ldc((const V4<T>*)w, i), ldc((const V4<T>*)z, i)}; }
    __device__ V f(const Raw& o) const { return {to_f64(o.u), to_f64(o.q)}; }
    __device__ void finish(cidx_t i, const Raw& raw, const Aux& o, const V&, const V& Av,
                           double (&acc)[ND]) const {
        V4<T> wn, zn;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            wn[k] = rep[k] ? (T)Av.a[k] : o.w[k];
            zn[k] = rep[k] ? (T)Av.b[k] : o.z[k];
        }
        stv(w, i, wn);
        stv(z, i, zn);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] += (double)o.r[k] * (double)raw.u[k];
            acc[4 + k] += (double)wn[k] * (double)raw.u[k];
            acc[8 + k] += (double)o.r[k] * (double)o.r[k];
        }
    }
};

template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_pcg_repb_kernel(TileGrid g, BCtl c, double a, double b,
                                                                  double cc, long long n, const V4<T>* u,
                                                                  const V4<T>* q, const V4<T>* r, V4<T>* w0,
                                                                  V4<T>* w1, V4<T>* z) {
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    const int* rp = c.rep + 4 * rb;
    if (!(rp[0] | rp[1] | rp[2] | rp[3])) return;
    __shared__ D2 lds[2][TY + 2][TX + 2];
    __shared__ double red[12 * NWAVES];
    __shared__ int lastflag;
    const long long off = (long long)rb * n;
    // w as the next iteration (iter[rb]) reads it
    V4<T>* w = (c.iter[rb] & 1) ? w0 : w1;
    PassBPcgRepB<T> P;
    P.a = a; P.b = b; P.c = cc;
    P.u = u + off; P.q = q + off; P.r = r + off; P.w = w + off; P.z = z + off;
#pragma unroll
    for (int k = 0; k < 4; k++) P.rep[k] = rp[k];
    double acc[12];
    run_tiles<PassBPcgRepB<T>, ZC>(P, g, lds, acc, true);
    if (!grid_sum<12>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q4 = 0; q4 < 4; q4++) {
        const int k = 4 * rb + q4;
        if (!c.rep[k]) continue;
        c.rep[k] = 0;
        const double gamma = acc[q4], delta = acc[4 + q4];
        const double e = sqrt(acc[8 + q4] / c.bb[k]);
        if (e > c.restart * c.errp[k]) {  // the recursion drifted: restart
            c.beta[k] = 0.0;
            c.alpha[k] = gamma / delta;
        } else {
            const double beta_n = gamma / c.gamma_prev[k];
            c.beta[k] = beta_n;
            c.alpha[k] = 1.0 / (delta / gamma - beta_n / c.alpha_prev[k]);
        }
        c.gamma[k] = gamma;
        c.delta[k] = delta;
        c.errp[k] = e;
        c.elast[k] = e;
    }
}

// ---- host ------------------------------------------------------------------
template <class T>
int pcg_setup(const BRun& r, int rb0, int count) {
    using C = compute_t<T>;
    const C a = (C)r.a, b = (C)r.b, cc = (C)r.cc, im = (C)r.im;
    auto V = [&](int s) { return static_cast<V4<T>*>(r.v[s]); };
    const KGeom g = bgeom("batch_pcg_init_kernel", batch_pcg_init_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
    SEDI_HIP(SEDI_BZC_LAUNCH("batch_pcg_init_kernel", batch_pcg_init_kernel, T, g, dim3(g.nblocks, count),
                             r.stream, g.tg, r.c, rb0, a, b, cc, r.n, (const V4<T>*)V(PCG_R), V(PCG_U),
                             V(PCG_W0), V(PCG_Z), V(PCG_Q), V(PCG_S), V(PCG_P), im));
    return SEDI_OK;
}

template <class T>
int pcg_iteration(const BRun& r) {
    using C = compute_t<T>;
    const C a = (C)r.a, b = (C)r.b, cc = (C)r.cc, im = (C)r.im;
    auto V = [&](int s) { return static_cast<V4<T>*>(r.v[s]); };
    const KGeom g = bgeom("batch_pcg_iter_kernel", batch_pcg_iter_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
    CgDirs<T> P4;
    for (int j = 0; j < 4; j++) P4.p[j] = j < CG_XG ? V(pcg_pslot(j)) : nullptr;
    SEDI_HIP(SEDI_BZC_LAUNCH("batch_pcg_iter_kernel", batch_pcg_iter_kernel, T, g, dim3(g.nblocks, r.nrb),
                             r.stream, g.tg, r.c, a, b, cc, r.n, V(PCG_W0), V(PCG_W1), V(SLOT_XLO), V(PCG_R),
                             V(PCG_U), V(PCG_Z), V(PCG_Q), V(PCG_S), P4, im));
    if (!r.mixed) return SEDI_OK;
    // x_lo and the pending updates into x_hi for the right-hand sides about to
    // be replaced and those that just finished, then the two replacement sweeps
    if (int rc = fold_lanes<T>(r)) return rc;
    const KGeom ga = bgeom("batch_pcg_repa_kernel", batch_pcg_repa_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
    SEDI_HIP(SEDI_BZC_LAUNCH("batch_pcg_repa_kernel", batch_pcg_repa_kernel, T, ga, dim3(ga.nblocks, r.nrb),
                             r.stream, ga.tg, r.c, r.a, r.b, r.cc, r.n, (const double*)r.xhi,
                             (const V4<T>*)V(SLOT_XLO), P4, (const V4<T>*)V(SLOT_B),
                             V(PCG_R), V(PCG_S), V(PCG_U), V(PCG_Q), r.im));
    const KGeom gb = bgeom("batch_pcg_repb_kernel", batch_pcg_repb_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
    SEDI_HIP(SEDI_BZC_LAUNCH("batch_pcg_repb_kernel", batch_pcg_repb_kernel, T, gb, dim3(gb.nblocks, r.nrb),
                             r.stream, gb.tg, r.c, r.a, r.b, r.cc, r.n, (const V4<T>*)V(PCG_U),
                             (const V4<T>*)V(PCG_Q), (const V4<T>*)V(PCG_R), V(PCG_W0), V(PCG_W1), V(PCG_Z)));
    return SEDI_OK;
}

template int pcg_setup<float>(const BRun&, int, int);
template int pcg_setup<double>(const BRun&, int, int);
template int pcg_iteration<float>(const BRun&);
template int pcg_iteration<double>(const BRun&);

}  // namespace sedi
