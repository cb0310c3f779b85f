// batched_bicg.hip -- batched BiCGStab: per RHS the recurrences, breakdown
// codes and absolute s-test of BiCgStab.hpp:117-264, in the three fused
// passes of the plain solver (direction | half step | update).
//
// Mixed form: FP32 vectors, FP64 arithmetic and dot products, the solution
// accumulated in FP64.  The FP64 replacement r = b - A x (batched.hip) runs
// after the update at each tenfold decrease of the residual and restarts the
// method there (r^ = r, p = r); the reliable-update analysis of Sleijpen &
// van der Vorst (1996) motivates replacing at decreases, the restart keeps
// the FP32 recurrences' loss of bi-orthogonality from building up.  The
// vectors r, p, v alternate between two buffers with the parity of the
// iteration, because each pass reads them at the halo while writing their
// successors.
#include "batched_common.h"

#include <cmath>

namespace sedi {

// ---- direction (BiCgStab.hpp:165-187): p = r + beta p - beta omega v --------
// (p = r at the first iteration), p^ = p.*iM, v = A p^, <v, r^>
template <class T>
struct PassBBicgDir {
    using C = compute_t<T>;
    using V = V4<C>;
    using S = V4<T>;
    static constexpr int ND = 4, D = SEDI_PREFETCH_BATCH;
    C a, b, c;
    const S *r, *pold, *vold, *rhat;
    S *pnew, *vnew;
    C im;
    V beta, cbw;
    bool first;
    struct Raw { S r, p, v; };
    struct Aux { S rhat; };
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(pold, i), ldc(vold, i)}; }
    __device__ Aux load_aux(cidx_t i) const { return {ldc(rhat, i)}; }
    __device__ S pdir(const Raw& q) const {  // :169
        return first ? q.r : cvt<T>(cvt<C>(q.r) + beta * cvt<C>(q.p) + cbw * cvt<C>(q.v));
    }
    __device__ V f(const Raw& q) const { return im * cvt<C>(pdir(q)); }
    __device__ void finish(cidx_t i, const Raw& q, const Aux& o, const V&, const V& v,
                           double (&acc)[ND]) const {
        const S vs = cvt<T>(v);
        stv(pnew, i, pdir(q));
        stv(vnew, i, vs);
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] += (double)vs[k] * (double)o.rhat[k];
    }
};

template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_bicg_dir_kernel(TileGrid g, BCtl c, compute_t<T> a,
                                                                  compute_t<T> b, compute_t<T> cc, long long n,
                                                                  const V4<T>* r0, const V4<T>* r1, V4<T>* p0,
                                                                  V4<T>* p1, V4<T>* v0, V4<T>* v1,
                                                                  const V4<T>* rhat, compute_t<T> im) {
    using C = compute_t<T>;
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<C> lds[2][TY + 2][TX + 2];
    __shared__ double red[4 * NWAVES];
    __shared__ int lastflag;
    const int it = c.iter[rb];
    const int cur = it & 1;
    const long long off = (long long)rb * n;
    V4<C> beta, cbw;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool d = c.done[4 * rb + k];
        beta[k] = d ? (C)0 : (C)c.beta[4 * rb + k];
        cbw[k] = d ? (C)0 : (C)c.cbw[4 * rb + k];
    }
    PassBBicgDir<T> P{a, b, cc, (cur ? r1 : r0) + off, (cur ? p0 : p1) + off, (cur ? v0 : v1) + off,
                      rhat + off, (cur ? p1 : p0) + off, (cur ? v1 : v0) + off, im, beta, cbw, it == 1};
    double acc[4];
    run_tiles<PassBBicgDir<T>, ZC>(P, g, lds, acc, cur == 1);
    if (!grid_sum<4>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        c.alpha[k] = c.done[k] ? 0.0 : c.rho[k] / acc[q];  // BiCgStab.hpp:187
    }
}

// ---- half step (BiCgStab.hpp:190-213): s = r - alpha v, s^ = s.*iM, --------
// t = A s^, <s,s>, <t,s>, <t,t>; s and t are not stored
template <class T>
struct PassBBicgHalf {
    using C = compute_t<T>;
    using V = V4<C>;
    using S = V4<T>;
    static constexpr int ND = 12, D = SEDI_PREFETCH_BATCH;
    C a, b, c;
    const S *r, *v;
    C im;
    V alpha;
    struct Raw { S r, v; };
    struct Aux {};
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(v, i)}; }
    __device__ Aux load_aux(cidx_t) const { return {}; }
    // s rounded to the storage type, as the update pass will see it
    __device__ S sval(const Raw& q) const { return cvt<T>(cvt<C>(q.r) - alpha * cvt<C>(q.v)); }
    __device__ V f(const Raw& q) const { return im * cvt<C>(sval(q)); }
    __device__ void finish(cidx_t, const Raw& q, const Aux&, const V&, const V& t, double (&acc)[ND]) const {
        const S sv = sval(q);
        const S ts = cvt<T>(t);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] += (double)sv[k] * (double)sv[k];
            acc[4 + k] += (double)ts[k] * (double)sv[k];
            acc[8 + k] += (double)ts[k] * (double)ts[k];
        }
    }
};

template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_bicg_half_kernel(TileGrid g, BCtl c, compute_t<T> a,
                                                                   compute_t<T> b, compute_t<T> cc, long long n,
                                                                   const V4<T>* r0, const V4<T>* r1,
                                                                   const V4<T>* v0, const V4<T>* v1,
                                                                   compute_t<T> im) {
    using C = compute_t<T>;
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<C> lds[2][TY + 2][TX + 2];
    __shared__ double red[12 * NWAVES];
    __shared__ int lastflag;
    const int it = c.iter[rb];
    const int cur = it & 1;
    const long long off = (long long)rb * n;
    V4<C> alpha;
#pragma unroll
    for (int k = 0; k < 4; k++) alpha[k] = (C)c.alpha[4 * rb + k];
    PassBBicgHalf<T> P{a, b, cc, (cur ? r1 : r0) + off, (cur ? v1 : v0) + off, im, alpha};
    double acc[12];
    run_tiles<PassBBicgHalf<T>, ZC>(P, g, lds, acc, cur == 0);
    if (!grid_sum<12>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        if (c.done[k]) continue;
        c.ss[k] = acc[q];                                        // :193-194
        if (acc[q] < c.tolv[k] * c.tolv[k]) c.sexit[k] = 1;      // :195 (absolute test)
        else c.omega[k] = acc[4 + q] / acc[8 + q];               // :213
    }
}

// ---- update (BiCgStab.hpp:217-224 and :160 of the next iteration): ---------
// s, s^, t recomputed; x = x + alpha p^ + omega s^; r = s - omega t; <r,r>,
// the next rho = <r, r^>.  A RHS that left at the s-test takes x + alpha p^
// only (:195-199).
//
// XD: x is accumulated in FP64 (four [N] arrays, one per RHS of the block)
// instead of in the storage type.  The mixed form uses it: BiCGStab runs
// hundreds of iterations between replacements, and an FP32 accumulator over
// that many updates loses about eps32 * sqrt(m) * |x_lo| per cell, which A
// amplifies to ~ eps32 * sqrt(m) * |r| / lambda_min in the replaced residual
// -- at 512^3 more than the drop that triggered the replacement, so the
// solver cannot converge (CG and pipelined CG replace every <= 50 and <= 10
// iterations, where the FP32 accumulator suffices).
template <class T, bool XD>
struct PassBBicgUpd {
    using C = compute_t<T>;
    using V = V4<C>;
    using S = V4<T>;
    using XV = typename std::conditional<XD, V4<double>, S>::type;
    static constexpr int ND = 8, D = SEDI_PREFETCH_BATCH;
    C a, b, c;
    const S *r, *v, *p, *rhat;
    S *x, *rnew;
    double* xh[4];  // XD: x of the block's four RHS
    C im;
    V alpha, omega;
    int sexit[4];
    struct Raw { S r, v; };
    struct Aux { S p; XV x; S rhat; };
    __device__ Raw load(cidx_t i) const { return {ldc(r, i), ldc(v, i)}; }
    __device__ Aux load_aux(cidx_t i) const {
        Aux o;
        o.p = ldc(p, i);
        if constexpr (XD) {
#pragma unroll
            for (int k = 0; k < 4; k++) o.x[k] = ldc((const double*)xh[k], i);
        } else {
            o.x = ldc((const S*)x, i);
        }
        o.rhat = ldc(rhat, i);
        return o;
    }
    __device__ S sval(const Raw& q) const { return cvt<T>(cvt<C>(q.r) - alpha * cvt<C>(q.v)); }
    __device__ V f(const Raw& q) const { return im * cvt<C>(sval(q)); }
    __device__ void finish(cidx_t i, const Raw& q, const Aux& o, const V& sh, const V& t,
                           double (&acc)[ND]) const {
        const V sv = cvt<C>(sval(q));
        const V ph = im * cvt<C>(o.p);
        const V xa = cvt<C>(o.x) + alpha * ph;
        const V xv = xa + omega * sh;  // combine3 (:217)
        const S rs = cvt<T>(sv - omega * t);
        if constexpr (XD) {
            const V xn = sel(sexit, xa, xv);
#pragma unroll
            for (int k = 0; k < 4; k++) stc(xh[k], i, (double)xn[k]);
        } else {
            stv(x, i, cvt<T>(sel(sexit, xa, xv)));
        }
        stv(rnew, i, rs);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[k] += (double)rs[k] * (double)rs[k];
            acc[4 + k] += (double)rs[k] * (double)o.rhat[k];
        }
    }
};

template <class T, bool XD, int ZC>
__device__ __forceinline__ void bicg_upd_body(TileGrid g, BCtl c, compute_t<T> a, compute_t<T> b, compute_t<T> cc, long long n,
    V4<T>* r0, V4<T>* r1, const V4<T>* v0, const V4<T>* v1, const V4<T>* p0, const V4<T>* p1,
    const V4<T>* rhat, V4<T>* x, double* xhi, compute_t<T> im) {
    using C = compute_t<T>;
    const int rb = blockIdx.y;
    if (c.rb_done[rb]) return;
    __shared__ V4<C> lds[2][TY + 2][TX + 2];
    __shared__ double red[8 * NWAVES];
    __shared__ int lastflag;
    const int it = c.iter[rb];
    const int cur = it & 1;
    const long long off = (long long)rb * n;
    PassBBicgUpd<T, XD> P;
    P.a = a; P.b = b; P.c = cc;
    P.r = (cur ? r1 : r0) + off;
    P.v = (cur ? v1 : v0) + off;
    P.p = (cur ? p1 : p0) + off;
    P.rhat = rhat + off;
    P.x = x + off;
#pragma unroll
    for (int k = 0; k < 4; k++) P.xh[k] = XD ? xhi + (long long)(4 * rb + k) * n : nullptr;
    P.rnew = (cur ? r0 : r1) + off;
    P.im = im;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int kk = 4 * rb + k;
        const bool d = c.done[kk];
        P.alpha[k] = d ? (C)0 : (C)c.alpha[kk];
        P.sexit[k] = !d && c.sexit[kk];
        P.omega[k] = (d || P.sexit[k]) ? (C)0 : (C)c.omega[kk];
    }
    double acc[8];
    run_tiles<PassBBicgUpd<T, XD>, ZC>(P, g, lds, acc, cur == 1);
    if (!grid_sum<8>(acc, c.partials + rb * c.pstride, c.counters + rb, red, &lastflag) ||
        threadIdx.x != 0)
        return;
    for (int q = 0; q < 4; q++) {
        const int k = 4 * rb + q;
        if (c.done[k]) continue;
        if (c.sexit[k]) { bbicg_outcome(c, k, it); continue; }  // :195-199
        const double err = sqrt(acc[q] / c.bb[k]);  // :223-227
        c.err[k] = err;
        c.hist[(long long)k * c.hstride + it - 1] = err;
        if (c.omega[k] == 0.0) { bbicg_outcome(c, k, it); continue; }  // :229-230
        if (!c.mixed && err <= c.tolv[k]) { bbicg_outcome(c, k, it); continue; }
        if (!c.mixed && it >= c.maxv[k]) { bbicg_outcome(c, k, c.maxv[k] + 1); continue; }
        c.rho_prev[k] = c.rho[k];  // :232
        c.rho[k] = acc[4 + q];     // next iteration's :160
        const double beta = (c.rho[k] / c.rho_prev[k]) * (c.alpha[k] / c.omega[k]);  // :167
        c.beta[k] = beta;
        c.cbw[k] = -beta * c.omega[k];  // the combine3 coefficient of :169
        if (c.mixed) {
            // the FP64 pass decides convergence, maxiter and rho breakdown;
            // the drop is measured from the largest residual since the last
            // replacement (BiCGStab's residual spikes are what the recursion
            // loses accuracy to: van der Vorst & Ye 2000)
            c.elast[k] = fmax(c.elast[k], err);
            c.rep[k] = (err <= c.tolv[k]) || (it % c.rep_every == 0) || (it >= c.maxv[k]) ||
                       (err < c.drop * c.elast[k]) || (c.rho[k] == 0.0);
        } else if (c.rho[k] == 0.0) {
            bbicg_outcome(c, k, it + 1);  // :163
        }
    }
    c.iter[rb] = it + 1;
    brb_update(c, rb);
}

#define SEDI_BICG_UPD_ARGS g, c, a, b, cc, n, r0, r1, v0, v1, p0, p1, rhat, x, xhi, im
template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_bicg_upd_kernel(TileGrid g, BCtl c, compute_t<T> a, compute_t<T> b, compute_t<T> cc, long long n,
    V4<T>* r0, V4<T>* r1, const V4<T>* v0, const V4<T>* v1, const V4<T>* p0, const V4<T>* p1,
    const V4<T>* rhat, V4<T>* x, double* xhi, compute_t<T> im) {
    bicg_upd_body<T, false, ZC>(SEDI_BICG_UPD_ARGS);
}
template <class T, int ZC>
__global__ __launch_bounds__(NTHREADS) void batch_bicg_upd_xd_kernel(TileGrid g, BCtl c, compute_t<T> a, compute_t<T> b, compute_t<T> cc, long long n,
    V4<T>* r0, V4<T>* r1, const V4<T>* v0, const V4<T>* v1, const V4<T>* p0, const V4<T>* p1,
    const V4<T>* rhat, V4<T>* x, double* xhi, compute_t<T> im) {
    bicg_upd_body<T, true, ZC>(SEDI_BICG_UPD_ARGS);
}
#undef SEDI_BICG_UPD_ARGS

// ---- host ------------------------------------------------------------------
template <class T>
int bicg_iteration(const BRun& r) {
    using C = compute_t<T>;
    const C a = (C)r.a, b = (C)r.b, cc = (C)r.cc, im = (C)r.im;
    auto V = [&](int s) { return static_cast<V4<T>*>(r.v[s]); };
    const KGeom g1 = bgeom("batch_bicg_dir_kernel", batch_bicg_dir_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
    const KGeom g2 = bgeom("batch_bicg_half_kernel", batch_bicg_half_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
    constexpr bool mixed_t = std::is_same<T, float>::value;  // FP32 storage: x in FP64
    SEDI_HIP(SEDI_BZC_LAUNCH("batch_bicg_dir_kernel", batch_bicg_dir_kernel, T, g1, dim3(g1.nblocks, r.nrb),
                             r.stream, g1.tg, r.c, a, b, cc, r.n, (const V4<T>*)V(BI_R0), (const V4<T>*)V(BI_R1),
                             V(BI_P0), V(BI_P1), V(BI_V0), V(BI_V1), (const V4<T>*)V(BI_RH), im));
    SEDI_HIP(SEDI_BZC_LAUNCH("batch_bicg_half_kernel", batch_bicg_half_kernel, T, g2, dim3(g2.nblocks, r.nrb),
                             r.stream, g2.tg, r.c, a, b, cc, r.n, (const V4<T>*)V(BI_R0), (const V4<T>*)V(BI_R1),
                             (const V4<T>*)V(BI_V0), (const V4<T>*)V(BI_V1), im));
    if constexpr (mixed_t) {
        const KGeom g3 = bgeom("batch_bicg_upd_kernel", batch_bicg_upd_xd_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
        SEDI_HIP(SEDI_BZC_LAUNCH("batch_bicg_upd_kernel", batch_bicg_upd_xd_kernel, T, g3,
                                 dim3(g3.nblocks, r.nrb), r.stream, g3.tg, r.c, a, b, cc, r.n, V(BI_R0), V(BI_R1),
                                 (const V4<T>*)V(BI_V0), (const V4<T>*)V(BI_V1), (const V4<T>*)V(BI_P0),
                                 (const V4<T>*)V(BI_P1), (const V4<T>*)V(BI_RH), V(SLOT_XLO), r.xhi, im));
    } else {
        const KGeom g3 = bgeom("batch_bicg_upd_kernel", batch_bicg_upd_kernel<T, 16>, r.shape, SEDI_PREFETCH_BATCH);
        SEDI_HIP(SEDI_BZC_LAUNCH("batch_bicg_upd_kernel", batch_bicg_upd_kernel, T, g3,
                                 dim3(g3.nblocks, r.nrb), r.stream, g3.tg, r.c, a, b, cc, r.n, V(BI_R0), V(BI_R1),
                                 (const V4<T>*)V(BI_V0), (const V4<T>*)V(BI_V1), (const V4<T>*)V(BI_P0),
                                 (const V4<T>*)V(BI_P1), (const V4<T>*)V(BI_RH), V(SLOT_XLO), (double*)nullptr,
                                 im));
    }
    // x is in FP64 (x_lo stays zero): the replacement needs no fold
    if (r.mixed) return replace_and_fold<T>(r, V(BI_R0), V(BI_RH), false);
    return SEDI_OK;
}

template int bicg_iteration<float>(const BRun&);
template int bicg_iteration<double>(const BRun&);

}  // namespace sedi
