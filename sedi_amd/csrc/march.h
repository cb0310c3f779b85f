// march.h -- the z-marching 7-point pass that every solver kernel is built on.
//
// A pass computes, for every cell, an input field f (a pointwise function of
// the solver vectors, e.g. p = r*iM + beta*p), applies the stencil to f, and
// hands (cell, f, A f) to the pass's pointwise update, which may write vectors
// and accumulate dot products.  f is never stored unless the pass writes it:
// neighbours come from registers (z) and an LDS plane (x, y).
//
// Tiling (gfx950, wave64): a workgroup is 4 waves = 256 lanes covering a tile
// of TX=64 cells along x (one wave row, 512 contiguous bytes per vector) by
// TY=4 rows along y, and marches over a chunk of ZC planes along z.  For each
// plane it evaluates f at its 256 cells plus the 2*64 + 2*4 in-plane halo
// cells (lanes 0..31 of every wave take half a y-halo row, lanes 32/33 the two
// x-halo cells of their row), stores them in a double-buffered LDS plane and
// needs one workgroup barrier per plane.  The z neighbours stay in registers.
//
// Boundary: f is 0 outside the grid (the reference's zero ghost cells,
// Cg.hpp:57-62; DistrAsyncStencil.hpp:74-86 with non-periodic neighbours,
// DomainDecomposition.hpp:61-80).
//
// The stencil expression is the reference's, left to right
// (Stencil.hpp:43-50): a*x0 - c*x(i1-1) - c*x(i2-1) - c*x(i3-1)
//                          - b*x(i1+1) - b*x(i2+1) - b*x(i3+1);
// the library is compiled with -ffp-contract=off so each product and
// difference rounds once, as in the reference build.
#pragma once

#include "sedi_internal.h"

namespace sedi {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NTHREADS = TX * TY;

struct TileGrid {
    int n1, n2, n3;
    int nbx, nby, nbz, zc;
    long long s2, s3;
    long long ntiles;
};

inline TileGrid make_tile_grid(const Shape& s, int zc) {
    TileGrid t;
    t.n1 = s.n1; t.n2 = s.n2; t.n3 = s.n3;
    t.nbx = (s.n1 + TX - 1) / TX;
    t.nby = (s.n2 + TY - 1) / TY;
    t.zc = zc;
    t.nbz = (s.n3 + zc - 1) / zc;
    t.s2 = s.n1;
    t.s3 = (long long)s.n1 * s.n2;
    t.ntiles = (long long)t.nbx * t.nby * t.nbz;
    return t;
}

// ---- wave / block reductions (fixed order => run-to-run deterministic) ----
__device__ __forceinline__ double wave_sum(double v) {
    // butterfly over 64 lanes; lane 0's result is used
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Sums ND per-thread values over the workgroup; result valid in thread 0.
template <int ND>
__device__ __forceinline__ void block_sum(double (&acc)[ND], double* red /* LDS [ND][4] */) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const double s = wave_sum(acc[d]);
        if (lane == 0) red[d * 4 + w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++)
            acc[d] = ((red[d * 4 + 0] + red[d * 4 + 1]) + red[d * 4 + 2]) + red[d * 4 + 3];
    }
}

// Cross-workgroup reduction: each workgroup publishes its ND partial sums; the
// last one to arrive sums all partials in a fixed order and returns true with
// the totals in thread 0's `tot` (the other threads of that workgroup see
// true too).  The arrival counter is re-armed by the last workgroup, so no
// memset is needed between launches.
template <int ND>
__device__ bool grid_sum(double (&acc)[ND], double* partials, unsigned* counter,
                         double* red /* LDS >= 4*ND doubles */, int* flag /* LDS */) {
    block_sum<ND>(acc, red);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) partials[(long long)d * nb + blockIdx.x] = acc[d];
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        *flag = (prev == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!*flag) return false;
    __threadfence();
    double loc[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += NTHREADS)
            s += __hip_atomic_load(partials + (long long)d * nb + b, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        loc[d] = s;
    }
    __syncthreads();  // `red` is reused
    block_sum<ND>(loc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) acc[d] = loc[d];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// ---- the march -------------------------------------------------------------
// Pass interface:
//   struct Cell;                                   values loaded at one cell
//   __device__ Cell load(long long idx) const;     loads what f needs
//   __device__ double f(const Cell&) const;         the stencil input field
//   __device__ void finish(long long idx, const Cell& c, double fc, double Af,
//                          double (&acc)[ND]) const;   pointwise update at a cell
//   static constexpr int ND;                        number of dot products
// and the stencil coefficients a, b, c.
template <class P>
__device__ __forceinline__ void march_tile(const P& p, const TileGrid& g, long long tile,
                                           double (&lds)[2][TY + 2][TX + 2],
                                           double (&acc)[P::ND]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long t_xy = (long long)g.nbx * g.nby;
    const int bz = (int)(tile / t_xy);
    const long long rem = tile - (long long)bz * t_xy;
    const int by = (int)(rem / g.nbx);
    const int bx = (int)(rem - (long long)by * g.nbx);
    const int i0 = bx * TX, j0 = by * TY;
    const int zs = bz * g.zc;
    const int ze = min(zs + g.zc, g.n3);

    const int i = i0 + lane, j = j0 + w;
    const bool c_in = (i < g.n1) && (j < g.n2);
    const long long cxy = (long long)i + (long long)j * g.s2;

    // the one extra (halo) cell this lane evaluates per plane
    int ex_i = 0, ex_j = 0, ex_r = 0, ex_c = 0;
    bool ex_act = false;
    if (lane < 32) {
        ex_j = (w < 2) ? j0 - 1 : j0 + TY;
        ex_i = i0 + (w & 1) * 32 + lane;
        ex_r = (w < 2) ? 0 : TY + 1;
        ex_c = 1 + (w & 1) * 32 + lane;
        ex_act = true;
    } else if (lane == 32) {
        ex_i = i0 - 1; ex_j = j; ex_r = w + 1; ex_c = 0; ex_act = true;
    } else if (lane == 33) {
        ex_i = i0 + TX; ex_j = j; ex_r = w + 1; ex_c = TX + 1; ex_act = true;
    }
    const bool ex_in = ex_act && ex_i >= 0 && ex_i < g.n1 && ex_j >= 0 && ex_j < g.n2;
    const long long exy = (long long)ex_i + (long long)ex_j * g.s2;

    typename P::Cell cm, cc, cn;
    double fm = 0.0, fc = 0.0, fn = 0.0;
    if (c_in && zs - 1 >= 0) { cm = p.load(cxy + (long long)(zs - 1) * g.s3); fm = p.f(cm); }
    if (c_in) { cc = p.load(cxy + (long long)zs * g.s3); fc = p.f(cc); }
    lds[0][w + 1][lane + 1] = fc;
    if (ex_act) {
        double fe = 0.0;
        if (ex_in) fe = p.f(p.load(exy + (long long)zs * g.s3));
        lds[0][ex_r][ex_c] = fe;
    }
    __syncthreads();

    for (int z = zs; z < ze; z++) {
        const int b = (z - zs) & 1;
        const bool has_next = z + 1 < g.n3;
        fn = 0.0;
        if (c_in && has_next) { cn = p.load(cxy + (long long)(z + 1) * g.s3); fn = p.f(cn); }
        if (z + 1 < ze) {
            lds[b ^ 1][w + 1][lane + 1] = fn;
            if (ex_act) {
                double fe = 0.0;
                if (ex_in) fe = p.f(p.load(exy + (long long)(z + 1) * g.s3));
                lds[b ^ 1][ex_r][ex_c] = fe;
            }
        }
        const double xm = lds[b][w + 1][lane];
        const double xp = lds[b][w + 1][lane + 2];
        const double ym = lds[b][w][lane + 1];
        const double yp = lds[b][w + 2][lane + 1];
        const double Af = p.a * fc - p.c * xm - p.c * ym - p.c * fm
                          - p.b * xp - p.b * yp - p.b * fn;
        if (c_in) p.finish(cxy + (long long)z * g.s3, cc, fc, Af, acc);
        __syncthreads();
        fm = fc; fc = fn; cc = cn;
    }
}

}  // namespace sedi
