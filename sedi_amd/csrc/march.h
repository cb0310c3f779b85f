// march.h -- the z-marching 7-point pass that every solver kernel is built on.
//
// A pass computes, for every cell, an input field f (a pointwise function of
// the solver vectors, e.g. p = r*iM + beta*p), applies the stencil to f, and
// hands (cell, f, A f) to the pass's pointwise update, which may write vectors
// and accumulate dot products.  f is never stored unless the pass writes it:
// neighbours come from registers (z) and an LDS plane (x, y).
//
// Tiling (gfx950, wave64): a workgroup has one wave per tile row; a tile is
// TX=64 cells along x (one wave row, 512 contiguous bytes per FP64 vector) by
// TYK rows along y (4; 8 for the passes whose registers allow two such
// workgroups per CU; 16 for the standalone stencil and, on large grids, the
// batched CG update pass), and marches over a chunk of ZC planes along z.  For
// each plane it evaluates f at its cells plus the 2*64 + 2*TYK in-plane halo
// cells (waves 0 and 1 take the y-halo rows, wave 2 the x-halo cells of every
// row; the other waves load no halo), stores them in a double-buffered LDS
// plane and needs one workgroup barrier per plane.  The z neighbours stay in
// registers.  The y halo costs 2/TYK extra rows per row.
//
// Latency: all loads are issued D planes before they are consumed, into a
// ring of registers whose slots are compile-time constants (the z loop is
// unrolled by the ring size), so every wave keeps D planes of loads in flight
// across the per-plane barrier.
//
// Tile order: workgroup b runs on XCD b % 8.  When the tile count allows, the
// tiles are renumbered so that each XCD gets a contiguous run of tiles (whole
// xy layers of a z chunk); a tile's x/y halo cells are then read by workgroups
// of the same XCD at about the same time and hit that XCD's L2.
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

#include <algorithm>
#include <type_traits>

namespace sedi {

constexpr int TX = 64;
constexpr int TY = 4;  // default tile height: rows per tile = waves per workgroup
constexpr int NTHREADS = TX * TY;
constexpr int NWAVES = NTHREADS / 64;
// Tall tiles (8 rows, 8 waves) halve the y-halo re-reads of a pass; kernels
// whose registers allow it (four waves per SIMD, 128 VGPRs, so that two
// workgroups fit on a CU) use them.
constexpr int TY_TALL = 8;
constexpr int NXCD = 8;

struct TileGrid {
    int n1, n2, n3;
    int nbx, nby, nbz, zc, ty;
    int tx;   // cells per tile row: TX (one per lane) or TX2 (a pair per lane)
    int xcd;  // 1: each XCD takes a contiguous run of tiles; 2: chunks of xg tiles
              // dealt round-robin to the XCDs; 0: tiles in launch order
    int xg;   // xcd == 2: tiles per chunk (whole rows of tiles along x)
    long long s2, s3;
    long long ntiles;
};

inline TileGrid make_tile_grid(const Shape& s, int zc, int ty = TY, int tx = TX) {
    TileGrid t;
    t.n1 = s.n1; t.n2 = s.n2; t.n3 = s.n3;
    t.tx = tx;
    t.nbx = (s.n1 + tx - 1) / tx;
    t.nby = (s.n2 + ty - 1) / ty;
    t.zc = zc;
    t.ty = ty;
    t.xcd = 1;
    t.xg = t.nbx;
    t.nbz = (s.n3 + zc - 1) / zc;
    t.s2 = s.n1;
    t.s3 = (long long)s.n1 * s.n2;
    t.ntiles = (long long)t.nbx * t.nby * t.nbz;
    return t;
}

// Launch geometry of one kernel on one grid: tiles and workgroup count.
struct KGeom {
    TileGrid tg;
    int nblocks;
};

// Chooses the z chunk so that the tiles fill the workgroups that are resident
// at once (`resident` = occupancy x CUs): the makespan of a pass is about
// rounds * (chunk + prologue) plane steps of one workgroup, where a round is
// one tile per resident workgroup and the prologue costs ~2 + D planes.
// The march is compiled for chunks of 8, 16 and 32 planes.
constexpr int ZC_CHOICES[3] = {8, 16, 32};

// The launch plan the autotuner may override per kernel and grid: "zc:<kernel>"
// the planes per tile, "sched:<kernel>" 1 for as many workgroups as the
// partial sums allow (up to 8 per CU) in launch order instead of the resident
// set walking XCD runs, 2 for one workgroup per tile in launch order (no
// grid-stride loop; the hardware dispatcher hands out tiles as workgroups
// retire) where the kernel has no partial sums to bound its workgroups
// (`one_per_tile` false: as 1), 3 and 4 as 1 but with the tiles dealt to the
// XCDs in chunks of 4 or 8 rows of tiles: a tile and its y neighbours then
// run on one XCD at about the same time, so the halo rows one of them loads
// are in that XCD's L2 for the others, while the eight XCDs still work on
// one region of memory.
inline KGeom tuned_geometry(const KGeom& planned, const Shape& s, int zc, int sched, int max_blocks,
                            bool one_per_tile = false) {
    KGeom k = planned;
    if (zc == 8 || zc == 16 || zc == 32) {
        k.tg = make_tile_grid(s, zc, planned.tg.ty, planned.tg.tx);
        k.nblocks = (int)std::min<long long>(k.tg.ntiles, planned.nblocks);
    }
    if (sched >= 1 && sched <= 4) {
        k.tg.xcd = 0;
        const long long cap = (sched == 2 && one_per_tile) ? k.tg.ntiles : (long long)max_blocks;
        k.nblocks = (int)std::min<long long>(k.tg.ntiles, cap);
        if (sched >= 3) {
            k.tg.xcd = 2;
            k.tg.xg = k.tg.nbx * std::min(k.tg.nby, sched == 3 ? 4 : 8);
            // the dispatcher deals workgroup b to XCD b % 8: the chunked order
            // needs every XCD to take whole rounds
            k.nblocks = std::max(NXCD, k.nblocks / NXCD * NXCD);
        }
    }
    return k;
}

inline KGeom plan_geometry(const Shape& s, int resident, int D, int ty = TY, int tx = TX) {
    const long long nxy = (long long)((s.n1 + tx - 1) / tx) * ((s.n2 + ty - 1) / ty);
    long long best_cost = -1;
    int best_zc = 16;
    for (int zc : ZC_CHOICES) {
        const long long nt = nxy * ((s.n3 + zc - 1) / zc);
        const long long rounds = (nt + resident - 1) / resident;
        const long long cost = rounds * (zc + 2 + D);
        if (best_cost < 0 || cost < best_cost || (cost == best_cost && zc > best_zc)) {
            best_cost = cost;
            best_zc = zc;
        }
    }
    KGeom k;
    k.tg = make_tile_grid(s, best_zc, ty, tx);
    k.nblocks = (int)std::min<long long>(k.tg.ntiles, resident);
    return k;
}

// ---- wave / block reductions (fixed order => run-to-run deterministic) ----
__device__ __forceinline__ double wave_sum(double v) {
    // butterfly over 64 lanes; lane 0's result is used
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Sums ND per-thread values over the workgroup; result valid in thread 0.
template <int ND, int NW = NWAVES>
__device__ __forceinline__ void block_sum(double (&acc)[ND], double* red /* LDS [ND][NW] */) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const double s = wave_sum(acc[d]);
        if (lane == 0) red[d * NW + w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // wave by wave, all ND sums at once: each sum keeps its order
        // (wave 0, 1, ..., NW-1) while at most 2*ND values are live (a fully
        // unrolled loop loads all ND*NW at once, which sets the register
        // peak of kernels with many dot products)
#pragma unroll
        for (int d = 0; d < ND; d++) acc[d] = red[d * NW];
#pragma unroll 1
        for (int v = 1; v < NW; v++) {
#pragma unroll
            for (int d = 0; d < ND; d++) acc[d] += red[d * NW + v];
        }
    }
}

// Cross-workgroup reduction: each workgroup publishes its ND partial sums; the
// last one to arrive sums all partials in a fixed order and returns true with
// the totals in thread 0's `acc` (the other threads of that workgroup see
// true too).  The arrival counter is re-armed by the last workgroup, so no
// memset is needed between launches.
//
// Publication without a device-wide fence: a release fence at agent scope
// writes back the whole L2 of the XCD (the XCDs' L2s are not coherent), once
// per workgroup.  Instead the partials are stored with agent-scope atomic
// stores (written through to the coherence point), the lane waits until they
// have completed (vmcnt 0) before the agent-scope arrival increment, and the
// last workgroup reads them with agent-scope atomic loads.
template <int ND, int NW = NWAVES>
__device__ bool grid_sum(double (&acc)[ND], double* partials, unsigned* counter,
                         double* red /* LDS >= ND*NW doubles */, int* flag /* LDS */) {
    block_sum<ND, NW>(acc, red);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
#ifdef SEDI_GRID_SUM_FENCE
#pragma unroll
        for (int d = 0; d < ND; d++) partials[(long long)d * nb + blockIdx.x] = acc[d];
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
#else
#pragma unroll
        for (int d = 0; d < ND; d++)
            __hip_atomic_store(partials + (long long)d * nb + blockIdx.x, acc[d], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);  // all counters 0: the stores have completed
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
#endif
        *flag = (prev == (unsigned)(nb - 1));
    }
    __syncthreads();
    if (!*flag) return false;
#ifdef SEDI_GRID_SUM_FENCE
    __threadfence();
#endif
    double loc[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += NW * 64)
            s += __hip_atomic_load(partials + (long long)d * nb + b, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        loc[d] = s;
    }
    __syncthreads();  // `red` is reused
    block_sum<ND, NW>(loc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) acc[d] = loc[d];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// Workgroup barrier that orders LDS only.  __syncthreads() is a workgroup
// fence on all memory and waits for every outstanding global load before the
// barrier, which would drain the prefetch ring once per plane; the march
// communicates between waves through LDS only, so it needs just this.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Tile handled by step t of the grid-stride loop.  With xcd set, each XCD's
// workgroups walk a contiguous run of tiles (halo rows shared in its L2);
// otherwise consecutive tiles go to consecutive workgroups, which keeps the
// XCDs' concurrent accesses within one region of memory (measured faster on
// grids much larger than the Infinity Cache; the autotuner decides).  `reverse` walks each
// XCD's run of tiles backwards: consecutive passes alternate direction, so a
// pass starts on the tiles the previous pass touched last, whose vectors are
// still in the 256 MB Infinity Cache.
__device__ __forceinline__ long long tile_of(const TileGrid& g, long long t, bool reverse) {
    if (g.xcd == 1 && (g.ntiles % NXCD) == 0 && (gridDim.x % NXCD) == 0) {
        const long long per = g.ntiles / NXCD;
        const long long k = t / NXCD;
        return (t % NXCD) * per + (reverse ? per - 1 - k : k);
    }
    const long long u = reverse ? g.ntiles - 1 - t : t;
    if (g.xcd == 2 && (g.ntiles % NXCD) == 0 && (gridDim.x % NXCD) == 0) {
        // step u runs on XCD u % 8 (as ntiles and the grid are multiples of
        // 8, for both walking directions); its k-th step takes tile k % xg of
        // chunk (k / xg) * 8 + u % 8.  Tiles past the last whole round of
        // 8 chunks keep their place.  A bijection of [0, ntiles).
        const long long span = (long long)g.xg * NXCD;
        const long long full = g.ntiles / span * span;
        if (u < full) {
            const long long k = u / NXCD;
            return ((k / g.xg) * NXCD + (u % NXCD)) * g.xg + (k % g.xg);
        }
    }
    return u;
}

// ---- 32-bit cell offsets ---------------------------------------------------
// The march hands passes a cell's index within one vector as 32 bits; every
// vector a pass touches is smaller than 4 GiB (checked on the host), so the
// byte offset index*sizeof(T) is also 32-bit and loads take the vector's base
// from scalar registers plus one 32-bit offset register, with no 64-bit
// address arithmetic per access.
using cidx_t = unsigned;

template <class T>
__device__ __forceinline__ T ldc(const T* base, cidx_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)(i * (unsigned)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void stc(T* base, cidx_t i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)(i * (unsigned)sizeof(T))) = v;
}
// Non-temporal store: for outputs that no later kernel of the same sweep reads.
__device__ __forceinline__ void stc_nt(double* base, cidx_t i, double v) {
    __builtin_nontemporal_store(v, reinterpret_cast<double*>(reinterpret_cast<char*>(base) +
                                                             (unsigned)(i * (unsigned)sizeof(double))));
}

// ---- the march -------------------------------------------------------------
// Pass interface (offsets are cell indices; the field f has type P::V, double
// for the plain solvers and a 4-RHS vector for the batched ones):
//   static constexpr int ND;                        number of dot products
//   static constexpr int D;                         prefetch distance (planes)
//   struct Raw;   Raw load(cidx_t) const;           what f needs at a cell
//   double f(const Raw&) const;                     the stencil input field
//   struct Aux;   Aux load_aux(cidx_t) const;       other inputs of the update
//   void finish(cidx_t, const Raw&, const Aux&, V f, V Af,
//               double (&acc)[ND]) const;           pointwise update at a cell
// and the stencil coefficients a, b, c.
//
// ZC (planes per tile) is a compile-time constant and the march over the
// tile's planes is fully unrolled: the prefetch ring's slots are then fixed
// registers with no loop back-edge to reconcile, and the compiler's wait
// counts keep D planes of loads in flight all the way through.  A tile that
// runs past the last plane still executes every step (clamped loads, zeroed
// values) but does not store.
template <class P, class = void>
struct value_of { using type = double; };
template <class P>
struct value_of<P, std::void_t<typename P::V>> { using type = typename P::V; };
template <class P>
using value_t = typename value_of<P>::type;

template <class P, int ZC, int TYK, bool EX>
__device__ __forceinline__ void march_tile(const P& p, const TileGrid& g, long long tile,
                                           value_t<P> (&lds)[2][TYK + 2][TX + 2],
                                           double (&acc)[P::ND]) {
    static_assert(TYK >= 3 && 2 * TYK <= 64, "three waves load the halo; wave 2 covers 2*TYK x-halo cells");
    using V = value_t<P>;
    const V zero{};
    constexpr int D = P::D;
    constexpr int R = D + 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long t_xy = (long long)g.nbx * g.nby;
    const int bz = (int)(tile / t_xy);
    const long long rem = tile - (long long)bz * t_xy;
    const int by = (int)(rem / g.nbx);
    const int bx = (int)(rem - (long long)by * g.nbx);
    const int i0 = bx * TX, j0 = by * TYK;
    const int zs = bz * ZC;
    const int n3 = g.n3;
    const cidx_t s3 = (cidx_t)g.s3;

    const int i = i0 + lane, j = j0 + w;
    const bool c_in = (i < g.n1) && (j < g.n2);
    // every load below is unconditional from a clamped, in-grid address and
    // out-of-grid values are replaced by 0 afterwards: the body is straight-line
    // code, so the compiler's wait counts on the prefetch ring stay exact
    const cidx_t cxy = (cidx_t)(min(i, g.n1 - 1) + min(j, g.n2 - 1) * g.n1);

    // the halo: waves 0 and 1 evaluate the y-halo rows below and above the
    // tile (one cell per lane), lanes 2r and 2r+1 of wave 2 the x-halo cells
    // left and right of tile row r; the other waves load no halo, so a plane
    // costs them one load per field.  EX (true for waves 0-2) is a compile-time
    // constant: each kind of wave runs straight-line code whose wait counts
    // keep exactly D planes of its own loads in flight (a run-time branch
    // around the halo loads made the compiler wait for the next plane's loads
    // as well in the halo waves, which every wave then waited for at the
    // plane's barrier)
    const int wu = __builtin_amdgcn_readfirstlane(w);
    int ex_i = i, ex_j = j, ex_r = w + 1, ex_c = lane + 1;
    bool ex_act = false;
    if (!EX) {
    } else if (wu == 0) {
        ex_j = j0 - 1; ex_r = 0; ex_act = true;
    } else if (wu == 1) {
        ex_j = j0 + TYK; ex_r = TYK + 1; ex_act = true;
    } else if (wu == 2 && lane < 2 * TYK) {
        ex_j = j0 + (lane >> 1);
        ex_r = (lane >> 1) + 1;
        ex_i = (lane & 1) ? i0 + TX : i0 - 1;
        ex_c = (lane & 1) ? TX + 1 : 0;
        ex_act = true;
    }
    const bool ex_in = ex_act && ex_i >= 0 && ex_i < g.n1 && ex_j >= 0 && ex_j < g.n2;
    const cidx_t exy = (cidx_t)(min(max(ex_i, 0), g.n1 - 1) + min(max(ex_j, 0), g.n2 - 1) * g.n1);

    // rings: f-inputs of plane zs+q-1 live in slot q % R, update inputs of
    // plane zs+q in slot q % R
    typename P::Raw rc[R], rx[R];
    typename P::Aux ra[R];
    auto zoff = [&](int zz) { return (cidx_t)min(max(zz, 0), n3 - 1) * s3; };
    auto ld_f = [&](int zz, int slot) {
        rc[slot] = p.load(cxy + zoff(zz));
        if (EX) rx[slot] = p.load(exy + zoff(zz));
    };
    auto ld_a = [&](int zz, int slot) { ra[slot] = p.load_aux(cxy + zoff(zz)); };

    // prologue: planes zs-1 and zs now, zs+1 .. zs+D in flight, update inputs
    // of planes zs .. zs+D-1 in flight
    ld_f(zs - 1, 0);
    ld_f(zs, 1 % R);
#pragma unroll
    for (int k = 1; k < D && k <= ZC; k++) ld_f(zs + k, (k + 1) % R);
#pragma unroll
    for (int k = 0; k < D && k < ZC; k++) ld_a(zs + k, k);
    V fm = (c_in && zs >= 1) ? p.f(rc[0]) : zero;
    V fc = (c_in && zs < n3) ? p.f(rc[1 % R]) : zero;
    typename P::Raw cc = rc[1 % R];
    const V fe0 = (ex_in && zs < n3) ? p.f(rx[1 % R]) : zero;
    lds[0][w + 1][lane + 1] = fc;
    if (ex_act) lds[0][ex_r][ex_c] = fe0;
    if (D <= ZC) ld_f(zs + D, (D + 1) % R);
    lds_barrier();

#pragma unroll
    for (int u = 0; u < ZC; u++) {
        const int z = zs + u;
        const int b = u & 1;
        // prefetch plane z+1+D (f-inputs) and plane z+D (update inputs)
        if (u + 1 + D <= ZC) ld_f(z + 1 + D, (u + 1) % R);
        if (u + D < ZC) ld_a(z + D, (u + D) % R);
        // f on plane z+1 (centre and halo) into the other LDS buffer
        const bool nxt = z + 1 < n3;
        const typename P::Raw cn = rc[(u + 2) % R];
        const V fn = (c_in && nxt) ? p.f(cn) : zero;
        if (u + 1 < ZC) {
            const V fe = (ex_in && nxt) ? p.f(rx[(u + 2) % R]) : zero;
            lds[b ^ 1][w + 1][lane + 1] = fn;
            if (ex_act) lds[b ^ 1][ex_r][ex_c] = fe;
        }
        const V xm = lds[b][w + 1][lane];
        const V xp = lds[b][w + 1][lane + 2];
        const V ym = lds[b][w][lane + 1];
        const V yp = lds[b][w + 2][lane + 1];
        const V Af = p.a * fc - p.c * xm - p.c * ym - p.c * fm
                     - p.b * xp - p.b * yp - p.b * fn;
        if (c_in && z < n3) p.finish(cxy + (cidx_t)z * s3, cc, ra[u % R], fc, Af, acc);
        lds_barrier();
        fm = fc;
        fc = fn;
        cc = cn;
    }
}

// Grid-stride loop over the tiles of a pass (tile order per tile_of).
template <class P, int ZC, int TYK = TY>
__device__ __forceinline__ void run_tiles(const P& p, const TileGrid& g,
                                          value_t<P> (&lds)[2][TYK + 2][TX + 2],
                                          double (&acc)[P::ND], bool reverse = false) {
#pragma unroll
    for (int d = 0; d < P::ND; d++) acc[d] = 0.0;
    // waves 0-2 load the halo (march_tile's EX); both bodies cross the same
    // barriers
    if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) < 3) {
        for (long long t = blockIdx.x; t < g.ntiles; t += gridDim.x)
            march_tile<P, ZC, TYK, true>(p, g, tile_of(g, t, reverse), lds, acc);
    } else {
        for (long long t = blockIdx.x; t < g.ntiles; t += gridDim.x)
            march_tile<P, ZC, TYK, false>(p, g, tile_of(g, t, reverse), lds, acc);
    }
}

// ---- the march with two cells per lane (FP64 pairs) -------------------------
// Lane l of a wave owns the adjacent cells i0+2l and i0+2l+1, so a tile row is
// TX2 = 128 cells and every access moves 16 bytes (one FP64 pair).  This needs
// n1 even (a pair never straddles a row, and the pair index stays even) and
// 16-byte aligned vectors; the host checks both.  A plane's second load is
// made by three waves only: waves 0 and 1 take the y-halo rows below and above
// the tile (one pair per lane), and lanes 2r and 2r+1 of wave 2 take the pairs
// left and right of tile row r, of which the cells i0-1 and i0+128 are used.
// The other waves issue one load per plane, where the one-cell march issues
// two.  The LDS row holds the tile in columns 2..129 (16-byte aligned pairs),
// cell i0-1 in column 1 and cell i0+128 in column 130.
//
// Pass interface (a pair of adjacent cells; offsets are the first cell's):
//   static constexpr int ND, D;
//   struct Raw;  Raw load(cidx_t) const;   double2 f(const Raw&) const;
//   struct Aux;  Aux load_aux(cidx_t) const;
//   void finish(cidx_t, const Raw&, const Aux&, double2 f, double2 Af, double (&acc)[ND]) const;
// The stencil is the reference's expression, left to right, for each cell.
constexpr int TX2 = 128;
constexpr int LX2 = TX2 + 4;

__device__ __forceinline__ double2 ldc2(const double* base, cidx_t i) {
    return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) + (unsigned)(i * 8u));
}
// Non-temporal store of a pair (the two stores merge into one 16-byte store).
__device__ __forceinline__ void stc2_nt(double* base, cidx_t i, double2 v) {
    double* q = reinterpret_cast<double*>(reinterpret_cast<char*>(base) + (unsigned)(i * 8u));
    __builtin_nontemporal_store(v.x, q);
    __builtin_nontemporal_store(v.y, q + 1);
}

template <class P, int ZC, int TYK, bool EX>
__device__ __forceinline__ void march_tile_x2(const P& p, const TileGrid& g, long long tile,
                                              double (&lds)[2][TYK + 2][LX2], double (&acc)[P::ND]) {
    static_assert(TYK >= 3 && 2 * TYK <= 64, "three waves load the halo; wave 2 covers 2*TYK x-halo pairs");
    constexpr int D = P::D;
    constexpr int R = D + 1;
    const double2 zero = {0.0, 0.0};
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long t_xy = (long long)g.nbx * g.nby;
    const int bz = (int)(tile / t_xy);
    const long long rem = tile - (long long)bz * t_xy;
    const int by = (int)(rem / g.nbx);
    const int bx = (int)(rem - (long long)by * g.nbx);
    const int i0 = bx * TX2, j0 = by * TYK, zs = bz * ZC;
    const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
    const cidx_t s3 = (cidx_t)g.s3;
    const int i = i0 + 2 * lane, j = j0 + w;
    const bool c_in = (i < n1) && (j < n2);
    // clamped, in-grid, even pair index (n1 is even)
    const cidx_t cxy = (cidx_t)(min(i, n1 - 2) + min(j, n2 - 1) * n1);

    int ei = i, ej = j, xr = 0, xs = 0;
    bool ex_in = false;
    if (!EX) {
    } else if (w == 0) {
        ej = j0 - 1;
        ex_in = i < n1 && ej >= 0;
    } else if (w == 1) {
        ej = j0 + TYK;
        ex_in = i < n1 && ej < n2;
    } else if (w == 2) {
        xr = lane >> 1;
        xs = lane & 1;
        ej = j0 + xr;
        ei = xs ? i0 + TX2 : i0 - 2;
        ex_in = lane < 2 * TYK && ej < n2 && ei >= 0 && ei < n1;
    }
    const cidx_t exy = (cidx_t)(min(max(ei, 0), n1 - 2) + min(max(ej, 0), n2 - 1) * n1);

    typename P::Raw rc[R], rx[R];
    typename P::Aux ra[R];
    auto zoff = [&](int zz) { return (cidx_t)min(max(zz, 0), n3 - 1) * s3; };
    auto ld_f = [&](int zz, int slot) {
        rc[slot] = p.load(cxy + zoff(zz));
        if (EX) rx[slot] = p.load(exy + zoff(zz));
    };
    auto ld_a = [&](int zz, int slot) { ra[slot] = p.load_aux(cxy + zoff(zz)); };
    auto put = [&](int buf, const double2& fcv, const double2& fe) {
        *reinterpret_cast<double2*>(&lds[buf][w + 1][2 + 2 * lane]) = fcv;
        if (!EX) {
        } else if (w == 0) {
            *reinterpret_cast<double2*>(&lds[buf][0][2 + 2 * lane]) = fe;
        } else if (w == 1) {
            *reinterpret_cast<double2*>(&lds[buf][TYK + 1][2 + 2 * lane]) = fe;
        } else if (w == 2 && lane < 2 * TYK) {
            if (xs) lds[buf][xr + 1][2 + TX2] = fe.x;
            else lds[buf][xr + 1][1] = fe.y;
        }
    };

    ld_f(zs - 1, 0);
    ld_f(zs, 1 % R);
#pragma unroll
    for (int k = 1; k < D && k <= ZC; k++) ld_f(zs + k, (k + 1) % R);
#pragma unroll
    for (int k = 0; k < D && k < ZC; k++) ld_a(zs + k, k);
    double2 fm = (c_in && zs >= 1) ? p.f(rc[0]) : zero;
    double2 fc = (c_in && zs < n3) ? p.f(rc[1 % R]) : zero;
    typename P::Raw cc = rc[1 % R];
    put(0, fc, (ex_in && zs < n3) ? p.f(rx[1 % R]) : zero);
    if (D <= ZC) ld_f(zs + D, (D + 1) % R);
    lds_barrier();

#pragma unroll
    for (int u = 0; u < ZC; u++) {
        const int z = zs + u;
        const int b = u & 1;
        if (u + 1 + D <= ZC) ld_f(z + 1 + D, (u + 1) % R);
        if (u + D < ZC) ld_a(z + D, (u + D) % R);
        const bool nxt = z + 1 < n3;
        const typename P::Raw cn = rc[(u + 2) % R];
        const double2 fn = (c_in && nxt) ? p.f(cn) : zero;
        if (u + 1 < ZC) put(b ^ 1, fn, (ex_in && nxt) ? p.f(rx[(u + 2) % R]) : zero);
        const double2 xl = *reinterpret_cast<const double2*>(&lds[b][w + 1][2 * lane]);      // cells 2l-2, 2l-1
        const double2 xh = *reinterpret_cast<const double2*>(&lds[b][w + 1][2 * lane + 4]);  // cells 2l+2, 2l+3
        const double2 ym = *reinterpret_cast<const double2*>(&lds[b][w][2 + 2 * lane]);
        const double2 yp = *reinterpret_cast<const double2*>(&lds[b][w + 2][2 + 2 * lane]);
        double2 Af;
        Af.x = p.a * fc.x - p.c * xl.y - p.c * ym.x - p.c * fm.x - p.b * fc.y - p.b * yp.x - p.b * fn.x;
        Af.y = p.a * fc.y - p.c * fc.x - p.c * ym.y - p.c * fm.y - p.b * xh.x - p.b * yp.y - p.b * fn.y;
        if (c_in && z < n3) p.finish(cxy + (cidx_t)z * s3, cc, ra[u % R], fc, Af, acc);
        lds_barrier();
        fm = fc;
        fc = fn;
        cc = cn;
    }
}

template <class P, int ZC, int TYK>
__device__ __forceinline__ void run_tiles_x2(const P& p, const TileGrid& g, double (&lds)[2][TYK + 2][LX2],
                                             double (&acc)[P::ND], bool reverse = false) {
#pragma unroll
    for (int d = 0; d < P::ND; d++) acc[d] = 0.0;
    if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) < 3) {
        for (long long t = blockIdx.x; t < g.ntiles; t += gridDim.x)
            march_tile_x2<P, ZC, TYK, true>(p, g, tile_of(g, t, reverse), lds, acc);
    } else {
        for (long long t = blockIdx.x; t < g.ntiles; t += gridDim.x)
            march_tile_x2<P, ZC, TYK, false>(p, g, tile_of(g, t, reverse), lds, acc);
    }
}

}  // namespace sedi
