/*
 * SolverHip.hpp -- the reference-side adapter: intel/sedi's solver and probe
 * interfaces implemented over libsedi_amd's C ABI (sedi_c.h), so the
 * reference's stochastic-estimator.cc runs its sample loop on one MI355X.
 *
 *   sedi_hip::Solver<Dtype, KIND>  replaces Cg / PipelinedCg / BiCgStab: same
 *       constructor (DomainDecomposition&, nghost, mode, Timers*) (Cg.hpp:48,
 *       PipelinedCg.hpp:48, BiCgStab.hpp:51), same result_t codes 0..3
 *       (Cg.hpp:31-34, BiCgStab.hpp:32-37) and solve(A, comm_halo, x, b, iM,
 *       maxiter, tol, &final_err, &final_iter) (Cg.hpp:70-82), printing the
 *       same "<Solver>: after iter=%d, err=%.10e" lines (Cg.hpp:192,
 *       PipelinedCg.hpp:228, BiCgStab.hpp:228) from sedi_solver_err_history.
 *   sedi_hip::Rng<Dtype>  replaces DistrRngStd (DistrRngStd.hpp:15-44):
 *       fill_plusminusone(Slice3d&) through sedi_rademacher_fill, one
 *       mt19937 stream per run (seed from SEDI_RNG_SEED, else
 *       std::random_device as the reference), call k taking draws
 *       [2Nk, 2N(k+1)).
 *
 * Two ways to use it:
 *   1. edit the typedefs of stochastic-estimator.cc:134-144 to
 *      `typedef sedi_hip::Solver<real_t, SEDI_CG> solver_t;` and :108 to
 *      `sedi_hip::Rng<real_t> rng(dd);` after #include "SolverHip.hpp";
 *   2. leave the file untouched and compile it with
 *      `-include SolverHip.hpp -DSEDI_HIP_AS_REFERENCE_SOLVERS`: the adapter
 *      then defines the include guards of Cg.hpp, PipelinedCg.hpp,
 *      BiCgStab.hpp and DistrRngStd.hpp and provides classes of those names
 *      (oracle/Makefile target ref_hip builds the reference's own main that
 *      way).
 *
 * The whole grid lives on one GPU: the run must have one rank (-p 1x1x1);
 * the halo exchange and the dot products' allreduces of the reference have
 * no counterpart there.  Vectors cross as dense i1-fastest arrays of the
 * rank's cells (ghosts excluded); the library stages them to the device.
 */
#ifndef SEDI_SOLVER_HIP_HPP
#define SEDI_SOLVER_HIP_HPP

#include <mpi.h>
#include <fenv.h>
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "DomainDecomposition.hpp"
#include "Grid3d.hpp"
#include "Slice3d.hpp"
#include "Stencil.hpp"
#include "CommHalo.hpp"
#include "Timers.hpp"
#include "sedi_c.h"

namespace sedi_hip {

// The reference traps FP exceptions (feenableexcept, stochastic-estimator.cc:
// 64-69) to catch NaN/Inf in its own arithmetic.  The GPU traps nothing, and
// the HIP runtime's host code (its threads and its teardown at exit) is not
// written for trapping, so library calls run with traps masked and the traps
// are masked again before the runtime shuts down.
inline void mask_traps_at_exit() { fedisableexcept(FE_ALL_EXCEPT); }

struct TrapsMasked {
    int saved;
    TrapsMasked() : saved(fegetexcept()) {
        static bool registered = false;
        if (!registered) {
            registered = true;
            atexit(mask_traps_at_exit);  // runs before the runtime's teardown
        }
        fedisableexcept(FE_ALL_EXCEPT);
    }
    ~TrapsMasked() {
        feclearexcept(FE_ALL_EXCEPT);
        if (saved) feenableexcept(saved);
    }
};

inline void check(int rc, const char *what) {
    if (rc < 0) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, sedi_last_error());
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
}

// Dense copy of a slice's cells, i1 fastest (Slice3d::element_ref order).
template <typename Dtype>
inline void gather(const Slice3d<Dtype> &s, std::vector<double> &out) {
    out.resize((size_t)s.n1 * s.n2 * s.n3);
    size_t k = 0;
    for (long i3 = 0; i3 < s.n3; i3++)
        for (long i2 = 0; i2 < s.n2; i2++)
            for (long i1 = 0; i1 < s.n1; i1++) out[k++] = (double)s.element_ref(i1, i2, i3);
}

template <typename Dtype>
inline void scatter(const std::vector<double> &in, const Slice3d<Dtype> &s) {
    size_t k = 0;
    for (long i3 = 0; i3 < s.n3; i3++)
        for (long i2 = 0; i2 < s.n2; i2++)
            for (long i1 = 0; i1 < s.n1; i1++) s.element_ref(i1, i2, i3) = (Dtype)in[k++];
}

inline const char *solver_name(int kind) {
    return kind == SEDI_CG ? "Cg" : (kind == SEDI_PCG ? "PipelinedCg" : "BiCgStab");
}

template <typename Dtype, int KIND>
class Solver {
  public:
    // the codes of Cg.hpp:31-34 and BiCgStab.hpp:32-37 (0..3)
    typedef enum {
        SOLUTION_FOUND = SEDI_SOLUTION_FOUND,
        REACHED_MAXITER = SEDI_REACHED_MAXITER,
        BREAKDOWN_RHO = SEDI_BREAKDOWN_RHO,
        BREAKDOWN_OMEGA = SEDI_BREAKDOWN_OMEGA
    } result_t;

    DomainDecomposition &dd;
    const int mode;

    Solver(DomainDecomposition &dd_, long nghost, int mode_, Timers *timers = 00) : dd(dd_), mode(mode_) {
        (void)nghost;
        (void)timers;
        if (dd.np != 1) {
            if (dd.myrank == 0)
                fprintf(stderr, "sedi_hip::Solver: the grid lives on one GPU; run with -p 1x1x1 on one rank\n");
            MPI_Abort(MPI_COMM_WORLD, 2);
        }
        if (dd.myrank == 0)
            printf("Initialized %s solver, sync_dotprod=%d, sync_stencil=%d\n", solver_name(KIND),
                   (mode & 2) != 0, (mode & 1) != 0);
    }

    ~Solver() {
        TrapsMasked m;
        if (h) sedi_solver_destroy(h);
    }

    result_t solve(Stencil<Dtype> &A, CommHalo<Dtype> &comm_halo, Grid3d<Dtype> &x, Grid3d<Dtype> &b,
                   Grid3d<Dtype> &iM, int maxiter, Dtype tol, Dtype *final_err, int *final_iter) {
        (void)comm_halo;
        if (!h || A.a != sa || A.b != sb || A.c != sc) {
            TrapsMasked m;
            if (h) sedi_solver_destroy(h);
            h = 0;
            const sedi_grid g = {dd.ngp1, dd.ngp2, dd.ngp3};
            const sedi_stencil st = {(double)A.a, (double)A.b, (double)A.c};
            check(sedi_solver_create(&g, &st, KIND, mode, &h), "sedi_solver_create");
            sa = A.a, sb = A.b, sc = A.c;
        }
        gather(x.rank_cells, xv);
        gather(b.rank_cells, bv);
        gather(iM.rank_cells, mv);
        // the estimator's iM is the constant 1/a (stochastic-estimator.cc:103):
        // then the kernels keep it in a register instead of streaming it
        bool constant = true;
        const double inv_a = 1.0 / (double)A.a;
        for (size_t i = 0; i < mv.size() && constant; i++) constant = mv[i] == inv_a;
        double err = 0.0;
        int iter = 0, code = 0, nh = 0;
        {
            TrapsMasked m;
            code = sedi_solve(h, bv.data(), xv.data(), constant ? 0 : mv.data(), maxiter, (double)tol, &err, &iter);
            check(code, "sedi_solve");
            nh = sedi_solver_err_history(h, 0, 0);
            hist.resize(nh > 0 ? nh : 0);
            if (nh > 0) sedi_solver_err_history(h, hist.data(), nh);
        }
        scatter(xv, x.rank_cells);
        if (dd.myrank == 0)
            for (int i = 0; i < nh; i++) printf("%s: after iter=%d, err=%.10e\n", solver_name(KIND), i + 1, hist[i]);
        *final_err = (Dtype)err;
        *final_iter = iter;
        return (result_t)code;
    }

  private:
    sedi_solver *h = 0;
    Dtype sa = 0, sb = 0, sc = 0;
    std::vector<double> xv, bv, mv, hist;
};

template <typename Dtype>
class Rng {
  public:
    const DomainDecomposition &dd;

    Rng(const DomainDecomposition &dd_) : dd(dd_), seed(pick_seed()) {}

    void fill_plusminusone(Slice3d<Dtype> &slice) const {
        assert(slice.n == dd.ngp);
        const int64_t n = (int64_t)slice.n1 * slice.n2 * slice.n3;
        buf.resize((size_t)n);
        {
            TrapsMasked m;
            check(sedi_rademacher_fill(seed, draws, n, buf.data()), "sedi_rademacher_fill");
        }
        draws += 2 * (uint64_t)n;  // two 32-bit outputs per cell (DistrRngStd.hpp:36-40)
        scatter(buf, slice);
    }

  private:
    static uint32_t pick_seed() {
        const char *s = getenv("SEDI_RNG_SEED");
        if (s && *s) return (uint32_t)strtoul(s, 0, 10);
        std::random_device rd;
        return rd();
    }
    const uint32_t seed;
    mutable uint64_t draws = 0;
    mutable std::vector<double> buf;
};

}  // namespace sedi_hip

#ifdef SEDI_HIP_AS_REFERENCE_SOLVERS
// Take the place of the reference's solver and RNG headers: their include
// guards are set so stochastic-estimator.cc's #includes of them are empty,
// and classes of their names forward to the adapter.
#define CG_H
#define PIPELINED_CG_H
#define BI_CG_STAB_H
#define DISTR_RNG_STD_H
template <typename Dtype>
class Cg : public sedi_hip::Solver<Dtype, SEDI_CG> {
  public:
    Cg(DomainDecomposition &dd, long nghost, int mode, Timers *t = 00) : sedi_hip::Solver<Dtype, SEDI_CG>(dd, nghost, mode, t) {}
};
template <typename Dtype>
class PipelinedCg : public sedi_hip::Solver<Dtype, SEDI_PCG> {
  public:
    PipelinedCg(DomainDecomposition &dd, long nghost, int mode, Timers *t = 00)
        : sedi_hip::Solver<Dtype, SEDI_PCG>(dd, nghost, mode, t) {}
};
template <typename Dtype>
class BiCgStab : public sedi_hip::Solver<Dtype, SEDI_BICGSTAB> {
  public:
    BiCgStab(DomainDecomposition &dd, long nghost, int mode, Timers *t = 00)
        : sedi_hip::Solver<Dtype, SEDI_BICGSTAB>(dd, nghost, mode, t) {}
};
template <typename Dtype>
class DistrRngStd : public sedi_hip::Rng<Dtype> {
  public:
    DistrRngStd(const DomainDecomposition &dd) : sedi_hip::Rng<Dtype>(dd) {}
};
#endif

#endif /* SEDI_SOLVER_HIP_HPP */
