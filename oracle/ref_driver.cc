// ref_driver.cc -- runs intel/sedi's own solver headers with seeded probes.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_ref.sh, in the
// development container only, against the unmodified reference headers
// (Stencil.hpp, Cg.hpp, PipelinedCg.hpp, BiCgStab.hpp, SliceOps.hpp,
// DistrAsync*.hpp, CommHalo.hpp, Grid3d.hpp, Slice3d.hpp,
// DomainDecomposition.hpp, Timers.hpp) into oracle/_ref/sedi_ref.  Its outputs
// become the fixtures that pin the C restatement, and it is the "reference"
// CPU baseline that bench.py times on the host cores.
//
// It differs from stochastic-estimator.cc:57-232 in three declared ways:
//   1. probes come from one std::mt19937(seed) stream per run instead of a
//      std::random_device re-seed per call (DistrRngStd.hpp:27-29); cell c of
//      sample k uses outputs 2Nk+2c and 2Nk+2c+1, and every rank walks the
//      global stream and keeps its own block, so the probes do not depend on -p;
//   2. the solver is a runtime flag (-S 1 BiCGStab, 2 CG, 3 PipelinedCg; the
//      SOLVER numbering of stochastic-estimator.cc:134-142);
//   3. with -o FILE, rank 0 writes per-sample iterations, final errors and
//      codes and the gathered D = t./q (the reference computes D at :216 but
//      never writes it).
// The sample loop, accumulation order, exit codes and output lines are the
// reference's.
#include <mpi.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <cmath>
#include <random>
#include <vector>
#include <unistd.h>
#include <fenv.h>

#include "DomainDecomposition.hpp"
#include "Grid3d.hpp"
#include "SliceOps.hpp"
#include "CommHalo.hpp"
#include "Stencil.hpp"
#include "DistrAsyncStencil.hpp"
#include "PipelinedCg.hpp"
#include "Cg.hpp"
#include "BiCgStab.hpp"
#include "Timers.hpp"

typedef double real_t;

// Seeded replacement for DistrRngStd (same fill_plusminusone method).
class SeededRademacher {
  public:
    SeededRademacher(const DomainDecomposition &dd_, uint32_t seed) : dd(dd_), eng(seed) {}
    void fill_plusminusone(Slice3d<real_t> &slice) {
        int p1, p2, p3;
        dd.rank_1d_to_3d(dd.myrank, &p1, &p2, &p3);
        const long o1 = p1 * dd.ngp1, o2 = p2 * dd.ngp2, o3 = p3 * dd.ngp3;
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        for (long g3 = 0; g3 < dd.ng3; g3++)
            for (long g2 = 0; g2 < dd.ng2; g2++)
                for (long g1 = 0; g1 < dd.ng1; g1++) {
                    const double x = dist(eng);
                    const long l1 = g1 - o1, l2 = g2 - o2, l3 = g3 - o3;
                    if (l1 >= 0 && l1 < dd.ngp1 && l2 >= 0 && l2 < dd.ngp2 && l3 >= 0 && l3 < dd.ngp3)
                        slice.element_ref(l1, l2, l3) = (x > 0.5 ? 1.0 : -1.0);
                }
    }
  private:
    const DomainDecomposition &dd;
    std::mt19937 eng;
};

struct Args {
    long ng1 = 0, ng2 = 0, ng3 = 0;
    int np1 = 0, np2 = 0, np3 = 0;
    int samples = 400;
    double tol = 1e-4;
    int maxiter = 1000;
    int mode = 0;
    bool ignore = false, no_fpe = false;
    int solver = 2;
    uint32_t seed = 1234;
    const char *out = nullptr;
};

static void usage(const char *prog) {
    fprintf(stderr, "Usage: %s -g NxNxN -p NxNxN [-s NSAMPLES] [-e TOL] [-m MAXITER] [-i] [-f MODE] [-F]"
                    " [-S SOLVER(1 BiCGStab,2 CG,3 PipelinedCg)] [-r SEED] [-o OUTFILE]\n", prog);
    exit(1);
}

static Args parse(int argc, char **argv) {
    Args a;
    bool have_g = false, have_p = false;
    int c;
    while ((c = getopt(argc, argv, "g:p:s:e:m:if:FS:r:o:")) != -1) {
        switch (c) {
        case 'g': have_g = sscanf(optarg, "%ldx%ldx%ld", &a.ng1, &a.ng2, &a.ng3) == 3; break;
        case 'p': have_p = sscanf(optarg, "%dx%dx%d", &a.np1, &a.np2, &a.np3) == 3; break;
        case 's': if (sscanf(optarg, "%d", &a.samples) != 1) usage(argv[0]); break;
        case 'e': if (sscanf(optarg, "%lf", &a.tol) != 1) usage(argv[0]); break;
        case 'm': if (sscanf(optarg, "%d", &a.maxiter) != 1) usage(argv[0]); break;
        case 'i': a.ignore = true; break;
        case 'f': if (sscanf(optarg, "%d", &a.mode) != 1) usage(argv[0]); break;
        case 'F': a.no_fpe = true; break;
        case 'S': a.solver = atoi(optarg); break;
        case 'r': a.seed = (uint32_t)strtoul(optarg, nullptr, 10); break;
        case 'o': a.out = optarg; break;
        default: usage(argv[0]);
        }
    }
    if (!have_g || !have_p || a.solver < 1 || a.solver > 3) usage(argv[0]);
    return a;
}

template <class solver_t>
static int run(const Args &args, DomainDecomposition &dd, Timers &timers) {
    Stencil<real_t> stencil(6.0, 1.0, 1.0);
    const long nghost = 1, ninner = nghost;
    CommHalo<real_t> comm_halo(dd, nghost, &timers);
    Grid3d<real_t> x_k(dd, nghost, ninner); x_k.all_cells.fill(0.0);
    Grid3d<real_t> v_k(dd, 0, ninner); v_k.all_cells.fill(0.0);
    Grid3d<real_t> t_k(dd, 0, ninner); t_k.all_cells.fill(0.0);
    Grid3d<real_t> q_k(dd, 0, ninner); q_k.all_cells.fill(0.0);
    Grid3d<real_t> D(dd, 0, ninner); D.all_cells.fill(0.0);
    Grid3d<real_t> precond(dd, 0, ninner);
    precond.all_cells.fill(1.0 / stencil.a);
    SeededRademacher rng(dd, args.seed);

    solver_t solver(dd, nghost, args.mode, &timers);
    typename solver_t::result_t outcome;
    int final_iter;
    real_t final_err;
    SliceOps<real_t> slice_ops(&timers);
    std::vector<int> iters(args.samples, 0), codes(args.samples, -1);
    std::vector<double> errs(args.samples, 0.0);

    MPI_Barrier(MPI_COMM_WORLD);
    timers.start(Timers::TIMER_GLOBAL);
    int maxi = 0, mini = args.maxiter + 1, done = 0;
    for (int k = 0; k < args.samples; k++) {
        if (dd.myrank == 0) printf("Solving for sample %d/%d...\n", k + 1, args.samples);
        timers.start(Timers::TIMER_COMPUTE_RNG);
        rng.fill_plusminusone(v_k.rank_cells);
        timers.stop(Timers::TIMER_COMPUTE_RNG);
        x_k.all_cells.fill(0.0);
        outcome = solver.solve(stencil, comm_halo, x_k, v_k, precond,
                               args.maxiter, args.tol, &final_err, &final_iter);
        iters[k] = final_iter; errs[k] = final_err; codes[k] = (int)outcome;
        done = k + 1;
        if (!args.ignore && outcome != 0) {
            fprintf(stderr, "Incomplete convergence in solver, errcode=%d\n", (int)outcome);
            fflush(stdout);
            MPI_Abort(MPI_COMM_WORLD, 100 + (int)outcome);
        }
        if (final_iter > maxi) maxi = final_iter;
        if (final_iter < mini) mini = final_iter;
        slice_ops.accumulate_product(x_k.rank_cells, v_k.rank_cells, t_k.rank_cells);
        slice_ops.accumulate_product(v_k.rank_cells, v_k.rank_cells, q_k.rank_cells);
    }
    MPI_Barrier(MPI_COMM_WORLD);
    timers.stop(Timers::TIMER_GLOBAL);
    slice_ops.divide(t_k.rank_cells, q_k.rank_cells, D.rank_cells);
    if (dd.myrank == 0) {
        printf("Total solve time: %.4e seconds, iteration count min=%d max=%d\n",
               timers.total_time[Timers::TIMER_GLOBAL], mini, maxi);
        timers.print_timers();
    }

    if (args.out) {
        // Gather D into global i1-fastest order on rank 0.
        std::vector<double> glob(dd.ng, 0.0), mine(dd.ng, 0.0);
        int p1, p2, p3;
        dd.rank_1d_to_3d(dd.myrank, &p1, &p2, &p3);
        for (long l3 = 0; l3 < dd.ngp3; l3++)
            for (long l2 = 0; l2 < dd.ngp2; l2++)
                for (long l1 = 0; l1 < dd.ngp1; l1++) {
                    const long g = (l1 + p1 * dd.ngp1) + dd.ng1 * ((l2 + p2 * dd.ngp2) + dd.ng2 * (l3 + p3 * dd.ngp3));
                    mine[g] = D.rank_cells.element_ref(l1, l2, l3);
                }
        MPI_Reduce(mine.data(), glob.data(), (int)dd.ng, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
        if (dd.myrank == 0) {
            FILE *f = fopen(args.out, "wb");
            if (!f) { perror("fopen"); MPI_Abort(MPI_COMM_WORLD, 2); }
            const int64_t hdr[5] = {dd.ng1, dd.ng2, dd.ng3, (int64_t)done, (int64_t)args.solver};
            fwrite(hdr, sizeof(int64_t), 5, f);
            fwrite(iters.data(), sizeof(int), done, f);
            fwrite(codes.data(), sizeof(int), done, f);
            fwrite(errs.data(), sizeof(double), done, f);
            fwrite(glob.data(), sizeof(double), dd.ng, f);
            fclose(f);
        }
    }
    return 0;
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    Args args = parse(argc, argv);
    if (!args.no_fpe) feenableexcept(FE_INVALID | FE_OVERFLOW | FE_DIVBYZERO);
    Timers timers;
    DomainDecomposition dd(args.ng1, args.ng2, args.ng3, args.np1, args.np2, args.np3);
    int rc = 0;
    if (args.solver == 1) rc = run<BiCgStab<real_t> >(args, dd, timers);
    else if (args.solver == 2) rc = run<Cg<real_t> >(args, dd, timers);
    else rc = run<PipelinedCg<real_t> >(args, dd, timers);
    MPI_Finalize();
    return rc;
}
