// ref_cases.cc -- one solve of intel/sedi's own solver code on a given system.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target ref_cases),
// in the development container only, against the unmodified reference headers
// into oracle/_ref/sedi_ref_cases.  It solves A x = b once, from x = 0, with
// Stencil<double>(a, b, c) (Stencil.hpp:18-19), the Jacobi vector 1/a (as
// stochastic-estimator.cc:102-103) and the solver chosen by -S (the SOLVER
// numbering of stochastic-estimator.cc:134-142), on one rank (-p 1x1x1), with
// FP exceptions trapped as stochastic-estimator.cc:64-69 does, and the mode
// bits of -f (stochastic-estimator.cc:139; on grids with a side of 1 or 2
// cells the asynchronous stencil's inner region is empty, which its cache
// blocking refuses, so such grids take -f 1: the synchronous stencil applies
// the same expression to every cell at once).  It prints the
// solver's own per-iteration lines and then one line
//     result code=<result_t> iter=<final_iter> err=<final_err %.17g>
// and the solution ("x v0 v1 ...", i1 fastest, %.17g),
// so the oracle's outcome logic (e.g. BiCGStab's breakdown codes and their
// order, BiCgStab.hpp:235-263) can be checked on systems built to reach it.
//
// Usage: sedi_ref_cases -g N1xN2xN3 -c A,B,C -b v0,v1,... [-e TOL] [-m MAXITER] [-S SOLVER] [-f MODE]
//        (b lists the n1*n2*n3 right-hand-side values, i1 fastest)
#include <mpi.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

static void usage(const char *prog) {
    fprintf(stderr, "Usage: %s -g N1xN2xN3 -c A,B,C -b v0,v1,... [-e TOL] [-m MAXITER] [-S SOLVER] [-f MODE]\n",
            prog);
    exit(1);
}

template <class solver_t>
static int run(DomainDecomposition &dd, Timers &timers, double a, double bb, double cc,
               const std::vector<double> &rhs, int maxiter, double tol, int mode) {
    Stencil<double> stencil(a, bb, cc);
    const long nghost = 1;
    CommHalo<double> comm_halo(dd, nghost, &timers);
    Grid3d<double> x(dd, nghost, nghost); x.all_cells.fill(0.0);
    Grid3d<double> b(dd, 0, nghost); b.all_cells.fill(0.0);
    Grid3d<double> iM(dd, 0, nghost); iM.all_cells.fill(1.0 / a);
    for (long i3 = 0; i3 < dd.ngp3; i3++)
        for (long i2 = 0; i2 < dd.ngp2; i2++)
            for (long i1 = 0; i1 < dd.ngp1; i1++)
                b.rank_cells.element_ref(i1, i2, i3) = rhs[i1 + dd.ng1 * (i2 + dd.ng2 * i3)];
    solver_t solver(dd, nghost, mode, &timers);
    int final_iter = -1;
    double final_err = -1.0;
    const int code = (int)solver.solve(stencil, comm_halo, x, b, iM, maxiter, tol, &final_err, &final_iter);
    printf("result code=%d iter=%d err=%.17g\n", code, final_iter, final_err);
    printf("x");
    for (long i3 = 0; i3 < dd.ngp3; i3++)
        for (long i2 = 0; i2 < dd.ngp2; i2++)
            for (long i1 = 0; i1 < dd.ngp1; i1++) printf(" %.17g", x.rank_cells.element_ref(i1, i2, i3));
    printf("\n");
    return 0;
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    long n1 = 0, n2 = 0, n3 = 0;
    double a = 6, bb = 1, cc = 1, tol = 1e-9;
    int maxiter = 100, solver = 1, mode = 0;
    std::vector<double> rhs;
    int opt;
    while ((opt = getopt(argc, argv, "g:c:b:e:m:S:f:")) != -1) {
        switch (opt) {
        case 'g': if (sscanf(optarg, "%ldx%ldx%ld", &n1, &n2, &n3) != 3) usage(argv[0]); break;
        case 'c': if (sscanf(optarg, "%lf,%lf,%lf", &a, &bb, &cc) != 3) usage(argv[0]); break;
        case 'b': {
            char *s = optarg;
            while (*s) {
                char *end = nullptr;
                rhs.push_back(strtod(s, &end));
                if (end == s) usage(argv[0]);
                s = (*end == ',') ? end + 1 : end;
            }
            break;
        }
        case 'e': tol = atof(optarg); break;
        case 'm': maxiter = atoi(optarg); break;
        case 'S': solver = atoi(optarg); break;
        case 'f': mode = atoi(optarg); break;
        default: usage(argv[0]);
        }
    }
    if (n1 <= 0 || n2 <= 0 || n3 <= 0 || (long)rhs.size() != n1 * n2 * n3 || solver < 1 || solver > 3) usage(argv[0]);
    feenableexcept(FE_INVALID | FE_OVERFLOW | FE_DIVBYZERO);
    Timers timers;
    DomainDecomposition dd(n1, n2, n3, 1, 1, 1);
    if (solver == 1) run<BiCgStab<double> >(dd, timers, a, bb, cc, rhs, maxiter, tol, mode);
    else if (solver == 2) run<Cg<double> >(dd, timers, a, bb, cc, rhs, maxiter, tol, mode);
    else run<PipelinedCg<double> >(dd, timers, a, bb, cc, rhs, maxiter, tol, mode);
    MPI_Finalize();
    return 0;
}
