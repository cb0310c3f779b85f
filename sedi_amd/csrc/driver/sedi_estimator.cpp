// sedi-estimator -- command-line drop-in for intel/sedi's stochastic-estimator.
//
// Same flags, defaults and output lines as stochastic-estimator.cc:57-232 with
// CmdLineParser.hpp:26-122:
//   -g NxNxN  global grid (required)        -p NxNxN  decomposition (required; on
//   -s N      samples (400)                           one GPU it is logical and
//   -e TOL    tolerance (1e-4)                        must divide -g, as
//   -m N      maxiter (1000)                          DomainDecomposition.hpp:34-36)
//   -i        ignore solver errors          -f MODE   solver mode bits (accepted)
//   -F        no FP exceptions (accepted; the GPU does not trap)
//   -V FILE / -E FILE  HDF5 validation archives
// Additions (the reference fixes these at compile time or not at all):
//   -S K      solver: 1 BiCGStab, 2 CG, 3 PipelinedCg (the SOLVER numbers; default 2)
//   -r SEED   mt19937 seed of the probe stream (default 1234)
//   -P PREC   fp64 (default) or mixed
//   -k K      probes solved together (default 1)
//   -R M      mixed precision: FP64 residual replacement every M iterations (default 50)
//   -o FILE   write D = t./q as raw little-endian float64, i1 fastest
// Exit status: 0, 1 on usage errors, 100+code when a solve fails without -i
// (the MPI_Abort code of stochastic-estimator.cc:183), 2 on a runtime error.
#include <sys/resource.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sedi_c.h"

namespace {

struct Options {
    long ng1 = 0, ng2 = 0, ng3 = 0;
    int np1 = 0, np2 = 0, np3 = 0;
    int samples = 400;
    double tol = 1e-4;
    int maxiter = 1000;
    int mode = 0;
    bool ignore = false, no_fpe = false;
    const char *vfile = nullptr, *efile = nullptr, *out = nullptr;
    int solver = SEDI_CG;
    uint32_t seed = 1234;
    int precision = SEDI_FP64;
    int batch = 1;
    int rep_every = 0;
};

const char* progname = "sedi-estimator";

[[noreturn]] void usage() {
    fprintf(stderr, "Usage: %s -g NxNxN -p NxNxN [-s NSAMPLES] [-e TOL] [-m MAXITER] [-i] [-f MODE] [-F]"
                    " [-V HDF5_FILE] [-E HDF5_FILE] [-S SOLVER] [-r SEED] [-P fp64|mixed] [-k K] [-R M] [-o FILE]\n",
            progname);
    fprintf(stderr, "\t-g NxNxN : size of the global 3D grid\n");
    fprintf(stderr, "\t-p NxNxN : decomposition into ranks (logical on one GPU)\n");
    fprintf(stderr, "\t-s NSAMPLES : number of samples for stochastic estimation\n");
    fprintf(stderr, "\t-e TOL : solver tolerance\n");
    fprintf(stderr, "\t-m MAXITER : solver maximum iterations\n");
    fprintf(stderr, "\t-i : ignore solver convergence errors and keep going\n");
    fprintf(stderr, "\t-f : [advanced] flag for solver operation mode\n");
    fprintf(stderr, "\t-F : don't enable floating point exceptions\n");
    fprintf(stderr, "\t-V HDF5_FILE : enable solver validation mode against HDF5 archive\n");
    fprintf(stderr, "\t-E HDF5_FILE : enable estimator validation mode against HDF5 archive\n");
    fprintf(stderr, "\t-S SOLVER : 1 BiCGStab, 2 CG (default), 3 PipelinedCg\n");
    fprintf(stderr, "\t-r SEED : seed of the mt19937 probe stream (default 1234)\n");
    fprintf(stderr, "\t-P PREC : fp64 (default) or mixed (FP32 vectors, FP64 reductions)\n");
    fprintf(stderr, "\t-k K : probes solved together (default 1)\n");
    fprintf(stderr, "\t-R M : mixed precision: FP64 residual replacement every M iterations (default 50)\n");
    fprintf(stderr, "\t-o FILE : write the estimated diagonal as raw float64\n");
    exit(1);
}

Options parse(int argc, char** argv) {
    Options o;
    const char *g = nullptr, *p = nullptr;
    int c;
    while ((c = getopt(argc, argv, "g:p:s:e:m:if:FV:E:S:r:P:k:R:o:")) != -1) {
        switch (c) {
        case 'g': g = optarg; break;
        case 'p': p = optarg; break;
        case 's': if (sscanf(optarg, "%d", &o.samples) != 1) usage(); break;
        case 'e': if (sscanf(optarg, "%lf", &o.tol) != 1) usage(); break;
        case 'm': if (sscanf(optarg, "%d", &o.maxiter) != 1) usage(); break;
        case 'i': o.ignore = true; break;
        case 'f': if (sscanf(optarg, "%d", &o.mode) != 1) usage(); break;
        case 'F': o.no_fpe = true; break;
        case 'V': o.vfile = optarg; break;
        case 'E': o.efile = optarg; break;
        case 'S': if (sscanf(optarg, "%d", &o.solver) != 1 || o.solver < 1 || o.solver > 3) usage(); break;
        case 'r': o.seed = (uint32_t)strtoul(optarg, nullptr, 10); break;
        case 'P':
            if (!strcmp(optarg, "fp64")) o.precision = SEDI_FP64;
            else if (!strcmp(optarg, "mixed")) o.precision = SEDI_FP32_MIXED;
            else usage();
            break;
        case 'k': if (sscanf(optarg, "%d", &o.batch) != 1 || o.batch < 1) usage(); break;
        case 'R': if (sscanf(optarg, "%d", &o.rep_every) != 1 || o.rep_every < 0) usage(); break;
        case 'o': o.out = optarg; break;
        default: usage();
        }
    }
    if (!g || sscanf(g, "%ldx%ldx%ld", &o.ng1, &o.ng2, &o.ng3) != 3) usage();
    if (!p || sscanf(p, "%dx%dx%d", &o.np1, &o.np2, &o.np3) != 3) usage();
    return o;
}

const char* solver_name(int k) {
    return k == SEDI_CG ? "Cg" : (k == SEDI_PCG ? "PipelinedCg" : "BiCgStab");
}

void print_line(void*, const char* line) {
    fputs(line, stdout);
    fputc('\n', stdout);
}

}  // namespace

int main(int argc, char** argv) {
    progname = argv[0];
    Options o = parse(argc, argv);
    if (o.ng1 <= 0 || o.ng2 <= 0 || o.ng3 <= 0 || o.np1 <= 0 || o.np2 <= 0 || o.np3 <= 0 ||
        o.ng1 % o.np1 || o.ng2 % o.np2 || o.ng3 % o.np3) {
        fprintf(stderr, "%s: -p %dx%dx%d must divide -g %ldx%ldx%ld\n", progname, o.np1, o.np2, o.np3,
                o.ng1, o.ng2, o.ng3);
        return 1;
    }
    printf("Global grid shape is (%ld, %ld, %ld)\n", o.ng1, o.ng2, o.ng3);
    printf("Domain decomposition across 1 GPU (logical -p %dx%dx%d)\n", o.np1, o.np2, o.np3);
    printf("Process block shape is (%ld, %ld, %ld)\n", o.ng1 / o.np1, o.ng2 / o.np2, o.ng3 / o.np3);
    if (o.vfile || o.efile) {
        printf("WARNING: validation mode selected, but compiled without HDF5 support: disabling!\n");
        o.vfile = o.efile = nullptr;
    }

    sedi_estimate_args a;
    memset(&a, 0, sizeof(a));
    a.grid = sedi_grid{o.ng1, o.ng2, o.ng3};
    a.stencil = sedi_stencil{6.0, 1.0, 1.0};  // stochastic-estimator.cc:74
    a.kind = o.solver;
    a.mode = o.mode;
    a.samples = o.samples;
    a.maxiter = o.maxiter;
    a.tol = o.tol;
    a.ignore_errors = o.ignore ? 1 : 0;
    a.seed = o.seed;
    a.precision = o.precision;
    a.batch_k = o.batch;
    a.rep_every = o.rep_every;

    sedi_estimator* est = nullptr;
    int rc = sedi_estimator_create(&a, &est);
    if (rc < 0) {
        fprintf(stderr, "%s: %s\n", progname, sedi_last_error());
        return 2;
    }
    printf("Initialized %s solver, sync_dotprod=%d, sync_stencil=%d\n", solver_name(o.solver),
           (o.mode & 2) != 0, (o.mode & 1) != 0);
    sedi_estimator_set_log(est, print_line, nullptr);
    const long n = o.ng1 * o.ng2 * o.ng3;
    std::vector<double> D(o.out ? n : 0);
    std::vector<int> iters(o.samples), codes(o.samples, -1);
    std::vector<double> errs(o.samples);
    double secs = 0.0;
    rc = sedi_estimator_run(est, o.out ? D.data() : nullptr, iters.data(), errs.data(), codes.data(), &secs);
    if (rc < 0) {
        fprintf(stderr, "%s: %s\n", progname, sedi_last_error());
        sedi_estimator_destroy(est);
        return 2;
    }
    fflush(stdout);
    if (rc >= 100) {  // stochastic-estimator.cc:181-184
        fprintf(stderr, "Incomplete convergence in solver, errcode=%d\n", rc - 100);
        sedi_estimator_destroy(est);
        return rc;
    }
    int mn = o.maxiter + 1, mx = 0;
    for (int k = 0; k < o.samples; k++) {
        if (iters[k] > mx) mx = iters[k];
        if (iters[k] < mn) mn = iters[k];
    }
    printf("Total solve time: %.4e seconds, iteration count min=%d max=%d\n", secs, mn, mx);
    printf("===========================================\n");
    printf("TIMER_GLOBAL                 = %.4e s\n", secs);
    printf("===========================================\n");
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    const double rss = ru.ru_maxrss / 1024.0;
    printf("Resident set size across ranks [MB]: avg=%.3e, tot=%.3e, min=%.3e, max=%.3e\n", rss, rss, rss, rss);
    if (o.out) {
        FILE* f = fopen(o.out, "wb");
        if (!f || fwrite(D.data(), sizeof(double), n, f) != (size_t)n) {
            fprintf(stderr, "%s: cannot write %s\n", progname, o.out);
            if (f) fclose(f);
            sedi_estimator_destroy(est);
            return 2;
        }
        fclose(f);
    }
    sedi_estimator_destroy(est);
    return 0;
}
