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
//   -V FILE / -E FILE  HDF5 validation archives (ValidateSolver.hpp,
//             ValidateEstimator.hpp; libhdf5 is loaded at run time)
// Additions (the reference fixes these at compile time or not at all):
//   -S K      solver: 1 BiCGStab, 2 CG, 3 PipelinedCg (the SOLVER numbers; default 2)
//   -r SEED   mt19937 seed of the probe stream (default 1234)
//   -P PREC   fp64 (default) or mixed
//   -k K      probes solved together (default 1)
//   -R M      mixed precision: FP64 residual replacement every M iterations (default 50)
//   -o FILE   write D = t./q as raw little-endian float64, i1 fastest
//   -X FILE   write an HDF5 validation archive of this run: x_k and v_k of the
//             last sample, diag_inv = D (the datasets -V and -E read)
// Exit status: 0, 1 on usage errors, 100+code when a solve fails without -i
// (the MPI_Abort code of stochastic-estimator.cc:183), 2 on a runtime error.
#include <sys/resource.h>
#include <unistd.h>

#include <cmath>
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
    const char* xfile = nullptr;
};

const char* progname = "sedi-estimator";

[[noreturn]] void usage() {
    fprintf(stderr, "Usage: %s -g NxNxN -p NxNxN [-s NSAMPLES] [-e TOL] [-m MAXITER] [-i] [-f MODE] [-F]"
                    " [-V HDF5_FILE] [-E HDF5_FILE] [-S SOLVER] [-r SEED] [-P fp64|mixed] [-k K] [-R M] [-o FILE] [-X HDF5_FILE]\n",
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
    fprintf(stderr, "\t-k K : probes solved together (default 1; 0 = the autotuned batch for this grid)\n");
    fprintf(stderr, "\t-R M : mixed precision: FP64 residual replacement every M iterations (default 50)\n");
    fprintf(stderr, "\t-o FILE : write the diagonal estimate as raw float64\n");
    fprintf(stderr, "\t-X HDF5_FILE : write a validation archive (x_k, v_k of the last sample, diag_inv)\n");
    exit(1);
}

Options parse(int argc, char** argv) {
    Options o;
    const char *g = nullptr, *p = nullptr;
    int c;
    while ((c = getopt(argc, argv, "g:p:s:e:m:if:FV:E:S:r:P:k:R:o:X:")) != -1) {
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
        case 'k': if (sscanf(optarg, "%d", &o.batch) != 1 || o.batch < 0) usage(); break;
        case 'R': if (sscanf(optarg, "%d", &o.rep_every) != 1 || o.rep_every < 0) usage(); break;
        case 'o': o.out = optarg; break;
        case 'X': o.xfile = optarg; break;
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

// ---- validation (ValidateSolver.hpp:30-113, ValidateEstimator.hpp:27-87) ----
struct Validation {
    long n = 0;
    bool solver = false, estimator = false, archive = false;
    std::vector<double> ref_x, ref_v, ref_D;  // i1 fastest
    double l2_sol = 0.0, linf_sol = 0.0;
    std::vector<double> last_x, last_v;
};

// Norms of the reference solution (ValidateSolver.hpp:62-73).
void solution_norms(Validation& v) {
    double l2 = 0.0, linf = 0.0;
    for (long i = 0; i < v.n; i++) {
        const double x = v.ref_x[i];
        l2 += x * x;
        if (fabs(x) > linf) linf = fabs(x);
    }
    v.l2_sol = sqrt(l2 / v.n);
    v.linf_sol = linf;
}

// sqrt(sum delta^2 / n) and max |delta| (ValidateSolver.hpp:99-111,
// ValidateEstimator.hpp:72-84; one rank, so the allreduces are the identity).
void errors(const double* a, const double* b, long n, double* l2, double* linf) {
    double e2 = 0.0, ei = 0.0;
    for (long i = 0; i < n; i++) {
        const double d = a[i] - b[i];
        e2 += d * d;
        if (fabs(d) > ei) ei = fabs(d);
    }
    *l2 = sqrt(e2 / n);
    *linf = ei;
}

void on_sample(void* user, int, const double* x, const double* v, const double* D) {
    Validation& val = *static_cast<Validation*>(user);
    double l2, linf;
    if (val.solver && x) {
        printf("VALIDATION: comparing to ref solution from validation dataset\n");
        errors(x, val.ref_x.data(), val.n, &l2, &linf);
        printf("VALIDATION: rel L2 error=%.4e, rel Linf error=%.4e\n", l2 / val.l2_sol, linf / val.linf_sol);
    }
    if (val.estimator && D) {
        errors(D, val.ref_D.data(), val.n, &l2, &linf);
        printf("VALIDATION_ESTIMATOR: L2 error=%.4e, Linf error=%.4e\n", l2, linf);
    }
    if (val.archive) {
        if (x) val.last_x.assign(x, x + val.n);
        if (v) val.last_v.assign(v, v + val.n);
    }
}

void print_line(void* user, const char* line) {
    fputs(line, stdout);
    fputc('\n', stdout);
    const Validation* val = static_cast<const Validation*>(user);
    if (val && val->solver && !strncmp(line, "Solving for sample", 18))
        printf("VALIDATION: filling RHS from validation dataset\n");  // ValidateSolver.hpp:94
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
    const long n = o.ng1 * o.ng2 * o.ng3;
    const sedi_grid grid{o.ng1, o.ng2, o.ng3};
    Validation val;
    val.n = n;
    if ((o.vfile || o.efile || o.xfile) && sedi_h5_available() != 1) {
        // stochastic-estimator.cc:119-124
        printf("WARNING: validation mode selected, but compiled without HDF5 support: disabling!\n");
        o.vfile = o.efile = o.xfile = nullptr;
    }
    if (o.vfile) {  // ValidateSolver.hpp:30-85
        val.ref_x.resize(n);
        val.ref_v.resize(n);
        if (sedi_h5_read(o.vfile, "x_k", &grid, val.ref_x.data()) < 0 ||
            sedi_h5_read(o.vfile, "v_k", &grid, val.ref_v.data()) < 0) {
            fprintf(stderr, "%s: %s\n", progname, sedi_last_error());
            return 2;
        }
        solution_norms(val);
        printf("VALIDATION: rank %d read arrays from %s\n", 0, o.vfile);
        val.solver = true;
    }
    if (o.efile) {  // ValidateEstimator.hpp:27-62
        val.ref_D.resize(n);
        if (sedi_h5_read(o.efile, "diag_inv", &grid, val.ref_D.data()) < 0) {
            fprintf(stderr, "%s: %s\n", progname, sedi_last_error());
            return 2;
        }
        printf("VALIDATION_ESTIMATOR: rank %d read arrays from %s\n", 0, o.efile);
        val.estimator = true;
    }
    if (val.solver && val.estimator) {  // stochastic-estimator.cc:127-131
        printf("WARNING: cannot enable both solver and estimator validation mode: disabling both!\n");
        val.solver = val.estimator = false;
    }
    val.archive = o.xfile != nullptr;

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
    sedi_estimator_set_log(est, print_line, &val);
    if (val.solver) rc = sedi_estimator_set_rhs(est, val.ref_v.data());
    const int want = (val.solver ? SEDI_HOOK_X : 0) | (val.estimator ? SEDI_HOOK_D : 0) |
                     (val.archive ? SEDI_HOOK_X | SEDI_HOOK_V : 0);
    if (rc >= 0 && want) rc = sedi_estimator_set_sample_hook(est, on_sample, want, &val);
    if (rc < 0) {
        fprintf(stderr, "%s: %s\n", progname, sedi_last_error());
        sedi_estimator_destroy(est);
        return 2;
    }
    const bool want_D = o.out || val.archive;
    std::vector<double> D(want_D ? n : 0);
    std::vector<int> iters(o.samples), codes(o.samples, -1);
    std::vector<double> errs(o.samples);
    double secs = 0.0;
    rc = sedi_estimator_run(est, want_D ? D.data() : nullptr, iters.data(), errs.data(), codes.data(), &secs);
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
    if (val.archive && o.samples > 0) {
        const char* names[3] = {"x_k", "v_k", "diag_inv"};
        const double* arrays[3] = {val.last_x.data(), val.last_v.data(), D.data()};
        if (sedi_h5_write(o.xfile, 3, names, arrays, &grid) < 0) {
            fprintf(stderr, "%s: %s\n", progname, sedi_last_error());
            sedi_estimator_destroy(est);
            return 2;
        }
    }
    sedi_estimator_destroy(est);
    return 0;
}
