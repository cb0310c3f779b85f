/*
 * sedi_oracle.h -- CPU restatement of intel/sedi's stochastic diag(A^-1) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libsedi_amd.so, the sedi-estimator
 * driver) never links or calls it.
 *
 * Build: oracle/Makefile, plain C99 with -ffp-contract=off so that every
 * a*x - c*y is two roundings, matching the HIP kernels (built the same way).
 *
 * Layout everywhere: dense n1*n2*n3 arrays, i1 fastest (Slice3d.hpp:30-32),
 * no ghost cells; cells outside the grid read as 0 (the zero ghosts of
 * Cg.hpp:57-62).
 */
#ifndef SEDI_ORACLE_H
#define SEDI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- probe RNG (DistrRngStd.hpp:22-42 + the seeded-stream convention) ---- */
typedef struct oracle_mt19937 {
    uint32_t mt[624];
    int idx;
} oracle_mt19937;

void     oracle_mt_seed(oracle_mt19937 *s, uint32_t seed);
uint32_t oracle_mt_next(oracle_mt19937 *s);
/* Continue from 624 given state words (oldest first); the next output is the
 * tempered word generated after them. */
void     oracle_mt_from_words(oracle_mt19937 *s, const uint32_t *words);
/* One std::uniform_real_distribution<double>(0,1) draw (libstdc++
 * generate_canonical<double,53>: two 32-bit outputs). */
double   oracle_uniform01(oracle_mt19937 *s);
/* Fill v[0..ncells) with +-1 from the stream mt19937(seed) after skipping
 * first_draw 32-bit outputs. Cell c consumes two outputs. */
void     oracle_rademacher_fill(uint32_t seed, uint64_t first_draw,
                                int64_t ncells, double *v);

/* ---- kernels (Stencil.hpp:38-54, SliceOps.hpp:111-125) ---- */
void   oracle_stencil_apply(int64_t n1, int64_t n2, int64_t n3,
                            double a, double b, double c,
                            const double *x, double *y);
/* Dot product in the reference's order for one or several fused dots
 * (DistrAsyncDotProds.hpp:73-98 + BlockSliceIterator.hpp:81-109). */
void   oracle_dots(int64_t n1, int64_t n2, int64_t n3, int ndot,
                   const double *const *xs, const double *const *ys,
                   double *out);

/* ---- solvers ---- */
enum { ORACLE_BICGSTAB = 1, ORACLE_CG = 2, ORACLE_PCG = 3 };

/* Solve A x = b from the initial guess in x.  iM is the inverse-diagonal
 * preconditioner (NULL => the constant 1/a).  err_hist (may be NULL) gets the
 * err printed after each iteration (length >= maxiter).  Returns the
 * reference's result code (0..3). */
int oracle_solve(int kind, int64_t n1, int64_t n2, int64_t n3,
                 double a, double b, double c,
                 const double *rhs, double *x, const double *iM,
                 int maxiter, double tol,
                 double *final_err, int *final_iter, double *err_hist);

/* The sample loop of stochastic-estimator.cc:158-216.  Probe k uses draws
 * [2Nk, 2N(k+1)) of mt19937(seed).  Returns 0, or 100+code of the first
 * failing solve when ignore_errors == 0 (the MPI_Abort code of :183). */
int oracle_estimate(int kind, int64_t n1, int64_t n2, int64_t n3,
                    double a, double b, double c,
                    uint32_t seed, int samples, int maxiter, double tol,
                    int ignore_errors,
                    double *D, int *iters, double *errs, int *codes);

/* ---- closed forms used to pin the above ---- */
/* Exact diag(A^-1) for the symmetric operator (b == c) at the listed cells,
 * by the separable sine eigen-decomposition.  cells[i] are linear indices. */
void oracle_exact_diag_inv(int64_t n1, int64_t n2, int64_t n3,
                           double a, double bc,
                           int64_t ncells, const int64_t *cells, double *out);

#ifdef __cplusplus
}
#endif
#endif
