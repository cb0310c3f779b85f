/*
 * sedi_oracle.c -- CPU restatement of intel/sedi's stochastic diag(A^-1) path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of it.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function cites the reference file:line it restates.  Compiled with
 * -ffp-contract=off (oracle/Makefile): the reference's expressions are
 * evaluated left to right with one rounding per operation.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   oracle_mt_*, oracle_uniform01, oracle_rademacher_fill
 *       - libstdc++ std::mt19937 + uniform_real_distribution<double> run by
 *         tests/golden/rng_golden.cc (the library routine the reference calls);
 *       - the C++ standard's known answer: the 10000th output of a
 *         default-seeded (5489) mt19937 is 4123659995.
 *   oracle_stencil_apply
 *       - A*1 = a - (#in-grid neighbours)*b (b == c), sine eigenvectors with
 *         lambda = a - 2b*sum(cos), and the compiled reference's Stencil::apply
 *         (oracle/_ref) through the solver fixtures.
 *   oracle_dots
 *       - long-double sums (summation order differs only at rounding level).
 *   oracle_solve / oracle_estimate
 *       - fixtures written by the compiled reference (oracle/_ref/sedi_ref,
 *         built from the unmodified reference headers): iteration counts,
 *         per-iteration err lines and D;
 *       - a dense solve at <= 6^3, the one-iteration convergence on an
 *         eigenvector right-hand side, and exact diag(A^-1) in expectation.
 *   oracle_exact_diag_inv
 *       - a dense inverse at n = 4..6 computed in the test, and the known
 *         corner value 0.185577 for n >= 6 (a = 6, b = c = 1).
 */
#include "sedi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* mt19937: Matsumoto & Nishimura, ACM TOMACS 8(1), 1998.  The reference     */
/* calls it as std::mt19937 (DistrRngStd.hpp:27-29).                          */
/* ------------------------------------------------------------------------ */
enum { MT_N = 624, MT_M = 397 };

void oracle_mt_seed(oracle_mt19937 *s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < MT_N; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = MT_N;
}

void oracle_mt_from_words(oracle_mt19937 *s, const uint32_t *words)
{
    memcpy(s->mt, words, sizeof(s->mt));
    s->idx = MT_N;
}

static void mt_regenerate(oracle_mt19937 *s)
{
    for (int i = 0; i < MT_N; i++) {
        uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % MT_N] & 0x7fffffffu);
        s->mt[i] = s->mt[(i + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    s->idx = 0;
}

uint32_t oracle_mt_next(oracle_mt19937 *s)
{
    if (s->idx >= MT_N) mt_regenerate(s);
    uint32_t y = s->mt[s->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* uniform_real_distribution<double>(0,1)(eng) as libstdc++ computes it:
 * generate_canonical<double,53> takes k = 2 outputs, sum = x0 + x1*2^32 in
 * double, ret = sum / 2^64, clamped below 1; then ret*(1-0)+0.
 * DistrRngStd.hpp:32 builds the distribution, :38 draws once per cell. */
double oracle_uniform01(oracle_mt19937 *s)
{
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; k++) {
        sum += (double)oracle_mt_next(s) * tmp;
        tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret * (1.0 - 0.0) + 0.0;
}

/* DistrRngStd::fill_plusminusone (DistrRngStd.hpp:22-42): cells in i3,i2,i1
 * order (i1 fastest), +1 iff the draw is > 0.5.  The reference seeds
 * std::random_device on every call (:27-29); here the stream is
 * mt19937(seed) and the fill starts after first_draw outputs. */
void oracle_rademacher_fill(uint32_t seed, uint64_t first_draw,
                            int64_t ncells, double *v)
{
    oracle_mt19937 s;
    oracle_mt_seed(&s, seed);
    for (uint64_t k = 0; k < first_draw; k++) (void)oracle_mt_next(&s);
    for (int64_t c = 0; c < ncells; c++) {
        double x = oracle_uniform01(&s);
        v[c] = (x > 0.5 ? 1.0 : -1.0);
    }
}

/* ------------------------------------------------------------------------ */
/* Stencil::apply_direct (Stencil.hpp:38-54): y = a*x0 - c*x(-1) per axis ... */
/* - b*x(+1) per axis, left to right, zero outside the grid (Cg.hpp:57-62).   */
/* ------------------------------------------------------------------------ */
void oracle_stencil_apply(int64_t n1, int64_t n2, int64_t n3,
                          double a, double b, double c,
                          const double *x, double *y)
{
    const int64_t s2 = n1, s3 = n1 * n2;
    for (int64_t i3 = 0; i3 < n3; i3++)
        for (int64_t i2 = 0; i2 < n2; i2++)
            for (int64_t i1 = 0; i1 < n1; i1++) {
                const int64_t id = i1 + i2 * s2 + i3 * s3;
                const double xm1 = i1 > 0 ? x[id - 1] : 0.0;
                const double xm2 = i2 > 0 ? x[id - s2] : 0.0;
                const double xm3 = i3 > 0 ? x[id - s3] : 0.0;
                const double xp1 = i1 + 1 < n1 ? x[id + 1] : 0.0;
                const double xp2 = i2 + 1 < n2 ? x[id + s2] : 0.0;
                const double xp3 = i3 + 1 < n3 ? x[id + s3] : 0.0;
                y[id] = a * x[id] - c * xm1 - c * xm2 - c * xm3
                        - b * xp1 - b * xp2 - b * xp3;
            }
}

/* ------------------------------------------------------------------------ */
/* Dot products.  A single dot sums the whole slice row by row               */
/* (SliceOps.hpp:111-125; one block, DistrAsyncDotProds.hpp:73).  Two or     */
/* more fused dots walk cache_aware_iterator(slice, 3 MiB, 2*ndot) blocks    */
/* (DistrAsyncDotProds.hpp:74-78,92-98; BlockSliceIterator.hpp:81-109).      */
/* Within a row the reference's `omp simd reduction` order is the compiler's */
/* choice; here each row is summed left to right.                            */
/* ------------------------------------------------------------------------ */
static void block_sizes(int64_t n1, int64_t n2, int64_t n3, int narrays,
                        int64_t *bs1, int64_t *bs2, int64_t *bs3)
{
    const long cache = 3L * 1024 * 1024;        /* Platform.hpp:5 */
    const float safety = 0.6f;                  /* BlockSliceIterator.hpp:82 */
    const long eff = (long)((float)(cache / narrays) * safety);
    const long nelem = eff / (long)sizeof(double);
    int64_t b1 = n1;
    for (int64_t nb = 1; nb <= n1; nb++) { b1 = n1 / nb; if (b1 < nelem) break; }
    int64_t b2 = 10;
    for (; b2 >= 1; b2--) if (b1 * b2 < nelem) break;
    int64_t b3 = n3;
    for (int64_t nb = 1; nb <= n3; nb++) { b3 = n3 / nb; if (b1 * b2 * b3 < nelem) break; }
    if (b2 < 1) b2 = 1;
    *bs1 = b1 > n1 ? n1 : b1;
    *bs2 = b2 > n2 ? n2 : b2;
    *bs3 = b3 > n3 ? n3 : b3;
}

void oracle_dots(int64_t n1, int64_t n2, int64_t n3, int ndot,
                 const double *const *xs, const double *const *ys, double *out)
{
    int64_t bs1 = n1, bs2 = n2, bs3 = n3;
    if (ndot >= 2) block_sizes(n1, n2, n3, 2 * ndot, &bs1, &bs2, &bs3);
    const int64_t s2 = n1, s3 = n1 * n2;
    for (int d = 0; d < ndot; d++) out[d] = 0.0;
    for (int64_t b3 = 0; b3 < n3; b3 += bs3)
        for (int64_t b2 = 0; b2 < n2; b2 += bs2)
            for (int64_t b1 = 0; b1 < n1; b1 += bs1) {
                const int64_t e1 = b1 + bs1 < n1 ? b1 + bs1 : n1;
                const int64_t e2 = b2 + bs2 < n2 ? b2 + bs2 : n2;
                const int64_t e3 = b3 + bs3 < n3 ? b3 + bs3 : n3;
                for (int d = 0; d < ndot; d++) {
                    double res = 0.0;
                    for (int64_t i3 = b3; i3 < e3; i3++)
                        for (int64_t i2 = b2; i2 < e2; i2++) {
                            /* Two-lane row sum: lane 0 takes even offsets,
                             * lane 1 odd ones; an odd tail joins lane 0; the
                             * lanes are then added to the running sum one at a
                             * time.  This is the order of the SSE2 (2 x f64)
                             * `omp simd reduction` that oracle/Makefile's build
                             * of SliceOps.hpp:117-121 produces with g++ 11. */
                            double l0 = 0.0, l1 = 0.0;
                            const int64_t base = i2 * s2 + i3 * s3;
                            const int64_t len = e1 - b1, pairs = len / 2;
                            const double *xr = xs[d] + base + b1, *yr = ys[d] + base + b1;
                            for (int64_t k = 0; k < pairs; k++) {
                                l0 += xr[2 * k] * yr[2 * k];
                                l1 += xr[2 * k + 1] * yr[2 * k + 1];
                            }
                            if (len & 1) l0 = xr[len - 1] * yr[len - 1] + l0;
                            res += l0;
                            res += l1;
                        }
                    out[d] += res;
                }
            }
}

static double dot1(int64_t n1, int64_t n2, int64_t n3, const double *x, const double *y)
{
    double r;
    oracle_dots(n1, n2, n3, 1, &x, &y, &r);
    return r;
}

/* SliceOps helpers (SliceOps.hpp:17-108), written in the reference's form. */
static void combine2(int64_t n, double a, const double *u, double b, const double *v, double *z)
{ for (int64_t i = 0; i < n; i++) z[i] = a * u[i] + b * v[i]; }
static void combine3(int64_t n, double a, const double *u, double b, const double *v,
                     double c, const double *w, double *z)
{ for (int64_t i = 0; i < n; i++) z[i] = a * u[i] + b * v[i] + c * w[i]; }
static void multiply(int64_t n, const double *u, const double *v, double *z)
{ for (int64_t i = 0; i < n; i++) z[i] = u[i] * v[i]; }

/* ------------------------------------------------------------------------ */
/* Cg::solve (Cg.hpp:70-207): Jacobi-preconditioned CG.                      */
/* ------------------------------------------------------------------------ */
static int cg_solve(int64_t n1, int64_t n2, int64_t n3, double a, double b, double c,
                    const double *rhs, double *x, const double *iM,
                    int maxiter, double tol, double *final_err, int *final_iter,
                    double *hist)
{
    const int64_t n = n1 * n2 * n3;
    double *r = malloc(n * sizeof(double)), *s = malloc(n * sizeof(double));
    double *u = malloc(n * sizeof(double)), *p = malloc(n * sizeof(double));

    double b_norm2 = dot1(n1, n2, n3, rhs, rhs);               /* :112-114 */
    if (b_norm2 == 0.0) b_norm2 = 1.0;                          /* :115 */
    oracle_stencil_apply(n1, n2, n3, a, b, c, x, r);            /* :119-121 */
    combine2(n, 1.0, rhs, -1.0, r, r);                          /* :123 */
    double r_norm2 = dot1(n1, n2, n3, r, r);                    /* :126-128 */
    double err = sqrt(r_norm2 / b_norm2);
    int ret, iter;
    if (err < tol) {                                            /* :130-134 */
        *final_iter = 0; *final_err = err; ret = 0; goto out;
    }
    double alpha = 0.0, beta = 0.0, s_dot_p = 0.0, r_dot_u = 0.0, r_dot_u_new = 0.0;
    multiply(n, r, iM, u);                                      /* :144 */
    memcpy(p, u, n * sizeof(double));                           /* :147 */
    for (iter = 1; iter < maxiter + 1; iter++) {                /* :150 */
        oracle_stencil_apply(n1, n2, n3, a, b, c, p, s);        /* :152-154 */
        if (iter == 1) r_dot_u = dot1(n1, n2, n3, r, u);        /* :157-159 */
        else r_dot_u = r_dot_u_new;                             /* :162 */
        s_dot_p = dot1(n1, n2, n3, s, p);                       /* :164-165 */
        alpha = r_dot_u / s_dot_p;                              /* :166 */
        combine2(n, 1.0, x, alpha, p, x);                       /* :169 */
        combine2(n, 1.0, r, -alpha, s, r);                      /* :172 */
        multiply(n, r, iM, u);                                  /* :175 */
        {
            const double *xs[2] = {r, r}, *ys[2] = {u, r};      /* :91-93 */
            double ru_rr[2];
            oracle_dots(n1, n2, n3, 2, xs, ys, ru_rr);          /* :179-182 */
            r_dot_u_new = ru_rr[0];
            r_norm2 = ru_rr[1];
        }
        beta = r_dot_u_new / r_dot_u;                           /* :185 */
        combine2(n, 1.0, u, beta, p, p);                        /* :188 */
        err = sqrt(r_norm2 / b_norm2);                          /* :191 */
        if (hist) hist[iter - 1] = err;
        if (err <= tol) break;                                  /* :193 */
    }
    if (err <= tol) {                                           /* :197-201 */
        *final_iter = iter; *final_err = err; ret = 0;
    } else {                                                    /* :203-206 */
        *final_iter = maxiter; *final_err = err; ret = 1;
    }
out:
    free(r); free(s); free(u); free(p);
    return ret;
}

/* ------------------------------------------------------------------------ */
/* PipelinedCg::solve (PipelinedCg.hpp:72-243), Ghysels-Vanroose pipelined CG */
/* with the reference's one-iteration lag of the convergence test.           */
/* ------------------------------------------------------------------------ */
static int pcg_solve(int64_t n1, int64_t n2, int64_t n3, double a, double b, double c,
                     const double *rhs, double *x, const double *iM,
                     int maxiter, double tol, double *final_err, int *final_iter,
                     double *hist)
{
    const int64_t n = n1 * n2 * n3;
    double *r = malloc(n * sizeof(double)), *w = malloc(n * sizeof(double));
    double *nn = malloc(n * sizeof(double)), *z = calloc(n, sizeof(double));
    double *q = calloc(n, sizeof(double)), *s = calloc(n, sizeof(double));
    double *p = calloc(n, sizeof(double)), *u = malloc(n * sizeof(double));
    double *m = malloc(n * sizeof(double));

    double b_norm2 = dot1(n1, n2, n3, rhs, rhs);               /* :106-108 */
    if (b_norm2 == 0.0) b_norm2 = 1.0;                          /* :109 */
    oracle_stencil_apply(n1, n2, n3, a, b, c, x, r);            /* :113-115 */
    combine2(n, 1.0, rhs, -1.0, r, r);                          /* :117 */
    double r_norm2 = dot1(n1, n2, n3, r, r);                    /* :120-122 */
    double err = sqrt(r_norm2 / b_norm2);
    int ret, iter;
    if (err < tol) {                                            /* :124-128 */
        *final_iter = 0; *final_err = err; ret = 0; goto out;
    }
    double alpha = 0.0, beta = 0.0, gamma = 0.0, gamma_old = 0.0, delta = 0.0;
    multiply(n, r, iM, u);                                      /* :138 */
    oracle_stencil_apply(n1, n2, n3, a, b, c, u, w);            /* :141-143 */
    /* z, q, s, p = 0 (:146-149): calloc'd above */
    for (iter = 1; iter < maxiter + 1; iter++) {                /* :152 */
        multiply(n, w, iM, m);                                  /* :158 */
        oracle_stencil_apply(n1, n2, n3, a, b, c, m, nn);       /* :161-164,184 */
        gamma_old = gamma;                                      /* :167 */
        {
            const double *xs[3] = {r, w, r}, *ys[3] = {u, u, r};/* :90-92 */
            double gdr[3];
            oracle_dots(n1, n2, n3, 3, xs, ys, gdr);            /* :155,170 */
            gamma = gdr[0]; delta = gdr[1]; r_norm2 = gdr[2];
        }
        if (iter > 1) {                                         /* :175-181 */
            beta = gamma / gamma_old;
            alpha = 1.0 / (delta / gamma - beta / alpha);
        } else {
            beta = 0.0;
            alpha = gamma / delta;
        }
        combine2(n, 1.0, m, beta, q, q);                        /* :200 */
        combine2(n, 1.0, w, beta, s, s);                        /* :203 */
        combine2(n, 1.0, u, beta, p, p);                        /* :206 */
        combine2(n, 1.0, nn, beta, z, z);                       /* :209 */
        combine2(n, 1.0, r, -alpha, s, r);                      /* :212 */
        combine2(n, 1.0, x, alpha, p, x);                       /* :215 */
        combine2(n, 1.0, u, -alpha, q, u);                      /* :218 */
        combine2(n, 1.0, w, -alpha, z, w);                      /* :221 */
        err = sqrt(r_norm2 / b_norm2);                          /* :227 */
        if (hist) hist[iter - 1] = err;
        if (err <= tol) break;                                  /* :229 */
    }
    if (err <= tol) {                                           /* :233-237 */
        *final_iter = iter; *final_err = err; ret = 0;
    } else {                                                    /* :239-242 */
        *final_iter = maxiter; *final_err = err; ret = 1;
    }
out:
    free(r); free(w); free(nn); free(z); free(q); free(s); free(p); free(u); free(m);
    return ret;
}

/* ------------------------------------------------------------------------ */
/* BiCgStab::solve (BiCgStab.hpp:77-264): right-preconditioned BiCGStab with */
/* the absolute ||s||^2 < tol^2 early exit and breakdown codes 2 and 3.      */
/* ------------------------------------------------------------------------ */
static int bicgstab_solve(int64_t n1, int64_t n2, int64_t n3, double a, double b, double c,
                          const double *rhs, double *x, const double *iM,
                          int maxiter, double tol, double *final_err, int *final_iter,
                          double *hist)
{
    const int64_t n = n1 * n2 * n3;
    double *r = malloc(n * sizeof(double)), *rh = malloc(n * sizeof(double));
    double *p = calloc(n, sizeof(double)), *s = calloc(n, sizeof(double));
    double *v = calloc(n, sizeof(double)), *t = calloc(n, sizeof(double));
    double *ph = calloc(n, sizeof(double)), *sh = calloc(n, sizeof(double));

    double b_norm2 = dot1(n1, n2, n3, rhs, rhs);               /* :124-126 */
    if (b_norm2 == 0.0) b_norm2 = 1.0;                          /* :127 */
    oracle_stencil_apply(n1, n2, n3, a, b, c, x, r);            /* :131-133 */
    combine2(n, 1.0, rhs, -1.0, r, r);                          /* :135 */
    double r_norm2 = dot1(n1, n2, n3, r, r);                    /* :138-140 */
    double err = sqrt(r_norm2 / b_norm2);
    int ret, iter;
    if (err < tol) {                                            /* :142-146 */
        *final_iter = 0; *final_err = err; ret = 0; goto out;
    }
    double omega = 1.0, alpha = 0.0, rho_prev = 0.0, rho = 0.0; /* :148-151 */
    memcpy(rh, r, n * sizeof(double));                          /* :154 */
    double s_norm2 = 0.0;                                       /* :157 */
    for (iter = 1; iter < maxiter + 1; iter++) {                /* :158 */
        rho = dot1(n1, n2, n3, r, rh);                          /* :160-161 */
        if (rho == 0.0) break;                                  /* :163 */
        if (iter > 1) {
            double beta = (rho / rho_prev) * (alpha / omega);   /* :167 */
            combine3(n, 1.0, r, beta, p, -beta * omega, v, p);  /* :169 */
        } else {
            memcpy(p, r, n * sizeof(double));                   /* :172 */
        }
        multiply(n, p, iM, ph);                                 /* :176 */
        oracle_stencil_apply(n1, n2, n3, a, b, c, ph, v);       /* :179-181 */
        double v_dot_rhat = dot1(n1, n2, n3, v, rh);            /* :184-186 */
        alpha = rho / v_dot_rhat;                               /* :187 */
        combine2(n, 1.0, r, -alpha, v, s);                      /* :190 */
        s_norm2 = dot1(n1, n2, n3, s, s);                       /* :193-194 */
        if (s_norm2 < tol * tol) {                              /* :195 */
            combine2(n, 1.0, x, alpha, ph, x);                  /* :197 */
            break;
        }
        multiply(n, s, iM, sh);                                 /* :202 */
        oracle_stencil_apply(n1, n2, n3, a, b, c, sh, t);       /* :205-207 */
        {
            const double *xs[2] = {t, t}, *ys[2] = {s, t};      /* :98-100 */
            double ts_tt[2];
            oracle_dots(n1, n2, n3, 2, xs, ys, ts_tt);          /* :210-212 */
            omega = ts_tt[0] / ts_tt[1];                        /* :213 */
        }
        combine3(n, 1.0, x, alpha, ph, omega, sh, x);           /* :217 */
        combine2(n, 1.0, s, -omega, t, r);                      /* :220 */
        r_norm2 = dot1(n1, n2, n3, r, r);                       /* :223-224 */
        err = sqrt(r_norm2 / b_norm2);                          /* :227 */
        if (hist) hist[iter - 1] = err;
        if (err <= tol) break;                                  /* :229 */
        if (omega == 0.0) break;                                /* :230 */
        rho_prev = rho;                                         /* :232 */
    }
    if (err <= tol) {                                           /* :236-240 */
        *final_iter = iter; *final_err = err; ret = 0;
    } else if (s_norm2 <= tol * tol) {                          /* :242-246 */
        *final_iter = iter; *final_err = sqrt(s_norm2 / b_norm2); ret = 0;
    } else if (omega == 0.0) {                                  /* :248-252 */
        *final_iter = iter; *final_err = err; ret = 3;
    } else if (rho == 0.0) {                                    /* :254-258 */
        *final_iter = iter; *final_err = err; ret = 2;
    } else {                                                    /* :260-263 */
        *final_iter = maxiter; *final_err = err; ret = 1;
    }
out:
    free(r); free(rh); free(p); free(s); free(v); free(t); free(ph); free(sh);
    return ret;
}

int oracle_solve(int kind, int64_t n1, int64_t n2, int64_t n3,
                 double a, double b, double c,
                 const double *rhs, double *x, const double *iM,
                 int maxiter, double tol,
                 double *final_err, int *final_iter, double *err_hist)
{
    const int64_t n = n1 * n2 * n3;
    double *iMv = NULL;
    if (!iM) {   /* precond.all_cells.fill(1.0/stencil.a), stochastic-estimator.cc:103 */
        iMv = malloc(n * sizeof(double));
        const double inv = 1.0 / a;
        for (int64_t i = 0; i < n; i++) iMv[i] = inv;
        iM = iMv;
    }
    int ret = -1;
    if (kind == ORACLE_CG)
        ret = cg_solve(n1, n2, n3, a, b, c, rhs, x, iM, maxiter, tol, final_err, final_iter, err_hist);
    else if (kind == ORACLE_PCG)
        ret = pcg_solve(n1, n2, n3, a, b, c, rhs, x, iM, maxiter, tol, final_err, final_iter, err_hist);
    else if (kind == ORACLE_BICGSTAB)
        ret = bicgstab_solve(n1, n2, n3, a, b, c, rhs, x, iM, maxiter, tol, final_err, final_iter, err_hist);
    free(iMv);
    return ret;
}

/* ------------------------------------------------------------------------ */
/* The estimator loop, stochastic-estimator.cc:158-216.                      */
/* ------------------------------------------------------------------------ */
int oracle_estimate(int kind, int64_t n1, int64_t n2, int64_t n3,
                    double a, double b, double c,
                    uint32_t seed, int samples, int maxiter, double tol,
                    int ignore_errors,
                    double *D, int *iters, double *errs, int *codes)
{
    const int64_t n = n1 * n2 * n3;
    double *x = malloc(n * sizeof(double)), *v = malloc(n * sizeof(double));
    double *t = calloc(n, sizeof(double)), *q = calloc(n, sizeof(double));
    oracle_mt19937 eng;
    oracle_mt_seed(&eng, seed);
    int status = 0;
    for (int k = 0; k < samples; k++) {
        for (int64_t i = 0; i < n; i++) {                       /* :165 */
            double u = oracle_uniform01(&eng);
            v[i] = (u > 0.5 ? 1.0 : -1.0);
        }
        memset(x, 0, n * sizeof(double));                       /* :175 */
        double fe = 0.0; int fi = 0;
        int code = oracle_solve(kind, n1, n2, n3, a, b, c, v, x, NULL,
                                maxiter, tol, &fe, &fi, NULL);  /* :176-179 */
        if (iters) iters[k] = fi;
        if (errs) errs[k] = fe;
        if (codes) codes[k] = code;
        if (!ignore_errors && code != 0) { status = 100 + code; break; } /* :181-184 */
        for (int64_t i = 0; i < n; i++) t[i] += x[i] * v[i];    /* :195 */
        for (int64_t i = 0; i < n; i++) q[i] += v[i] * v[i];    /* :198 */
    }
    for (int64_t i = 0; i < n; i++) D[i] = t[i] / q[i];         /* :216 */
    free(x); free(v); free(t); free(q);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Exact diag(A^-1) for A = a*I - bc*(sum of +-1 shifts), Dirichlet-0.       */
/* Eigenvectors are products of sqrt(2/(n+1)) sin(pi k i/(n+1)), eigenvalues */
/* a - 2bc*sum_d cos(pi k_d/(n_d+1)).                                        */
/* ------------------------------------------------------------------------ */
void oracle_exact_diag_inv(int64_t n1, int64_t n2, int64_t n3, double a, double bc,
                           int64_t ncells, const int64_t *cells, double *out)
{
    const int64_t nd[3] = {n1, n2, n3};
    double *cs[3], *ws[3];
    for (int d = 0; d < 3; d++) {
        cs[d] = malloc(nd[d] * sizeof(double));
        ws[d] = malloc(nd[d] * sizeof(double));
        for (int64_t k = 1; k <= nd[d]; k++)
            cs[d][k - 1] = 2.0 * bc * cos(M_PI * (double)k / (double)(nd[d] + 1));
    }
    for (int64_t ic = 0; ic < ncells; ic++) {
        const int64_t id = cells[ic];
        const int64_t ii[3] = {id % n1 + 1, (id / n1) % n2 + 1, id / (n1 * n2) + 1};
        for (int d = 0; d < 3; d++)
            for (int64_t k = 1; k <= nd[d]; k++) {
                const double sn = sin(M_PI * (double)(k * ii[d]) / (double)(nd[d] + 1));
                ws[d][k - 1] = 2.0 / (double)(nd[d] + 1) * sn * sn;
            }
        long double acc = 0.0L;
        for (int64_t k3 = 0; k3 < n3; k3++)
            for (int64_t k2 = 0; k2 < n2; k2++) {
                const double w23 = ws[1][k2] * ws[2][k3];
                const double l23 = a - cs[1][k2] - cs[2][k3];
                long double row = 0.0L;
                for (int64_t k1 = 0; k1 < n1; k1++)
                    row += (long double)(ws[0][k1] * w23) / (long double)(l23 - cs[0][k1]);
                acc += row;
            }
        out[ic] = (double)acc;
    }
    for (int d = 0; d < 3; d++) { free(cs[d]); free(ws[d]); }
}
