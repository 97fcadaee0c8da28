/*
 * psi_oracle_omp.c -- OpenMP-threaded restatement of congrad_streamfunction,
 * for the benchmark's all-cores CPU baseline and for checks at full size
 * (2048^2, 4096^2) where the serial oracle takes minutes per solve.
 *
 * TEST INFRASTRUCTURE ONLY: the same algorithm as orc_congrad in
 * psi_oracle.c (solve_streamfct.f90:153-327) with the same tests in the same
 * order, but the grid loops fused into three passes per iteration, each
 * split over OpenMP threads by rows:
 *   1. p = Z r + beta/beta_prev p with the cyclic halo set row by row
 *      (:245-251; setcyclic_xy only copies within a row);
 *   2. A p and p.Ap (:254-261);
 *   3. x += alpha p, r -= alpha A p (:276-281), max|p| (:283) and, for the
 *      next iteration, Z r and (Z r).r (:226-231).
 * apply_op sums each point's non-zero planes in the reference's order
 * (solve_pressure.f90:254-274 visits jj outermost, then ii: S, W, C, E, N),
 * so A p is bit-identical to the serial oracle; the dot products and maxima
 * are per-thread partial sums combined by OpenMP, whose grouping differs from
 * the serial j-outer sums, so iterates differ in the last bits and iteration
 * counts by a few.  tests/test_oracle.py checks it against the serial oracle.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, nx) ((size_t)((j) + 1) * (size_t)((nx) + 4) + (size_t)((i) + 1))

void orc_setcyclic(int nx, int ny, int cyclic_x, double *p1);

int orc_threads(void) { return omp_get_max_threads(); }

/* one row's cyclic halo (parallel_none.f90:188-211 restricted to row j) */
static inline void cyclic_row(int nx, int j, double *a)
{
    for (int i = 1; i <= 2; ++i) {
        a[IX(nx + i, j, nx)] = a[IX(i, j, nx)];
        a[IX(1 - i, j, nx)] = a[IX(nx - i + 1, j, nx)];
    }
}

static inline double apply_pt(const double *cS, const double *cW, const double *cC,
                              const double *cE, const double *cN, const double *p, int nx,
                              int i, int j)
{
    const size_t q = IX(i, j, nx);
    double r = 0.0;
    r = r + cS[q] * p[IX(i, j - 1, nx)];
    r = r + cW[q] * p[IX(i - 1, j, nx)];
    r = r + cC[q] * p[q];
    r = r + cE[q] * p[IX(i + 1, j, nx)];
    r = r + cN[q] * p[IX(i, j + 1, nx)];
    return r;
}

/* Same contract as orc_congrad (without the trace). */
int orc_congrad_omp(int nx, int ny, int cyclic_x, const double *cf, const double *Z,
                    const double *forc, double *sol, double congr_epsilon,
                    int congr_max_iterations, int *iterations, int *converged,
                    double *est_error)
{
    const size_t plane = (size_t)(nx + 4) * (size_t)(ny + 4);
    const double *cS = cf + 1 * plane, *cW = cf + 3 * plane, *cC = cf + 4 * plane;
    const double *cE = cf + 5 * plane, *cN = cf + 7 * plane;
    double *res = calloc(plane, sizeof(double));
    double *Zres = calloc(plane, sizeof(double));
    double *ss = calloc(plane, sizeof(double));
    double *As = calloc(plane, sizeof(double));
    double estimated_error = 0.0, betakm1 = 1.0, betak, betak_next = 0.0, betak_min = 0.0;
    double betaquot, s_dot_As, smax, alpha, step, step1 = 0.0, convergence_rate, zresmax;
    int n, normal = 0;

    orc_setcyclic(nx, ny, cyclic_x, sol);                                   /* :191-192 */
    zresmax = 0.0;
    betak_next = 0.0;
#pragma omp parallel for schedule(static) reduction(max : zresmax) reduction(+ : betak_next)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {                                     /* :196-213 */
            const size_t q = IX(i, j, nx);
            const double r = forc[q] - apply_pt(cS, cW, cC, cE, cN, sol, nx, i, j);
            const double z = Z[q] * r;
            res[q] = r;
            Zres[q] = z;
            const double a = fabs(z);
            zresmax = a > zresmax ? a : zresmax;
            betak_next += z * r;
        }
    n = 0;
    if (100.0 * zresmax < congr_epsilon) {
        estimated_error = 100.0 * zresmax;
        normal = 1;
        goto done;
    }
    for (n = 1; n <= congr_max_iterations; ++n) {                           /* :223 */
        betak = betak_next;
        if (n == 1) {
            betak_min = fabs(betak);
        } else if (n > 2) {
            betak_min = (betak_min < fabs(betak)) ? betak_min : fabs(betak);
            if (fabs(betak) > 100.0 * betak_min) goto done;                 /* :236-238 */
        }
        betaquot = betak / betakm1;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j) {
            for (int i = 1; i <= nx; ++i)
                ss[IX(i, j, nx)] = Zres[IX(i, j, nx)] + betaquot * ss[IX(i, j, nx)];
            if (cyclic_x) cyclic_row(nx, j, ss);
        }
        s_dot_As = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s_dot_As)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = IX(i, j, nx);
                const double a = apply_pt(cS, cW, cC, cE, cN, ss, nx, i, j);
                As[q] = a;
                s_dot_As += ss[q] * a;
            }
        if (fabs(s_dot_As) < fabs(betak) * (double)1.e-10f) {              /* :262 */
            smax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : smax)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const double a = fabs(ss[IX(i, j, nx)]);
                    smax = a > smax ? a : smax;
                }
            estimated_error = 100.0 * smax;
            normal = 1;
            goto done;
        }
        alpha = betak / s_dot_As;
        smax = 0.0;
        betak_next = 0.0;
#pragma omp parallel for schedule(static) reduction(max : smax) reduction(+ : betak_next)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = IX(i, j, nx);
                const double s = ss[q];
                sol[q] = sol[q] + alpha * s;
                const double r = res[q] - alpha * As[q];
                res[q] = r;
                const double z = Z[q] * r;
                Zres[q] = z;
                betak_next += z * r;
                const double a = fabs(s);
                smax = a > smax ? a : smax;
            }
        step = fabs(alpha) * smax;                                          /* :288 */
        if (n == 1) {
            step1 = step;
            estimated_error = step;
            if (step < congr_epsilon) { normal = 1; goto done; }
        } else if (step < congr_epsilon) {
            convergence_rate = exp(log(step / step1) / (double)(n - 1));
            estimated_error = step * convergence_rate / (1.0 - convergence_rate);
            if (estimated_error < congr_epsilon) { normal = 1; goto done; }
        }
        betakm1 = betak;
    }
done:
    *iterations = n;
    *converged = normal;
    *est_error = estimated_error;
    free(res); free(Zres); free(ss); free(As);
    if (normal) return 0;
    return (estimated_error != estimated_error) ? 2 : 1;
}
