/*
 * psi_oracle_omp.c -- OpenMP-threaded restatement of congrad_streamfunction
 * for the CPU baseline of the benchmark.
 *
 * TEST INFRASTRUCTURE ONLY (timing): the same algorithm as orc_congrad in
 * psi_oracle.c (solve_streamfct.f90:153-327), with the grid loops split over
 * OpenMP threads by rows.  apply_op sums each point's planes in the
 * reference's order (solve_pressure.f90:254-274 visits jj outermost, then
 * ii: S, W, C, E, N of the non-zero planes), so A p is bit-identical to the
 * serial oracle; the dot products and maxima are OpenMP reductions, whose
 * grouping differs from the serial j-outer sums, so iterates can differ in
 * the last bits.  The parity tests use the serial oracle; this variant is
 * checked against it (tests/test_oracle.py) and timed.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, nx) ((size_t)((j) + 1) * (size_t)((nx) + 4) + (size_t)((i) + 1))

void orc_setcyclic(int nx, int ny, int cyclic_x, double *p1);

static void apply_op_omp(int nx, int ny, const double *cf, const double *p, double *res)
{
    const size_t plane = (size_t)(nx + 4) * (size_t)(ny + 4);
    const double *cS = cf + 1 * plane, *cW = cf + 3 * plane, *cC = cf + 4 * plane;
    const double *cE = cf + 5 * plane, *cN = cf + 7 * plane;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const size_t q = IX(i, j, nx);
            double r = 0.0;
            r = r + cS[q] * p[IX(i, j - 1, nx)];
            r = r + cW[q] * p[IX(i - 1, j, nx)];
            r = r + cC[q] * p[q];
            r = r + cE[q] * p[IX(i + 1, j, nx)];
            r = r + cN[q] * p[IX(i, j + 1, nx)];
            res[q] = r;
        }
}

static double dot_omp(int nx, int ny, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) s += a[IX(i, j, nx)] * b[IX(i, j, nx)];
    return s;
}

static double absmax_omp(int nx, int ny, const double *a)
{
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(max : s)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const double v = fabs(a[IX(i, j, nx)]);
            s = v > s ? v : s;
        }
    return s;
}

int orc_threads(void) { return omp_get_max_threads(); }

/* Same contract as orc_congrad (without the trace). */
int orc_congrad_omp(int nx, int ny, int cyclic_x, const double *cf, const double *Z,
                    const double *forc, double *sol, double congr_epsilon,
                    int congr_max_iterations, int *iterations, int *converged,
                    double *est_error)
{
    const size_t plane = (size_t)(nx + 4) * (size_t)(ny + 4);
    double *res = calloc(plane, sizeof(double));
    double *Zres = calloc(plane, sizeof(double));
    double *ss = calloc(plane, sizeof(double));
    double *As = calloc(plane, sizeof(double));
    double estimated_error = 0.0, betakm1 = 1.0, betak, betak_min = 0.0, betaquot;
    double s_dot_As, smax, alpha, step, step1 = 0.0, convergence_rate, zresmax;
    int n, normal = 0;

    orc_setcyclic(nx, ny, cyclic_x, sol);
    apply_op_omp(nx, ny, cf, sol, res);
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            const size_t q = IX(i, j, nx);
            res[q] = forc[q] - res[q];
            Zres[q] = Z[q] * res[q];
        }
    n = 0;
    zresmax = absmax_omp(nx, ny, Zres);
    if (100.0 * zresmax < congr_epsilon) {
        estimated_error = 100.0 * zresmax;
        normal = 1;
        goto done;
    }
    for (n = 1; n <= congr_max_iterations; ++n) {
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) Zres[IX(i, j, nx)] = Z[IX(i, j, nx)] * res[IX(i, j, nx)];
        betak = dot_omp(nx, ny, Zres, res);
        if (n == 1) {
            betak_min = fabs(betak);
        } else if (n > 2) {
            betak_min = (betak_min < fabs(betak)) ? betak_min : fabs(betak);
            if (fabs(betak) > 100.0 * betak_min) goto done;
        }
        betaquot = betak / betakm1;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                ss[IX(i, j, nx)] = Zres[IX(i, j, nx)] + betaquot * ss[IX(i, j, nx)];
        orc_setcyclic(nx, ny, cyclic_x, ss);
        apply_op_omp(nx, ny, cf, ss, As);
        s_dot_As = dot_omp(nx, ny, ss, As);
        if (fabs(s_dot_As) < fabs(betak) * (double)1.e-10f) {
            estimated_error = 100.0 * absmax_omp(nx, ny, ss);
            normal = 1;
            goto done;
        }
        alpha = betak / s_dot_As;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = IX(i, j, nx);
                sol[q] = sol[q] + alpha * ss[q];
                res[q] = res[q] - alpha * As[q];
            }
        smax = absmax_omp(nx, ny, ss);
        step = fabs(alpha) * smax;
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
