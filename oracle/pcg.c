/*
 * Oracle / CPU baseline -- TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's
 * cpu_baseline leg, never by the product).
 *
 * FP64 Jacobi-preconditioned CG on a CSR matrix in the row-eliminated layout of the
 * reference's GPU assembler (`gpu/gpu_assembler.py:147-165`), restating the KSP solve of
 * `numerics/solvers.py:322-372` with CG + Jacobi in place of GMRES(50) + bjacobi (the matrix
 * is SPD on the free rows once x0 holds the Dirichlet values).  Stopping test
 * ||b - A x||_2 <= rtol ||b||_2 on the recursive residual, as oracle/solve.py:pcg_jacobi.
 * OpenMP over rows; `threads` <= 0 uses the OpenMP default.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void spmv(int n, const int32_t* ip, const int32_t* ix, const double* a, const double* x,
                 double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t j = ip[i]; j < ip[i + 1]; ++j) s += a[j] * x[ix[j]];
        y[i] = s;
    }
}

static double dotp(int n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* returns iterations (>= 0), or -1 on breakdown; x in/out; rel_out = final ||r||/||b|| */
int oracle_pcg(int n, const int32_t* ip, const int32_t* ix, const double* a, const double* b,
               double* x, double rtol, int max_it, int threads, double* rel_out) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* p = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    double* dinv = (double*)malloc(sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int32_t j = ip[i]; j < ip[i + 1]; ++j)
            if (ix[j] == i) d = a[j];
        dinv[i] = 1.0 / d;
    }
    spmv(n, ip, ix, a, x, q);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
    const double bn = sqrt(dotp(n, b, b));
    const double tol = rtol * bn;
    double rn = sqrt(dotp(n, r, r));
    int it = 0;
    if (rn > tol) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            z[i] = dinv[i] * r[i];
            p[i] = z[i];
        }
        double rz = dotp(n, r, z);
        while (it < max_it) {
            spmv(n, ip, ix, a, p, q);
            const double pq = dotp(n, p, q);
            if (!(pq > 0.0)) {
                it = -1;
                break;
            }
            const double al = rz / pq;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                x[i] += al * p[i];
                r[i] -= al * q[i];
            }
            ++it;
            rn = sqrt(dotp(n, r, r));
            if (rn <= tol) break;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
            const double rzn = dotp(n, r, z);
            const double be = rzn / rz;
            rz = rzn;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
        }
    }
    *rel_out = bn > 0.0 ? rn / bn : rn;
    free(r);
    free(z);
    free(p);
    free(q);
    free(dinv);
    return it;
}
