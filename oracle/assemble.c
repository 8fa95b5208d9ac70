/*
 * Oracle C restatement of the per-step CPU work of one species (test infrastructure and the
 * timed CPU baseline only, like the rest of oracle/):
 *
 *   oracle_assemble   -- A = sum_e De_e K_e + (inv_dt + alpha_e) M_e scattered into the CSR
 *                        values through the scatter map, lumped (dJ/24 on the diagonal) or
 *                        consistent (dJ (1 + delta_ij) / 120), then row-only Dirichlet rows
 *                        (`gpu/gpu_assembler.py:60-139,147-165,472-532`);
 *   oracle_rhs        -- b = M (inv_dt u_old + f) per cell, lumped or consistent, then
 *                        b[bc] = bc_value (`gpu/gpu_assembler.py:176-242,534-603`).
 *
 * Per-cell De_e and alpha_e are given (the _to_cell_avg rule is applied by the caller,
 * oracle/assembly.py cell_average).  The reference scatters with FP64 atomicAdd; with more
 * than one thread this does the same (OpenMP atomics, so the summation order varies at the
 * 1e-16 level), with one thread it adds in cell order.  Checked against the numpy oracle
 * (oracle/assembly.py assemble_values / assemble_rhs) by tests/test_oracle_solve.py.
 */
#include <math.h>
#include <omp.h>
#include <string.h>

static void cell_matrix(const int* conn, const double* xyz, long e, double De, double al,
                        double inv_dt, int lumped, double A[16], double* dJ_out) {
    const double* p0 = xyz + 3 * (long)conn[4 * e];
    const double* p1 = xyz + 3 * (long)conn[4 * e + 1];
    const double* p2 = xyz + 3 * (long)conn[4 * e + 2];
    const double* p3 = xyz + 3 * (long)conn[4 * e + 3];
    double e1[3], e2[3], e3[3], c[4][3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = p1[k] - p0[k];
        e2[k] = p2[k] - p0[k];
        e3[k] = p3[k] - p0[k];
    }
    /* rows of adj(J): grad N_k = c_k / det for k = 1..3, grad N_0 = -(c_1 + c_2 + c_3) / det */
    c[1][0] = e2[1] * e3[2] - e2[2] * e3[1];
    c[1][1] = e2[2] * e3[0] - e2[0] * e3[2];
    c[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
    c[2][0] = e3[1] * e1[2] - e3[2] * e1[1];
    c[2][1] = e3[2] * e1[0] - e3[0] * e1[2];
    c[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
    c[3][0] = e1[1] * e2[2] - e1[2] * e2[1];
    c[3][1] = e1[2] * e2[0] - e1[0] * e2[2];
    c[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
    for (int k = 0; k < 3; ++k) c[0][k] = -(c[1][k] + c[2][k] + c[3][k]);
    const double det = e1[0] * c[1][0] + e1[1] * c[1][1] + e1[2] * c[1][2];
    const double dJ = fabs(det);
    const double w = De / (6.0 * dJ); /* (dJ De / 6) grad_i . grad_j = De c_i . c_j / (6 dJ) */
    const double mf = inv_dt + al;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double M = lumped ? (i == j ? dJ * 5.0 / 120.0 : 0.0)
                                    : dJ * (i == j ? 2.0 : 1.0) / 120.0;
            A[4 * i + j] = w * (c[i][0] * c[j][0] + c[i][1] * c[j][1] + c[i][2] * c[j][2]) +
                           mf * M;
        }
    if (dJ_out) *dJ_out = dJ;
}

void oracle_assemble(int C, const int* conn, const double* xyz, const double* De_c,
                     const double* al_c, double inv_dt, int lumped, const int* smap,
                     const int* indptr, const int* indices, int nnz, const int* bc, int n_bc,
                     double* values, int threads) {
    memset(values, 0, (size_t)nnz * sizeof(double));
    if (threads <= 1) {
        for (long e = 0; e < C; ++e) {
            double A[16];
            cell_matrix(conn, xyz, e, De_c[e], al_c[e], inv_dt, lumped, A, NULL);
            for (int k = 0; k < 16; ++k) values[smap[16 * e + k]] += A[k];
        }
    } else {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long e = 0; e < C; ++e) {
            double A[16];
            cell_matrix(conn, xyz, e, De_c[e], al_c[e], inv_dt, lumped, A, NULL);
            for (int k = 0; k < 16; ++k) {
#pragma omp atomic
                values[smap[16 * e + k]] += A[k];
            }
        }
    }
    for (int k = 0; k < n_bc; ++k) {
        const int row = bc[k];
        for (int j = indptr[row]; j < indptr[row + 1]; ++j)
            values[j] = (indices[j] == row) ? 1.0 : 0.0;
    }
}

void oracle_rhs(int C, int V, const int* conn, const double* xyz, const double* u,
                const double* f, double inv_dt, int lumped, const int* bc, int n_bc,
                double bc_value, double* b, int threads) {
    memset(b, 0, (size_t)V * sizeof(double));
    const int nt = threads < 1 ? 1 : threads;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (long e = 0; e < C; ++e) {
        double A[16], dJ;
        /* only |det J| is needed: the mass of a unit-free cell matrix */
        cell_matrix(conn, xyz, e, 0.0, 0.0, 0.0, lumped, A, &dJ);
        const int n[4] = {conn[4 * e], conn[4 * e + 1], conn[4 * e + 2], conn[4 * e + 3]};
        double add[4];
        if (lumped) {
            const double Ld = dJ * 5.0 / 120.0;
            for (int i = 0; i < 4; ++i) add[i] = Ld * (inv_dt * u[n[i]] + (f ? f[n[i]] : 0.0));
        } else {
            const double base = dJ / 120.0;
            const double su = u[n[0]] + u[n[1]] + u[n[2]] + u[n[3]];
            const double sf = f ? f[n[0]] + f[n[1]] + f[n[2]] + f[n[3]] : 0.0;
            for (int i = 0; i < 4; ++i)
                add[i] = base * inv_dt * (su + u[n[i]]) + base * (sf + (f ? f[n[i]] : 0.0));
        }
        for (int i = 0; i < 4; ++i) {
            if (nt > 1) {
#pragma omp atomic
                b[n[i]] += add[i];
            } else {
                b[n[i]] += add[i];
            }
        }
    }
    for (int k = 0; k < n_bc; ++k) b[bc[k]] = bc_value;
}
