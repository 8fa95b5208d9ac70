// FP64 CSR assembly, Dirichlet rows and RHS (`gpu/gpu_assembler.py:60-242,472-603`).
//
// The reference scatters the 16 entries of every cell with FP64 atomicAdd through a scatter
// map.  Here the scatter is conflict-free and deterministic: one pass writes the 16 entries
// of every cell, a second pass gives every CSR entry the sum of its contributions, read
// through the contribution lists that the set-up sort produced (ascending cell order).
#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

__global__ void k_elem_A(int32_t C, int32_t V, const int32_t* __restrict__ conn,
                         const double* __restrict__ xyz, const double* __restrict__ D, int Dmode,
                         double Dval, const double* __restrict__ A, int Amode, double Aval,
                         double inv_dt, int lumped, double* __restrict__ Ae) {
    int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= C) return;
    const double* X = xyz;
    const double* Y = xyz + V;
    const double* Z = xyz + 2 * (int64_t)V;
    const int n0 = conn[4 * e], n1 = conn[4 * e + 1], n2 = conn[4 * e + 2], n3 = conn[4 * e + 3];
    Vec3<double> p0{X[n0], Y[n0], Z[n0]};
    Vec3<double> e1{X[n1] - p0.x, Y[n1] - p0.y, Z[n1] - p0.z};
    Vec3<double> e2{X[n2] - p0.x, Y[n2] - p0.y, Z[n2] - p0.z};
    Vec3<double> e3{X[n3] - p0.x, Y[n3] - p0.y, Z[n3] - p0.z};
    // rows of adj(J): grad N_k = c_k / det (k = 1..3), grad N_0 = -(c1 + c2 + c3) / det
    Vec3<double> c[4];
    c[1] = cross(e2, e3);
    c[2] = cross(e3, e1);
    c[3] = cross(e1, e2);
    c[0] = Vec3<double>{-(c[1].x + c[2].x + c[3].x), -(c[1].y + c[2].y + c[3].y),
                        -(c[1].z + c[2].z + c[3].z)};
    const double det = dot(e1, c[1]);
    const double dJ = fabs(det);
    const double De = cell_coef(D, Dmode, Dval, e, conn);
    const double al = cell_coef(A, Amode, Aval, e, conn);
    // K_ij = (dJ De / 6) grad_i . grad_j = De (c_i . c_j) / (6 dJ)
    const double w = De / (6.0 * dJ);
    const double mf = inv_dt + al;
    double* out = Ae + 16 * (int64_t)e;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double M;
            if (lumped)
                M = (i == j) ? dJ * 5.0 / 120.0 : 0.0;
            else
                M = dJ * ((i == j) ? 2.0 : 1.0) / 120.0;
            out[4 * i + j] = w * dot(c[i], c[j]) + mf * M;
        }
}

__global__ void k_gather_csr(int32_t nnz, const int32_t* __restrict__ cptr,
                             const int32_t* __restrict__ cslot, const double* __restrict__ Ae,
                             double* __restrict__ values) {
    int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    double s = 0.0;
    for (int32_t k = cptr[p]; k < cptr[p + 1]; ++k) s += Ae[cslot[k]];
    values[p] = s;
}

__global__ void k_bc_rows(int32_t n_bc, const int32_t* __restrict__ bc,
                          const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                          double* __restrict__ values) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bc) return;
    const int32_t row = bc[k];
    for (int32_t j = indptr[row]; j < indptr[row + 1]; ++j)
        values[j] = (indices[j] == row) ? 1.0 : 0.0;
}

__global__ void k_rhs(int32_t V, const int32_t* __restrict__ conn,
                      const double* __restrict__ detJ, const int32_t* __restrict__ diag_pos,
                      const int32_t* __restrict__ cptr, const int32_t* __restrict__ cslot,
                      const double* __restrict__ lm, const double* __restrict__ u,
                      const double* __restrict__ f, double inv_dt, int lumped,
                      double* __restrict__ b) {
    int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= V) return;
    const double fr = f ? f[r] : 0.0;
    if (lumped) {
        b[r] = lm[r] * (inv_dt * u[r] + fr);
        return;
    }
    // consistent: b_r = sum_e dJ/120 [inv_dt (sum_e u + u_r) + (sum_e f + f_r)]
    const int32_t p = diag_pos[r];
    double s = 0.0;
    for (int32_t k = cptr[p]; k < cptr[p + 1]; ++k) {
        const int32_t e = cslot[k] >> 4;
        const int n0 = conn[4 * e], n1 = conn[4 * e + 1], n2 = conn[4 * e + 2], n3 = conn[4 * e + 3];
        const double su = ((u[n0] + u[n1]) + u[n2]) + u[n3];
        const double sf = f ? ((f[n0] + f[n1]) + f[n2]) + f[n3] : 0.0;
        const double base = detJ[e] / 120.0;
        s += (base * inv_dt) * (su + u[r]) + base * (sf + fr);
    }
    b[r] = s;
}

// lumped RHS of up to three species in one pass over the mass diagonal (the same arithmetic
// as k_rhs per species)
struct RhsBatch {
    const double* u[3];
    const double* f[3];
    double* b[3];
};
__global__ void k_rhs_lumped_batch(int32_t V, int S, const double* __restrict__ lm, RhsBatch a,
                                   double inv_dt) {
    int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= V) return;
    const double mr = lm[r];
    for (int s = 0; s < S; ++s) {
        const double fr = a.f[s] ? a.f[s][r] : 0.0;
        a.b[s][r] = mr * (inv_dt * a.u[s][r] + fr);
    }
}

// Per-cell source -> per-vertex values: the reference's per-cell branch of assemble_rhs
// (`gpu/gpu_assembler.py:560-571`), f_vert[v] = f[searchsorted(elem_nodes, v)] with CuPy's
// wrap-around fancy indexing.  elem_nodes (the flat 4C connectivity) is not sorted, so the
// index is whatever CuPy's per-key bisection lands on; restated here per vertex: 4C when v is
// above the last entry, else the left bisection over [0, 4C - 1].  Each key is searched on its
// own, as CuPy's kernel does (numpy's searchsorted reuses the previous key's bounds for
// increasing keys and can land elsewhere on an unsorted array).
__global__ void k_cell_source_to_vertex(int32_t C, int32_t V, const int32_t* __restrict__ conn,
                                        const double* __restrict__ fc, double* __restrict__ fv) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t n = 4 * (int64_t)C;
    int64_t p;
    if (v > conn[n - 1]) {
        p = n;
    } else {
        int64_t lo = 0, hi = n - 1;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (conn[mid] < v)
                lo = mid + 1;
            else
                hi = mid;
        }
        p = hi;
    }
    fv[v] = fc[p % C];
}

__global__ void k_bc_values(int32_t n_bc, const int32_t* __restrict__ bc, double value,
                            double* __restrict__ b) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bc) return;
    b[bc[k]] = value;
}

// Dirichlet indices come from the caller's device memory: copy them back and range-check
// before any kernel indexes with them.
// Range check of a Dirichlet index list (a device-to-host copy).  Under operator reuse, whose
// contract is that the arrays named do not change, a list (address, length) passes once per
// reuse epoch.
void validate_bc_list(const dsv_mesh* m, const int32_t* bc, int32_t n_bc) {
    if (n_bc <= 0) return;
    if (!bc) throw ArgError{"bc_dofs is NULL but n_bc > 0"};
    const std::pair<const int32_t*, int32_t> key(bc, n_bc);
    if (m->op_reuse && m->bc_ok_epoch == m->op_epoch)
        for (const auto& k : m->bc_ok)
            if (k == key) return;
    std::vector<int32_t> h(n_bc);
    DSV_HIP(hipMemcpy(h.data(), bc, (size_t)n_bc * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int32_t v : h)
        if (v < 0 || v >= m->V) throw ArgError{"bc dof index out of range"};
    if (m->op_reuse) {
        if (m->bc_ok_epoch != m->op_epoch) {
            m->bc_ok.clear();
            m->bc_ok_epoch = m->op_epoch;
        }
        m->bc_ok.push_back(key);
    }
}

void check_species_coefs(const dsv_mesh* m, const dsv_species& s) {
    if (s.D && s.D_len != m->C && s.D_len != m->V)
        throw ArgError{"array length " + std::to_string(s.D_len) +
                       " matches neither n_cells=" + std::to_string(m->C) +
                       " nor n_dofs=" + std::to_string(m->V)};
    if (s.alpha && s.alpha_len != m->C && s.alpha_len != m->V)
        throw ArgError{"array length " + std::to_string(s.alpha_len) +
                       " matches neither n_cells=" + std::to_string(m->C) +
                       " nor n_dofs=" + std::to_string(m->V)};
    if (s.n_bc < 0 || (s.n_bc > 0 && !s.bc_dofs)) throw ArgError{"invalid bc_dofs / n_bc"};
    validate_bc_list(m, s.bc_dofs, s.n_bc);
}

int coef_mode(const double* p, int32_t len, const dsv_mesh* m) {
    if (!p) return COEF_SCALAR;
    return len == m->C ? COEF_CELL : COEF_VERT;
}

void assemble_into(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped, double* values,
                   hipStream_t st) {
    check_species_coefs(m, s);
    if (!values) throw ArgError{"dsv_assemble: NULL values"};
    if (!m->Ae) m->Ae = dalloc<double>(16 * (size_t)m->C);
    hipLaunchKernelGGL(k_elem_A, dim3(cdiv(m->C, 256)), dim3(256), 0, st, m->C, m->V, m->conn,
                       m->xyz, s.D, coef_mode(s.D, s.D_len, m), s.D_value, s.alpha,
                       coef_mode(s.alpha, s.alpha_len, m), s.alpha ? 0.0 : s.alpha_value, inv_dt,
                       lumped ? 1 : 0, m->Ae);
    hipLaunchKernelGGL(k_gather_csr, dim3(cdiv(m->nnz, 256)), dim3(256), 0, st, (int32_t)m->nnz,
                       m->cptr, m->cslot, m->Ae, values);
    if (s.n_bc > 0)
        hipLaunchKernelGGL(k_bc_rows, dim3(cdiv(s.n_bc, 128)), dim3(128), 0, st, s.n_bc,
                           s.bc_dofs, m->indptr, m->indices, values);
    DSV_HIP(hipGetLastError());
}

}  // namespace dsv

using namespace dsv;

extern "C" {

int dsv_assemble(dsv_mesh* m, const dsv_species* sp, double inv_dt, int32_t lumped,
                 double* values, void* stream) {
    return guard([&] {
        if (!m || !sp) throw ArgError{"dsv_assemble: NULL argument"};
        DSV_HIP(hipSetDevice(m->device));
        assemble_into(m, *sp, inv_dt, lumped, values, (hipStream_t)stream);
    });
}

int dsv_rhs(dsv_mesh* m, const double* u_old, const double* f, int32_t f_len, double inv_dt,
            int32_t lumped, const int32_t* bc_dofs, int32_t n_bc, double bc_value, double* b,
            void* stream) {
    return guard([&] {
        if (!m || !u_old || !b) throw ArgError{"dsv_rhs: NULL argument"};
        if (f && f_len != m->V && f_len != m->C)
            throw ArgError{"f length " + std::to_string(f_len) + " unrecognised"};
        if (n_bc < 0 || (n_bc > 0 && !bc_dofs)) throw ArgError{"dsv_rhs: invalid bc_dofs"};
        DSV_HIP(hipSetDevice(m->device));
        validate_bc_list(m, bc_dofs, n_bc);
        hipStream_t st = (hipStream_t)stream;
        if (f && f_len == m->C && f_len != m->V) {
            if (!m->fvert) m->fvert = dalloc<double>(m->V);
            hipLaunchKernelGGL(k_cell_source_to_vertex, dim3(cdiv(m->V, 256)), dim3(256), 0, st,
                               m->C, m->V, m->conn, f, m->fvert);
            f = m->fvert;
        }
        hipLaunchKernelGGL(k_rhs, dim3(cdiv(m->V, 256)), dim3(256), 0, st, m->V, m->conn, m->detJ,
                           m->diag_pos, m->cptr, m->cslot, m->lumped_m, u_old, f, inv_dt,
                           lumped ? 1 : 0, b);
        if (n_bc > 0)
            hipLaunchKernelGGL(k_bc_values, dim3(cdiv(n_bc, 128)), dim3(128), 0, st, n_bc, bc_dofs,
                               bc_value, b);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_rhs_batch(dsv_mesh* m, int32_t n_species, const dsv_species* sp,
                  const double* const* u_old, const double* const* f, double inv_dt,
                  int32_t lumped, double* const* b, void* stream) {
    return guard([&] {
        if (!m || !sp || !u_old || !b) throw ArgError{"dsv_rhs_batch: NULL argument"};
        if (n_species < 1 || n_species > 3) throw ArgError{"n_species must be 1..3"};
        for (int s = 0; s < n_species; ++s) {
            if (!u_old[s] || !b[s]) throw ArgError{"dsv_rhs_batch: NULL u_old or b"};
            if (sp[s].n_bc < 0 || (sp[s].n_bc > 0 && !sp[s].bc_dofs))
                throw ArgError{"dsv_rhs_batch: invalid bc_dofs"};
        }
        DSV_HIP(hipSetDevice(m->device));
        for (int s = 0; s < n_species; ++s) validate_bc_list(m, sp[s].bc_dofs, sp[s].n_bc);
        hipStream_t st = (hipStream_t)stream;
        const unsigned gv = cdiv(m->V, 256);
        if (lumped) {
            RhsBatch a{};
            for (int s = 0; s < n_species; ++s) {
                a.u[s] = u_old[s];
                a.f[s] = f ? f[s] : nullptr;
                a.b[s] = b[s];
            }
            hipLaunchKernelGGL(k_rhs_lumped_batch, dim3(gv), dim3(256), 0, st, m->V, n_species,
                               m->lumped_m, a, inv_dt);
        } else {
            for (int s = 0; s < n_species; ++s)
                hipLaunchKernelGGL(k_rhs, dim3(gv), dim3(256), 0, st, m->V, m->conn, m->detJ,
                                   m->diag_pos, m->cptr, m->cslot, m->lumped_m, u_old[s],
                                   f ? f[s] : nullptr, inv_dt, 0, b[s]);
        }
        for (int s = 0; s < n_species; ++s)
            if (sp[s].n_bc > 0)
                hipLaunchKernelGGL(k_bc_values, dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0, st,
                                   sp[s].n_bc, sp[s].bc_dofs, sp[s].bc_value, b[s]);
        DSV_HIP(hipGetLastError());
    });
}

}  // extern "C"
