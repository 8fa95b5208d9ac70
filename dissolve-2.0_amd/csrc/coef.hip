// Upstream coefficient fields of the transport step, on the GPU:
//   - the nodal gradient of a P1 field, the volume-weighted average of the cells' constant
//     gradients (`gpu/gpu_assembler.py:251-307,646-671`), and |grad u| with the reference's
//     1e-12 regularisation (`physics/interface_velocity.py:65-87`);
//   - the regularised-delta ORR sink of O2 and the OH source (`dissolve.py:866-875,923-936`);
//   - the film factor of the effective diffusivities and its medium-masked form
//     (`physics/interface_velocity.py:90-155`).
//
// The reference scatters every cell's weighted gradient to its 4 vertices with FP64 atomics.
// Here the gradient is regrouped by graph edge (the stencil of kernels_common.h): the set-up
// sums, per CSR entry (r, j), w_c grad N_j^c over the cells holding both vertices in ascending
// cell order (the entry's contribution list, sorted by the pattern build), and every later
// gradient is one pass over the CSR rows, u_j - u_r times three FP64 weights per entry, with no
// atomics, so every result is bit-reproducible.  A cell's shape gradients are c_k / det J for
// k = 1..3 and -(c1 + c2 + c3) / det J for corner 0, with c1 = e2 x e3, c2 = e3 x e1,
// c3 = e1 x e2 (e_k = x_k - x_0), from the FP64 coordinates.
#include <cmath>

#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

// Set-up of the gradient stencil, one row per thread: G_rj for every entry of row r (0 on the
// diagonal) and W_r from the diagonal entry's cells.  A contribution slot is cell * 16 +
// (row corner) * 4 + (column corner).
__global__ void k_grad_stencil(int32_t V, int64_t nnz, const int32_t* __restrict__ conn,
                               const double* __restrict__ xyz, const double* __restrict__ detJ,
                               const int32_t* __restrict__ indptr,
                               const int32_t* __restrict__ diag_pos,
                               const int32_t* __restrict__ cptr, const int32_t* __restrict__ cslot,
                               double* __restrict__ G, double* __restrict__ W) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= V) return;
    const double* X = xyz;
    const double* Y = xyz + V;
    const double* Z = xyz + 2 * (size_t)V;
    const int32_t pd = diag_pos[r];
    for (int32_t p = indptr[r]; p < indptr[r + 1]; ++p) {
        double gx = 0.0, gy = 0.0, gz = 0.0, ws = 0.0;
        for (int32_t k = cptr[p]; k < cptr[p + 1]; ++k) {
            const int32_t c = cslot[k] >> 4, b = cslot[k] & 3;
            const double w = detJ[c] / 6.0;
            ws += w;
            if (p == pd) continue;
            const int32_t n0 = conn[4 * (size_t)c], n1 = conn[4 * (size_t)c + 1];
            const int32_t n2 = conn[4 * (size_t)c + 2], n3 = conn[4 * (size_t)c + 3];
            const Vec3<double> x0{X[n0], Y[n0], Z[n0]};
            const Vec3<double> e1{X[n1] - x0.x, Y[n1] - x0.y, Z[n1] - x0.z};
            const Vec3<double> e2{X[n2] - x0.x, Y[n2] - x0.y, Z[n2] - x0.z};
            const Vec3<double> e3{X[n3] - x0.x, Y[n3] - x0.y, Z[n3] - x0.z};
            const Vec3<double> c1 = cross(e2, e3), c2 = cross(e3, e1), c3 = cross(e1, e2);
            const double s = w / dot(e1, c1);
            const Vec3<double> gn = b == 1 ? c1 : b == 2 ? c2 : b == 3 ? c3
                : Vec3<double>{-(c1.x + c2.x + c3.x), -(c1.y + c2.y + c3.y), -(c1.z + c2.z + c3.z)};
            gx += s * gn.x;
            gy += s * gn.y;
            gz += s * gn.z;
        }
        G[p] = gx;
        G[nnz + p] = gy;
        G[2 * nnz + p] = gz;
        if (p == pd) W[r] = ws;
    }
}

// grad u and |grad u| = sqrt(|grad u|^2 + 1e-12): the quad's lanes 0, 1, 2 store the
// components, lane 3 the norm
__global__ __launch_bounds__(256) void k_nodal_gradient(int32_t V, int64_t nnz,
                                                        const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices,
                                                        const double* __restrict__ G,
                                                        const double* __restrict__ W,
                                                        const double* __restrict__ u,
                                                        double* __restrict__ grad,
                                                        double* __restrict__ gnorm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t r = (int32_t)(t / GS_LANES);
    const int lane = (int)(t % GS_LANES);
    if (r >= V) return;
    const double* const uu[1] = {u};
    Vec3<double> g[1];
    stencil_gradients<1>(r, lane, nnz, indptr, indices, G, W, uu, g);
    if (lane < 3) {
        if (grad) grad[3 * (size_t)r + lane] = lane == 0 ? g[0].x : lane == 1 ? g[0].y : g[0].z;
    } else if (gnorm) {
        gnorm[r] = sqrt(g[0].x * g[0].x + g[0].y * g[0].y + g[0].z * g[0].z + 1e-12);
    }
}

void ensure_grad_stencil(dsv_mesh* m) {
    if (m->gsten) return;
    double* G = dalloc<double>(3 * (size_t)m->nnz);
    double* W = nullptr;
    try {
        W = dalloc<double>((size_t)m->V);
    } catch (...) {
        dfree(G);
        throw;
    }
    hipLaunchKernelGGL(k_grad_stencil, dim3(cdiv(m->V, 128)), dim3(128), 0, 0, m->V, m->nnz,
                       m->conn, m->xyz, m->detJ, m->indptr, m->diag_pos, m->cptr, m->cslot, G, W);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        dfree(G);
        dfree(W);
        DSV_HIP(e);
    }
    m->gsten = G;
    m->gsw = W;
}

// delta_eps(phi) = (1 + cos(pi phi / eps)) / (2 eps) for |phi| < eps, else 0
__device__ __forceinline__ double delta_eps(double phi, double eps) {
    return fabs(phi) < eps ? (1.0 / (2.0 * eps)) * (1.0 + cos(M_PI * phi / eps)) : 0.0;
}

__global__ void k_transport_coefficients(int32_t V, const double* __restrict__ phi,
                                         const double* __restrict__ gnorm,
                                         const double* __restrict__ F, const double* __restrict__ Cl,
                                         const double* __restrict__ O2, dsv_coef_params p,
                                         double* __restrict__ alpha_o2, double* __restrict__ f_oh,
                                         double* __restrict__ film, double* __restrict__ medium) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double ph = phi[v];
    const double dl = delta_eps(ph, p.eps) * gnorm[v];
    const double Fv = F ? F[v] : 0.0;
    const double in_medium = ph <= 0.0 ? 1.0 : 0.0;
    if (alpha_o2) alpha_o2[v] = dl * p.k_orr;
    if (f_oh) {
        const double cl = Cl ? Cl[v] : 0.0;
        const double o2 = O2 ? O2[v] : 0.0;
        f_oh[v] = in_medium * p.kd * Fv * cl * cl + 4.0 * dl * p.k_orr * o2;
    }
    if (film || medium) {
        const double sat = np_min(np_max(Fv / fmax(p.Fmax, 1e-20), 0.0), 1.0);  // numpy clip
        const double ff = (1.0 - sat) + sat * p.film_floor;
        if (film) film[v] = ff;
        if (medium) medium[v] = ff * in_medium;
    }
}

// Effective diffusivities: DeZn = D0[0] film_factor, DeCl / DeOH / DeO2 = D0[1..3]
// medium_factor, the factors as in k_transport_coefficients
struct Diff4 {
    double d[4];
};
__global__ void k_effective_diffusion(int32_t V, const double* __restrict__ phi,
                                      const double* __restrict__ F, dsv_coef_params p, Diff4 D0,
                                      double* __restrict__ DeZn, double* __restrict__ DeCl,
                                      double* __restrict__ DeOH, double* __restrict__ DeO2) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double Fv = F ? F[v] : 0.0;
    const double sat = np_min(np_max(Fv / fmax(p.Fmax, 1e-20), 0.0), 1.0);  // numpy clip
    const double ff = (1.0 - sat) + sat * p.film_floor;
    const double medium = ff * (phi[v] <= 0.0 ? 1.0 : 0.0);
    if (DeZn) DeZn[v] = ff * D0.d[0];
    if (DeCl) DeCl[v] = medium * D0.d[1];
    if (DeOH) DeOH[v] = medium * D0.d[2];
    if (DeO2) DeO2[v] = medium * D0.d[3];
}

}  // namespace dsv

using namespace dsv;

extern "C" {

int dsv_nodal_gradient(dsv_mesh* m, const double* u, double* grad, double* grad_norm,
                       void* stream) {
    return guard([&] {
        if (!m || !u) throw ArgError{"dsv_nodal_gradient: NULL mesh or u"};
        if (!grad && !grad_norm) throw ArgError{"dsv_nodal_gradient: no output"};
        DSV_HIP(hipSetDevice(m->device));
        hipStream_t st = (hipStream_t)stream;
        ensure_grad_stencil(m);
        hipLaunchKernelGGL(k_nodal_gradient, dim3(cdiv((int64_t)m->V * GS_LANES, 256)), dim3(256),
                           0, st, m->V, m->nnz, m->indptr, m->indices, m->gsten, m->gsw, u, grad,
                           grad_norm);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_transport_coefficients(dsv_mesh* m, const double* phi, const double* grad_norm,
                               const double* F_film, const double* Cl, const double* O2,
                               const dsv_coef_params* params, double* alpha_o2, double* f_oh,
                               double* film_factor, double* medium_factor, void* stream) {
    return guard([&] {
        if (!m || !phi || !grad_norm || !params)
            throw ArgError{"dsv_transport_coefficients: NULL mesh, phi, grad_norm or params"};
        if (!(params->eps > 0.0)) throw ArgError{"dsv_transport_coefficients: eps must be > 0"};
        DSV_HIP(hipSetDevice(m->device));
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(k_transport_coefficients, dim3(cdiv(m->V, 256)), dim3(256), 0, st, m->V,
                           phi, grad_norm, F_film, Cl, O2, *params, alpha_o2, f_oh, film_factor,
                           medium_factor);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_effective_diffusion(dsv_mesh* m, const double* phi, const double* F_film,
                            const dsv_coef_params* params, const double* D0, double* DeZn,
                            double* DeCl, double* DeOH, double* DeO2, void* stream) {
    return guard([&] {
        if (!m || !phi || !params || !D0)
            throw ArgError{"dsv_effective_diffusion: NULL mesh, phi, params or D0"};
        Diff4 d{{D0[0], D0[1], D0[2], D0[3]}};
        DSV_HIP(hipSetDevice(m->device));
        hipLaunchKernelGGL(k_effective_diffusion, dim3(cdiv(m->V, 256)), dim3(256), 0,
                           (hipStream_t)stream, m->V, phi, F_film, *params, d, DeZn, DeCl, DeOH,
                           DeO2);
        DSV_HIP(hipGetLastError());
    });
}

}  // extern "C"
