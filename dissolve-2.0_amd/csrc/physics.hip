// The per-step physics around the transport solves, on the GPU:
//   - Zn2+ coefficients and the scaffold Dirichlet list (`dissolve.py:740-778`);
//   - the pointwise film update (`dissolve.py:797-833`);
//   - under-relaxation (`numerics/solvers.py:278-279`);
//   - the interface recession velocity (`physics/interface_velocity.py:622-785`, the default
//     nodal-O2 branch), the level-set advance (`dissolve.py:684-690`) and Sussman redistancing
//     (`domain/mesh_utils.py:450-542`) with the mean cell diameter (`:559-575`).
//
// Everything is elementwise except the gradients, which run over the CSR rows with the nodal
// gradient's edge stencil (kernels_common.h, built once per mesh by coef.hip): no atomics, so
// every result is bit-reproducible.  The scaffold list is a stream compaction in three fixed-order passes.
#include <cmath>

#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

// ---- Zn2+ coefficients: alpha_zn = [phi <= 0] kf (1 - F / Fmax), src_zn = [phi <= 0] kd F Cl^2
__global__ void k_zinc_coefficients(int32_t V, const double* __restrict__ phi,
                                    const double* __restrict__ F, const double* __restrict__ Cl,
                                    dsv_film_params p, double* __restrict__ alpha_zn,
                                    double* __restrict__ src_zn) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double medium = phi[v] <= 0.0 ? 1.0 : 0.0;
    const double Fv = F ? F[v] : 0.0;
    const double cl = Cl ? Cl[v] : 0.0;
    if (alpha_zn) alpha_zn[v] = (medium * p.kf) * (1.0 - Fv / p.Fmax);
    if (src_zn) src_zn[v] = ((medium * p.kd) * Fv) * (cl * cl);
}

// ---- stream compaction of {v : phi_v > threshold}, ascending
constexpr int SEL_THREADS = 256;

__global__ __launch_bounds__(SEL_THREADS) void k_select_count(int32_t V,
                                                              const double* __restrict__ phi,
                                                              double thr,
                                                              int32_t* __restrict__ counts) {
    const int32_t v = blockIdx.x * SEL_THREADS + threadIdx.x;
    const int pred = v < V && phi[v] > thr;
    const int n = __syncthreads_count(pred);
    if (threadIdx.x == 0) counts[blockIdx.x] = n;
}

// exclusive scan of nb block counts by one workgroup (fixed order), total in offs[nb]
__global__ __launch_bounds__(1024) void k_select_scan(int32_t nb, int32_t* __restrict__ offs) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int lo = t * per, hi = min(nb, lo + per);
    int32_t s = 0;
    for (int i = lo; i < hi; ++i) s += offs[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int32_t run = 0;
        for (int i = 0; i < 1024; ++i) {
            const int32_t c = part[i];
            part[i] = run;
            run += c;
        }
        offs[nb] = run;
    }
    __syncthreads();
    int32_t run = part[t];
    for (int i = lo; i < hi; ++i) {
        const int32_t c = offs[i];
        offs[i] = run;
        run += c;
    }
}

__global__ __launch_bounds__(SEL_THREADS) void k_select_write(int32_t V,
                                                              const double* __restrict__ phi,
                                                              double thr,
                                                              const int32_t* __restrict__ offs,
                                                              int32_t* __restrict__ out) {
    __shared__ int32_t wave_n[SEL_THREADS / 64];
    const int32_t v = blockIdx.x * SEL_THREADS + threadIdx.x;
    const bool pred = v < V && phi[v] > thr;
    const uint64_t mask = __ballot(pred);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int before = __popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) wave_n[wave] = __popcll(mask);
    __syncthreads();
    int32_t base = offs[blockIdx.x];
    for (int w = 0; w < wave; ++w) base += wave_n[w];
    if (pred) out[base + before] = v;
}

// ---- film: F = clip((F_old inv_dt + src) / (inv_dt + alpha), 0, Fmax) with
// alpha = [phi <= 0] (kf Zn / Fmax + kd Cl^2), src = [phi <= 0] kf Zn
__global__ void k_film_update(int32_t V, const double* __restrict__ phi,
                              const double* __restrict__ Zn, const double* __restrict__ Cl,
                              const double* __restrict__ Fold, dsv_film_params p, double inv_dt,
                              double* __restrict__ Fnew) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double medium = phi[v] <= 0.0 ? 1.0 : 0.0;
    const double zn = Zn[v], cl = Cl[v];
    const double alpha = medium * ((p.kf * zn) / p.Fmax + p.kd * (cl * cl));
    const double src = (medium * p.kf) * zn;
    const double F = (Fold[v] * inv_dt + src) / (inv_dt + alpha);
    // numpy's clip(F, 0, Fmax): NaN stays NaN
    Fnew[v] = F < 0.0 ? 0.0 : (F > p.Fmax ? p.Fmax : F);
}

__global__ void k_under_relax(int32_t n, const double* __restrict__ u_new,
                              const double* __restrict__ u_prev, double omega,
                              double* __restrict__ u) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u[i] = omega * u_new[i] + (1.0 - omega) * u_prev[i];
}

// Gradients of up to three P1 fields at vertex r: the volume-weighted average of the incident
// cells' constant gradients (the same arithmetic as k_nodal_gradient), sharing each cell's
// geometry between the fields.
// v_n = min(max(v_Zn, v_ORR, v_O2), 0) inside |phi| < 3 h, 0 outside, with
// n = grad phi / sqrt(|grad phi|^2 + 1e-12), v_Zn = -max(DeZn, 0) |grad Zn . n| / (Zn_sol - Zn_sat),
// v_O2 = -2 max(DeO2, 0) |grad O2 . n| / Zn_sol, v_ORR = -(2 / Zn_sol) k_orr max(O2, 0); a term
// whose fields are NULL (Zn and DeZn, or DeO2) is left out of the max
__global__ __launch_bounds__(256) void k_interface_velocity(
    int32_t V, int64_t nnz, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const double* __restrict__ G, const double* __restrict__ W, const double* phi,
    const double* Zn, const double* O2, const double* __restrict__ DeZn,
    const double* __restrict__ DeO2, dsv_velocity_params p, double denom_zn,
    double* __restrict__ vel) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t r = (int32_t)(t / GS_LANES);
    const int lane = (int)(t % GS_LANES);
    if (r >= V) return;
    const double ph = phi[r];
    if (!(fabs(ph) < 3.0 * p.h)) {  // the row's four lanes leave together
        if (lane == 0) vel[r] = 0.0;
        return;
    }
    const double* const u[3] = {phi, Zn ? Zn : phi, O2};  // (an omitted Zn term reads phi)
    Vec3<double> g[3];
    stencil_gradients<3>(r, lane, nnz, indptr, indices, G, W, u, g);
    if (lane != 0) return;
    const double gmag = sqrt(g[0].x * g[0].x + g[0].y * g[0].y + g[0].z * g[0].z + 1e-12);
    const double nx = g[0].x / gmag, ny = g[0].y / gmag, nz = g[0].z / gmag;
    const double dzn = g[1].x * nx + g[1].y * ny + g[1].z * nz;
    const double do2 = g[2].x * nx + g[2].y * ny + g[2].z * nz;
    const double o2 = np_max(O2[r], 0.0);
    const double v_zn = Zn ? -np_max(DeZn[r], 0.0) * fabs(dzn) / denom_zn : -INFINITY;
    const double v_o2d = DeO2 ? -2.0 * np_max(DeO2[r], 0.0) * fabs(do2) / p.zn_solid : -INFINITY;
    const double v_orr = -(2.0 / p.zn_solid) * p.k_orr * o2;
    vel[r] = np_min(np_max(v_zn, np_max(v_orr, v_o2d)), 0.0);
}

__global__ void k_levelset_advance(int32_t V, const double* __restrict__ phi_old,
                                   const double* __restrict__ vel, double dt,
                                   double* __restrict__ phi) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    phi[v] = phi_old[v] + vel[v] * dt;
}

// One Sussman pseudo-time iteration, Jacobi-style (every vertex from the same phi_in):
// outside the frozen band |phi0| < 3 h, phi_out = phi_in + dtau sign (1 - |grad phi_in|) with
// dtau = 0.5 h / max(|grad phi_in|, 0.5) and sign = phi0 / sqrt(phi0^2 + (1.5 h)^2)
__global__ __launch_bounds__(256) void k_redistance_iter(
    int32_t V, int64_t nnz, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const double* __restrict__ G, const double* __restrict__ W, const double* phi_in,
    const double* __restrict__ phi0, double h, double* __restrict__ phi_out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t r = (int32_t)(t / GS_LANES);
    const int lane = (int)(t % GS_LANES);
    if (r >= V) return;
    const double p0 = phi0[r];
    if (fabs(p0) < 3.0 * h) {  // the row's four lanes leave together
        if (lane == 0) phi_out[r] = phi_in[r];
        return;
    }
    const double* const u[1] = {phi_in};
    Vec3<double> g[1];
    stencil_gradients<1>(r, lane, nnz, indptr, indices, G, W, u, g);
    if (lane != 0) return;
    const double gm = sqrt(g[0].x * g[0].x + g[0].y * g[0].y + g[0].z * g[0].z + 1e-12);
    const double eps_sign = 1.5 * h;
    const double sgn = p0 / sqrt(p0 * p0 + eps_sign * eps_sign);
    const double dtau = 0.5 * h / np_max(gm, 0.5);
    phi_out[r] = phi_in[r] + dtau * sgn * (1.0 - gm);
}

// per-workgroup partial sums of the longest edge of each cell (fixed order)
constexpr int HAVG_THREADS = 256;
__global__ __launch_bounds__(HAVG_THREADS) void k_cell_diameter_partials(
    int32_t C, int32_t V, const int32_t* __restrict__ conn, const double* __restrict__ xyz,
    double* __restrict__ partials) {
    __shared__ double red[HAVG_THREADS / 64];
    const double* X = xyz;
    const double* Y = xyz + V;
    const double* Z = xyz + 2 * (size_t)V;
    double s = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)HAVG_THREADS + threadIdx.x; e < C;
         e += (int64_t)gridDim.x * HAVG_THREADS) {
        int32_t n[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) n[k] = conn[4 * e + k];
        double L2 = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a + 1; b < 4; ++b) {
                const double dx = X[n[a]] - X[n[b]], dy = Y[n[a]] - Y[n[b]],
                             dz = Z[n[a]] - Z[n[b]];
                L2 = np_max(L2, dx * dx + dy * dy + dz * dz);
            }
        s += sqrt(L2);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < HAVG_THREADS / 64; ++w) t += red[w];
        partials[blockIdx.x] = t;
    }
}

// ---- scaffold volumes: per-workgroup partial sums of a per-cell quantity times |det J| / 6,
// summed on the host in workgroup order (deterministic)
constexpr int VOL_THREADS = 256;
constexpr int VOL_MAX_WG = 1024;

__device__ __forceinline__ double smoothed_heaviside(double p, double eps) {
    return p < -eps ? 0.0
                    : (p > eps ? 1.0 : 0.5 * (1.0 + p / eps + sin(M_PI * p / eps) / M_PI));
}

// Volume of the corner tetrahedron cut off at corner a by the plane phi = 0, as a fraction of
// the cell: the product of the edge parameters a / (a - v_k) over the other three corners
__device__ __forceinline__ double corner_fraction(const double (&v)[4], int a) {
    double r = 1.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k != a) r *= v[a] / (v[a] - v[k]);
    return r;
}

__device__ __forceinline__ double tet_volume(const Vec3<double>& a, const Vec3<double>& b,
                                             const Vec3<double>& c, const Vec3<double>& d) {
    const Vec3<double> u{b.x - a.x, b.y - a.y, b.z - a.z}, v{c.x - a.x, c.y - a.y, c.z - a.z},
        w{d.x - a.x, d.y - a.y, d.z - a.z};
    return fabs(dot(u, cross(v, w))) / 6.0;
}

// Exact fraction of the cell where the linear field with vertex values v is > 0 (marching
// tetrahedra in the reference cell, affine invariant): 0 / 1 for one-signed cells, a corner
// tetrahedron for one positive (one negative) vertex, and for two and two the prism between
// the cut and the positive edge split into three tetrahedra
__device__ double solid_fraction(const double (&v)[4]) {
    int npos = 0, first_pos = -1, first_neg = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (v[k] > 0.0) {
            ++npos;
            if (first_pos < 0) first_pos = k;
        } else if (first_neg < 0) {
            first_neg = k;
        }
    }
    double f;
    if (npos == 0) {
        f = 0.0;
    } else if (npos == 4) {
        f = 1.0;
    } else if (npos == 1) {
        f = corner_fraction(v, first_pos);
    } else if (npos == 3) {
        f = 1.0 - corner_fraction(v, first_neg);
    } else {
        // positives first, each group in index order
        double pv[2], nv[2];
        int ip = 0, in = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (v[k] > 0.0)
                pv[ip++] = v[k];
            else
                nv[in++] = v[k];
        }
        const double pa = pv[0], pb = pv[1], na = nv[0], nb = nv[1];
        const Vec3<double> A{0.0, 0.0, 0.0}, B{1.0, 0.0, 0.0};
        const Vec3<double> q1{0.0, pa / (pa - na), 0.0};
        const Vec3<double> q2{0.0, 0.0, pa / (pa - nb)};
        const Vec3<double> q3{1.0 - pb / (pb - na), pb / (pb - na), 0.0};
        const Vec3<double> q4{1.0 - pb / (pb - nb), 0.0, pb / (pb - nb)};
        f = 6.0 * (tet_volume(A, q1, q2, B) + tet_volume(q1, q2, B, q3) + tet_volume(q2, B, q3, q4));
    }
    return f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
}

// mode 0: the smoothed Heaviside H_eps(phi) at the 4-point degree-2 rule (barycentric
// (alpha, beta, beta, beta), weights 1/4); mode 1: the exact solid fraction.  frac (optional)
// receives the per-cell value.
__global__ __launch_bounds__(VOL_THREADS) void k_scaffold_volume(
    int32_t C, const int32_t* __restrict__ conn, const double* __restrict__ detJ,
    const double* __restrict__ phi, int mode, double eps, double* __restrict__ frac,
    double* __restrict__ partials) {
    __shared__ double red[VOL_THREADS / 64];
    const double qa = (5.0 + 3.0 * sqrt(5.0)) / 20.0, qb = (5.0 - sqrt(5.0)) / 20.0;
    double s = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)VOL_THREADS + threadIdx.x; e < C;
         e += (int64_t)gridDim.x * VOL_THREADS) {
        const int4 n = reinterpret_cast<const int4*>(conn)[e];
        const double v[4] = {phi[n.x], phi[n.y], phi[n.z], phi[n.w]};
        double f;
        if (mode == 0) {
            const double sum = ((v[0] + v[1]) + v[2]) + v[3];
            double h = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) h += smoothed_heaviside((qa - qb) * v[k] + qb * sum, eps);
            f = h / 4.0;
        } else {
            f = solid_fraction(v);
        }
        if (frac) frac[e] = f;
        s += f * (detJ[e] / 6.0);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < VOL_THREADS / 64; ++w) t += red[w];
        partials[blockIdx.x] = t;
    }
}

static double* physics_scratch(dsv_mesh* m, int which) {
    // 0, 1: V doubles each (redistancing phi0 and ping-pong); 2: the scaffold offsets
    if (!m->phys[which]) m->phys[which] = dalloc<double>((size_t)m->V + 1);
    return m->phys[which];
}

}  // namespace dsv

using namespace dsv;

extern "C" {

int dsv_zinc_coefficients(dsv_mesh* m, const double* phi, const double* F_film, const double* Cl,
                          const dsv_film_params* params, double* alpha_zn, double* src_zn,
                          int32_t* scaffold, int32_t* n_scaffold, void* stream) {
    return guard([&] {
        if (!m || !phi || !params) throw ArgError{"dsv_zinc_coefficients: NULL argument"};
        if (!(params->Fmax != 0.0)) throw ArgError{"dsv_zinc_coefficients: Fmax must be != 0"};
        if ((scaffold == nullptr) != (n_scaffold == nullptr))
            throw ArgError{"dsv_zinc_coefficients: scaffold and n_scaffold go together"};
        DSV_HIP(hipSetDevice(m->device));
        hipStream_t st = (hipStream_t)stream;
        const unsigned gv = cdiv(m->V, 256);
        if (alpha_zn || src_zn)
            hipLaunchKernelGGL(k_zinc_coefficients, dim3(gv), dim3(256), 0, st, m->V, phi, F_film,
                               Cl, *params, alpha_zn, src_zn);
        if (scaffold) {
            const int32_t nb = (int32_t)cdiv(m->V, SEL_THREADS);
            auto* offs = reinterpret_cast<int32_t*>(physics_scratch(m, 2));
            hipLaunchKernelGGL(k_select_count, dim3(nb), dim3(SEL_THREADS), 0, st, m->V, phi, 0.0,
                               offs);
            hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(1024), 0, st, nb, offs);
            hipLaunchKernelGGL(k_select_write, dim3(nb), dim3(SEL_THREADS), 0, st, m->V, phi, 0.0,
                               offs, scaffold);
            DSV_HIP(hipMemcpyAsync(n_scaffold, offs + nb, sizeof(int32_t), hipMemcpyDeviceToHost,
                                   st));
            DSV_HIP(hipStreamSynchronize(st));
        }
        DSV_HIP(hipGetLastError());
    });
}

int dsv_film_update(dsv_mesh* m, const double* phi, const double* Zn, const double* Cl,
                    const double* F_old, const dsv_film_params* params, double inv_dt,
                    double* F_new, void* stream) {
    return guard([&] {
        if (!m || !phi || !Zn || !Cl || !F_old || !params || !F_new)
            throw ArgError{"dsv_film_update: NULL argument"};
        DSV_HIP(hipSetDevice(m->device));
        hipLaunchKernelGGL(k_film_update, dim3(cdiv(m->V, 256)), dim3(256), 0, (hipStream_t)stream,
                           m->V, phi, Zn, Cl, F_old, *params, inv_dt, F_new);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_under_relax(int32_t n, const double* u_new, const double* u_prev, double omega,
                    double* u, void* stream) {
    return guard([&] {
        if (n < 0 || (n > 0 && (!u_new || !u_prev || !u)))
            throw ArgError{"dsv_under_relax: invalid argument"};
        if (n == 0) return;
        hipLaunchKernelGGL(k_under_relax, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream,
                           n, u_new, u_prev, omega, u);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_interface_velocity(dsv_mesh* m, const double* phi, const double* Zn, const double* O2,
                           const double* DeZn, const double* DeO2,
                           const dsv_velocity_params* params, double* v, void* stream) {
    return guard([&] {
        if (!m || !phi || !O2 || !params || !v || (!Zn != !DeZn))
            throw ArgError{"dsv_interface_velocity: NULL argument (Zn and DeZn are NULL together "
                           "or neither)"};
        if (!(params->zn_solid != 0.0)) throw ArgError{"dsv_interface_velocity: zn_solid == 0"};
        double denom = params->zn_solid - params->zn_sat;
        if (fabs(denom) < 1e-30) denom = params->zn_solid;
        DSV_HIP(hipSetDevice(m->device));
        ensure_grad_stencil(m);
        hipLaunchKernelGGL(k_interface_velocity, dim3(cdiv((int64_t)m->V * GS_LANES, 256)),
                           dim3(256), 0, (hipStream_t)stream, m->V, m->nnz, m->indptr, m->indices,
                           m->gsten, m->gsw, phi, Zn, O2, DeZn, DeO2, *params, denom, v);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_levelset_advance(dsv_mesh* m, const double* phi_old, const double* v, double dt,
                         double* phi, void* stream) {
    return guard([&] {
        if (!m || !phi_old || !v || !phi) throw ArgError{"dsv_levelset_advance: NULL argument"};
        DSV_HIP(hipSetDevice(m->device));
        hipLaunchKernelGGL(k_levelset_advance, dim3(cdiv(m->V, 256)), dim3(256), 0,
                           (hipStream_t)stream, m->V, phi_old, v, dt, phi);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_mesh_h_avg(dsv_mesh* m, double* h_avg) {
    return guard([&] {
        if (!m || !h_avg) throw ArgError{"dsv_mesh_h_avg: NULL argument"};
        DSV_HIP(hipSetDevice(m->device));
        if (!(m->h_avg > 0.0)) {
            const int nb = (int)std::min<int64_t>(1024, cdiv(m->C, HAVG_THREADS));
            double* part = dalloc<double>(nb);
            hipLaunchKernelGGL(k_cell_diameter_partials, dim3(nb), dim3(HAVG_THREADS), 0, 0, m->C,
                               m->V, m->conn, m->xyz, part);
            DSV_HIP(hipGetLastError());
            std::vector<double> h(nb);
            const hipError_t e = hipMemcpy(h.data(), part, nb * sizeof(double),
                                           hipMemcpyDeviceToHost);
            dfree(part);
            DSV_HIP(e);
            double s = 0.0;
            for (double x : h) s += x;
            m->h_avg = s / m->C;
        }
        *h_avg = m->h_avg;
    });
}

int dsv_scaffold_volume(dsv_mesh* m, const double* phi, int32_t exact, double eps, double* frac,
                        double* volume, void* stream) {
    return guard([&] {
        if (!m || !phi || !volume) throw ArgError{"dsv_scaffold_volume: NULL argument"};
        if (!exact && !(eps > 0.0)) throw ArgError{"dsv_scaffold_volume: eps must be > 0"};
        DSV_HIP(hipSetDevice(m->device));
        hipStream_t st = (hipStream_t)stream;
        const int nb = (int)std::min<int64_t>(VOL_MAX_WG, cdiv(m->C, VOL_THREADS));
        double* part = physics_scratch(m, 1);  // V + 1 >= nb doubles? checked below
        if ((size_t)nb > (size_t)m->V + 1) throw ArgError{"dsv_scaffold_volume: mesh too small"};
        hipLaunchKernelGGL(k_scaffold_volume, dim3(nb), dim3(VOL_THREADS), 0, st, m->C, m->conn,
                           m->detJ, phi, exact ? 1 : 0, eps, frac, part);
        DSV_HIP(hipGetLastError());
        std::vector<double> h(nb);
        DSV_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        DSV_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (double x : h) s += x;
        *volume = s;
    });
}

int dsv_redistance(dsv_mesh* m, double* phi, int32_t num_iters, double h, void* stream) {
    return guard([&] {
        if (!m || !phi) throw ArgError{"dsv_redistance: NULL argument"};
        if (num_iters < 0) throw ArgError{"dsv_redistance: num_iters < 0"};
        DSV_HIP(hipSetDevice(m->device));
        if (!(h > 0.0)) {
            double ha = 0.0;
            if (dsv_mesh_h_avg(m, &ha) != DSV_OK) throw ArgError{dsv_last_error()};
            h = ha;
        }
        if (num_iters == 0) return;
        ensure_grad_stencil(m);
        hipStream_t st = (hipStream_t)stream;
        double* phi0 = physics_scratch(m, 0);
        double* tmp = physics_scratch(m, 1);
        DSV_HIP(hipMemcpyAsync(phi0, phi, (size_t)m->V * sizeof(double), hipMemcpyDeviceToDevice,
                               st));
        const double* in = phi0;
        for (int it = 0; it < num_iters; ++it) {
            // ping-pong so that the last iteration writes phi
            double* out = ((num_iters - 1 - it) % 2 == 0) ? phi : tmp;
            hipLaunchKernelGGL(k_redistance_iter, dim3(cdiv((int64_t)m->V * GS_LANES, 256)),
                               dim3(256), 0, st, m->V, m->nnz, m->indptr, m->indices, m->gsten,
                               m->gsw, in, phi0, h, out);
            in = out;
        }
        DSV_HIP(hipGetLastError());
    });
}

}  // extern "C"
