// Device helpers shared by the libdsv kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dsv_internal.h"

namespace dsv {

static inline unsigned cdiv(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <class T>
struct Vec3 {
    T x, y, z;
};

template <class T>
__device__ __forceinline__ Vec3<T> cross(const Vec3<T>& a, const Vec3<T>& b) {
    return Vec3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class T>
__device__ __forceinline__ T dot(const Vec3<T>& a, const Vec3<T>& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Lane exchange by DPP (VALU, no LDS round trip), gfx9-family controls: quad_perm xor 1 and
// xor 2, row_half_mirror, row_mirror, row_bcast:15 and row_bcast:31.  Lanes of rows left out
// by ROW_MASK read 0.
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141,
              DPP_ROW_MIRROR = 0x140, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// v + the value of lane (lane ^ 1), resp. (lane ^ 2): the sum over a vertex's 2 or 4 lanes
template <class T>
__device__ __forceinline__ T add_lane_xor1(T v) {
    return v + dpp_move<DPP_QUAD_XOR1>(v);
}
template <class T>
__device__ __forceinline__ T add_lane_xor2(T v) {
    return v + dpp_move<DPP_QUAD_XOR2>(v);
}

// Sum over the 64 lanes of a wavefront; every lane gets the total.  Quad, half-row and row
// sums by DPP, then rows 0+1 and 2+3 by row broadcasts, lane 63 holding the total, which
// readlane hands to every lane.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<DPP_QUAD_XOR1>(v);
    v += dpp_move<DPP_QUAD_XOR2>(v);
    v += dpp_move<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_move<DPP_ROW_MIRROR>(v);
    v += dpp_move<DPP_ROW_BCAST15, 0xA>(v);
    v += dpp_move<DPP_ROW_BCAST31, 0xC>(v);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Coefficient of cell e for a caller-order coefficient array (COEF_CELL: length C,
// COEF_VERT: length V, averaged over the cell's corners in connectivity order) or a scalar
// (`gpu/gpu_assembler.py:452-470` _to_cell_avg averages nodal coefficients the same way).
__device__ __forceinline__ double cell_coef(const double* __restrict__ p, int mode, double val,
                                            int e, const int32_t* __restrict__ conn) {
    if (mode == COEF_SCALAR) return val;
    if (mode == COEF_CELL) return p[e];
    const int n0 = conn[4 * e], n1 = conn[4 * e + 1], n2 = conn[4 * e + 2], n3 = conn[4 * e + 3];
    return (((p[n0] + p[n1]) + p[n2]) + p[n3]) * 0.25;
}

// Edge stencil of the volume-weighted nodal gradient.  The cell gradient of a P1 field is
// sum_j u_j grad N_j^c, and sum_j grad N_j^c = 0, so the average over the cells c holding r
// with weights w_c = |det J_c| / 6 regroups by graph edge:
//   grad u(r) = sum_{j != r} G_rj (u_j - u_r) / W_r,
//   G_rj = sum_{c holding r and j} w_c grad N_j^c,   W_r = sum_{c holding r} w_c.
// G is stored per CSR entry as three FP64 arrays (x, y, z, each nnz long, 0 on the diagonal)
// and W per vertex.  GS_LANES lanes share a row, striding over its entries; the quad's DPP sum
// leaves the row's totals in every lane of the quad.
constexpr int GS_LANES = 4;
template <int NF>
__device__ __forceinline__ void stencil_gradients(int32_t r, int lane, int64_t nnz,
                                                  const int32_t* __restrict__ indptr,
                                                  const int32_t* __restrict__ indices,
                                                  const double* __restrict__ G,
                                                  const double* __restrict__ W,
                                                  const double* const (&u)[NF],
                                                  Vec3<double> (&g)[NF]) {
    double ur[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        ur[f] = u[f][r];
        g[f] = Vec3<double>{0.0, 0.0, 0.0};
    }
    const int32_t p1 = indptr[r + 1];
    for (int32_t p = indptr[r] + lane; p < p1; p += GS_LANES) {
        const int32_t j = indices[p];
        const double gx = G[p], gy = G[nnz + p], gz = G[2 * nnz + p];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double d = u[f][j] - ur[f];
            g[f].x += gx * d;
            g[f].y += gy * d;
            g[f].z += gz * d;
        }
    }
    const double w = W[r];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        g[f].x = add_lane_xor2(add_lane_xor1(g[f].x)) / w;
        g[f].y = add_lane_xor2(add_lane_xor1(g[f].y)) / w;
        g[f].z = add_lane_xor2(add_lane_xor1(g[f].z)) / w;
    }
}

// Lumped backward-Euler right-hand side of one vertex, b = m (inv_dt u + f): the one expression
// every kernel that forms it uses, so that they all round alike
__device__ __forceinline__ double lumped_rhs(double m, double inv_dt, double u, double f) {
    return m * (inv_dt * u + f);
}

// numpy's maximum / minimum: NaN if either operand is NaN (fmax / fmin drop a NaN operand)
__device__ __forceinline__ double np_max(double a, double b) {
    return (a != a || b != b) ? a + b : (a >= b ? a : b);
}
__device__ __forceinline__ double np_min(double a, double b) {
    return (a != a || b != b) ? a + b : (a <= b ? a : b);
}

}  // namespace dsv
