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

// Sum over the 64 lanes of a wavefront; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
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

}  // namespace dsv
