// Device helpers shared by the libdsv kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Block-wide sums of N values (fixed order: lanes, then waves in index order).  `scratch`
// must hold 16*N doubles (up to 1024 threads).  Result valid in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < N; ++i) scratch[i * 16 + wid] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += scratch[i * 16 + w];
            v[i] = s;
        }
    __syncthreads();
}

__device__ __forceinline__ void store_partial(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_partial(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every block calls this after thread 0 stored its partials.  Returns true in exactly one
// block -- the last to arrive -- after which that block may read every block's partials.
// The ticket is reset by the last block for the next launch.
__device__ __forceinline__ bool last_block_arrive(unsigned* ticket, unsigned nblocks) {
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == nblocks - 1) ? 1 : 0;
        if (is_last) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (is_last) __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return is_last != 0;
}

// Sum partials[q*stride + b] over b < nb for q < NQ in a fixed order, using the whole block.
// Result valid in thread 0.
template <int NQ>
__device__ __forceinline__ void reduce_partials(const double* partials, int stride, int nb,
                                                double (&out)[NQ], double* scratch) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s += load_partial(partials + q * stride + b);
        out[q] = s;
    }
    block_sum<NQ>(out, scratch);
}

}  // namespace dsv
