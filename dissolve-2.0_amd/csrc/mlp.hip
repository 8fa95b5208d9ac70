// Learned CG warm start: a per-vertex MLP 7 -> 64 -> 64 -> 1 (tanh) that predicts the next
// backward-Euler increment of a species from local features; inference, MSE gradient and
// Adam on the GPU.  No counterpart in the reference (its solves start from the previous
// solution, `numerics/solvers.py:322-372`).
//
// Parameters (FP32, flat): W1[64][7], b1[64], W2[64][64], b2[64], w3[64], b3.
// Features of vertex v of species s at step n (row of X, 7 FP32 values):
//   u_n, du/sigma, clamp(phi/eps, -4, 4)/4, min(|grad phi|, 4)/4, log(D)/2, alpha*dt, dirichlet
// with du = u_n - u_{n-1} and sigma = max_v |du| (a per-species, per-step scale); the network
// predicts a per-vertex quantity in units of sigma (the stepper trains it on the part of the
// next increment that linear extrapolation misses).
//
// The gradient is computed without atomics: workgroups of 64 lanes take 64-sample chunks, lane
// j owns row j of every weight matrix and accumulates over the chunks it sees in a fixed
// order; a second kernel sums the per-workgroup gradients in workgroup order.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>

#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

constexpr int MLP_IN = 7;
constexpr int MLP_H = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = OFF_B3 + 1;  // 4737
constexpr int GRAD_WG = 64;             // lanes (and samples per chunk) of the gradient kernel
constexpr int GRAD_MAX_WG = 1024;

static_assert(MLP_NPARAM == DSV_MLP_NPARAM, "parameter count of the C-ABI");

__device__ __forceinline__ void load_params(const float* __restrict__ p, float* sp) {
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
}

// forward of one sample with parameters in LDS; h1 (and h2 when given) kept for backprop
__device__ __forceinline__ float mlp_forward(const float* sp, const float (&x)[MLP_IN],
                                             float (&h1)[MLP_H], float* h2_out) {
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float a = sp[OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a += sp[OFF_W1 + j * MLP_IN + i] * x[i];
        h1[j] = tanhf(a);
    }
    float out = sp[OFF_B3];
    for (int j = 0; j < MLP_H; ++j) {
        float a = sp[OFF_B2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) a += sp[OFF_W2 + j * MLP_H + k] * h1[k];
        const float h = tanhf(a);
        if (h2_out) h2_out[j] = h;
        out += sp[OFF_W3 + j] * h;
    }
    return out;
}

// ---------------------------------------------------------------------------------------
// Inference on the matrix cores.  A wavefront takes 32 samples at a time and computes both
// layers as transposed products H^T = W X^T with v_mfma_f32_32x32x2f32 (FP32 in, FP32
// accumulate): rows are hidden units, columns the 32 samples.  In the 32 x 32 result, lane l
// holds column n = l % 32 and, in register r, row 8 (r / 4) + 4 (l / 32) + r % 4.  Layer 2
// sums over the hidden units in exactly that order: at k-step (T, r) the two lane halves
// supply the units they already hold, and the A operand reads the matching columns of W2, so
// the first layer's outputs feed the second without moving.  Layer 1 pads its 7 inputs to 8.
typedef float f32x16 __attribute__((ext_vector_type(16)));

// tanh(a) = 1 - 2 / (1 + e^(2a)), with the hardware exponential and reciprocal (absolute error
// about 1e-7; +-1 at overflow of the exponential)
__device__ __forceinline__ float tanh_fast(float a) {
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * a));
}

// LDS layout of the inference kernel: W1 rows padded to 9 and W2 rows to 65 floats, so that the
// 32 rows one operand load touches fall into distinct banks
constexpr int FW_W1 = 0;                       // [64][9]
constexpr int FW_W2 = FW_W1 + MLP_H * 9;       // [64][65]
constexpr int FW_B1 = FW_W2 + MLP_H * 65;
constexpr int FW_B2 = FW_B1 + MLP_H;
constexpr int FW_W3 = FW_B2 + MLP_H;
constexpr int FW_B3 = FW_W3 + MLP_H;
constexpr int FW_N = FW_B3 + 1;

__device__ __forceinline__ int unit_of(int T, int r, int half) {
    return 32 * T + 8 * (r >> 2) + 4 * half + (r & 3);
}

__global__ void __launch_bounds__(256)
k_mlp_forward(const float* __restrict__ params, const float* __restrict__ X, int n,
              float* __restrict__ out) {
    __shared__ float sw[FW_N];
    for (int i = threadIdx.x; i < MLP_H * 9; i += blockDim.x) {
        const int m = i / 9, k = i % 9;
        sw[FW_W1 + i] = k < MLP_IN ? params[OFF_W1 + m * MLP_IN + k] : 0.0f;
    }
    for (int i = threadIdx.x; i < MLP_H * 65; i += blockDim.x) {
        const int m = i / 65, k = i % 65;
        sw[FW_W2 + i] = k < MLP_H ? params[OFF_W2 + m * MLP_H + k] : 0.0f;
    }
    for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        sw[FW_B1 + i] = params[OFF_B1 + i];
        sw[FW_B2 + i] = params[OFF_B2 + i];
        sw[FW_W3 + i] = params[OFF_W3 + i];
    }
    if (threadIdx.x == 0) sw[FW_B3] = params[OFF_B3];
    __syncthreads();
    const int lane = threadIdx.x & 63, col = lane & 31, half = lane >> 5;
    const int wpb = blockDim.x >> 6;
    const int stride = gridDim.x * wpb * 32;
    for (int base = (blockIdx.x * wpb + (threadIdx.x >> 6)) * 32; base < n; base += stride) {
        const int smp = base + col;
        const bool live = smp < n;
        // layer 1: rows 32 T + m of W1 X^T, K = 8 in steps of 2 (lane half = k)
        f32x16 h1[2];
#pragma unroll
        for (int T = 0; T < 2; ++T)
#pragma unroll
            for (int r = 0; r < 16; ++r) h1[T][r] = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int f = 2 * t + half;
            const float xb = (live && f < MLP_IN) ? X[(size_t)smp * MLP_IN + f] : 0.0f;
#pragma unroll
            for (int T = 0; T < 2; ++T)
                h1[T] = __builtin_amdgcn_mfma_f32_32x32x2f32(sw[FW_W1 + (32 * T + col) * 9 + f],
                                                             xb, h1[T], 0, 0, 0);
        }
#pragma unroll
        for (int T = 0; T < 2; ++T)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                h1[T][r] = tanh_fast(h1[T][r] + sw[FW_B1 + unit_of(T, r, half)]);
        // layer 2: K runs over the (T, r) register pairs, unit_of(T, r, half) per lane half
        f32x16 h2[2];
#pragma unroll
        for (int M = 0; M < 2; ++M)
#pragma unroll
            for (int r = 0; r < 16; ++r) h2[M][r] = 0.0f;
#pragma unroll
        for (int T = 0; T < 2; ++T)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int u = unit_of(T, r, half);
#pragma unroll
                for (int M = 0; M < 2; ++M)
                    h2[M] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                        sw[FW_W2 + (32 * M + col) * 65 + u], h1[T][r], h2[M], 0, 0, 0);
            }
        // output: b3 + sum over the units of w3 tanh(layer 2 + b2); the two lane halves hold
        // the two halves of a sample's units
        float acc = 0.0f;
#pragma unroll
        for (int M = 0; M < 2; ++M)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int u = unit_of(M, r, half);
                acc += sw[FW_W3 + u] * tanh_fast(h2[M][r] + sw[FW_B2 + u]);
            }
        acc += __shfl_xor(acc, 32, 64);
        if (half == 0 && live) out[smp] = acc + sw[FW_B3];
    }
}

// Gradient of L = mean_i (f(x_i) - y_i)^2 (partials per workgroup: NPARAM values + loss sum)
__global__ void __launch_bounds__(GRAD_WG)
k_mlp_grad(const float* __restrict__ params, const float* __restrict__ X,
           const float* __restrict__ y, int n, float* __restrict__ gpart) {
    __shared__ float sp[MLP_NPARAM];
    __shared__ float sx[GRAD_WG][MLP_IN];
    __shared__ float sh1[GRAD_WG][MLP_H + 1];  // +1: lane-strided rows without bank conflicts
    __shared__ float sh2[GRAD_WG][MLP_H + 1];  // h2, then dL/da2, then dL/da1
    __shared__ float sg[GRAD_WG];
    load_params(params, sp);
    const int t = threadIdx.x;  // sample slot in a chunk, and the row j = t owned for gradients
    const float inv_n = 1.0f / (float)n;
    float gW1[MLP_IN], gW2[MLP_H];
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) gW1[i] = 0.f;
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) gW2[k] = 0.f;
    const int nchunks = (n + GRAD_WG - 1) / GRAD_WG;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int i = c * GRAD_WG + t;
        const bool live = i < n;
        float x[MLP_IN];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) {
            x[k] = live ? X[(size_t)i * MLP_IN + k] : 0.f;
            sx[t][k] = x[k];
        }
        float h1[MLP_H];
        const float out = mlp_forward(sp, x, h1, sh2[t]);
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) sh1[t][k] = h1[k];
        const float r = live ? out - y[i] : 0.f;
        loss += r * r;
        const float g = 2.f * r * inv_n;  // dL/dout of this sample
        sg[t] = g;
        __syncthreads();
        // row t of the output layer, and dL/da2 of sample t (in place of h2 after use)
        {
            float s = 0.f;
            for (int q = 0; q < GRAD_WG; ++q) s += sg[q] * sh2[q][t];
            gw3 += s;
            if (t == 0)
                for (int q = 0; q < GRAD_WG; ++q) gb3 += sg[q];
        }
        __syncthreads();
        for (int j = 0; j < MLP_H; ++j) {
            const float h = sh2[t][j];
            sh2[t][j] = g * sp[OFF_W3 + j] * (1.f - h * h);
        }
        __syncthreads();
        // row t of W2 / b2
        for (int q = 0; q < GRAD_WG; ++q) {
            const float d = sh2[q][t];
            gb2 += d;
#pragma unroll
            for (int k = 0; k < MLP_H; ++k) gW2[k] += d * sh1[q][k];
        }
        // dL/da1 of sample t (from its own dL/da2 row), then stored over that row
        float d1[MLP_H];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) {
            float s = 0.f;
            for (int j = 0; j < MLP_H; ++j) s += sp[OFF_W2 + j * MLP_H + k] * sh2[t][j];
            d1[k] = s * (1.f - h1[k] * h1[k]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) sh2[t][k] = d1[k];
        __syncthreads();
        // row t of W1 / b1
        for (int q = 0; q < GRAD_WG; ++q) {
            const float d = sh2[q][t];
            gb1 += d;
#pragma unroll
            for (int k = 0; k < MLP_IN; ++k) gW1[k] += d * sx[q][k];
        }
        __syncthreads();  // LDS rows are refilled by the next chunk
    }
    float* gp = gpart + (size_t)blockIdx.x * (MLP_NPARAM + 1);
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) gp[OFF_W1 + t * MLP_IN + k] = gW1[k];
    gp[OFF_B1 + t] = gb1;
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) gp[OFF_W2 + t * MLP_H + k] = gW2[k];
    gp[OFF_B2 + t] = gb2;
    gp[OFF_W3 + t] = gw3;
    // loss: lane sums in lane order by lane 0
    sg[t] = loss;
    __syncthreads();
    if (t == 0) {
        float l = 0.f;
        for (int q = 0; q < GRAD_WG; ++q) l += sg[q];
        gp[OFF_B3] = gb3;
        gp[MLP_NPARAM] = l;
    }
}

// grad[p] = sum over workgroups of the partials, in workgroup order; loss = mean square error
__global__ void k_mlp_grad_reduce(const float* __restrict__ gpart, int nwg, int n,
                                  float* __restrict__ grad, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NPARAM) return;
    float s = 0.f;
    for (int w = 0; w < nwg; ++w) s += gpart[(size_t)w * (MLP_NPARAM + 1) + p];
    if (p < MLP_NPARAM)
        grad[p] = s;
    else if (loss)
        *loss = s / (float)n;
}

__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// |x| maximum over n values as float bits (non-negative floats order like their bits)
// max |a - b| as FP32 bits (non-negative floats order like their bits): wave and workgroup
// maxima first, one atomic per workgroup (max is order-independent, so the result is exact)
__global__ void __launch_bounds__(256) k_absmax_bits(int n, const double* __restrict__ a,
                                                     const double* __restrict__ b,
                                                     unsigned* __restrict__ out) {
    __shared__ unsigned wmax[4];
    unsigned best = 0u;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = (float)fabs(a[i] - b[i]);
        best = max(best, __float_as_uint(d));
    }
    for (int off = 32; off > 0; off >>= 1) best = max(best, (unsigned)__shfl_xor((int)best, off, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) best = max(best, wmax[w]);
        atomicMax(out, best);
    }
}

__global__ void k_mlp_features(int n, const double* __restrict__ u, const double* __restrict__ u_prev,
                               const double* __restrict__ phi, const double* __restrict__ gphi,
                               const double* __restrict__ alpha, double alpha_value, double D,
                               double dt, double eps, const uint8_t* __restrict__ bcmask,
                               const unsigned* __restrict__ sigma_bits, float* __restrict__ X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sig = fmaxf(__uint_as_float(*sigma_bits), 1e-30f);
    float* x = X + (size_t)i * MLP_IN;
    x[0] = (float)u[i];
    x[1] = (float)(u[i] - u_prev[i]) / sig;
    x[2] = fminf(fmaxf((float)(phi[i] / eps), -4.f), 4.f) * 0.25f;
    x[3] = fminf((float)gphi[i], 4.f) * 0.25f;
    x[4] = (float)(0.5 * log(D));
    x[5] = (float)((alpha ? alpha[i] : alpha_value) * dt);
    x[6] = bcmask && bcmask[i] ? 1.f : 0.f;
}

// Static scratch of dsv_mlp_grad_noalloc: the partial gradients of at most NOALLOC_WG
// workgroups (4.85 MB), held by every GPU the library runs on until the process ends.  Every
// call writes all the partials it reads, so nothing needs clearing; calls that overlap on one
// GPU share this array and race.
constexpr int NOALLOC_WG = 256;
constexpr int NOALLOC_MAX_N = NOALLOC_WG * GRAD_WG;  // 16384 samples
static __device__ __align__(256) float scratch[(size_t)NOALLOC_WG * (MLP_NPARAM + 1)];

// Device address of `scratch` on the current device, looked up once per device (the array's
// name is a host-side stand-in, not a device address).
static float* scratch_on_current_device() {
    static std::mutex mu;
    static std::unordered_map<int, float*> addr;
    int dev = 0;
    DSV_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    DSV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ArgError{"dsv_mlp_grad_noalloc: no device address for the scratch array"};
    addr[dev] = static_cast<float*>(p);
    return static_cast<float*>(p);
}

// The learned warm start's directions in FP64: dirs[s][i] = out[s V + i] (sigma_s tau), from one
// forward pass over the rows of all species
struct DirOut {
    double* d[3];
};
__global__ void k_mlp_directions(int32_t V, const float* __restrict__ out,
                                 const float* __restrict__ sigma, double tau, DirOut dirs) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= V) return;
    dirs.d[s][i] = (double)out[(size_t)s * V + i] * ((double)sigma[s] * tau);
}

}  // namespace dsv

using namespace dsv;

extern "C" {

int dsv_mlp_forward(const float* params, const float* X, int32_t n, float* out, void* stream) {
    return guard([&] {
        if (!params || !X || !out || n < 0) throw ArgError{"dsv_mlp_forward: bad arguments"};
        if (n == 0) return;
        // four wavefronts of 32-sample blocks per workgroup, at most 1024 workgroups
        const unsigned grid = std::min<unsigned>(cdiv(n, 128), 1024);
        hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(256), 0, (hipStream_t)stream, params, X,
                           n, out);
        DSV_HIP(hipGetLastError());
    });
}

size_t dsv_mlp_grad_workspace(int32_t n) {
    const int chunks = (int)cdiv(n > 0 ? n : 1, GRAD_WG);
    const int nwg = chunks < GRAD_MAX_WG ? chunks : GRAD_MAX_WG;
    return (size_t)nwg * (MLP_NPARAM + 1) * sizeof(float);
}

int dsv_mlp_grad(const float* params, const float* X, const float* y, int32_t n, float* grad,
                 float* loss, void* workspace, size_t workspace_bytes, void* stream) {
    return guard([&] {
        if (!params || !X || !y || !grad || !workspace || n < 1)
            throw ArgError{"dsv_mlp_grad: bad arguments"};
        if (workspace_bytes < dsv_mlp_grad_workspace(n))
            throw ArgError{"dsv_mlp_grad: workspace smaller than dsv_mlp_grad_workspace(n)"};
        const int chunks = (int)cdiv(n, GRAD_WG);
        const int nwg = std::min(chunks, GRAD_MAX_WG);
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(GRAD_WG), 0, st, params, X, y, n,
                           (float*)workspace);
        hipLaunchKernelGGL(k_mlp_grad_reduce, dim3(cdiv(MLP_NPARAM + 1, 256)), dim3(256), 0, st,
                           (const float*)workspace, nwg, n, grad, loss);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_mlp_grad_noalloc(const float* params, const float* X, const float* y, int32_t n,
                         float* grad, float* loss, void* stream) {
    return guard([&] {
        if (!params || !X || !y || !grad || n < 1)
            throw ArgError{"dsv_mlp_grad_noalloc: bad arguments"};
        if (n > NOALLOC_MAX_N)
            throw ArgError{"dsv_mlp_grad_noalloc: n = " + std::to_string(n) +
                           " exceeds the static scratch (at most " +
                           std::to_string(NOALLOC_MAX_N) + " samples); use dsv_mlp_grad"};
        float* work = scratch_on_current_device();
        const int nwg = (int)cdiv(n, GRAD_WG);  // <= NOALLOC_WG, the same as dsv_mlp_grad's
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(GRAD_WG), 0, st, params, X, y, n, work);
        hipLaunchKernelGGL(k_mlp_grad_reduce, dim3(cdiv(MLP_NPARAM + 1, 256)), dim3(256), 0, st,
                           (const float*)work, nwg, n, grad, loss);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_mlp_adam(float* params, const float* grad, float* m, float* v, int32_t step, float lr,
                 float beta1, float beta2, float eps, void* stream) {
    return guard([&] {
        if (!params || !grad || !m || !v || step < 1) throw ArgError{"dsv_mlp_adam: bad arguments"};
        const float c1 = 1.f - std::pow(beta1, (float)step);
        const float c2 = 1.f - std::pow(beta2, (float)step);
        hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv(MLP_NPARAM, 256)), dim3(256), 0, (hipStream_t)stream,
                           params, grad, m, v, MLP_NPARAM, lr, beta1, beta2, eps, c1, c2);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_mlp_directions(int32_t n_species, int32_t V, const float* out, const float* sigma,
                       double tau, double* const* dirs, void* stream) {
    return guard([&] {
        if (n_species < 1 || n_species > 3 || V < 1 || !out || !sigma || !dirs)
            throw ArgError{"dsv_mlp_directions: bad arguments"};
        DirOut d{};
        for (int s = 0; s < n_species; ++s) {
            if (!dirs[s]) throw ArgError{"dsv_mlp_directions: NULL direction"};
            d.d[s] = dirs[s];
        }
        hipLaunchKernelGGL(k_mlp_directions, dim3(cdiv(V, 256), n_species), dim3(256), 0,
                           (hipStream_t)stream, V, out, sigma, tau, d);
        DSV_HIP(hipGetLastError());
    });
}

int dsv_mlp_features(int32_t n, const double* u, const double* u_prev, const double* phi,
                     const double* grad_phi, const double* alpha, double alpha_value, double D,
                     double dt, double eps, const uint8_t* bc_mask, float* X, float* sigma,
                     void* stream) {
    return guard([&] {
        if (n < 1 || !u || !u_prev || !phi || !grad_phi || !X || !sigma)
            throw ArgError{"dsv_mlp_features: bad arguments"};
        if (!(D > 0.0) || !(dt > 0.0) || !(eps > 0.0))
            throw ArgError{"dsv_mlp_features: D, dt and eps must be > 0"};
        hipStream_t st = (hipStream_t)stream;
        unsigned* bits = reinterpret_cast<unsigned*>(sigma);  // max |du| as float bits
        DSV_HIP(hipMemsetAsync(bits, 0, sizeof(unsigned), st));
        hipLaunchKernelGGL(k_absmax_bits, dim3(std::min<unsigned>(cdiv(n, 256), 128)), dim3(256), 0,
                           st, n, u, u_prev, bits);
        hipLaunchKernelGGL(k_mlp_features, dim3(cdiv(n, 256)), dim3(256), 0, st, n, u, u_prev, phi,
                           grad_phi, alpha, alpha_value, D, dt, eps, bc_mask, bits, X);
        DSV_HIP(hipGetLastError());
    });
}

}  // extern "C"
