// Matrix-free multi-species apply and the batched Jacobi-PCG solver.
//
// Replaces CeedDiffusionMassOp.apply (`gpu/ceed_operators.py:341-343`, whose QFunction is a
// different operator -- the apply here reproduces the ASSEMBLED operator of
// `gpu/gpu_assembler.py:60-165`) and the KSP solve of _solve_gpu_asm
// (`numerics/solvers.py:322-372`, GMRES(50)+bjacobi on a CSR that is copied through the host
// every step).
//
// Matrix-free layout (internal Morton vertex numbering):
//   block b of the partition owns vertices [b*NV, b*NV+NV) and lists every cell touching them;
//   one workgroup per block recomputes each cell's Jacobian from SoA coordinates, applies the
//   4x4 cell operator of all species at once and accumulates the owned rows in LDS (no global
//   atomics, no colouring); owned rows are then written with coalesced stores.  Workgroup ids
//   are remapped so that the 8 XCDs each get a contiguous (spatially compact) run of blocks,
//   which keeps halo vertex gathers in the XCD's own L2.
//
// CG: per iteration two kernels -- (1) the apply, computing p = z + beta p_old on the fly
// for gathered vertices, writing p and q = A p for owned vertices and the block partials of
// p.q; (2) the vertex update x += alpha p, r -= alpha q, z = r/diag with partials of r.z and
// r.r.  The last workgroup to finish each kernel reduces the partials in a fixed order and
// writes alpha / beta / the convergence flags, so an iteration never returns to the host;
// chunks of iterations are captured in a hipGraph and replayed until every species is done.
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

void check_species_coefs(const dsv_mesh* m, const dsv_species& s);
int coef_mode(const double* p, int32_t len, const dsv_mesh* m);
void assemble_into(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped, double* values,
                   hipStream_t st);

constexpr int MAXS = 3;

enum MFMode { MF_APPLY = 0, MF_DIAG = 1, MF_CG = 2, MF_RESID = 3 };

// status codes in CGState::status
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_BREAKDOWN = 2, ST_MAXIT = 3, ST_IDLE = 4 };

struct CGState {
    double rz[MAXS], alpha[MAXS], beta[MAXS], rr[MAXS], bb[MAXS], tol2[MAXS];
    int active[MAXS], iters[MAXS], status[MAXS];
    int max_it;
    unsigned ticket[4];
};

template <class T>
struct SpeciesP {
    const T* D;  // COEF_CELL: caller cell order; COEF_VERT: vertex order of the kernel
    const T* A;
    const uint8_t* bc;  // per-vertex Dirichlet mask (vertex order of the kernel) or null
    T Dval, Aval;
    int Dmode, Amode;
};

template <class T>
struct OpParams {
    SpeciesP<T> sp[MAXS];
    T inv_dt;
    int lumped;
    int any_cell;  // some coefficient is per cell: entries need their caller cell id
};

// vectors of the kernels, per species (vertex order of the kernel)
template <class T>
struct Vecs {
    T* x[MAXS];      // APPLY/RESID input x ; CG: z
    T* y[MAXS];      // APPLY output y ; DIAG output dinv ; CG: q ; RESID: r
    T* pold[MAXS];   // CG
    T* pnew[MAXS];   // CG
    const double* b[MAXS];  // RESID right-hand side (FP64)
};

struct Geo {
    int V, NV, nb, per_xcd;
    const int32_t* bptr;
    const int4* bconn;
    const int32_t* belem;
};

template <class T>
__device__ __forceinline__ T coef_at(const T* p, int mode, T val, int cell, const int4& n) {
    if (mode == COEF_SCALAR) return val;
    if (mode == COEF_CELL) return p[cell];
    return (((p[n.x] + p[n.y]) + p[n.z]) + p[n.w]) * T(0.25);
}

template <class T>
__device__ __forceinline__ void lds_add(T* a, T v) {
    atomicAdd(a, v);
}

// ---------------------------------------------------------------------------------------
// matrix-free kernel, all modes
template <class T, int S, int MODE>
__global__ void __launch_bounds__(1024)
k_mf(Geo g, const T* __restrict__ X, const T* __restrict__ Y, const T* __restrict__ Z,
     const OpParams<T>* __restrict__ prm, Vecs<T> v, CGState* __restrict__ st,
     double* __restrict__ partials) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* acc = reinterpret_cast<T*>(smem_raw);
    __shared__ double red[16 * 2 * MAXS];
    const int bid = blockIdx.x;
    const int b = (bid & 7) * g.per_xcd + (bid >> 3);
    const int nbl = gridDim.x;

    bool act[S];
    T beta[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = (MODE == MF_CG) ? (st->active[s] != 0) : true;
        beta[s] = (MODE == MF_CG) ? (T)st->beta[s] : T(0);
    }
    bool any = false;
#pragma unroll
    for (int s = 0; s < S; ++s) any |= act[s];
    if (MODE == MF_CG && !any) return;  // every species converged: nothing to do

    double part[S];
    double part2[S];
#pragma unroll
    for (int s = 0; s < S; ++s) part[s] = part2[s] = 0.0;

    const T inv_dt = prm->inv_dt;
    const int lumped = prm->lumped;

    if (b < g.nb) {
        const int v0 = b * g.NV;
        const int nv = min(g.NV, g.V - v0);
        for (int i = threadIdx.x; i < S * g.NV; i += blockDim.x) acc[i] = T(0);
        __syncthreads();
        const int kend = g.bptr[b + 1];
        for (int k = g.bptr[b] + threadIdx.x; k < kend; k += blockDim.x) {
            const int4 n = g.bconn[k];
            const Vec3<T> p0{X[n.x], Y[n.x], Z[n.x]};
            const Vec3<T> e1{X[n.y] - p0.x, Y[n.y] - p0.y, Z[n.y] - p0.z};
            const Vec3<T> e2{X[n.z] - p0.x, Y[n.z] - p0.y, Z[n.z] - p0.z};
            const Vec3<T> e3{X[n.w] - p0.x, Y[n.w] - p0.y, Z[n.w] - p0.z};
            const Vec3<T> c1 = cross(e2, e3), c2 = cross(e3, e1), c3 = cross(e1, e2);
            const T adet = fabs(dot(e1, c1));
            const T r6 = T(1) / (T(6) * adet);
            const int l0 = n.x - v0, l1 = n.y - v0, l2 = n.z - v0, l3 = n.w - v0;
            const bool o0 = (unsigned)l0 < (unsigned)nv, o1 = (unsigned)l1 < (unsigned)nv;
            const bool o2 = (unsigned)l2 < (unsigned)nv, o3 = (unsigned)l3 < (unsigned)nv;
            const int cell = prm->any_cell ? g.belem[k] : 0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                const SpeciesP<T>& sp = prm->sp[s];
                const T De = coef_at(sp.D, sp.Dmode, sp.Dval, cell, n);
                const T al = coef_at(sp.A, sp.Amode, sp.Aval, cell, n);
                const T w = De * r6;
                const T mf = inv_dt + al;
                if (MODE == MF_DIAG) {
                    const Vec3<T> c0{-(c1.x + c2.x + c3.x), -(c1.y + c2.y + c3.y),
                                     -(c1.z + c2.z + c3.z)};
                    const T md = lumped ? mf * adet * T(5.0 / 120.0) : mf * adet * T(2.0 / 120.0);
                    T* a = acc + s * g.NV;
                    if (o0) lds_add(a + l0, w * dot(c0, c0) + md);
                    if (o1) lds_add(a + l1, w * dot(c1, c1) + md);
                    if (o2) lds_add(a + l2, w * dot(c2, c2) + md);
                    if (o3) lds_add(a + l3, w * dot(c3, c3) + md);
                } else {
                    T u0, u1, u2, u3;
                    if (MODE == MF_CG) {
                        const T* zz = v.x[s];
                        const T* pp = v.pold[s];
                        const T bt = beta[s];
                        u0 = zz[n.x] + bt * pp[n.x];
                        u1 = zz[n.y] + bt * pp[n.y];
                        u2 = zz[n.z] + bt * pp[n.z];
                        u3 = zz[n.w] + bt * pp[n.w];
                    } else {
                        const T* xx = v.x[s];
                        u0 = xx[n.x];
                        u1 = xx[n.y];
                        u2 = xx[n.z];
                        u3 = xx[n.w];
                    }
                    const T d1 = u1 - u0, d2 = u2 - u0, d3 = u3 - u0;
                    const Vec3<T> gr{d1 * c1.x + d2 * c2.x + d3 * c3.x,
                                     d1 * c1.y + d2 * c2.y + d3 * c3.y,
                                     d1 * c1.z + d2 * c2.z + d3 * c3.z};
                    const T y1 = w * dot(c1, gr), y2 = w * dot(c2, gr), y3 = w * dot(c3, gr);
                    const T y0 = -(y1 + y2 + y3);
                    T m0, m1, m2, m3;
                    if (lumped) {
                        const T ml = mf * adet * T(5.0 / 120.0);
                        m0 = ml * u0;
                        m1 = ml * u1;
                        m2 = ml * u2;
                        m3 = ml * u3;
                    } else {
                        const T mc = mf * adet * T(1.0 / 120.0);
                        const T su = ((u0 + u1) + u2) + u3;
                        m0 = mc * (su + u0);
                        m1 = mc * (su + u1);
                        m2 = mc * (su + u2);
                        m3 = mc * (su + u3);
                    }
                    T* a = acc + s * g.NV;
                    if (o0) lds_add(a + l0, y0 + m0);
                    if (o1) lds_add(a + l1, y1 + m1);
                    if (o2) lds_add(a + l2, y2 + m2);
                    if (o3) lds_add(a + l3, y3 + m3);
                }
            }
        }
        __syncthreads();
        for (int lv = threadIdx.x; lv < nv; lv += blockDim.x) {
            const int vtx = v0 + lv;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                const uint8_t* bc = prm->sp[s].bc;
                const bool isbc = bc && bc[vtx];
                T val = acc[s * g.NV + lv];
                if (MODE == MF_APPLY) {
                    if (isbc) val = v.x[s][vtx];
                    v.y[s][vtx] = val;
                } else if (MODE == MF_DIAG) {
                    v.y[s][vtx] = isbc ? T(1) : T(1) / val;
                } else if (MODE == MF_CG) {
                    const T p = v.x[s][vtx] + beta[s] * v.pold[s][vtx];
                    if (isbc) val = p;
                    v.pnew[s][vtx] = p;
                    v.y[s][vtx] = val;
                    part[s] += (double)p * (double)val;
                } else {  // MF_RESID
                    if (isbc) val = v.x[s][vtx];
                    const double bv = v.b[s][vtx];
                    const double r = bv - (double)val;
                    v.y[s][vtx] = (T)r;
                    part[s] += r * r;
                    part2[s] += bv * bv;
                }
            }
        }
    }
    if (MODE == MF_APPLY || MODE == MF_DIAG) return;

    // block partials -> last block -> scalars
    double tot[2 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = part[s];
        tot[S + s] = part2[s];
    }
    block_sum<2 * S>(tot, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 2 * S; ++q) store_partial(partials + q * nbl + bid, tot[q]);
    if (!last_block_arrive(&st->ticket[0], nbl)) return;
    double all[2 * S];
    reduce_partials<2 * S>(partials, nbl, nbl, all, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (MODE == MF_CG) {
                if (!act[s]) continue;
                const double pq = all[s];
                if (!(pq > 0.0) || !isfinite(pq)) {
                    st->alpha[s] = 0.0;
                    st->active[s] = 0;
                    st->status[s] = ST_BREAKDOWN;
                } else {
                    st->alpha[s] = st->rz[s] / pq;
                }
            } else {
                st->rr[s] = all[s];
                st->bb[s] = all[S + s];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// CSR (caller order) kernels of the plain FP64 path: CG apply and residual.
constexpr int CSR_LANES = 8;

template <int S, int MODE>
__global__ void __launch_bounds__(512)
k_csr(int V, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
      const double* const* __restrict__ vals_unused, Vecs<double> v, const double* __restrict__ A0,
      const double* __restrict__ A1, const double* __restrict__ A2, CGState* __restrict__ st,
      double* __restrict__ partials) {
    __shared__ double red[16 * 2 * MAXS];
    const double* Av[3] = {A0, A1, A2};
    const int nbl = gridDim.x;
    bool act[S];
    double beta[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = (MODE == MF_CG) ? (st->active[s] != 0) : true;
        beta[s] = (MODE == MF_CG) ? st->beta[s] : 0.0;
    }
    bool any = false;
#pragma unroll
    for (int s = 0; s < S; ++s) any |= act[s];
    if (MODE == MF_CG && !any) return;
    double part[S], part2[S];
#pragma unroll
    for (int s = 0; s < S; ++s) part[s] = part2[s] = 0.0;

    const int lane = threadIdx.x & (CSR_LANES - 1);
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) / CSR_LANES;
    if (row < V) {
        const int js = indptr[row], je = indptr[row + 1];
        double acc[S];
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] = 0.0;
        for (int j = js + lane; j < je; j += CSR_LANES) {
            const int c = indices[j];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                double u;
                if (MODE == MF_CG)
                    u = v.x[s][c] + beta[s] * v.pold[s][c];
                else
                    u = v.x[s][c];
                acc[s] += Av[s][j] * u;
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int off = CSR_LANES / 2; off > 0; off >>= 1) acc[s] += __shfl_xor(acc[s], off, 64);
        if (lane == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                if (MODE == MF_CG) {
                    const double p = v.x[s][row] + beta[s] * v.pold[s][row];
                    v.pnew[s][row] = p;
                    v.y[s][row] = acc[s];
                    part[s] += p * acc[s];
                } else {
                    const double bv = v.b[s][row];
                    const double r = bv - acc[s];
                    v.y[s][row] = r;
                    part[s] += r * r;
                    part2[s] += bv * bv;
                }
            }
        }
    }
    double tot[2 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = part[s];
        tot[S + s] = part2[s];
    }
    block_sum<2 * S>(tot, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 2 * S; ++q) store_partial(partials + q * nbl + blockIdx.x, tot[q]);
    if (!last_block_arrive(&st->ticket[0], nbl)) return;
    double all[2 * S];
    reduce_partials<2 * S>(partials, nbl, nbl, all, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (MODE == MF_CG) {
                if (!act[s]) continue;
                const double pq = all[s];
                if (!(pq > 0.0) || !isfinite(pq)) {
                    st->alpha[s] = 0.0;
                    st->active[s] = 0;
                    st->status[s] = ST_BREAKDOWN;
                } else {
                    st->alpha[s] = st->rz[s] / pq;
                }
            } else {
                st->rr[s] = all[s];
                st->bb[s] = all[S + s];
            }
        }
    }
}

// CG vector update: x += alpha p; r -= alpha q; z = r * dinv; partials r.z, r.r.
// INIT variant (alpha ignored): r = (T) rin; x = 0 when zero_x; z = r * dinv.
template <class T, int S, bool INIT>
__global__ void __launch_bounds__(256)
k_update(int V, T* const* __restrict__ unused, Vecs<T> v, T* x0, T* x1, T* x2, T* r0, T* r1,
         T* r2, const T* dv0, const T* dv1, const T* dv2, const double* rin0, const double* rin1,
         const double* rin2, int zero_x, CGState* __restrict__ st, double* __restrict__ partials) {
    __shared__ double red[16 * 2 * MAXS];
    T* xs[3] = {x0, x1, x2};
    T* rs[3] = {r0, r1, r2};
    const T* ds[3] = {dv0, dv1, dv2};
    const double* ri[3] = {rin0, rin1, rin2};
    const int nbl = gridDim.x;
    bool act[S];
    T alpha[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = st->active[s] != 0;
        alpha[s] = INIT ? T(0) : (T)st->alpha[s];
    }
    bool any = false;
#pragma unroll
    for (int s = 0; s < S; ++s) any |= act[s];
    if (!any) return;
    double prz[S], prr[S];
#pragma unroll
    for (int s = 0; s < S; ++s) prz[s] = prr[s] = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!act[s]) continue;
            T r;
            if (INIT) {
                r = (T)ri[s][i];
                if (zero_x) xs[s][i] = T(0);
                v.pold[s][i] = T(0);
            } else {
                const T a = alpha[s];
                xs[s][i] += a * v.pnew[s][i];
                r = rs[s][i] - a * v.y[s][i];
            }
            rs[s][i] = r;
            const T z = r * ds[s][i];
            v.x[s][i] = z;
            prz[s] += (double)r * (double)z;
            prr[s] += (double)r * (double)r;
        }
    }
    double tot[2 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = prz[s];
        tot[S + s] = prr[s];
    }
    block_sum<2 * S>(tot, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 2 * S; ++q) store_partial(partials + q * nbl + blockIdx.x, tot[q]);
    if (!last_block_arrive(&st->ticket[1], nbl)) return;
    double all[2 * S];
    reduce_partials<2 * S>(partials, nbl, nbl, all, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!act[s]) continue;
            const double rz = all[s], rr = all[S + s];
            if (INIT) {
                st->rz[s] = rz;
                st->beta[s] = 0.0;
                st->rr[s] = rr;
                if (rr <= st->tol2[s]) {
                    st->active[s] = 0;
                    st->status[s] = ST_CONVERGED;
                }
                continue;
            }
            st->iters[s] += 1;
            st->rr[s] = rr;
            if (!isfinite(rr)) {
                st->active[s] = 0;
                st->status[s] = ST_BREAKDOWN;
            } else if (rr <= st->tol2[s]) {
                st->active[s] = 0;
                st->status[s] = ST_CONVERGED;
            } else if (st->iters[s] >= st->max_it) {
                st->active[s] = 0;
                st->status[s] = ST_MAXIT;
            } else {
                st->beta[s] = rz / st->rz[s];
                st->rz[s] = rz;
            }
        }
    }
}

// x64 += (double) d over the species in `mask` (refinement correction)
template <class T, int S>
__global__ void k_add_correction(int V, double* x0, double* x1, double* x2, const T* d0,
                                 const T* d1, const T* d2, int mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    double* xs[3] = {x0, x1, x2};
    const T* ds[3] = {d0, d1, d2};
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (mask & (1 << s)) xs[s][i] += (double)ds[s][i];
}

// caller order -> internal order (and back), with conversion
template <class Ti, class To>
__global__ void k_gather_perm(int V, const int32_t* __restrict__ perm, const Ti* __restrict__ in,
                              To* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    out[i] = (To)in[perm[i]];
}
template <class Ti, class To>
__global__ void k_scatter_perm(int V, const int32_t* __restrict__ perm, const Ti* __restrict__ in,
                               To* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    out[perm[i]] = (To)in[i];
}
template <class Ti, class To>
__global__ void k_convert(int64_t n, const Ti* __restrict__ in, To* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (To)in[i];
}
__global__ void k_set_bc_mask(int n_bc, const int32_t* __restrict__ bc,
                              const int32_t* __restrict__ iperm, uint8_t* __restrict__ mask) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bc) return;
    mask[iperm ? iperm[bc[k]] : bc[k]] = 1;
}
__global__ void k_set_bc_values(int n_bc, const int32_t* __restrict__ bc,
                                const int32_t* __restrict__ iperm, double value,
                                double* __restrict__ x) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bc) return;
    x[iperm ? iperm[bc[k]] : bc[k]] = value;
}
__global__ void k_diag_inv_csr(int V, const int32_t* __restrict__ diag_pos,
                               const double* __restrict__ A, double* __restrict__ dinv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    dinv[i] = 1.0 / A[diag_pos[i]];
}

// ---------------------------------------------------------------------------------------
// workspace

struct Workspace {
    int method = 0, S = 0, V = 0;
    hipStream_t stream = nullptr;
    CGState* st = nullptr;
    CGState* h_st = nullptr;  // pinned
    double* partials = nullptr;
    int max_grid = 0;
    // coefficient copies (target precision) and masks
    void* Dc[MAXS] = {};
    void* Ac[MAXS] = {};
    uint8_t* bcmask[MAXS] = {};
    void* prm = nullptr;   // OpParams<T> of the CG precision, on device
    void* prm2 = nullptr;  // OpParams<double> of the refinement residual (MF32_IR64)
    void* Dc2[MAXS] = {};
    void* Ac2[MAXS] = {};
    // vectors (FP64 unless noted)
    double* b64[MAXS] = {};
    double* x64[MAXS] = {};
    double* r64[MAXS] = {};
    // CG vectors in the CG precision
    void* xcg[MAXS] = {};
    void* rcg[MAXS] = {};
    void* zcg[MAXS] = {};
    void* qcg[MAXS] = {};
    void* pa[MAXS] = {};
    void* pb[MAXS] = {};
    void* dinv[MAXS] = {};
    double* csr[MAXS] = {};  // CSR64 values
    // graphs: [parity of p buffers]
    hipGraphExec_t graph[2] = {};
    int graph_chunk = 0;
    int graph_valid = 0;
};

static size_t tsize(int method) { return method == DSV_METHOD_MF32_IR64 ? sizeof(float) : sizeof(double); }

static void free_ws(Workspace* w) {
    if (!w) return;
    for (int k = 0; k < 2; ++k)
        if (w->graph[k]) (void)hipGraphExecDestroy(w->graph[k]);
    dfree(w->st);
    if (w->h_st) (void)hipHostFree(w->h_st);
    dfree(w->partials);
    for (int s = 0; s < MAXS; ++s) {
        dfree(w->Dc[s]);
        dfree(w->Ac[s]);
        dfree(w->Dc2[s]);
        dfree(w->Ac2[s]);
        dfree(w->bcmask[s]);
        dfree(w->b64[s]);
        dfree(w->x64[s]);
        dfree(w->r64[s]);
        dfree(w->xcg[s]);
        dfree(w->rcg[s]);
        dfree(w->zcg[s]);
        dfree(w->qcg[s]);
        dfree(w->pa[s]);
        dfree(w->pb[s]);
        dfree(w->dinv[s]);
        dfree(w->csr[s]);
    }
    dfree(w->prm);
    dfree(w->prm2);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
}

void free_workspaces(dsv_mesh* m) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 4; ++b) {
            free_ws(m->ws[a][b]);
            m->ws[a][b] = nullptr;
        }
    dfree(m->Ae);
}

static int mf_grid(const dsv_mesh* m) { return 8 * ((m->part.nb + 7) / 8); }
static int csr_grid(const dsv_mesh* m) { return (int)cdiv((int64_t)m->V * CSR_LANES, 512); }
static int upd_grid(const dsv_mesh* m) { return std::max(1, std::min((int)cdiv(m->V, 256 * 4), 4096)); }

static Workspace* get_ws(dsv_mesh* m, int method, int S) {
    Workspace*& w = m->ws[method][S];
    if (w) return w;
    w = new Workspace();
    w->method = method;
    w->S = S;
    w->V = m->V;
    try {
        DSV_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
        w->st = dalloc<CGState>(1);
        DSV_HIP(hipMemset(w->st, 0, sizeof(CGState)));
        DSV_HIP(hipHostMalloc((void**)&w->h_st, sizeof(CGState), hipHostMallocDefault));
        w->max_grid = std::max({mf_grid(m), csr_grid(m), upd_grid(m)});
        w->partials = dalloc<double>((size_t)2 * MAXS * w->max_grid);
        const size_t ts = tsize(method);
        const size_t V = m->V;
        for (int s = 0; s < S; ++s) {
            w->bcmask[s] = dalloc<uint8_t>(V);
            w->b64[s] = dalloc<double>(V);
            w->x64[s] = dalloc<double>(V);
            w->r64[s] = dalloc<double>(V);
            DSV_HIP(hipMalloc(&w->xcg[s], V * ts));
            DSV_HIP(hipMalloc(&w->rcg[s], V * ts));
            DSV_HIP(hipMalloc(&w->zcg[s], V * ts));
            DSV_HIP(hipMalloc(&w->qcg[s], V * ts));
            DSV_HIP(hipMalloc(&w->pa[s], V * ts));
            DSV_HIP(hipMalloc(&w->pb[s], V * ts));
            DSV_HIP(hipMalloc(&w->dinv[s], V * ts));
            if (method == DSV_METHOD_CSR64) w->csr[s] = dalloc<double>(m->nnz);
        }
        DSV_HIP(hipMalloc(&w->prm, std::max(sizeof(OpParams<float>), sizeof(OpParams<double>))));
        DSV_HIP(hipMalloc(&w->prm2, sizeof(OpParams<double>)));
    } catch (...) {
        free_ws(w);
        w = nullptr;
        throw;
    }
    return w;
}

// Copies / converts the coefficients of every species into the workspace (precision T,
// vertex arrays in internal order) and writes the device OpParams.
template <class T>
static void prepare_op(dsv_mesh* m, Workspace* w, int S, const dsv_species* sp, double inv_dt,
                       int lumped, bool internal, hipStream_t st, bool second = false) {
    OpParams<T> h{};
    h.inv_dt = (T)inv_dt;
    h.lumped = lumped ? 1 : 0;
    h.any_cell = 0;
    for (int s = 0; s < S; ++s) {
        SpeciesP<T>& q = h.sp[s];
        q.Dmode = coef_mode(sp[s].D, sp[s].D_len, m);
        q.Amode = coef_mode(sp[s].alpha, sp[s].alpha_len, m);
        q.Dval = (T)sp[s].D_value;
        q.Aval = sp[s].alpha ? T(0) : (T)sp[s].alpha_value;
        h.any_cell |= (q.Dmode == COEF_CELL || q.Amode == COEF_CELL) ? 1 : 0;
        const double* src[2] = {sp[s].D, sp[s].alpha};
        const int32_t lens[2] = {sp[s].D_len, sp[s].alpha_len};
        void** dst[2] = {second ? &w->Dc2[s] : &w->Dc[s], second ? &w->Ac2[s] : &w->Ac[s]};
        const T** out[2] = {&q.D, &q.A};
        for (int c = 0; c < 2; ++c) {
            *out[c] = nullptr;
            if (!src[c]) continue;
            const size_t n = (lens[c] == m->C) ? (size_t)m->C : (size_t)m->V;
            if (!*dst[c]) DSV_HIP(hipMalloc(dst[c], std::max((size_t)m->C, (size_t)m->V) * sizeof(T)));
            T* d = (T*)*dst[c];
            if (lens[c] == m->V && internal)
                hipLaunchKernelGGL((k_gather_perm<double, T>), dim3(cdiv(m->V, 256)), dim3(256), 0, st,
                                   m->V, m->perm, src[c], d);
            else
                hipLaunchKernelGGL((k_convert<double, T>), dim3(cdiv(n, 256)), dim3(256), 0, st,
                                   (int64_t)n, src[c], d);
            *out[c] = d;
        }
        if (!second) DSV_HIP(hipMemsetAsync(w->bcmask[s], 0, m->V, st));
        if (sp[s].n_bc > 0) {
            if (!second)
                hipLaunchKernelGGL(k_set_bc_mask, dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0, st,
                                   sp[s].n_bc, sp[s].bc_dofs,
                                   internal ? m->iperm : (const int32_t*)nullptr, w->bcmask[s]);
            q.bc = w->bcmask[s];
        } else {
            q.bc = nullptr;
        }
    }
    DSV_HIP(hipMemcpyAsync(second ? w->prm2 : w->prm, &h, sizeof(h), hipMemcpyHostToDevice, st));
    DSV_HIP(hipStreamSynchronize(st));  // h lives on this stack frame
    DSV_HIP(hipGetLastError());
}

static Geo make_geo(const dsv_mesh* m) {
    Geo g;
    g.V = m->V;
    g.NV = m->part.NV;
    g.nb = m->part.nb;
    g.per_xcd = (m->part.nb + 7) / 8;
    g.bptr = m->part.bptr;
    g.bconn = m->part.bconn;
    g.belem = m->part.belem;
    return g;
}

template <class T, int S, int MODE>
static void launch_mf(const dsv_mesh* m, const void* prm, const Vecs<T>& v, CGState* st,
                      double* partials, hipStream_t stream) {
    const Geo g = make_geo(m);
    const T* X = nullptr;
    if (sizeof(T) == 4)
        X = (const T*)m->xyz32;
    else
        X = (const T*)m->xyz64;
    const size_t lds = (size_t)S * m->part.NV * sizeof(T);
    hipLaunchKernelGGL((k_mf<T, S, MODE>), dim3(mf_grid(m)), dim3(m->part.threads), lds, stream, g, X,
                       X + m->V, X + 2 * (size_t)m->V, (const OpParams<T>*)prm, v, st, partials);
}

template <class T, int MODE>
static void launch_mf_s(int S, const dsv_mesh* m, const void* prm, const Vecs<T>& v, CGState* st,
                        double* partials, hipStream_t stream) {
    switch (S) {
        case 1: launch_mf<T, 1, MODE>(m, prm, v, st, partials, stream); break;
        case 2: launch_mf<T, 2, MODE>(m, prm, v, st, partials, stream); break;
        default: launch_mf<T, 3, MODE>(m, prm, v, st, partials, stream); break;
    }
}

template <int MODE>
static void launch_csr_s(int S, const dsv_mesh* m, Workspace* w, const Vecs<double>& v,
                         hipStream_t stream) {
    const int grid = csr_grid(m);
    const double* A0 = w->csr[0];
    const double* A1 = S > 1 ? w->csr[1] : A0;
    const double* A2 = S > 2 ? w->csr[2] : A0;
    switch (S) {
        case 1:
            hipLaunchKernelGGL((k_csr<1, MODE>), dim3(grid), dim3(512), 0, stream, m->V, m->indptr,
                               m->indices, (const double* const*)nullptr, v, A0, A1, A2, w->st,
                               w->partials);
            break;
        case 2:
            hipLaunchKernelGGL((k_csr<2, MODE>), dim3(grid), dim3(512), 0, stream, m->V, m->indptr,
                               m->indices, (const double* const*)nullptr, v, A0, A1, A2, w->st,
                               w->partials);
            break;
        default:
            hipLaunchKernelGGL((k_csr<3, MODE>), dim3(grid), dim3(512), 0, stream, m->V, m->indptr,
                               m->indices, (const double* const*)nullptr, v, A0, A1, A2, w->st,
                               w->partials);
            break;
    }
}

template <class T, bool INIT>
static void launch_update(int S, const dsv_mesh* m, Workspace* w, const Vecs<T>& v,
                          const double* const* rin, int zero_x, hipStream_t stream) {
    T* x[3];
    T* r[3];
    const T* d[3];
    const double* ri[3];
    for (int s = 0; s < 3; ++s) {
        const int q = s < S ? s : 0;
        x[s] = (T*)w->xcg[q];
        r[s] = (T*)w->rcg[q];
        d[s] = (const T*)w->dinv[q];
        ri[s] = rin ? rin[q] : nullptr;
    }
    const int grid = upd_grid(m);
    switch (S) {
        case 1:
            hipLaunchKernelGGL((k_update<T, 1, INIT>), dim3(grid), dim3(256), 0, stream, m->V,
                               (T* const*)nullptr, v, x[0], x[1], x[2], r[0], r[1], r[2], d[0], d[1],
                               d[2], ri[0], ri[1], ri[2], zero_x, w->st, w->partials);
            break;
        case 2:
            hipLaunchKernelGGL((k_update<T, 2, INIT>), dim3(grid), dim3(256), 0, stream, m->V,
                               (T* const*)nullptr, v, x[0], x[1], x[2], r[0], r[1], r[2], d[0], d[1],
                               d[2], ri[0], ri[1], ri[2], zero_x, w->st, w->partials);
            break;
        default:
            hipLaunchKernelGGL((k_update<T, 3, INIT>), dim3(grid), dim3(256), 0, stream, m->V,
                               (T* const*)nullptr, v, x[0], x[1], x[2], r[0], r[1], r[2], d[0], d[1],
                               d[2], ri[0], ri[1], ri[2], zero_x, w->st, w->partials);
            break;
    }
}

// vectors of one CG iteration: z in x[], q in y[], p_old / p_new alternate between pa / pb
template <class T>
static Vecs<T> cg_vecs(Workspace* w, int S, int parity) {
    Vecs<T> v{};
    for (int s = 0; s < MAXS; ++s) {
        const int q = s < S ? s : 0;
        v.x[s] = (T*)w->zcg[q];
        v.y[s] = (T*)w->qcg[q];
        v.pold[s] = (T*)(parity ? w->pb[q] : w->pa[q]);
        v.pnew[s] = (T*)(parity ? w->pa[q] : w->pb[q]);
        v.b[s] = nullptr;
    }
    return v;
}

// one CG iteration (apply + update) with p buffers of the given parity
template <class T>
static void cg_iteration(dsv_mesh* m, Workspace* w, int S, int parity, hipStream_t stream) {
    Vecs<T> v = cg_vecs<T>(w, S, parity);
    if constexpr (std::is_same<T, double>::value) {
        if (w->method == DSV_METHOD_CSR64)
            launch_csr_s<MF_CG>(S, m, w, v, stream);
        else
            launch_mf_s<T, MF_CG>(S, m, w->prm, v, w->st, w->partials, stream);
    } else {
        launch_mf_s<T, MF_CG>(S, m, w->prm, v, w->st, w->partials, stream);
    }
    launch_update<T, false>(S, m, w, v, nullptr, 0, stream);
}

// Runs CG iterations until no species is active; returns the number of iterations launched.
// Chunks of an even number of iterations are captured once into a hipGraph and replayed.
template <class T>
static int run_cg(dsv_mesh* m, Workspace* w, int S, int max_it, hipStream_t stream) {
    const int chunk = std::max(2, m->graph_chunk & ~1);
    int launched = 0;
    if (m->use_graphs && (!w->graph_valid || w->graph_chunk != chunk)) {
        if (w->graph[0]) (void)hipGraphExecDestroy(w->graph[0]);
        w->graph[0] = nullptr;
        hipGraph_t gr;
        DSV_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < chunk; ++i) cg_iteration<T>(m, w, S, i & 1, stream);
        DSV_HIP(hipStreamEndCapture(stream, &gr));
        DSV_HIP(hipGraphInstantiate(&w->graph[0], gr, nullptr, nullptr, 0));
        DSV_HIP(hipGraphDestroy(gr));
        w->graph_valid = 1;
        w->graph_chunk = chunk;
    }
    while (true) {
        DSV_HIP(hipMemcpyAsync(w->h_st, w->st, sizeof(CGState), hipMemcpyDeviceToHost, stream));
        DSV_HIP(hipStreamSynchronize(stream));
        bool any = false;
        for (int s = 0; s < S; ++s) any |= (w->h_st->active[s] != 0);
        if (!any || launched >= max_it + chunk) break;
        if (m->use_graphs) {
            DSV_HIP(hipGraphLaunch(w->graph[0], stream));
        } else {
            for (int i = 0; i < chunk; ++i) cg_iteration<T>(m, w, S, i & 1, stream);
        }
        launched += chunk;
    }
    return launched;
}

// host-side state setup before a CG run
static void set_state(Workspace* w, int S, const double* tol2, const int* active, int max_it,
                      bool reset_iters, hipStream_t stream) {
    DSV_HIP(hipMemcpyAsync(w->h_st, w->st, sizeof(CGState), hipMemcpyDeviceToHost, stream));
    DSV_HIP(hipStreamSynchronize(stream));
    CGState h = *w->h_st;
    for (int s = 0; s < MAXS; ++s) {
        h.tol2[s] = s < S ? tol2[s] : 0.0;
        h.active[s] = s < S ? active[s] : 0;
        h.status[s] = (s < S && active[s]) ? ST_RUNNING : ST_IDLE;
        if (reset_iters) h.iters[s] = 0;
        h.beta[s] = 0.0;
    }
    h.max_it = max_it;
    h.ticket[0] = h.ticket[1] = 0;
    *w->h_st = h;
    DSV_HIP(hipMemcpyAsync(w->st, w->h_st, sizeof(CGState), hipMemcpyHostToDevice, stream));
}

static CGState read_state(Workspace* w, hipStream_t stream) {
    DSV_HIP(hipMemcpyAsync(w->h_st, w->st, sizeof(CGState), hipMemcpyDeviceToHost, stream));
    DSV_HIP(hipStreamSynchronize(stream));
    return *w->h_st;
}


// ---------------------------------------------------------------------------------------
// residual r64 = b - A x64 (FP64, vertex order of the method) -> st->rr, st->bb
static void residual64(dsv_mesh* m, Workspace* w, int S, hipStream_t stream) {
    Vecs<double> v{};
    for (int s = 0; s < MAXS; ++s) {
        const int q = s < S ? s : 0;
        v.x[s] = w->x64[q];
        v.y[s] = w->r64[q];
        v.b[s] = w->b64[q];
    }
    if (w->method == DSV_METHOD_CSR64)
        launch_csr_s<MF_RESID>(S, m, w, v, stream);
    else
        launch_mf_s<double, MF_RESID>(S, m, w->method == DSV_METHOD_MF32_IR64 ? w->prm2 : w->prm,
                                      v, w->st, w->partials, stream);
}

}  // namespace dsv

namespace dsv {

static void check_solve_args(const dsv_mesh* m, int S, const dsv_species* sp, bool need_bx) {
    if (!m || !sp) throw ArgError{"NULL mesh or species"};
    if (S < 1 || S > MAXS) throw ArgError{"n_species must be 1, 2 or 3"};
    for (int s = 0; s < S; ++s) {
        check_species_coefs(m, sp[s]);  // also range-checks the Dirichlet indices
        if (need_bx && (!sp[s].b || !sp[s].x)) throw ArgError{"species b and x must be set"};
    }
}

template <class T>
static void add_correction(int S, dsv_mesh* m, Workspace* w, int mask, hipStream_t st) {
    double* x[3];
    const T* d[3];
    for (int s = 0; s < 3; ++s) {
        const int q = s < S ? s : 0;
        x[s] = w->x64[q];
        d[s] = (const T*)w->xcg[q];
    }
    const unsigned grid = cdiv(m->V, 256);
    switch (S) {
        case 1: hipLaunchKernelGGL((k_add_correction<T, 1>), dim3(grid), dim3(256), 0, st, m->V, x[0], x[1], x[2], d[0], d[1], d[2], mask); break;
        case 2: hipLaunchKernelGGL((k_add_correction<T, 2>), dim3(grid), dim3(256), 0, st, m->V, x[0], x[1], x[2], d[0], d[1], d[2], mask); break;
        default: hipLaunchKernelGGL((k_add_correction<T, 3>), dim3(grid), dim3(256), 0, st, m->V, x[0], x[1], x[2], d[0], d[1], d[2], mask); break;
    }
}

// inner CG on A d = r64 (d starts at 0), T = CG precision
template <class T>
static void inner_cg(dsv_mesh* m, Workspace* w, int S, const double* tol2, const int* active,
                     int max_it, hipStream_t st) {
    set_state(w, S, tol2, active, max_it, false, st);
    Vecs<T> v = cg_vecs<T>(w, S, 0);
    const double* rin[3];
    for (int s = 0; s < 3; ++s) rin[s] = w->r64[s < S ? s : 0];
    launch_update<T, true>(S, m, w, v, rin, 1, st);
    DSV_HIP(hipGetLastError());
    run_cg<T>(m, w, S, max_it, st);
}

static void solve_impl(dsv_mesh* m, int S, const dsv_species* sp, double inv_dt, int lumped,
                       int method, double rtol, int max_it, hipStream_t user, dsv_stats* stats) {
    if (method < 0 || method > 2) throw ArgError{"unknown method"};
    if (!(rtol > 0.0)) throw ArgError{"rtol must be > 0"};
    if (max_it < 1) throw ArgError{"max_it must be >= 1"};
    check_solve_args(m, S, sp, true);
    DSV_HIP(hipSetDevice(m->device));
    DSV_HIP(hipStreamSynchronize(user));
    Workspace* w = get_ws(m, method, S);
    hipStream_t st = w->stream;
    const auto t0 = std::chrono::steady_clock::now();
    const bool internal = method != DSV_METHOD_CSR64;
    const unsigned gv = cdiv(m->V, 256);
    for (int s = 0; s < S; ++s) {
        if (internal) {
            hipLaunchKernelGGL((k_gather_perm<double, double>), dim3(gv), dim3(256), 0, st, m->V, m->perm, sp[s].b, w->b64[s]);
            hipLaunchKernelGGL((k_gather_perm<double, double>), dim3(gv), dim3(256), 0, st, m->V, m->perm, (const double*)sp[s].x, w->x64[s]);
        } else {
            DSV_HIP(hipMemcpyAsync(w->b64[s], sp[s].b, m->V * sizeof(double), hipMemcpyDeviceToDevice, st));
            DSV_HIP(hipMemcpyAsync(w->x64[s], sp[s].x, m->V * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        if (sp[s].n_bc > 0)
            hipLaunchKernelGGL(k_set_bc_values, dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0, st, sp[s].n_bc,
                               sp[s].bc_dofs, internal ? m->iperm : (const int32_t*)nullptr, sp[s].bc_value,
                               w->x64[s]);
    }
    // operator set-up: assembled values / coefficients and the Jacobi diagonal
    if (method == DSV_METHOD_CSR64) {
        for (int s = 0; s < S; ++s) {
            assemble_into(m, sp[s], inv_dt, lumped, w->csr[s], st);
            hipLaunchKernelGGL(k_diag_inv_csr, dim3(gv), dim3(256), 0, st, m->V, m->diag_pos, w->csr[s],
                               (double*)w->dinv[s]);
        }
    } else if (method == DSV_METHOD_MF64) {
        prepare_op<double>(m, w, S, sp, inv_dt, lumped, true, st);
        Vecs<double> v{};
        for (int s = 0; s < MAXS; ++s) v.y[s] = (double*)w->dinv[s < S ? s : 0];
        launch_mf_s<double, MF_DIAG>(S, m, w->prm, v, w->st, w->partials, st);
    } else {
        prepare_op<float>(m, w, S, sp, inv_dt, lumped, true, st);
        prepare_op<double>(m, w, S, sp, inv_dt, lumped, true, st, true);
        Vecs<float> v{};
        for (int s = 0; s < MAXS; ++s) v.y[s] = (float*)w->dinv[s < S ? s : 0];
        launch_mf_s<float, MF_DIAG>(S, m, w->prm, v, w->st, w->partials, st);
    }
    DSV_HIP(hipGetLastError());

    double tol_abs[MAXS] = {0, 0, 0}, rel0[MAXS] = {0, 0, 0}, relf[MAXS] = {0, 0, 0};
    int conv[MAXS] = {0, 0, 0}, outer[MAXS] = {0, 0, 0}, dead[MAXS] = {0, 0, 0};
    const int zero_act[MAXS] = {0, 0, 0};
    const double zero_tol[MAXS] = {0, 0, 0};
    set_state(w, S, zero_tol, zero_act, max_it, true, st);
    const int max_rounds = method == DSV_METHOD_MF32_IR64 ? m->max_outer : 4;
    for (int round = 0;; ++round) {
        residual64(m, w, S, st);
        DSV_HIP(hipGetLastError());
        CGState h = read_state(w, st);
        bool all = true;
        for (int s = 0; s < S; ++s) {
            const double bn = std::sqrt(h.bb[s]), rn = std::sqrt(h.rr[s]);
            if (round == 0) {
                tol_abs[s] = rtol * bn;
                rel0[s] = bn > 0 ? rn / bn : rn;
            }
            relf[s] = bn > 0 ? rn / bn : rn;
            conv[s] = (rn <= tol_abs[s]) ? 1 : 0;
            if (h.status[s] == ST_BREAKDOWN || h.iters[s] >= max_it) dead[s] = 1;
            all &= (conv[s] || dead[s]);
        }
        if (all || round >= max_rounds) break;
        double tol2[MAXS] = {0, 0, 0};
        int act[MAXS] = {0, 0, 0};
        int mask = 0;
        for (int s = 0; s < S; ++s) {
            if (conv[s] || dead[s]) continue;
            const double rn = std::sqrt(h.rr[s]);
            double t = method == DSV_METHOD_MF32_IR64 ? std::max(m->inner_rtol * rn, 0.3 * tol_abs[s])
                                                      : 0.9 * tol_abs[s];
            tol2[s] = t * t;
            act[s] = 1;
            mask |= 1 << s;
            outer[s] += 1;
        }
        if (method == DSV_METHOD_MF32_IR64) {
            inner_cg<float>(m, w, S, tol2, act, max_it, st);
            add_correction<float>(S, m, w, mask, st);
        } else {
            inner_cg<double>(m, w, S, tol2, act, max_it, st);
            add_correction<double>(S, m, w, mask, st);
        }
        DSV_HIP(hipGetLastError());
    }
    CGState h = read_state(w, st);
    for (int s = 0; s < S; ++s) {
        if (internal)
            hipLaunchKernelGGL((k_scatter_perm<double, double>), dim3(gv), dim3(256), 0, st, m->V, m->perm,
                               (const double*)w->x64[s], sp[s].x);
        else
            DSV_HIP(hipMemcpyAsync(sp[s].x, w->x64[s], m->V * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    DSV_HIP(hipStreamSynchronize(st));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (int s = 0; s < S; ++s) {
            stats->iterations[s] = h.iters[s];
            stats->outer[s] = outer[s];
            stats->converged[s] = conv[s];
            stats->rel_residual[s] = relf[s];
            stats->initial_rel_residual[s] = rel0[s];
        }
        stats->solve_ms = ms;
    }
}

template <class T>
static void apply_impl(dsv_mesh* m, int S, const dsv_species* sp, double inv_dt, int lumped,
                       const double* const* x, double* const* y, hipStream_t user) {
    const int method = sizeof(T) == 4 ? DSV_METHOD_MF32_IR64 : DSV_METHOD_MF64;
    Workspace* w = get_ws(m, method, S);
    hipStream_t st = w->stream;
    DSV_HIP(hipStreamSynchronize(user));
    prepare_op<T>(m, w, S, sp, inv_dt, lumped, true, st);
    const unsigned gv = cdiv(m->V, 256);
    Vecs<T> v{};
    for (int s = 0; s < MAXS; ++s) {
        const int q = s < S ? s : 0;
        v.x[s] = (T*)w->zcg[q];
        v.y[s] = (T*)w->qcg[q];
    }
    for (int s = 0; s < S; ++s)
        hipLaunchKernelGGL((k_gather_perm<double, T>), dim3(gv), dim3(256), 0, st, m->V, m->perm, x[s], v.x[s]);
    launch_mf_s<T, MF_APPLY>(S, m, w->prm, v, w->st, w->partials, st);
    for (int s = 0; s < S; ++s)
        hipLaunchKernelGGL((k_scatter_perm<T, double>), dim3(gv), dim3(256), 0, st, m->V, m->perm,
                           (const T*)v.y[s], y[s]);
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipStreamSynchronize(st));
}

template <class T>
static double time_impl(dsv_mesh* m, int S, const dsv_species* sp, double inv_dt, int lumped,
                        int mode, int n_rep, hipStream_t stream) {
    const int method = sizeof(T) == 4 ? DSV_METHOD_MF32_IR64 : DSV_METHOD_MF64;
    Workspace* w = get_ws(m, method, S);
    DSV_HIP(hipStreamSynchronize(stream));
    prepare_op<T>(m, w, S, sp, inv_dt, lumped, true, w->stream);
    const unsigned gv = cdiv(m->V, 256);
    for (int s = 0; s < S; ++s) {
        // z = x (internal order) when x is given, else ones; p_old = 0
        if (sp[s].x)
            hipLaunchKernelGGL((k_gather_perm<double, T>), dim3(gv), dim3(256), 0, w->stream, m->V, m->perm,
                               (const double*)sp[s].x, (T*)w->zcg[s]);
        else
            hipLaunchKernelGGL((k_convert<double, T>), dim3(gv), dim3(256), 0, w->stream, (int64_t)m->V,
                               m->lumped_m, (T*)w->zcg[s]);
        DSV_HIP(hipMemsetAsync(w->pa[s], 0, m->V * sizeof(T), w->stream));
    }
    const double one[MAXS] = {0, 0, 0};
    const int act[MAXS] = {1, 1, 1};
    set_state(w, S, one, act, 1 << 30, true, w->stream);
    CGState h = read_state(w, w->stream);
    for (int s = 0; s < S; ++s) h.rz[s] = 1.0;
    *w->h_st = h;
    DSV_HIP(hipMemcpyAsync(w->st, w->h_st, sizeof(CGState), hipMemcpyHostToDevice, w->stream));
    DSV_HIP(hipStreamSynchronize(w->stream));
    Vecs<T> v = cg_vecs<T>(w, S, 0);
    auto launch = [&] {
        if (mode == 1)
            launch_mf_s<T, MF_CG>(S, m, w->prm, v, w->st, w->partials, stream);
        else
            launch_mf_s<T, MF_APPLY>(S, m, w->prm, v, w->st, w->partials, stream);
    };
    for (int i = 0; i < 3; ++i) launch();
    hipEvent_t e0, e1;
    DSV_HIP(hipEventCreate(&e0));
    DSV_HIP(hipEventCreate(&e1));
    DSV_HIP(hipEventRecord(e0, stream));
    for (int i = 0; i < n_rep; ++i) launch();
    DSV_HIP(hipEventRecord(e1, stream));
    DSV_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DSV_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    DSV_HIP(hipGetLastError());
    return (double)ms / n_rep;
}

}  // namespace dsv

using namespace dsv;

extern "C" {

int dsv_apply(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
              int32_t lumped, const double* const* x, double* const* y, int32_t precision,
              void* stream) {
    return guard([&] {
        check_solve_args(m, n_species, sp, false);
        if (!x || !y) throw ArgError{"dsv_apply: NULL x or y"};
        for (int s = 0; s < n_species; ++s)
            if (!x[s] || !y[s]) throw ArgError{"dsv_apply: NULL x[s] or y[s]"};
        if (precision != 32 && precision != 64) throw ArgError{"precision must be 32 or 64"};
        DSV_HIP(hipSetDevice(m->device));
        if (precision == 32)
            apply_impl<float>(m, n_species, sp, inv_dt, lumped, x, y, (hipStream_t)stream);
        else
            apply_impl<double>(m, n_species, sp, inv_dt, lumped, x, y, (hipStream_t)stream);
    });
}

int dsv_solve(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
              int32_t lumped, int32_t method, double rtol, int32_t max_it, void* stream,
              dsv_stats* stats) {
    int rc = guard([&] {
        solve_impl(m, n_species, sp, inv_dt, lumped, method, rtol, max_it, (hipStream_t)stream, stats);
    });
    return rc;
}

int dsv_set_solver_options(dsv_mesh* m, double inner_rtol, int32_t max_outer,
                           int32_t graph_chunk, int32_t use_graphs) {
    return guard([&] {
        if (!m) throw ArgError{"NULL mesh"};
        if (!(inner_rtol > 0.0 && inner_rtol < 1.0)) throw ArgError{"inner_rtol must be in (0, 1)"};
        if (max_outer < 1 || max_outer > 100) throw ArgError{"max_outer must be in [1, 100]"};
        if (graph_chunk < 2 || graph_chunk > 256) throw ArgError{"graph_chunk must be in [2, 256]"};
        m->inner_rtol = inner_rtol;
        m->max_outer = max_outer;
        m->graph_chunk = graph_chunk;
        m->use_graphs = use_graphs ? 1 : 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 4; ++b)
                if (m->ws[a][b]) m->ws[a][b]->graph_valid = 0;
    });
}

int dsv_time_apply(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t precision, int32_t mode, int32_t n_rep, void* stream,
                   double* ms) {
    return guard([&] {
        check_solve_args(m, n_species, sp, false);
        if (!ms || n_rep < 1) throw ArgError{"dsv_time_apply: bad arguments"};
        if (precision != 32 && precision != 64) throw ArgError{"precision must be 32 or 64"};
        DSV_HIP(hipSetDevice(m->device));
        *ms = precision == 32
                  ? time_impl<float>(m, n_species, sp, inv_dt, lumped, mode, n_rep, (hipStream_t)stream)
                  : time_impl<double>(m, n_species, sp, inv_dt, lumped, mode, n_rep, (hipStream_t)stream);
    });
}

}  // extern "C"
