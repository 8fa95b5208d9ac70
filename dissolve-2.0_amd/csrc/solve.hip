// Matrix-free multi-species apply and the batched Jacobi-PCG solver.
//
// Replaces CeedDiffusionMassOp.apply (`gpu/ceed_operators.py:341-343`, whose QFunction is a
// different operator -- the apply here reproduces the ASSEMBLED operator of
// `gpu/gpu_assembler.py:60-165`) and the KSP solve of _solve_gpu_asm
// (`numerics/solvers.py:322-372`, GMRES(50)+bjacobi on a CSR that is copied through the host
// every step).
//
// Matrix-free layout (internal Morton vertex numbering):
//   partition block b owns vertices [b*NV, b*NV+NV) and lists every cell touching them; a
//   workgroup recomputes each listed cell's Jacobian from SoA coordinates, applies the 4x4
//   cell operator of all species at once and accumulates the owned rows in LDS (no global
//   atomics, no colouring); owned rows are then written with coalesced stores.  The 8 XCDs
//   each get a contiguous (spatially compact) run of partition blocks, which keeps the halo
//   vertex gathers in the XCD's own L2.
//
// CG: per iteration two kernels -- (A) the apply: p = z + beta p_old computed on the fly for
// gathered vertices, p and q = A p written for owned vertices, block partials of p.q;
// (U) the vertex update: x += alpha p, r -= alpha q, z = r/diag, block partials of r.z and
// r.r.  Scalars never go through the host and no kernel uses a fence or a global atomic:
// every workgroup of a consumer kernel reduces ALL partials of the producer kernel in the
// same fixed order, so all workgroups derive identical alpha / beta / convergence decisions;
// workgroup 0 records them in the state block for the host.  Chunks of iterations are
// captured in a hipGraph and replayed until every species has converged.
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
constexpr int MAX_PROD = 1024;  // max workgroups of a partial-producing kernel

enum MFMode { MF_APPLY = 0, MF_DIAG = 1, MF_CG = 2, MF_RESID = 3 };

// status codes in CGState::status
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_BREAKDOWN = 2, ST_MAXIT = 3, ST_IDLE = 4 };

// Written by workgroup 0 of the kernel named in the comment, read by later kernels / host.
struct CGState {
    double tol2[MAXS];     // host
    double rz[2][MAXS];    // A_k -> rz[k&1] = r_k.z_k
    double rr[MAXS];       // A_k -> r_k.r_k (host)
    int act[2][MAXS];      // A_k -> act[k&1]; host sets act[1] before INIT
    int dead[MAXS];        // U_k (breakdown)
    int iters[MAXS];       // U_k
    int status[MAXS];      // A_k / U_k
    int k;                 // U_k increments; host resets
    int max_it;            // host
    int any_active;        // A_k (host polls)
    int pad;
};

template <class T>
struct SpeciesP {
    const T* D;  // COEF_CELL: partition-entry order; COEF_VERT: vertex order of the kernel
    const T* A;
    const uint8_t* bc;  // per-vertex Dirichlet mask (vertex order of the kernel) or null
    T Dval, Aval;
    int Dmode, Amode;
    int Dslot, Aslot;   // LDS slot of a per-vertex coefficient staged by the MF kernel
};

template <class T>
struct OpParams {
    SpeciesP<T> sp[MAXS];
    T inv_dt;
    int lumped;
    int n_vslots;  // per-vertex coefficient arrays staged in LDS
};

template <class T>
struct alignas(4 * sizeof(T)) Vec4 {
    T x, y, z, w;
};
template <class T>
__device__ __forceinline__ T comp(const Vec4<T>& a, int i) {
    return i == 0 ? a.x : (i == 1 ? a.y : (i == 2 ? a.z : a.w));
}

// vectors of the kernels, per species (vertex order of the kernel)
template <class T>
struct Vecs {
    T* x[MAXS];             // APPLY/RESID input x ; CG: z ; UPDATE: z out
    T* y[MAXS];             // APPLY output y ; DIAG output dinv ; CG: q ; RESID: r
    T* pold[MAXS];          // CG
    T* pnew[MAXS];          // CG
    const double* b[MAXS];  // RESID right-hand side (FP64)
};

// vectors of the update kernel
template <class T>
struct UVecs {
    T* x[MAXS];      // correction d
    T* r[MAXS];
    T* z[MAXS];
    const T* dinv[MAXS];
    const T* p[MAXS];     // p_k (written by A_k)
    const T* q[MAXS];
    T* pzero[MAXS];       // INIT: p_old buffer to clear
    const double* rin[MAXS];  // INIT: FP64 residual
};

struct Geo {
    int V, NV, nb, per_xcd, max_local, max_entries, max_inc;
    const int32_t* bptr;
    const ushort4* bloc;
    const int32_t* hptr;
    const int32_t* hlist;
    const int32_t* vptr;
    const ushort4* bslot;
};

// coefficient of partition entry k: scalar, per cell (entry order) or the mean of the four
// vertex values staged in LDS slot `slot`
template <class T>
__device__ __forceinline__ T coef_at(const T* __restrict__ p, int mode, T val, int k,
                                     const T* __restrict__ scv, int slot, int ml,
                                     const ushort4& l) {
    if (mode == COEF_SCALAR) return val;
    if (mode == COEF_CELL) return p[k];
    const T* c = scv + slot * ml;
    return (((c[l.x] + c[l.y]) + c[l.z]) + c[l.w]) * T(0.25);
}

// Block sums of N values, result in EVERY thread (fixed order: lanes, then waves).
template <int N>
__device__ __forceinline__ void block_sum_all(double (&v)[N], double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < N; ++i) scratch[i * 16 + wid] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += scratch[i * 16 + w];
        v[i] = s;
    }
    __syncthreads();
}

// Sum partials[q*nprod + b] over b < nprod for q < NQ; identical result in every thread of
// every workgroup that calls it.
template <int NQ>
__device__ __forceinline__ void consume_partials(const double* __restrict__ partials, int nprod,
                                                 double (&out)[NQ], double* scratch) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nprod; b += blockDim.x) s += partials[q * nprod + b];
        out[q] = s;
    }
    block_sum_all<NQ>(out, scratch);
}

// ---------------------------------------------------------------------------------------
// matrix-free kernel, all modes.  CG mode = A_k with parity `par` = k & 1.
//
// Per partition block: (1) stage the block's local vertices (owned range, then halo) into
// LDS -- coordinates, the species input (x, or p = z + beta p_old in CG mode) and any
// per-vertex coefficient; (2) every lane takes cells of the block's entry list, reads its
// four vertices from LDS, recomputes the Jacobian, applies the 4x4 operator of all species
// and adds the rows of owned vertices into LDS accumulators; (3) owned rows are written
// out with coalesced stores (identity rows at Dirichlet vertices).
template <class T, int S, int MODE>
__global__ void __launch_bounds__(1024)
k_mf(Geo g, const T* __restrict__ X, const T* __restrict__ Y, const T* __restrict__ Z,
     OpParams<T> prm, Vecs<T> v, CGState* __restrict__ st, const double* __restrict__ pin,
     int nin, double* __restrict__ pout, int par) {
    extern __shared__ __align__(32) unsigned char smem_raw[];
    const int ml = g.max_local;
    Vec4<T>* sgeo = reinterpret_cast<Vec4<T>*>(smem_raw);
    Vec4<T>* spv = sgeo + ml;
    Vec4<T>* ys = spv + ml;                               // owned-corner rows
    T* scv = reinterpret_cast<T*>(ys + g.max_inc);        // per-vertex coefficients
    __shared__ double red[16 * 2 * MAXS];
    const int nwg = gridDim.x;

    bool act[S];
    T beta[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = true;
        beta[s] = T(0);
    }
    if constexpr (MODE == MF_CG) {
        bool prev[S];
        bool anyprev = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            prev[s] = st->act[par ^ 1][s] != 0;
            anyprev |= prev[s];
        }
        if (!anyprev) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                for (int s = 0; s < S; ++s) st->act[par][s] = 0;
                st->any_active = 0;
            }
            return;
        }
        double sums[2 * S];
        consume_partials<2 * S>(pin, nin, sums, red);  // U_{k-1}: r.z, r.r per species
        const int k = st->k;
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double rz = sums[s], rr = sums[S + s];
            const bool dead = st->dead[s] != 0;
            act[s] = prev[s] && !dead && (rr > st->tol2[s]) && (k < st->max_it) && isfinite(rr);
            beta[s] = (k == 0 || !act[s]) ? T(0) : (T)(rz / st->rz[par ^ 1][s]);
            any |= act[s];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int s = 0; s < S; ++s) {
                st->act[par][s] = act[s] ? 1 : 0;
                if (!prev[s]) continue;
                st->rz[par][s] = sums[s];
                st->rr[s] = sums[S + s];
                if (!act[s] && st->dead[s] == 0)
                    st->status[s] = (sums[S + s] <= st->tol2[s]) ? ST_CONVERGED : ST_MAXIT;
            }
            st->any_active = any ? 1 : 0;
        }
        if (!any) return;
    }

    double part[S];
    double part2[S];
#pragma unroll
    for (int s = 0; s < S; ++s) part[s] = part2[s] = 0.0;

    const T inv_dt = prm.inv_dt;
    const int lumped = prm.lumped;
    // XCD x (= workgroup id mod 8) owns partition blocks [x*per_xcd, (x+1)*per_xcd)
    const int xcd = blockIdx.x & 7;
    const int wl = blockIdx.x >> 3;
    const int wpx = nwg >> 3;  // workgroups per XCD
    const int bbeg = xcd * g.per_xcd;
    const int bend = min(g.nb, bbeg + g.per_xcd);
    for (int b = bbeg + wl; b < bend; b += wpx) {
        const int v0 = b * g.NV;
        const int nv = min(g.NV, g.V - v0);
        const int h0 = g.hptr[b];
        const int nloc = nv + (g.hptr[b + 1] - h0);
        // (1) stage local vertices
        for (int i = threadIdx.x; i < nloc; i += blockDim.x) {
            const int gv = i < nv ? v0 + i : g.hlist[h0 + i - nv];
            sgeo[i] = Vec4<T>{X[gv], Y[gv], Z[gv], T(0)};
            if constexpr (MODE != MF_DIAG) {
                T pv[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!act[s]) continue;
                    if constexpr (MODE == MF_CG)
                        pv[s] = v.x[s][gv] + beta[s] * v.pold[s][gv];
                    else
                        pv[s] = v.x[s][gv];
                }
                spv[i] = Vec4<T>{pv[0], pv[1], pv[2], pv[3]};
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (prm.sp[s].Dslot >= 0) scv[prm.sp[s].Dslot * ml + i] = prm.sp[s].D[gv];
                if (prm.sp[s].Aslot >= 0) scv[prm.sp[s].Aslot * ml + i] = prm.sp[s].A[gv];
            }
        }
        const int i0 = g.vptr[v0];
        __syncthreads();
        // (2) cells: one lane per listed cell computes the rows of its four corners for every
        // species and stores the rows of owned corners in their own LDS slots (no atomics)
        const int kend = g.bptr[b + 1];
        for (int k = g.bptr[b] + threadIdx.x; k < kend; k += blockDim.x) {
            const ushort4 l = g.bloc[k];
            const ushort4 sl4 = g.bslot[k];
            const Vec4<T> q0 = sgeo[l.x], q1 = sgeo[l.y], q2 = sgeo[l.z], q3 = sgeo[l.w];
            Vec4<T> w0{}, w1{}, w2{}, w3{};
            if constexpr (MODE != MF_DIAG) {
                w0 = spv[l.x];
                w1 = spv[l.y];
                w2 = spv[l.z];
                w3 = spv[l.w];
            }
            const Vec3<T> e1{q1.x - q0.x, q1.y - q0.y, q1.z - q0.z};
            const Vec3<T> e2{q2.x - q0.x, q2.y - q0.y, q2.z - q0.z};
            const Vec3<T> e3{q3.x - q0.x, q3.y - q0.y, q3.z - q0.z};
            const Vec3<T> c1 = cross(e2, e3), c2 = cross(e3, e1), c3 = cross(e1, e2);
            const T adet = fabs(dot(e1, c1));
            const T r6 = T(1) / (T(6) * adet);
            T y[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int s = 0; s < 4; ++s) y[i][s] = T(0);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                const T De = coef_at(prm.sp[s].D, prm.sp[s].Dmode, prm.sp[s].Dval, k, scv,
                                     prm.sp[s].Dslot, ml, l);
                const T al = coef_at(prm.sp[s].A, prm.sp[s].Amode, prm.sp[s].Aval, k, scv,
                                     prm.sp[s].Aslot, ml, l);
                const T w = De * r6;
                const T mf = inv_dt + al;
                if constexpr (MODE == MF_DIAG) {
                    const Vec3<T> c0{-(c1.x + c2.x + c3.x), -(c1.y + c2.y + c3.y),
                                     -(c1.z + c2.z + c3.z)};
                    const T md = lumped ? mf * adet * T(5.0 / 120.0) : mf * adet * T(2.0 / 120.0);
                    y[0][s] = w * dot(c0, c0) + md;
                    y[1][s] = w * dot(c1, c1) + md;
                    y[2][s] = w * dot(c2, c2) + md;
                    y[3][s] = w * dot(c3, c3) + md;
                } else {
                    const T u0 = comp(w0, s), u1 = comp(w1, s), u2 = comp(w2, s), u3 = comp(w3, s);
                    const T d1 = u1 - u0, d2 = u2 - u0, d3 = u3 - u0;
                    const Vec3<T> gr{d1 * c1.x + d2 * c2.x + d3 * c3.x,
                                     d1 * c1.y + d2 * c2.y + d3 * c3.y,
                                     d1 * c1.z + d2 * c2.z + d3 * c3.z};
                    const T y1 = w * dot(c1, gr), y2 = w * dot(c2, gr), y3 = w * dot(c3, gr);
                    const T y0 = -(y1 + y2 + y3);
                    if (lumped) {
                        const T mlp = mf * adet * T(5.0 / 120.0);
                        y[0][s] = y0 + mlp * u0;
                        y[1][s] = y1 + mlp * u1;
                        y[2][s] = y2 + mlp * u2;
                        y[3][s] = y3 + mlp * u3;
                    } else {
                        const T mc = mf * adet * T(1.0 / 120.0);
                        const T su = ((u0 + u1) + u2) + u3;
                        y[0][s] = y0 + mc * (su + u0);
                        y[1][s] = y1 + mc * (su + u1);
                        y[2][s] = y2 + mc * (su + u2);
                        y[3][s] = y3 + mc * (su + u3);
                    }
                }
            }
            if (sl4.x != 0xFFFF) ys[sl4.x] = Vec4<T>{y[0][0], y[0][1], y[0][2], y[0][3]};
            if (sl4.y != 0xFFFF) ys[sl4.y] = Vec4<T>{y[1][0], y[1][1], y[1][2], y[1][3]};
            if (sl4.z != 0xFFFF) ys[sl4.z] = Vec4<T>{y[2][0], y[2][1], y[2][2], y[2][3]};
            if (sl4.w != 0xFFFF) ys[sl4.w] = Vec4<T>{y[3][0], y[3][1], y[3][2], y[3][3]};
        }
        __syncthreads();
        // (3) owned rows: sum the vertex's slots (fixed order), identity rows at Dirichlet
        // vertices, coalesced stores
        for (int lv = threadIdx.x; lv < nv; lv += blockDim.x) {
            const int vtx = v0 + lv;
            T acc[4] = {T(0), T(0), T(0), T(0)};
            const int jb = g.vptr[vtx] - i0, je = g.vptr[vtx + 1] - i0;
            for (int j = jb; j < je; ++j) {
                const Vec4<T> r = ys[j];
                acc[0] += r.x;
                acc[1] += r.y;
                acc[2] += r.z;
            }
            Vec4<T> own{};
            if constexpr (MODE != MF_DIAG) own = spv[lv];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                const uint8_t* __restrict__ bc = prm.sp[s].bc;
                const bool isbc = bc && bc[vtx];
                T val = acc[s];
                if constexpr (MODE == MF_APPLY) {
                    if (isbc) val = comp(own, s);
                    v.y[s][vtx] = val;
                } else if constexpr (MODE == MF_DIAG) {
                    v.y[s][vtx] = isbc ? T(1) : T(1) / val;
                } else if constexpr (MODE == MF_CG) {
                    const T p = comp(own, s);
                    if (isbc) val = p;
                    v.pnew[s][vtx] = p;
                    v.y[s][vtx] = val;
                    part[s] += (double)p * (double)val;
                } else {  // MF_RESID
                    if (isbc) val = comp(own, s);
                    const double bv = v.b[s][vtx];
                    const double r = bv - (double)val;
                    v.y[s][vtx] = (T)r;
                    part[s] += r * r;
                    part2[s] += bv * bv;
                }
            }
        }
        __syncthreads();  // LDS is reused by the next partition block
    }
    if constexpr (MODE == MF_CG || MODE == MF_RESID) {
        constexpr int NQ = MODE == MF_CG ? S : 2 * S;
        double tot[NQ];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            tot[s] = part[s];
            if constexpr (MODE == MF_RESID) tot[S + s] = part2[s];
        }
        block_sum_all<NQ>(tot, red);
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < NQ; ++q) pout[q * nwg + blockIdx.x] = tot[q];
    }
}

// ---------------------------------------------------------------------------------------
// CSR (caller order) kernels of the plain FP64 path: CG apply (A_k) and residual.
constexpr int CSR_LANES = 8;
constexpr int CSR_THREADS = 512;
constexpr int CSR_ROWS = CSR_THREADS / CSR_LANES;

template <int S, int MODE>
__global__ void __launch_bounds__(CSR_THREADS)
k_csr(int V, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
      const double* __restrict__ A0, const double* __restrict__ A1, const double* __restrict__ A2,
      Vecs<double> v, CGState* __restrict__ st, const double* __restrict__ pin, int nin,
      double* __restrict__ pout, int par) {
    __shared__ double red[16 * 2 * MAXS];
    const int nwg = gridDim.x;
    bool act[S];
    double beta[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = true;
        beta[s] = 0.0;
    }
    if constexpr (MODE == MF_CG) {
        bool prev[S];
        bool anyprev = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            prev[s] = st->act[par ^ 1][s] != 0;
            anyprev |= prev[s];
        }
        if (!anyprev) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                for (int s = 0; s < S; ++s) st->act[par][s] = 0;
                st->any_active = 0;
            }
            return;
        }
        double sums[2 * S];
        consume_partials<2 * S>(pin, nin, sums, red);
        const int k = st->k;
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double rz = sums[s], rr = sums[S + s];
            act[s] = prev[s] && st->dead[s] == 0 && (rr > st->tol2[s]) && (k < st->max_it) &&
                     isfinite(rr);
            beta[s] = (k == 0 || !act[s]) ? 0.0 : rz / st->rz[par ^ 1][s];
            any |= act[s];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int s = 0; s < S; ++s) {
                st->act[par][s] = act[s] ? 1 : 0;
                if (!prev[s]) continue;
                st->rz[par][s] = sums[s];
                st->rr[s] = sums[S + s];
                if (!act[s] && st->dead[s] == 0)
                    st->status[s] = (sums[S + s] <= st->tol2[s]) ? ST_CONVERGED : ST_MAXIT;
            }
            st->any_active = any ? 1 : 0;
        }
        if (!any) return;
    }
    const double* __restrict__ Av[3] = {A0, A1, A2};
    double part[S], part2[S];
#pragma unroll
    for (int s = 0; s < S; ++s) part[s] = part2[s] = 0.0;
    const int lane = threadIdx.x & (CSR_LANES - 1);
    for (int row = blockIdx.x * CSR_ROWS + (int)(threadIdx.x / CSR_LANES); row < V;
         row += nwg * CSR_ROWS) {
        const int js = indptr[row], je = indptr[row + 1];
        double acc[S];
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] = 0.0;
        for (int j = js + lane; j < je; j += CSR_LANES) {
            const int c = indices[j];
            double a[S], u[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                a[s] = act[s] ? Av[s][j] : 0.0;
                if constexpr (MODE == MF_CG)
                    u[s] = act[s] ? v.x[s][c] + beta[s] * v.pold[s][c] : 0.0;
                else
                    u[s] = v.x[s][c];
            }
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] += a[s] * u[s];
        }
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int off = CSR_LANES / 2; off > 0; off >>= 1) acc[s] += __shfl_xor(acc[s], off, 64);
        if (lane == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                if constexpr (MODE == MF_CG) {
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
    constexpr int NQ = MODE == MF_CG ? S : 2 * S;
    double tot[NQ];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = part[s];
        if constexpr (MODE != MF_CG) tot[S + s] = part2[s];
    }
    block_sum_all<NQ>(tot, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) pout[q * nwg + blockIdx.x] = tot[q];
}

// U_k: x += alpha p; r -= alpha q; z = r * dinv; partials r.z, r.r (parity par = k & 1).
// INIT (U_{-1}): r = (T) rin; x = 0; p_old = 0; z = r * dinv; partials.
template <class T, int S, bool INIT>
__global__ void __launch_bounds__(256)
k_update(int V, UVecs<T> v, CGState* __restrict__ st, const double* __restrict__ pin, int nin,
         double* __restrict__ pout, int par) {
    __shared__ double red[16 * 2 * MAXS];
    const int nwg = gridDim.x;
    bool act[S];
    T alpha[S];
    bool any = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = st->act[par][s] != 0;
        any |= act[s];
        alpha[s] = T(0);
    }
    if (!any) return;
    if constexpr (!INIT) {
        double pq[S];
        consume_partials<S>(pin, nin, pq, red);  // A_k: p.q per species
        bool bad[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            bad[s] = act[s] && (!(pq[s] > 0.0) || !isfinite(pq[s]));
            alpha[s] = (act[s] && !bad[s]) ? (T)(st->rz[par][s] / pq[s]) : T(0);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int s = 0; s < S; ++s) {
                if (bad[s]) {
                    st->dead[s] = 1;
                    st->status[s] = ST_BREAKDOWN;
                }
                if (act[s] && !bad[s]) st->iters[s] += 1;
            }
            st->k += 1;
        }
#pragma unroll
        for (int s = 0; s < S; ++s) act[s] = act[s] && !bad[s];
    }
    double prz[S], prr[S];
#pragma unroll
    for (int s = 0; s < S; ++s) prz[s] = prr[s] = 0.0;
    const int stride = nwg * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
        T r[S], d[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!act[s]) continue;
            d[s] = v.dinv[s][i];
            if constexpr (INIT) {
                r[s] = (T)v.rin[s][i];
            } else {
                const T pi = v.p[s][i], qi = v.q[s][i], xi = v.x[s][i], ri = v.r[s][i];
                v.x[s][i] = xi + alpha[s] * pi;
                r[s] = ri - alpha[s] * qi;
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!act[s]) continue;
            if constexpr (INIT) {
                v.x[s][i] = T(0);
                v.pzero[s][i] = T(0);
            }
            const T z = r[s] * d[s];
            v.r[s][i] = r[s];
            v.z[s][i] = z;
            prz[s] += (double)r[s] * (double)z;
            prr[s] += (double)r[s] * (double)r[s];
        }
    }
    double tot[2 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = prz[s];
        tot[S + s] = prr[s];
    }
    block_sum_all<2 * S>(tot, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 2 * S; ++q) pout[q * nwg + blockIdx.x] = tot[q];
}

// x64 += (double) d over the species in `mask` (refinement correction)
template <class T, int S>
__global__ void k_add_correction(int V, double* __restrict__ x0, double* __restrict__ x1,
                                 double* __restrict__ x2, const T* __restrict__ d0,
                                 const T* __restrict__ d1, const T* __restrict__ d2, int mask) {
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
    double* pA = nullptr;     // partials of A kernels (S per workgroup)
    double* pU[2] = {};       // partials of U kernels by parity (2S per workgroup)
    double* pR = nullptr;     // partials of the residual kernel (2S per workgroup)
    double* h_pR = nullptr;   // pinned copy
    // coefficient copies (target precision) and masks
    void* Dc[MAXS] = {};
    void* Ac[MAXS] = {};
    void* Dc2[MAXS] = {};
    void* Ac2[MAXS] = {};
    uint8_t* bcmask[MAXS] = {};
    unsigned char prm[sizeof(OpParams<double>)] = {};   // OpParams of the CG precision
    unsigned char prm2[sizeof(OpParams<double>)] = {};  // OpParams<double> of the residual
    // FP64 vectors
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
    // captured CG chunk and the parameters it was captured with
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0;
    int graph_valid = 0;
    unsigned char graph_prm[sizeof(OpParams<double>)] = {};
};

static size_t tsize(int method) { return method == DSV_METHOD_MF32_IR64 ? sizeof(float) : sizeof(double); }

static void free_ws(Workspace* w) {
    if (!w) return;
    if (w->graph) (void)hipGraphExecDestroy(w->graph);
    dfree(w->st);
    if (w->h_st) (void)hipHostFree(w->h_st);
    if (w->h_pR) (void)hipHostFree(w->h_pR);
    dfree(w->pA);
    dfree(w->pU[0]);
    dfree(w->pU[1]);
    dfree(w->pR);
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

// launch grids (every partial-producing kernel has at most MAX_PROD workgroups)
static int mf_grid(const dsv_mesh* m) {
    const int per = (m->part.nb + 7) / 8;
    return 8 * std::min(per, MAX_PROD / 8);
}
static int csr_grid(const dsv_mesh* m) { return std::min((int)cdiv(m->V, CSR_ROWS), MAX_PROD); }
static int upd_grid(const dsv_mesh* m) { return std::max(1, std::min((int)cdiv(m->V, 256 * 4), MAX_PROD)); }

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
        DSV_HIP(hipHostMalloc((void**)&w->h_pR, sizeof(double) * 2 * MAXS * MAX_PROD,
                              hipHostMallocDefault));
        w->pA = dalloc<double>((size_t)MAXS * MAX_PROD);
        w->pU[0] = dalloc<double>((size_t)2 * MAXS * MAX_PROD);
        w->pU[1] = dalloc<double>((size_t)2 * MAXS * MAX_PROD);
        w->pR = dalloc<double>((size_t)2 * MAXS * MAX_PROD);
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
    } catch (...) {
        free_ws(w);
        w = nullptr;
        throw;
    }
    return w;
}

// Copies / converts the coefficients of every species into the workspace (precision T):
// per-vertex arrays in the kernel's vertex order, per-cell arrays in partition-entry order
// (matrix-free) so that the MF kernels read them with coalesced loads.  Fills the kernel
// parameter block, including the LDS slots of staged per-vertex coefficients.
template <class T>
static void prepare_op(dsv_mesh* m, Workspace* w, int S, const dsv_species* sp, double inv_dt,
                       int lumped, bool internal, hipStream_t st, bool second = false) {
    OpParams<T> h{};
    h.inv_dt = (T)inv_dt;
    h.lumped = lumped ? 1 : 0;
    h.n_vslots = 0;
    const size_t cap = std::max({(size_t)m->C, (size_t)m->V, (size_t)m->part.n_entries});
    for (int s = 0; s < S; ++s) {
        SpeciesP<T>& q = h.sp[s];
        q.Dmode = coef_mode(sp[s].D, sp[s].D_len, m);
        q.Amode = coef_mode(sp[s].alpha, sp[s].alpha_len, m);
        q.Dval = (T)sp[s].D_value;
        q.Aval = sp[s].alpha ? T(0) : (T)sp[s].alpha_value;
        q.Dslot = q.Dmode == COEF_VERT ? h.n_vslots++ : -1;
        q.Aslot = q.Amode == COEF_VERT ? h.n_vslots++ : -1;
        const double* src[2] = {sp[s].D, sp[s].alpha};
        const int32_t lens[2] = {sp[s].D_len, sp[s].alpha_len};
        void** dst[2] = {second ? &w->Dc2[s] : &w->Dc[s], second ? &w->Ac2[s] : &w->Ac[s]};
        const T** out[2] = {&q.D, &q.A};
        for (int c = 0; c < 2; ++c) {
            *out[c] = nullptr;
            if (!src[c]) continue;
            if (!*dst[c]) DSV_HIP(hipMalloc(dst[c], cap * sizeof(T)));
            T* d = (T*)*dst[c];
            const int mode = c == 0 ? q.Dmode : q.Amode;
            if (mode == COEF_CELL && internal)
                hipLaunchKernelGGL((k_gather_perm<double, T>), dim3(cdiv(m->part.n_entries, 256)), dim3(256),
                                   0, st, (int)m->part.n_entries, m->part.belem, src[c], d);
            else if (mode == COEF_VERT && internal)
                hipLaunchKernelGGL((k_gather_perm<double, T>), dim3(cdiv(m->V, 256)), dim3(256), 0, st,
                                   m->V, m->perm, src[c], d);
            else
                hipLaunchKernelGGL((k_convert<double, T>), dim3(cdiv(lens[c], 256)), dim3(256), 0, st,
                                   (int64_t)lens[c], src[c], d);
            *out[c] = d;
        }
        if (!second) {
            DSV_HIP(hipMemsetAsync(w->bcmask[s], 0, m->V, st));
            if (sp[s].n_bc > 0)
                hipLaunchKernelGGL(k_set_bc_mask, dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0, st,
                                   sp[s].n_bc, sp[s].bc_dofs,
                                   internal ? m->iperm : (const int32_t*)nullptr, w->bcmask[s]);
        }
        q.bc = sp[s].n_bc > 0 ? w->bcmask[s] : nullptr;
    }
    std::memcpy(second ? w->prm2 : w->prm, &h, sizeof(h));
    DSV_HIP(hipGetLastError());
}

static Geo make_geo(const dsv_mesh* m) {
    Geo g;
    g.V = m->V;
    g.NV = m->part.NV;
    g.nb = m->part.nb;
    g.per_xcd = (m->part.nb + 7) / 8;
    g.max_local = m->part.max_local;
    g.max_entries = m->part.max_entries;
    g.max_inc = m->part.max_inc;
    g.bptr = m->part.bptr;
    g.bloc = m->part.bloc;
    g.hptr = m->part.hptr;
    g.hlist = m->part.hlist;
    g.vptr = m->part.vptr;
    g.bslot = m->part.bslot;
    return g;
}

static int max_lds_per_block(int device) {
    static int cached[64] = {};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    int v = 0;
    DSV_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    if (device >= 0 && device < 64) cached[device] = v;
    return v;
}

// LDS bytes of the MF kernel: staged geometry + species values (Vec4 each) and per-vertex
// coefficient slots of every local vertex, and one Vec4 row per owned cell corner
template <class T>
static size_t mf_lds_bytes(const dsv_mesh* m, int S, int n_vslots) {
    (void)S;
    return (size_t)m->part.max_local * (2 * sizeof(Vec4<T>) + n_vslots * sizeof(T)) +
           (size_t)m->part.max_inc * sizeof(Vec4<T>);
}

template <class T, int S, int MODE>
static void launch_mf(const dsv_mesh* m, const unsigned char* prm, const Vecs<T>& v, CGState* st,
                      const double* pin, int nin, double* pout, int par, hipStream_t stream) {
    const Geo g = make_geo(m);
    const T* X = sizeof(T) == 4 ? (const T*)m->xyz32 : (const T*)m->xyz64;
    OpParams<T> p;
    std::memcpy(&p, prm, sizeof(p));
    const size_t lds = mf_lds_bytes<T>(m, S, p.n_vslots);
    if (lds + 1024 > (size_t)max_lds_per_block(m->device))
        throw ArgError{"matrix-free partition needs " + std::to_string(lds) +
                       " B of LDS per workgroup; use fewer vertices_per_block"};
    hipLaunchKernelGGL((k_mf<T, S, MODE>), dim3(mf_grid(m)), dim3(m->part.threads), lds, stream, g, X,
                       X + m->V, X + 2 * (size_t)m->V, p, v, st, pin, nin, pout, par);
}

template <class T, int MODE>
static void launch_mf_s(int S, const dsv_mesh* m, const unsigned char* prm, const Vecs<T>& v,
                        CGState* st, const double* pin, int nin, double* pout, int par,
                        hipStream_t stream) {
    switch (S) {
        case 1: launch_mf<T, 1, MODE>(m, prm, v, st, pin, nin, pout, par, stream); break;
        case 2: launch_mf<T, 2, MODE>(m, prm, v, st, pin, nin, pout, par, stream); break;
        default: launch_mf<T, 3, MODE>(m, prm, v, st, pin, nin, pout, par, stream); break;
    }
}

template <int MODE>
static void launch_csr_s(int S, const dsv_mesh* m, Workspace* w, const Vecs<double>& v,
                         const double* pin, int nin, double* pout, int par, hipStream_t stream) {
    const int grid = csr_grid(m);
    const double* A0 = w->csr[0];
    const double* A1 = S > 1 ? w->csr[1] : A0;
    const double* A2 = S > 2 ? w->csr[2] : A0;
#define DSV_CSR_LAUNCH(SS)                                                                        \
    hipLaunchKernelGGL((k_csr<SS, MODE>), dim3(grid), dim3(CSR_THREADS), 0, stream, m->V, m->indptr, \
                       m->indices, A0, A1, A2, v, w->st, pin, nin, pout, par)
    switch (S) {
        case 1: DSV_CSR_LAUNCH(1); break;
        case 2: DSV_CSR_LAUNCH(2); break;
        default: DSV_CSR_LAUNCH(3); break;
    }
#undef DSV_CSR_LAUNCH
}

template <class T>
static UVecs<T> upd_vecs(Workspace* w, int S, int parity) {
    UVecs<T> u{};
    for (int s = 0; s < MAXS; ++s) {
        const int q = s < S ? s : 0;
        u.x[s] = (T*)w->xcg[q];
        u.r[s] = (T*)w->rcg[q];
        u.z[s] = (T*)w->zcg[q];
        u.dinv[s] = (const T*)w->dinv[q];
        u.p[s] = (const T*)(parity ? w->pa[q] : w->pb[q]);  // p_new of A_k with this parity
        u.q[s] = (const T*)w->qcg[q];
        u.pzero[s] = (T*)w->pa[q];  // INIT clears the p_old of A_0 (parity 0 reads pa)
        u.rin[s] = w->r64[q];
    }
    return u;
}

template <class T, bool INIT>
static void launch_update(int S, const dsv_mesh* m, Workspace* w, int parity, hipStream_t stream) {
    const UVecs<T> u = upd_vecs<T>(w, S, parity);
    const int grid = upd_grid(m);
    const int nin = w->method == DSV_METHOD_CSR64 ? csr_grid(m) : mf_grid(m);
    double* pout = w->pU[parity];
    switch (S) {
        case 1:
            hipLaunchKernelGGL((k_update<T, 1, INIT>), dim3(grid), dim3(256), 0, stream, m->V, u, w->st,
                               w->pA, nin, pout, parity);
            break;
        case 2:
            hipLaunchKernelGGL((k_update<T, 2, INIT>), dim3(grid), dim3(256), 0, stream, m->V, u, w->st,
                               w->pA, nin, pout, parity);
            break;
        default:
            hipLaunchKernelGGL((k_update<T, 3, INIT>), dim3(grid), dim3(256), 0, stream, m->V, u, w->st,
                               w->pA, nin, pout, parity);
            break;
    }
}

// vectors of A_k: z in x[], q in y[], p_old / p_new alternate between pa / pb by parity
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

// one CG iteration (A_k + U_k) with parity k & 1
template <class T>
static void cg_iteration(dsv_mesh* m, Workspace* w, int S, int parity, hipStream_t stream) {
    const Vecs<T> v = cg_vecs<T>(w, S, parity);
    const int nU = upd_grid(m);
    if constexpr (std::is_same<T, double>::value) {
        if (w->method == DSV_METHOD_CSR64) {
            launch_csr_s<MF_CG>(S, m, w, v, w->pU[parity ^ 1], nU, w->pA, parity, stream);
            launch_update<T, false>(S, m, w, parity, stream);
            return;
        }
    }
    launch_mf_s<T, MF_CG>(S, m, w->prm, v, w->st, w->pU[parity ^ 1], nU, w->pA, parity, stream);
    launch_update<T, false>(S, m, w, parity, stream);
}

// Runs CG iterations until no species is active.  Chunks of an even number of iterations
// are captured once into a hipGraph (re-captured when the operator parameters change) and
// replayed; the host polls one state word per chunk.
template <class T>
static int run_cg(dsv_mesh* m, Workspace* w, int S, int max_it, hipStream_t stream) {
    const int chunk = std::max(2, m->graph_chunk & ~1);
    int launched = 0;
    if (m->use_graphs && (!w->graph_valid || w->graph_chunk != chunk ||
                          std::memcmp(w->graph_prm, w->prm, sizeof(w->prm)) != 0)) {
        if (w->graph) (void)hipGraphExecDestroy(w->graph);
        w->graph = nullptr;
        w->graph_valid = 0;
        hipGraph_t gr;
        DSV_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < chunk; ++i) cg_iteration<T>(m, w, S, i & 1, stream);
        DSV_HIP(hipStreamEndCapture(stream, &gr));
        DSV_HIP(hipGraphInstantiate(&w->graph, gr, nullptr, nullptr, 0));
        DSV_HIP(hipGraphDestroy(gr));
        w->graph_valid = 1;
        w->graph_chunk = chunk;
        std::memcpy(w->graph_prm, w->prm, sizeof(w->prm));
    }
    while (true) {
        if (m->use_graphs) {
            DSV_HIP(hipGraphLaunch(w->graph, stream));
        } else {
            for (int i = 0; i < chunk; ++i) cg_iteration<T>(m, w, S, i & 1, stream);
        }
        launched += chunk;
        DSV_HIP(hipMemcpyAsync(&w->h_st->any_active, &w->st->any_active, sizeof(int),
                               hipMemcpyDeviceToHost, stream));
        DSV_HIP(hipStreamSynchronize(stream));
        if (!w->h_st->any_active || launched > max_it + 2 * chunk) break;
    }
    return launched;
}

static CGState read_state(Workspace* w, hipStream_t stream) {
    DSV_HIP(hipMemcpyAsync(w->h_st, w->st, sizeof(CGState), hipMemcpyDeviceToHost, stream));
    DSV_HIP(hipStreamSynchronize(stream));
    return *w->h_st;
}

static void write_state(Workspace* w, const CGState& h, hipStream_t stream) {
    *w->h_st = h;
    DSV_HIP(hipMemcpyAsync(w->st, w->h_st, sizeof(CGState), hipMemcpyHostToDevice, stream));
    DSV_HIP(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------------------
// residual r64 = b - A x64 (FP64, vertex order of the method); returns rr and bb per species
static void residual64(dsv_mesh* m, Workspace* w, int S, hipStream_t stream, double* rr,
                       double* bb) {
    Vecs<double> v{};
    for (int s = 0; s < MAXS; ++s) {
        const int q = s < S ? s : 0;
        v.x[s] = w->x64[q];
        v.y[s] = w->r64[q];
        v.b[s] = w->b64[q];
    }
    int nprod;
    if (w->method == DSV_METHOD_CSR64) {
        launch_csr_s<MF_RESID>(S, m, w, v, nullptr, 0, w->pR, 0, stream);
        nprod = csr_grid(m);
    } else {
        launch_mf_s<double, MF_RESID>(S, m, w->method == DSV_METHOD_MF32_IR64 ? w->prm2 : w->prm, v,
                                      w->st, nullptr, 0, w->pR, 0, stream);
        nprod = mf_grid(m);
    }
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipMemcpyAsync(w->h_pR, w->pR, sizeof(double) * 2 * S * nprod, hipMemcpyDeviceToHost,
                           stream));
    DSV_HIP(hipStreamSynchronize(stream));
    for (int s = 0; s < S; ++s) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < nprod; ++i) {
            a += w->h_pR[s * nprod + i];
            b += w->h_pR[(S + s) * nprod + i];
        }
        rr[s] = a;
        bb[s] = b;
    }
}

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
    CGState h = read_state(w, st);
    for (int s = 0; s < MAXS; ++s) {
        h.tol2[s] = s < S ? tol2[s] : 0.0;
        h.act[1][s] = (s < S && active[s]) ? 1 : 0;
        h.act[0][s] = 0;
        h.dead[s] = 0;
        h.iters[s] = 0;
        h.status[s] = (s < S && active[s]) ? ST_RUNNING : ST_IDLE;
    }
    h.k = 0;
    h.max_it = max_it;
    h.any_active = 1;
    write_state(w, h, st);
    launch_update<T, true>(S, m, w, 1, st);  // INIT = U_{-1}
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
        launch_mf_s<double, MF_DIAG>(S, m, w->prm, v, w->st, nullptr, 0, nullptr, 0, st);
    } else {
        prepare_op<float>(m, w, S, sp, inv_dt, lumped, true, st);
        prepare_op<double>(m, w, S, sp, inv_dt, lumped, true, st, true);
        Vecs<float> v{};
        for (int s = 0; s < MAXS; ++s) v.y[s] = (float*)w->dinv[s < S ? s : 0];
        launch_mf_s<float, MF_DIAG>(S, m, w->prm, v, w->st, nullptr, 0, nullptr, 0, st);
    }
    DSV_HIP(hipGetLastError());

    double tol_abs[MAXS] = {0, 0, 0}, rel0[MAXS] = {0, 0, 0}, relf[MAXS] = {0, 0, 0};
    int conv[MAXS] = {0, 0, 0}, outer[MAXS] = {0, 0, 0}, dead[MAXS] = {0, 0, 0};
    int iters_total[MAXS] = {0, 0, 0};
    const int max_rounds = method == DSV_METHOD_MF32_IR64 ? m->max_outer : 4;
    for (int round = 0;; ++round) {
        double rr[MAXS] = {0, 0, 0}, bb[MAXS] = {0, 0, 0};
        residual64(m, w, S, st, rr, bb);
        bool all = true;
        for (int s = 0; s < S; ++s) {
            const double bn = std::sqrt(bb[s]), rn = std::sqrt(rr[s]);
            if (round == 0) {
                tol_abs[s] = rtol * bn;
                rel0[s] = bn > 0 ? rn / bn : rn;
            }
            relf[s] = bn > 0 ? rn / bn : rn;
            conv[s] = (rn <= tol_abs[s]) ? 1 : 0;
            if (iters_total[s] >= max_it || !std::isfinite(rn)) dead[s] = 1;
            all &= (conv[s] || dead[s]);
        }
        if (all || round >= max_rounds) break;
        double tol2[MAXS] = {0, 0, 0};
        int act[MAXS] = {0, 0, 0};
        int mask = 0;
        for (int s = 0; s < S; ++s) {
            if (conv[s] || dead[s]) continue;
            const double rn = std::sqrt(rr[s]);
            double t = method == DSV_METHOD_MF32_IR64 ? std::max(m->inner_rtol * rn, 0.3 * tol_abs[s])
                                                      : 0.9 * tol_abs[s];
            tol2[s] = t * t;
            act[s] = 1;
            mask |= 1 << s;
            outer[s] += 1;
        }
        // remaining iteration budget of the most advanced active species
        int budget = max_it;
        for (int s = 0; s < S; ++s)
            if (act[s]) budget = std::min(budget, max_it - iters_total[s]);
        if (method == DSV_METHOD_MF32_IR64) {
            inner_cg<float>(m, w, S, tol2, act, std::max(1, budget), st);
            add_correction<float>(S, m, w, mask, st);
        } else {
            inner_cg<double>(m, w, S, tol2, act, std::max(1, budget), st);
            add_correction<double>(S, m, w, mask, st);
        }
        DSV_HIP(hipGetLastError());
        // the state's iteration counters restart in every inner solve; accumulate them
        const CGState h2 = read_state(w, st);
        for (int s = 0; s < S; ++s) {
            if (!act[s]) continue;
            iters_total[s] += h2.iters[s];
            if (h2.status[s] == ST_BREAKDOWN) dead[s] = 1;
        }
    }
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
            stats->iterations[s] = iters_total[s];
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
    launch_mf_s<T, MF_APPLY>(S, m, w->prm, v, w->st, nullptr, 0, nullptr, 0, st);
    for (int s = 0; s < S; ++s)
        hipLaunchKernelGGL((k_scatter_perm<T, double>), dim3(gv), dim3(256), 0, st, m->V, m->perm,
                           (const T*)v.y[s], y[s]);
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipStreamSynchronize(st));
}

// Times the fused kernel on `stream`: mode 0 = apply, mode 1 = CG form (A_k, parity 0, with
// every species active and beta = 0 so that repeated launches are identical).
template <class T>
static double time_impl(dsv_mesh* m, int S, const dsv_species* sp, double inv_dt, int lumped,
                        int mode, int n_rep, hipStream_t stream) {
    const int method = sizeof(T) == 4 ? DSV_METHOD_MF32_IR64 : DSV_METHOD_MF64;
    Workspace* w = get_ws(m, method, S);
    DSV_HIP(hipStreamSynchronize(stream));
    prepare_op<T>(m, w, S, sp, inv_dt, lumped, true, w->stream);
    const unsigned gv = cdiv(m->V, 256);
    for (int s = 0; s < S; ++s) {
        // z = x (internal order) when x is given, else the lumped mass; p_old = 0
        if (sp[s].x)
            hipLaunchKernelGGL((k_gather_perm<double, T>), dim3(gv), dim3(256), 0, w->stream, m->V, m->perm,
                               (const double*)sp[s].x, (T*)w->zcg[s]);
        else
            hipLaunchKernelGGL((k_convert<double, T>), dim3(gv), dim3(256), 0, w->stream, (int64_t)m->V,
                               m->lumped_m, (T*)w->zcg[s]);
        DSV_HIP(hipMemsetAsync(w->pa[s], 0, m->V * sizeof(T), w->stream));
    }
    // state: every species active before A_0 (act[1]), k = 0, tolerance 0; the U partials
    // the kernel consumes hold r.r = 1 per species
    CGState h = read_state(w, w->stream);
    for (int s = 0; s < MAXS; ++s) {
        h.act[1][s] = s < S ? 1 : 0;
        h.tol2[s] = 0.0;
        h.dead[s] = 0;
    }
    h.k = 0;
    h.max_it = 1 << 30;
    write_state(w, h, w->stream);
    std::vector<double> pu((size_t)2 * MAXS * MAX_PROD, 0.0);
    const int nU = upd_grid(m);
    for (int s = 0; s < S; ++s) pu[(size_t)(S + s) * nU] = 1.0;
    DSV_HIP(hipMemcpy(w->pU[1], pu.data(), pu.size() * sizeof(double), hipMemcpyHostToDevice));
    const Vecs<T> v = cg_vecs<T>(w, S, 0);
    auto launch = [&] {
        if (mode == 1)
            launch_mf_s<T, MF_CG>(S, m, w->prm, v, w->st, w->pU[1], nU, w->pA, 0, stream);
        else
            launch_mf_s<T, MF_APPLY>(S, m, w->prm, v, w->st, nullptr, 0, nullptr, 0, stream);
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
    return guard([&] {
        solve_impl(m, n_species, sp, inv_dt, lumped, method, rtol, max_it, (hipStream_t)stream, stats);
    });
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
