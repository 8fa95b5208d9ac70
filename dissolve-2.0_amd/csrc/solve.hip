// Matrix-free multi-species apply and the batched Jacobi-PCG solvers.
//
// Replaces CeedDiffusionMassOp.apply (`gpu/ceed_operators.py:341-343`, whose QFunction is a
// different operator -- the apply here reproduces the ASSEMBLED operator of
// `gpu/gpu_assembler.py:60-165`) and the KSP solve of _solve_gpu_asm
// (`numerics/solvers.py:322-372`, GMRES(50)+bjacobi on a CSR that is copied through the host
// every step).
//
// Data layout: every solver vector is species-interleaved -- one unpadded record SV<T,S> per
// vertex holding the S species -- so one load fetches every species of a vertex.  Coordinates are (x, y, z, 0) records.
//
// Matrix-free apply (internal Morton vertex numbering), one workgroup per partition block:
//   (1) stage the block's local vertices (owned range, then halo) in LDS: coordinates and the
//       species input;
//   (2) L lanes per owned vertex walk the vertex's incident cells (sliced-ELL lists of the
//       other three corners' local indices), recompute each cell's geometry from LDS and
//       accumulate the vertex's row of the cell operator for every species in registers --
//       no atomics, one barrier, a fixed summation order -- then write the row (identity
//       rows at Dirichlet vertices).  The 8 XCDs each take a contiguous run of partition
//       blocks, which keeps halo reads in the XCD's own L2.
//
// Matrix-free CG (methods MF64 and MF32_IR64): Jacobi-PCG in the Chronopoulos-Gear form, ONE
// kernel per iteration.  Iteration i reduces gamma_i = (r_i, u_i), delta_i = (w_i, u_i) and
// rho_i = (r_i, r_i) from the previous kernel's block partials, then computes
//   beta = gamma_i / gamma_{i-1}, alpha = gamma_i / (delta_i - beta gamma_i / alpha_{i-1}),
//   s_i = w_i + beta s_{i-1}, r_{i+1} = r_i - alpha s_i, u_{i+1} = r_{i+1} / diag(A)
// for every local vertex while staging it (halo copies are recomputed from the same inputs,
// bit-identical to the owner's), updates p_i = u_i + beta p_{i-1} and x_{i+1} = x_i + alpha p_i
// for the owned ones, and applies w_{i+1} = A u_{i+1}, writing the partials of iteration i+1.
// r, w and s alternate between two buffers so that halo reads never see this kernel's writes.
// No kernel uses a global atomic, and the iterations need no fence: every workgroup reduces
// ALL partials of the previous kernel in the same fixed order, so all derive identical
// alpha / beta / convergence decisions; workgroup 0 records them, and the last kernel of a
// chunk publishes them to the host (one system-scope fence, one thread).  Chunks of
// iterations are launched (or replayed from captured hipGraphs) until every species is done.
//
// CSR CG (method CSR64, the plain FP64 path): standard Jacobi-PCG with two kernels per
// iteration, the CSR apply (A_k) and the vector update (U_k), with the same fence-free
// reductions.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <atomic>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

void check_species_coefs(const dsv_mesh* m, const dsv_species& s);
int coef_mode(const double* p, int32_t len, const dsv_mesh* m);
void assemble_into(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped, double* values,
                   hipStream_t st);

constexpr int MAXS = 3;
constexpr int MAXD = 8;  // warm-start directions of dsv_solve_warm
constexpr int MAX_PROD = 2048;  // max workgroups of a partial-producing kernel

enum MFMode { MF_APPLY = 0, MF_DIAG = 1, MF_RESID = 3, MF_MASS = 4 };

// status codes in CGState::status
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_BREAKDOWN = 2, ST_MAXIT = 3, ST_IDLE = 4 };

// Written by workgroup 0 of the kernel named in the comment, read by later kernels / host.
// Entries indexed [par] are written by the kernel with parity par and read by the next one.
struct CGState {
    double tol2[MAXS];     // host
    double rz[2][MAXS];    // CSR A_k -> r_k.z_k ; MF K_i -> gamma_i
    double alp[2][MAXS];   // MF K_i -> alpha_i
    double rr[MAXS];       // A_k / K_i -> r.r (host)
    int act[2][MAXS];      // A_k / K_i; host sets act[1] before the first iteration
    int dead[MAXS];        // CSR U_k (breakdown)
    int iters[MAXS];       // U_k / K_i
    int status[MAXS];      // A_k / U_k / K_i
    int k;                 // CSR: U_k increments; host resets
    int kk[2];             // MF: K_i -> i + 1; host sets kk[1] = 0
    int max_it;            // host
    int any_active;        // A_k / K_i (host polls)
    int seq;               // poll snapshots only: set last, once the rest is written
};

// The residual kernel's norms and, on request, the CG state, published into pinned host
// memory by k_resid_publish: the host spins on `seq` (set last) instead of copying them back
// and synchronising the stream.
struct ResidPub {
    double rr[MAXS];
    double bb[MAXS];
    CGState state;
    int seq;
};

// species-interleaved vertex record: the S values of one vertex, unpadded (12 B for three
// FP32 species), so that every byte a vector pass moves is a species value
template <class T, int S>
struct SVs;
template <class T>
struct alignas(sizeof(T)) SVs<T, 1> {
    T v[1];
};
template <class T>
struct alignas(2 * sizeof(T)) SVs<T, 2> {
    T v[2];
};
template <class T>
struct alignas(sizeof(T)) SVs<T, 3> {
    T v[3];
};
static_assert(sizeof(SVs<float, 3>) == 12 && sizeof(SVs<double, 3>) == 24, "unpadded records");
template <class T, int S>
using SV = SVs<T, S>;

template <class T>
struct alignas(4 * sizeof(T)) Vec4 {
    T x, y, z, w;
};

// Operator coefficients of one species.  Generic form: the cell's diffusivity D_e and
// reaction coefficient alpha_e per ELL entry (cell-averaged as in assembly).  Fast form
// (lumped mass and a scalar D for every species): D/6 as a scalar, the whole mass term folded
// into the per-vertex diagonal `mdiag`.
template <class T>
struct SpeciesP {
    const T* D;  // per ELL entry (generic form); NULL for a scalar D (D1) or when only the
                 // stored form is prepared (OpParams::no_generic)
    const T* A;  // per ELL entry; NULL for a scalar reaction coefficient (A0)
    T A0;        // the reaction coefficient where A is NULL
    T D6;        // fast forms: D / 6
    T D1;        // stored-geometry form: D (1 when W holds the species' own K(D) weights)
    const T* W;  // stored-geometry form with per-species weights: K(D_s) per ELL entry
};

template <class T>
struct OpParams {
    SpeciesP<T> sp[MAXS];
    const uint8_t* bcbits;  // per vertex: bit s set at the Dirichlet vertices of species s
    const void* mdiag;      // fast form: sum over cells of (inv_dt + alpha_e) |det J_e| / 24
    T inv_dt;
    int lumped;
    int fast;
    int stored;  // stored-geometry form (lumped mass): apply, residual, diagonal, CG
    int persp;   // stored form with per-species weights (some species has a D field)
    int no_generic;  // a D field's per-entry values were not built: generic kernels (other than
                     // the mass diagonal) must not run with these parameters
};

// vectors of the apply-type kernels (vertex order of the kernel)
template <class T, int S>
struct MFVec {
    const SV<T, S>* x;       // APPLY / RESID: x
    SV<T, S>* y;             // APPLY: y ; DIAG: dinv ; RESID: r ; MASS: mass diagonal
    const SV<double, S>* b;  // RESID
};

// vectors of the matrix-free CG kernel (internal order); R, W, S alternate by parity
template <class T, int S>
struct CGVec {
    const SV<T, S>* dinv;
    SV<T, S>* R[2];
    SV<T, S>* W[2];
    SV<T, S>* Sd[2];
    SV<T, S>* P;
    SV<T, S>* X;               // correction
    const SV<double, S>* rin;  // INIT: FP64 residual
};

// vectors of the CSR path (caller order)
template <int S>
struct CSRVec {
    const SV<double, S>* x;     // A_k: z ; RESID: x
    SV<double, S>* y;           // A_k: q ; RESID: r
    const SV<double, S>* pold;  // A_k
    SV<double, S>* pnew;        // A_k
    const SV<double, S>* b;     // RESID
};
template <int S>
struct UVec {
    SV<double, S>* x;  // correction d
    SV<double, S>* r;
    SV<double, S>* z;
    const SV<double, S>* dinv;
    const SV<double, S>* p;    // p_k (written by A_k)
    const SV<double, S>* q;
    SV<double, S>* pzero;      // INIT: p_old of A_0, cleared
    const SV<double, S>* rin;  // INIT: residual
};

struct Geo {
    int V, NV, L, nb, per_xcd, max_local;
    const int32_t* hptr;
    const int32_t* lvl;
    const uint16_t* lperm;
    const int32_t* sptr;
    const uint32_t* ell;
    const int32_t* eptr;
    const uint16_t* elperm;
    const uint16_t* eidx;
    const float* ewt;
    const double* ewt64;
};

template <class T>
__device__ __forceinline__ T recip(T x);
template <>
__device__ __forceinline__ float recip<float>(float x) {
    return __builtin_amdgcn_rcpf(x);  // v_rcp_f32, 1 ulp: the FP32 inner operator only
}
template <>
__device__ __forceinline__ double recip<double>(double x) {
    return 1.0 / x;
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
// every workgroup that calls it.  All NQ loads of a step are issued together.  Partials are
// FP64, or FP32 (P = float) where the consumer is the FP32 inner CG: every workgroup of the
// consuming kernel reads all of them, so their size sets the cost of the reduction.
template <int NQ, class P = double>
__device__ __forceinline__ void consume_partials(const P* __restrict__ partials, int nprod,
                                                 double (&out)[NQ], double* scratch) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = 0.0;
    for (int b = threadIdx.x; b < nprod; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[q] += (double)partials[q * nprod + b];
    block_sum_all<NQ>(out, scratch);
}

// Writes the NQ block sums of `tot` as this workgroup's partials.
template <int NQ, class P = double>
__device__ __forceinline__ void write_partials(double (&tot)[NQ], double* scratch,
                                               P* __restrict__ pout) {
    block_sum_all<NQ>(tot, scratch);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) pout[q * gridDim.x + blockIdx.x] = (P)tot[q];
}

// ---------------------------------------------------------------------------------------
// matrix-free rows
//
// Row of owned vertex 0 of a cell with edges e_k = x_k - x_0 (k = 1..3), c1 = e2 x e3,
// c2 = e3 x e1, c3 = e1 x e2, |det| = |e1 . c1|, cs = c1 + c2 + c3: the stiffness row is
// K_0k = D c0.c_k / (6 |det|) with c0 = -cs, so (K u)_0 = -(D/6) sum_k g_k (u_k - u_0) with
// g_k = cs.c_k / |det|; the diagonal is (D/6) cs.cs / |det|.  Lumped mass |det|/24 on the
// diagonal, consistent mass |det| (1 + delta_0k) / 120.

enum RowKind { ROW_MATVEC = 0, ROW_DIAG = 1, ROW_MASS = 2 };

// Row sums of the owned vertex served by lane slot `ls` of block b, summed over its L lanes
// (all lanes of the wave must call this).  Generic form: the full rows (matvec, diagonal or
// lumped mass diagonal).  Fast form: matvec acc[s] = sum_k g_k (u_k - u_0), diagonal
// acc[0] = sum cs.cs / |det| (species-independent); finish_fast turns these into rows.
// Sets `ov` (owned local index), `live`, `own` (the vertex's input).
template <class T, int S, int KIND, bool FAST>
__device__ __forceinline__ void lane_rows(const Geo& g, const OpParams<T>& prm,
                                          const Vec4<T>* sgeo, const SV<T, S>* spv, int b, int nv,
                                          int ls, int& ov, bool& live, SV<T, S>& own, T (&acc)[S]) {
    const int L = g.L;
    const int vps = 64 / L;
    const int spb = g.NV / vps;
    const int lane = ls & 63;
    const int rank = ls / L;  // < NV: ls < NV * L
    const int* sp0 = g.sptr + b * spb + (ls >> 6);
    const int jbeg = sp0[0] + lane;
    const int width = (sp0[1] - sp0[0]) >> 6;
    live = rank < nv;
    const int ovl = (int)g.lperm[b * g.NV + rank];
    ov = live ? ovl : 0;
    const Vec4<T> q0 = sgeo[ov];
    own = SV<T, S>{};
    if constexpr (KIND == ROW_MATVEC) own = spv[ov];
    const T inv_dt = prm.inv_dt;
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = T(0);
    for (int t0 = 0; t0 < width; t0 += 4) {
        // the list words of 4 steps are loaded together (clamped in-bounds addresses, steps
        // past the list read as padding), off the dependent chain
        uint32_t pk4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t wv = g.ell[jbeg + 64 * min(t0 + u, width - 1)];
            pk4[u] = t0 + u < width ? wv : ELL_PAD;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t pk = pk4[u];
            if (pk == ELL_PAD) continue;  // padding, also every entry of a lane past nv
            const int j = jbeg + 64 * (t0 + u);
            const int la = pk & 1023, lb = (pk >> 10) & 1023, lc = pk >> 20;
            SV<T, S> w1{}, w2{}, w3{};
            if constexpr (KIND == ROW_MATVEC) {
                w1 = spv[la];
                w2 = spv[lb];
                w3 = spv[lc];
            }
            const Vec4<T> q1 = sgeo[la], q2 = sgeo[lb], q3 = sgeo[lc];
            const Vec3<T> e1{q1.x - q0.x, q1.y - q0.y, q1.z - q0.z};
            const Vec3<T> e2{q2.x - q0.x, q2.y - q0.y, q2.z - q0.z};
            const Vec3<T> e3{q3.x - q0.x, q3.y - q0.y, q3.z - q0.z};
            const Vec3<T> c1 = cross(e2, e3), c2 = cross(e3, e1), c3 = cross(e1, e2);
            const T adet = fabs(dot(e1, c1));
            const T ri = recip(adet);
            const Vec3<T> cs{c1.x + c2.x + c3.x, c1.y + c2.y + c3.y, c1.z + c2.z + c3.z};
            if constexpr (KIND == ROW_DIAG && FAST) {
                acc[0] += dot(cs, cs) * ri;
            } else if constexpr (KIND != ROW_MATVEC) {
                const T gd = dot(cs, cs) * ri;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const T mf = inv_dt + (prm.sp[s].A ? prm.sp[s].A[j] : prm.sp[s].A0);
                    if constexpr (KIND == ROW_MASS)
                        acc[s] += mf * adet * T(5.0 / 120.0);
                    else
                        acc[s] += (prm.sp[s].D ? prm.sp[s].D[j] : prm.sp[s].D1) * T(1.0 / 6.0) * gd +
                                  (prm.lumped ? mf * adet * T(5.0 / 120.0) : mf * adet * T(2.0 / 120.0));
                }
            } else {
                const T g1 = dot(cs, c1) * ri, g2 = dot(cs, c2) * ri, g3 = dot(cs, c3) * ri;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const T u0 = own.v[s], u1 = w1.v[s], u2 = w2.v[s], u3 = w3.v[s];
                    const T gsum = g1 * (u1 - u0) + g2 * (u2 - u0) + g3 * (u3 - u0);
                    if constexpr (FAST) {
                        acc[s] += gsum;
                    } else {
                        const T mf = inv_dt + (prm.sp[s].A ? prm.sp[s].A[j] : prm.sp[s].A0);
                        const T kr = -(prm.sp[s].D ? prm.sp[s].D[j] : prm.sp[s].D1) * T(1.0 / 6.0) *
                                     gsum;
                        if (prm.lumped)
                            acc[s] += kr + mf * adet * T(5.0 / 120.0) * u0;
                        else
                            acc[s] += kr + mf * adet * T(1.0 / 120.0) * ((((u0 + u1) + u2) + u3) + u0);
                    }
                }
            }
        }
    }
    // sum the partial rows of the L adjacent lanes of a vertex (pairwise, so every lane holds
    // the same bits)
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (L >= 2) acc[s] = add_lane_xor1(acc[s]);
        if (L >= 4) acc[s] = add_lane_xor2(acc[s]);
    }
}

// fast form: rows from the geometric sums, -(D/6) sum + mdiag u (diagonal: (D/6) sum + mdiag)
template <class T, int S, int KIND>
__device__ __forceinline__ void finish_fast(const OpParams<T>& prm, int vtx, const SV<T, S>& own,
                                            T (&acc)[S]) {
    const SV<T, S> md = reinterpret_cast<const SV<T, S>*>(prm.mdiag)[vtx];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if constexpr (KIND == ROW_DIAG)
            acc[s] = prm.sp[s].D6 * acc[0] + md.v[s];
        else
            acc[s] = -prm.sp[s].D6 * acc[s] + md.v[s] * own.v[s];
    }
}

// Stored-geometry rows: acc[s] = sum over graph neighbours w of K_vw(D=1) (u_w - u_v) for the
// owned vertex of lane slot `ls`, summed over its L lanes (all lanes of the wave must call
// this).  Padding entries point at the vertex itself with weight 0, so there is no branch.
template <class T>
__device__ __forceinline__ const T* edge_weights(const Geo& g);
template <>
__device__ __forceinline__ const float* edge_weights<float>(const Geo& g) {
    return g.ewt;
}
template <>
__device__ __forceinline__ const double* edge_weights<double>(const Geo& g) {
    return g.ewt64;
}

template <class T, int S, bool PERSP = false, bool WSUM = false>
__device__ __forceinline__ void edge_rows(const Geo& g, const OpParams<T>& prm,
                                          const SV<T, S>* spv, int b, int nv, int ls, int& ov,
                                          bool& live, SV<T, S>& own, T (&acc)[S]) {
    const T* __restrict__ ewt = edge_weights<T>(g);
    const int L = g.L;
    const int vps = 64 / L;
    const int spb = g.NV / vps;
    const int lane = ls & 63;
    const int rank = ls / L;  // < NV: ls < NV * L
    const int* sp0 = g.eptr + b * spb + (ls >> 6);
    const int jbeg = sp0[0] + lane;
    const int width = (sp0[1] - sp0[0]) >> 6;
    live = rank < nv;
    const int ovl = (int)g.elperm[b * g.NV + rank];
    ov = live ? ovl : 0;
    if constexpr (!WSUM) own = spv[ov];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = T(0);
    for (int t0 = 0; t0 < width; t0 += 4) {
        int ix[4];
        T wt[PERSP ? S : 1][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = jbeg + 64 * min(t0 + u, width - 1);
            const bool in = t0 + u < width;
            ix[u] = in ? (int)g.eidx[j] : ov;
#pragma unroll
            for (int q = 0; q < (PERSP ? S : 1); ++q) {
                const T wv = PERSP ? prm.sp[q].W[j] : ewt[j];
                wt[q][u] = in ? wv : T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if constexpr (WSUM) {
#pragma unroll
                for (int s = 0; s < S; ++s) acc[s] += wt[PERSP ? s : 0][u];
            } else {
                const SV<T, S> nb = spv[ix[u]];
#pragma unroll
                for (int s = 0; s < S; ++s) acc[s] += wt[PERSP ? s : 0][u] * (nb.v[s] - own.v[s]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (L >= 2) acc[s] = add_lane_xor1(acc[s]);
        if (L >= 4) acc[s] = add_lane_xor2(acc[s]);
    }
}

// Workgroup -> partition blocks: XCD x (= workgroup id mod 8) owns blocks
// [x*per_xcd, (x+1)*per_xcd), taken round-robin by its workgroups.
struct BlockRange {
    int beg, end, step;
};
__device__ __forceinline__ BlockRange my_blocks(const Geo& g) {
    const int xcd = blockIdx.x & 7;
    const int bbeg = xcd * g.per_xcd;
    return BlockRange{bbeg + (int)(blockIdx.x >> 3), min(g.nb, bbeg + g.per_xcd),
                      (int)(gridDim.x >> 3)};
}

constexpr int MF_THREADS_MAX = 512;  // workgroup sizes 256 and 512

// ---------------------------------------------------------------------------------------
// matrix-free operator kernel.  MODE: APPLY y = A x; DIAG y = 1 / diag(A); RESID r = b - A x
// (FP64 partials of r.r and b.b); MASS (generic form only) y = lumped (inv_dt + alpha) mass
// diagonal, the `mdiag` of the fast form.
template <class T, int S, int MODE, bool FAST, bool STORED = false, bool PERSP = false>
__global__ void __launch_bounds__(MF_THREADS_MAX)
k_mf(Geo g, const Vec4<T>* __restrict__ xyz, OpParams<T> prm, MFVec<T, S> v,
     double* __restrict__ pout) {
    static_assert(!(FAST && MODE == MF_MASS), "the mass diagonal is built by the generic form");
    static_assert(!STORED || (MODE != MF_MASS && FAST),
                  "stored geometry: apply, residual and diagonal");
    static_assert(STORED || !PERSP, "per-species weights belong to the stored form");
    constexpr int KIND = MODE == MF_DIAG ? ROW_DIAG : (MODE == MF_MASS ? ROW_MASS : ROW_MATVEC);
    extern __shared__ __align__(32) unsigned char smem_raw[];
    Vec4<T>* sgeo = reinterpret_cast<Vec4<T>*>(smem_raw);
    // the stored form stages no coordinates
    SV<T, S>* spv = STORED ? reinterpret_cast<SV<T, S>*>(smem_raw)
                           : reinterpret_cast<SV<T, S>*>(sgeo + g.max_local);
    __shared__ double red[16 * 2 * MAXS];
    const int ml = g.max_local;

    double part[S], part2[S];
#pragma unroll
    for (int s = 0; s < S; ++s) part[s] = part2[s] = 0.0;
    const BlockRange br = my_blocks(g);
    for (int b = br.beg; b < br.end; b += br.step) {
        const int v0 = b * g.NV;
        const int nv = min(g.NV, g.V - v0);
        const int nloc = nv + (g.hptr[b + 1] - g.hptr[b]);
        for (int i = threadIdx.x; i < nloc; i += blockDim.x) {
            const int gv = i < nv ? v0 + i : g.lvl[(size_t)b * ml + i];  // owned: contiguous
            if constexpr (!STORED) sgeo[i] = xyz[gv];
            if constexpr (KIND == ROW_MATVEC) spv[i] = v.x[gv];
        }
        __syncthreads();
        const int nls = ((nv + 64 / g.L - 1) / (64 / g.L)) * 64;
        for (int ls = threadIdx.x; ls < nls; ls += blockDim.x) {
            int ov;
            bool live;
            SV<T, S> own;
            T acc[S];
            if constexpr (STORED)
                edge_rows<T, S, PERSP, KIND == ROW_DIAG>(g, prm, spv, b, nv, ls, ov, live, own,
                                                         acc);
            else
                lane_rows<T, S, KIND, FAST>(g, prm, sgeo, spv, b, nv, ls, ov, live, own, acc);
            if (!live || (ls & (g.L - 1)) != 0) continue;
            const int vtx = v0 + ov;
            if constexpr (STORED) {
                // rows D1 sum_w w_vw (x_w - x_v) + mdiag x_v; diagonal -D1 sum_w w_vw + mdiag
                // (K 1 = 0: the diagonal of K is minus the sum of its off-diagonal entries)
                const SV<T, S> md = reinterpret_cast<const SV<T, S>*>(prm.mdiag)[vtx];
#pragma unroll
                for (int s = 0; s < S; ++s)
                    acc[s] = KIND == ROW_DIAG ? md.v[s] - prm.sp[s].D1 * acc[s]
                                              : prm.sp[s].D1 * acc[s] + md.v[s] * own.v[s];
            } else if constexpr (FAST) {
                finish_fast<T, S, KIND>(prm, vtx, own, acc);
            }
            const unsigned bcb = prm.bcbits ? (unsigned)prm.bcbits[vtx] : 0u;
            SV<T, S> out{};
            if constexpr (MODE == MF_RESID) {
                const SV<double, S> bv = v.b[vtx];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const T val = ((bcb >> s) & 1u) ? own.v[s] : acc[s];
                    const double r = bv.v[s] - (double)val;
                    out.v[s] = (T)r;
                    part[s] += r * r;
                    part2[s] += bv.v[s] * bv.v[s];
                }
            } else if constexpr (MODE == MF_APPLY) {
#pragma unroll
                for (int s = 0; s < S; ++s) out.v[s] = ((bcb >> s) & 1u) ? own.v[s] : acc[s];
            } else if constexpr (MODE == MF_DIAG) {
#pragma unroll
                for (int s = 0; s < S; ++s) out.v[s] = ((bcb >> s) & 1u) ? T(1) : T(1) / acc[s];
            } else {  // MF_MASS
#pragma unroll
                for (int s = 0; s < S; ++s) out.v[s] = acc[s];
            }
            v.y[vtx] = out;
        }
        __syncthreads();  // LDS is reused by the next partition block
    }
    if constexpr (MODE == MF_RESID) {
        double tot[2 * S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            tot[s] = part[s];
            tot[S + s] = part2[s];
        }
        write_partials<2 * S>(tot, red, pout);
    }
}

// ---------------------------------------------------------------------------------------
// matrix-free Chronopoulos-Gear PCG iteration K_i (parity par = i & 1), or the INIT kernel
// (r_0 = (T) rin for the solved species, 0 otherwise; x = p = s_{-1} = 0; u_0 = r_0 / diag;
// w_0 = A u_0).  Partials per workgroup: gamma, delta, rho of the next iteration.
// The state snapshot the host polls, written by workgroup 0 of a chunk's last iteration into
// coherent pinned host memory: the fields, a system-scope fence, then seq = 1, which the host
// waits for (it clears seq before queuing the chunk).
__device__ __forceinline__ void publish_state(const CGState* __restrict__ st,
                                              CGState* __restrict__ poll) {
    if (!poll) return;
    CGState snap = *st;
    snap.seq = 0;
    *poll = snap;
    __threadfence_system();
    *reinterpret_cast<volatile int*>(&poll->seq) = 1;
}

template <class T, int S, bool FAST, bool INIT, bool STORED, bool PERSP = false>
__global__ void __launch_bounds__(MF_THREADS_MAX)
k_cg(Geo g, const Vec4<T>* __restrict__ xyz, OpParams<T> prm, CGVec<T, S> v,
     CGState* __restrict__ st, const T* __restrict__ pin, int nin, T* __restrict__ pout,
     int par, CGState* __restrict__ poll) {
    extern __shared__ __align__(32) unsigned char smem_raw[];
    Vec4<T>* sgeo = reinterpret_cast<Vec4<T>*>(smem_raw);
    // the stored form stages no coordinates
    SV<T, S>* spv = STORED ? reinterpret_cast<SV<T, S>*>(smem_raw)
                           : reinterpret_cast<SV<T, S>*>(sgeo + g.max_local);
    __shared__ double red[16 * 3 * MAXS];
    const int ml = g.max_local;

    bool act[S];
    T alpha[S], beta[S];
    bool any = false;
    const int rd = INIT ? 0 : par, wr = INIT ? 0 : par ^ 1;  // buffers read / written
    if constexpr (INIT) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            act[s] = st->act[1][s] != 0;
            any |= act[s];
            alpha[s] = beta[s] = T(0);
        }
        if (!any) return;
    } else {
        bool prev[S];
        bool anyprev = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            prev[s] = st->act[par ^ 1][s] != 0;
            anyprev |= prev[s];
        }
        const int k = st->kk[par ^ 1];
        if (!anyprev) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                for (int s = 0; s < S; ++s) st->act[par][s] = 0;
                st->kk[par] = k;
                st->any_active = 0;
                publish_state(st, poll);
            }
            return;
        }
        double sums[3 * S];
        consume_partials<3 * S, T>(pin, nin, sums, red);  // gamma_i, delta_i, rho_i
        double gam[S], alp[S];
        bool bad[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double gm = sums[s], dl = sums[S + s], rho = sums[2 * S + s];
            act[s] = prev[s] && (rho > st->tol2[s]) && (k < st->max_it) && isfinite(rho);
            double bb = 0.0, den = dl;
            if (k > 0) {
                bb = gm / st->rz[par ^ 1][s];
                den = dl - bb * gm / st->alp[par ^ 1][s];
            }
            bad[s] = act[s] && (!(gm > 0.0) || !(den > 0.0) || !isfinite(den) || !isfinite(bb));
            act[s] = act[s] && !bad[s];
            gam[s] = gm;
            alp[s] = act[s] ? gm / den : 0.0;
            alpha[s] = (T)alp[s];
            beta[s] = act[s] ? (T)bb : T(0);
            any |= act[s];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int s = 0; s < S; ++s) {
                st->act[par][s] = act[s] ? 1 : 0;
                if (!prev[s]) continue;
                st->rz[par][s] = gam[s];
                st->alp[par][s] = alp[s];
                st->rr[s] = sums[2 * S + s];
                if (act[s])
                    st->iters[s] += 1;
                else
                    st->status[s] = bad[s] ? ST_BREAKDOWN
                                           : (sums[2 * S + s] <= st->tol2[s] ? ST_CONVERGED : ST_MAXIT);
            }
            st->kk[par] = k + 1;
            st->any_active = any ? 1 : 0;
            publish_state(st, poll);
        }
        if (!any) return;
    }

    double pg[S], pd[S], pr[S];
#pragma unroll
    for (int s = 0; s < S; ++s) pg[s] = pd[s] = pr[s] = 0.0;
    const BlockRange br = my_blocks(g);
    for (int b = br.beg; b < br.end; b += br.step) {
        const int v0 = b * g.NV;
        const int nv = min(g.NV, g.V - v0);
        const int nloc = nv + (g.hptr[b + 1] - g.hptr[b]);
        // (1) stage u_{i+1} of every local vertex; owned vertices also advance p, x, s, r
        for (int i = threadIdx.x; i < nloc; i += blockDim.x) {
            const int gv = i < nv ? v0 + i : g.lvl[(size_t)b * ml + i];  // owned: contiguous
            if constexpr (!STORED) sgeo[i] = xyz[gv];
            const SV<T, S> d = v.dinv[gv];
            SV<T, S> rn{}, un{}, sn{};
            if constexpr (INIT) {
                const SV<double, S> ri = v.rin[gv];
#pragma unroll
                for (int s = 0; s < S; ++s) rn.v[s] = act[s] ? (T)ri.v[s] : T(0);
            } else {
                const SV<T, S> r = v.R[rd][gv], w = v.W[rd][gv], so = v.Sd[rd][gv];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    sn.v[s] = w.v[s] + beta[s] * so.v[s];
                    rn.v[s] = r.v[s] - alpha[s] * sn.v[s];
                }
                if (i < nv) {
                    SV<T, S> p = v.P[gv], x = v.X[gv];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        p.v[s] = r.v[s] * d.v[s] + beta[s] * p.v[s];
                        if (act[s]) x.v[s] += alpha[s] * p.v[s];
                    }
                    v.P[gv] = p;
                    v.X[gv] = x;
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) un.v[s] = rn.v[s] * d.v[s];
            spv[i] = un;
            if (i < nv) {
                v.R[wr][gv] = rn;
                v.Sd[wr][gv] = sn;  // INIT: s_{-1} = 0
                if constexpr (INIT) {
                    v.P[gv] = SV<T, S>{};
                    v.X[gv] = SV<T, S>{};
                }
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (act[s]) {
                        pg[s] += (double)rn.v[s] * (double)un.v[s];
                        pr[s] += (double)rn.v[s] * (double)rn.v[s];
                    }
            }
        }
        __syncthreads();
        // (2) w_{i+1} = A u_{i+1} for the owned vertices
        const int nls = ((nv + 64 / g.L - 1) / (64 / g.L)) * 64;
        for (int ls = threadIdx.x; ls < nls; ls += blockDim.x) {
            int ov;
            bool live;
            SV<T, S> own;
            T acc[S];
            if constexpr (STORED) {
                static_assert(std::is_same<T, float>::value && FAST, "stored geometry: FP32 fast form");
                edge_rows<T, S, PERSP>(g, prm, spv, b, nv, ls, ov, live, own, acc);
            } else {
                lane_rows<T, S, ROW_MATVEC, FAST>(g, prm, sgeo, spv, b, nv, ls, ov, live, own, acc);
            }
            if (!live || (ls & (g.L - 1)) != 0) continue;
            const int vtx = v0 + ov;
            if constexpr (STORED) {
                // D K(D=1) u + mdiag u
                const SV<T, S> md = reinterpret_cast<const SV<T, S>*>(prm.mdiag)[vtx];
#pragma unroll
                for (int s = 0; s < S; ++s) acc[s] = prm.sp[s].D1 * acc[s] + md.v[s] * own.v[s];
            } else if constexpr (FAST) {
                finish_fast<T, S, ROW_MATVEC>(prm, vtx, own, acc);
            }
            const unsigned bcb = prm.bcbits ? (unsigned)prm.bcbits[vtx] : 0u;
            SV<T, S> out{};
#pragma unroll
            for (int s = 0; s < S; ++s) {
                out.v[s] = ((bcb >> s) & 1u) ? own.v[s] : acc[s];
                if (act[s]) pd[s] += (double)out.v[s] * (double)own.v[s];
            }
            v.W[wr][vtx] = out;
        }
        __syncthreads();  // LDS is reused by the next partition block
    }
    double tot[3 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = pg[s];
        tot[S + s] = pd[s];
        tot[2 * S + s] = pr[s];
    }
    write_partials<3 * S, T>(tot, red, pout);
}

// ---------------------------------------------------------------------------------------
// Stored-geometry FP32 CG iteration K_i (i > 0), the same recurrences as k_cg, scheduled for
// latency: the loads that do not depend on this iteration's alpha / beta -- the vectors of the
// first PRE local vertices of each thread, the halo vertices' local-list entries, the
// edge-slice headers of the first PRE lane slots -- are issued before the previous
// iteration's partials are reduced, and those of the next partition block before the current
// one's rows are computed, so that the reduction, the gathers and the headers overlap.  Owned
// vertices are the contiguous range [v0, v0 + nv) and are addressed directly.  Local vertices
// or lane slots past PRE per thread (large blocks) are loaded in place.
constexpr int PRE = 2;

template <int S>
struct StagePre {
    int gv[PRE];
    SV<float, S> d[PRE], r[PRE], w[PRE], so[PRE], p[PRE], x[PRE];
};
struct SlotPre {
    int jbeg[PRE], width[PRE], ov[PRE];
};

template <int S>
__device__ __forceinline__ void stage_loads(const Geo& g, const CGVec<float, S>& v, int rd, int b,
                                            int i, int& gv, SV<float, S>& d, SV<float, S>& r,
                                            SV<float, S>& w, SV<float, S>& so, SV<float, S>& p,
                                            SV<float, S>& x) {
    const int v0 = b * g.NV;
    const int nv = min(g.NV, g.V - v0);
    gv = i < nv ? v0 + i : g.lvl[(size_t)b * g.max_local + i];
    d = v.dinv[gv];
    r = v.R[rd][gv];
    w = v.W[rd][gv];
    so = v.Sd[rd][gv];
    if (i < nv) {
        p = v.P[gv];
        x = v.X[gv];
    }
}

template <int S>
__device__ __forceinline__ void prefetch_stage(const Geo& g, const CGVec<float, S>& v, int rd,
                                               int b, StagePre<S>& pf) {
    const int v0 = b * g.NV;
    const int nv = min(g.NV, g.V - v0);
    const int nloc = nv + (g.hptr[b + 1] - g.hptr[b]);
#pragma unroll
    for (int j = 0; j < PRE; ++j) {
        const int i = threadIdx.x + j * blockDim.x;
        if (i < nloc)
            stage_loads<S>(g, v, rd, b, i, pf.gv[j], pf.d[j], pf.r[j], pf.w[j], pf.so[j], pf.p[j],
                           pf.x[j]);
    }
}

__device__ __forceinline__ void slot_header(const Geo& g, int b, int nv, int ls, int& jbeg,
                                            int& width, int& ov) {
    const int vps = 64 / g.L;
    const int spb = g.NV / vps;
    const int* sp0 = g.eptr + b * spb + (ls >> 6);
    jbeg = sp0[0] + (ls & 63);
    width = (sp0[1] - sp0[0]) >> 6;
    const int rank = ls / g.L;
    ov = rank < nv ? (int)g.elperm[b * g.NV + rank] : -1;
}

__device__ __forceinline__ void prefetch_slots(const Geo& g, int b, SlotPre& sp) {
    const int nv = min(g.NV, g.V - b * g.NV);
    const int nls = ((nv + 64 / g.L - 1) / (64 / g.L)) * 64;
#pragma unroll
    for (int j = 0; j < PRE; ++j) {
        const int ls = threadIdx.x + j * blockDim.x;
        if (ls < nls) slot_header(g, b, nv, ls, sp.jbeg[j], sp.width[j], sp.ov[j]);
    }
}

// sum over the slot's edges of K_vw(D=1) (u_w - u_v), reduced over the L lanes of the vertex;
// ov < 0 marks a lane past the block's vertices (it still takes part in the lane reduction)
template <int S, bool PERSP>
__device__ __forceinline__ void slot_rows(const Geo& g, const OpParams<float>& prm,
                                          const SV<float, S>* spv, int jbeg, int width, int ov,
                                          SV<float, S>& own, float (&acc)[S]) {
    const int o = ov < 0 ? 0 : ov;
    own = spv[o];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0f;
    for (int t0 = 0; t0 < width; t0 += 4) {
        int ix[4];
        float wt[PERSP ? S : 1][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = jbeg + 64 * min(t0 + u, width - 1);
            const bool in = t0 + u < width;
            ix[u] = in ? (int)g.eidx[j] : o;
#pragma unroll
            for (int q = 0; q < (PERSP ? S : 1); ++q) {
                const float wv = PERSP ? prm.sp[q].W[j] : g.ewt[j];
                wt[q][u] = in ? wv : 0.0f;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const SV<float, S> nb = spv[ix[u]];
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] += wt[PERSP ? s : 0][u] * (nb.v[s] - own.v[s]);
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (g.L >= 2) acc[s] = add_lane_xor1(acc[s]);
        if (g.L >= 4) acc[s] = add_lane_xor2(acc[s]);
    }
}

template <int S, bool PERSP>
__global__ void __launch_bounds__(MF_THREADS_MAX)
k_cg_st(Geo g, OpParams<float> prm, CGVec<float, S> v, CGState* __restrict__ st,
        const float* __restrict__ pin, int nin, float* __restrict__ pout, int par,
        CGState* __restrict__ poll) {
    using T = float;
    extern __shared__ __align__(32) unsigned char smem_raw[];
    SV<T, S>* spv = reinterpret_cast<SV<T, S>*>(smem_raw);
    __shared__ double red[16 * 3 * MAXS];
    const int rd = par, wr = par ^ 1;

    bool prev[S];
    bool anyprev = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        prev[s] = st->act[par ^ 1][s] != 0;
        anyprev |= prev[s];
    }
    const int k = st->kk[par ^ 1];
    if (!anyprev) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int s = 0; s < S; ++s) st->act[par][s] = 0;
            st->kk[par] = k;
            st->any_active = 0;
            publish_state(st, poll);
        }
        return;
    }
    const BlockRange br = my_blocks(g);
    StagePre<S> pf;
    SlotPre sl;
    if (br.beg < br.end) {
        prefetch_stage<S>(g, v, rd, br.beg, pf);
        prefetch_slots(g, br.beg, sl);
    }

    // scalars of this iteration from the previous kernel's partials (as in k_cg)
    bool act[S];
    T alpha[S], beta[S];
    bool any = false;
    {
        double sums[3 * S];
        consume_partials<3 * S, T>(pin, nin, sums, red);  // gamma_i, delta_i, rho_i
        double gam[S], alp[S];
        bool bad[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double gm = sums[s], dl = sums[S + s], rho = sums[2 * S + s];
            act[s] = prev[s] && (rho > st->tol2[s]) && (k < st->max_it) && isfinite(rho);
            double bb = 0.0, den = dl;
            if (k > 0) {
                bb = gm / st->rz[par ^ 1][s];
                den = dl - bb * gm / st->alp[par ^ 1][s];
            }
            bad[s] = act[s] && (!(gm > 0.0) || !(den > 0.0) || !isfinite(den) || !isfinite(bb));
            act[s] = act[s] && !bad[s];
            gam[s] = gm;
            alp[s] = act[s] ? gm / den : 0.0;
            alpha[s] = (T)alp[s];
            beta[s] = act[s] ? (T)bb : T(0);
            any |= act[s];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int s = 0; s < S; ++s) {
                st->act[par][s] = act[s] ? 1 : 0;
                if (!prev[s]) continue;
                st->rz[par][s] = gam[s];
                st->alp[par][s] = alp[s];
                st->rr[s] = sums[2 * S + s];
                if (act[s])
                    st->iters[s] += 1;
                else
                    st->status[s] = bad[s] ? ST_BREAKDOWN
                                           : (sums[2 * S + s] <= st->tol2[s] ? ST_CONVERGED : ST_MAXIT);
            }
            st->kk[par] = k + 1;
            st->any_active = any ? 1 : 0;
            publish_state(st, poll);
        }
    }
    if (!any) return;

    double pg[S], pd[S], pr[S];
#pragma unroll
    for (int s = 0; s < S; ++s) pg[s] = pd[s] = pr[s] = 0.0;
    for (int b = br.beg; b < br.end; b += br.step) {
        const int v0 = b * g.NV;
        const int nv = min(g.NV, g.V - v0);
        const int nloc = nv + (g.hptr[b + 1] - g.hptr[b]);
        // (1) stage u_{i+1}; owned vertices also advance p, x, s, r
        auto stage = [&](int i, int gv, const SV<T, S>& d, const SV<T, S>& r, const SV<T, S>& w,
                         const SV<T, S>& so, SV<T, S> p, SV<T, S> x) {
            SV<T, S> rn, un, sn;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                sn.v[s] = w.v[s] + beta[s] * so.v[s];
                rn.v[s] = r.v[s] - alpha[s] * sn.v[s];
                un.v[s] = rn.v[s] * d.v[s];
            }
            spv[i] = un;
            if (i < nv) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    p.v[s] = r.v[s] * d.v[s] + beta[s] * p.v[s];
                    if (act[s]) x.v[s] += alpha[s] * p.v[s];
                }
                v.P[gv] = p;
                v.X[gv] = x;
                v.R[wr][gv] = rn;
                v.Sd[wr][gv] = sn;
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (act[s]) {
                        pg[s] += (double)rn.v[s] * (double)un.v[s];
                        pr[s] += (double)rn.v[s] * (double)rn.v[s];
                    }
            }
        };
#pragma unroll
        for (int j = 0; j < PRE; ++j) {
            const int i = threadIdx.x + j * blockDim.x;
            if (i < nloc) stage(i, pf.gv[j], pf.d[j], pf.r[j], pf.w[j], pf.so[j], pf.p[j], pf.x[j]);
        }
        for (int i = threadIdx.x + PRE * blockDim.x; i < nloc; i += blockDim.x) {
            int gv;
            SV<T, S> d, r, w, so, p{}, x{};
            stage_loads<S>(g, v, rd, b, i, gv, d, r, w, so, p, x);
            stage(i, gv, d, r, w, so, p, x);
        }
        // the next block's staging inputs are read while this block's rows are computed
        const int bn = b + br.step;
        if (bn < br.end) prefetch_stage<S>(g, v, rd, bn, pf);
        __syncthreads();
        // (2) w_{i+1} = A u_{i+1} for the owned vertices
        const int nls = ((nv + 64 / g.L - 1) / (64 / g.L)) * 64;
        auto rows = [&](int ls, int jbeg, int width, int ov) {
            SV<T, S> own;
            T acc[S];
            slot_rows<S, PERSP>(g, prm, spv, jbeg, width, ov, own, acc);
            if (ov < 0 || (ls & (g.L - 1)) != 0) return;
            const int vtx = v0 + ov;
            const SV<T, S> md = reinterpret_cast<const SV<T, S>*>(prm.mdiag)[vtx];
            const unsigned bcb = prm.bcbits ? (unsigned)prm.bcbits[vtx] : 0u;
            SV<T, S> out;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const T a = prm.sp[s].D1 * acc[s] + md.v[s] * own.v[s];
                out.v[s] = ((bcb >> s) & 1u) ? own.v[s] : a;
                if (act[s]) pd[s] += (double)out.v[s] * (double)own.v[s];
            }
            v.W[wr][vtx] = out;
        };
#pragma unroll
        for (int j = 0; j < PRE; ++j) {
            const int ls = threadIdx.x + j * blockDim.x;
            if (ls < nls) rows(ls, sl.jbeg[j], sl.width[j], sl.ov[j]);
        }
        for (int ls = threadIdx.x + PRE * blockDim.x; ls < nls; ls += blockDim.x) {
            int jbeg, width, ov;
            slot_header(g, b, nv, ls, jbeg, width, ov);
            rows(ls, jbeg, width, ov);
        }
        if (bn < br.end) prefetch_slots(g, bn, sl);
        __syncthreads();  // LDS is reused by the next partition block
    }
    double tot[3 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = pg[s];
        tot[S + s] = pd[s];
        tot[2 * S + s] = pr[s];
    }
    write_partials<3 * S, T>(tot, red, pout);
}

// ---------------------------------------------------------------------------------------
// CSR (caller order) kernels of the plain FP64 path: CG apply (A_k) and residual.
constexpr int CSR_LANES = 8;
constexpr int CSR_THREADS = 512;
constexpr int CSR_ROWS = CSR_THREADS / CSR_LANES;

// A_k preamble of the CSR CG: reduce U_{k-1}'s partials, decide which species continue and
// their beta.  Returns false when no species is active.
template <int S>
__device__ __forceinline__ bool cg_begin(CGState* __restrict__ st, const double* __restrict__ pin,
                                         int nin, int par, bool (&act)[S], double (&beta)[S],
                                         double* red) {
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
        return false;
    }
    double sums[2 * S];
    consume_partials<2 * S>(pin, nin, sums, red);  // U_{k-1}: r.z, r.r per species
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
    return any;
}

enum CSRMode { CSR_CG = 0, CSR_RESID = 1 };

template <int S, int MODE>
__global__ void __launch_bounds__(CSR_THREADS)
k_csr(int V, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
      const double* __restrict__ A0, const double* __restrict__ A1, const double* __restrict__ A2,
      CSRVec<S> v, CGState* __restrict__ st, const double* __restrict__ pin, int nin,
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
    if constexpr (MODE == CSR_CG) {
        if (!cg_begin<S>(st, pin, nin, par, act, beta, red)) return;
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
            const SV<double, S> z = v.x[c];
            SV<double, S> po{};
            if constexpr (MODE == CSR_CG) po = v.pold[c];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                const double u = MODE == CSR_CG ? z.v[s] + beta[s] * po.v[s] : z.v[s];
                acc[s] += Av[s][j] * u;
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int off = CSR_LANES / 2; off > 0; off >>= 1) acc[s] += __shfl_xor(acc[s], off, 64);
        if (lane == 0) {
            if constexpr (MODE == CSR_CG) {
                const SV<double, S> z = v.x[row];
                const SV<double, S> po = v.pold[row];
                SV<double, S> p{}, q{};
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    p.v[s] = z.v[s] + beta[s] * po.v[s];
                    q.v[s] = acc[s];
                    if (act[s]) part[s] += p.v[s] * acc[s];
                }
                v.pnew[row] = p;
                v.y[row] = q;
            } else {
                const SV<double, S> bv = v.b[row];
                SV<double, S> r{};
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    r.v[s] = bv.v[s] - acc[s];
                    part[s] += r.v[s] * r.v[s];
                    part2[s] += bv.v[s] * bv.v[s];
                }
                v.y[row] = r;
            }
        }
    }
    constexpr int NQ = MODE == CSR_CG ? S : 2 * S;
    double tot[NQ];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = part[s];
        if constexpr (MODE != CSR_CG) tot[S + s] = part2[s];
    }
    write_partials<NQ>(tot, red, pout);
}

// U_k: x += alpha p; r -= alpha q; z = r * dinv; partials r.z, r.r (parity par = k & 1).
// INIT (U_{-1}): r = rin (0 for species not solved); x = 0; p_old = 0; z = r * dinv.
template <int S, bool INIT>
__global__ void __launch_bounds__(256)
k_update(int V, UVec<S> v, CGState* __restrict__ st, const double* __restrict__ pin, int nin,
         double* __restrict__ pout, int par) {
    __shared__ double red[16 * 2 * MAXS];
    const int nwg = gridDim.x;
    bool act[S];
    double alpha[S];
    bool any = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        act[s] = st->act[par][s] != 0;
        any |= act[s];
        alpha[s] = 0.0;
    }
    if (!any) return;
    if constexpr (!INIT) {
        double pq[S];
        consume_partials<S>(pin, nin, pq, red);  // A_k: p.q per species
        bool bad[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            bad[s] = act[s] && (!(pq[s] > 0.0) || !isfinite(pq[s]));
            alpha[s] = (act[s] && !bad[s]) ? st->rz[par][s] / pq[s] : 0.0;
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
        const SV<double, S> d = v.dinv[i];
        SV<double, S> r{}, z{}, x{};
        if constexpr (INIT) {
            const SV<double, S> ri = v.rin[i];
#pragma unroll
            for (int s = 0; s < S; ++s) r.v[s] = act[s] ? ri.v[s] : 0.0;
            v.pzero[i] = SV<double, S>{};
        } else {
            const SV<double, S> p = v.p[i], q = v.q[i];
            x = v.x[i];
            r = v.r[i];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!act[s]) continue;
                x.v[s] += alpha[s] * p.v[s];
                r.v[s] -= alpha[s] * q.v[s];
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            z.v[s] = r.v[s] * d.v[s];
            if (act[s]) {
                prz[s] += r.v[s] * z.v[s];
                prr[s] += r.v[s] * r.v[s];
            }
        }
        v.x[i] = x;
        v.r[i] = r;
        v.z[i] = z;
    }
    double tot[2 * S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        tot[s] = prz[s];
        tot[S + s] = prr[s];
    }
    write_partials<2 * S>(tot, red, pout);
}

// x64 += (double) d over the species in `mask` (refinement correction)
template <class T, int S>
__global__ void k_add_correction(int V, SV<double, S>* __restrict__ x,
                                 const SV<T, S>* __restrict__ d, int mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    SV<double, S> a = x[i];
    const SV<T, S> c = d[i];
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (mask & (1 << s)) a.v[s] += (double)c.v[s];
    x[i] = a;
}

// caller-order species arrays <-> species-interleaved records (internal order when perm)
template <class To, int S>
__global__ void k_pack(int V, const int32_t* __restrict__ perm, const double* __restrict__ s0,
                       const double* __restrict__ s1, const double* __restrict__ s2,
                       SV<To, S>* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int c = perm ? perm[i] : i;
    const double* src[3] = {s0, s1, s2};
    SV<To, S> r{};
#pragma unroll
    for (int s = 0; s < S; ++s) r.v[s] = src[s] ? (To)src[s][c] : To(0);
    out[i] = r;
}
// the FP64 start of a matrix-free solve: x in internal order with the Dirichlet values of each
// species (bit s of bcbits) in place
struct BcVals {
    double v[3];
};
template <int S>
__global__ void k_pack_start(int V, const int32_t* __restrict__ perm, const double* __restrict__ s0,
                             const double* __restrict__ s1, const double* __restrict__ s2,
                             const uint8_t* __restrict__ bcbits, BcVals bv,
                             SV<double, S>* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int c = perm[i];
    const unsigned bits = bcbits[i];
    const double* src[3] = {s0, s1, s2};
    SV<double, S> r{};
#pragma unroll
    for (int s = 0; s < S; ++s) r.v[s] = ((bits >> s) & 1u) ? bv.v[s] : src[s][c];
    out[i] = r;
}
// k_pack of the right-hand sides and k_pack_start of the start in one pass
template <int S>
__global__ void k_pack_bx(int V, const int32_t* __restrict__ perm, const double* __restrict__ b0,
                          const double* __restrict__ b1, const double* __restrict__ b2,
                          const double* __restrict__ s0, const double* __restrict__ s1,
                          const double* __restrict__ s2, const uint8_t* __restrict__ bcbits,
                          BcVals bv, SV<double, S>* __restrict__ bout,
                          SV<double, S>* __restrict__ xout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int c = perm[i];
    const unsigned bits = bcbits[i];
    const double* bsrc[3] = {b0, b1, b2};
    const double* src[3] = {s0, s1, s2};
    SV<double, S> rb, rx;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        rb.v[s] = bsrc[s][c];
        rx.v[s] = ((bits >> s) & 1u) ? bv.v[s] : src[s][c];
    }
    bout[i] = rb;
    xout[i] = rx;
}
// k_pack_bx with the lumped right-hand sides formed on the way (dsv_step_warm_history): per
// species b = lm (inv_dt u + f) from the start u (caller order), b = bc value on Dirichlet
// vertices, written to the caller's b and packed; the start packed with its Dirichlet values
template <int S>
__global__ void k_pack_rhs_bx(int V, const int32_t* __restrict__ perm,
                              const double* __restrict__ lm, double inv_dt,
                              double* __restrict__ b0, double* __restrict__ b1,
                              double* __restrict__ b2, const double* __restrict__ s0,
                              const double* __restrict__ s1, const double* __restrict__ s2,
                              const double* __restrict__ f0, const double* __restrict__ f1,
                              const double* __restrict__ f2, const uint8_t* __restrict__ bcbits,
                              BcVals bv, SV<double, S>* __restrict__ bout,
                              SV<double, S>* __restrict__ xout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int c = perm[i];
    const unsigned bits = bcbits[i];
    double* bdst[3] = {b0, b1, b2};
    const double* src[3] = {s0, s1, s2};
    const double* fs[3] = {f0, f1, f2};
    const double mc = lm[c];
    SV<double, S> rb, rx;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double u = src[s][c];
        const bool d = (bits >> s) & 1u;
        const double b = d ? bv.v[s] : lumped_rhs(mc, inv_dt, u, fs[s] ? fs[s][c] : 0.0);
        bdst[s][c] = b;
        rb.v[s] = b;
        rx.v[s] = d ? bv.v[s] : u;
    }
    bout[i] = rb;
    xout[i] = rx;
}
template <class Ti, int S>
__global__ void k_unpack(int V, const int32_t* __restrict__ perm, const SV<Ti, S>* __restrict__ in,
                         double* __restrict__ d0, double* __restrict__ d1, double* __restrict__ d2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int c = perm ? perm[i] : i;
    double* dst[3] = {d0, d1, d2};
    const SV<Ti, S> r = in[i];
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (dst[s]) dst[s][c] = (double)r.v[s];
}
// ---------------------------------------------------------------------------------------
// Species narrowing of the batched inner CG: the species sel[0..S2) of a workspace of S
// species continue in a workspace of S2 < S species (kernels of fewer species move fewer bytes
// and reduce fewer partials).  k_narrow copies their per-vertex operator data and either the
// FP64 right-hand side (before the first iteration) or, between chunks, the iteration vectors
// read by the next iteration (parity 0), its partials (zero-padded to the narrowed grid, so the
// sums, taken in the same per-thread order, are bit-identical) and the CG state.  k_widen copies
// the correction and the counters back.
template <class T, int S>
__device__ __forceinline__ T pick(const SV<T, S>& r, int s) {
    T x = r.v[0];
#pragma unroll
    for (int k = 1; k < S; ++k)
        if (s == k) x = r.v[k];
    return x;
}
template <class T, int S, int S2>
__device__ __forceinline__ SV<T, S2> pick_all(const SV<T, S>& r, const int (&sel)[S2]) {
    SV<T, S2> o;
#pragma unroll
    for (int i = 0; i < S2; ++i) o.v[i] = pick<T, S>(r, sel[i]);
    return o;
}

constexpr int NARROW_VECS = 5;  // X, P, R[0], W[0], Sd[0]

template <class T, int S, int S2>
struct NarrowArgs {
    int sel[S2];
    const SV<T, S>* dinv;
    SV<T, S2>* dinv2;
    const SV<T, S>* mdiag;  // null when the operator form has no mass diagonal
    SV<T, S2>* mdiag2;
    const uint8_t* bc;
    uint8_t* bc2;
    const SV<double, S>* rin;  // before the first iteration (else null)
    SV<double, S2>* rin2;
    const SV<T, S>* vec[NARROW_VECS];  // between chunks (else null)
    SV<T, S2>* vec2[NARROW_VECS];
    const T* pin;
    int nin;
    T* pin2;
    int nin2;
    const CGState* st;  // between chunks (else null)
    CGState* st2;
};

template <class T, int S, int S2>
__global__ void k_narrow(int V, NarrowArgs<T, S, S2> a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nth = gridDim.x * blockDim.x;
    for (int v = tid; v < V; v += nth) {
        a.dinv2[v] = pick_all<T, S, S2>(a.dinv[v], a.sel);
        if (a.mdiag) a.mdiag2[v] = pick_all<T, S, S2>(a.mdiag[v], a.sel);
        if (a.bc) {
            const unsigned b = a.bc[v];
            unsigned b2 = 0u;
#pragma unroll
            for (int i = 0; i < S2; ++i) b2 |= ((b >> a.sel[i]) & 1u) << i;
            a.bc2[v] = (uint8_t)b2;
        }
        if (a.rin) a.rin2[v] = pick_all<double, S, S2>(a.rin[v], a.sel);
        if (a.st)
#pragma unroll
            for (int k = 0; k < NARROW_VECS; ++k) a.vec2[k][v] = pick_all<T, S, S2>(a.vec[k][v], a.sel);
    }
    if (!a.st) return;
    // partials q * nin + b with q = c * S + s (c: gamma, delta, rho)
    for (int j = tid; j < 3 * S2 * a.nin2; j += nth) {
        const int q2 = j / a.nin2, b = j - q2 * a.nin2;
        const int c = q2 / S2, i = q2 - c * S2;
        a.pin2[j] = b < a.nin ? a.pin[(c * S + a.sel[i]) * a.nin + b] : T(0);
    }
    if (tid == 0) {
        const CGState& s = *a.st;
        CGState o = s;
        for (int i = 0; i < MAXS; ++i) {
            const bool in = i < S2;
            const int f = in ? a.sel[i] : 0;
            o.tol2[i] = in ? s.tol2[f] : 0.0;
            o.rr[i] = in ? s.rr[f] : 0.0;
            o.dead[i] = in ? s.dead[f] : 0;
            o.iters[i] = in ? s.iters[f] : 0;
            o.status[i] = in ? s.status[f] : ST_IDLE;
            for (int p = 0; p < 2; ++p) {
                o.rz[p][i] = in ? s.rz[p][f] : 0.0;
                o.alp[p][i] = in ? s.alp[p][f] : 0.0;
                o.act[p][i] = in ? s.act[p][f] : 0;
            }
        }
        *a.st2 = o;
    }
}

template <class T, int S, int S2>
__global__ void k_widen(int V, int s0, int s1, SV<T, S>* __restrict__ X,
                        const SV<T, S2>* __restrict__ X2, CGState* __restrict__ st,
                        const CGState* __restrict__ st2) {
    const int sel[2] = {s0, s1};
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < V) {
        SV<T, S> x = X[tid];
        const SV<T, S2> x2 = X2[tid];
#pragma unroll
        for (int k = 0; k < S; ++k)
#pragma unroll
            for (int i = 0; i < S2; ++i)
                if (sel[i] == k) x.v[k] = x2.v[i];
        X[tid] = x;
    }
    if (tid == 0) {
        for (int i = 0; i < S2; ++i) {
            st->iters[sel[i]] = st2->iters[i];
            st->status[sel[i]] = st2->status[i];
            st->rr[sel[i]] = st2->rr[i];
        }
        for (int s = 0; s < MAXS; ++s) st->act[0][s] = st->act[1][s] = 0;
        st->any_active = 0;
    }
}

// stored-form weights of a species with a D field: entry e takes the CSR value at its position
// (0 for padding entries)
template <class T>
__global__ void k_gather_weights(int64_t n, const int32_t* __restrict__ emap,
                                 const double* __restrict__ kvals, T* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int32_t j = emap[e];
    out[e] = j >= 0 ? (T)kvals[j] : T(0);
}

// per-ELL-entry coefficient of the entry's cell (0 for padding entries)
template <class T>
__global__ void k_inc_coef(int64_t n, const int32_t* __restrict__ ell_cell,
                           const int32_t* __restrict__ conn, const double* __restrict__ p, int mode,
                           double val, T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t e = ell_cell[i];
    out[i] = e >= 0 ? (T)cell_coef(p, mode, val, e, conn) : T(0);
}
__global__ void k_set_bc_bits(int n_bc, const int32_t* __restrict__ bc,
                              const int32_t* __restrict__ iperm, int s,
                              uint8_t* __restrict__ bits) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bc) return;
    const int v = iperm ? iperm[bc[k]] : bc[k];
    // one launch per species: concurrent writers of a byte set the same bit
    bits[v] = (uint8_t)(bits[v] | (1u << s));
}
template <int S>
__global__ void k_set_bc_values(int n_bc, const int32_t* __restrict__ bc,
                                const int32_t* __restrict__ iperm, int s, double value,
                                SV<double, S>* __restrict__ x) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bc) return;
    x[iperm ? iperm[bc[k]] : bc[k]].v[s] = value;
}
template <int S>
__global__ void k_diag_inv_csr(int V, const int32_t* __restrict__ diag_pos,
                               const double* __restrict__ A, int s,
                               SV<double, S>* __restrict__ dinv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    dinv[i].v[s] = 1.0 / A[diag_pos[i]];
}

// ---------------------------------------------------------------------------------------
// workspace

// species-record vectors of a workspace (CG precision unless noted)
enum VecId {
    VB_B64 = 0, VB_X64, VB_R64,  // FP64
    VB_DINV,
    VB_X,                        // correction (both CG forms)
    VB_V1, VB_V2, VB_V3, VB_V4, VB_V5, VB_V6, VB_V7,
    VB_COUNT
};
// CSR CG: r, z, q, p_a, p_b = V1..V5.  MF CG: R[2] = V1,V2; W[2] = V3,V4; S[2] = V5,V6; P = V7.

// Inputs an operator coefficient set / the Dirichlet bits were built from: with operator reuse
// on, a call whose inputs match skips the rebuild.
struct OpKey {
    int valid = 0;
    unsigned epoch = 0;
    int S = 0, lumped = 0, form = 0;
    double inv_dt = 0.0;
    const void* ptr[MAXS][3] = {};
    int32_t len[MAXS][3] = {};
    double val[MAXS][2] = {};
};
static bool same_key(const OpKey& a, const OpKey& b) {
    if (!a.valid || !b.valid || a.epoch != b.epoch || a.S != b.S || a.lumped != b.lumped ||
        a.form != b.form || a.inv_dt != b.inv_dt)
        return false;
    for (int s = 0; s < a.S; ++s)
        for (int c = 0; c < 3; ++c)
            if (a.ptr[s][c] != b.ptr[s][c] || a.len[s][c] != b.len[s][c] ||
                (c < 2 && a.val[s][c] != b.val[s][c]))
                return false;
    return true;
}
static OpKey make_key(const dsv_mesh* m, int S, const dsv_species* sp, double inv_dt, int lumped) {
    OpKey k;
    k.valid = 1;
    k.epoch = m->op_epoch;
    k.S = S;
    k.lumped = lumped ? 1 : 0;
    k.form = m->operator_form;
    k.inv_dt = inv_dt;
    for (int s = 0; s < S; ++s) {
        k.ptr[s][0] = sp[s].D;
        k.ptr[s][1] = sp[s].alpha;
        k.ptr[s][2] = sp[s].bc_dofs;
        k.len[s][0] = sp[s].D_len;
        k.len[s][1] = sp[s].alpha_len;
        k.len[s][2] = sp[s].n_bc;
        k.val[s][0] = sp[s].D_value;
        k.val[s][1] = sp[s].alpha_value;
    }
    return k;
}

struct WarmOut;

struct Workspace {
    int method = 0, S = 0, V = 0;
    hipStream_t stream = nullptr;
    CGState* st = nullptr;
    CGState* h_st = nullptr;  // pinned
    double* pA = nullptr;     // CSR: partials of A_k (S per workgroup)
    double* pU[2] = {};       // CSR: partials of U_k; MF: partials of K_i (3S per workgroup,
                              // in the CG precision)
    double* pR = nullptr;     // partials of the residual kernel (2S per workgroup)
    double* h_pR = nullptr;   // pinned copy
    CGState* h_poll = nullptr;  // pinned state snapshots after the two chunks in flight
    CGState* d_poll = nullptr;  // their device address (the MF CG kernels write them)
    hipEvent_t ev[2] = {};
    hipEvent_t ev_io[2] = {};  // ordering with the caller's stream: its work in, the solution out
    // operator coefficients: [0] kernels of the CG precision, [1] FP64 residual of the
    // refinement path; per-ELL-entry D then alpha per species, and the fast form's mass
    // diagonal (species records)
    void* coef[2][2][MAXS] = {};
    size_t coef_cap[2][2][MAXS] = {};
    void* mdiag[2] = {};
    uint8_t* bcbits = nullptr;
    unsigned char prm[sizeof(OpParams<double>)] = {};   // OpParams of the CG precision
    unsigned char prm2[sizeof(OpParams<double>)] = {};  // OpParams<double> of the residual
    void* vb[VB_COUNT] = {};
    double* csr[MAXS] = {};  // CSR64 values per species
    void* wd[2 * MAXD] = {};  // FP64 warm-start directions d_k and A d_k (dsv_solve_warm)
    // the previous dsv_solve_warm call's directions: buffer pair per direction, the arrays
    // each was formed from, its history count, and the key of the operator applied to them
    int wmap[MAXD] = {};
    const double* wpair[MAXD][MAXS][2] = {};
    int wnhist = 0;
    int wnextra = 0;
    int wvalid = 0;
    OpKey wkey;
    const double* wout_prev[MAXS] = {};  // the solution arrays the last solve wrote
    double* pW = nullptr;    // partials of the warm start's dot and norm kernels
    double* pT = nullptr;    // their totals
    double* gsave[2] = {};   // Gram matrices of the last call [gcur] and the one before
    int gcur = 0;
    int gnd = 0;             // directions of the saved one
    WarmOut* wout = nullptr;    // its coefficients and norms (device) ...
    WarmOut* h_wout = nullptr;  // ... published by the acceptance kernel into pinned memory
    WarmOut* d_wout_pub = nullptr;  // (h_wout's device address)
    ResidPub* h_rpub = nullptr;     // the residual norms and CG state, published (pinned)
    ResidPub* d_rpub = nullptr;     // (h_rpub's device address)
    void* fw[2][MAXS] = {};  // stored form: per-species weights K(D_s) of sets 0 / 1
    double* kvs[MAXS] = {};  // CSR values of K(D_s) (nnz) per species, assembled to build them
    // operator reuse: the inputs of coefficient sets 0 / 1, of the Dirichlet bits, and whether
    // the Jacobi diagonal (VB_DINV) belongs to set 0
    OpKey key[2];
    OpKey bckey;
    int dinv_valid = 0;
    // persistent grid of the MF CG kernels and the kernel variant it was sized for
    int cg_grid = 0;
    int cg_grid_key = 0;
    // CG iterations per species of the last solve's first inner solve [0] and of a later one
    // [1] (0: none), the first chunks' guide for the next solve
    int prior[2][MAXS] = {};
    // captured CG chunk and the parameters it was captured with
    hipGraphExec_t graph[2] = {};       // per poll slot (the MF kernels publish into it)
    hipGraphExec_t graph_tail[2] = {};  // the short chunk used near convergence
    hipGraphExec_t graph_tail_np = nullptr;  // the short chunk without the state publication
    int graph_chunk = 0;
    int graph_valid = 0;
    unsigned char graph_prm[sizeof(OpParams<double>)] = {};
    // low-mode subspace (dsv_build_low_modes): lm_n FP16 vectors per species (element (j, s, v)
    // at lm_index; internal order, 0 at Dirichlet vertices and in the padding to lm_vp), the
    // inverse of W^T A W per species under the FP64 operator, the operator key they belong
    // to, and per-solve scratch
    int lm_n = 0;
    int lm_vp = 0;              // V rounded up to a multiple of LM_GROUP
    _Float16* lm_w = nullptr;
    double* lm_ginv = nullptr;  // [S][lm_n][lm_n]
    double* lm_part = nullptr;  // dot partials
    double* lm_c = nullptr;     // coefficients [lm_n][S], then their dot totals
    double* lm_y = nullptr;     // the correction W c (V FP64 species records)
    OpKey lm_key;
};

static size_t rec_bytes(size_t t, int S) { return t * S; }

static void free_ws(Workspace* w) {
    if (!w) return;
    for (int i = 0; i < 2; ++i) {
        if (w->graph[i]) (void)hipGraphExecDestroy(w->graph[i]);
        if (w->graph_tail[i]) (void)hipGraphExecDestroy(w->graph_tail[i]);
    }
    if (w->graph_tail_np) (void)hipGraphExecDestroy(w->graph_tail_np);
    dfree(w->st);
    if (w->h_st) (void)hipHostFree(w->h_st);
    if (w->h_pR) (void)hipHostFree(w->h_pR);
    if (w->h_poll) (void)hipHostFree(w->h_poll);
    if (w->h_rpub) (void)hipHostFree(w->h_rpub);
    for (int i = 0; i < 2; ++i) {
        if (w->ev[i]) (void)hipEventDestroy(w->ev[i]);
        if (w->ev_io[i]) (void)hipEventDestroy(w->ev_io[i]);
    }
    dfree(w->pA);
    dfree(w->pU[0]);
    dfree(w->pU[1]);
    dfree(w->pR);
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c)
            for (int s = 0; s < MAXS; ++s) dfree(w->coef[a][c][s]);
    dfree(w->mdiag[0]);
    dfree(w->mdiag[1]);
    for (int s = 0; s < MAXS; ++s) dfree(w->csr[s]);
    for (int i = 0; i < 2 * MAXD; ++i) dfree(w->wd[i]);
    dfree(w->pW);
    dfree(w->pT);
    dfree(w->gsave[0]);
    dfree(w->gsave[1]);
    dfree(w->wout);
    if (w->h_wout) (void)hipHostFree(w->h_wout);
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < MAXS; ++q) dfree(w->fw[a][q]);
    for (auto& k : w->kvs) dfree(k);
    dfree(w->lm_w);
    dfree(w->lm_ginv);
    dfree(w->lm_part);
    dfree(w->lm_c);
    dfree(w->lm_y);
    dfree(w->bcbits);
    for (int i = 0; i < VB_COUNT; ++i) dfree(w->vb[i]);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
}

void free_workspaces(dsv_mesh* m) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 4; ++b) {
            free_ws(m->ws[a][b]);
            m->ws[a][b] = nullptr;
            free_ws(m->nws[a][b]);
            m->nws[a][b] = nullptr;
        }
}

static int max_lds_per_block(int device) {
    static int cached[64] = {};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    int v = 0;
    DSV_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    if (device >= 0 && device < 64) cached[device] = v;
    return v;
}

// LDS bytes of the MF kernels: coordinates and species input of every local vertex
template <class T>
static size_t mf_lds_bytes(const Partition& P, int S) {
    return (size_t)P.max_local * (sizeof(Vec4<T>) + rec_bytes(sizeof(T), S));
}

// launch grids (every partial-producing kernel has at most MAX_PROD workgroups)
static int mf_grid(const Partition& P) {
    const int per = (P.nb + 7) / 8;
    return 8 * std::min(per, MAX_PROD / 8);
}
static int csr_grid(const dsv_mesh* m) { return std::min((int)cdiv(m->V, CSR_ROWS), 1024); }
static int upd_grid(const dsv_mesh* m) {
    return std::max(1, std::min((int)cdiv(m->V, 1024), 1024));
}

static bool is_fp32(int method) { return method == DSV_METHOD_MF32_IR64; }

static Workspace* get_ws(dsv_mesh* m, int method, int S, bool narrowed = false) {
    Workspace*& w = narrowed ? m->nws[method][S] : m->ws[method][S];
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
        DSV_HIP(hipHostMalloc((void**)&w->h_poll, 2 * sizeof(CGState), hipHostMallocCoherent));
        DSV_HIP(hipHostGetDevicePointer((void**)&w->d_poll, w->h_poll, 0));
        DSV_HIP(hipHostMalloc((void**)&w->h_rpub, sizeof(ResidPub), hipHostMallocCoherent));
        DSV_HIP(hipHostGetDevicePointer((void**)&w->d_rpub, w->h_rpub, 0));
        for (int i = 0; i < 2; ++i) {
            DSV_HIP(hipEventCreateWithFlags(&w->ev[i], hipEventDisableTiming));
            DSV_HIP(hipEventCreateWithFlags(&w->ev_io[i], hipEventDisableTiming));
        }
        w->pA = dalloc<double>((size_t)MAXS * MAX_PROD);
        w->pU[0] = dalloc<double>((size_t)3 * MAXS * MAX_PROD);
        w->pU[1] = dalloc<double>((size_t)3 * MAXS * MAX_PROD);
        w->pR = dalloc<double>((size_t)2 * MAXS * MAX_PROD);
        const size_t V = m->V;
        const size_t rt = rec_bytes(is_fp32(method) ? sizeof(float) : sizeof(double), S);
        const size_t r64 = rec_bytes(sizeof(double), S);
        w->bcbits = dalloc<uint8_t>(V);
        for (int i = 0; i < VB_COUNT; ++i) {
            if (method == DSV_METHOD_CSR64 && (i == VB_V6 || i == VB_V7)) continue;
            DSV_HIP(hipMalloc(&w->vb[i], V * (i <= VB_R64 ? r64 : rt)));
        }
        if (method == DSV_METHOD_CSR64)
            for (int s = 0; s < S; ++s) w->csr[s] = dalloc<double>(m->nnz);
    } catch (...) {
        free_ws(w);
        w = nullptr;
        throw;
    }
    return w;
}

// Dirichlet bits (internal vertex order): bit s at the Dirichlet vertices of species s
static void prepare_bc(dsv_mesh* m, Workspace* w, int S, const dsv_species* sp, hipStream_t st) {
    OpKey k = make_key(m, S, sp, 0.0, 0);
    if (m->op_reuse && same_key(k, w->bckey)) return;
    w->bckey = k;
    w->bckey.valid = m->op_reuse;
    DSV_HIP(hipMemsetAsync(w->bcbits, 0, m->V, st));
    for (int s = 0; s < S; ++s)
        if (sp[s].n_bc > 0)
            hipLaunchKernelGGL(k_set_bc_bits, dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0, st,
                               sp[s].n_bc, sp[s].bc_dofs, m->iperm, s, w->bcbits);
    DSV_HIP(hipGetLastError());
}

static Geo make_geo(const dsv_mesh* m, const Partition& P) {
    Geo g;
    g.V = m->V;
    g.NV = P.NV;
    g.L = P.L;
    g.nb = P.nb;
    g.per_xcd = (P.nb + 7) / 8;
    g.max_local = P.max_local;
    g.hptr = P.hptr;
    g.lvl = P.lvl;
    g.lperm = P.lperm;
    g.sptr = P.sptr;
    g.ell = P.ell;
    g.eptr = P.eptr;
    g.elperm = P.elperm;
    g.eidx = P.eidx;
    g.ewt = P.ewt;
    g.ewt64 = P.ewt64;
    return g;
}

template <class T>
static const Vec4<T>* coords(const dsv_mesh* m) {
    return sizeof(T) == 4 ? (const Vec4<T>*)(const void*)m->xyz32
                          : (const Vec4<T>*)(const void*)m->xyz64;
}

template <class T, int S>
static size_t checked_lds(const dsv_mesh* m, const Partition& P, bool stored = false) {
    const size_t lds = stored ? (size_t)P.max_local * rec_bytes(sizeof(T), S)
                              : mf_lds_bytes<T>(P, S);
    if (lds + 1024 > (size_t)max_lds_per_block(m->device))
        throw ArgError{"matrix-free partition needs " + std::to_string(lds) +
                       " B of LDS per workgroup; use fewer vertices_per_block"};
    return lds;
}

template <class T, int S, int MODE>
static void launch_mf(const dsv_mesh* m, const Partition& P, const unsigned char* prm,
                      const MFVec<T, S>& v, double* pout, hipStream_t stream) {
    const Geo g = make_geo(m, P);
    OpParams<T> p;
    std::memcpy(&p, prm, sizeof(p));
    if constexpr (MODE == MF_APPLY || MODE == MF_RESID || MODE == MF_DIAG) {
        if (p.stored) {
            const size_t lds_st = checked_lds<T, S>(m, P, true);
            if (p.persp)
                hipLaunchKernelGGL((k_mf<T, S, MODE, true, true, true>), dim3(mf_grid(P)),
                                   dim3(P.threads), lds_st, stream, g, coords<T>(m), p, v, pout);
            else
                hipLaunchKernelGGL((k_mf<T, S, MODE, true, true, false>), dim3(mf_grid(P)),
                                   dim3(P.threads), lds_st, stream, g, coords<T>(m), p, v, pout);
            return;
        }
    }
    if (MODE != MF_MASS && p.no_generic)
        throw std::logic_error("generic operator kernel on parameters prepared for the stored form");
    const size_t lds = checked_lds<T, S>(m, P);
    if constexpr (MODE != MF_MASS) {
        if (p.fast) {
            hipLaunchKernelGGL((k_mf<T, S, MODE, true>), dim3(mf_grid(P)), dim3(P.threads), lds,
                               stream, g, coords<T>(m), p, v, pout);
            return;
        }
    }
    hipLaunchKernelGGL((k_mf<T, S, MODE, false>), dim3(mf_grid(P)), dim3(P.threads), lds, stream, g,
                       coords<T>(m), p, v, pout);
}

// Builds the operator parameters of precision T in coefficient set `set`: per-ELL-entry cell
// coefficients (D and alpha, cell-averaged like the assembly) and, for the fast form (lumped
// mass, scalar D for every species), the mass diagonal computed by the generic MASS kernel.
// Operator parameters of set `set` (0: the CG precision; 1: the FP64 residual of mf32_ir64).
// Returns true when it assembled the K(D_s) of the species with a D field into w->kvs (which the
// set-1 call of the same solve then gathers again, k_ready, instead of assembling them twice).
template <class T, int S>
static bool prepare_op(dsv_mesh* m, Workspace* w, const Partition& P, const dsv_species* sp,
                       double inv_dt, int lumped, hipStream_t st, int set, bool k_ready = false) {
    const OpKey key = make_key(m, S, sp, inv_dt, lumped);
    if (m->op_reuse && same_key(key, w->key[set])) return false;  // prm / prm2 still hold the set
    w->key[set] = key;
    w->key[set].valid = m->op_reuse;
    if (set == 0) w->dinv_valid = 0;
    OpParams<T> h{};
    h.inv_dt = (T)inv_dt;
    h.lumped = lumped ? 1 : 0;
    h.bcbits = w->bcbits;
    bool fast = lumped != 0;
    for (int s = 0; s < S; ++s) fast = fast && sp[s].D == nullptr;
    const bool stored = lumped != 0 && m->operator_form == DSV_FORM_STORED && P.eptr;
    // Where only stored-form kernels run on these parameters (the FP32 CG, and the FP64 residual
    // and apply of set 1), a D field needs no per-entry values: the generic kernels are then
    // refused (no_generic), except the mass diagonal, which reads the reaction alone.
    const bool d_entries = !(stored && (std::is_same<T, float>::value || set == 1));
    const size_t n = (size_t)P.n_entries;
    for (int s = 0; s < S; ++s) {
        const double* src[2] = {sp[s].D, sp[s].alpha};
        const int32_t len[2] = {sp[s].D_len, sp[s].alpha_len};
        const T** out[2] = {&h.sp[s].D, &h.sp[s].A};
        for (int c = 0; c < 2; ++c) {
            *out[c] = nullptr;
            if (!src[c]) continue;  // a scalar: D1 / A0 in the kernels
            if (c == 0 && !d_entries) {
                h.no_generic = 1;
                continue;
            }
            void*& buf = w->coef[set][c][s];
            size_t& cap = w->coef_cap[set][c][s];
            if (cap < n * sizeof(T)) {
                dfree(buf);
                cap = 0;
                DSV_HIP(hipMalloc(&buf, std::max<size_t>(1, n) * sizeof(T)));
                cap = n * sizeof(T);
            }
            hipLaunchKernelGGL((k_inc_coef<T>), dim3(cdiv(n, 256)), dim3(256), 0, st, (int64_t)n,
                               P.ell_cell, m->conn, src[c], coef_mode(src[c], len[c], m), 0.0,
                               (T*)buf);
            *out[c] = (const T*)buf;
        }
        h.sp[s].A0 = (T)(sp[s].alpha ? 0.0 : sp[s].alpha_value);
        h.sp[s].D6 = (T)(sp[s].D_value / 6.0);
        h.sp[s].D1 = (T)sp[s].D_value;
    }
    h.fast = 0;
    h.stored = 0;
    h.persp = 0;
    h.mdiag = nullptr;
    if (fast || stored) {
        if (!w->mdiag[set]) DSV_HIP(hipMalloc(&w->mdiag[set], (size_t)m->V * rec_bytes(sizeof(T), S)));
        MFVec<T, S> v{};
        v.y = (SV<T, S>*)w->mdiag[set];
        unsigned char hb[sizeof(OpParams<double>)] = {};
        std::memcpy(hb, &h, sizeof(h));
        launch_mf<T, S, MF_MASS>(m, P, hb, v, nullptr, st);
        h.fast = fast ? 1 : 0;
        h.mdiag = w->mdiag[set];
        h.stored = stored ? 1 : 0;
    }
    bool assembled = false;
    if (stored && !fast) {
        // some species has a D field: every species gets its own weights, K(D_s) per ELL entry
        // for a field (its off-diagonal CSR values, no mass, reaction or Dirichlet rows, gathered
        // through the entries' CSR positions), the unit weights times D for a scalar
        h.persp = 1;
        const size_t ne = (size_t)P.n_edge_entries;
        for (int s = 0; s < S; ++s) {
            if (sp[s].D == nullptr) {
                h.sp[s].W = (const T*)(const void*)(sizeof(T) == 4 ? (const void*)P.ewt
                                                                   : (const void*)P.ewt64);
                continue;
            }
            if (!w->kvs[s]) w->kvs[s] = dalloc<double>((size_t)m->nnz);
            if (!w->fw[set][s]) DSV_HIP(hipMalloc(&w->fw[set][s], std::max<size_t>(1, ne) * sizeof(T)));
            if (!k_ready) {
                dsv_species ks{};
                ks.D = sp[s].D;
                ks.D_len = sp[s].D_len;
                ks.D_value = sp[s].D_value;
                assemble_into(m, ks, 0.0, 1, w->kvs[s], st);
                assembled = true;
            }
            hipLaunchKernelGGL((k_gather_weights<T>), dim3(cdiv(ne, 256)), dim3(256), 0, st,
                               (int64_t)ne, P.emap, w->kvs[s], (T*)w->fw[set][s]);
            h.sp[s].W = (const T*)w->fw[set][s];
            h.sp[s].D1 = T(1);
        }
    }
    std::memcpy(set ? w->prm2 : w->prm, &h, sizeof(h));
    DSV_HIP(hipGetLastError());
    return assembled;
}

template <class T, int S>
static CGVec<T, S> cg_vecs(Workspace* w) {
    CGVec<T, S> v{};
    v.dinv = (const SV<T, S>*)w->vb[VB_DINV];
    v.R[0] = (SV<T, S>*)w->vb[VB_V1];
    v.R[1] = (SV<T, S>*)w->vb[VB_V2];
    v.W[0] = (SV<T, S>*)w->vb[VB_V3];
    v.W[1] = (SV<T, S>*)w->vb[VB_V4];
    v.Sd[0] = (SV<T, S>*)w->vb[VB_V5];
    v.Sd[1] = (SV<T, S>*)w->vb[VB_V6];
    v.P = (SV<T, S>*)w->vb[VB_V7];
    v.X = (SV<T, S>*)w->vb[VB_X];
    v.rin = (const SV<double, S>*)w->vb[VB_R64];
    return v;
}

// Persistent grid of the MF CG kernels: the partition's workgroup count capped at the number of
// workgroups of the iteration kernel K_i the device keeps resident (its occupancy times the
// CUs), so that no second round of workgroups queues behind the first.  INIT uses the same grid
// (K_0 reads its partials).  Sized once per workspace and kernel variant.
template <class T, int S>
static int cg_grid(dsv_mesh* m, Workspace* w) {
    const Partition& P = m->part;
    OpParams<T> p;
    std::memcpy(&p, w->prm, sizeof(p));
    const bool stored = std::is_same<T, float>::value && p.stored;
    const int key = stored ? (p.persp ? 4 : 1) : (p.fast ? 2 : 3);
    if (w->cg_grid_key == key) return w->cg_grid;
    const void* f = nullptr;
    size_t lds = 0;
    if constexpr (std::is_same<T, float>::value) {
        if (stored) {
            f = p.persp ? reinterpret_cast<const void*>(&k_cg_st<S, true>)
                        : reinterpret_cast<const void*>(&k_cg_st<S, false>);
            lds = checked_lds<T, S>(m, P, true);
        }
    }
    if (!f) {
        f = p.fast ? reinterpret_cast<const void*>(&k_cg<T, S, true, false, false>)
                   : reinterpret_cast<const void*>(&k_cg<T, S, false, false, false>);
        lds = checked_lds<T, S>(m, P);
    }
    int per_cu = 0, ncu = 0;
    DSV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, P.threads, lds));
    DSV_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device));
    int n = mf_grid(P);
    const int cap = (per_cu * ncu) & ~7;  // a multiple of 8: the XCD mapping
    if (cap >= 8) n = std::min(n, cap);
    w->cg_grid = n;
    w->cg_grid_key = key;
    return n;
}

// MF CG kernel: INIT (writes the partials K_0 reads, pU[0]) or K_i with parity par (which
// publishes the state to `poll` when it is not null)
template <class T, int S, bool INIT>
static void launch_cg(dsv_mesh* m, Workspace* w, int par, hipStream_t stream,
                      CGState* poll = nullptr) {
    const Partition& P = m->part;
    const Geo g = make_geo(m, P);
    OpParams<T> p;
    std::memcpy(&p, w->prm, sizeof(p));
    const CGVec<T, S> v = cg_vecs<T, S>(w);
    const int n = cg_grid<T, S>(m, w);
    const T* pin = INIT ? nullptr : (const T*)w->pU[par];
    T* pout = (T*)(INIT ? w->pU[0] : w->pU[par ^ 1]);
    if constexpr (std::is_same<T, float>::value) {
        if (p.stored) {
            const size_t lds_st = checked_lds<T, S>(m, P, true);
            if constexpr (!INIT) {
                if (p.persp)
                    hipLaunchKernelGGL((k_cg_st<S, true>), dim3(n), dim3(P.threads), lds_st,
                                       stream, g, p, v, w->st, pin, n, pout, par, poll);
                else
                    hipLaunchKernelGGL((k_cg_st<S, false>), dim3(n), dim3(P.threads), lds_st,
                                       stream, g, p, v, w->st, pin, n, pout, par, poll);
            } else {
                if (p.persp)
                    hipLaunchKernelGGL((k_cg<T, S, true, INIT, true, true>), dim3(n),
                                       dim3(P.threads), lds_st, stream, g, coords<T>(m), p, v,
                                       w->st, pin, n, pout, par, poll);
                else
                    hipLaunchKernelGGL((k_cg<T, S, true, INIT, true, false>), dim3(n),
                                       dim3(P.threads), lds_st, stream, g, coords<T>(m), p, v,
                                       w->st, pin, n, pout, par, poll);
            }
            return;
        }
    }
    if (p.no_generic)
        throw std::logic_error("generic CG kernel on parameters prepared for the stored form");
    const size_t lds = checked_lds<T, S>(m, P);
    if (p.fast)
        hipLaunchKernelGGL((k_cg<T, S, true, INIT, false>), dim3(n), dim3(P.threads), lds, stream, g,
                           coords<T>(m), p, v, w->st, pin, n, pout, par, poll);
    else
        hipLaunchKernelGGL((k_cg<T, S, false, INIT, false>), dim3(n), dim3(P.threads), lds, stream, g,
                           coords<T>(m), p, v, w->st, pin, n, pout, par, poll);
}

template <int S, int MODE>
static void launch_csr(const dsv_mesh* m, Workspace* w, const CSRVec<S>& v, const double* pin,
                       int nin, double* pout, int par, hipStream_t stream) {
    const double* A0 = w->csr[0];
    const double* A1 = S > 1 ? w->csr[1] : A0;
    const double* A2 = S > 2 ? w->csr[2] : A0;
    hipLaunchKernelGGL((k_csr<S, MODE>), dim3(csr_grid(m)), dim3(CSR_THREADS), 0, stream, m->V,
                       m->indptr, m->indices, A0, A1, A2, v, w->st, pin, nin, pout, par);
}

template <int S, bool INIT>
static void launch_update(dsv_mesh* m, Workspace* w, int parity, hipStream_t stream) {
    UVec<S> u{};
    u.x = (SV<double, S>*)w->vb[VB_X];
    u.r = (SV<double, S>*)w->vb[VB_V1];
    u.z = (SV<double, S>*)w->vb[VB_V2];
    u.dinv = (const SV<double, S>*)w->vb[VB_DINV];
    u.p = (const SV<double, S>*)(parity ? w->vb[VB_V4] : w->vb[VB_V5]);  // p_new of A_k
    u.q = (const SV<double, S>*)w->vb[VB_V3];
    u.pzero = (SV<double, S>*)w->vb[VB_V4];  // INIT clears the p_old of A_0 (parity 0)
    u.rin = (const SV<double, S>*)w->vb[VB_R64];
    hipLaunchKernelGGL((k_update<S, INIT>), dim3(upd_grid(m)), dim3(256), 0, stream, m->V, u, w->st,
                       w->pA, csr_grid(m), w->pU[parity], parity);
}

// one CSR CG iteration (A_k + U_k) with parity k & 1: z in x, q in y, p_old / p_new alternate
template <int S>
static void csr_iteration(dsv_mesh* m, Workspace* w, int parity, hipStream_t stream) {
    CSRVec<S> v{};
    v.x = (const SV<double, S>*)w->vb[VB_V2];
    v.y = (SV<double, S>*)w->vb[VB_V3];
    v.pold = (const SV<double, S>*)(parity ? w->vb[VB_V5] : w->vb[VB_V4]);
    v.pnew = (SV<double, S>*)(parity ? w->vb[VB_V4] : w->vb[VB_V5]);
    launch_csr<S, CSR_CG>(m, w, v, w->pU[parity ^ 1], upd_grid(m), w->pA, parity, stream);
    launch_update<S, false>(m, w, parity, stream);
}

// one CG iteration of the workspace's method (iteration parity `parity`); the matrix-free
// kernel publishes its state to `poll` when that is not null
template <class T, int S>
static void cg_iteration(dsv_mesh* m, Workspace* w, int parity, hipStream_t stream,
                         CGState* poll = nullptr) {
    if constexpr (std::is_same<T, double>::value) {
        if (w->method == DSV_METHOD_CSR64) {
            csr_iteration<S>(m, w, parity, stream);
            return;
        }
    }
    launch_cg<T, S, false>(m, w, parity, stream, poll);
}

// Poll history of run_cg handed to the narrowed run that continues it (species remapped)
struct PollSeed {
    CGState last{};  // state at the most recent completed chunk end
    double rr0[MAXS] = {0, 0, 0};  // r.r and iteration counts at the chunk end before it
    int it0[MAXS] = {0, 0, 0};
    bool have_last = false, have0 = false;
    int pending = 0;   // iterations queued after `last` ahead of the narrowed run
    int launched = 0;  // iterations launched so far
    int prior[MAXS] = {0, 0, 0};  // iterations the species took in the previous such solve
};

// A narrowed run starts only if it is predicted to need at least this many more iterations
// (its copies cost about one iteration of the wider kernel)
constexpr int NARROW_MIN = 4;

template <class T, int S>
static int run_cg(dsv_mesh* m, Workspace* w, int max_it, hipStream_t stream,
                  const PollSeed* seed = nullptr);

// The workspace of the S2 species sel[] of w, with its operator parameters: those species'
// coefficients, and its own mass diagonal and Dirichlet bits (which k_narrow fills).
template <class T, int S, int S2>
static Workspace* narrow_ws(dsv_mesh* m, Workspace* w, const int* sel) {
    Workspace* w2 = get_ws(m, w->method, S2, true);
    OpParams<T> p, p2;
    std::memcpy(&p, w->prm, sizeof(p));
    p2 = p;
    for (int i = 0; i < MAXS; ++i) p2.sp[i] = i < S2 ? p.sp[sel[i]] : SpeciesP<T>{};
    if (p.mdiag && !w2->mdiag[0])
        DSV_HIP(hipMalloc(&w2->mdiag[0], (size_t)m->V * rec_bytes(sizeof(T), S2)));
    p2.mdiag = p.mdiag ? w2->mdiag[0] : nullptr;
    p2.bcbits = p.bcbits ? w2->bcbits : nullptr;
    std::memset(w2->prm, 0, sizeof(w2->prm));
    std::memcpy(w2->prm, &p2, sizeof(p2));
    return w2;
}

template <class T, int S, int S2>
static NarrowArgs<T, S, S2> narrow_args(Workspace* w, Workspace* w2, const int* sel) {
    NarrowArgs<T, S, S2> a{};
    for (int i = 0; i < S2; ++i) a.sel[i] = sel[i];
    OpParams<T> p;
    std::memcpy(&p, w->prm, sizeof(p));
    a.dinv = (const SV<T, S>*)w->vb[VB_DINV];
    a.dinv2 = (SV<T, S2>*)w2->vb[VB_DINV];
    a.mdiag = (const SV<T, S>*)p.mdiag;
    a.mdiag2 = p.mdiag ? (SV<T, S2>*)w2->mdiag[0] : nullptr;
    a.bc = p.bcbits;
    a.bc2 = p.bcbits ? w2->bcbits : nullptr;
    return a;
}

template <class T, int S, int S2>
static void launch_narrow(dsv_mesh* m, const NarrowArgs<T, S, S2>& a, hipStream_t st) {
    bool ok = a.dinv && a.dinv2 && (!a.mdiag || a.mdiag2) && (!a.bc || a.bc2) && (!a.rin || a.rin2);
    if (a.st)
        for (int k = 0; k < NARROW_VECS; ++k) ok = ok && a.vec[k] && a.vec2[k];
    ok = ok && (!a.st || (a.st2 && a.pin && a.pin2 && a.nin2 >= a.nin));
    if (!ok) throw std::runtime_error("species narrowing: a workspace vector is missing");
    hipLaunchKernelGGL((k_narrow<T, S, S2>), dim3(cdiv(m->V, 256)), dim3(256), 0, st, m->V, a);
    DSV_HIP(hipGetLastError());
}

template <class T, int S, int S2>
static void widen(dsv_mesh* m, Workspace* w, Workspace* w2, const int* sel, hipStream_t st) {
    static_assert(S2 <= 2, "k_widen takes two species");
    hipLaunchKernelGGL((k_widen<T, S, S2>), dim3(cdiv(m->V, 256)), dim3(256), 0, st, m->V, sel[0],
                       S2 > 1 ? sel[1] : 0, (SV<T, S>*)w->vb[VB_X], (const SV<T, S2>*)w2->vb[VB_X],
                       w->st, (const CGState*)w2->st);
    DSV_HIP(hipGetLastError());
}

static CGState remap_state(const CGState& s, const int* sel, int S2) {
    CGState o = s;
    for (int i = 0; i < MAXS; ++i) {
        const bool in = i < S2;
        const int f = in ? sel[i] : 0;
        o.tol2[i] = in ? s.tol2[f] : 0.0;
        o.rr[i] = in ? s.rr[f] : 0.0;
        o.iters[i] = in ? s.iters[f] : 0;
        o.status[i] = in ? s.status[f] : ST_IDLE;
    }
    return o;
}

// Continues run_cg on w with the kernel of the S2 species sel[] (the chunk queued last on w
// leaves the next iteration at parity 0).  Returns false, with nothing queued, when the
// narrowed kernel's persistent grid is smaller than w's: its partials could not take w's.
template <class T, int S, int S2>
static bool narrow_mid(dsv_mesh* m, Workspace* w, const int* sel, int max_it, hipStream_t st,
                       const PollSeed& seed) {
    if (w->method == DSV_METHOD_CSR64) return false;
    Workspace* w2 = narrow_ws<T, S, S2>(m, w, sel);
    const int n = cg_grid<T, S>(m, w), n2 = cg_grid<T, S2>(m, w2);
    if (n2 < n) return false;
    NarrowArgs<T, S, S2> a = narrow_args<T, S, S2>(w, w2, sel);
    const CGVec<T, S> v = cg_vecs<T, S>(w);
    const CGVec<T, S2> v2 = cg_vecs<T, S2>(w2);
    const SV<T, S>* in[NARROW_VECS] = {v.X, v.P, v.R[0], v.W[0], v.Sd[0]};
    SV<T, S2>* out[NARROW_VECS] = {v2.X, v2.P, v2.R[0], v2.W[0], v2.Sd[0]};
    for (int k = 0; k < NARROW_VECS; ++k) {
        a.vec[k] = in[k];
        a.vec2[k] = out[k];
    }
    a.pin = (const T*)w->pU[0];
    a.nin = n;
    a.pin2 = (T*)w2->pU[0];
    a.nin2 = n2;
    a.st = w->st;
    a.st2 = w2->st;
    launch_narrow<T, S, S2>(m, a, st);
    m->stat_narrowings += 1;
    PollSeed s2 = seed;
    s2.last = remap_state(seed.last, sel, S2);
    for (int i = 0; i < MAXS; ++i) {
        s2.rr0[i] = i < S2 ? seed.rr0[sel[i]] : 0.0;
        s2.it0[i] = i < S2 ? seed.it0[sel[i]] : 0;
        s2.prior[i] = i < S2 ? seed.prior[sel[i]] : 0;
    }
    run_cg<T, S2>(m, w2, max_it, st, &s2);
    widen<T, S, S2>(m, w, w2, sel, st);
    return true;
}

// Spins until a kernel sets the published word `seq` in pinned host memory.  The spin checks
// the stream now and then, so that work that ends without publishing is an error, not a hang.
static void wait_published(volatile int* seq, hipStream_t stream, const char* what) {
    for (long i = 1;; ++i) {
        if (*seq != 0) break;
        if ((i & 1023) == 0) {
            const hipError_t q = hipStreamQuery(stream);
            if (q == hipSuccess) {
                if (*seq != 0) break;
                throw std::runtime_error(std::string(what) + " ended without publishing its result");
            }
            if (q != hipErrorNotReady) DSV_HIP(q);
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
}

// Runs CG iterations until no species is active.  Chunks of an even number of iterations are
// launched directly or, with use_graphs, replayed from hipGraphs captured once (re-captured
// when the operator parameters change); the host polls one published state per chunk.  Once
// some species have finished and the rest are predicted to need NARROW_MIN or more iterations,
// the rest continue with the kernel of fewer species (narrow_mid).
template <class T, int S>
static int run_cg(dsv_mesh* m, Workspace* w, int max_it, hipStream_t stream, const PollSeed* seed) {
    const int chunk = std::max(2, m->graph_chunk & ~1);
    const int tail = std::max(2, (chunk / 4) & ~1);
    // the MF kernels publish the state the host polls themselves; the CSR path copies it
    const bool kpoll = w->method != DSV_METHOD_CSR64;
    int launched = seed ? seed->launched : 0;
    if (m->use_graphs && (!w->graph_valid || w->graph_chunk != chunk ||
                          std::memcmp(w->graph_prm, w->prm, sizeof(w->prm)) != 0)) {
        for (hipGraphExec_t* ge : {&w->graph[0], &w->graph_tail[0], &w->graph[1], &w->graph_tail[1],
                                   &w->graph_tail_np}) {
            if (*ge) (void)hipGraphExecDestroy(*ge);
            *ge = nullptr;
        }
        w->graph_valid = 0;
        // captured on the workspace's own stream, which has no work of its own queued (a
        // narrowed workspace is captured while the wider run's chunks are in flight on `stream`)
        // per poll slot a full and a short chunk whose last kernel publishes into that slot,
        // and a short chunk that publishes nothing (leading parts of a longer short-chunk run)
        for (int slot = 0; slot < 3; ++slot)
            for (int which = 0; which < 2; ++which) {
                if (slot == 2 && which == 0) continue;
                const int len = which ? tail : chunk;
                CGState* poll = kpoll && slot < 2 ? &w->d_poll[slot] : nullptr;
                hipGraph_t gr;
                DSV_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < len; ++i)
                    cg_iteration<T, S>(m, w, i & 1, w->stream, i == len - 1 ? poll : nullptr);
                DSV_HIP(hipStreamEndCapture(w->stream, &gr));
                hipGraphExec_t* ge = slot == 2 ? &w->graph_tail_np
                                               : (which ? &w->graph_tail[slot] : &w->graph[slot]);
                DSV_HIP(hipGraphInstantiate(ge, gr, nullptr, nullptr, 0));
                DSV_HIP(hipGraphDestroy(gr));
            }
        w->graph_valid = 1;
        w->graph_chunk = chunk;
        std::memcpy(w->graph_prm, w->prm, sizeof(w->prm));
    }
    // Chunk k+1 is queued before the host waits for chunk k's state, so the device never idles
    // on the poll; once every species is done the queued chunk's kernels return at their first
    // check.  From the r.r of the species at two chunk ends the host predicts how many
    // iterations remain, and queues short chunks once less than a full chunk is predicted, so
    // that few no-op iterations follow convergence.
    auto enqueue = [&](int slot, int len) {
        if (kpoll) w->h_poll[slot].seq = 0;
        if (m->use_graphs) {
            if (len == chunk)
                DSV_HIP(hipGraphLaunch(w->graph[slot], stream));
            else
                for (int i = 0; i < len; i += tail)  // only the last part publishes
                    DSV_HIP(hipGraphLaunch(i + tail < len ? w->graph_tail_np : w->graph_tail[slot],
                                           stream));
        } else {
            for (int i = 0; i < len; ++i)
                cg_iteration<T, S>(m, w, i & 1, stream,
                                   kpoll && i == len - 1 ? &w->d_poll[slot] : nullptr);
        }
        launched += len;
        if (!kpoll) {
            DSV_HIP(hipMemcpyAsync(&w->h_poll[slot], w->st, sizeof(CGState), hipMemcpyDeviceToHost,
                                   stream));
            DSV_HIP(hipEventRecord(w->ev[slot], stream));
        }
        return len;
    };
    // Waits for the chunk of `slot`: its event (CSR path), or the seq word its last kernel
    // publishes (no event in the stream between chunks).  The spin checks the stream now and
    // then, so that a chunk that ends without publishing is an error, not a hang.
    auto wait_slot = [&](int slot) {
        if (!kpoll) {
            DSV_HIP(hipEventSynchronize(w->ev[slot]));
            return;
        }
        wait_published(&w->h_poll[slot].seq, stream, "a CG chunk");
    };
    PollSeed h;  // `last`, the chunk end before it, and iterations queued after `last`
    if (seed) h = *seed;
    // Iterations each running species still needs after `inflight` more than `last` has
    // (negative: predicted done within them), from its r.r at the last two chunk ends or,
    // before that, from the iterations it took in the previous such solve: the fewest and the
    // most over the species; known = false while some species has neither.
    struct Pred {
        bool known;       // for every running species
        bool any;         // for some running species (fewest: over those)
        double fewest, most;
    };
    auto predict = [&](int inflight) -> Pred {
        Pred p{h.have_last, false, 1e300, -1e300};
        if (!p.known) return p;
        for (int s = 0; s < S; ++s) {
            if (h.last.status[s] != ST_RUNNING) continue;
            double r = 0.0;
            const int dn = h.last.iters[s] - h.it0[s];
            if (h.have0 && dn > 0 && h.last.rr[s] > 0.0 && h.rr0[s] > h.last.rr[s]) {
                const double rate = std::log(h.last.rr[s] / h.rr0[s]) / dn;  // per iteration, < 0
                r = std::log(h.last.tol2[s] / h.last.rr[s]) / rate - inflight;
            } else if (h.prior[s] > 0) {
                r = h.prior[s] - h.last.iters[s] - inflight;
            } else {
                p.known = false;
                continue;
            }
            p.any = true;
            p.fewest = std::min(p.fewest, r);
            p.most = std::max(p.most, r);
        }
        return p;
    };
    bool may_narrow = S > 1 && m->narrow != 0 && w->method != DSV_METHOD_CSR64;
    // A chunk ends near the next predicted event: the last species finishing (few no-op
    // iterations follow convergence) or, where narrowing may follow, the first one finishing
    // (the rest move to the narrower kernel soon after).  Lengths are multiples of the short
    // chunk up to a full one.
    auto next_len = [&](int inflight) {
        const Pred p = predict(inflight);
        if (!(may_narrow ? p.any : p.known)) return chunk;
        const double e = may_narrow ? p.fewest : p.most;
        if (e >= chunk) return chunk;
        const int n = std::max(1, (int)std::ceil(std::max(e, 0.0) / tail));
        return std::min(chunk, n * tail);
    };
    int len_q[2] = {0, 0};
    len_q[0] = enqueue(0, next_len(h.pending));
    for (int k = 0;; ++k) {
        const int cur = k & 1, nxt = cur ^ 1;
        // chunk k (in flight) runs len_q[cur] more iterations before chunk k + 1
        const int inflight = len_q[cur] + h.pending;
        len_q[nxt] = launched <= max_it + 2 * chunk ? enqueue(nxt, next_len(inflight)) : 0;
        wait_slot(cur);
        if (h.have_last) {
            for (int s = 0; s < S; ++s) {
                h.rr0[s] = h.last.rr[s];
                h.it0[s] = h.last.iters[s];
            }
            h.have0 = true;
        }
        h.last = w->h_poll[cur];  // copied before the next enqueue reuses this slot
        h.have_last = true;
        h.pending = 0;
        if (!h.last.any_active || launched > max_it + 2 * chunk) break;
        if constexpr (S > 1) {
            if (may_narrow) {
                int sel[MAXS] = {0, 0, 0}, n = 0;
                for (int s = 0; s < S; ++s)
                    if (h.last.status[s] == ST_RUNNING) sel[n++] = s;
                const Pred p = predict(len_q[nxt]);
                if (n > 0 && n < S && (!p.known || p.most >= NARROW_MIN)) {
                    PollSeed sd = h;
                    sd.pending = len_q[nxt];
                    sd.launched = launched;
                    bool done = false;
                    if (n == 1) {
                        done = narrow_mid<T, S, 1>(m, w, sel, max_it, stream, sd);
                    } else {
                        if constexpr (S == 3) done = narrow_mid<T, 3, 2>(m, w, sel, max_it, stream, sd);
                    }
                    if (done) break;
                    may_narrow = false;
                }
            }
        }
    }
    m->stat_launches += launched - (seed ? seed->launched : 0);
    return launched;
}

// Start of an inner solve: tolerances, active species and the iteration limit, written into the
// CG state on the stream (no host round trip)
struct CGStart {
    double tol2[MAXS];
    int active[MAXS];
    int max_it;
};
__global__ void k_cg_start(CGState* __restrict__ st, CGStart a) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int s = 0; s < MAXS; ++s) {
        st->tol2[s] = a.tol2[s];
        st->act[1][s] = a.active[s];
        st->act[0][s] = 0;
        st->dead[s] = 0;
        st->iters[s] = 0;
        st->status[s] = a.active[s] ? ST_RUNNING : ST_IDLE;
    }
    st->k = 0;
    st->kk[0] = st->kk[1] = 0;
    st->max_it = a.max_it;
    st->any_active = 1;
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
// Totals of the residual kernel's partials (r.r, then b.b, of every species; nprod each), one
// wavefront per quantity in a fixed order (lane-strided sums, then the DPP wave sum), and with
// `st` a snapshot of the CG state, written into the pinned slot `pub`; `seq` is set last.
// One workgroup of 64 * 2S threads.
template <int S>
__global__ __launch_bounds__(64 * 2 * MAXS) void k_resid_publish(const double* __restrict__ pR,
                                                                int nprod,
                                                                const CGState* __restrict__ st,
                                                                ResidPub* __restrict__ pub) {
    __shared__ double tot[2 * MAXS];
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (q < 2 * S) {  // wave-uniform
        double a = 0.0;
        for (int i = lane; i < nprod; i += 64) a += pR[(size_t)q * nprod + i];
        a = wave_sum(a);
        if (lane == 0) tot[q] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int s = 0; s < S; ++s) {
            pub->rr[s] = tot[s];
            pub->bb[s] = tot[S + s];
        }
        if (st) {
            CGState snap = *st;
            snap.seq = 0;
            pub->state = snap;
        }
        __threadfence_system();
        *reinterpret_cast<volatile int*>(&pub->seq) = 1;
    }
}

// residual r64 = b - A x64 (FP64, vertex order of the method); returns rr and bb per species
template <int S>
static int residual64_launch(dsv_mesh* m, Workspace* w, hipStream_t stream) {
    int nprod;
    if (w->method == DSV_METHOD_CSR64) {
        CSRVec<S> v{};
        v.x = (const SV<double, S>*)w->vb[VB_X64];
        v.y = (SV<double, S>*)w->vb[VB_R64];
        v.b = (const SV<double, S>*)w->vb[VB_B64];
        launch_csr<S, CSR_RESID>(m, w, v, nullptr, 0, w->pR, 0, stream);
        nprod = csr_grid(m);
    } else {
        MFVec<double, S> v{};
        v.x = (const SV<double, S>*)w->vb[VB_X64];
        v.y = (SV<double, S>*)w->vb[VB_R64];
        v.b = (const SV<double, S>*)w->vb[VB_B64];
        launch_mf<double, S, MF_RESID>(m, m->part, w->method == DSV_METHOD_MF32_IR64 ? w->prm2 : w->prm,
                                       v, w->pR, stream);
        nprod = mf_grid(m->part);
    }
    DSV_HIP(hipGetLastError());
    return nprod;
}

// (with `state`, the CG state left by the previous inner solve comes back in the same round trip)
template <int S>
static void residual64(dsv_mesh* m, Workspace* w, hipStream_t stream, double* rr, double* bb,
                       CGState* state = nullptr) {
    const int nprod = residual64_launch<S>(m, w, stream);
    volatile int* seq = &w->h_rpub->seq;
    *seq = 0;  // every earlier publication has been waited for
    hipLaunchKernelGGL((k_resid_publish<S>), dim3(1), dim3(64 * 2 * S), 0, stream, w->pR, nprod,
                       state ? (const CGState*)w->st : nullptr, w->d_rpub);
    DSV_HIP(hipGetLastError());
    wait_published(seq, stream, "the residual");
    const ResidPub& p = *w->h_rpub;
    if (state) *state = p.state;
    for (int s = 0; s < S; ++s) {
        rr[s] = p.rr[s];
        bb[s] = p.bb[s];
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

template <class T, int S>
static void inner_cg(dsv_mesh* m, Workspace* w, const double* tol2, const int* active, int max_it,
                     const double* rr0, const int* prior, hipStream_t st);

// inner CG of the S2 species sel[] of w (the others are not active) in the narrowed workspace
template <class T, int S, int S2>
static void inner_narrow(dsv_mesh* m, Workspace* w, const int* sel, const double* tol2,
                         int max_it, const double* rr0, const int* prior, hipStream_t st) {
    if (w->method == DSV_METHOD_CSR64) throw std::logic_error("inner_narrow: CSR method");
    Workspace* w2 = narrow_ws<T, S, S2>(m, w, sel);
    NarrowArgs<T, S, S2> a = narrow_args<T, S, S2>(w, w2, sel);
    a.rin = (const SV<double, S>*)w->vb[VB_R64];
    a.rin2 = (SV<double, S2>*)w2->vb[VB_R64];
    launch_narrow<T, S, S2>(m, a, st);
    m->stat_narrowings += 1;
    double t2[MAXS] = {0, 0, 0}, r2[MAXS] = {0, 0, 0};
    int a2[MAXS] = {0, 0, 0}, p2[MAXS] = {0, 0, 0};
    for (int i = 0; i < S2; ++i) {
        t2[i] = tol2[sel[i]];
        a2[i] = 1;
        r2[i] = rr0[sel[i]];
        p2[i] = prior[sel[i]];
    }
    inner_cg<T, S2>(m, w2, t2, a2, max_it, r2, p2, st);
    widen<T, S, S2>(m, w, w2, sel, st);
}

// inner CG on A d = r64 (d starts at 0), T = CG precision; the matrix-free methods solve only
// the active species' system when some are not active (inner_narrow).  rr0: r64.r64 per
// species (the CG's first r.r); prior: iterations per species of the previous such solve
// (0 = none), the first chunks' guide.
template <class T, int S>
static void inner_cg(dsv_mesh* m, Workspace* w, const double* tol2, const int* active, int max_it,
                     const double* rr0, const int* prior, hipStream_t st) {
    if constexpr (S > 1) {
        if (m->narrow && w->method != DSV_METHOD_CSR64) {
            int sel[MAXS] = {0, 0, 0}, n = 0;
            for (int s = 0; s < S; ++s)
                if (active[s]) sel[n++] = s;
            if (n == 1) {
                inner_narrow<T, S, 1>(m, w, sel, tol2, max_it, rr0, prior, st);
                return;
            }
            if constexpr (S == 3) {
                if (n == 2) {
                    inner_narrow<T, 3, 2>(m, w, sel, tol2, max_it, rr0, prior, st);
                    return;
                }
            }
        }
    }
    CGStart a{};
    for (int s = 0; s < MAXS; ++s) {
        a.tol2[s] = s < S ? tol2[s] : 0.0;
        a.active[s] = (s < S && active[s]) ? 1 : 0;
    }
    a.max_it = max_it;
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(64), 0, st, w->st, a);
    if (w->method == DSV_METHOD_CSR64)
        launch_update<S, true>(m, w, 1, st);  // U_{-1}
    else
        launch_cg<T, S, true>(m, w, 0, st);
    DSV_HIP(hipGetLastError());
    PollSeed seed;
    seed.have_last = true;  // the state before the first iteration
    for (int s = 0; s < MAXS; ++s) {
        const bool on = s < S && active[s];
        seed.last.status[s] = on ? ST_RUNNING : ST_IDLE;
        seed.last.rr[s] = on ? rr0[s] : 0.0;
        seed.last.tol2[s] = on ? tol2[s] : 0.0;
        seed.prior[s] = on ? prior[s] : 0;
    }
    run_cg<T, S>(m, w, max_it, st, &seed);
}

// Galerkin warm start folded into a solve (dsv_solve_warm): directions from past solutions
// (hist) and explicit ones (extra); outputs as dsv_warm_start's.
struct WarmArgs {
    int n_hist = 0;
    const double* const* hist = nullptr;
    bool from_hist = false;  // x0 = hist_0, d_k = hist_k - hist_(k+1) (n_hist + 1 arrays)
    int n_extra = 0;
    const double* const* extra = nullptr;
    double* coef = nullptr;
    double* r_before = nullptr;
    double* r_after = nullptr;
    // set by the solve: the low-mode correction follows the Galerkin start (its rtol decides
    // which species already meet the tolerance)
    bool low_modes = false;
    double rtol = 0.0;
    // dsv_step_warm_history: the lumped right-hand sides are formed from hist_0 (and the
    // per-vertex sources rhs_f, nullable) into the species' b while the solve packs them
    bool form_rhs = false;
    const double* const* rhs_f = nullptr;
};
template <int S>
static void fused_warm(dsv_mesh* m, Workspace* w, const dsv_species* sp, const WarmArgs& wa,
                       int nres, double* rr, double* bb, hipStream_t st);
static bool low_modes_valid(const dsv_mesh* m, const Workspace* w);

// The solution arrays a solve is about to write: the warm-start directions of another
// workspace may have been formed from them, and this workspace's next call must not take a
// direction formed from them (their contents change).  A solve that formed directions
// (`formed`) keeps them and records its outputs for that check; any other solve drops the
// directions of its workspace too, since they may name arrays it writes.
static void note_outputs(dsv_mesh* m, Workspace* w, double* const* out, bool formed) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 4; ++b)
            if (m->ws[a][b] && (m->ws[a][b] != w || !formed)) m->ws[a][b]->wvalid = 0;
    if (w)
        for (int s = 0; s < MAXS; ++s) w->wout_prev[s] = out[s];
}

// Runs on the workspace's stream, ordered after the work queued on `caller` and, at return,
// ahead of any later work on `caller` (events; the host waits only for the residual norms).
template <int S>
static void solve_impl(dsv_mesh* m, const dsv_species* sp, double inv_dt, int lumped, int method,
                       double rtol, int max_it, dsv_stats* stats, hipStream_t caller,
                       const WarmArgs* warm = nullptr) {
    Workspace* w = get_ws(m, method, S);
    hipStream_t st = w->stream;
    const auto t0 = std::chrono::steady_clock::now();
    DSV_HIP(hipEventRecord(w->ev_io[0], caller));
    DSV_HIP(hipStreamWaitEvent(st, w->ev_io[0], 0));
    m->stat_launches = 0;
    m->stat_narrowings = 0;
    m->stat_dir_reused = 0;
    m->stat_gram_reused = 0;
    m->stat_low_modes = 0;
    const bool internal = method != DSV_METHOD_CSR64;
    const int32_t* perm = internal ? m->perm : nullptr;
    const unsigned gv = cdiv(m->V, 256);
    const double* bs[3] = {sp[0].b, S > 1 ? sp[1].b : nullptr, S > 2 ? sp[2].b : nullptr};
    const double* xs[3] = {sp[0].x, S > 1 ? sp[1].x : nullptr, S > 2 ? sp[2].x : nullptr};
    if (warm && warm->from_hist)
        for (int s = 0; s < S; ++s) xs[s] = warm->hist[s];  // the start is step n of the history
    SV<double, S>* x64 = (SV<double, S>*)w->vb[VB_X64];
    if (internal) {
        // the right-hand sides, and the start with its Dirichlet values from the Dirichlet
        // bits of the operator set-up, in one pass
        prepare_bc(m, w, S, sp, st);
        BcVals bv{};
        for (int s = 0; s < S; ++s) bv.v[s] = sp[s].bc_value;
        if (warm && warm->form_rhs) {
            const double* const* f = warm->rhs_f;
            hipLaunchKernelGGL((k_pack_rhs_bx<S>), dim3(gv), dim3(256), 0, st, m->V, perm,
                               (const double*)m->lumped_m, inv_dt, const_cast<double*>(bs[0]),
                               const_cast<double*>(bs[1]), const_cast<double*>(bs[2]), xs[0],
                               xs[1], xs[2], f ? f[0] : nullptr, f && S > 1 ? f[1] : nullptr,
                               f && S > 2 ? f[2] : nullptr, (const uint8_t*)w->bcbits, bv,
                               (SV<double, S>*)w->vb[VB_B64], x64);
        } else {
            hipLaunchKernelGGL((k_pack_bx<S>), dim3(gv), dim3(256), 0, st, m->V, perm, bs[0],
                               bs[1], bs[2], xs[0], xs[1], xs[2], (const uint8_t*)w->bcbits, bv,
                               (SV<double, S>*)w->vb[VB_B64], x64);
        }
    } else {
        hipLaunchKernelGGL((k_pack<double, S>), dim3(gv), dim3(256), 0, st, m->V, perm, bs[0],
                           bs[1], bs[2], (SV<double, S>*)w->vb[VB_B64]);
        hipLaunchKernelGGL((k_pack<double, S>), dim3(gv), dim3(256), 0, st, m->V, perm, xs[0],
                           xs[1], xs[2], x64);
        for (int s = 0; s < S; ++s)
            if (sp[s].n_bc > 0)
                hipLaunchKernelGGL((k_set_bc_values<S>), dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0,
                                   st, sp[s].n_bc, sp[s].bc_dofs, (const int32_t*)nullptr, s,
                                   sp[s].bc_value, x64);
    }
    // operator set-up: assembled values / coefficients and the Jacobi diagonal
    if (method == DSV_METHOD_CSR64) {
        SV<double, S>* dinv = (SV<double, S>*)w->vb[VB_DINV];
        for (int s = 0; s < S; ++s) {
            assemble_into(m, sp[s], inv_dt, lumped, w->csr[s], st);
            hipLaunchKernelGGL((k_diag_inv_csr<S>), dim3(gv), dim3(256), 0, st, m->V, m->diag_pos,
                               w->csr[s], s, dinv);
        }
    } else if (method == DSV_METHOD_MF64) {
        prepare_op<double, S>(m, w, m->part, sp, inv_dt, lumped, st, 0);
        if (!w->dinv_valid) {
            MFVec<double, S> v{};
            v.y = (SV<double, S>*)w->vb[VB_DINV];
            launch_mf<double, S, MF_DIAG>(m, m->part, w->prm, v, nullptr, st);
            w->dinv_valid = m->op_reuse;
        }
    } else {
        const bool k_ready = prepare_op<float, S>(m, w, m->part, sp, inv_dt, lumped, st, 0);
        prepare_op<double, S>(m, w, m->part, sp, inv_dt, lumped, st, 1, k_ready);
        if (!w->dinv_valid) {
            MFVec<float, S> v{};
            v.y = (SV<float, S>*)w->vb[VB_DINV];
            launch_mf<float, S, MF_DIAG>(m, m->part, w->prm, v, nullptr, st);
            w->dinv_valid = m->op_reuse;
        }
    }
    DSV_HIP(hipGetLastError());

    double tol_abs[MAXS] = {0, 0, 0}, rel0[MAXS] = {0, 0, 0}, relf[MAXS] = {0, 0, 0};
    int conv[MAXS] = {0, 0, 0}, outer[MAXS] = {0, 0, 0}, dead[MAXS] = {0, 0, 0};
    int iters_total[MAXS] = {0, 0, 0};
    int prev_act[MAXS] = {0, 0, 0};  // species of the previous inner solve (none in the first sweep)
    const int max_rounds = method == DSV_METHOD_MF32_IR64 ? m->max_outer : 4;
    for (int round = 0;; ++round) {
        double rr[MAXS] = {0, 0, 0}, bb[MAXS] = {0, 0, 0};
        CGState h2{};
        const bool warm0 = round == 0 && warm && warm->n_hist + warm->n_extra > 0;
        if (warm0) {
            // x64 and r64 move to the Galerkin start (rr: its residual); should that already
            // meet the tolerance, the true residual confirms it
            WarmArgs wa = *warm;
            wa.low_modes = method == DSV_METHOD_MF32_IR64 && low_modes_valid(m, w);
            wa.rtol = rtol;
            fused_warm<S>(m, w, sp, wa, residual64_launch<S>(m, w, st), rr, bb, st);
            bool met = true;
            for (int s = 0; s < S; ++s) met &= std::sqrt(rr[s]) <= rtol * std::sqrt(bb[s]);
            if (met) residual64<S>(m, w, st, rr, bb);
        } else {
            residual64<S>(m, w, st, rr, bb, round > 0 ? &h2 : nullptr);
        }
        if (round > 0) {
            // the state's iteration counters restart in every inner solve; accumulate them
            for (int s = 0; s < S; ++s) {
                w->prior[std::min(round - 1, 1)][s] = prev_act[s] ? h2.iters[s] : 0;
                if (!prev_act[s]) continue;
                iters_total[s] += h2.iters[s];
                if (h2.status[s] == ST_BREAKDOWN) dead[s] = 1;
            }
        }
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
            // The FP32 sweep's target is floored at half the FP64 tolerance: below that, its
            // recursive residual parts from the true one (FP32 operator rounding amplified by
            // the condition number), and on the 806k-tet disc the true residual after the sweep
            // ends near 0.85 rtol ||b|| for floors 0.3 and 0.5 alike, the latter with 7% fewer
            // iterations.
            const double rn = std::sqrt(rr[s]);
            const double t = method == DSV_METHOD_MF32_IR64
                                 ? std::max(m->inner_rtol * rn, 0.5 * tol_abs[s])
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
        const int* prior = w->prior[std::min(round, 1)];
        if (method == DSV_METHOD_MF32_IR64) {
            inner_cg<float, S>(m, w, tol2, act, std::max(1, budget), rr, prior, st);
            hipLaunchKernelGGL((k_add_correction<float, S>), dim3(gv), dim3(256), 0, st, m->V, x64,
                               (const SV<float, S>*)w->vb[VB_X], mask);
        } else {
            inner_cg<double, S>(m, w, tol2, act, std::max(1, budget), rr, prior, st);
            hipLaunchKernelGGL((k_add_correction<double, S>), dim3(gv), dim3(256), 0, st, m->V, x64,
                               (const SV<double, S>*)w->vb[VB_X], mask);
        }
        DSV_HIP(hipGetLastError());
        for (int s = 0; s < S; ++s) prev_act[s] = act[s];
    }
    double* xo[3] = {sp[0].x, S > 1 ? sp[1].x : nullptr, S > 2 ? sp[2].x : nullptr};
    hipLaunchKernelGGL((k_unpack<double, S>), dim3(gv), dim3(256), 0, st, m->V, perm,
                       (const SV<double, S>*)x64, xo[0], xo[1], xo[2]);
    DSV_HIP(hipGetLastError());
    note_outputs(m, w, xo, warm && warm->n_hist + warm->n_extra > 0);
    // the solution is ordered on the caller's stream (no host wait for the unpack)
    DSV_HIP(hipEventRecord(w->ev_io[1], st));
    DSV_HIP(hipStreamWaitEvent(caller, w->ev_io[1], 0));
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
        stats->cg_launches = m->stat_launches;
        stats->narrowings = m->stat_narrowings;
        stats->dirs_reused = m->stat_dir_reused;
        stats->gram_reused = m->stat_gram_reused;
        stats->low_modes = m->stat_low_modes;
    }
}

// ||b - A x||^2 and ||b||^2 per species for x = sp[s].x with x[bc] = bc_value (FP64,
// matrix-free), the check dsv_solve makes before its first iteration
template <int S>
static void residual_norms_impl(dsv_mesh* m, const dsv_species* sp, double inv_dt, int lumped,
                               double* rr, double* bb) {
    Workspace* w = get_ws(m, DSV_METHOD_MF64, S);
    hipStream_t st = w->stream;
    const unsigned gv = cdiv(m->V, 256);
    const double* bs[3] = {sp[0].b, S > 1 ? sp[1].b : nullptr, S > 2 ? sp[2].b : nullptr};
    const double* xs[3] = {sp[0].x, S > 1 ? sp[1].x : nullptr, S > 2 ? sp[2].x : nullptr};
    SV<double, S>* x64 = (SV<double, S>*)w->vb[VB_X64];
    hipLaunchKernelGGL((k_pack<double, S>), dim3(gv), dim3(256), 0, st, m->V, m->perm, bs[0], bs[1],
                       bs[2], (SV<double, S>*)w->vb[VB_B64]);
    hipLaunchKernelGGL((k_pack<double, S>), dim3(gv), dim3(256), 0, st, m->V, m->perm, xs[0], xs[1],
                       xs[2], x64);
    for (int s = 0; s < S; ++s)
        if (sp[s].n_bc > 0)
            hipLaunchKernelGGL((k_set_bc_values<S>), dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0, st,
                               sp[s].n_bc, sp[s].bc_dofs, m->iperm, s, sp[s].bc_value, x64);
    prepare_bc(m, w, S, sp, st);
    prepare_op<double, S>(m, w, m->part, sp, inv_dt, lumped, st, 0);
    residual64<S>(m, w, st, rr, bb);
}

// ---------------------------------------------------------------------------------------
// Galerkin warm start: per species, x = x0 + sum_k c_k d_k with c minimising the A-energy
// norm of the error over x0 + span{d_k}.  The directions are zero at Dirichlet vertices and
// x0 holds the Dirichlet values there, so only the free block of A -- symmetric positive
// definite -- enters, and c solves (d_j . A d_k) c_k = d_j . r0 with r0 = b - A x0.

constexpr int PROJ_THREADS = 256;
constexpr int PROJ_GRID_F = 1024;  // workgroups of dsv_solve_warm's norm and update kernels, and
                                  // the most its dot kernels (whose epilogue writes up to 44
                                  // partials per workgroup and species) take
constexpr int PROJ_GRID = 256;  // dsv_warm_start's S * 9 * PROJ_GRID partials (3 directions at
                                // most) fit the residual partial buffer
static_assert(MAXS * 9 * PROJ_GRID <= 2 * MAXS * MAX_PROD, "projection partials");

template <int S>
struct ProjVec {
    const SV<double, S>* d[MAXD];
    const SV<double, S>* Ad[MAXD];
    const SV<double, S>* r0;
    const SV<double, S>* x0;
    SV<double, S>* x;
    double c[MAXD][MAXS];
    const double* cdev;  // when set: the coefficients c[k][s] at cdev[k * MAXS + s] (device)
};

template <int S, int ND>
__device__ __forceinline__ void proj_coefs(const ProjVec<S>& pv, double (&c)[ND][S]) {
#pragma unroll
    for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int s = 0; s < S; ++s) c[k][s] = pv.cdev ? pv.cdev[k * MAXS + s] : pv.c[k][s];
}

// Gram quantities of a species, in this order: d_j . A d_k (j <= k, row by row), then
// d_j . r0 (NP of them)
template <int ND>
constexpr int gram_row(int j) {  // index of d_j . A d_j
    return j * ND - j * (j - 1) / 2;
}
template <int ND>
constexpr int gram_rows(int j0, int j1) {  // quantities of rows [j0, j1): their G and g entries
    return gram_row<ND>(j1) - gram_row<ND>(j0) + (j1 - j0);
}

// Partials of the Gram rows [J0, J1) of species blockIdx.y: d_j . A d_k (k >= j) and
// d_j . r0, at the species' quantity index (s * NP + q) * gridDim.x + blockIdx.x.  One species
// per workgroup row keeps the accumulators of eight directions in registers (the three
// species of a vertex column run on the same XCD, so its records are read from HBM once), and
// every quantity's sum has the same order whichever rows a launch covers.
template <int S, int ND, int J0, int J1>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_dots(int V, ProjVec<S> pv,
                                                           double* __restrict__ pout) {
    constexpr int NP = ND * (ND + 1) / 2 + ND;
    constexpr int NR = gram_rows<ND>(J0, J1);
    constexpr int NG = gram_row<ND>(J1) - gram_row<ND>(J0);  // of them d_j . A d_k
    __shared__ double red[NR * 16];
    const int s = blockIdx.y;
    double tot[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) tot[q] = 0.0;
    // d_j . A d_k = A d_j . d_k on the free rows (the directions vanish at Dirichlet rows), so
    // rows [J0, J1) read their own A d_j and the directions, not every A d_k.  Entry (j, k >= j)
    // is always summed as A d_j . d_k: the same bits as the shifted rows of k_proj_dots_shift.
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        double Ad[J1 - J0];
#pragma unroll
        for (int j = J0; j < J1; ++j) Ad[j - J0] = pv.Ad[j][i].v[s];
        const double r = pv.r0[i].v[s];
#pragma unroll
        for (int k = J0; k < ND; ++k) {
            const double d = pv.d[k][i].v[s];
#pragma unroll
            for (int j = J0; j < J1; ++j)
                if (j <= k) tot[gram_row<ND>(j) - gram_row<ND>(J0) + (k - j)] += Ad[j - J0] * d;
            if (k < J1) tot[NG + (k - J0)] += d * r;
        }
    }
    block_sum_all<NR>(tot, red);
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int q = 0; q < NG; ++q)
        pout[(s * NP + gram_row<ND>(J0) + q) * gridDim.x + blockIdx.x] = tot[q];
#pragma unroll
    for (int j = J0; j < J1; ++j)
        pout[(s * NP + ND * (ND + 1) / 2 + j) * gridDim.x + blockIdx.x] = tot[NG + (j - J0)];
}

// Gram quantities when directions 2.. are the previous call's 1.. (shifted history): only
// rows 0 and 1 of d_j . A d_k and every d_j . r0 are new, 3 ND - 1 per species, in the order
// d_0 . A d_k (k < ND), d_1 . A d_k (1 <= k < ND), d_j . r0 per species.  The rows
// are summed as A d_0 . d_k and A d_1 . d_k (ND + 3 vectors read), each sum in the same
// per-thread and per-workgroup order as in k_proj_dots, so the values are the same bits.
template <int S, int ND>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_dots_shift(int V, ProjVec<S> pv,
                                                                 double* __restrict__ pout) {
    // every species of a vertex in one thread, so that each record is read once (a workgroup
    // row per species read each 8-byte value of a record on its own, 2.5x the bytes from HBM
    // once the vectors outgrow the caches); per species the sums keep k_proj_dots's order
    constexpr int NQ = 3 * ND - 1;
    __shared__ double red[S * NQ * 16];
    double tot[S * NQ];
#pragma unroll
    for (int q = 0; q < S * NQ; ++q) tot[q] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        const SV<double, S> r = pv.r0[i], Ad0 = pv.Ad[0][i], Ad1 = pv.Ad[1][i];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const SV<double, S> d = pv.d[k][i];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                tot[s * NQ + k] += Ad0.v[s] * d.v[s];
                if (k >= 1) tot[s * NQ + ND + k - 1] += Ad1.v[s] * d.v[s];
                tot[s * NQ + 2 * ND - 1 + k] += d.v[s] * r.v[s];
            }
        }
    }
    // block sums as in block_sum_all (wave, then the waves in order), one value at a time so
    // that the epilogue does not hold every wave sum at once; thread q totals quantity q
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < S * NQ; ++q) {
        const double v = wave_sum(tot[q]);
        if (lane == 0) red[q * 16 + wid] = v;
    }
    __syncthreads();
    static_assert(S * NQ <= PROJ_THREADS, "one thread per quantity");
    const int q = threadIdx.x;
    if (q >= S * NQ) return;
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[q * 16 + w];
    pout[q * gridDim.x + blockIdx.x] = t;
}

// the Gram rows of ND directions (grid x S workgroups): one launch up to six directions, rows
// [0, 4) and [4, ND) beyond, so that the accumulators stay in registers
template <int S, int ND>
static void launch_proj_dots(const dsv_mesh* m, const ProjVec<S>& pv, double* pout, int grid,
                             hipStream_t st) {
    if constexpr (ND <= 6) {
        hipLaunchKernelGGL((k_proj_dots<S, ND, 0, ND>), dim3(grid, S), dim3(PROJ_THREADS), 0, st,
                           m->V, pv, pout);
    } else {
        hipLaunchKernelGGL((k_proj_dots<S, ND, 0, 4>), dim3(grid, S), dim3(PROJ_THREADS), 0, st,
                           m->V, pv, pout);
        hipLaunchKernelGGL((k_proj_dots<S, ND, 4, ND>), dim3(grid, S), dim3(PROJ_THREADS), 0, st,
                           m->V, pv, pout);
    }
}

// x = x0 + sum_k c_k d_k, and partials of ||r0 - sum_k c_k A d_k||^2 per species
template <int S, int ND>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_combine(int V, ProjVec<S> pv,
                                                              double* __restrict__ pout) {
    __shared__ double red[S * 16];
    double tot[S];
#pragma unroll
    for (int s = 0; s < S; ++s) tot[s] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        SV<double, S> x = pv.x0[i], r = pv.r0[i];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const SV<double, S> d = pv.d[k][i], Ad = pv.Ad[k][i];
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (pv.c[k][s] != 0.0) {  // a dropped direction may hold NaN: 0 * NaN = NaN
                    x.v[s] += pv.c[k][s] * d.v[s];
                    r.v[s] -= pv.c[k][s] * Ad.v[s];
                }
        }
        pv.x[i] = x;
#pragma unroll
        for (int s = 0; s < S; ++s) tot[s] += r.v[s] * r.v[s];
    }
    write_partials<S>(tot, red, pout);
}

// ||r0 - sum_k c_k A d_k||^2 per species, as partials (nothing written)
template <int S, int ND>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_norms(int V, ProjVec<S> pv,
                                                            double* __restrict__ pout) {
    __shared__ double red[S * 16];
    double c[ND][S];
    proj_coefs<S, ND>(pv, c);
    double tot[S];
#pragma unroll
    for (int s = 0; s < S; ++s) tot[s] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        SV<double, S> r = pv.r0[i];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const SV<double, S> Ad = pv.Ad[k][i];
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (c[k][s] != 0.0) r.v[s] -= c[k][s] * Ad.v[s];  // skip dropped directions
        }
#pragma unroll
        for (int s = 0; s < S; ++s) tot[s] += r.v[s] * r.v[s];
    }
    write_partials<S>(tot, red, pout);
}

// d = a - b per species (caller order -> internal order; NULL b: d = a), 0 where bit s of
// bcbits (internal order) marks a Dirichlet vertex of species s
template <int S>
__global__ void k_pack_diff(int V, const int32_t* __restrict__ perm, const double* a0,
                            const double* a1, const double* a2, const double* b0, const double* b1,
                            const double* b2, const uint8_t* __restrict__ bcbits,
                            SV<double, S>* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const int c = perm ? perm[i] : i;
    const unsigned bits = bcbits[i];
    const double* A[3] = {a0, a1, a2};
    const double* B[3] = {b0, b1, b2};
    SV<double, S> r{};
#pragma unroll
    for (int s = 0; s < S; ++s)
        r.v[s] = ((bits >> s) & 1u) ? 0.0 : (A[s] ? A[s][c] : 0.0) - (B[s] ? B[s][c] : 0.0);
    out[i] = r;
}

// Solves G c = g for the N x N SPD Gram matrix G (row-major) by LDL^T, dropping a direction
// whose pivot falls below rel_tol of its own energy d.Ad (numerically dependent on the
// previous ones, or zero); a dropped direction gets c = 0.
template <int N>
__host__ __device__ static void solve_gram(const double* G, const double* g, double* c,
                                           double rel_tol) {
    double L[N][N] = {}, Dg[N] = {}, y[N] = {};
    bool keep[N] = {};
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double dj = G[j * N + j];
#pragma unroll
        for (int k = 0; k < j; ++k)
            if (keep[k]) dj -= L[j][k] * L[j][k] * Dg[k];
        const bool finite = dj - dj == 0.0;  // false for an infinity or a NaN
        keep[j] = G[j * N + j] > 0.0 && finite && dj > rel_tol * G[j * N + j];
        Dg[j] = keep[j] ? dj : 1.0;
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            double v = G[i * N + j];
#pragma unroll
            for (int k = 0; k < j; ++k)
                if (keep[k]) v -= L[i][k] * L[j][k] * Dg[k];
            L[i][j] = keep[j] ? v / dj : 0.0;
        }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double v = g[j];
#pragma unroll
        for (int k = 0; k < j; ++k)
            if (keep[k]) v -= L[j][k] * y[k];
        y[j] = keep[j] ? v : 0.0;
    }
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        double v = y[j] / Dg[j];
#pragma unroll
        for (int i = j + 1; i < N; ++i)
            if (keep[i]) v -= L[i][j] * c[i];
        c[j] = keep[j] ? v : 0.0;
    }
}

// Totals of partials, one workgroup per quantity: out[q] = sum over b < nprod of
// partials[q * nprod + b] (fixed order), so that a single-workgroup consumer reads totals,
// not every producer's partial.
__global__ __launch_bounds__(PROJ_THREADS) void k_sum_partials(const double* __restrict__ partials,
                                                              int nprod, double* __restrict__ out) {
    __shared__ double red[16];
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nprod; b += blockDim.x) v[0] += partials[(size_t)blockIdx.x * nprod + b];
    block_sum_all<1>(v, red);
    if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// k_sum_partials over two sources in one launch: quantities [0, q1) from p1 (n1 partials
// each), the following ones from p2 (n2 each), in the same order per quantity
__global__ __launch_bounds__(PROJ_THREADS) void k_sum_partials2(const double* __restrict__ p1,
                                                               int n1, int q1,
                                                               const double* __restrict__ p2,
                                                               int n2, double* __restrict__ out) {
    __shared__ double red[16];
    const bool first = (int)blockIdx.x < q1;
    const int nprod = first ? n1 : n2;
    const double* p = first ? p1 + (size_t)blockIdx.x * n1 : p2 + (size_t)(blockIdx.x - q1) * n2;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nprod; b += blockDim.x) v[0] += p[b];
    block_sum_all<1>(v, red);
    if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// Device side of the warm start folded into a solve: the Gram systems, then the acceptance
// test, with no host round trip in between.
struct WarmOut {
    double c[MAXD][MAXS];  // coefficients of the Gram solves
    double rr0[MAXS];   // ||r||^2 of the given start and ||b||^2 (the solve's first residual)
    double bb[MAXS];
    double rr1[MAXS];   // ||r||^2 of the start taken
    int take[MAXS];     // 1 where the combination was taken (else the start stays, c counts 0)
    int seq;            // the pinned copy only: set last by the acceptance kernel
};

// One Gram solve per species (thread s), from the totals of the dot partials: the full
// quantities (k_proj_dots), or with SHIFT the new rows of k_proj_dots_shift and, for
// j, k >= 2, the previous call's d_(j-1) . A d_(k-1) from gprev.  The Gram matrix is saved to
// gout ([species][MAXD][MAXD]) for the next call.
template <int S, int ND, bool SHIFT>
__global__ __launch_bounds__(64) void k_proj_solve(const double* __restrict__ tot,
                                                  const double* __restrict__ gprev,
                                                  double* __restrict__ gout,
                                                  WarmOut* __restrict__ out) {
    constexpr int NP = ND * (ND + 1) / 2 + ND;
    constexpr int NQ = 3 * ND - 1;
    if (threadIdx.x >= S) return;
    const int s = threadIdx.x;
    double G[ND * ND], g[ND], c[ND];
    if constexpr (SHIFT) {
        const double* q = tot + s * NQ;
#pragma unroll
        for (int k = 0; k < ND; ++k) G[k] = G[k * ND] = q[k];
#pragma unroll
        for (int k = 1; k < ND; ++k) G[ND + k] = G[k * ND + 1] = q[ND + k - 1];
#pragma unroll
        for (int j = 2; j < ND; ++j)
#pragma unroll
            for (int k = 2; k < ND; ++k) G[j * ND + k] = gprev[(s * MAXD + j - 1) * MAXD + k - 1];
#pragma unroll
        for (int j = 0; j < ND; ++j) g[j] = q[2 * ND - 1 + j];
    } else {
        const double* q = tot + s * NP;
        int t = 0;
#pragma unroll
        for (int j = 0; j < ND; ++j)
#pragma unroll
            for (int k = j; k < ND; ++k) G[j * ND + k] = G[k * ND + j] = q[t++];
#pragma unroll
        for (int j = 0; j < ND; ++j) g[j] = q[t++];
    }
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int k = 0; k < ND; ++k) gout[(s * MAXD + j) * MAXD + k] = G[j * ND + k];
    solve_gram<ND>(G, g, c, 1e-10);
#pragma unroll
    for (int k = 0; k < ND; ++k) out->c[k][s] = c[k];
}

// Takes the combination of a species only where some direction moved it and its FP64
// residual fell (the given start is the fallback), from the totals of the norm partials and
// of the residual kernel's (r.r then b.b per species), then updates in place:
// x += sum_k c_k d_k, r -= sum_k c_k A d_k for the species taken.  Every workgroup makes the
// same decision from the same totals; workgroup 0 records it.
// With `pub` (pinned host memory), workgroup 0 also publishes the outcome there, the
// coefficients zero for the species not taken, and sets its `seq` last.
template <int S, int ND>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_accept_update(
    int V, ProjVec<S> pv, SV<double, S>* __restrict__ r, const double* __restrict__ tot,
    WarmOut* __restrict__ out, WarmOut* __restrict__ pub) {
    double c[ND][S];
    proj_coefs<S, ND>(pv, c);
    const double* n1 = tot;      // ||r0 - sum_k c_k A d_k||^2 per species
    const double* r0 = tot + S;  // the residual kernel's r.r, then b.b, per species
    const bool rec = blockIdx.x == 0 && threadIdx.x == 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        bool moved = false;  // every direction dropped: nothing to take
#pragma unroll
        for (int k = 0; k < ND; ++k) moved |= c[k][s] != 0.0;
        const double a = sqrt(r0[s]), b = sqrt(n1[s]);
        const bool take = moved && b - b == 0.0 && b < a;
        if (!take)
#pragma unroll
            for (int k = 0; k < ND; ++k) c[k][s] = 0.0;
        if (rec) {
            out->take[s] = take ? 1 : 0;
            out->rr0[s] = r0[s];
            out->bb[s] = r0[S + s];
            out->rr1[s] = take ? n1[s] : r0[s];
            if (pub) {
                pub->take[s] = take ? 1 : 0;
                pub->rr0[s] = r0[s];
                pub->bb[s] = r0[S + s];
                pub->rr1[s] = take ? n1[s] : r0[s];
#pragma unroll
                for (int k = 0; k < ND; ++k) pub->c[k][s] = c[k][s];
            }
        }
    }
    if (rec && pub) {
        __threadfence_system();
        *reinterpret_cast<volatile int*>(&pub->seq) = 1;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        SV<double, S> x = pv.x[i], rv = r[i];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const SV<double, S> d = pv.d[k][i], Ad = pv.Ad[k][i];
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (c[k][s] != 0.0) {  // dropped or rejected: keep x and r exactly (NaN-safe)
                    x.v[s] += c[k][s] * d.v[s];
                    rv.v[s] -= c[k][s] * Ad.v[s];
                }
        }
        pv.x[i] = x;
        r[i] = rv;
    }
}

template <int S, int ND>
static void warm_start_impl(dsv_mesh* m, const dsv_species* sp, double inv_dt, int lumped,
                            const double* const* dirs, double* coef, double* r_before,
                            double* r_after) {
    constexpr int NP = ND * (ND + 1) / 2 + ND;
    double rr[MAXS] = {}, bb[MAXS] = {};
    residual_norms_impl<S>(m, sp, inv_dt, lumped, rr, bb);  // x0 -> X64, r0 -> R64
    Workspace* w = get_ws(m, DSV_METHOD_MF64, S);
    hipStream_t st = w->stream;
    const unsigned gv = cdiv(m->V, 256);
    ProjVec<S> pv{};
    for (int k = 0; k < ND; ++k) {
        auto* d = (SV<double, S>*)w->vb[VB_V1 + 2 * k];
        auto* Ad = (SV<double, S>*)w->vb[VB_V2 + 2 * k];
        const double* const* dk = dirs + (size_t)k * S;
        hipLaunchKernelGGL((k_pack<double, S>), dim3(gv), dim3(256), 0, st, m->V, m->perm, dk[0],
                           S > 1 ? dk[1] : nullptr, S > 2 ? dk[2] : nullptr, d);
        for (int s = 0; s < S; ++s)
            if (sp[s].n_bc > 0)
                hipLaunchKernelGGL((k_set_bc_values<S>), dim3(cdiv(sp[s].n_bc, 128)), dim3(128), 0,
                                   st, sp[s].n_bc, sp[s].bc_dofs, m->iperm, s, 0.0, d);
        MFVec<double, S> v{};
        v.x = d;
        v.y = Ad;
        launch_mf<double, S, MF_APPLY>(m, m->part, w->prm, v, nullptr, st);
        pv.d[k] = d;
        pv.Ad[k] = Ad;
    }
    pv.r0 = (const SV<double, S>*)w->vb[VB_R64];
    pv.x0 = (const SV<double, S>*)w->vb[VB_X64];
    pv.x = (SV<double, S>*)w->vb[VB_X];
    launch_proj_dots<S, ND>(m, pv, w->pR, PROJ_GRID, st);
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipMemcpyAsync(w->h_pR, w->pR, sizeof(double) * S * NP * PROJ_GRID,
                           hipMemcpyDeviceToHost, st));
    DSV_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s) {
        double q[NP];
        for (int j = 0; j < NP; ++j) {
            double a = 0.0;
            for (int b = 0; b < PROJ_GRID; ++b) a += w->h_pR[(size_t)(s * NP + j) * PROJ_GRID + b];
            q[j] = a;
        }
        double G[ND * ND], g[ND], c[ND];
        int t = 0;
        for (int j = 0; j < ND; ++j)
            for (int k = j; k < ND; ++k) G[j * ND + k] = G[k * ND + j] = q[t++];
        for (int j = 0; j < ND; ++j) g[j] = q[t++];
        solve_gram<ND>(G, g, c, 1e-10);
        for (int k = 0; k < ND; ++k) pv.c[k][s] = c[k];
    }
    hipLaunchKernelGGL((k_proj_combine<S, ND>), dim3(PROJ_GRID), dim3(PROJ_THREADS), 0, st, m->V,
                       pv, w->pR);
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipMemcpyAsync(w->h_pR, w->pR, sizeof(double) * S * PROJ_GRID, hipMemcpyDeviceToHost,
                           st));
    DSV_HIP(hipStreamSynchronize(st));
    double* out[3] = {nullptr, nullptr, nullptr};
    for (int s = 0; s < S; ++s) {
        double a = 0.0;
        for (int b = 0; b < PROJ_GRID; ++b) a += w->h_pR[(size_t)s * PROJ_GRID + b];
        const double r0 = std::sqrt(rr[s]), r1 = std::sqrt(a);
        bool moved = false;  // every direction dropped: nothing to take
        for (int k = 0; k < ND; ++k) moved |= pv.c[k][s] != 0.0;
        const bool take = moved && std::isfinite(r1) && r1 < r0;  // the start is the fallback
        if (take) out[s] = sp[s].x;
        if (r_before) r_before[s] = r0;
        if (r_after) r_after[s] = take ? r1 : r0;
        if (coef)
            for (int k = 0; k < ND; ++k) coef[k * S + s] = take ? pv.c[k][s] : 0.0;
    }
    if (out[0] || out[1] || out[2]) {
        note_outputs(m, nullptr, out, false);
        hipLaunchKernelGGL((k_unpack<double, S>), dim3(gv), dim3(256), 0, st, m->V, m->perm,
                           (const SV<double, S>*)pv.x, out[0], out[1], out[2]);
        DSV_HIP(hipGetLastError());
        DSV_HIP(hipStreamSynchronize(st));
    }
}

// Galerkin warm start of dsv_solve_warm after the residual kernel of the solve's first round
// (its partials in w->pR, nres per quantity): rr and bb come back as that residual's norms,
// rr updated to the start taken; one host round trip.
// ---------------------------------------------------------------------------------------
// Low-mode subspace of the MF32_IR64 operator (dsv_build_low_modes).  After the warm start's
// history directions, the error of a start lies mostly in the smooth modes that Jacobi-PCG
// resolves last.  A subspace W approximating the eigenvectors of the lowest eigenvalues of
// D^-1 A (D = diag A; vectors 0 at Dirichlet vertices) takes them out once per solve by the
// Galerkin correction x += W c, (W^T A W) c = W^T r, which minimises the A-norm of the start's
// error over start + span W.  W is built once per operator by Chebyshev-filtered subspace
// iteration: a seeded random block; passes of a degree-m Chebyshev polynomial in P = D^-1 A
// that damps [a, b] (b a Gershgorin bound of P, a the block's largest Ritz value of the pass
// before, b / 64 at the first pass) and so raises the low end of the spectrum against it;
// Rayleigh-Ritz in the D inner product after each pass.  The solves read W in FP16, each vector
// scaled by a power of two to a largest entry in [1/2, 1); W^T A W is formed from the rounded
// vectors, so the correction is the exact Galerkin step of the subspace they span.
constexpr int LM_MAX = 64;        // modes per species at most (a multiple of 4)
constexpr int LM_GUARD = 16;      // block vectors beyond the modes kept
constexpr int LM_BLOCK_MAX = LM_MAX + LM_GUARD;
constexpr int LM_THREADS = 256;
constexpr int LM_CHUNKS = 240;    // vertex chunks (workgroups along x) of the build's dot kernels
constexpr int LM_DOT_CHUNKS = 1024;  // workgroups of the correction's dots at most
// The FP16 subspace is stored in groups of LM_GROUP vertices: all modes and species of a group
// are contiguous (vector j, species s of the group in LM_GROUP consecutive halves), so that
// the workgroup that takes a group streams one contiguous block.
constexpr int LM_GROUP = 256;
__host__ __device__ __forceinline__ size_t lm_index(int n, int S, int j, int s, int v) {
    return (((size_t)(v / LM_GROUP) * n + j) * S + s) * LM_GROUP + (v % LM_GROUP);
}
// workgroups of the correction's dot kernel: one per vertex group, up to LM_DOT_CHUNKS
static int lm_dot_chunks(int Vp) { return std::max(1, std::min(LM_DOT_CHUNKS, Vp / LM_GROUP)); }
typedef _Float16 lm_half4 __attribute__((ext_vector_type(4)));

// uniform in (-1, 1) from a 64-bit key: the splitmix64 finaliser (Steele, Lea and Flood 2014)
__device__ __forceinline__ float lm_uniform(uint64_t key) {
    uint64_t z = key + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)((double)(z >> 40) * (2.0 / 16777216.0) - 1.0);
}

template <int S>
__global__ void k_lm_random(int V, int nb, uint64_t seed, const uint8_t* __restrict__ bcbits,
                            SV<float, S>* __restrict__ X) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)V * nb) return;
    const int v = (int)(t % V);
    const unsigned bcb = bcbits ? (unsigned)bcbits[v] : 0u;
    SV<float, S> x;
#pragma unroll
    for (int s = 0; s < S; ++s)
        x.v[s] = ((bcb >> s) & 1u) ? 0.0f
                                   : lm_uniform(seed * 0x100000001B3ull + (uint64_t)t * S + s);
    X[t] = x;
}

// part[((i nb + j) S + s) gridDim.x + blockIdx.x]: the chunk's sum of X_i Y_j per species
// (divided by dinv with `dinv`: the D inner product) over the 4 x 4 tile blockIdx.y of (i, j)
template <int S, class TY>
__global__ __launch_bounds__(LM_THREADS) void k_lm_block_dots(int V, int nb,
                                                              const SV<float, S>* __restrict__ X,
                                                              const SV<TY, S>* __restrict__ Y,
                                                              const SV<float, S>* __restrict__ dinv,
                                                              double* __restrict__ part) {
    const int nt = nb / 4;
    const int i0 = (blockIdx.y / nt) * 4, j0 = (blockIdx.y % nt) * 4;
    double acc[4][4][S];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int s = 0; s < S; ++s) acc[a][b][s] = 0.0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        SV<float, S> xi[4];
        SV<TY, S> yj[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xi[q] = X[(size_t)(i0 + q) * V + v];
            yj[q] = Y[(size_t)(j0 + q) * V + v];
        }
        double wgt[S];
        if (dinv) {
            const SV<float, S> d = dinv[v];
#pragma unroll
            for (int s = 0; s < S; ++s) wgt[s] = 1.0 / (double)d.v[s];
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) wgt[s] = 1.0;
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const double xa = (double)xi[a].v[s] * wgt[s];
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b][s] += xa * (double)yj[b].v[s];
            }
    }
    __shared__ double red[16 * MAXS][LM_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 16 * S; ++q) {
        const double t = wave_sum(acc[q / (4 * S)][(q / S) % 4][q % S]);
        if (lane == 0) red[q][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < 16 * S) {
        const int q = threadIdx.x;
        double t = 0.0;
        for (int w = 0; w < LM_THREADS / 64; ++w) t += red[q][w];
        const int a = q / (4 * S), b = (q / S) % 4, sp = q % S;
        part[(((size_t)(i0 + a) * nb + (j0 + b)) * S + sp) * gridDim.x + blockIdx.x] = t;
    }
}

// out[o] = sum over the chunks of part[o nchunk + i], in chunk order
__global__ void k_lm_sum(int n_out, int nchunk, const double* __restrict__ part,
                         double* __restrict__ out) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_out) return;
    double t = 0.0;
    for (int i = 0; i < nchunk; ++i) t += part[(size_t)o * nchunk + i];
    out[o] = t;
}

// Xout_j = sum_i C_s[i][j] X_i per species (C: [S][nb][nb] FP32), j < nout
template <int S>
__global__ __launch_bounds__(LM_THREADS) void k_lm_rotate(int V, int nb, int nout,
                                                          const SV<float, S>* __restrict__ X,
                                                          const float* __restrict__ C,
                                                          SV<float, S>* __restrict__ Xout) {
    __shared__ float cs[LM_BLOCK_MAX * LM_BLOCK_MAX];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    for (int s = 0; s < S; ++s) {
        __syncthreads();
        for (int i = threadIdx.x; i < nb * nout; i += blockDim.x)
            cs[i] = C[((size_t)s * nb + i / nout) * nb + i % nout];
        __syncthreads();
        if (v < V) {
            float xin[LM_BLOCK_MAX];
#pragma unroll
            for (int i = 0; i < LM_BLOCK_MAX; ++i) xin[i] = i < nb ? X[(size_t)i * V + v].v[s] : 0.0f;
            for (int j = 0; j < nout; ++j) {
                float o = 0.0f;
#pragma unroll
                for (int i = 0; i < LM_BLOCK_MAX; ++i)
                    if (i < nb) o += xin[i] * cs[i * nout + j];
                Xout[(size_t)j * V + v].v[s] = o;
            }
        }
    }
}

template <int S>
__global__ void k_lm_to_double(int V, const SV<float, S>* __restrict__ x,
                               SV<double, S>* __restrict__ y) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const SV<float, S> a = x[v];
    SV<double, S> b;
#pragma unroll
    for (int s = 0; s < S; ++s) b.v[s] = (double)a.v[s];
    y[v] = b;
}

// mx[j S + s] = max_v |X_j(v)_s| over the vectors j of blockIdx.x
template <int S>
__global__ __launch_bounds__(256) void k_lm_colmax(int V, const SV<float, S>* __restrict__ X,
                                                  float* __restrict__ mx) {
    const int j = blockIdx.x;
    float a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) a[s] = 0.0f;
    for (int v = threadIdx.x; v < V; v += blockDim.x) {
        const SV<float, S> t = X[(size_t)j * V + v];
#pragma unroll
        for (int s = 0; s < S; ++s) a[s] = fmaxf(a[s], fabsf(t.v[s]));
    }
    __shared__ float red[MAXS][256];
#pragma unroll
    for (int s = 0; s < S; ++s) red[s][threadIdx.x] = a[s];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
#pragma unroll
            for (int s = 0; s < S; ++s)
                red[s][threadIdx.x] = fmaxf(red[s][threadIdx.x], red[s][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < S) mx[j * S + threadIdx.x] = red[threadIdx.x][0];
}

// The FP16 subspace: W[lm_index(j, s, v)] = fp16(X_j(v)_s sc[j S + s]) for v < V, 0 in the
// padding up to Vp; X_j(v) is replaced by the same rounded values in FP32 (the vectors the
// Gram matrix W^T A W is formed from).  One thread per (j, v < Vp).
template <int S>
__global__ void k_lm_to_half(int V, int Vp, int n, const float* __restrict__ sc,
                             SV<float, S>* __restrict__ X, _Float16* __restrict__ W) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * Vp) return;
    const int j = (int)(t / Vp), v = (int)(t % Vp);
    if (v >= V) {
#pragma unroll
        for (int s = 0; s < S; ++s) W[lm_index(n, S, j, s, v)] = (_Float16)0.0f;
        return;
    }
    SV<float, S> x = X[(size_t)j * V + v];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const _Float16 h = (_Float16)(x.v[s] * sc[j * S + s]);
        W[lm_index(n, S, j, s, v)] = h;
        x.v[s] = (float)h;
    }
    X[(size_t)j * V + v] = x;
}

// per-block maximum over the rows of sum_j |a_ij| / a_ii (the Gershgorin bound of D^-1 A)
__global__ __launch_bounds__(256) void k_lm_gershgorin(int V, const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ diag_pos,
                                                      const double* __restrict__ val,
                                                      double* __restrict__ part) {
    double mx = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < V; r += gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int32_t p = indptr[r]; p < indptr[r + 1]; ++p) t += fabs(val[p]);
        mx = fmax(mx, t / fabs(val[diag_pos[r]]));
    }
    __shared__ double red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The per-solve kernels of the correction: a workgroup of LM_WAVES wavefronts takes one vertex
// group (64 quads) per pass, wavefront w the modes [LM_JW w, LM_JW (w + 1)), a lane one quad
// (one 8-byte FP16 load per mode and species), so that each vertex's r (or the quad's update)
// is handled by one workgroup for every mode.
constexpr int LM_WAVES = 4;
constexpr int LM_JW = LM_MAX / LM_WAVES;
constexpr int LM_DOT_WAVES = 8;  // the dot kernel: more, narrower wavefronts (occupancy)
constexpr int LM_DOT_JW = LM_MAX / LM_DOT_WAVES;

// the correction's dots: part[(j S + s) gridDim.x + blockIdx.x] = the workgroup's sum of
// w_j(v) r(v) over its quads
template <int S>
__global__ __launch_bounds__(64 * LM_DOT_WAVES) void k_lm_dots(int V, int Vp, int n,
                                                          const _Float16* __restrict__ W,
                                                          const SV<double, S>* __restrict__ r,
                                                          double* __restrict__ part) {
    const int lane = threadIdx.x & 63, j0 = (threadIdx.x >> 6) * LM_DOT_JW;
    if (j0 >= n) return;  // wave-uniform (no barrier follows)
    double acc[LM_DOT_JW][S];
#pragma unroll
    for (int j = 0; j < LM_DOT_JW; ++j)
#pragma unroll
        for (int s = 0; s < S; ++s) acc[j][s] = 0.0;
    static_assert(LM_GROUP == 4 * 64, "a lane takes one quad of the group");
    for (int g = blockIdx.x; g < Vp / LM_GROUP; g += gridDim.x) {
        const int q = g * 64 + lane;
        double rv[4][S];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int v = 4 * q + k;
            if (v < V) {
                const SV<double, S> t = r[v];
#pragma unroll
                for (int s = 0; s < S; ++s) rv[k][s] = t.v[s];
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) rv[k][s] = 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < LM_DOT_JW; ++j)
            if (j0 + j < n) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const lm_half4 h =
                        *reinterpret_cast<const lm_half4*>(W + lm_index(n, S, j0 + j, s, 4 * q));
                    acc[j][s] += (double)h[0] * rv[0][s] + (double)h[1] * rv[1][s] +
                                 (double)h[2] * rv[2][s] + (double)h[3] * rv[3][s];
                }
            }
    }
#pragma unroll
    for (int i = 0; i < LM_DOT_JW * S; ++i) {
        const double t = wave_sum(acc[i / S][i % S]);
        if (lane == 0 && j0 + i / S < n) part[((size_t)j0 * S + i) * gridDim.x + blockIdx.x] = t;
    }
}

// t[o] = the sum of the dot partials of coefficient o = j S + s, one wave per coefficient
__global__ __launch_bounds__(64) void k_lm_tot(int nchunk, const double* __restrict__ part,
                                               double* __restrict__ t) {
    const int o = blockIdx.x, lane = threadIdx.x;
    double a = 0.0;
    for (int i = lane; i < nchunk; i += 64) a += part[(size_t)o * nchunk + i];
    a = wave_sum(a);
    if (lane == 0) t[o] = a;
}

// c[o] = (Ginv_s t_s)_j for o = j S + s, one wave per coefficient (lane i takes term i)
template <int S>
__global__ __launch_bounds__(64) void k_lm_solve(int n, const double* __restrict__ ginv,
                                                 const double* __restrict__ t,
                                                 double* __restrict__ c) {
    const int o = blockIdx.x, lane = threadIdx.x;
    const int j = o / S, s = o % S;
    double a = lane < n ? ginv[((size_t)s * n + j) * n + lane] * t[lane * S + s] : 0.0;
    a = wave_sum(a);
    if (lane == 0) c[o] = a;
}

// y = W c (FP64), x += y.  Each wavefront sums its modes' terms for the workgroup's 64 quads;
// wavefront k then adds the LM_WAVES parts (in wavefront order) for vertex k of every quad.
// A species whose start already meets the tolerance (||r||^2 <= rtol^2 ||b||^2 from the warm
// start's outputs) or whose coefficients are not finite takes c = 0: every workgroup decides
// it from the same values.
template <int S>
__global__ __launch_bounds__(64 * LM_WAVES) void k_lm_update(int V, int Vp, int n,
                                                            const _Float16* __restrict__ W,
                                                            const double* __restrict__ c,
                                                            const WarmOut* __restrict__ wout,
                                                            double rtol2,
                                                            SV<double, S>* __restrict__ x,
                                                            SV<double, S>* __restrict__ y) {
    static_assert(LM_WAVES == 4, "wavefront k finishes vertex k of each quad");
    __shared__ double cs[LM_MAX * MAXS];
    __shared__ double red[LM_WAVES][4 * MAXS][64];
    __shared__ int ok[MAXS];
    if (threadIdx.x < MAXS) ok[threadIdx.x] = 1;
    __syncthreads();
    for (int i = threadIdx.x; i < n * S; i += blockDim.x) {
        const double v = c[i];
        cs[i] = v;
        if (!isfinite(v)) ok[i % S] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j0 = wv * LM_JW;
    const int q = blockIdx.x * 64 + lane;
    double acc[4][S];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int s = 0; s < S; ++s) acc[k][s] = 0.0;
    if (q < (Vp >> 2)) {
#pragma unroll
        for (int j = 0; j < LM_JW; ++j)
            if (j0 + j < n) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const lm_half4 h =
                        *reinterpret_cast<const lm_half4*>(W + lm_index(n, S, j0 + j, s, 4 * q));
                    const double cj = cs[(j0 + j) * S + s];
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc[k][s] += cj * (double)h[k];
                }
            }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int s = 0; s < S; ++s) red[wv][k * S + s][lane] = acc[k][s];
    __syncthreads();
    const int v = 4 * q + wv;
    if (v >= V) return;
    SV<double, S> yv, xv = x[v];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < LM_WAVES; ++w) t += red[w][wv * S + s][lane];
        const bool run = ok[s] && wout->rr1[s] > rtol2 * wout->bb[s];
        yv.v[s] = run ? t : 0.0;
        xv.v[s] += yv.v[s];
    }
    y[v] = yv;
    x[v] = xv;
}

static bool low_modes_valid(const dsv_mesh* m, const Workspace* w) {
    return w->lm_n > 0 && m->op_reuse && w->method == DSV_METHOD_MF32_IR64 &&
           same_key(w->key[1], w->lm_key);
}

// the correction of x64 / r64 and the publication of the new residual norms (ResidPub)
template <int S>
static void launch_low_mode_correction(dsv_mesh* m, Workspace* w, double rtol, hipStream_t st) {
    const int n = w->lm_n;
    const _Float16* W = w->lm_w;
    const int Vp = w->lm_vp;
    auto* r64 = (SV<double, S>*)w->vb[VB_R64];
    auto* x64 = (SV<double, S>*)w->vb[VB_X64];
    const int nchunk = lm_dot_chunks(Vp);
    hipLaunchKernelGGL((k_lm_dots<S>), dim3(nchunk), dim3(64 * LM_DOT_WAVES), 0, st, m->V, Vp, n, W,
                       (const SV<double, S>*)r64, w->lm_part);
    double* tot = w->lm_c + LM_MAX * MAXS;
    hipLaunchKernelGGL(k_lm_tot, dim3(n * S), dim3(64), 0, st, nchunk, (const double*)w->lm_part,
                       tot);
    hipLaunchKernelGGL((k_lm_solve<S>), dim3(n * S), dim3(64), 0, st, n,
                       (const double*)w->lm_ginv, (const double*)tot, w->lm_c);
    hipLaunchKernelGGL((k_lm_update<S>), dim3(cdiv(Vp / 4, 64)), dim3(64 * LM_WAVES), 0, st, m->V,
                       Vp, n, W,
                       (const double*)w->lm_c, (const WarmOut*)w->wout, rtol * rtol, x64,
                       (SV<double, S>*)w->lm_y);
    // r64 -= A y (the residual kernel with b = r64 read and written in place by each vertex)
    MFVec<double, S> v{};
    v.x = (const SV<double, S>*)w->lm_y;
    v.y = r64;
    v.b = r64;
    launch_mf<double, S, MF_RESID>(m, m->part, w->prm2, v, w->pR, st);
    hipLaunchKernelGGL((k_resid_publish<S>), dim3(1), dim3(64 * 2 * S), 0, st, w->pR,
                       mf_grid(m->part), (const CGState*)nullptr, w->d_rpub);
    DSV_HIP(hipGetLastError());
}

template <int S, int ND>
static void fused_warm_nd(dsv_mesh* m, Workspace* w, const dsv_species* sp, const WarmArgs& wa,
                          int nres, double* rr, double* bb, hipStream_t st) {
    const unsigned gv = cdiv(m->V, 256);
    const unsigned char* prm64 = w->method == DSV_METHOD_MF32_IR64 ? w->prm2 : w->prm;
    if (!w->pW) w->pW = dalloc<double>((size_t)MAXS * (MAXD * (MAXD + 1) / 2 + MAXD) * PROJ_GRID_F);
    if (!w->wout) w->wout = dalloc<WarmOut>(1);
    if (!w->pT) w->pT = dalloc<double>((size_t)MAXS * (MAXD * (MAXD + 1) / 2 + MAXD));
    if (!w->h_wout) {
        DSV_HIP(hipHostMalloc((void**)&w->h_wout, sizeof(WarmOut), hipHostMallocCoherent));
        DSV_HIP(hipHostGetDevicePointer((void**)&w->d_wout_pub, w->h_wout, 0));
    }
    // Internal direction k: the explicit ones first (k < ne), new every call, then the history
    // increments h = k - ne, a[s] - b[s] per species: d_0 = x - u_{n-1}, d_h = u_{n-h} -
    // u_{n-h-1} (from_hist: d_h = hist_h - hist_(h+1), x = hist_0).  The caller's order (and
    // that of `coef`) is the history first, then the explicit ones.
    // Under operator reuse (whose contract keeps the arrays named unchanged), a history
    // direction h >= 1 whose pair of arrays is the previous call's internal pair k - 1 is that
    // call's direction, and its A d is taken from there -- unless one of the two arrays is a
    // solution the last solve wrote (as the previous start x is when the start is not taken
    // from the history).  With the same FP64 operator and a history shifted by one step, that
    // leaves one apply per call for the history (two with the start given in x), and with at
    // most one explicit direction the Gram rows of directions 2.. shift as well.
    const int ne = wa.n_extra;
    const double* a[MAXD][MAXS] = {};
    const double* b[MAXD][MAXS] = {};
    int api[MAXD];  // internal -> caller's index
    for (int k = 0; k < ND; ++k) {
        const int h = k - ne;
        api[k] = h < 0 ? wa.n_hist + k : h;
        for (int s = 0; s < S; ++s) {
            if (h < 0) {
                a[k][s] = wa.extra[(size_t)k * S + s];
            } else if (wa.from_hist) {
                a[k][s] = wa.hist[(size_t)h * S + s];
                b[k][s] = wa.hist[(size_t)(h + 1) * S + s];
            } else {
                a[k][s] = h == 0 ? sp[s].x : wa.hist[(size_t)(h - 1) * S + s];
                b[k][s] = wa.hist[(size_t)h * S + s];
            }
        }
    }
    const OpKey& opkey = w->key[w->method == DSV_METHOD_MF32_IR64 ? 1 : 0];
    const bool can = m->op_reuse && w->wvalid && same_key(opkey, w->wkey);
    auto written = [&](const double* p) {
        for (int s = 0; s < MAXS; ++s)
            if (p && p == w->wout_prev[s]) return true;
        return false;
    };
    int map[MAXD];  // direction -> buffer pair (wd[2p], wd[2p + 1])
    bool reuse[MAXD] = {}, busy[MAXD] = {};
    for (int k = 1; k < ND && can; ++k) {
        // explicit directions and the newest increment are never reused; the previous call's
        // direction k - 1 must be its history increment h - 1
        const int h = k - ne;
        if (h < 1 || ne != w->wnextra || h - 1 >= w->wnhist) continue;
        bool same = true;
        for (int s = 0; s < S; ++s)
            same = same && a[k][s] == w->wpair[k - 1][s][0] && b[k][s] == w->wpair[k - 1][s][1] &&
                   !written(a[k][s]) && !written(b[k][s]);
        if (!same) continue;
        reuse[k] = true;
        map[k] = w->wmap[k - 1];
        busy[map[k]] = true;
    }
    ProjVec<S> pv{};
    for (int k = 0; k < ND; ++k) {
        if (!reuse[k]) {
            int p = 0;
            while (busy[p]) ++p;
            busy[p] = true;
            map[k] = p;
        }
        const int p = map[k];
        for (int j = 0; j < 2; ++j)
            if (!w->wd[2 * p + j])
                DSV_HIP(hipMalloc(&w->wd[2 * p + j], (size_t)m->V * rec_bytes(sizeof(double), S)));
        auto* d = (SV<double, S>*)w->wd[2 * p];
        auto* Ad = (SV<double, S>*)w->wd[2 * p + 1];
        pv.d[k] = d;
        pv.Ad[k] = Ad;
        if (reuse[k]) continue;
        hipLaunchKernelGGL((k_pack_diff<S>), dim3(gv), dim3(256), 0, st, m->V, m->perm, a[k][0],
                           a[k][1], a[k][2], b[k][0], b[k][1], b[k][2], w->bcbits, d);
        MFVec<double, S> v{};
        v.x = d;
        v.y = Ad;
        launch_mf<double, S, MF_APPLY>(m, m->part, prm64, v, nullptr, st);
    }
    m->stat_dir_reused += [&] {
        int n = 0;
        for (int k = 0; k < ND; ++k) n += reuse[k];
        return n;
    }();
    w->wvalid = m->op_reuse;
    w->wkey = opkey;
    w->wnhist = wa.n_hist;
    w->wnextra = ne;
    for (int k = 0; k < ND; ++k) {
        w->wmap[k] = map[k];
        for (int s = 0; s < S; ++s) {
            w->wpair[k][s][0] = a[k][s];
            w->wpair[k][s][1] = b[k][s];
        }
    }
    pv.r0 = (const SV<double, S>*)w->vb[VB_R64];
    pv.x = (SV<double, S>*)w->vb[VB_X64];
    pv.cdev = &w->wout->c[0][0];
    constexpr int NP = ND * (ND + 1) / 2 + ND;
    // the previous call's Gram matrix serves for directions 2.. when they are its 1..
    for (int i = 0; i < 2; ++i)
        if (!w->gsave[i]) w->gsave[i] = dalloc<double>((size_t)MAXS * MAXD * MAXD);
    bool gshift = ND >= 3 && w->gnd >= ND - 1;
    for (int k = 2; k < ND; ++k) gshift = gshift && reuse[k];
    const double* gprev = w->gsave[w->gcur];
    double* gout = w->gsave[w->gcur ^ 1];
    // Gram workgroups: 256 (the epilogue dominates at a few hundred thousand vertices), up to
    // PROJ_GRID_F on larger meshes so that enough loads are in flight; the shifted and full
    // Gram kernels use the same grid, so that each sum runs in the same order either way
    const int pg = std::min(PROJ_GRID_F, std::max(PROJ_GRID, (int)((cdiv(m->V, 2048) + 7) / 8 * 8)));
    if constexpr (ND >= 3) {
        if (gshift) {
            hipLaunchKernelGGL((k_proj_dots_shift<S, ND>), dim3(pg), dim3(PROJ_THREADS), 0, st,
                               m->V, pv, w->pW);
            hipLaunchKernelGGL(k_sum_partials, dim3(S * (3 * ND - 1)), dim3(PROJ_THREADS), 0, st,
                               w->pW, pg, w->pT);
            hipLaunchKernelGGL((k_proj_solve<S, ND, true>), dim3(1), dim3(64), 0, st, w->pT,
                               gprev, gout, w->wout);
        }
    }
    if (!gshift) {
        launch_proj_dots<S, ND>(m, pv, w->pW, pg, st);
        hipLaunchKernelGGL(k_sum_partials, dim3(S * NP), dim3(PROJ_THREADS), 0, st, w->pW, pg,
                           w->pT);
        hipLaunchKernelGGL((k_proj_solve<S, ND, false>), dim3(1), dim3(64), 0, st, w->pT, gprev,
                           gout, w->wout);
    }
    w->gcur ^= 1;
    w->gnd = ND;
    m->stat_gram_reused += gshift ? 1 : 0;
    hipLaunchKernelGGL((k_proj_norms<S, ND>), dim3(PROJ_GRID_F), dim3(PROJ_THREADS), 0, st, m->V,
                       pv, w->pW);
    hipLaunchKernelGGL(k_sum_partials2, dim3(3 * S), dim3(PROJ_THREADS), 0, st, w->pW,
                       PROJ_GRID_F, S, w->pR, nres, w->pT);
    volatile int* seq = &w->h_wout->seq;
    *seq = 0;  // every earlier publication has been waited for
    hipLaunchKernelGGL((k_proj_accept_update<S, ND>), dim3(PROJ_GRID_F), dim3(PROJ_THREADS), 0, st,
                       m->V, pv, (SV<double, S>*)w->vb[VB_R64], w->pT, w->wout, w->d_wout_pub);
    DSV_HIP(hipGetLastError());
    if (wa.low_modes) {
        // the low-mode correction of the start taken; its residual norms arrive after the
        // warm start's outputs, so one wait covers both
        volatile int* rseq = &w->h_rpub->seq;
        *rseq = 0;
        launch_low_mode_correction<S>(m, w, wa.rtol, st);
        wait_published(rseq, st, "the low-mode correction");
    } else {
        wait_published(seq, st, "the warm start");
    }
    const WarmOut& o = *w->h_wout;
    m->stat_low_modes = 0;
    for (int s = 0; s < S; ++s) {
        const double r1 = wa.low_modes ? w->h_rpub->rr[s] : o.rr1[s];
        if (wa.low_modes && r1 != o.rr1[s]) ++m->stat_low_modes;
        rr[s] = r1;
        bb[s] = o.bb[s];
        if (wa.r_before) wa.r_before[s] = std::sqrt(o.rr0[s]);
        if (wa.r_after) wa.r_after[s] = std::sqrt(r1);
        if (wa.coef)
            for (int k = 0; k < ND; ++k) wa.coef[api[k] * S + s] = o.take[s] ? o.c[k][s] : 0.0;
    }
}

template <int S>
static void fused_warm(dsv_mesh* m, Workspace* w, const dsv_species* sp, const WarmArgs& wa,
                       int nres, double* rr, double* bb, hipStream_t st) {
    switch (wa.n_hist + wa.n_extra) {
        case 1: fused_warm_nd<S, 1>(m, w, sp, wa, nres, rr, bb, st); break;
        case 2: fused_warm_nd<S, 2>(m, w, sp, wa, nres, rr, bb, st); break;
        case 3: fused_warm_nd<S, 3>(m, w, sp, wa, nres, rr, bb, st); break;
        case 4: fused_warm_nd<S, 4>(m, w, sp, wa, nres, rr, bb, st); break;
        case 5: fused_warm_nd<S, 5>(m, w, sp, wa, nres, rr, bb, st); break;
        case 6: fused_warm_nd<S, 6>(m, w, sp, wa, nres, rr, bb, st); break;
        case 7: fused_warm_nd<S, 7>(m, w, sp, wa, nres, rr, bb, st); break;
        default: fused_warm_nd<S, 8>(m, w, sp, wa, nres, rr, bb, st); break;
    }
}

template <class T, int S>
static void apply_impl(dsv_mesh* m, const dsv_species* sp, double inv_dt, int lumped,
                       const double* const* x, double* const* y) {
    const int method = sizeof(T) == 4 ? DSV_METHOD_MF32_IR64 : DSV_METHOD_MF64;
    Workspace* w = get_ws(m, method, S);
    hipStream_t st = w->stream;
    prepare_bc(m, w, S, sp, st);
    prepare_op<T, S>(m, w, m->part, sp, inv_dt, lumped, st, 0);
    const unsigned gv = cdiv(m->V, 256);
    MFVec<T, S> v{};
    v.x = (const SV<T, S>*)w->vb[VB_V1];
    v.y = (SV<T, S>*)w->vb[VB_V2];
    hipLaunchKernelGGL((k_pack<T, S>), dim3(gv), dim3(256), 0, st, m->V, m->perm, x[0],
                       S > 1 ? x[1] : nullptr, S > 2 ? x[2] : nullptr, (SV<T, S>*)w->vb[VB_V1]);
    launch_mf<T, S, MF_APPLY>(m, m->part, w->prm, v, nullptr, st);
    hipLaunchKernelGGL((k_unpack<T, S>), dim3(gv), dim3(256), 0, st, m->V, m->perm,
                       (const SV<T, S>*)w->vb[VB_V2], y[0], S > 1 ? y[1] : nullptr,
                       S > 2 ? y[2] : nullptr);
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipStreamSynchronize(st));
}

// Times one kernel on `stream`: mode 0 = matrix-free apply; mode 1 = one matrix-free CG
// iteration K_0 (every species active, partials gamma = delta = rho = 1, so that repeated
// launches take identical decisions).
template <class T, int S>
static double time_impl(dsv_mesh* m, const dsv_species* sp, double inv_dt, int lumped, int mode,
                        int n_rep, hipStream_t stream) {
    const int method = sizeof(T) == 4 ? DSV_METHOD_MF32_IR64 : DSV_METHOD_MF64;
    Workspace* w = get_ws(m, method, S);
    hipStream_t ws = w->stream;
    prepare_bc(m, w, S, sp, ws);
    prepare_op<T, S>(m, w, m->part, sp, inv_dt, lumped, ws, 0);
    const unsigned gv = cdiv(m->V, 256);
    // inputs: x (internal order) when every x is given, else the lumped mass; the diagonal
    const bool have_x = sp[0].x && (S < 2 || sp[1].x) && (S < 3 || sp[2].x);
    const double* x0 = have_x ? sp[0].x : m->lumped_m;
    const double* x1 = S > 1 ? (have_x ? sp[1].x : m->lumped_m) : nullptr;
    const double* x2 = S > 2 ? (have_x ? sp[2].x : m->lumped_m) : nullptr;
    MFVec<T, S> dv{};
    dv.y = (SV<T, S>*)w->vb[VB_DINV];
    launch_mf<T, S, MF_DIAG>(m, m->part, w->prm, dv, nullptr, ws);
    for (int id : {VB_V1, VB_V3, VB_V5, VB_V7, VB_X})
        hipLaunchKernelGGL((k_pack<T, S>), dim3(gv), dim3(256), 0, ws, m->V, m->perm, x0, x1, x2,
                           (SV<T, S>*)w->vb[id]);
    CGState h = read_state(w, ws);
    for (int s = 0; s < MAXS; ++s) {
        h.act[1][s] = s < S ? 1 : 0;
        h.tol2[s] = 0.0;
    }
    h.kk[1] = 0;
    h.max_it = 1 << 30;
    write_state(w, h, ws);
    std::vector<T> pu((size_t)3 * MAXS * MAX_PROD, T(0));  // partials in the CG precision
    const int n = cg_grid<T, S>(m, w);
    for (int q = 0; q < 3 * S; ++q) pu[(size_t)q * n] = T(1);
    DSV_HIP(hipMemcpy(w->pU[0], pu.data(), pu.size() * sizeof(T), hipMemcpyHostToDevice));
    MFVec<T, S> v{};
    v.x = (const SV<T, S>*)w->vb[VB_V1];
    v.y = (SV<T, S>*)w->vb[VB_V2];
    auto launch = [&](hipStream_t s) {
        if (mode == 1)
            launch_cg<T, S, false>(m, w, 0, s);
        else
            launch_mf<T, S, MF_APPLY>(m, m->part, w->prm, v, nullptr, s);
    };
    DSV_HIP(hipStreamSynchronize(ws));
    for (int i = 0; i < 3; ++i) launch(stream);
    DSV_HIP(hipStreamSynchronize(stream));
    // the n_rep launches are captured in one graph and replayed, so that host launch overhead
    // does not show in a short kernel's time (the solver replays its iterations the same way)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DSV_HIP(hipStreamBeginCapture(ws, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < n_rep; ++i) launch(ws);
    DSV_HIP(hipStreamEndCapture(ws, &graph));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    try {
        DSV_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        DSV_HIP(hipGraphLaunch(exec, ws));  // warm-up replay
        DSV_HIP(hipEventCreate(&e0));
        DSV_HIP(hipEventCreate(&e1));
        DSV_HIP(hipEventRecord(e0, ws));
        DSV_HIP(hipGraphLaunch(exec, ws));
        DSV_HIP(hipEventRecord(e1, ws));
        DSV_HIP(hipEventSynchronize(e1));
        DSV_HIP(hipEventElapsedTime(&ms, e0, e1));
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (exec) (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    DSV_HIP(hipGetLastError());
    return (double)ms / n_rep;
}

}  // namespace dsv

using namespace dsv;

// ---- the low-mode subspace's build (host side)

// eigen-decomposition of a symmetric n x n matrix (row-major, overwritten) by cyclic Jacobi
// rotations: eigenvalues in ev, eigenvectors the columns of Q (row-major n x n)
static void lm_sym_eig(int n, std::vector<double>& A, std::vector<double>& ev,
                       std::vector<double>& Q) {
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double a = A[(size_t)i * n + j];
                tot += a * a;
                if (j > i) off += a * a;
            }
        if (!(off > 1e-30 * tot)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
                const double t =
                    (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - sn * akq;
                    A[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - sn * aqk;
                    A[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double qkp = Q[(size_t)k * n + p], qkq = Q[(size_t)k * n + q];
                    Q[(size_t)k * n + p] = c * qkp - sn * qkq;
                    Q[(size_t)k * n + q] = sn * qkp + c * qkq;
                }
            }
    }
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = A[(size_t)i * n + i];
}

// Rayleigh-Ritz of one species' block: G c = lambda M c (nb x nb Gram matrices of A and D).
// M = U diag(sig) U^T, B = U sig^-1/2 over the directions with sig > 1e-12 max sig, H = B^T G B
// = Q diag(lambda) Q^T; C = B Q (nb x nb row-major, columns in ascending lambda, columns past
// the rank 0 with lambda = +inf)
static void lm_rayleigh_ritz(int nb, const double* G, const double* M, std::vector<double>& C,
                             std::vector<double>& lam) {
    std::vector<double> Ms((size_t)nb * nb), sig, U;
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j)
            Ms[(size_t)i * nb + j] = 0.5 * (M[(size_t)i * nb + j] + M[(size_t)j * nb + i]);
    lm_sym_eig(nb, Ms, sig, U);
    double smax = 0.0;
    for (double x : sig) smax = std::max(smax, x);
    std::vector<int> keep;
    for (int k = 0; k < nb; ++k)
        if (sig[k] > 1e-12 * smax) keep.push_back(k);
    const int r = (int)keep.size();
    std::vector<double> B((size_t)nb * r);
    for (int i = 0; i < nb; ++i)
        for (int k = 0; k < r; ++k)
            B[(size_t)i * r + k] = U[(size_t)i * nb + keep[k]] / std::sqrt(sig[keep[k]]);
    std::vector<double> GB((size_t)nb * r, 0.0), H((size_t)r * r, 0.0);
    for (int i = 0; i < nb; ++i)
        for (int k = 0; k < r; ++k) {
            double a = 0.0;
            for (int j = 0; j < nb; ++j) a += 0.5 * (G[(size_t)i * nb + j] + G[(size_t)j * nb + i]) * B[(size_t)j * r + k];
            GB[(size_t)i * r + k] = a;
        }
    for (int k = 0; k < r; ++k)
        for (int l = 0; l < r; ++l) {
            double a = 0.0;
            for (int i = 0; i < nb; ++i) a += B[(size_t)i * r + k] * GB[(size_t)i * r + l];
            H[(size_t)k * r + l] = a;
        }
    for (int k = 0; k < r; ++k)
        for (int l = k + 1; l < r; ++l) {
            const double a = 0.5 * (H[(size_t)k * r + l] + H[(size_t)l * r + k]);
            H[(size_t)k * r + l] = H[(size_t)l * r + k] = a;
        }
    std::vector<double> hv, Qh;
    lm_sym_eig(r, H, hv, Qh);
    std::vector<int> order(r);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return hv[a] < hv[b]; });
    C.assign((size_t)nb * nb, 0.0);
    lam.assign(nb, std::numeric_limits<double>::infinity());
    for (int jj = 0; jj < r; ++jj) {
        const int j = order[jj];
        lam[jj] = hv[j];
        for (int i = 0; i < nb; ++i) {
            double a = 0.0;
            for (int k = 0; k < r; ++k) a += B[(size_t)i * r + k] * Qh[(size_t)k * r + j];
            C[(size_t)i * nb + jj] = a;
        }
    }
}

// one Chebyshev step on P y = dinv (A y), AY = A Y, per species s with centre c[s] and half
// width e[s]: first, out = (P Y - c Y) / e; otherwise out = 2 (P Y - c Y) / e - out (the
// iterate before, overwritten)
struct LmCheb {
    float c[MAXS], e[MAXS];
};

template <int S>
__global__ void k_lm_cheb_s(int V, const SV<float, S>* __restrict__ dinv,
                            const SV<float, S>* __restrict__ AY, const SV<float, S>* __restrict__ Y,
                            SV<float, S>* __restrict__ out, LmCheb ce, int first) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const SV<float, S> d = dinv[v], a = AY[v], y = Y[v];
    SV<float, S> o;
    if (first) {
#pragma unroll
        for (int s = 0; s < S; ++s) o.v[s] = (d.v[s] * a.v[s] - ce.c[s] * y.v[s]) / ce.e[s];
    } else {
        const SV<float, S> p = out[v];
#pragma unroll
        for (int s = 0; s < S; ++s)
            o.v[s] = 2.0f * (d.v[s] * a.v[s] - ce.c[s] * y.v[s]) / ce.e[s] - p.v[s];
    }
    out[v] = o;
}

// Gram matrices of a block: out[(i nb + j) S + s] (host), from the block-dot partials
template <int S, class TY>
static void lm_gram(dsv_mesh* m, int nb, const SV<float, S>* X, const SV<TY, S>* Y,
                    const SV<float, S>* dinv, double* part, double* dtot, std::vector<double>& out,
                    hipStream_t st) {
    const int nt = nb / 4;
    hipLaunchKernelGGL((k_lm_block_dots<S, TY>), dim3(LM_CHUNKS, nt * nt), dim3(LM_THREADS), 0, st,
                       m->V, nb, X, Y, dinv, part);
    const int n_out = nb * nb * S;
    hipLaunchKernelGGL(k_lm_sum, dim3(cdiv(n_out, 256)), dim3(256), 0, st, n_out, LM_CHUNKS,
                       (const double*)part, dtot);
    DSV_HIP(hipGetLastError());
    out.resize(n_out);
    DSV_HIP(hipMemcpyAsync(out.data(), dtot, n_out * sizeof(double), hipMemcpyDeviceToHost, st));
    DSV_HIP(hipStreamSynchronize(st));
}

template <int S>
static void build_low_modes(dsv_mesh* m, const dsv_species* sp, double inv_dt, int lumped, int n,
                            int degree, int passes, uint64_t seed, double* ritz,
                            hipStream_t caller) {
    Workspace* w = get_ws(m, DSV_METHOD_MF32_IR64, S);
    hipStream_t st = w->stream;
    DSV_HIP(hipEventRecord(w->ev_io[0], caller));
    DSV_HIP(hipStreamWaitEvent(st, w->ev_io[0], 0));
    w->lm_n = 0;
    dfree(w->lm_w);
    dfree(w->lm_ginv);
    dfree(w->lm_part);
    dfree(w->lm_c);
    dfree(w->lm_y);
    if (n == 0) return;
    const size_t V = m->V;
    // the operator, set up as a MF32_IR64 solve sets it up
    prepare_bc(m, w, S, sp, st);
    const bool k_ready = prepare_op<float, S>(m, w, m->part, sp, inv_dt, lumped, st, 0);
    prepare_op<double, S>(m, w, m->part, sp, inv_dt, lumped, st, 1, k_ready);
    if (!w->dinv_valid) {
        MFVec<float, S> v{};
        v.y = (SV<float, S>*)w->vb[VB_DINV];
        launch_mf<float, S, MF_DIAG>(m, m->part, w->prm, v, nullptr, st);
        w->dinv_valid = m->op_reuse;
    }
    const auto* dinv = (const SV<float, S>*)w->vb[VB_DINV];
    const int nb = n + LM_GUARD;
    std::vector<void*> tmp;
    auto take = [&](size_t bytes) {
        void* p = nullptr;
        DSV_HIP(hipMalloc(&p, std::max<size_t>(bytes, 1)));
        tmp.push_back(p);
        return p;
    };
    try {
        // Gershgorin bound of D^-1 A per species from its assembled FP64 matrix
        double bmax[MAXS], a_cut[MAXS];
        {
            auto* vals = (double*)take(m->nnz * sizeof(double));
            auto* gp = (double*)take(256 * sizeof(double));
            std::vector<double> h(256);
            for (int s = 0; s < S; ++s) {
                assemble_into(m, sp[s], inv_dt, lumped, vals, st);
                hipLaunchKernelGGL(k_lm_gershgorin, dim3(256), dim3(256), 0, st, m->V, m->indptr,
                                   m->diag_pos, (const double*)vals, gp);
                DSV_HIP(hipGetLastError());
                DSV_HIP(hipMemcpyAsync(h.data(), gp, 256 * sizeof(double), hipMemcpyDeviceToHost, st));
                DSV_HIP(hipStreamSynchronize(st));
                double b = 0.0;
                for (double x : h) b = std::max(b, x);
                if (!(b > 0.0) || !std::isfinite(b)) throw ArgError{"dsv_build_low_modes: non-finite operator"};
                bmax[s] = 1.02 * b;
                a_cut[s] = bmax[s] / 64.0;
            }
        }
        const size_t blk = (size_t)nb * V;
        auto* X = (SV<float, S>*)take(blk * sizeof(SV<float, S>));
        auto* X2 = (SV<float, S>*)take(blk * sizeof(SV<float, S>));
        auto* AX = (SV<float, S>*)take(blk * sizeof(SV<float, S>));
        auto* Yb = (SV<float, S>*)take(V * sizeof(SV<float, S>));
        auto* AY = (SV<float, S>*)take(V * sizeof(SV<float, S>));
        auto* part = (double*)take((size_t)nb * nb * S * LM_CHUNKS * sizeof(double));
        auto* dtot = (double*)take((size_t)nb * nb * S * sizeof(double));
        auto* dC = (float*)take((size_t)S * nb * nb * sizeof(float));
        hipLaunchKernelGGL((k_lm_random<S>), dim3(cdiv((int64_t)V * nb, 256)), dim3(256), 0, st,
                           m->V, nb, seed, (const uint8_t*)w->bcbits, X);
        DSV_HIP(hipGetLastError());
        auto apply32 = [&](const SV<float, S>* x, SV<float, S>* y) {
            MFVec<float, S> v{};
            v.x = x;
            v.y = y;
            launch_mf<float, S, MF_APPLY>(m, m->part, w->prm, v, nullptr, st);
        };
        std::vector<double> Gh, Mh, lamv[MAXS];
        for (int pass = 0; pass < passes; ++pass) {
            LmCheb ce{};
            // the filter raises an eigenvalue near 0 by T_m(1 + 2 a / (b - a)) against those in
            // [a, b]; m is capped so that this stays near 1e4, which leaves the block's weaker
            // low-mode components above FP32 rounding until the Rayleigh-Ritz step
            int m_pass = degree;
            for (int s = 0; s < S; ++s) {
                ce.c[s] = (float)(0.5 * (bmax[s] + a_cut[s]));
                ce.e[s] = (float)(0.5 * (bmax[s] - a_cut[s]));
                const double g = std::acosh(1.0 + 2.0 * a_cut[s] / (bmax[s] - a_cut[s]));
                if (g > 0.0) m_pass = std::min(m_pass, std::max(4, (int)(9.0 / g)));
            }
            for (int j = 0; j < nb; ++j) {
                SV<float, S>* y0 = X + (size_t)j * V;
                SV<float, S>* cur = Yb;
                SV<float, S>* prev = y0;
                apply32(y0, AY);
                hipLaunchKernelGGL((k_lm_cheb_s<S>), dim3(cdiv(m->V, 256)), dim3(256), 0, st, m->V,
                                   dinv, (const SV<float, S>*)AY, (const SV<float, S>*)y0, Yb, ce, 1);
                for (int k = 2; k <= m_pass; ++k) {
                    apply32(cur, AY);
                    hipLaunchKernelGGL((k_lm_cheb_s<S>), dim3(cdiv(m->V, 256)), dim3(256), 0, st,
                                       m->V, dinv, (const SV<float, S>*)AY,
                                       (const SV<float, S>*)cur, prev, ce, 0);
                    std::swap(cur, prev);
                }
                if (cur != y0)
                    DSV_HIP(hipMemcpyAsync(y0, cur, V * sizeof(SV<float, S>),
                                           hipMemcpyDeviceToDevice, st));
                DSV_HIP(hipGetLastError());
            }
            // Rayleigh-Ritz in the D inner product
            for (int j = 0; j < nb; ++j) apply32(X + (size_t)j * V, AX + (size_t)j * V);
            lm_gram<S, float>(m, nb, X, X, dinv, part, dtot, Mh, st);
            lm_gram<S, float>(m, nb, X, AX, nullptr, part, dtot, Gh, st);
            std::vector<float> Ch((size_t)S * nb * nb);
            for (int s = 0; s < S; ++s) {
                std::vector<double> Gs((size_t)nb * nb), Ms((size_t)nb * nb), C;
                for (int i = 0; i < nb; ++i)
                    for (int j = 0; j < nb; ++j) {
                        Gs[(size_t)i * nb + j] = Gh[((size_t)i * nb + j) * S + s];
                        Ms[(size_t)i * nb + j] = Mh[((size_t)i * nb + j) * S + s];
                    }
                lm_rayleigh_ritz(nb, Gs.data(), Ms.data(), C, lamv[s]);
                for (size_t q = 0; q < C.size(); ++q) Ch[(size_t)s * nb * nb + q] = (float)C[q];
                const double a = lamv[s][nb - 1];
                if (std::isfinite(a) && a > 0.0 && a < 0.5 * bmax[s]) a_cut[s] = a;
            }
            DSV_HIP(hipMemcpyAsync(dC, Ch.data(), Ch.size() * sizeof(float), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL((k_lm_rotate<S>), dim3(cdiv(m->V, LM_THREADS)), dim3(LM_THREADS), 0,
                               st, m->V, nb, nb, (const SV<float, S>*)X, (const float*)dC, X2);
            DSV_HIP(hipGetLastError());
            std::swap(X, X2);
        }
        // W: the first n vectors (ascending Ritz values), each scaled by a power of two to a
        // largest entry in [1/2, 1) and rounded to FP16; W^T A W of the rounded vectors under
        // the FP64 operator
        const int Vp = (int)((V + LM_GROUP - 1) / LM_GROUP * LM_GROUP);
        std::vector<float> sc((size_t)n * S);  // outlives its asynchronous upload
        {
            auto* dmx = (float*)take((size_t)2 * n * S * sizeof(float));
            hipLaunchKernelGGL((k_lm_colmax<S>), dim3(n), dim3(256), 0, st, m->V,
                               (const SV<float, S>*)X, dmx);
            DSV_HIP(hipGetLastError());
            DSV_HIP(hipMemcpyAsync(sc.data(), dmx, sc.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            DSV_HIP(hipStreamSynchronize(st));
            for (float& x : sc) {
                if (!std::isfinite(x)) throw ArgError{"dsv_build_low_modes: non-finite subspace"};
                int e = 0;
                (void)std::frexp(x, &e);  // x = f 2^e, f in [1/2, 1)
                x = x > 0.0f ? std::ldexp(1.0f, -e) : 1.0f;
            }
            DSV_HIP(hipMemcpyAsync(dmx + (size_t)n * S, sc.data(), sc.size() * sizeof(float),
                                   hipMemcpyHostToDevice, st));
            w->lm_w = dalloc<_Float16>((size_t)n * S * Vp);
            hipLaunchKernelGGL((k_lm_to_half<S>), dim3(cdiv((int64_t)n * Vp, 256)), dim3(256), 0,
                               st, m->V, Vp, n, (const float*)(dmx + (size_t)n * S), X, w->lm_w);
            DSV_HIP(hipGetLastError());
        }
        auto* Wd = (SV<double, S>*)take(V * sizeof(SV<double, S>));
        auto* AWd = (SV<double, S>*)take((size_t)n * V * sizeof(SV<double, S>));
        const auto* W = (const SV<float, S>*)X;
        for (int j = 0; j < n; ++j) {
            hipLaunchKernelGGL((k_lm_to_double<S>), dim3(cdiv(m->V, 256)), dim3(256), 0, st, m->V,
                               W + (size_t)j * V, Wd);
            MFVec<double, S> v{};
            v.x = Wd;
            v.y = AWd + (size_t)j * V;
            launch_mf<double, S, MF_APPLY>(m, m->part, w->prm2, v, nullptr, st);
        }
        std::vector<double> Gw;
        lm_gram<S, double>(m, n, W, AWd, nullptr, part, dtot, Gw, st);
        // (W^T A W)^-1 per species, by its eigen-decomposition (directions with eigenvalues
        // below 1e-12 of the largest get no weight)
        std::vector<double> ginv((size_t)S * n * n, 0.0);
        for (int s = 0; s < S; ++s) {
            std::vector<double> Gs((size_t)n * n), ev, Q;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    Gs[(size_t)i * n + j] = 0.5 * (Gw[((size_t)i * n + j) * S + s] +
                                                   Gw[((size_t)j * n + i) * S + s]);
            lm_sym_eig(n, Gs, ev, Q);
            double emax = 0.0;
            for (double x : ev) emax = std::max(emax, x);
            for (int k = 0; k < n; ++k) {
                if (!(ev[k] > 1e-12 * emax)) continue;
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j)
                        ginv[((size_t)s * n + i) * n + j] +=
                            Q[(size_t)i * n + k] * Q[(size_t)j * n + k] / ev[k];
            }
            if (ritz)
                for (int k = 0; k < n; ++k) ritz[(size_t)s * n + k] = lamv[s].empty() ? 0.0 : lamv[s][k];
        }
        w->lm_ginv = dalloc<double>(ginv.size());
        DSV_HIP(hipMemcpyAsync(w->lm_ginv, ginv.data(), ginv.size() * sizeof(double),
                               hipMemcpyHostToDevice, st));
        w->lm_part = dalloc<double>((size_t)n * S * lm_dot_chunks(Vp));
        w->lm_c = dalloc<double>((size_t)2 * LM_MAX * MAXS);  // coefficients, then totals
        w->lm_y = dalloc<double>(V * S);
        DSV_HIP(hipStreamSynchronize(st));
        for (void* p : tmp) (void)hipFree(p);
        tmp.clear();
        w->lm_key = w->key[1];
        w->lm_vp = Vp;
        w->lm_n = n;
    } catch (...) {
        (void)hipStreamSynchronize(st);
        for (void* p : tmp) (void)hipFree(p);
        dfree(w->lm_w);
        dfree(w->lm_ginv);
        dfree(w->lm_part);
        dfree(w->lm_c);
        dfree(w->lm_y);
        w->lm_n = 0;
        throw;
    }
    DSV_HIP(hipEventRecord(w->ev_io[1], st));
    DSV_HIP(hipStreamWaitEvent(caller, w->ev_io[1], 0));
}

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
        DSV_HIP(hipStreamSynchronize((hipStream_t)stream));
        switch (n_species * 100 + precision) {
            case 132: apply_impl<float, 1>(m, sp, inv_dt, lumped, x, y); break;
            case 232: apply_impl<float, 2>(m, sp, inv_dt, lumped, x, y); break;
            case 332: apply_impl<float, 3>(m, sp, inv_dt, lumped, x, y); break;
            case 164: apply_impl<double, 1>(m, sp, inv_dt, lumped, x, y); break;
            case 264: apply_impl<double, 2>(m, sp, inv_dt, lumped, x, y); break;
            default: apply_impl<double, 3>(m, sp, inv_dt, lumped, x, y); break;
        }
    });
}

int dsv_solve(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
              int32_t lumped, int32_t method, double rtol, int32_t max_it, void* stream,
              dsv_stats* stats) {
    return guard([&] {
        if (method < 0 || method > 2) throw ArgError{"unknown method"};
        if (!(rtol > 0.0)) throw ArgError{"rtol must be > 0"};
        if (max_it < 1) throw ArgError{"max_it must be >= 1"};
        check_solve_args(m, n_species, sp, true);
        DSV_HIP(hipSetDevice(m->device));
        const hipStream_t cs = (hipStream_t)stream;
        switch (n_species) {
            case 1: solve_impl<1>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs); break;
            case 2: solve_impl<2>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs); break;
            default: solve_impl<3>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs); break;
        }
    });
}

int dsv_solve_warm(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t method, double rtol, int32_t max_it, int32_t n_hist,
                   const double* const* hist, int32_t n_extra, const double* const* extra,
                   void* stream, dsv_stats* stats, double* coef, double* r_before,
                   double* r_after) {
    return guard([&] {
        if (method != DSV_METHOD_MF32_IR64 && method != DSV_METHOD_MF64)
            throw ArgError{"dsv_solve_warm: method must be matrix-free (MF32_IR64 or MF64)"};
        if (!(rtol > 0.0)) throw ArgError{"rtol must be > 0"};
        if (max_it < 1) throw ArgError{"max_it must be >= 1"};
        if (n_hist < 0 || n_extra < 0 || n_hist + n_extra > MAXD)
            throw ArgError{"dsv_solve_warm: n_hist + n_extra must be in [0, 8]"};
        if ((n_hist > 0 && !hist) || (n_extra > 0 && !extra))
            throw ArgError{"dsv_solve_warm: NULL hist or extra"};
        check_solve_args(m, n_species, sp, true);
        for (int i = 0; i < n_hist * n_species; ++i)
            if (!hist[i]) throw ArgError{"dsv_solve_warm: NULL history vector"};
        DSV_HIP(hipSetDevice(m->device));
        const hipStream_t cs = (hipStream_t)stream;
        WarmArgs wa;
        wa.n_hist = n_hist;
        wa.hist = hist;
        wa.n_extra = n_extra;
        wa.extra = extra;
        wa.coef = coef;
        wa.r_before = r_before;
        wa.r_after = r_after;
        switch (n_species) {
            case 1: solve_impl<1>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs, &wa); break;
            case 2: solve_impl<2>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs, &wa); break;
            default: solve_impl<3>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs, &wa); break;
        }
    });
}

static void warm_history_impl(dsv_mesh* m, int32_t n_species, const dsv_species* sp,
                              double inv_dt, int32_t lumped, int32_t method, double rtol,
                              int32_t max_it, int32_t n_hist, const double* const* hist,
                              int32_t n_extra, const double* const* extra, void* stream,
                              dsv_stats* stats, double* coef, double* r_before, double* r_after,
                              bool form_rhs, const double* const* f) {
    if (method != DSV_METHOD_MF32_IR64 && method != DSV_METHOD_MF64)
        throw ArgError{"dsv_solve_warm_history: method must be matrix-free (MF32_IR64 or MF64)"};
    if (!(rtol > 0.0)) throw ArgError{"rtol must be > 0"};
    if (max_it < 1) throw ArgError{"max_it must be >= 1"};
    if (n_hist < 0 || n_extra < 0 || n_hist + n_extra > MAXD)
        throw ArgError{"dsv_solve_warm_history: n_hist + n_extra must be in [0, 8]"};
    if (!hist || (n_extra > 0 && !extra))
        throw ArgError{"dsv_solve_warm_history: NULL hist or extra"};
    check_solve_args(m, n_species, sp, true);
    for (int i = 0; i < (n_hist + 1) * n_species; ++i)
        if (!hist[i]) throw ArgError{"dsv_solve_warm_history: NULL history vector"};
    for (int s = 0; s < n_species; ++s)
        for (int i = 0; i < (n_hist + 1) * n_species; ++i)
            if (hist[i] == sp[s].x)
                throw ArgError{"dsv_solve_warm_history: a solution array is also in the history"};
    DSV_HIP(hipSetDevice(m->device));
    const hipStream_t cs = (hipStream_t)stream;
    WarmArgs wa;
    wa.n_hist = n_hist;
    wa.hist = hist;
    wa.from_hist = true;
    wa.n_extra = n_extra;
    wa.extra = extra;
    wa.coef = coef;
    wa.r_before = r_before;
    wa.r_after = r_after;
    wa.form_rhs = form_rhs;
    wa.rhs_f = f;
    switch (n_species) {
        case 1: solve_impl<1>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs, &wa); break;
        case 2: solve_impl<2>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs, &wa); break;
        default: solve_impl<3>(m, sp, inv_dt, lumped, method, rtol, max_it, stats, cs, &wa); break;
    }
}

int dsv_solve_warm_history(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                           int32_t lumped, int32_t method, double rtol, int32_t max_it,
                           int32_t n_hist, const double* const* hist, int32_t n_extra,
                           const double* const* extra, void* stream, dsv_stats* stats,
                           double* coef, double* r_before, double* r_after) {
    return guard([&] {
        warm_history_impl(m, n_species, sp, inv_dt, lumped, method, rtol, max_it, n_hist, hist,
                          n_extra, extra, stream, stats, coef, r_before, r_after, false, nullptr);
    });
}

int dsv_step_warm_history(dsv_mesh* m, int32_t n_species, const dsv_species* sp,
                          const double* const* f, double inv_dt, int32_t lumped, int32_t method,
                          double rtol, int32_t max_it, int32_t n_hist, const double* const* hist,
                          int32_t n_extra, const double* const* extra, void* stream,
                          dsv_stats* stats, double* coef, double* r_before, double* r_after) {
    return guard([&] {
        if (!lumped) throw ArgError{"dsv_step_warm_history: lumped mass only"};
        if (!m || !sp || !hist) throw ArgError{"dsv_step_warm_history: NULL argument"};
        if (n_species < 1 || n_species > MAXS) throw ArgError{"n_species must be 1, 2 or 3"};
        if (n_hist < 0 || n_hist > MAXD)
            throw ArgError{"dsv_step_warm_history: n_hist must be in [0, 8]"};
        // b is written: it must be none of the arrays the step reads
        for (int s = 0; s < n_species; ++s) {
            if (!sp[s].b) throw ArgError{"dsv_step_warm_history: species b and x must be set"};
            for (int i = 0; i < (n_hist + 1) * n_species; ++i)
                if (hist[i] == sp[s].b)
                    throw ArgError{"dsv_step_warm_history: a right-hand side is also in the history"};
            for (int q = 0; q < n_species; ++q)
                if (sp[q].x == sp[s].b || (f && f[q] == sp[s].b))
                    throw ArgError{"dsv_step_warm_history: a right-hand side aliases x or f"};
        }
        warm_history_impl(m, n_species, sp, inv_dt, lumped, method, rtol, max_it, n_hist, hist,
                          n_extra, extra, stream, stats, coef, r_before, r_after, true, f);
    });
}

int dsv_build_low_modes(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                        int32_t lumped, int32_t n_modes, int32_t degree, int32_t passes,
                        uint64_t seed, double* ritz, void* stream) {
    return guard([&] {
        check_solve_args(m, n_species, sp, false);
        if (n_modes < 0 || n_modes > LM_MAX || n_modes % 4 != 0)
            throw ArgError{"dsv_build_low_modes: n_modes must be a multiple of 4 in [0, 64]"};
        if (n_modes > 0 && (degree < 1 || degree > 1000 || passes < 1 || passes > 50))
            throw ArgError{"dsv_build_low_modes: degree must be in [1, 1000] and passes in [1, 50]"};
        if (n_modes > 0 && !m->op_reuse)
            throw ArgError{"dsv_build_low_modes: operator reuse is off (dsv_set_operator_reuse): "
                           "the next solve would not keep the operator the subspace belongs to"};
        if (n_modes > 0 && !(inv_dt >= 0.0)) throw ArgError{"inv_dt must be >= 0"};
        DSV_HIP(hipSetDevice(m->device));
        const hipStream_t cs = (hipStream_t)stream;
        switch (n_species) {
            case 1: build_low_modes<1>(m, sp, inv_dt, lumped, n_modes, degree, passes, seed, ritz, cs); break;
            case 2: build_low_modes<2>(m, sp, inv_dt, lumped, n_modes, degree, passes, seed, ritz, cs); break;
            default: build_low_modes<3>(m, sp, inv_dt, lumped, n_modes, degree, passes, seed, ritz, cs); break;
        }
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
                for (Workspace* w : {m->ws[a][b], m->nws[a][b]})
                    if (w) w->graph_valid = 0;
    });
}

int dsv_set_species_narrowing(dsv_mesh* m, int32_t enable) {
    return guard([&] {
        if (!m) throw ArgError{"NULL mesh"};
        m->narrow = enable ? 1 : 0;
    });
}

int dsv_residual_norms(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                       int32_t lumped, double* rnorm, double* bnorm, void* stream) {
    return guard([&] {
        check_solve_args(m, n_species, sp, true);
        if (!rnorm) throw ArgError{"dsv_residual_norms: NULL rnorm"};
        DSV_HIP(hipSetDevice(m->device));
        DSV_HIP(hipStreamSynchronize((hipStream_t)stream));
        double rr[MAXS] = {0, 0, 0}, bb[MAXS] = {0, 0, 0};
        switch (n_species) {
            case 1: residual_norms_impl<1>(m, sp, inv_dt, lumped, rr, bb); break;
            case 2: residual_norms_impl<2>(m, sp, inv_dt, lumped, rr, bb); break;
            default: residual_norms_impl<3>(m, sp, inv_dt, lumped, rr, bb); break;
        }
        for (int s = 0; s < n_species; ++s) {
            rnorm[s] = std::sqrt(rr[s]);
            if (bnorm) bnorm[s] = std::sqrt(bb[s]);
        }
    });
}

int dsv_warm_start(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t n_dir, const double* const* dirs, double* coef,
                   double* r_before, double* r_after, void* stream) {
    return guard([&] {
        check_solve_args(m, n_species, sp, true);
        if (n_dir < 1 || n_dir > 3) throw ArgError{"n_dir must be 1, 2 or 3"};
        if (!dirs) throw ArgError{"dsv_warm_start: NULL dirs"};
        DSV_HIP(hipSetDevice(m->device));
        DSV_HIP(hipStreamSynchronize((hipStream_t)stream));
        switch (n_species * 10 + n_dir) {
#define DSV_WS_CASE(S_, N_)                                                                  \
    case S_ * 10 + N_:                                                                     \
        warm_start_impl<S_, N_>(m, sp, inv_dt, lumped, dirs, coef, r_before, r_after); \
        break;
            DSV_WS_CASE(1, 1) DSV_WS_CASE(1, 2) DSV_WS_CASE(1, 3)
            DSV_WS_CASE(2, 1) DSV_WS_CASE(2, 2) DSV_WS_CASE(2, 3)
            DSV_WS_CASE(3, 1) DSV_WS_CASE(3, 2) DSV_WS_CASE(3, 3)
#undef DSV_WS_CASE
            default: throw ArgError{"n_species must be 1, 2 or 3"};
        }
    });
}

int dsv_set_operator_reuse(dsv_mesh* m, int32_t reuse) {
    return guard([&] {
        if (!m) throw ArgError{"NULL mesh"};
        m->op_reuse = reuse ? 1 : 0;
        m->op_epoch += 1;  // discards every operator built so far
    });
}

int dsv_set_operator_form(dsv_mesh* m, int32_t form) {
    return guard([&] {
        if (!m) throw ArgError{"NULL mesh"};
        if (form != DSV_FORM_RECOMPUTE && form != DSV_FORM_STORED)
            throw ArgError{"form must be DSV_FORM_RECOMPUTE or DSV_FORM_STORED"};
        m->operator_form = form;
    });
}

int dsv_time_apply(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t precision, int32_t mode, int32_t n_rep, void* stream,
                   double* ms) {
    return guard([&] {
        check_solve_args(m, n_species, sp, false);
        if (!ms || n_rep < 1) throw ArgError{"dsv_time_apply: bad arguments"};
        if (precision != 32 && precision != 64) throw ArgError{"precision must be 32 or 64"};
        if (mode < 0 || mode > 2)
            throw ArgError{"mode must be 0 (apply), 1 (CG iteration) or 2 (CSR assembly)"};
        DSV_HIP(hipSetDevice(m->device));
        DSV_HIP(hipStreamSynchronize((hipStream_t)stream));
        hipStream_t s = (hipStream_t)stream;
        if (mode == 2) {
            if (sp[0].n_bc > 0) throw ArgError{"dsv_time_apply: mode 2 times no Dirichlet rows"};
            *ms = time_assembly(m, sp[0], inv_dt, lumped, n_rep, s);
            return;
        }
        switch (n_species * 100 + precision) {
            case 132: *ms = time_impl<float, 1>(m, sp, inv_dt, lumped, mode, n_rep, s); break;
            case 232: *ms = time_impl<float, 2>(m, sp, inv_dt, lumped, mode, n_rep, s); break;
            case 332: *ms = time_impl<float, 3>(m, sp, inv_dt, lumped, mode, n_rep, s); break;
            case 164: *ms = time_impl<double, 1>(m, sp, inv_dt, lumped, mode, n_rep, s); break;
            case 264: *ms = time_impl<double, 2>(m, sp, inv_dt, lumped, mode, n_rep, s); break;
            default: *ms = time_impl<double, 3>(m, sp, inv_dt, lumped, mode, n_rep, s); break;
        }
    });
}

}  // extern "C"
