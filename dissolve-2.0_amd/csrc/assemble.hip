// FP64 CSR assembly, Dirichlet rows and RHS (`gpu/gpu_assembler.py:60-242,472-603`).
//
// The reference scatters the 16 entries of every cell with FP64 atomicAdd through a scatter
// map.  Here the scatter is a segmented reduction staged in LDS, with no atomics and no
// element-matrix round trip through HBM.  A workgroup takes a block of consecutive CSR rows
// whose incident (row, cell) pairs number at most SEG_THREADS; lane l takes pair l, recomputes
// the cell geometry from FP64 vertex records, and leaves the four entries of the row's line of
// the element matrix in LDS.  Then every CSR entry of the block sums its contributions -- the
// LDS values the set-up listed for it (`loc`), in ascending cell order -- and the block's values
// leave as one coalesced stream.  The sums run in the same order as a per-entry gather over the
// sorted contribution list, so the result is deterministic.
//
// A scalar D needs no geometry per call: the values stream from the unit stiffness K1 and the
// unit mass matrices kept on the mesh (k_assemble_scalar, and k_assemble_scalar_valpha for a
// per-vertex reaction field with lumped mass).
#include "dsv_internal.h"
#include "kernels_common.h"

namespace dsv {

constexpr int SEG_THREADS = 256;
constexpr int SEG_PER_LANE = 1;                          // (row, cell) pairs per lane
constexpr int SEG_PAIRS = SEG_THREADS * SEG_PER_LANE;    // pairs per workgroup at most

// Caller-order vertex records (x, y, z, a) in FP64 (a = a per-vertex field, 0 when a is NULL):
// one 32-byte record per corner instead of three scattered coordinate loads.
__global__ void k_vertex_records(int32_t V, const double* __restrict__ xyz,
                                 const double* __restrict__ a, double4* __restrict__ rec) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    rec[v] = double4{xyz[v], xyz[V + (int64_t)v], xyz[2 * (int64_t)V + v], a ? a[v] : 0.0};
}

// Line i of the element matrix of cell e (the row of its corner i) into T[j * SEG_PAIRS + l],
// from the cell's vertex records p[0..3]
__device__ __forceinline__ void element_line(int32_t e, int i, const double4 (&p)[4],
                                             const double (&av)[4], int l,
                                             const int32_t* __restrict__ conn,
                                             const double* __restrict__ D, int Dmode, double Dval,
                                             const double* __restrict__ A, int Amode, double Aval,
                                             double inv_dt, int lumped, double* T) {
    const double4 p0 = p[0];
    const Vec3<double> e1{p[1].x - p0.x, p[1].y - p0.y, p[1].z - p0.z};
    const Vec3<double> e2{p[2].x - p0.x, p[2].y - p0.y, p[2].z - p0.z};
    const Vec3<double> e3{p[3].x - p0.x, p[3].y - p0.y, p[3].z - p0.z};
    // rows of adj(J): grad N_k = c_k / det (k = 1..3), grad N_0 = -(c1 + c2 + c3) / det
    Vec3<double> c[4];
    c[1] = cross(e2, e3);
    c[2] = cross(e3, e1);
    c[3] = cross(e1, e2);
    c[0] = Vec3<double>{-(c[1].x + c[2].x + c[3].x), -(c[1].y + c[2].y + c[3].y),
                        -(c[1].z + c[2].z + c[3].z)};
    // c_i as a 0/1-weighted sum (exact): a dynamic index or a select chain over c[] would be
    // lowered to a scratch-memory array
    const double s0 = i == 0, s1 = i == 1, s2 = i == 2, s3 = i == 3;
    const Vec3<double> ci{s0 * c[0].x + s1 * c[1].x + s2 * c[2].x + s3 * c[3].x,
                          s0 * c[0].y + s1 * c[1].y + s2 * c[2].y + s3 * c[3].y,
                          s0 * c[0].z + s1 * c[1].z + s2 * c[2].z + s3 * c[3].z};
    const double det = dot(e1, c[1]);
    const double dJ = fabs(det);
    const double De = cell_coef(D, Dmode, Dval, e, conn);
    // a per-vertex reaction field (loaded with the records), averaged in connectivity order
    const double al = Amode == COEF_VERT ? (((av[0] + av[1]) + av[2]) + av[3]) * 0.25
                                         : cell_coef(A, Amode, Aval, e, conn);
    // K_ij = (dJ De / 6) grad_i . grad_j = De (c_i . c_j) / (6 dJ)
    const double w = De / (6.0 * dJ);
    const double mf = inv_dt + al;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double M;
        if (lumped)
            M = (i == j) ? dJ * 5.0 / 120.0 : 0.0;
        else
            M = dJ * ((i == j) ? 2.0 : 1.0) / 120.0;
        T[j * SEG_PAIRS + l] = w * dot(ci, c[j]) + mf * M;
    }
}

// desc[b] = (first pair, number of pairs, first CSR entry, end CSR entry) of workgroup b.  A
// block's pairs contribute 4 entries each, so it has at most 4 SEG_PAIRS CSR entries and
// 4 SEG_PAIRS contributions.
__global__ __launch_bounds__(SEG_THREADS) void k_assemble_seg(
    const int4* __restrict__ desc, const int32_t* __restrict__ vinc,
    const int32_t* __restrict__ conn, const double4* __restrict__ rec,
    const int32_t* __restrict__ cptr, const uint16_t* __restrict__ loc,
    const double* __restrict__ D, int Dmode, double Dval, const double* __restrict__ A, int Amode,
    double Aval, double inv_dt, int lumped, double* __restrict__ values) {
    __shared__ double T[4 * SEG_PAIRS];               // T[j * SEG_PAIRS + pair]
    __shared__ int32_t cp[4 * SEG_PAIRS + 1];         // the block's contribution ranges
    __shared__ uint16_t lc[4 * SEG_PAIRS];            // the block's contribution locations
    // XCD-aware order: workgroup b runs on XCD b mod 8, and XCD x takes a contiguous run of
    // row blocks, so that neighbouring blocks (which share cells and vertices) meet in one L2
    const int nb = gridDim.x, x = blockIdx.x & 7, q8 = nb >> 3, r8 = nb & 7;
    const int4 d = desc[x * q8 + min(x, r8) + (int)(blockIdx.x >> 3)];
    const int l = threadIdx.x;
    const int4* conn4 = reinterpret_cast<const int4*>(conn);
    // the pairs' incidence entries, and (independent of them) the block's contribution lists,
    // staged in LDS with coalesced loads
    int32_t ent[SEG_PER_LANE];
#pragma unroll
    for (int h = 0; h < SEG_PER_LANE; ++h)
        ent[h] = l + h * SEG_THREADS < d.y ? vinc[d.x + l + h * SEG_THREADS] : -1;
    const int ne = d.w - d.z;
    for (int k = l; k <= ne; k += SEG_THREADS) cp[k] = cptr[d.z + k];
    __syncthreads();
    const int32_t kb = cp[0], nk = cp[ne] - kb;
    for (int k = l; k < nk; k += SEG_THREADS) lc[k] = loc[kb + k];
    // every lane's loads in flight together: the cells' vertices, then their records
    int4 n[SEG_PER_LANE];
#pragma unroll
    for (int h = 0; h < SEG_PER_LANE; ++h)
        n[h] = conn4[ent[h] >= 0 ? ent[h] >> 2 : 0];
    double4 p[SEG_PER_LANE][4];
    double av[SEG_PER_LANE][4];
#pragma unroll
    for (int h = 0; h < SEG_PER_LANE; ++h) {
        p[h][0] = rec[n[h].x];
        p[h][1] = rec[n[h].y];
        p[h][2] = rec[n[h].z];
        p[h][3] = rec[n[h].w];
        if (Amode == COEF_VERT) {
            av[h][0] = A[n[h].x];
            av[h][1] = A[n[h].y];
            av[h][2] = A[n[h].z];
            av[h][3] = A[n[h].w];
        } else {
            av[h][0] = av[h][1] = av[h][2] = av[h][3] = 0.0;
        }
    }
#pragma unroll
    for (int h = 0; h < SEG_PER_LANE; ++h)
        if (ent[h] >= 0)
            element_line(ent[h] >> 2, ent[h] & 3, p[h], av[h], l + h * SEG_THREADS, conn, D,
                         Dmode, Dval, A, Amode, Aval, inv_dt, lumped, T);
    __syncthreads();
    for (int q = l; q < ne; q += SEG_THREADS) {
        // in ascending contribution order, four LDS locations and values in flight at a time
        double sum = 0.0;
        int k = cp[q] - kb;
        const int k1 = cp[q + 1] - kb;
        for (; k + 4 <= k1; k += 4) {
            const int u0 = lc[k], u1 = lc[k + 1], u2 = lc[k + 2], u3 = lc[k + 3];
            const double t0 = T[u0], t1 = T[u1], t2 = T[u2], t3 = T[u3];
            sum = (((sum + t0) + t1) + t2) + t3;
        }
        for (; k < k1; ++k) sum += T[lc[k]];
        values[d.z + q] = sum;
    }
}

// loc[k] of contribution k (cell e, corner i of the row's vertex, corner jj of the column's):
// jj * SEG_PAIRS + the pair's index in its workgroup, the pair (row vertex, e) found by
// bisection in the vertex's ascending incidence list
__global__ void k_contribution_locations(int64_t n, const int32_t* __restrict__ cslot,
                                         const int32_t* __restrict__ conn,
                                         const int32_t* __restrict__ vinc_ptr,
                                         const int32_t* __restrict__ vinc,
                                         const int32_t* __restrict__ vblk,
                                         const int32_t* __restrict__ rblk,
                                         uint16_t* __restrict__ loc) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t slot = cslot[k];
    const int32_t e = slot >> 4, i = (slot >> 2) & 3, jj = slot & 3;
    const int32_t v = conn[4 * (int64_t)e + i];
    const int32_t key = 4 * e + i;
    int32_t lo = vinc_ptr[v], hi = vinc_ptr[v + 1] - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (vinc[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int32_t pair = lo - vinc_ptr[rblk[vblk[v]]];
    loc[k] = (uint16_t)(jj * SEG_PAIRS + pair);
}

int coef_mode(const double* p, int32_t len, const dsv_mesh* m);

static void launch_row_assembly(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped,
                                double* values, hipStream_t st) {
    hipLaunchKernelGGL(k_assemble_seg, dim3(m->n_rblk), dim3(SEG_THREADS), 0, st,
                       (const int4*)m->rdesc, m->vinc, m->conn, (const double4*)m->vrec0,
                       m->cptr, m->loc, s.D, coef_mode(s.D, s.D_len, m), s.D_value, s.alpha,
                       coef_mode(s.alpha, s.alpha_len, m), s.alpha ? 0.0 : s.alpha_value, inv_dt,
                       lumped ? 1 : 0, values);
}

// Scalar coefficients (D and alpha both scalars): A = D K1 + (inv_dt + alpha) M1 entry by
// entry, K1 = K(D = 1) and M1 the mass matrix (consistent or lumped) being assembled once per
// mesh by k_assemble_seg; one coalesced pass over three FP64 nnz arrays.
__global__ __launch_bounds__(256) void k_assemble_scalar(int64_t nnz, const double* __restrict__ K1,
                                                        const double* __restrict__ M1, double D,
                                                        double mf, double* __restrict__ values) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) values[p] = D * K1[p] + mf * M1[p];
}

// Scalar D, a per-vertex reaction field alpha and lumped mass.  The cell's reaction coefficient
// is the mean of its corners' alpha, so the lumped reaction mass of row i regroups by graph
// edge: sum over cells c holding i of (|det J_c| / 24) (sum_k alpha_k / 4) = sum_k W_ik alpha_k,
// W_ik = sum over cells holding i and k of |det J_c| / 96, which is 1.25 M_ik (k != i) and
// 0.625 M_ii of the consistent unit mass M.  Off the diagonal the entries are D K1; the diagonal
// is D K1 + inv_dt lm_i + (W alpha)_i.  VA_LANES lanes (one DPP row) share a CSR row, so that
// most rows take one entry per lane and every load of a lane is independent of the others.
// Workgroup b runs on XCD b mod 8, and XCD x takes a contiguous run of rows, so that the alpha
// of a row's neighbours is mostly in that XCD's own L2.
constexpr int VA_LANES = 16;
__global__ __launch_bounds__(256) void k_assemble_scalar_valpha(
    int32_t V, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const int32_t* __restrict__ diag_pos, const double* __restrict__ K1,
    const double* __restrict__ Mc, const double* __restrict__ lm, const double* __restrict__ alpha,
    double D, double inv_dt, double* __restrict__ values) {
    const int nb = gridDim.x, x = blockIdx.x & 7, q8 = nb >> 3, r8 = nb & 7;
    const int64_t blk = x * q8 + min(x, r8) + (int)(blockIdx.x >> 3);
    const int64_t t = blk * blockDim.x + threadIdx.x;
    const int32_t r = (int32_t)(t / VA_LANES);
    const int lane = (int)(t % VA_LANES);
    if (r >= V) return;  // the whole DPP row (it shares r)
    double acc = 0.0;
    const int32_t p1 = indptr[r + 1];
    for (int32_t p = indptr[r] + lane; p < p1; p += VA_LANES) {
        const int32_t k = indices[p];
        acc += ((k == r) ? 0.625 : 1.25) * Mc[p] * alpha[k];
        if (k != r) values[p] = D * K1[p];
    }
    // the sum over the 16 lanes: quads, then half-rows, then the row (every lane gets it)
    acc = add_lane_xor2(add_lane_xor1(acc));
    acc += dpp_move<DPP_ROW_HALF_MIRROR>(acc);
    acc += dpp_move<DPP_ROW_MIRROR>(acc);
    if (lane == 0) {
        const int32_t dp = diag_pos[r];
        values[dp] = D * K1[dp] + inv_dt * lm[r] + acc;
    }
}

static void ensure_row_assembly(dsv_mesh* m);

// K1 and the mass matrix of the given lumping, once per mesh (null stream, synchronised)
static void ensure_unit_operators(dsv_mesh* m, int lumped) {
    const int l = lumped ? 1 : 0;
    if (m->k1 && m->m1[l]) return;
    ensure_row_assembly(m);
    double* k = nullptr;
    double* mm = nullptr;
    try {
        if (!m->k1) {
            k = dalloc<double>(m->nnz);
            dsv_species u{};
            u.D_value = 1.0;
            launch_row_assembly(m, u, 0.0, 1, k, 0);
        }
        if (!m->m1[l]) {
            mm = dalloc<double>(m->nnz);
            dsv_species u{};
            launch_row_assembly(m, u, 1.0, l, mm, 0);
        }
        DSV_HIP(hipGetLastError());
        DSV_HIP(hipStreamSynchronize(0));
    } catch (...) {
        dfree(k);
        dfree(mm);
        throw;
    }
    if (k) m->k1 = k;
    if (mm) m->m1[l] = mm;
}

// The assembly of species s: the streaming forms for a scalar D (with a scalar reaction, or a
// per-vertex one under lumped mass), the row assembly otherwise
static void launch_assembly(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped,
                            double* values, hipStream_t st) {
    if (!s.D && !s.alpha) {
        ensure_unit_operators(m, lumped);
        hipLaunchKernelGGL(k_assemble_scalar, dim3(cdiv(m->nnz, 256)), dim3(256), 0, st, m->nnz,
                           (const double*)m->k1, (const double*)m->m1[lumped ? 1 : 0], s.D_value,
                           inv_dt + s.alpha_value, values);
    } else if (!s.D && lumped && coef_mode(s.alpha, s.alpha_len, m) == COEF_VERT) {
        ensure_unit_operators(m, 0);  // K1 and the consistent unit mass
        hipLaunchKernelGGL(k_assemble_scalar_valpha, dim3(cdiv((int64_t)m->V * VA_LANES, 256)),
                           dim3(256), 0, st, m->V, m->indptr, m->indices, m->diag_pos,
                           (const double*)m->k1, (const double*)m->m1[0], m->lumped_m, s.alpha,
                           s.D_value, inv_dt, values);
    } else {
        launch_row_assembly(m, s, inv_dt, lumped, values, st);
    }
}

// Set-up of the assembly, once per mesh: the incident cells of every vertex (entry
// cell * 4 + local corner, ascending cell id; host counting sort), the row blocks of the
// workgroups (at most SEG_THREADS incident pairs each) and the contribution locations.
static void ensure_row_assembly(dsv_mesh* m) {
    if (m->rblk) return;
    const int32_t V = m->V, C = m->C;
    std::vector<int32_t> ptr((size_t)V + 1, 0), inc(4 * (size_t)C);
    for (int64_t k = 0; k < 4 * (int64_t)C; ++k) ++ptr[m->h_conn[k] + 1];
    for (int32_t v = 0; v < V; ++v) ptr[v + 1] += ptr[v];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < C; ++e)
        for (int i = 0; i < 4; ++i) inc[fill[m->h_conn[4 * e + i]]++] = (int32_t)(4 * e + i);
    std::vector<int32_t> blk{0}, vblk(V);
    int32_t r = 0;
    while (r < V) {
        const int32_t start = r;
        while (r < V && ptr[r + 1] - ptr[start] <= SEG_PAIRS) vblk[r++] = (int32_t)blk.size() - 1;
        if (r == start) throw ArgError{"dsv_assemble: a vertex has more incident cells than a "
                                       "workgroup takes"};
        blk.push_back(r);
    }
    std::vector<int32_t> ip((size_t)V + 1);
    DSV_HIP(hipMemcpy(ip.data(), m->indptr, ((size_t)V + 1) * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
    std::vector<int32_t> desc;
    for (size_t b = 0; b + 1 < blk.size(); ++b) {
        const int32_t r0 = blk[b], r1 = blk[b + 1];
        desc.insert(desc.end(), {ptr[r0], ptr[r1] - ptr[r0], ip[r0], ip[r1]});
    }
    auto up = [](const std::vector<int32_t>& h) {
        int32_t* d = dalloc<int32_t>(h.size());
        DSV_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        return d;
    };
    int32_t* d_vblk = nullptr;
    try {
        m->vrec0 = dalloc<double>(4 * (size_t)V);
        hipLaunchKernelGGL(k_vertex_records, dim3(cdiv(V, 256)), dim3(256), 0, 0, V, m->xyz,
                           (const double*)nullptr, (double4*)m->vrec0);
        m->rdesc = up(desc);
        m->vinc_ptr = up(ptr);
        m->vinc = up(inc);
        d_vblk = up(vblk);
        m->rblk = up(blk);
        m->n_rblk = (int32_t)blk.size() - 1;
        const int64_t n = 16 * (int64_t)C;
        m->loc = dalloc<uint16_t>(n);
        hipLaunchKernelGGL(k_contribution_locations, dim3(cdiv(n, 256)), dim3(256), 0, 0, n,
                           m->cslot, m->conn, m->vinc_ptr, m->vinc, d_vblk, m->rblk, m->loc);
        DSV_HIP(hipGetLastError());
        DSV_HIP(hipDeviceSynchronize());
    } catch (...) {
        // everything is built again on the next call
        dfree(d_vblk);
        dfree(m->vrec0);
        dfree(m->rdesc);
        dfree(m->vinc_ptr);
        dfree(m->vinc);
        dfree(m->rblk);
        dfree(m->loc);
        throw;
    }
    dfree(d_vblk);
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
        a.b[s][r] = lumped_rhs(mr, inv_dt, a.u[s][r], fr);
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
    ensure_row_assembly(m);
    launch_assembly(m, s, inv_dt, lumped, values, st);
    if (s.n_bc > 0)
        hipLaunchKernelGGL(k_bc_rows, dim3(cdiv(s.n_bc, 128)), dim3(128), 0, st, s.n_bc,
                           s.bc_dofs, m->indptr, m->indices, values);
    DSV_HIP(hipGetLastError());
}

// Average device time (ms) of one launch of the assembly of species s (no Dirichlet rows; the
// kernel assemble_into takes for its coefficients) into a scratch buffer: n_rep launches
// captured in one graph and replayed between two events.
double time_assembly(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped, int n_rep,
                     hipStream_t stream) {
    check_species_coefs(m, s);
    ensure_row_assembly(m);
    double* out = dalloc<double>(m->nnz);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    auto launch = [&](hipStream_t st) { launch_assembly(m, s, inv_dt, lumped, out, st); };
    try {
        for (int i = 0; i < 3; ++i) launch(stream);
        DSV_HIP(hipStreamSynchronize(stream));
        DSV_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < n_rep; ++i) launch(stream);
        DSV_HIP(hipStreamEndCapture(stream, &graph));
        DSV_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        DSV_HIP(hipGraphLaunch(exec, stream));
        DSV_HIP(hipEventCreate(&e0));
        DSV_HIP(hipEventCreate(&e1));
        DSV_HIP(hipEventRecord(e0, stream));
        DSV_HIP(hipGraphLaunch(exec, stream));
        DSV_HIP(hipEventRecord(e1, stream));
        DSV_HIP(hipEventSynchronize(e1));
        DSV_HIP(hipEventElapsedTime(&ms, e0, e1));
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        dfree(out);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    dfree(out);
    return (double)ms / n_rep;
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
