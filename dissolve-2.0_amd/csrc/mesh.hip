// Mesh set-up: validation, per-cell geometry, CSR pattern + scatter map, vertex renumbering
// and the matrix-free partition.  Replaces GPUSparseAssembler.__init__
// (`gpu/gpu_assembler.py:369-450`), whose pattern and scatter map are Python set/dict loops,
// with a GPU radix sort of the 16 (row, col) keys of every cell.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "dsv_internal.h"

namespace dsv {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------------------
// pattern kernels

__global__ void k_make_keys(int64_t n_slots, const int32_t* __restrict__ conn, uint64_t V,
                            uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    int64_t e = s >> 4;
    int i = (int)((s >> 2) & 3), j = (int)(s & 3);
    uint64_t r = (uint64_t)conn[4 * e + i];
    uint64_t c = (uint64_t)conn[4 * e + j];
    keys[s] = r * V + c;
    vals[s] = (int32_t)s;
}

__global__ void k_flags(int64_t n, const uint64_t* __restrict__ keys, int32_t* __restrict__ flag) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    flag[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

__global__ void k_unique(int64_t n, uint64_t V, const uint64_t* __restrict__ keys,
                         const int32_t* __restrict__ slot, const int32_t* __restrict__ pos1,
                         int32_t* __restrict__ smap, int32_t* __restrict__ indices,
                         int32_t* __restrict__ cptr, int32_t* __restrict__ indptr,
                         int32_t* __restrict__ diag_pos) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int32_t p = pos1[k] - 1;
    smap[slot[k]] = p;
    uint64_t key = keys[k];
    bool first = (k == 0 || keys[k - 1] != key);
    if (!first) return;
    uint64_t r = key / V, c = key % V;
    indices[p] = (int32_t)c;
    cptr[p] = (int32_t)k;
    if (k == 0 || keys[k - 1] / V != r) indptr[r] = p;
    if (r == c) diag_pos[r] = p;
}

__global__ void k_detj(int32_t C, const int32_t* __restrict__ conn, const double* __restrict__ xyz,
                       int32_t V, double* __restrict__ detJ, double* __restrict__ Gflat,
                       double* __restrict__ JinvT) {
    int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= C) return;
    const double* X = xyz;
    const double* Y = xyz + V;
    const double* Z = xyz + 2 * (int64_t)V;
    int n0 = conn[4 * e], n1 = conn[4 * e + 1], n2 = conn[4 * e + 2], n3 = conn[4 * e + 3];
    // J columns are the edge vectors x_k - x_0
    double J[3][3] = {{X[n1] - X[n0], X[n2] - X[n0], X[n3] - X[n0]},
                      {Y[n1] - Y[n0], Y[n2] - Y[n0], Y[n3] - Y[n0]},
                      {Z[n1] - Z[n0], Z[n2] - Z[n0], Z[n3] - Z[n0]}};
    double a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    double a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    double a10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double a11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    double a12 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    double a20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double a21 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    double a22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double det = J[0][0] * a00 + J[0][1] * a10 + J[0][2] * a20;
    detJ[e] = fabs(det);
    if (Gflat) {
        double id = 1.0 / det;
        double Ji[3][3] = {{a00 * id, a01 * id, a02 * id},
                           {a10 * id, a11 * id, a12 * id},
                           {a20 * id, a21 * id, a22 * id}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double g = 0.0;
                for (int k = 0; k < 3; ++k) g += Ji[r][k] * Ji[c][k];
                Gflat[9 * (int64_t)e + 3 * r + c] = g;
                JinvT[9 * (int64_t)e + 3 * r + c] = Ji[c][r];
            }
    }
}

// lumped nodal mass m_r = sum over cells containing r of |detJ|/24, in ascending cell order
// (the contributions of the diagonal CSR entry are exactly the cells containing r).
__global__ void k_lumped_mass(int32_t V, const int32_t* __restrict__ diag_pos,
                              const int32_t* __restrict__ cptr, const int32_t* __restrict__ cslot,
                              const double* __restrict__ detJ, double* __restrict__ m) {
    int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= V) return;
    int32_t p = diag_pos[r];
    double s = 0.0;
    for (int32_t k = cptr[p]; k < cptr[p + 1]; ++k) s += detJ[cslot[k] >> 4] * (5.0 / 120.0);
    m[r] = s;
}

// internal-order AoS coordinates (x, y, z, 0): one 16/32-byte load per vertex
template <class T>
__global__ void k_aos4_internal(int32_t V, const int32_t* __restrict__ perm,
                                const double* __restrict__ xyz, T* __restrict__ out) {
    int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int32_t c = perm[v];
    for (int d = 0; d < 3; ++d) out[4 * (int64_t)v + d] = (T)xyz[(int64_t)d * V + c];
    out[4 * (int64_t)v + 3] = T(0);
}

static inline unsigned grid_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

static uint64_t spread3(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Morton (Z-order) renumbering of the vertices: contiguous internal index ranges are then
// compact in space, which is what the matrix-free partition relies on.
static void morton_order(const double* xyz_aos, int32_t V, std::vector<int32_t>& perm) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t v = 0; v < V; ++v)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], xyz_aos[3 * (int64_t)v + d]);
            hi[d] = std::max(hi[d], xyz_aos[3 * (int64_t)v + d]);
        }
    double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<uint64_t> key(V);
    const double scale = (double)((1 << 21) - 1) / span;
    for (int32_t v = 0; v < V; ++v) {
        uint64_t q[3];
        for (int d = 0; d < 3; ++d)
            q[d] = (uint64_t)std::llround((xyz_aos[3 * (int64_t)v + d] - lo[d]) * scale);
        key[v] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    }
    perm.resize(V);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int32_t a, int32_t b) { return key[a] < key[b]; });
}

void free_partition(Partition& p) {
    dfree(p.hptr);
    dfree(p.hlist);
    dfree(p.lvl);
    dfree(p.lperm);
    dfree(p.sptr);
    dfree(p.ell);
    dfree(p.ell_cell);
    dfree(p.eptr);
    dfree(p.elperm);
    dfree(p.eidx);
    dfree(p.ewt);
    dfree(p.ewt64);
    dfree(p.emap);
    p.n_edge_entries = 0;
    p.nb = 0;
    p.n_entries = 0;
    p.n_halo = 0;
    p.max_local = 0;
}

// K(D = 1) (stiffness of unit diffusivity, no mass, no Dirichlet rows) and the CSR pattern on
// the host, assembled once by the deterministic FP64 assembly
static void ensure_unit_stiffness(dsv_mesh* m) {
    if (!m->h_kgeo.empty()) return;
    dsv_species unit{};
    unit.D_value = 1.0;
    double* kd = dalloc<double>(m->nnz);
    try {
        assemble_into(m, unit, 0.0, 1, kd, 0);
        DSV_HIP(hipDeviceSynchronize());
        m->h_kgeo.resize(m->nnz);
        m->h_indptr.resize((size_t)m->V + 1);
        m->h_indices.resize(m->nnz);
        DSV_HIP(hipMemcpy(m->h_kgeo.data(), kd, m->nnz * sizeof(double), hipMemcpyDeviceToHost));
        DSV_HIP(hipMemcpy(m->h_indptr.data(), m->indptr, ((size_t)m->V + 1) * sizeof(int32_t),
                          hipMemcpyDeviceToHost));
        DSV_HIP(hipMemcpy(m->h_indices.data(), m->indices, m->nnz * sizeof(int32_t),
                          hipMemcpyDeviceToHost));
    } catch (...) {
        dfree(kd);
        throw;
    }
    dfree(kd);
}

template <class T>
static T* upload(const std::vector<T>& h) {
    T* d = dalloc<T>(std::max<size_t>(1, h.size()));
    if (!h.empty()) DSV_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Block b owns internal vertices [b*NV, min((b+1)*NV, V)).  Every owned vertex gets the
// cells incident to it (ascending caller cell id) with the other three corners renumbered
// locally (owned first, then the sorted halo), dealt round-robin to its L lanes; vertices are
// ranked by decreasing degree so that the lanes of a slice have similar list lengths.
void build_partition(dsv_mesh* m, int NV, int L) {
    if (NV < 64 || NV > 1024 || NV % 64 != 0)
        throw ArgError{"vertices_per_block must be a multiple of 64 in [64, 1024]"};
    if (L != 1 && L != 2 && L != 4) throw ArgError{"lanes_per_vertex must be 1, 2 or 4"};
    if (NV * L > 1024) throw ArgError{"vertices_per_block * lanes_per_vertex must be <= 1024"};
    const int32_t C = m->C, V = m->V;
    const int nb = (V + NV - 1) / NV;
    const int vps = 64 / L;        // vertices per slice
    const int spb = NV / vps;      // slices per block
    const int32_t* cn = m->h_conn.data();
    const int32_t* ip = m->h_iperm.data();
    // vertex -> incident cells, internal vertex order, ascending caller cell id
    std::vector<int32_t> vinc((size_t)V + 1, 0);
    for (int64_t k = 0; k < 4 * (int64_t)C; ++k) vinc[ip[cn[k]] + 1]++;
    for (int32_t v = 0; v < V; ++v) vinc[v + 1] += vinc[v];
    std::vector<int32_t> inc_cell(4 * (size_t)C);
    {
        std::vector<int32_t> fill(vinc.begin(), vinc.end() - 1);
        for (int64_t e = 0; e < C; ++e)
            for (int k = 0; k < 4; ++k) inc_cell[fill[ip[cn[4 * e + k]]]++] = (int32_t)e;
    }
    std::vector<int32_t> hptr(nb + 1, 0), hlist, halo;
    std::vector<uint16_t> lperm((size_t)nb * NV, 0);
    std::vector<int32_t> sptr((size_t)nb * spb + 1, 0);
    std::vector<uint32_t> ell;
    std::vector<int32_t> ell_cell;
    std::vector<int> order;
    int max_local = 0;
    // stored-geometry operator: K(D = 1) per graph edge, assembled once (FP64, deterministic)
    ensure_unit_stiffness(m);
    const int32_t* kp = m->h_indptr.data();
    const int32_t* ki = m->h_indices.data();
    const double* kv = m->h_kgeo.data();
    std::vector<int32_t> perm_h(V);
    for (int32_t c = 0; c < V; ++c) perm_h[ip[c]] = c;
    std::vector<int32_t> eptr((size_t)nb * spb + 1, 0);
    std::vector<uint16_t> elperm((size_t)nb * NV, 0), eidx;
    std::vector<float> ewt;
    std::vector<double> ewt64;
    std::vector<int32_t> emap;
    std::vector<int> eorder;
    for (int b = 0; b < nb; ++b) {
        const int v0 = b * NV, nv = std::min(NV, V - v0);
        halo.clear();
        for (int i = 0; i < nv; ++i)
            for (int32_t j = vinc[v0 + i]; j < vinc[v0 + i + 1]; ++j)
                for (int k = 0; k < 4; ++k) {
                    const int32_t g = ip[cn[4 * (int64_t)inc_cell[j] + k]];
                    if (g < v0 || g >= v0 + nv) halo.push_back(g);
                }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        const int nloc = nv + (int)halo.size();
        if (nloc > ELL_MAX_LOCAL)
            throw ArgError{"partition block has " + std::to_string(nloc) + " local vertices (max " +
                           std::to_string(ELL_MAX_LOCAL) + "); use fewer vertices_per_block"};
        max_local = std::max(max_local, nloc);
        auto local = [&](int32_t g) -> uint32_t {
            if (g >= v0 && g < v0 + nv) return (uint32_t)(g - v0);
            return (uint32_t)(nv + (std::lower_bound(halo.begin(), halo.end(), g) - halo.begin()));
        };
        order.resize(nv);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
            return vinc[v0 + a + 1] - vinc[v0 + a] > vinc[v0 + c + 1] - vinc[v0 + c];
        });
        for (int r = 0; r < nv; ++r) lperm[(size_t)b * NV + r] = (uint16_t)order[r];
        for (int s = 0; s < spb; ++s) {
            const size_t sidx = (size_t)b * spb + s;
            const int r0 = vps * s, r1 = std::min(nv, vps * s + vps);  // degree ranks
            int width = 0;
            for (int r = r0; r < r1; ++r) {
                const int v = v0 + order[r];
                width = std::max(width, (vinc[v + 1] - vinc[v] + L - 1) / L);
            }
            const size_t base = ell.size();
            if (base + 64 * (size_t)width >= (size_t)INT32_MAX)
                throw ArgError{"matrix-free incidence lists too large for int32 offsets"};
            sptr[sidx + 1] = (int32_t)(base + 64 * (size_t)width);
            ell.resize(base + 64 * (size_t)width, ELL_PAD);
            ell_cell.resize(base + 64 * (size_t)width, -1);
            for (int r = r0; r < r1; ++r) {
                const int v = v0 + order[r];
                for (int32_t i = 0; i < vinc[v + 1] - vinc[v]; ++i) {
                    const int32_t e = inc_cell[vinc[v] + i];
                    uint32_t pk = 0;
                    int f = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int32_t g = ip[cn[4 * (int64_t)e + k]];
                        if (g == v) continue;
                        pk |= local(g) << (10 * f++);
                    }
                    const int lane = (r - r0) * L + i % L;
                    const size_t at = base + 64 * (size_t)(i / L) + lane;
                    ell[at] = pk;
                    ell_cell[at] = e;
                }
            }
        }
        // edge lists of the block's owned vertices (graph neighbours w != v, CSR order)
        auto ndeg = [&](int i) { return kp[perm_h[v0 + i] + 1] - kp[perm_h[v0 + i]] - 1; };
        eorder.resize(nv);
        std::iota(eorder.begin(), eorder.end(), 0);
        std::stable_sort(eorder.begin(), eorder.end(), [&](int a, int c) { return ndeg(a) > ndeg(c); });
        for (int r = 0; r < nv; ++r) elperm[(size_t)b * NV + r] = (uint16_t)eorder[r];
        for (int s = 0; s < spb; ++s) {
            const size_t sidx = (size_t)b * spb + s;
            const int r0 = vps * s, r1 = std::min(nv, vps * s + vps);
            int width = 0;
            for (int r = r0; r < r1; ++r) width = std::max(width, (ndeg(eorder[r]) + L - 1) / L);
            const size_t base = eidx.size();
            if (base + 64 * (size_t)width >= (size_t)INT32_MAX)
                throw ArgError{"stored-geometry lists too large for int32 offsets"};
            eptr[sidx + 1] = (int32_t)(base + 64 * (size_t)width);
            eidx.resize(base + 64 * (size_t)width, 0);
            ewt.resize(base + 64 * (size_t)width, 0.0f);
            ewt64.resize(base + 64 * (size_t)width, 0.0);
            emap.resize(base + 64 * (size_t)width, -1);
            for (int r = r0; r < r1; ++r) {
                const int ov = eorder[r];
                const int c = perm_h[v0 + ov];
                for (int k = 0; k < L; ++k)  // padding: the lane's own vertex, weight 0
                    for (int t = 0; t < width; ++t)
                        eidx[base + 64 * (size_t)t + (r - r0) * L + k] = (uint16_t)ov;
                int i = 0;
                for (int32_t j = kp[c]; j < kp[c + 1]; ++j) {
                    if (ki[j] == c) continue;
                    const size_t at = base + 64 * (size_t)(i / L) + (r - r0) * L + i % L;
                    eidx[at] = (uint16_t)local(ip[ki[j]]);
                    ewt[at] = (float)kv[j];
                    ewt64[at] = kv[j];
                    emap[at] = j;
                    ++i;
                }
            }
        }
        hlist.insert(hlist.end(), halo.begin(), halo.end());
        if ((int64_t)hlist.size() >= (int64_t)INT32_MAX) throw ArgError{"halo lists too large"};
        hptr[b + 1] = (int32_t)hlist.size();
    }
    // padded local vertex lists: one load per staged vertex, no dependence on the halo ranges
    std::vector<int32_t> lvl((size_t)nb * max_local);
    for (int b = 0; b < nb; ++b) {
        const int v0 = b * NV, nv = std::min(NV, V - v0);
        const int nh = hptr[b + 1] - hptr[b];
        int32_t* row = lvl.data() + (size_t)b * max_local;
        for (int i = 0; i < max_local; ++i)
            row[i] = i < nv ? v0 + i : (i < nv + nh ? hlist[hptr[b] + i - nv] : v0);
    }
    Partition p;
    p.NV = NV;
    p.L = L;
    p.nb = nb;
    p.threads = m->part.threads ? m->part.threads : 256;
    p.n_entries = (int64_t)ell.size();
    p.n_edge_entries = (int64_t)eidx.size();
    p.n_halo = (int64_t)hlist.size();
    p.max_local = max_local;
    try {
        p.hptr = upload(hptr);
        p.hlist = upload(hlist);
        p.lvl = upload(lvl);
        p.eptr = upload(eptr);
        p.elperm = upload(elperm);
        p.eidx = upload(eidx);
        p.ewt = upload(ewt);
        p.ewt64 = upload(ewt64);
        p.emap = upload(emap);
        p.lperm = upload(lperm);
        p.sptr = upload(sptr);
        p.ell = upload(ell);
        p.ell_cell = upload(ell_cell);
    } catch (...) {
        free_partition(p);
        throw;
    }
    free_partition(m->part);
    m->part = p;
}

void ensure_geometry_views(dsv_mesh* m) {
    if (m->Gflat) return;
    m->Gflat = dalloc<double>(9 * (size_t)m->C);
    m->JinvT = dalloc<double>(9 * (size_t)m->C);
    hipLaunchKernelGGL(k_detj, dim3(grid_for(m->C, 256)), dim3(256), 0, 0, m->C, m->conn,
                       m->xyz, m->V, m->detJ, m->Gflat, m->JinvT);
    DSV_HIP(hipGetLastError());
    DSV_HIP(hipDeviceSynchronize());
}

static void destroy_mesh(dsv_mesh* m) {
    if (!m) return;
    free_workspaces(m);
    free_partition(m->part);
    dfree(m->conn);
    dfree(m->xyz);
    dfree(m->detJ);
    dfree(m->Gflat);
    dfree(m->JinvT);
    dfree(m->gsten);
    dfree(m->gsw);
    dfree(m->k1);
    dfree(m->m1[0]);
    dfree(m->m1[1]);
    dfree(m->indptr);
    dfree(m->indices);
    dfree(m->diag_pos);
    dfree(m->smap);
    dfree(m->cptr);
    dfree(m->cslot);
    dfree(m->lumped_m);
    dfree(m->vinc_ptr);
    dfree(m->vinc);
    dfree(m->rblk);
    dfree(m->loc);
    dfree(m->vrec0);
    dfree(m->rdesc);
    dfree(m->fvert);
    for (auto& q : m->phys) dfree(q);
    dfree(m->perm);
    dfree(m->iperm);
    dfree(m->xyz32);
    dfree(m->xyz64);
    delete m;
}

static void create_mesh(dsv_mesh** out, int32_t C, int32_t V, const int32_t* conn,
                        const double* xyz, int32_t NV) {
    if (!out || !conn || !xyz) throw ArgError{"dsv_mesh_create: NULL argument"};
    if (C <= 0 || V <= 0) throw ArgError{"dsv_mesh_create: n_cells and n_verts must be > 0"};
    if ((int64_t)C * 16 >= (int64_t)INT32_MAX)
        throw ArgError{"dsv_mesh_create: 16*n_cells must fit in int32"};
    // validate indices, usage and orientation-independent non-degeneracy on the host
    std::vector<uint8_t> used(V, 0);
    for (int64_t k = 0; k < 4 * (int64_t)C; ++k) {
        int32_t v = conn[k];
        if (v < 0 || v >= V) throw ArgError{"dsv_mesh_create: conn index out of range"};
        used[v] = 1;
    }
    for (int32_t v = 0; v < V; ++v)
        if (!used[v]) {
            set_error("dsv_mesh_create: vertex " + std::to_string(v) + " is used by no cell");
            throw ArgError{"dsv_mesh_create: vertex " + std::to_string(v) + " is used by no cell"};
        }
    for (int64_t e = 0; e < C; ++e) {
        const double* p0 = xyz + 3 * (int64_t)conn[4 * e];
        double J[3][3];
        for (int k = 0; k < 3; ++k) {
            const double* pk = xyz + 3 * (int64_t)conn[4 * e + k + 1];
            for (int d = 0; d < 3; ++d) J[d][k] = pk[d] - p0[d];
        }
        double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (det == 0.0 || !std::isfinite(det))
            throw ArgError{"dsv_mesh_create: cell " + std::to_string(e) + " is degenerate"};
    }

    dsv_mesh* m = new dsv_mesh();
    try {
        DSV_HIP(hipGetDevice(&m->device));
        m->C = C;
        m->V = V;
        m->h_conn.assign(conn, conn + 4 * (int64_t)C);
        m->conn = dalloc<int32_t>(4 * (size_t)C);
        DSV_HIP(hipMemcpy(m->conn, conn, 4 * (size_t)C * sizeof(int32_t), hipMemcpyHostToDevice));
        std::vector<double> soa(3 * (size_t)V);
        for (int32_t v = 0; v < V; ++v)
            for (int d = 0; d < 3; ++d) soa[(size_t)d * V + v] = xyz[3 * (int64_t)v + d];
        m->xyz = dalloc<double>(3 * (size_t)V);
        DSV_HIP(hipMemcpy(m->xyz, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
        m->detJ = dalloc<double>(C);
        hipLaunchKernelGGL(k_detj, dim3(grid_for(C, 256)), dim3(256), 0, 0, C, m->conn, m->xyz, V,
                           m->detJ, (double*)nullptr, (double*)nullptr);
        DSV_HIP(hipGetLastError());

        // ---- CSR pattern: sort the 16C (row, col) keys, stable in the slot index ----
        const int64_t ns = 16 * (int64_t)C;
        uint64_t* kin = dalloc<uint64_t>(ns);
        uint64_t* kout = dalloc<uint64_t>(ns);
        int32_t* vin = dalloc<int32_t>(ns);
        m->cslot = dalloc<int32_t>(ns);
        hipLaunchKernelGGL(k_make_keys, dim3(grid_for(ns, 256)), dim3(256), 0, 0, ns, m->conn,
                           (uint64_t)V, kin, vin);
        DSV_HIP(hipGetLastError());
        int end_bit = 1;
        while (end_bit < 64 && ((uint64_t)1 << end_bit) <= (uint64_t)V * (uint64_t)V) ++end_bit;
        size_t tmp_bytes = 0;
        DSV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, vin, m->cslot,
                                                   (int)ns, 0, end_bit));
        int32_t* flag = dalloc<int32_t>(ns);
        int32_t* pos1 = dalloc<int32_t>(ns);
        size_t scan_bytes = 0;
        DSV_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, flag, pos1, (int)ns));
        void* tmp = dalloc<char>(std::max(tmp_bytes, scan_bytes));
        DSV_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, m->cslot,
                                                   (int)ns, 0, end_bit));
        hipLaunchKernelGGL(k_flags, dim3(grid_for(ns, 256)), dim3(256), 0, 0, ns, kout, flag);
        DSV_HIP(hipGetLastError());
        DSV_HIP(hipcub::DeviceScan::InclusiveSum(tmp, scan_bytes, flag, pos1, (int)ns));
        int32_t nnz = 0;
        DSV_HIP(hipMemcpy(&nnz, pos1 + ns - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
        m->nnz = nnz;
        m->smap = dalloc<int32_t>(ns);
        m->indices = dalloc<int32_t>(nnz);
        m->cptr = dalloc<int32_t>((size_t)nnz + 1);
        m->indptr = dalloc<int32_t>((size_t)V + 1);
        m->diag_pos = dalloc<int32_t>(V);
        hipLaunchKernelGGL(k_unique, dim3(grid_for(ns, 256)), dim3(256), 0, 0, ns, (uint64_t)V,
                           kout, m->cslot, pos1, m->smap, m->indices, m->cptr, m->indptr,
                           m->diag_pos);
        DSV_HIP(hipGetLastError());
        int32_t tail_c = (int32_t)ns, tail_r = nnz;
        DSV_HIP(hipMemcpy(m->cptr + nnz, &tail_c, sizeof(int32_t), hipMemcpyHostToDevice));
        DSV_HIP(hipMemcpy(m->indptr + V, &tail_r, sizeof(int32_t), hipMemcpyHostToDevice));
        DSV_HIP(hipDeviceSynchronize());
        dfree(kin);
        dfree(kout);
        dfree(vin);
        dfree(flag);
        dfree(pos1);
        char* tmpc = (char*)tmp;
        dfree(tmpc);

        m->lumped_m = dalloc<double>(V);
        hipLaunchKernelGGL(k_lumped_mass, dim3(grid_for(V, 256)), dim3(256), 0, 0, V, m->diag_pos,
                           m->cptr, m->cslot, m->detJ, m->lumped_m);
        DSV_HIP(hipGetLastError());

        // ---- internal numbering + partition ----
        std::vector<int32_t> perm;
        morton_order(xyz, V, perm);
        m->h_iperm.resize(V);
        for (int32_t v = 0; v < V; ++v) m->h_iperm[perm[v]] = v;
        m->perm = dalloc<int32_t>(V);
        m->iperm = dalloc<int32_t>(V);
        DSV_HIP(hipMemcpy(m->perm, perm.data(), V * sizeof(int32_t), hipMemcpyHostToDevice));
        DSV_HIP(hipMemcpy(m->iperm, m->h_iperm.data(), V * sizeof(int32_t), hipMemcpyHostToDevice));
        m->xyz32 = dalloc<float>(4 * (size_t)V);
        m->xyz64 = dalloc<double>(4 * (size_t)V);
        hipLaunchKernelGGL(k_aos4_internal<float>, dim3(grid_for(V, 256)), dim3(256), 0, 0, V,
                           m->perm, m->xyz, m->xyz32);
        hipLaunchKernelGGL(k_aos4_internal<double>, dim3(grid_for(V, 256)), dim3(256), 0, 0, V,
                           m->perm, m->xyz, m->xyz64);
        DSV_HIP(hipGetLastError());
        if (NV > 0) {
            build_partition(m, NV, 4);
        } else {
            // default: the largest block size whose blocks fit the 10-bit local indices
            for (int nv = 256;; nv /= 2) {
                try {
                    build_partition(m, nv, 4);
                    break;
                } catch (const ArgError&) {
                    if (nv == 64) throw;
                }
            }
        }
        DSV_HIP(hipDeviceSynchronize());
    } catch (...) {
        destroy_mesh(m);
        throw;
    }
    *out = m;
}

}  // namespace dsv

using namespace dsv;

extern "C" {

const char* dsv_last_error(void) { return g_last_error.c_str(); }

int dsv_device_count(int32_t* n) {
    return guard([&] {
        int c = 0;
        DSV_HIP(hipGetDeviceCount(&c));
        *n = c;
    });
}

int dsv_mesh_create(dsv_mesh** out, int32_t n_cells, int32_t n_verts, const int32_t* conn,
                    const double* xyz, int32_t vertices_per_block) {
    int rc = guard([&] { create_mesh(out, n_cells, n_verts, conn, xyz, vertices_per_block); });
    return (rc == DSV_E_ARG && std::string(dsv_last_error()).find("degenerate") != std::string::npos)
               ? DSV_E_MESH
           : (rc == DSV_E_ARG && std::string(dsv_last_error()).find("used by no cell") != std::string::npos)
               ? DSV_E_MESH
               : rc;
}

int dsv_mesh_destroy(dsv_mesh* m) {
    return guard([&] { destroy_mesh(m); });
}

int dsv_mesh_info(const dsv_mesh* m, int64_t* info) {
    return guard([&] {
        if (!m || !info) throw ArgError{"dsv_mesh_info: NULL argument"};
        info[0] = m->C;
        info[1] = m->V;
        info[2] = m->nnz;
        info[3] = m->part.nb;
        info[4] = m->part.NV;
        info[5] = m->part.n_entries;
        info[6] = m->part.n_halo;
        info[7] = m->part.max_local;
    });
}

int dsv_mesh_view(const dsv_mesh* mc, int32_t which, void** ptr) {
    return guard([&] {
        dsv_mesh* m = const_cast<dsv_mesh*>(mc);
        if (!m || !ptr) throw ArgError{"dsv_mesh_view: NULL argument"};
        switch (which) {
            case 0: *ptr = m->indptr; break;
            case 1: *ptr = m->indices; break;
            case 2: *ptr = m->smap; break;
            case 3: *ptr = m->detJ; break;
            case 4: ensure_geometry_views(m); *ptr = m->Gflat; break;
            case 5: ensure_geometry_views(m); *ptr = m->JinvT; break;
            case 6: *ptr = m->conn; break;
            default: throw ArgError{"dsv_mesh_view: unknown view"};
        }
    });
}

int dsv_mesh_copy_view(const dsv_mesh* mc, int32_t which, void* dst, void* stream) {
    return guard([&] {
        dsv_mesh* m = const_cast<dsv_mesh*>(mc);
        if (!m || !dst) throw ArgError{"dsv_mesh_copy_view: NULL argument"};
        DSV_HIP(hipSetDevice(m->device));
        void* src = nullptr;
        size_t bytes = 0;
        switch (which) {
            case 0: src = m->indptr; bytes = ((size_t)m->V + 1) * 4; break;
            case 1: src = m->indices; bytes = (size_t)m->nnz * 4; break;
            case 2: src = m->smap; bytes = (size_t)m->C * 16 * 4; break;
            case 3: src = m->detJ; bytes = (size_t)m->C * 8; break;
            case 4: ensure_geometry_views(m); src = m->Gflat; bytes = (size_t)m->C * 72; break;
            case 5: ensure_geometry_views(m); src = m->JinvT; bytes = (size_t)m->C * 72; break;
            case 6: src = m->conn; bytes = (size_t)m->C * 16; break;
            default: throw ArgError{"dsv_mesh_copy_view: unknown view"};
        }
        DSV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
        DSV_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

int dsv_mesh_set_partition(dsv_mesh* m, int32_t vertices_per_block, int32_t block_threads,
                           int32_t lanes_per_vertex) {
    return guard([&] {
        if (!m) throw ArgError{"dsv_mesh_set_partition: NULL mesh"};
        if (block_threads != 256 && block_threads != 512)
            throw ArgError{"block_threads must be 256 or 512"};
        free_workspaces(m);  // workspaces and captured graphs depend on the launch shape
        if (vertices_per_block != m->part.NV || lanes_per_vertex != m->part.L)
            build_partition(m, vertices_per_block, lanes_per_vertex);
        m->part.threads = block_threads;
    });
}

}  // extern "C"
