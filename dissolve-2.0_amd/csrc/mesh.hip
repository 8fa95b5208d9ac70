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

template <class T>
__global__ void k_soa_internal(int32_t V, const int32_t* __restrict__ perm,
                               const double* __restrict__ xyz, T* __restrict__ out) {
    int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int32_t c = perm[v];
    for (int d = 0; d < 3; ++d) out[(int64_t)d * V + v] = (T)xyz[(int64_t)d * V + c];
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
    dfree(p.bptr);
    dfree(p.bloc);
    dfree(p.belem);
    dfree(p.hptr);
    dfree(p.hlist);
    dfree(p.vptr);
    dfree(p.bslot);
    p.nb = 0;
    p.n_entries = 0;
    p.n_halo = 0;
    p.max_local = 0;
}

// Block b owns internal vertices [b*NV, min((b+1)*NV, V)); its entry list holds every cell
// with at least one owned vertex, in ascending caller cell order, with the cell's vertices
// renumbered locally (owned first, then the sorted halo).
void build_partition(dsv_mesh* m, int NV) {
    if (NV < 32 || NV > 2048) throw ArgError{"vertices_per_block must be in [32, 2048]"};
    const int32_t C = m->C, V = m->V;
    const int nb = (V + NV - 1) / NV;
    std::vector<int64_t> cnt(nb + 1, 0);
    const int32_t* cn = m->h_conn.data();
    const int32_t* ip = m->h_iperm.data();
    auto blocks_of = [&](int64_t e, int* bl) {
        int n = 0;
        for (int k = 0; k < 4; ++k) {
            int b = ip[cn[4 * e + k]] / NV;
            bool seen = false;
            for (int q = 0; q < n; ++q) seen |= (bl[q] == b);
            if (!seen) bl[n++] = b;
        }
        return n;
    };
    for (int64_t e = 0; e < C; ++e) {
        int bl[4];
        int n = blocks_of(e, bl);
        for (int q = 0; q < n; ++q) cnt[bl[q] + 1]++;
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    const int64_t ne = cnt[nb];
    if (ne >= (int64_t)INT32_MAX) throw ArgError{"partition too large for int32 entry indices"};
    std::vector<int32_t> bptr(nb + 1);
    for (int b = 0; b <= nb; ++b) bptr[b] = (int32_t)cnt[b];
    std::vector<int4> gconn(ne);
    std::vector<int32_t> belem(ne);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < C; ++e) {
        int bl[4];
        int n = blocks_of(e, bl);
        int4 c4 = make_int4(ip[cn[4 * e]], ip[cn[4 * e + 1]], ip[cn[4 * e + 2]], ip[cn[4 * e + 3]]);
        for (int q = 0; q < n; ++q) {
            int64_t k = fill[bl[q]]++;
            gconn[k] = c4;
            belem[k] = (int32_t)e;
        }
    }
    // halo lists and local numbering
    std::vector<int32_t> hptr(nb + 1, 0);
    std::vector<int32_t> vptr((size_t)V + 1, 0);
    std::vector<ushort4> bslot(ne);
    std::vector<int32_t> vcount;
    int max_entries = 0, max_inc = 0;
    std::vector<int32_t> hlist;
    std::vector<ushort4> bloc(ne);
    std::vector<int32_t> halo;
    int max_local = 0;
    for (int b = 0; b < nb; ++b) {
        const int v0 = b * NV, nv = std::min(NV, V - v0);
        halo.clear();
        for (int64_t k = bptr[b]; k < bptr[b + 1]; ++k) {
            const int32_t g4[4] = {gconn[k].x, gconn[k].y, gconn[k].z, gconn[k].w};
            for (int j = 0; j < 4; ++j)
                if (g4[j] < v0 || g4[j] >= v0 + nv) halo.push_back(g4[j]);
        }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        const int nloc = nv + (int)halo.size();
        if (nloc > 65535) throw ArgError{"partition block has more than 65535 local vertices"};
        max_local = std::max(max_local, nloc);
        for (int64_t k = bptr[b]; k < bptr[b + 1]; ++k) {
            const int32_t g4[4] = {gconn[k].x, gconn[k].y, gconn[k].z, gconn[k].w};
            unsigned short l4[4];
            for (int j = 0; j < 4; ++j) {
                const int32_t g = g4[j];
                if (g >= v0 && g < v0 + nv)
                    l4[j] = (unsigned short)(g - v0);
                else
                    l4[j] = (unsigned short)(nv + (std::lower_bound(halo.begin(), halo.end(), g) -
                                                   halo.begin()));
            }
            bloc[k] = make_ushort4(l4[0], l4[1], l4[2], l4[3]);
        }
        // owned-vertex incidence: every corner of every cell is owned by exactly one block
        const int64_t nbe = bptr[b + 1] - bptr[b];
        if (nbe > 65535) throw ArgError{"partition block lists more than 65535 cells"};
        max_entries = std::max(max_entries, (int)nbe);
        vcount.assign(nv + 1, 0);
        for (int64_t e = 0; e < nbe; ++e) {
            const ushort4 l = bloc[bptr[b] + e];
            const unsigned short l4[4] = {l.x, l.y, l.z, l.w};
            for (int j = 0; j < 4; ++j)
                if (l4[j] < nv) vcount[l4[j] + 1]++;
        }
        for (int i = 0; i < nv; ++i) vcount[i + 1] += vcount[i];
        const int32_t base = vptr[v0];
        for (int i = 0; i <= nv; ++i) vptr[v0 + i] = base + vcount[i];
        max_inc = std::max(max_inc, vcount[nv]);
        if (vcount[nv] > 65535) throw ArgError{"partition block has more than 65535 owned corners"};
        for (int64_t e = 0; e < nbe; ++e) {
            const ushort4 l = bloc[bptr[b] + e];
            const unsigned short l4[4] = {l.x, l.y, l.z, l.w};
            unsigned short s4[4];
            for (int j = 0; j < 4; ++j)
                s4[j] = l4[j] < nv ? (unsigned short)(vcount[l4[j]]++) : (unsigned short)0xFFFF;
            bslot[bptr[b] + e] = make_ushort4(s4[0], s4[1], s4[2], s4[3]);
        }
        hlist.insert(hlist.end(), halo.begin(), halo.end());
        if ((int64_t)hlist.size() >= (int64_t)INT32_MAX) throw ArgError{"halo lists too large"};
        hptr[b + 1] = (int32_t)hlist.size();
    }
    Partition p;
    p.NV = NV;
    p.nb = nb;
    p.threads = m->part.threads ? m->part.threads : 256;
    p.n_entries = ne;
    p.n_halo = (int64_t)hlist.size();
    p.max_local = max_local;
    p.max_entries = max_entries;
    p.max_inc = max_inc;
    if ((int64_t)vptr[V] != 4 * (int64_t)C) throw ArgError{"internal error: incidence count"};
    p.vptr = dalloc<int32_t>((size_t)V + 1);
    p.bslot = dalloc<ushort4>(ne);
    DSV_HIP(hipMemcpy(p.vptr, vptr.data(), ((size_t)V + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    DSV_HIP(hipMemcpy(p.bslot, bslot.data(), ne * sizeof(ushort4), hipMemcpyHostToDevice));
    p.bptr = dalloc<int32_t>(nb + 1);
    p.bloc = dalloc<ushort4>(ne);
    p.belem = dalloc<int32_t>(ne);
    p.hptr = dalloc<int32_t>(nb + 1);
    p.hlist = dalloc<int32_t>(std::max<size_t>(1, hlist.size()));
    DSV_HIP(hipMemcpy(p.bptr, bptr.data(), (nb + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    DSV_HIP(hipMemcpy(p.bloc, bloc.data(), ne * sizeof(ushort4), hipMemcpyHostToDevice));
    DSV_HIP(hipMemcpy(p.belem, belem.data(), ne * sizeof(int32_t), hipMemcpyHostToDevice));
    DSV_HIP(hipMemcpy(p.hptr, hptr.data(), (nb + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!hlist.empty())
        DSV_HIP(hipMemcpy(p.hlist, hlist.data(), hlist.size() * sizeof(int32_t), hipMemcpyHostToDevice));
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
    dfree(m->indptr);
    dfree(m->indices);
    dfree(m->diag_pos);
    dfree(m->smap);
    dfree(m->cptr);
    dfree(m->cslot);
    dfree(m->lumped_m);
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
        m->xyz32 = dalloc<float>(3 * (size_t)V);
        m->xyz64 = dalloc<double>(3 * (size_t)V);
        hipLaunchKernelGGL(k_soa_internal<float>, dim3(grid_for(V, 256)), dim3(256), 0, 0, V,
                           m->perm, m->xyz, m->xyz32);
        hipLaunchKernelGGL(k_soa_internal<double>, dim3(grid_for(V, 256)), dim3(256), 0, 0, V,
                           m->perm, m->xyz, m->xyz64);
        DSV_HIP(hipGetLastError());
        build_partition(m, NV > 0 ? NV : 128);
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

int dsv_mesh_set_partition(dsv_mesh* m, int32_t vertices_per_block, int32_t block_threads) {
    return guard([&] {
        if (!m) throw ArgError{"dsv_mesh_set_partition: NULL mesh"};
        if (block_threads != 64 && block_threads != 128 && block_threads != 256 &&
            block_threads != 512 && block_threads != 1024)
            throw ArgError{"block_threads must be 64, 128, 256, 512 or 1024"};
        free_workspaces(m);  // workspaces and captured graphs depend on the launch shape
        m->part.threads = block_threads;
        if (vertices_per_block != m->part.NV) build_partition(m, vertices_per_block);
    });
}

}  // extern "C"
