// Internal declarations of libdsv (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <utility>
#include <vector>

#include "../../include/dsv.h"

namespace dsv {

void set_error(const std::string& msg);

struct HipError {
    hipError_t code;
    std::string where;
};

#define DSV_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) throw ::dsv::HipError{_e, std::string(#expr) + " at " + \
                                                            __FILE__ + ":" +           \
                                                            std::to_string(__LINE__)}; \
    } while (0)

struct ArgError {
    std::string msg;
};

// Runs f(), turning exceptions into a status code + dsv_last_error message.
template <class F>
int guard(F&& f) {
    try {
        f();
        return DSV_OK;
    } catch (const ArgError& e) {
        set_error(e.msg);
        return DSV_E_ARG;
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return e.code == hipErrorOutOfMemory ? DSV_E_NOMEM : DSV_E_HIP;
    } catch (const std::exception& e) {
        set_error(e.what());
        return DSV_E_HIP;
    }
}

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    DSV_HIP(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// Coefficient access modes inside kernels.
enum CoefMode : int { COEF_SCALAR = 0, COEF_CELL = 1, COEF_VERT = 2 };

// Matrix-free partition: vertices (in internal Morton order) are cut into contiguous blocks
// of NV (a multiple of 64); block b owns vertices [b*NV, (b+1)*NV).  A block's local vertex
// numbering: owned vertices first (local = internal - b*NV), then its halo (the other
// vertices of the cells touching an owned vertex, ascending internal id), at most
// ELL_MAX_LOCAL in all.
//
// Vertex-centric incidence lists in sliced ELL: the owned vertices of a block are taken in
// order of decreasing degree (`lperm`: rank -> owned local index) and each gets L lanes
// (L = 1, 2 or 4); lane k of a vertex takes its incident cells k, k+L, k+2L, ...  Every 64
// consecutive lanes (64/L vertices) form a slice whose width is the longest lane list in it.
// Entry (t, lane) of slice s sits at ell[sptr[s] + 64 t + lane] and packs the local indices
// of the other three corners of that cell (10 bits each); padding entries are ELL_PAD.
constexpr int ELL_MAX_LOCAL = 1024;
constexpr uint32_t ELL_PAD = 0xFFFFFFFFu;

struct Partition {
    int NV = 0;
    int L = 4;                    // lanes per owned vertex
    int nb = 0;
    int threads = 256;
    int max_local = 0;            // max over blocks of owned + halo vertices
    int64_t n_entries = 0;        // ELL entries including padding
    int64_t n_halo = 0;
    int32_t* hptr = nullptr;      // nb+1 halo ranges
    int32_t* hlist = nullptr;     // halo vertices (internal ids)
    int32_t* lvl = nullptr;       // nb*max_local: local -> internal vertex (owned, halo, padding
                                  // repeats the block's first vertex)
    uint16_t* lperm = nullptr;    // nb*NV: degree rank -> owned local index
    int32_t* sptr = nullptr;      // nb*(NV*L/64)+1 slice offsets (multiples of 64)
    uint32_t* ell = nullptr;      // packed corner triples
    int32_t* ell_cell = nullptr;  // caller cell id per entry, -1 for padding
    // Stored-geometry operator: the same sliced-ELL layout over the graph neighbours w != v
    // of every owned vertex (lanes ranked by neighbour count, `elperm`), holding the local
    // index of w and the FP32 weight K_vw of the unit-diffusivity stiffness (padding: the
    // lane's own vertex with weight 0).
    int64_t n_edge_entries = 0;
    int32_t* eptr = nullptr;      // nb*(NV*L/64)+1 slice offsets
    uint16_t* elperm = nullptr;   // nb*NV: neighbour-count rank -> owned local index
    uint16_t* eidx = nullptr;     // local neighbour index per entry
    float* ewt = nullptr;         // K_vw(D = 1) per entry
    double* ewt64 = nullptr;      // the same weights in FP64 (FP64 residual and apply)
    int32_t* emap = nullptr;      // CSR position (caller order) of each entry, -1 for padding
};

struct Workspace;  // solver buffers, defined in solve.hip

}  // namespace dsv

struct dsv_mesh {
    int device = 0;
    int32_t C = 0, V = 0;
    int64_t nnz = 0;
    // caller numbering
    int32_t* conn = nullptr;      // 4C
    double* xyz = nullptr;        // 3V SoA (x[V], y[V], z[V])
    double* detJ = nullptr;       // C  |detJ|
    double* Gflat = nullptr;      // 9C (lazy)
    double* JinvT = nullptr;      // 9C (lazy)
    int32_t* indptr = nullptr;    // V+1
    int32_t* indices = nullptr;   // nnz
    int32_t* diag_pos = nullptr;  // V
    int32_t* smap = nullptr;      // 16C scatter map (cell*16+i*4+j -> CSR position)
    int32_t* cptr = nullptr;      // nnz+1 contribution ranges per CSR entry
    int32_t* cslot = nullptr;     // 16C slots sorted by (row, col, slot)
    double* lumped_m = nullptr;   // V  sum over incident cells of |detJ|/24
    int32_t* vinc_ptr = nullptr;  // V+1 incident-cell ranges per vertex (row assembly)
    int32_t* vinc = nullptr;      // 4C  cell * 4 + local corner, ascending cell per vertex
    int32_t* rblk = nullptr;      // n_rblk+1 first rows of the assembly workgroups
    int32_t n_rblk = 0;
    uint16_t* loc = nullptr;      // 16C LDS location of every contribution (assembly)
    double* vrec0 = nullptr;      // 4V  caller-order (x, y, z, 0) FP64 records (assembly)
    int32_t* rdesc = nullptr;     // 4 n_rblk  workgroup descriptors of the assembly
    double* fvert = nullptr;      // V  per-vertex values of a per-cell RHS source (scratch)
    double* phys[3] = {};         // V+1 each: redistancing phi0 / ping-pong, scaffold offsets
    double* gsten = nullptr;      // 3 nnz  nodal-gradient edge stencil G_rj (x, y, z) (lazy)
    double* gsw = nullptr;        // V  its cell-volume weights W_r (lazy)
    // unit operators in CSR order for scalar-coefficient assembly (lazy): K(D = 1) and the mass
    // matrix M (inv_dt = 1, no reaction), consistent [0] and lumped [1]
    double* k1 = nullptr;         // nnz
    double* m1[2] = {};           // nnz each
    double h_avg = 0.0;           // mean longest cell edge (computed once, dsv_mesh_h_avg)
    // internal (Morton) numbering for the matrix-free path
    int32_t* perm = nullptr;      // internal -> caller
    int32_t* iperm = nullptr;     // caller -> internal
    float* xyz32 = nullptr;       // 4V AoS (x, y, z, 0) internal
    double* xyz64 = nullptr;      // 4V AoS (x, y, z, 0) internal
    dsv::Partition part;          // matrix-free kernels (autotuned)
    // unit-diffusivity stiffness K(D = 1) in CSR order (FP64, host), built once for the
    // stored-geometry operator
    std::vector<double> h_kgeo;
    std::vector<int32_t> h_indptr, h_indices;
    std::vector<int32_t> h_conn;  // caller numbering
    std::vector<int32_t> h_iperm;
    // solver options
    double inner_rtol = 1e-3;
    int max_outer = 12;
    int graph_chunk = 8;  // CG iterations per poll chunk
    int use_graphs = 0;   // direct launches: on MI355X they outrun hipGraph replay here
    int operator_form = DSV_FORM_STORED;  // inner FP32 CG operator
    int op_reuse = 0;        // dsv_set_operator_reuse
    int narrow = 1;          // dsv_set_species_narrowing
    int stat_launches = 0;   // counters of the solve in progress (dsv_stats)
    int stat_narrowings = 0;
    int stat_dir_reused = 0;
    int stat_gram_reused = 0;
    int stat_low_modes = 0;  // species whose start the low-mode correction moved
    unsigned op_epoch = 0;   // bumped by every dsv_set_operator_reuse call
    // Dirichlet lists (address, length) range-checked in epoch bc_ok_epoch (operator reuse)
    mutable std::vector<std::pair<const int32_t*, int32_t>> bc_ok;
    mutable unsigned bc_ok_epoch = ~0u;
    dsv::Workspace* ws[3][4] = {};  // [method][n_species]
    dsv::Workspace* nws[3][4] = {};  // inner CG narrowed to fewer species: [method][n_species]
};

namespace dsv {
void build_partition(dsv_mesh* m, int NV, int L);
void assemble_into(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped, double* values,
                   hipStream_t st);
double time_assembly(dsv_mesh* m, const dsv_species& s, double inv_dt, int lumped, int n_rep,
                     hipStream_t stream);
void free_partition(Partition& p);
void free_workspaces(dsv_mesh* m);
void ensure_geometry_views(dsv_mesh* m);
void ensure_grad_stencil(dsv_mesh* m);
}  // namespace dsv
