/*
 * dsv.h -- C-ABI of libdsv.so, the MI355X (gfx950) transport hot path of Dissolve 2.0:
 * P1-tet backward-Euler diffusion-reaction assembly into CSR, the matrix-free apply of the
 * same operator, and the Jacobi-PCG solve for up to three species at once.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative DSV_E* code on failure; the message
 *     of the last failure on the calling thread is returned by dsv_last_error().  Nothing is
 *     thrown across the ABI.
 *   - Mesh arrays given to dsv_mesh_create are HOST pointers and are copied.  Every other
 *     array argument is a DEVICE pointer (memory of the device the mesh was created on),
 *     in the caller's vertex / cell numbering.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Calls are stream-ordered.
 *     dsv_solve and dsv_solve_warm run after the work queued on `stream` and leave the
 *     solution ordered ahead of later work on it (they return once convergence is known,
 *     with the final copy into x still queued); dsv_residual_norms and dsv_warm_start
 *     synchronise the stream before returning.
 *   - Cell / vertex coefficient arrays follow the reference's _to_cell_avg rule: length C is
 *     taken per cell, length V is averaged over the 4 vertices of each cell.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * Dissolve 2.0 repository).
 */
#ifndef DSV_H
#define DSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSV_OK 0
#define DSV_E_ARG (-1)      /* invalid argument (length, index, NULL, bad mode) */
#define DSV_E_MESH (-2)     /* degenerate cell (detJ == 0) or unused vertex */
#define DSV_E_HIP (-3)      /* HIP runtime error */
#define DSV_E_NOMEM (-4)    /* device allocation failed */
#define DSV_E_BREAKDOWN (-5) /* CG breakdown (p.Ap <= 0 or non-finite) */

/* operator forms of the inner FP32 matrix-free CG (dsv_set_operator_form) */
#define DSV_FORM_RECOMPUTE 0  /* element Jacobians recomputed on the fly from coordinates */
#define DSV_FORM_STORED 1     /* FP32 stored geometry: unit-diffusivity stiffness weights */

/* solve methods */
#define DSV_METHOD_CSR64 0    /* assembled FP64 CSR + FP64 Jacobi-PCG ("plain FP64 path") */
#define DSV_METHOD_MF32_IR64 1 /* FP32 matrix-free Jacobi-PCG inside FP64 iterative refinement */
#define DSV_METHOD_MF64 2     /* FP64 matrix-free Jacobi-PCG */

typedef struct dsv_mesh dsv_mesh;

/* One species of a batched operator / solve.  Coefficient pointers are device arrays of
 * length `*_len` (C or V); a NULL pointer means the scalar `*_value` everywhere (the
 * reference passes None for alpha -> 0, `gpu/gpu_assembler.py:459-460`). */
typedef struct dsv_species {
    const double* D;        /* diffusion coefficient De */
    int32_t D_len;
    double D_value;
    const double* alpha;    /* linear reaction coefficient */
    int32_t alpha_len;
    double alpha_value;
    const int32_t* bc_dofs; /* Dirichlet vertices (device, caller numbering) or NULL */
    int32_t n_bc;
    double bc_value;
    const double* b;        /* right-hand side (dsv_solve) */
    double* x;              /* initial guess in, solution out (dsv_solve) */
} dsv_species;

typedef struct dsv_stats {
    int32_t iterations[3];       /* CG iterations (inner FP32 iterations for MF32_IR64) */
    int32_t outer[3];            /* refinement sweeps (MF32_IR64), 1 otherwise */
    int32_t converged[3];        /* ||b - A x||_2 <= rtol ||b||_2 in FP64 */
    double rel_residual[3];      /* final true FP64 ||b - A x|| / ||b|| */
    double initial_rel_residual[3];
    double solve_ms;             /* host wall time of the call */
    int32_t cg_launches;         /* CG iterations launched, all species together (includes
                                    the no-op ones queued after convergence) */
    int32_t narrowings;          /* inner solves continued with fewer species */
    int32_t dirs_reused;         /* dsv_solve_warm: directions whose A d came from the
                                    previous call (history shifted by one step) */
    int32_t gram_reused;         /* dsv_solve_warm: 1 when the Gram entries of those
                                    directions came from the previous call too */
    int32_t low_modes;           /* dsv_solve_warm: species whose start the low-mode
                                    correction (dsv_build_low_modes) moved */
} dsv_stats;

const char* dsv_last_error(void);
int dsv_device_count(int32_t* n);

/* --- mesh / geometry / pattern: GPUSparseAssembler.__init__ (`gpu/gpu_assembler.py:369-450`) --- */

/* Copies conn (C x 4, caller numbering) and xyz (V x 3) to the current device, checks the
 * indices and that no cell is degenerate, and builds the CSR pattern, the scatter map, the
 * per-cell geometry and the matrix-free partition.  vertices_per_block: a multiple of 64 in
 * [64, 1024] whose blocks have at most 1024 local (owned + halo) vertices, or <= 0 for the
 * largest of 256, 128, 64 that satisfies this. */
int dsv_mesh_create(dsv_mesh** out, int32_t n_cells, int32_t n_verts, const int32_t* conn,
                    const double* xyz, int32_t vertices_per_block);
int dsv_mesh_destroy(dsv_mesh* m);
/* sizes (8 values): n_cells, n_verts, nnz, number of matrix-free blocks, vertices per block,
 * number of incidence-list entries of the partition (sliced ELL, padding included), total
 * halo vertices, max local vertices of a block */
int dsv_mesh_info(const dsv_mesh* m, int64_t* info8);

/* Device views owned by the mesh (valid until dsv_mesh_destroy): the attributes
 * d_indptr, d_indices, d_scatter_map, d_detJ, d_Gflat, d_JinvT, d_elem_nodes of
 * GPUSparseAssembler (`gpu/gpu_assembler.py:388-438`). which: 0 indptr (int32 V+1),
 * 1 indices (int32 nnz), 2 scatter_map (int32 16C), 3 detJ (f64 C), 4 Gflat (f64 9C),
 * 5 JinvT (f64 9C), 6 elem_nodes (int32 4C). */
int dsv_mesh_view(const dsv_mesh* m, int32_t which, void** ptr);
/* Copies view `which` (see dsv_mesh_view) into the caller's device buffer dst. */
int dsv_mesh_copy_view(const dsv_mesh* m, int32_t which, void* dst, void* stream);

/* Repartitions the matrix-free blocks (vertices_per_block, constraints as in
 * dsv_mesh_create; DSV_E_ARG leaves the previous partition in place), the lanes that share
 * one vertex's incident cells (lanes_per_vertex: 1, 2 or 4, with vertices_per_block *
 * lanes_per_vertex <= 1024; default 4) and sets the workgroup size (256 or 512) of the
 * matrix-free kernels; used by the autotuner.  No counterpart in the reference (launch-shape
 * tuning of this library). */
int dsv_mesh_set_partition(dsv_mesh* m, int32_t vertices_per_block, int32_t block_threads,
                           int32_t lanes_per_vertex);

/* --- assembly: GPUSparseAssembler.assemble (`gpu/gpu_assembler.py:472-532`) ---
 * values (device, nnz, pattern order) = A = (inv_dt + alpha_e) M_e + De_e K_e summed over
 * cells, lumped (M = dJ/24 I) when lumped != 0 else consistent, then Dirichlet rows
 * (row zeroed, diagonal 1; `gpu/gpu_assembler.py:147-165`).  Deterministic (no atomics).
 * With D and alpha both scalars (no arrays) the values are D K1 + (inv_dt + alpha) M1 entry by
 * entry, from K(D = 1) and the unit mass matrix of the lumping, kept on the mesh after the
 * first such call (nnz doubles each). */
int dsv_assemble(dsv_mesh* m, const dsv_species* sp, double inv_dt, int32_t lumped,
                 double* values, void* stream);

/* --- RHS: GPUSparseAssembler.assemble_rhs (`gpu/gpu_assembler.py:534-603`) ---
 * b = M (inv_dt u_old + f) with f per vertex (f_len == V), per cell (f_len == C, V != C) or
 * absent (f == NULL), then b[bc] = bc_value.  A per-cell f takes the reference's per-cell
 * branch (`:560-571`): vertex v gets f[p mod C], p = CuPy's searchsorted of v in the flat
 * caller connectivity (4C entries, not sorted): 4C if v > conn[4C-1], else the left bisection
 * over [0, 4C-1], each vertex searched on its own. */
int dsv_rhs(dsv_mesh* m, const double* u_old, const double* f, int32_t f_len, double inv_dt,
            int32_t lumped, const int32_t* bc_dofs, int32_t n_bc, double bc_value, double* b,
            void* stream);

/* dsv_rhs for n_species (1..3) species in one call (the reference calls assemble_rhs once per
 * species, `dissolve.py:787,895,946` via `numerics/solvers.py:348-353`): b[s] from u_old[s]
 * and f[s] (per vertex, length V; f or f[s] NULL = absent), with the Dirichlet list and value
 * of sp[s] (its other fields are not read).  The same values as dsv_rhs, bit for bit; the
 * lumped mass makes one pass for all species. */
int dsv_rhs_batch(dsv_mesh* m, int32_t n_species, const dsv_species* sp,
                  const double* const* u_old, const double* const* f, double inv_dt,
                  int32_t lumped, double* const* b, void* stream);

/* --- matrix-free apply: CeedDiffusionMassOp.apply (`gpu/ceed_operators.py:341-343`) ---
 * y_s = A_s x_s for n_species species sharing the mesh, the SAME operator as dsv_assemble
 * (including the Dirichlet identity rows).  precision: 64 -> FP64 arithmetic, 32 -> FP32
 * geometry and arithmetic (x, y stay FP64 at the boundary).  x[s], y[s] are device arrays. */
int dsv_apply(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
              int32_t lumped, const double* const* x, double* const* y, int32_t precision,
              void* stream);

/* --- solve: _solve_gpu_asm's KSP solve (`numerics/solvers.py:287-372`) ---
 * For each species s: solve A_s x_s = b_s with Jacobi-PCG to ||b - A x||_2 <= rtol ||b||_2
 * (FP64 true residual), starting from sp[s].x (Dirichlet entries are set to bc_value first).
 * All species advance in the same kernels. */
int dsv_solve(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
              int32_t lumped, int32_t method, double rtol, int32_t max_it, void* stream,
              dsv_stats* stats);

/* ||b_s - A_s x_s||_2 (rnorm, host array of n_species) and ||b_s||_2 (bnorm, may be NULL)
 * for x_s = sp[s].x with its Dirichlet entries replaced by bc_value -- the residual dsv_solve
 * starts from; FP64, matrix-free.  Used to compare candidate initial guesses. */
int dsv_residual_norms(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                       int32_t lumped, double* rnorm, double* bnorm, void* stream);

/* Galerkin warm start (no counterpart in the reference, whose solves start from the previous
 * solution, `numerics/solvers.py:329,366-367`).  For each species s, with x0 = sp[s].x
 * (Dirichlet entries set to bc_value) and the n_dir (1..3) directions d_k = dirs[k*n_species
 * + s] (device arrays, caller numbering, taken as 0 at Dirichlet vertices; NULL = zero),
 * computes x = x0 + sum_k c_k d_k with c minimising the A_s-energy norm of the error over
 * x0 + span{d_k}: (d_j . A d_k) c = d_j . (b - A x0), FP64, directions numerically dependent
 * on earlier ones dropped (c_k = 0).  sp[s].x is replaced by x only when its FP64 residual
 * 2-norm is below that of x0, so x0 stays the fallback.  coef (host, n_dir*n_species, may be
 * NULL) receives c (all 0 where x0 was kept, layout as dirs); r_before / r_after (host,
 * n_species, may be NULL) the residual norms of x0 and of the start left in sp[s].x. */
int dsv_warm_start(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t n_dir, const double* const* dirs, double* coef,
                   double* r_before, double* r_after, void* stream);

/* dsv_solve preceded by the Galerkin warm start of dsv_warm_start, folded into the solve (the
 * start's FP64 residual is the solve's first; no further synchronisation).  The directions are
 * the differences of past solutions -- hist[j*n_species + s] is species s at step n-1-j, the
 * start sp[s].x is step n, so d_0 = x - hist_0, d_j = hist_(j-1) - hist_j (n_hist of them) --
 * followed by n_extra explicit directions extra[k*n_species + s] (NULL = zero);
 * n_hist + n_extra <= 8.  While operator reuse is on (dsv_set_operator_reuse), a history
 * direction k >= 2 formed from the same two arrays as the previous call's direction k - 1,
 * with the same operator, takes its A d from that call (the reuse contract keeps the arrays
 * unchanged): a caller that shifts its history by one step pays one apply per call for d_0 and
 * one for d_1.  The combination is taken per species only where it lowers the FP64
 * residual (the start stays the fallback); coef (n_dir*n_species), r_before and r_after
 * (n_species) are as in dsv_warm_start and may be NULL.  Matrix-free methods only
 * (DSV_METHOD_MF32_IR64, DSV_METHOD_MF64).  No counterpart in the reference. */
int dsv_solve_warm(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t method, double rtol, int32_t max_it, int32_t n_hist,
                   const double* const* hist, int32_t n_extra, const double* const* extra,
                   void* stream, dsv_stats* stats, double* coef, double* r_before,
                   double* r_after);

/* dsv_solve_warm with the start taken from the history (no counterpart in the reference).
 * hist holds n_hist + 1 steps: hist[j*n_species + s] is species s at step n-j, the start is
 * hist_0 (step n), and the n_hist history directions are d_j = hist_j - hist_(j+1).  sp[s].x
 * receives the solution; its content on entry is not read, and it must not be one of the
 * history arrays.  Otherwise as dsv_solve_warm.  A caller that keeps its last solutions in
 * rotating buffers needs no copy of the start, and under operator reuse every history
 * direction but d_0 is the previous call's direction shifted by one (a direction formed
 * from an array that a solve has written since is not taken, and a solve without directions
 * on the mesh in between drops them all): one FP64 apply per call.
 * The start, the directions and so the results are the same bits as dsv_solve_warm's with
 * x = hist_0 and the history hist_1 ... */
int dsv_solve_warm_history(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                           int32_t lumped, int32_t method, double rtol, int32_t max_it,
                           int32_t n_hist, const double* const* hist, int32_t n_extra,
                           const double* const* extra, void* stream, dsv_stats* stats,
                           double* coef, double* r_before, double* r_after);

/* One lumped backward-Euler step in one call (no counterpart in the reference, whose species
 * blocks form each right-hand side with GPUSparseAssembler.assemble_rhs,
 * `gpu/gpu_assembler.py:534-603`, then solve): dsv_rhs_batch's right-hand sides from the start
 * hist_0 and the per-vertex sources f[s] (f or any f[s] NULL: none), written into sp[s].b, then
 * dsv_solve_warm_history.  The right-hand sides are formed while the solve packs its vectors,
 * with the same arithmetic, so b and the solution are the same bits as the two calls give.
 * lumped must be nonzero; sp[s].b must be none of hist, x or f. */
int dsv_step_warm_history(dsv_mesh* m, int32_t n_species, const dsv_species* sp,
                          const double* const* f, double inv_dt, int32_t lumped, int32_t method,
                          double rtol, int32_t max_it, int32_t n_hist, const double* const* hist,
                          int32_t n_extra, const double* const* extra, void* stream,
                          dsv_stats* stats, double* coef, double* r_before, double* r_after);

/* Low-mode subspace of an operator, for the Galerkin-started MF32_IR64 solves (no counterpart in
 * the reference, whose solves start from the previous solution; for callers that step one
 * operator many times).  For the operator of `species` (the arrays, inv_dt and lumping a later
 * dsv_solve_warm / dsv_solve_warm_history with DSV_METHOD_MF32_IR64 passes), builds n_modes
 * vectors per species (a multiple of 4, at most 64; 0 drops the subspace) approximating the
 * eigenvectors of the lowest eigenvalues of D^-1 A (D = diag A) on the non-Dirichlet vertices:
 * Chebyshev-filtered subspace iteration from a block of n_modes + 16 random vectors (`seed`),
 * `passes` passes of a degree-`degree` filter, Rayleigh-Ritz after each.  While operator reuse
 * keeps that operator (dsv_set_operator_reuse on, no epoch change, the same arrays), every such
 * solve then corrects the start its Galerkin warm start took by the Galerkin projection on the
 * subspace, x += W c with (W^T A W) c = W^T r, which minimises the A-norm of the start's error
 * over start + span W; a species whose start already meets rtol is left as it is, and the true
 * FP64 residual still decides convergence.  `ritz` (n_species x n_modes, nullable) receives the
 * last Rayleigh-Ritz values.  Device memory kept: 2 n_modes V n_species bytes (FP16); about three
 * times (n_modes + 16) V n_species * 4 bytes more during the build.  Needs operator reuse on. */
int dsv_build_low_modes(dsv_mesh* m, int32_t n_species, const dsv_species* species, double inv_dt,
                        int32_t lumped, int32_t n_modes, int32_t degree, int32_t passes,
                        uint64_t seed, double* ritz, void* stream);

/* Inner-solve controls of DSV_METHOD_MF32_IR64 (relative inner tolerance, max refinement
 * sweeps), the number of CG iterations per polled chunk, and whether chunks are replayed from
 * captured hipGraphs (use_graphs != 0) or launched directly.  Defaults: 1e-3, 12, 8, 0. */
int dsv_set_solver_options(dsv_mesh* m, double inner_rtol, int32_t max_outer,
                           int32_t graph_chunk, int32_t use_graphs);

/* Species narrowing of the batched matrix-free inner CG (DSV_METHOD_MF32_IR64, DSV_METHOD_MF64;
 * no counterpart in the reference, which solves each species on its own).  With enable != 0
 * (the default), an inner solve in which some species are not active, or in which some finish
 * while the others are predicted to need several more iterations, continues the remaining
 * species with the kernel of fewer species: it moves fewer bytes and reduces fewer partials per
 * iteration.  The copies between the two keep every species' iterates; only the order of
 * later partial sums may differ (the narrowed kernel may run another persistent grid), so the
 * solutions agree to rounding, not bit for bit. */
int dsv_set_species_narrowing(dsv_mesh* m, int32_t enable);

/* Operator reuse (no counterpart in the reference, which re-assembles every call).  With
 * reuse != 0, dsv_solve, dsv_apply, dsv_residual_norms and dsv_warm_start keep the operator
 * data they build (per-entry coefficients, mass diagonal, Jacobi diagonal, Dirichlet bits) and
 * skip rebuilding it when a later call names the same coefficient and Dirichlet arrays (same
 * pointers and lengths), the same scalar D / alpha values, inv_dt, mass lumping and operator
 * form.  Dirichlet index lists are range-checked once per (address, length) while reuse is on.
 * The caller guarantees that the contents of those arrays, and of the past solutions given to
 * dsv_solve_warm, do not change while reuse is on;
 * every call of this function discards what was built (call it again after changing a
 * coefficient array in place).  Default: off. */
int dsv_set_operator_reuse(dsv_mesh* m, int32_t reuse);

/* Operator form of the FP32 inner CG of DSV_METHOD_MF32_IR64 when every species has a scalar
 * D and the mass is lumped (otherwise the generic recomputing form is used):
 * DSV_FORM_STORED (default) applies D K(D=1) + diag((inv_dt + alpha) M) with K(D=1) stored
 * per graph edge in FP32; DSV_FORM_RECOMPUTE recomputes every element Jacobian.  No
 * counterpart in the reference (its operator choice is libCEED's). */
int dsv_set_operator_form(dsv_mesh* m, int32_t form);

/* Average device duration (ms) of one launch of the fused n_species matrix-free kernel: three
 * warm-up launches on `stream`, then n_rep launches captured in one hipGraph and replayed on
 * the library's stream of the mesh between two HIP events.  mode 0 times the plain apply
 * y = A x (x = sp[s].x, or the lumped mass when NULL); mode 1 times one fused CG iteration as
 * dsv_solve launches it (vector updates, w = A u and the partials of the next iteration); mode 2
 * times dsv_assemble's kernel for sp[0] (FP64 CSR values, no Dirichlet rows: sp[0].n_bc must be
 * 0) on `stream`.  The FP32 kernels use the operator form set by dsv_set_operator_form. */
int dsv_time_apply(dsv_mesh* m, int32_t n_species, const dsv_species* sp, double inv_dt,
                   int32_t lumped, int32_t precision, int32_t mode, int32_t n_rep, void* stream,
                   double* ms);

/* --- upstream coefficient fields of the transport step --- */

/* Nodal gradient of a P1 field u (device, V): per vertex, the average of the constant gradients
 * of its incident cells weighted by cell volume |det J| / 6 -- GPUSparseAssembler.nodal_gradient
 * (`gpu/gpu_assembler.py:251-307,646-671`).  Instead of scattering with atomics, the mesh keeps
 * an edge stencil (three FP64 weights per CSR entry and one per vertex, 24 nnz + 8 V bytes,
 * built on the first gradient call of any kind) and every call is one pass over the CSR rows,
 * bit-reproducible.  grad (V x 3, interleaved gx, gy, gz) and grad_norm
 * (sqrt(gx^2 + gy^2 + gz^2 + 1e-12), compute_norm_grad_phi, `physics/interface_velocity.py:65-87`)
 * may each be NULL, not both. */
int dsv_nodal_gradient(dsv_mesh* m, const double* u, double* grad, double* grad_norm,
                       void* stream);

typedef struct dsv_coef_params {
    double eps;         /* half-width of the regularised delta band */
    double k_orr;       /* ORR rate */
    double kd;          /* film-dissolution coefficient of the OH source */
    double Fmax;        /* film saturation */
    double film_floor;  /* porosity * constrictivity / tortuosity^2 */
} dsv_coef_params;

/* Per vertex (device arrays, length V; every output may be NULL):
 *   alpha_o2 = delta_eps(phi) |grad phi| k_orr                          (`dissolve.py:866-875`)
 *   f_oh = [phi <= 0] kd F Cl^2 + 4 delta_eps(phi) |grad phi| k_orr O2  (`dissolve.py:923-936`)
 *   film_factor = (1 - s) + s film_floor, s = clip(F / Fmax, 0, 1), and
 *   medium_factor = film_factor [phi <= 0]       (`physics/interface_velocity.py:90-155`: the
 *   effective diffusivities are D_Zn0 film_factor and D_i0 medium_factor for Cl, OH, O2)
 * with delta_eps(phi) = (1 + cos(pi phi / eps)) / (2 eps) for |phi| < eps, else 0.  F_film, Cl and
 * O2 may be NULL (taken as 0). */
int dsv_transport_coefficients(dsv_mesh* m, const double* phi, const double* grad_norm,
                               const double* F_film, const double* Cl, const double* O2,
                               const dsv_coef_params* params, double* alpha_o2, double* f_oh,
                               double* film_factor, double* medium_factor, void* stream);

/* compute_effective_diffusion (`physics/interface_velocity.py:90-155`), per vertex, with
 * film_factor and medium_factor as above (params: Fmax, film_floor; eps, k_orr, kd unused):
 *   DeZn = D0[0] film_factor,  DeCl = D0[1] medium_factor,  DeOH = D0[2] medium_factor,
 *   DeO2 = D0[3] medium_factor.
 * D0 is a host array of four (the reference's argument order Zn, Cl, OH, O2); F_film may be NULL
 * (taken as 0) and every output may be NULL.  Unlike the reference, it raises no bounds errors
 * itself: the Python wrapper checks the fields the reference checks. */
int dsv_effective_diffusion(dsv_mesh* m, const double* phi, const double* F_film,
                            const dsv_coef_params* params, const double* D0, double* DeZn,
                            double* DeCl, double* DeOH, double* DeO2, void* stream);

/* --- the per-step physics around the solves: Zn2+ coefficients, film, under-relaxation,
 * interface velocity, level-set advance and redistancing (device arrays of length V in the
 * caller's vertex numbering) --- */

typedef struct dsv_film_params {
    double kf;    /* film formation rate */
    double kd;    /* film dissolution coefficient */
    double Fmax;  /* film saturation */
} dsv_film_params;

/* Zn2+ coefficients of the Zn species block (`dissolve.py:740-778`), per vertex:
 *   alpha_zn = [phi <= 0] kf (1 - F / Fmax),  src_zn = [phi <= 0] kd F Cl^2
 * (F_film, Cl NULL -> 0; alpha_zn, src_zn may be NULL), and the scaffold Dirichlet vertices
 * {v : phi_v > 0} in ascending order into scaffold (device, capacity V) with their number in
 * *n_scaffold (host; the call then synchronises the stream).  scaffold and n_scaffold are both
 * NULL or both given. */
int dsv_zinc_coefficients(dsv_mesh* m, const double* phi, const double* F_film, const double* Cl,
                          const dsv_film_params* params, double* alpha_zn, double* src_zn,
                          int32_t* scaffold, int32_t* n_scaffold, void* stream);

/* The film update of the film block (`dissolve.py:797-833`), the lumped pointwise solution of
 * (1/dt + alpha) F = F_old / dt + src:
 *   F_new = clip((F_old inv_dt + src) / (inv_dt + alpha), 0, Fmax),
 *   alpha = [phi <= 0] (kf Zn / Fmax + kd Cl^2),  src = [phi <= 0] kf Zn.
 * F_new may alias F_old. */
int dsv_film_update(dsv_mesh* m, const double* phi, const double* Zn, const double* Cl,
                    const double* F_old, const dsv_film_params* params, double inv_dt,
                    double* F_new, void* stream);

/* u = omega u_new + (1 - omega) u_prev for n values, solve_transport's under-relaxation
 * (`numerics/solvers.py:278-279`); u may alias either input. */
int dsv_under_relax(int32_t n, const double* u_new, const double* u_prev, double omega,
                    double* u, void* stream);

typedef struct dsv_velocity_params {
    double h;         /* mean cell size: the interface band is |phi| < 3 h */
    double k_orr;     /* ORR rate */
    double zn_solid;  /* molar density of solid Zn, [Zn]_sol */
    double zn_sat;    /* saturation concentration [Zn]_sat (the Zn Dirichlet value) */
} dsv_velocity_params;

/* compute_interface_velocity_paper (`physics/interface_velocity.py:622-785`, the default branch:
 * nodal O2, no exact-ORR override, no probe): per vertex, inside |phi| < 3 h,
 *   v = min(max(v_Zn, v_ORR, v_O2), 0),  n = grad phi / sqrt(|grad phi|^2 + 1e-12),
 *   v_Zn = -max(DeZn, 0) |grad Zn . n| / (zn_solid - zn_sat)  (zn_solid if that is ~0),
 *   v_O2 = -2 max(DeO2, 0) |grad O2 . n| / zn_solid,  v_ORR = -(2 / zn_solid) k_orr max(O2, 0),
 * and 0 outside; the gradients are the nodal gradients of dsv_nodal_gradient.  Zn and DeZn NULL
 * (together) leave v_Zn out, DeO2 NULL leaves v_O2 out: with both, v = min(v_ORR, 0) in the band,
 * the ORR law alone (`physics/interface_velocity.py:741,749`). */
int dsv_interface_velocity(dsv_mesh* m, const double* phi, const double* Zn, const double* O2,
                           const double* DeZn, const double* DeO2,
                           const dsv_velocity_params* params, double* v, void* stream);

/* The algebraic level-set update phi = phi_old + v dt (`dissolve.py:684-690`); phi may alias
 * phi_old. */
int dsv_levelset_advance(dsv_mesh* m, const double* phi_old, const double* v, double dt,
                         double* phi, void* stream);

/* Mean over cells of the longest edge (DOLFINx mesh.h of a tetrahedron), the h_avg of
 * `domain/mesh_utils.py:559-575`; computed once per mesh, synchronous. */
int dsv_mesh_h_avg(dsv_mesh* m, double* h_avg);

/* Sussman redistancing, `redistance` (`domain/mesh_utils.py:450-542`), in place: num_iters
 * Jacobi pseudo-time iterations phi += dtau sign (1 - |grad phi|) outside the band
 * |phi0| < 3 h that stays frozen, dtau = 0.5 h / max(|grad phi|, 0.5),
 * sign = phi0 / sqrt(phi0^2 + (1.5 h)^2), phi0 the input; h <= 0 takes dsv_mesh_h_avg. */
int dsv_redistance(dsv_mesh* m, double* phi, int32_t num_iters, double h, void* stream);

/* Scaffold volume of the level set phi (per-vertex P1 field, device).  exact == 0:
 * GPUSparseAssembler.scaffold_volume (`gpu/gpu_assembler.py:604-643`), the integral of the
 * smoothed Heaviside H_eps(phi) = 0 / 1 / (1 + phi/eps + sin(pi phi/eps)/pi)/2 at the 4-point
 * degree-2 rule (barycentric (a, b, b, b), a = (5 + 3 sqrt 5)/20, b = (5 - sqrt 5)/20, weights
 * 1/4).  exact != 0: compute_scaffold_volume_exact (`domain/mesh_utils.py:608-630`), the exact
 * volume of {phi > 0} by marching tetrahedra (`analysis/mechanics_fe.py:45-82`, the cell
 * fraction clipped to [0, 1]).  frac (device, length C, may be NULL) receives the per-cell
 * quadrature mean of H_eps, resp. the cell's solid fraction; *volume (host) the sum over cells of
 * that value times |det J| / 6, in a fixed order.  Synchronises the stream. */
int dsv_scaffold_volume(dsv_mesh* m, const double* phi, int32_t exact, double eps, double* frac,
                        double* volume, void* stream);

/* --- learned warm start (no counterpart in the reference, whose solves start from the
 * previous solution): a per-vertex MLP 7 -> 64 -> 64 -> 1 with tanh activations, FP32
 * parameters packed as W1[64][7], b1[64], W2[64][64], b2[64], w3[64], b3.  Device pointers. */
#define DSV_MLP_NPARAM 4737
#define DSV_MLP_NFEAT 7

/* Features (X: n x 7 FP32) of one species at step n, per vertex (caller numbering): u_n,
 * (u_n - u_prev)/sigma, clamp(phi/eps, -4, 4)/4, min(|grad phi|, 4)/4, log(D)/2,
 * alpha dt (alpha NULL -> alpha_value), Dirichlet flag (bc_mask NULL -> 0).  sigma (one device
 * FP32) = max |u_n - u_prev|; the network's output times sigma is the predicted increment. */
int dsv_mlp_features(int32_t n, const double* u, const double* u_prev, const double* phi,
                     const double* grad_phi, const double* alpha, double alpha_value, double D,
                     double dt, double eps, const uint8_t* bc_mask, float* X, float* sigma,
                     void* stream);
/* out[i] = f(X[i]) for n samples */
int dsv_mlp_forward(const float* params, const float* X, int32_t n, float* out, void* stream);
/* The learned directions in FP64 from one forward pass over n_species blocks of V rows:
 * dirs[s][i] = out[s V + i] * (sigma[s] * tau), sigma the species' dsv_mlp_features scales
 * (device FP32) and tau the training targets' RMS; no counterpart in the reference. */
int dsv_mlp_directions(int32_t n_species, int32_t V, const float* out, const float* sigma,
                       double tau, double* const* dirs, void* stream);
/* Bytes of device workspace dsv_mlp_grad needs for n samples. */
size_t dsv_mlp_grad_workspace(int32_t n);
/* grad = dL/dparams and *loss = L of L = mean_i (f(X[i]) - y[i])^2 (loss may be NULL);
 * deterministic (fixed summation order, no atomics). */
int dsv_mlp_grad(const float* params, const float* X, const float* y, int32_t n, float* grad,
                 float* loss, void* workspace, size_t workspace_bytes, void* stream);
/* dsv_mlp_grad that allocates nothing: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving or training
 * loops, harnesses that forbid hipMalloc).  The partial gradients go to a static device array
 * of the library (4.85 MB per GPU, held until the process ends) instead of a workspace, so
 * n is limited to 16384 samples (DSV_E_ARG beyond that).  Results are the bits of dsv_mlp_grad
 * for the same inputs.  Calls that overlap on one GPU share that array and race. */
int dsv_mlp_grad_noalloc(const float* params, const float* X, const float* y, int32_t n,
                         float* grad, float* loss, void* stream);
/* One Adam step (step >= 1 counts from 1) on all DSV_MLP_NPARAM parameters. */
int dsv_mlp_adam(float* params, const float* grad, float* m, float* v, int32_t step, float lr,
                 float beta1, float beta2, float eps, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DSV_H */
