/*
 * emerge_amd -- C ABI of the MI355X-native Nedelec-2 frequency-domain /
 * eigenmode hot path.
 *
 * Conventions
 *   - Every pointer named d_* is device memory (HBM) owned by the caller.
 *     Complex numbers are interleaved (re, im) doubles or floats.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls are asynchronous on that stream unless stated otherwise.
 *   - Return value: 0 on success, negative on error; emg_last_error() gives
 *     the message of the last failing call on the calling thread.
 *   - Handles (emg_pattern) are owned by the caller and freed with
 *     emg_pattern_destroy.  One handle may be used from one thread at a time.
 *
 * DOF numbering is the reference's [edge mode 1 | face mode 1 | edge mode 2 |
 * face mode 2] (fem/elements/nedelec2.py:46-62) over "entities" numbered
 * [edges | faces].  The global matrices are 2x2 block matrices whose four
 * blocks share one entity-adjacency pattern, so N = 2*n_ent and
 * nnz = 4*nnz_ent; columns of every CSR row are sorted ascending.
 */
#ifndef EMERGE_AMD_H
#define EMERGE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emg_pattern emg_pattern;

/* Message of the last failed call on this thread ("" if none). */
const char *emg_last_error(void);

/* Symbolic CSR pattern, built once per mesh on the GPU.
 * Replaces the index half of scipy's coo_matrix((data,(rows,cols))).tocsr()
 * in tet_mass_stiffness_matrices (fem/physics/edm/optimized_assembly.py:59-62)
 * and the rows/cols fill of _matrix_builder (:109-112).
 *   d_tet_ent: (nt, 10) int32, per canonical tet (vertices sorted by global
 *              id) its 6 edge entities in local order (0,1)(0,2)(0,3)(1,2)
 *              (1,3)(2,3) and its 4 face entities (0,1,2)(0,2,3)(0,1,3)(1,2,3).
 *   Synchronises `stream` before returning (sizes are read back).            */
int emg_pattern_build(const int32_t *d_tet_ent, int64_t nt, int64_t n_ent, void *stream, emg_pattern **out);

/* Sizes of a pattern: rows N = 2*n_ent, nnz = 4*nnz_ent. */
int emg_pattern_info(const emg_pattern *pat, int64_t *n_rows, int64_t *nnz, int64_t *nnz_ent, int64_t *n_incidences);

/* Write the global CSR index arrays: d_rowptr (N+1) and d_cols (nnz), int32. */
int emg_pattern_csr(const emg_pattern *pat, int32_t *d_rowptr, int32_t *d_cols, void *stream);

void emg_pattern_destroy(emg_pattern *pat);

/* The pattern's autotuned fused-assembly grid: the first fused emg_assemble_iso on a
 * pattern times 4, 8, 16, 32 and 64 workgroups per CU (HIP events, 2 runs each after one
 * untimed, so that call synchronises its stream) and keeps the fastest; every grid gives the
 * same values bit for bit.  h_wg_per_cu: the choice (0 before the first fused assembly);
 * h_ms[5]: milliseconds per assembly in that order.  EMG_ASM_WG_PER_CU overrides the search.
 * No reference counterpart (its element loop is a Numba prange).                       */
int emg_pattern_assembly_tuning(const emg_pattern *pat, int *h_wg_per_cu, float *h_ms);

/* Flags for emg_assemble_*. */
#define EMG_EXACT_FACE_MASS 1 /* use the exact face-face mass integral instead of
                                 the reference's coefficient (fem/mth/tet.py:1036) */
#define EMG_FORCE_TWO_PHASE 2 /* isotropic assembly through per-tet element blocks in
                                 HBM instead of the fused LDS kernel (same results) */
#define EMG_REFERENCE_GEOMETRY 4 /* isotropic fused path only: d_tets hold each tet in the
                                 reference's vertex order, and the vertex gradients and volume
                                 come from the reference's absolute-coordinate cofactor formulas
                                 (fem/mth/optimized.py:343-369, no fused multiply-adds), so the
                                 matrices carry the reference's own rounding; the default uses
                                 edge vectors relative to one vertex (more accurate away from
                                 the origin) */

/* Numeric assembly of the curl-curl matrix E (Dmat, with 1/mu) and the mass
 * matrix B (Fmat, with eps) into the pattern's CSR value arrays, complex128.
 * Replaces _matrix_builder + ned2_tet_stiff_mass + the value half of tocsr()
 * (fem/physics/edm/optimized_assembly.py:43-116, fem/mth/tet.py:854-1088).
 * Isotropic materials: one complex scalar per tet.
 *   d_nodes (nv,3) f64; d_tets (nt,4) int32 canonical (sorted);
 *   d_eps (nt) c128 (eps_r); d_nu (nt) c128 (1/mu_r);
 *   d_E, d_B (nnz) c128 outputs.                                             */
int emg_assemble_iso(const emg_pattern *pat, const double *d_nodes, const int32_t *d_tets,
                     const double *d_eps, const double *d_nu, int flags,
                     double *d_E, double *d_B, void *stream);

/* nbatch meshes that share the pattern's tets (same topology, own node coordinates and
 * materials: geometry variants, parameter studies) assembled in one pair of fused launches:
 * mesh b's nodes at d_nodes + 3 nv b (nv x 3 float64 each), eps and nu at + nt b (complex128),
 * E and B values at + nnz b.  Each mesh's values equal those of emg_assemble_iso on it.  No
 * reference counterpart (the reference assembles one mesh per call); flags:
 * EMG_EXACT_FACE_MASS, EMG_REFERENCE_GEOMETRY.                                            */
int emg_assemble_iso_batched(const emg_pattern *pat, int nbatch, int64_t nv, const double *d_nodes,
                             const int32_t *d_tets, const double *d_eps, const double *d_nu, int flags, double *d_E,
                             double *d_B, void *stream);

/* Same with full 3x3 complex tensors per tet: d_mm (nt,3,3) c128 multiplies
 * the mass term (eps_r), d_ms (nt,3,3) c128 the curl term (the reference
 * passes matinv(mu_r), fem/physics/edm/optimized_assembly.py:103-107).       */
int emg_assemble_tensor(const emg_pattern *pat, const double *d_nodes, const int32_t *d_tets,
                        const double *d_mm, const double *d_ms, int flags,
                        double *d_E, double *d_B, void *stream);

/* ------------------------------------------------------------------ solver
 * Right-preconditioned restarted GMRES for K x = b on a pattern's CSR,
 * replacing SolveRoutine.solve (fem/solver.py:405-469).  PEC DOFs (mask 0)
 * are eliminated implicitly (x = 0 there, as solver.py:429,465 leaves them).
 * precision 64: complex128 GMRES.  precision 32: complex64 GMRES inside FP64
 * iterative refinement; the returned residual is always the FP64
 * ||b - K x|| / ||b|| over the free DOFs.  Preconditioner: inverse 2x2 blocks
 * coupling the two modes of every edge / face (entity e <-> e + N/2).       */
typedef struct emg_solver emg_solver;
typedef struct {
    int iterations;   /* Krylov iterations over all restarts / refinements */
    int refinements;  /* FP64 refinement steps (precision 32) */
    int converged;    /* residual <= tol */
    int reserved;
    double residual;  /* FP64 relative residual */
} emg_solve_info;

int emg_solver_create(int64_t n, int64_t nnz, const int32_t *d_rowptr, const int32_t *d_cols, int restart,
                      int precision, emg_solver **out);
/* K values (nnz complex128, caller-owned, must outlive the solves) and the
 * free-DOF mask (n uint8, 1 = solve, 0 = PEC; NULL = all free). */
int emg_solver_set_matrix(emg_solver *s, const double *d_vals, const uint8_t *d_mask, void *stream);
/* Gradient correction of the smoother (no counterpart in the reference, whose
 * direct solvers need none): G (n x np2, CSR, at most 8 entries per row) maps
 * second-order Lagrange nodes to the Nedelec DOFs (the null space of the
 * curl-curl matrix), G^T (np2 x n) is its transpose; values float64,
 * caller-owned, must outlive the solver; columns touching masked DOFs must be
 * left out.  From the next emg_solver_set_matrix on, the preconditioner is
 * z = BJ(r) + G diag(G^T K G)^{-1} G^T r.  np2 = 0 switches it off.          */
int emg_solver_set_gradient(emg_solver *s, int64_t np2, const int32_t *d_g_rowptr, const int32_t *d_g_cols,
                            const double *d_g_vals, const int32_t *d_gt_rowptr, const int32_t *d_gt_cols,
                            const double *d_gt_vals, void *stream);
/* Learned smoother weights (no counterpart in the reference): complex128 w_e
 * per entity (n/2) multiplies that entity's 2x2 block inverse, w_p per P2
 * node (np2) its entry of diag(G^T K G)^{-1}; either may be NULL.  Pointers are
 * kept (caller-owned) and applied by every following emg_solver_set_matrix.  */
int emg_solver_set_weights(emg_solver *s, const double *d_we, const double *d_wp);
/* The smoother's two parts for r (complex128): d_u = BJ(r) (n) and, when the
 * gradient correction is set and d_t is not NULL, d_t = diag(G^T K G)^{-1}
 * G^T r (np2), with the tables of the last emg_solver_set_matrix.           */
int emg_solver_smoother_parts(emg_solver *s, const double *d_r, double *d_u, double *d_t, void *stream);
/* Polynomial acceleration of the smoother (no counterpart in the reference,
 * whose direct solvers need none): with degree d > 1 the preconditioner
 * becomes M p(K M), 1 - z p(z) = prod_i (1 - z c_i), for the d complex128
 * coefficients c_i (h_coef: 2 d doubles, host, copied; applied in order,
 * normally the reciprocals of Leja-ordered harmonic Ritz values of K M).
 * d = 1 switches it off.  Each preconditioner application then costs d - 1
 * extra SpMVs and smoother applications.                                    */
int emg_solver_set_poly(emg_solver *s, int degree, const double *h_coef);
/* d (<= 16) steps of Arnoldi on K M (complex128, masked K, the smoother of the
 * last emg_solver_set_matrix) from d_v0 (n complex128); h_out (host) receives
 * the (d + 1) x d Hessenberg matrix column-major as complex128.  Uses the
 * solver's Krylov storage: not during a solve.  Synchronises `stream`.       */
int emg_solver_arnoldi_probe(emg_solver *s, int d, const double *d_v0, double *h_out, void *stream);
/* y = mask.(K (mask.x)) as the solver's Arnoldi steps form it, with the matrix of
 * the last emg_solver_set_matrix: precision 64 complex128, precision 32 the
 * solver's complex64 copy (x, y complex64).  No counterpart in the reference
 * (scipy's CSR product inside its solvers); exported to time the product. */
int emg_solver_apply(emg_solver *s, int precision, const void *d_x, void *d_y, void *stream);

/* The solver's autotuned SpMV shape: on its first emg_solver_set_matrix the solver times
 * the blocked SpMV with 4, 8, 16 and 32 lanes per entity (HIP events, 5 launches each, in
 * complex128 and, for precision 32, complex64) and keeps the fastest.  h_lanes[2]: the
 * chosen lanes (complex128, complex64); h_ms[8]: the measured milliseconds per launch
 * (4 per precision, in the order 4, 8, 16, 32).  EMG_SPMV_BLK_LANES overrides the search.
 * No reference counterpart (PARDISO has no shape to tune).                              */
int emg_solver_spmv_tuning(emg_solver *s, int *h_lanes, float *h_ms);

/* The solver's autotuned smoother shape: with the gradient correction set, the first
 * emg_solver_set_matrix also times the smoother's G^T product with 4, 8 and 16 lanes per
 * row (HIP events, 5 launches each, in the precision of the solver's steps) and keeps the
 * fastest.  h_lanes[1]: the choice (0 before tuning or without a gradient: 8 is used);
 * h_ms[3]: milliseconds per launch for 4, 8, 16.  EMG_GT_LANES overrides the search.
 * No reference counterpart (its solvers are direct).                                     */
int emg_solver_smoother_tuning(emg_solver *s, int *h_lanes, float *h_ms);
/* d_x: initial guess in, solution out (n complex128). Synchronises `stream`. */
int emg_solver_solve(emg_solver *s, const double *d_b, double *d_x, double tol, int maxiter, int max_refine,
                     emg_solve_info *info, void *stream);
void emg_solver_destroy(emg_solver *s);

/* K = E - s*B on the pattern (complex128 values, nnz each). */
int emg_affine_combine(int64_t nnz, const double *d_E, const double *d_B, double s, double *d_K, void *stream);
/* d_pos[i] = CSR position of (d_qr[i], d_qc[i]) or -1 (binary search in sorted rows). */
int emg_csr_find(int64_t m, const int32_t *d_rowptr, const int32_t *d_cols, const int32_t *d_qr, const int32_t *d_qc,
                 int32_t *d_pos, void *stream);
/* d_K[d_pos[i]] += (gre + j gim) * d_v[i] for i < m (complex128, atomic). */
int emg_scatter_add(int64_t m, const int32_t *d_pos, const double *d_v, double gre, double gim, double *d_K, void *stream);

/* complex128 vector primitives (used by the shift-invert Arnoldi driver that
 * replaces scipy eigs/eigsh, fem/solver.py:343-357 and the deprecated 3D call
 * at fem/physics/edm/emfreq3d.py:838-848).  V holds nv vectors of n entries
 * back to back; d_work is scratch of ceil(nv/8)*1024*16 + 1024 doubles.    */
int emg_spmv(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_vals, const uint8_t *d_mask,
             const double *d_x, double *d_y, void *stream);
/* The same product for a matrix on an entity pattern (the CSR of
 * emg_pattern_csr: rows (0, e) and (1, e) share their column entities, the
 * mode-1 columns of a row follow its mode-0 columns): both rows of an entity in
 * one pass, each column index read once per four entries.  nnz = rowptr[n].
 * No counterpart in the reference (scipy's CSR product).                     */
int emg_spmv_blocked(int64_t n, int64_t nnz, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_vals,
                     const uint8_t *d_mask, const double *d_x, double *d_y, void *stream);
int emg_mdotc(int64_t n, int nv, const double *d_V, const double *d_w, double *d_h, double *d_work, void *stream);
int emg_maxpy(int64_t n, int nv, const double *d_V, const double *d_h, double *d_w, double *d_norm, double *d_work,
              void *stream);
int emg_mcomb(int64_t n, int nv, const double *d_V, const double *d_y, double *d_x, void *stream);
/* y (+)= A x, A real CSR (float64 values), x and y complex128. */
int emg_csr_real_apply(int64_t nrows, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_vals,
                       const double *d_x, double *d_y, int accumulate, void *stream);

/* ------------------------------------------------------------------ boundary terms
 * Robin (port / absorbing boundary) triangle terms, replacing
 * ned2_tri_stiff_force and ned2_tri_stiff (fem/mth/tri.py:575-796) as called
 * from assemble_robin_bc_excited / assemble_robin_bc
 * (fem/physics/edm/assembler.py:100-144).
 *   d_lxy  (ntri, 6) f64: local x0,x1,x2,y0,y1,y2 of each triangle (vertices in
 *          ascending global id), in the frame the reference uses;
 *   d_elen (ntri, 3) f64 edge lengths of edges (0,1),(1,2),(0,2), or NULL to use
 *          the local distances (force variant);
 *   d_uinc (ntri, nq, 2) c128 incident field (x, y local) at the nq points of
 *          d_dpts (4, nq: weights, then barycentric rows), or NULL;
 *   d_B    (ntri, 64) f64: the 8x8 block for gamma = 1;
 *   d_b    (ntri, 8) c128 source vector, or NULL.                             */
int emg_robin_blocks(int64_t ntri, const double *d_lxy, const double *d_elen, const double *d_uinc, int nq,
                     const double *d_dpts, double *d_B, double *d_b, void *stream);

/* Nedelec-2 field of nsol solution vectors (d_x: nsol x N complex128) at npts
 * points given by tet id (d_ptet) and barycentric coordinates (d_bary, npts x 4,
 * canonical vertex order).  d_E: nsol x npts x 3 complex128.  Replaces
 * ned2_tet_interp (fem/mth/tet.py:371-497).                                   */
int emg_interp_points(int64_t npts, const int32_t *d_ptet, const double *d_bary, const double *d_nodes,
                      const int32_t *d_tets, const int32_t *d_tet_ent, int64_t n_ent, int nsol,
                      const double *d_x, double *d_E, void *stream);

/* curl of the Nedelec-2 field of nsol solutions at tet-tagged points (arguments as
 * emg_interp_points), times d_cscale[tet] (nt complex128, NULL = 1): d_H nsol x npts x 3
 * complex128.  Replaces ned2_tet_interp_curl (fem/mth/tet.py:499-626), used by
 * EMDataSet.interpolate for H (fem/physics/edm/emdata.py:181-199).  Points with tet
 * -1 (outside the mesh) get zero, as in the reference (and in emg_interp_points).  */
int emg_interp_curl_points(int64_t npts, const int32_t *d_ptet, const double *d_bary, const double *d_nodes,
                           const int32_t *d_tets, const int32_t *d_tet_ent, int64_t n_ent, int nsol,
                           const double *d_x, const double *d_cscale, double *d_H, void *stream);
/* Tet holding each of npts points (d_pts npts x 3) and its barycentric coordinates
 * (d_bary npts x 4, canonical vertex order), by a uniform grid of cells: d_grid = origin
 * (3), inverse cell size (3), cell counts (3, as doubles); d_cell_ptr (ncells+1) and
 * d_cell_tets list the tets overlapping each cell in ascending id.  The last tet that
 * holds a point wins, with the reference's tolerances (fem/mth/tet.py:423-426); -1 if
 * none.  Replaces the brute-force search over all tets in ned2_tet_interp.          */
int emg_locate_points(int64_t npts, const double *d_pts, const double *d_grid, const int32_t *d_cell_ptr,
                      const int32_t *d_cell_tets, const double *d_nodes, const int32_t *d_tets, int32_t *d_tet,
                      double *d_bary, void *stream);

/* Stratton-Chu far field (fem/physics/edm/sc.py:27-180): ns surface samples at d_pts
 * (ns x 3) with area-weighted normals d_wn (ns x 3) carrying E and H (ns x 3 complex128);
 * for ndir directions r = (cos t cos p, cos t sin p, sin t) (d_theta, d_phi, radians):
 *   E(r) = -j k0/(4 pi) sum_j r x [(w_j x E_j - Z0 r x (w_j x H_j)) exp(j k0 r.v_j)],
 *   H(r) = r x E(r) / Z0; d_Eout, d_Hout ndir x 3 complex128.  FP64 (the reference
 * evaluates the same sum in complex64).                                          */
int emg_farfield(int64_t ns, const double *d_pts, const double *d_wn, const double *d_E, const double *d_H,
                 int64_t ndir, const double *d_theta, const double *d_phi, double k0, double *d_Eout,
                 double *d_Hout, void *stream);

/* ---- mesh topology on the GPU (fem/mesh3d.py:224-355, fem/elements/nedelec2.py:32-70) ---- */
typedef struct emg_topology emg_topology;

/* Topology of nt tets over nv vertices (d_tets: 4 x nt int64, row k = the tets' k-th vertex, any
 * vertex order; d_nodes: 3 x nv float64 or NULL, for the edge lengths): unique edges and faces
 * numbered in lexicographic order of their sorted vertex ids, by stable radix sorts of 64-bit
 * keys.  Replaces Mesh3D.update's set/dict loops (mesh3d.py:252-336) and the entity numbering
 * of Nedelec2 (nedelec2.py:46-62).  Fails (-1) on a vertex id out of range or repeated in a
 * tet, and on a triangle shared by more than two tets; nv < 2^21.                        */
int emg_topology_build(int64_t nv, int64_t nt, const int64_t *d_tets, const double *d_nodes, emg_topology **out,
                       void *stream);
int emg_topology_sizes(const emg_topology *h, int64_t *ne, int64_t *nf);
/* Copies (each destination optional, NULL to skip): edges 2 x ne, tris 3 x nf (int64, sorted
 * vertex ids); tets_canonical nt x 4 and tet_ent nt x 10 (int32: 6 edges in the local order
 * (0,1)(0,2)(0,3)(1,2)(1,3)(2,3), then ne + 4 faces (0,1,2)(0,2,3)(0,1,3)(1,2,3) of the sorted
 * tet); tet_to_edge and its sign 6 x nt and tet_to_tri and its sign 4 x nt in the reference's
 * local orders on the tet's own vertex order (mesh3d.py:292-305, the sign of _tri_ordering as
 * written); tri_to_tet 2 x nf (tets in ascending order, -1 padded); tri_to_edge 3 x nf (edges
 * (i1,i2), (i2,i3), (i1,i3)); edge_lengths ne float64.                                      */
int emg_topology_get(const emg_topology *h, int64_t *d_edges, int64_t *d_tris, int32_t *d_tets_canonical,
                     int32_t *d_tet_ent, int64_t *d_tet_to_edge, int64_t *d_tet_to_edge_sign, int64_t *d_tet_to_tri,
                     int64_t *d_tet_to_tri_sign, int64_t *d_tri_to_tet, int64_t *d_tri_to_edge,
                     double *d_edge_lengths, void *stream);
void emg_topology_destroy(emg_topology *h);

/* ---- port-face modal analysis (fem/physics/edm/emfreq3d.py:201-364) ---- */
#define EMG_PORT_FACE_BLOCK 1 /* fill the face-face entries of Dtt, which the reference's
                                 generalized_matrix_GQ leaves zero (nedeleclegrange2.py:350-358) */

/* Port element matrices and their dense assembly.  Replaces generalized_matrix_GQ and
 * _matrix_builder (fem/physics/edm/nedeleclegrange2.py:224-418) with the NedelecLegrange2
 * DOF order of fem/elements/nedleg2.py:54-61.  ntri triangles, d_verts (ntri x 2 x 3: x0 x1
 * x2 y0 y1 y2 in the face frame, ascending vertex numbers), d_Ms = ur^-1 and d_Mm = er per
 * triangle (ntri x 3 x 3 complex128, row-major), k0 in rad/m.  With d_A/d_B (n x n complex128,
 * column-major, zeroed by the caller) every entry is added at (d_t2r[t,i], d_t2r[t,j]) of
 * d_t2r (ntri x 14 int64: reduced DOF or -1 for a removed DOF); with d_Ael/d_Bel the
 * element matrices are also written (ntri x 14 x 14, row-major).                       */
int emg_port_assemble(int64_t ntri, const double *d_verts, const double *d_Ms, const double *d_Mm, double k0,
                      int flags, const int64_t *d_t2r, int64_t n, double *d_A, double *d_B, double *d_Ael,
                      double *d_Bel, void *stream);

/* The face field at np points d_pts (np x 2, face frame): ned2_tri_interp_full and
 * ned2_tri_interp_curl (fem/mth/tri.py:228-401), the last triangle holding a point wins,
 * points outside every triangle get zeros.  d_tris (ntri x 3, entries < nnodes), d_t2f
 * (ntri x 14, entries < nfield) int64, d_nodes (nnodes x 2) in the face frame, d_field
 * (nfield) complex128; a point whose triangle has an index outside those sizes gets NaN.  d_E (np x 3 complex128) =
 * d_basis (3 x 3 row-major, local -> global) applied to (Ex, Ey, Ez); with d_H, d_H =
 * hconst * basis (diag(d_diadic[t]) * (dEz/dy + j beta Ey, dEz/dx - j beta Ex, curl_z Et)),
 * d_diadic ntri x 3 x 3 complex128 indexed by the triangle number as the reference does
 * (NedelecLegrange2.interpolate_Hf, fem/elements/nedleg2.py:92-105).                   */
int emg_port_field(int64_t np, const double *d_pts, int64_t ntri, const int64_t *d_tris, int64_t nnodes,
                   const double *d_nodes, const int64_t *d_t2f, int64_t nfield, const double *d_field, double beta,
                   const double *d_diadic,
                   const double *d_basis, double hconst_re, double hconst_im, double *d_E, double *d_H,
                   void *stream);

/* d_M = d_A - sigma d_B (n x n complex128, column-major; sigma = sigma_re + j sigma_im) and
 * its rocSOLVER LU in place (d_ipiv: n int32).  Replaces the dense LAPACK factorisation
 * inside scipy's eig (SolverLAPACK.eig, fem/solver.py:311-341); fails if the shifted
 * matrix is exactly singular.                                                         */
int emg_port_factor(int64_t n, const double *d_A, const double *d_B, double sigma_re, double sigma_im, double *d_M,
                    int *d_ipiv, void *stream);

/* d_Y = (A - sigma B)^{-1} B d_X for k columns (n x k complex128, column-major) from the LU
 * of emg_port_factor: one step of inverse iteration on k vectors.                     */
int emg_port_inverse_step(int64_t n, const double *d_M, const int *d_ipiv, const double *d_B, int64_t k,
                          const double *d_X, double *d_Y, void *stream);

/* d_T = (A - sigma B)^{-1} d_B (n x n complex128, column-major) from the LU of
 * emg_port_factor: the shift-invert operator of the port pencil, formed once (one zgetrs with
 * n right-hand sides) so that each Arnoldi step is one matrix-vector product.           */
int emg_port_operator(int64_t n, const double *d_M, const int *d_ipiv, const double *d_B, double *d_T,
                      void *stream);

/* m steps of Arnoldi on d_T (emg_port_operator): d_V (n x (m+1)
 * complex128, column 0 the start vector, normalised here) receives the orthonormal basis,
 * d_H ((m+1) x m complex128, column-major) the Hessenberg matrix; a breakdown leaves zero
 * columns and a zero subdiagonal entry.  Classical Gram-Schmidt, two passes.             */
int emg_port_krylov(int64_t n, const double *d_T, int m, double *d_V, double *d_H, void *stream);

/* C = alpha op(A) op(B) + beta C, complex128 column-major, op 0 = N, 1 = T, 2 = C
 * (rocBLAS zgemm): the Rayleigh-Ritz projections and Ritz vectors of the port solve.   */
int emg_zgemm(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha_re, double alpha_im,
              const double *d_A, int64_t lda, const double *d_B, int64_t ldb, double beta_re, double beta_im,
              double *d_C, int64_t ldc, void *stream);

/* ------------------------------------------------------------------ batched sweep
 * A block of nvec (frequency, port) pairs over the shared pattern; replaces the
 * per-(frequency, port) assemble + solve loop of frequency_domain
 * (fem/physics/edm/emfreq3d.py:658-694, assembler.py:312-388) where the sweep
 * driver certifies trial solutions.  Multi-vectors are interleaved complex128:
 * X[row * nvec + j].  nvec is 1, 2, 4, 8 or 16.
 *
 * emg_sweep_residual:
 *   r_j = mask.(c_j Q[qidx_j] - (E - s_j B) x_j - Zp_j),  d_norms[2j] = ||r_j||^2,
 *   d_norms[2j+1] = ||mask.c_j Q[qidx_j]||^2.  x_j must vanish on masked DOFs.
 *   d_s (nvec) f64 = k0^2; d_Q (nq, n) c128 port sources for force scale 1;
 *   d_c (nvec) c128 force scales; d_prow_of_row (n) int32 index into Zp or -1
 *   (NULL together with d_Zp when there are no Robin rows); d_R (n, nvec) or
 *   NULL; d_work >= 4096 * nvec * 2 doubles.  E and B must be on an entity
 *   pattern (emg_pattern_csr), n = 2 * entities: both rows of an entity are
 *   formed together.                                                           */
int emg_sweep_residual(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_E,
                       const double *d_B, const uint8_t *d_mask, int nvec, const double *d_s, const double *d_X,
                       const double *d_Q, const int32_t *d_qidx, const double *d_c, const int32_t *d_prow_of_row,
                       const double *d_Zp, double *d_R, double *d_norms, double *d_work, void *stream);
/* The same residual with no workspace argument and no allocation: the
 * per-block partial sums live in the library's static 1 MiB device array
 * (one per GPU, held until the process ends).  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP stream
 * capture, tight serving loops).  Limits as emg_sweep_residual (nvec <= 16);
 * larger requests are refused.  Calls that overlap on one GPU race on the
 * array: serialise them (one stream, or events).                             */
int emg_sweep_residual_noalloc(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const double *d_E,
                               const double *d_B, const uint8_t *d_mask, int nvec, const double *d_s,
                               const double *d_X, const double *d_Q, const int32_t *d_qidx, const double *d_c,
                               const int32_t *d_prow_of_row, const double *d_Zp, double *d_R, double *d_norms,
                               void *stream);
/* Robin rows: Zp[i, j] = sum_e g[term_e, j] * val_e * X[col_e, j] over the
 * entries d_ptr[i] .. d_ptr[i+1] of port row i (B_p(f) = gamma_p(f) P_p,
 * assembler.py:382-383).  d_g (nterm, nvec) c128.                              */
int emg_port_apply(int64_t nprow, int nvec, const int32_t *d_ptr, const int32_t *d_pcol, const int32_t *d_pterm,
                   const double *d_pval, const double *d_g, const double *d_X, double *d_Zp, void *stream);
/* X[row, j] = sum_i V[i, row] Y[j, i]: d_V (m, n) c128 basis, d_Y (nvec, m) c128,
 * nvec * m <= 2048.                                                            */
int emg_basis_combine(int64_t n, int m, const double *d_V, int nvec, const double *d_Y, double *d_X, void *stream);

/* ------------------------------------------------------------------ numbering maps
 * The reference numbers edges and faces in Python set order
 * (fem/mesh3d.py:252-271); these move matrices and vectors between a caller's
 * numbering and the library's on the GPU, so tet_mass_stiffness_matrices
 * (fem/physics/edm/optimized_assembly.py:43-64), assemble_freq_matrix
 * (fem/physics/edm/assembler.py:312-388) and SolveRoutine.solve
 * (fem/solver.py:405-469) keep the caller's numbering.  d_p (n int32): the
 * library index of caller DOF i; d_pinv: its inverse.
 *
 * emg_csr_permute: caller-order CSR of a matrix on the library CSR (d_rowptr,
 *   d_cols, n rows): d_rowptr_out (n+1), d_cols_out (nnz, sorted per row) and
 *   d_map (nnz): the library entry of every caller entry.  Synchronises.
 * emg_csr_map: d_pos (nnz of the caller CSR) = library position of every entry
 *   of a caller-order CSR (d_rowptr_c, d_cols_c), -1 where the pattern lacks it.
 * emg_gather_c128:  d_dst[k] = d_src[d_idx[k]] (0 where d_idx[k] is outside [0, n_src)).
 * emg_scatter_c128: d_dst[d_idx[k]] = d_src[k] for d_idx[k] in [0, n_dst); nonzero
 *   values with no target are counted into *d_dropped (uint64, may be NULL).   */
int emg_csr_permute(int64_t n, const int32_t *d_rowptr, const int32_t *d_cols, const int32_t *d_p,
                    const int32_t *d_pinv, int32_t *d_rowptr_out, int32_t *d_cols_out, int32_t *d_map, void *stream);
int emg_csr_map(int64_t n, const int32_t *d_rowptr_c, const int32_t *d_cols_c, const int32_t *d_p,
                const int32_t *d_rowptr, const int32_t *d_cols, int32_t *d_pos, void *stream);
int emg_gather_c128(int64_t n, const int32_t *d_idx, int64_t n_src, const double *d_src, double *d_dst, void *stream);
int emg_scatter_c128(int64_t n, const int32_t *d_idx, int64_t n_dst, const double *d_src, double *d_dst,
                     unsigned long long *d_dropped, void *stream);

/* ------------------------------------------------------------------ learned smoother
 * Kernels of the graph network that predicts the smoother weights from mesh
 * features (no counterpart in the reference).  Activations float32 row-major,
 * layer weights float16 (k x m row-major), float32 accumulation.
 *   gemm:      C[n x m] = A[n x k] W + bias + beta C          (k <= 128)
 *   gemm_nt:   out[n x k] = D[n x m] W^T + beta out           (m <= 128)
 *   gemm_tn:   dW[k x m] += A^T D, dbias[m] += column sums of D (either NULL-able bias)
 *   agg:       Y[v] = rs[v] sum_{u in N(v)} cs[u] X[u]  over a vertex CSR (rs, cs NULL = 1)
 *   relu:      forward H = relu(P) + Hprev (d_dH NULL), backward dP = dH . (P > 0)
 *   readout:   forward O[e, c] = bias[c] + sum_q Hp[v_q(e), off + q C + c] (K <= 3 slots),
 *              backward (d_dO given): dHp and dbias accumulate (atomics)
 *   weights:   forward (d_dw NULL) w = 1 + O[., 0] + j O[., 1] (mode 0), exp(O[., 0]) (mode 1) or
 *              exp(0.5 tanh(O[., 0])) (mode 2),
 *              complex128; backward dO from d(Re w, Im w) (float32 pairs)
 *   weight_grad: dO[i] += 2 (Re, Im) sum_s scale_s conj(a_s[i]) b_s[i] (+ DOF i + stride)   */
int emg_gnn_gemm(int64_t n, int k, int m, const float *d_A, const void *d_W, const float *d_bias, float beta,
                 float *d_C, void *stream);
int emg_gnn_gemm_nt(int64_t n, int m, int k, const float *d_D, const void *d_W, float beta, float *d_out, void *stream);
int emg_gnn_gemm_tn(int64_t n, int k, int m, const float *d_A, const float *d_D, float *d_dW, float *d_dbias,
                    void *stream);
int emg_gnn_agg(int64_t nv, int C, const int32_t *d_ptr, const int32_t *d_idx, const float *d_rs, const float *d_cs,
                const float *d_X, float *d_Y, void *stream);
int emg_gnn_relu(int64_t n, const float *d_P, const float *d_Hprev, float *d_H, const float *d_dH, float *d_dP,
                 void *stream);
int emg_gnn_readout(int64_t ne, int K, int C, const int32_t *d_ev, const float *d_Hp, int ldh, int off,
                    const float *d_bias, float *d_O, const float *d_dO, float *d_dHp, float *d_dbias, void *stream);
int emg_gnn_weights(int64_t n, int mode, const float *d_O, double *d_w, const float *d_dw, float *d_dO,
                    void *stream);
int emg_gnn_weight_grad(int64_t n, int ns, int64_t ld, int64_t stride, const double *d_a, const double *d_b,
                        const double *d_scale, float *d_dO, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EMERGE_AMD_H */
