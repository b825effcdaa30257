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

/* Flags for emg_assemble_*. */
#define EMG_EXACT_FACE_MASS 1 /* use the exact face-face mass integral instead of
                                 the reference's coefficient (fem/mth/tet.py:1036) */
#define EMG_FORCE_TWO_PHASE 2 /* isotropic assembly through per-tet element blocks in
                                 HBM instead of the fused LDS kernel (same results) */

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

/* Same with full 3x3 complex tensors per tet: d_mm (nt,3,3) c128 multiplies
 * the mass term (eps_r), d_ms (nt,3,3) c128 the curl term (the reference
 * passes matinv(mu_r), fem/physics/edm/optimized_assembly.py:103-107).       */
int emg_assemble_tensor(const emg_pattern *pat, const double *d_nodes, const int32_t *d_tets,
                        const double *d_mm, const double *d_ms, int flags,
                        double *d_E, double *d_B, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EMERGE_AMD_H */
