/*
 * gridfoam_hip.h -- C ABI of the MI355X (gfx950) pressure-velocity path.
 *
 * The reference (ricosjp/gridfoam) has no FFI: its hot path sits behind
 * Python plugin interfaces.  Every entry point below names the reference
 * interface it replaces (file:line, relative to the reference checkout).
 *
 * Conventions
 *   - every function is extern "C", takes device pointers, element counts
 *     and a hipStream_t, and returns 0 on success or a negative code;
 *     gf_last_error() returns a thread-local message for the last failure;
 *   - time-step calls allocate no device memory; only the set-up helpers
 *     (gf_octree_build, gf_exclusive_scan_i32, gf_fill_nbr_lists,
 *     gf_hang_cell_lists) allocate temporary buffers, once per grid;
 *   - fields are row-major: a cell field with nc components is [C][nc];
 *   - all FV assembly is FP64; Krylov solves run in FP64 or in FP32 inside
 *     an FP64 iterative-refinement loop.
 */
#ifndef GRIDFOAM_HIP_H
#define GRIDFOAM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *gf_stream_t; /* == hipStream_t */

#define GF_OK 0
#define GF_EINVAL (-1)
#define GF_EHIP (-2)
#define GF_ENOMEM (-3)

/* Device-resident face-neighbour operator.  Filled by the host from
 * buffers it owns (gridfoam_amd/grid.py); passed by pointer to every
 * call.  Replaces the index/geometry tensors of GridBase
 * (src/gridfoam/core/grid/base.py:19-275) and the cached FaceGeometry
 * (src/gridfoam/fv/kernels/face_geometry.py:20-72).
 *
 * Face-order contract: internal faces are grouped by owner in ascending
 * owner order (own_row[c] .. own_row[c+1] are cell c's owned faces).  The
 * neighbour lies on the +axis side of the owner.  Boundary faces may come
 * in any order (per-cell lists bnd_row / bnd_face).  Faces in another order (fluxel's order is not
 * specified) are brought into this order by a stable sort by owner, e.g.
 * gridfoam_amd.grid.canonical_face_order; gf_mesh_validate checks it. */
typedef struct gf_mesh {
    int32_t n_cells, n_faces, n_bnd, n_hang, n_hang_cells, n_slices;
    int64_t nnz_pad; /* padded entries of the sliced-ELL operator */
    /* topology */
    const int32_t *owner, *neighbour, *axis; /* [F]  neighbour on +axis side */
    const int32_t *bnd_owner, *bnd_dir;      /* [B]  dir = 2*axis + (side>0) */
    const int32_t *own_row;                  /* [C+1] faces owned by a cell */
    const int32_t *nbr_row, *nbr_face;       /* [C+1],[F] faces seen from the neighbour */
    const int32_t *bnd_row;                  /* [C+1] boundary faces of a cell */
    const int32_t *sell_ptr;                 /* [n_slices+1] slice start (entries) */
    const int32_t *sell_col;                 /* [nnz_pad] column (cell) */
    const int32_t *sell_code;                /* [nnz_pad] f: upper[f]; -f-1: lower[f]; INT32_MIN: pad */
    /* geometry (FP64) */
    const double *centers, *sizes, *volumes; /* [C,3],[C,3],[C] */
    const double *w, *mag_sf, *delta;        /* [F] */
    const double *bnd_d, *bnd_mag_d, *bnd_mag_sf, *bnd_sign; /* [B,3],[B],[B],[B] */
    const double *ata_inv;                   /* [C,9] */
    /* hanging (2:1) faces */
    const int32_t *hang_face;                /* [H] face index */
    const double *hang_offset, *hang_move_o, *hang_move_n; /* [H,3] */
    const int32_t *hang_cell;                /* [HC] cells adjacent to a hanging face */
    const int32_t *hc_row, *hc_code;         /* [HC+1], h (owner side) or -h-1 (neighbour side) */
    /* solver row order of the symmetric one-rhs PCG (gridfoam_amd/solver_layout.py;
     * n_bricks = 0 / NULL: not built, the grid's order is used): 64-cell
     * same-level 4x4x4 bricks first (lane = in-brick Morton index), then the
     * remaining cells in the grid's order */
    int32_t n_bricks, n_expl;
    int64_t pnnz_pad;                        /* padded entries of the solver-order sliced ELL */
    const int32_t *row_perm, *row_iperm;     /* [C] solver row -> cell, cell -> solver row */
    const int32_t *brick_nbr;                /* [nb*6] brick across side 2a+s: >=0 brick, -1 none, -2-e explicit block e */
    const int32_t *brick_vp;                 /* [nb*3*64] +axis face code per lane (f: upper[f]; INT32_MIN none) */
    const int32_t *brick_vm;                 /* [nb*3*16] -axis face code toward a neighbour brick (-f-1: lower[f]) */
    const int32_t *expl_col, *expl_code;     /* [ne*64] explicit 2:1 slots: solver column, face code */
    const int32_t *psell_ptr, *psell_col, *psell_code; /* sliced ELL over solver rows */
    /* boundary faces of each cell: bnd_face[bnd_row[c] .. bnd_row[c+1]) (any
     * boundary-face order: domain faces, then the two sides of immersed faces) */
    const int32_t *bnd_face;                 /* [B] */
    const uint8_t *bnd_snap;                 /* [B] immersed side within the Gibou distance (NULL: none) */
} gf_mesh;

const char *gf_last_error(void);
int gf_device_count(int *count);
/* sizeof(gf_mesh), sizeof(gf_assembly), sizeof(gf_krylov_cfg): lets a
 * binding check its struct layouts against the library (no GPU needed) */
int gf_abi_sizes(int64_t out[3]);

/* ---- segmented PCG for batched serving.  No counterpart in the reference
 * (it has no serving mode): several cases concatenated into one mesh form a
 * block-diagonal operator whose rows of case s are [seg_row[s], seg_row[s+1])
 * (every segment at least one row; seg_row on the device and on the host,
 * row_seg [C] the segment of each row).  Each segment runs the Jacobi PCG of
 * solvers/cg.py:75-174 with its own scalars and its own threshold
 * max(atol, rtol |r0|) (solvers/base.py:140-162).  x [C] holds the initial
 * guess and receives the solution; per segment: iterations, (|r0|, |r|),
 * converged (host arrays).  workspace: device bytes from
 * gf_seg_pcg_workspace; n_active_host: 2 n_seg pinned host words for the polls. */
int gf_seg_pcg_workspace(int32_t n_cells, int32_t n_seg, int64_t nnz_pad, int64_t *bytes);
int gf_seg_pcg(const gf_mesh *m, int32_t n_seg, const int32_t *seg_row, const int32_t *seg_row_host,
               const int32_t *row_seg, const double *diag, const double *upper, const double *lower,
               const double *b, double *x, int32_t precond, double atol, double rtol, int32_t max_iter,
               int32_t check_every, void *workspace, int64_t workspace_bytes, int32_t *iters_out,
               double *res_out, uint8_t *conv_out, int32_t *n_active_host, gf_stream_t stream);
/* ---- segmented BiCGSTAB for the momentum systems of a batch.  No
 * counterpart in the reference (serving only): the three components of
 * every segment run the right-preconditioned BiCGSTAB of
 * solvers/bicgstab.py:76-272 (restarts on breakdown, 2-norm) with their
 * own scalars and thresholds.  b, x: [C][3] row-major (x: initial guess in,
 * solution out); per (segment, component) s*3+c: iterations, (|r0|, final
 * residual), converged.  workspace from gf_seg_bicgstab_workspace;
 * n_active_host: 2 n_seg pinned host words. */
int gf_seg_bicgstab_workspace(int32_t n_cells, int32_t n_seg, int64_t nnz_pad, int64_t *bytes);
int gf_seg_bicgstab(const gf_mesh *m, int32_t n_seg, const int32_t *seg_row_host, const double *diag,
                    const double *upper, const double *lower, const double *b, double *x, int32_t precond,
                    double atol, double rtol, int32_t max_iter, int32_t max_restart, int32_t check_every,
                    void *workspace, int64_t workspace_bytes, int32_t *iters_out, double *res_out,
                    uint8_t *conv_out, int32_t *n_active_host, gf_stream_t stream);

/* ---- axis-projected immersed boundary (replaces the AP payload fluxel's
 * builder hands to AxisProjectedGrid, src/gridfoam/core/grid/axis_projected.py:163-211;
 * contract in docs/source/contributor_guide/architecture/data_contracts.rst:181-231)
 * for analytic bodies.  kind 0: plane normal to `axis` at p[0] (open surface);
 * 1: box p[0..2] - p[3..5]; 2: prism along z over the outline
 * r(theta) = p[2] (1 + sum_k p[2+k] cos k theta + p[10+k] sin k theta), k = 1..8,
 * centred at (p[0], p[1]); 3: bluff body, the box p[0..2] - p[3..5] (nose at
 * x = p[0]) with nose edges rounded by p[6], the top rear edge cut by a plane
 * of outward normal (p[7], 0, p[8]) through the top p[9] ahead of the rear,
 * and `axis` (<= 4) vertical struts of radius p[10] at (p[12+2k], p[13+2k])
 * reaching p[11] below the body.  Both sides of a cut face carry the body's
 * patch. */
typedef struct gf_shape {
    int32_t kind, axis, patch, pad;
    double p[20];
} gf_shape;
/* Per internal face: cut or not, axis distances owner/neighbour centre ->
 * body, near-boundary flags (distance / dx <= dx / ref_length), patch ids
 * (-1 where uncut) and the body-outward unit normals at the two hits
 * ([3*F] each; they take the place of the surface-mesh anchor ids fluxel
 * hands to the force sampling, post/forces/sampling.py:140-166); at most 16
 * bodies (host array). */
int gf_ap_cut(int32_t n_faces, const int32_t *owner, const int32_t *neighbour, const int32_t *axis,
              const double *centers, const double *sizes, const double *face_centers, int32_t n_shapes,
              const gf_shape *shapes, double ref_length, uint8_t *is_immersed, double *dist_owner,
              double *dist_neighbour, uint8_t *near_owner, uint8_t *near_neighbour, int32_t *patch_owner,
              int32_t *patch_neighbour, double *normal_owner, double *normal_neighbour, gf_stream_t stream);

/* ---- octree face-neighbour operator (replaces fluxel-cfd's
 * FluxelManager.build_axis_projected_mesh, src/gridfoam/core/grid/factory.py:32-65,
 * consumed at src/gridfoam/core/grid/axis_projected.py:97-161) ---------- */
typedef struct gf_octree gf_octree;
/* counts_out = {n_cells, n_faces, n_bnd_faces, lmax} */
int gf_octree_build(const double lower[3], const double upper[3], const int32_t root[3],
                    int32_t n_regions, const double *rmin, const double *rmax,
                    const int32_t *rlevel, int32_t n_leaf, gf_octree **out,
                    int64_t counts_out[4], gf_stream_t stream);
/* The same build with fluxel's surface refinement (meta/config.py:100-106,
 * `target_level`: the finest level around the input surface; passed to the
 * mesher at core/grid/factory.py:52-64) for analytic bodies: before the
 * region boxes, every leaf below target_level whose centre lies within half
 * its diagonal of a body's surface (|G| of gf_shapes' implicit functions) is
 * split, to a fixed point.  Bodies of kind 0 (plane), 1 (box) and 3 (bluff
 * body); at most 16. */
int gf_octree_build_surface(const double lower[3], const double upper[3], const int32_t root[3],
                            int32_t n_regions, const double *rmin, const double *rmax,
                            const int32_t *rlevel, int32_t n_shapes, const gf_shape *shapes,
                            int32_t target_level, int32_t n_leaf, gf_octree **out,
                            int64_t counts_out[4], gf_stream_t stream);
/* device-to-device copy into caller buffers */
int gf_octree_fetch(const gf_octree *t, int32_t *level, int32_t *ijk, int32_t *f_owner,
                    int32_t *f_nbr, int32_t *f_axis, int32_t *b_owner, int32_t *b_dir,
                    gf_stream_t stream);
void gf_octree_free(gf_octree *t);

/* ---- grid geometry (axis_projected.py:130-161, :715-811) ---------------- */
int gf_grid_geometry(int32_t n_cells, int32_t n_faces, int32_t n_bnd, const double lower[3],
                     const double upper[3], const int32_t root[3], const int32_t *level,
                     const int32_t *ijk, const int32_t *owner, const int32_t *nbr,
                     const int32_t *axis, const int32_t *b_owner, const int32_t *b_dir,
                     double *centers, double *sizes, double *volumes, double *face_centers,
                     double *sf, double *bnd_centers, double *bnd_sf, gf_stream_t stream);

/* ---- face geometry + operator layout (face_geometry.py:107-262) ---------
 * Phase 1 computes per-face weights and hanging flags and the per-cell row
 * counts; the host scans counts (gf_exclusive_scan) and calls phase 2. */
int gf_face_geometry(const gf_mesh *m, const double *face_centers, double *w, double *mag_sf,
                     double *delta, uint8_t *is_hang, double h_min, gf_stream_t stream);
int gf_row_counts(const gf_mesh *m, int32_t *own_cnt, int32_t *nbr_cnt, int32_t *bnd_cnt,
                  gf_stream_t stream);
/* Checks the face-order contract above (index ranges; internal faces
 * grouped by owner, ascending; own_row when set; every boundary face once in
 * the bnd_row/bnd_face lists of its owner when set; neighbour on the +axis
 * side when centers is set).  flags_out (host) = OR of 1 index, 2 internal
 * order, 8 row lists, 16 neighbour side;
 * GF_EINVAL with a message when any is set.  Guards the boundary that
 * replaces GridBase's face lists (core/grid/base.py:19-275). */
int gf_mesh_validate(const gf_mesh *m, int32_t *flags_out, gf_stream_t stream);
int gf_exclusive_scan_i32(const int32_t *in, int32_t *out, int32_t n, gf_stream_t stream);
int gf_fill_nbr_lists(const gf_mesh *m, int32_t *nbr_face, gf_stream_t stream);
/* boundary faces grouped by owner (ascending face index in a group) into
 * bnd_face, after gf_row_counts / scan gave bnd_row */
int gf_fill_bnd_lists(const gf_mesh *m, int32_t *bnd_face, gf_stream_t stream);
int gf_sell_widths(const gf_mesh *m, int32_t *slice_len, gf_stream_t stream);
int gf_sell_fill(const gf_mesh *m, int32_t *sell_col, int32_t *sell_code, gf_stream_t stream);
int gf_lsq_ata_inv(const gf_mesh *m, double *ata_inv, gf_stream_t stream);
int gf_hang_geometry(const gf_mesh *m, const double *face_centers, double *offset,
                     double *move_o, double *move_n, gf_stream_t stream);
int gf_hang_cell_lists(const gf_mesh *m, const int32_t *cell_to_hc, int32_t *hc_code,
                       gf_stream_t stream);

/* ---- boundary state (boundary_ops.py:311-326, :545-593) -----------------
 * bc arrays are per boundary face: frac [B], ref_v [B,nc], ref_g [B,nc],
 * covered [B] (1 when a non-empty boundary condition owns the face). */
int gf_bc_state(const gf_mesh *m, int32_t nc, const double *psi, const double *frac,
                const double *ref_v, const double *ref_g, const uint8_t *covered,
                double *psi_b, gf_stream_t stream);
/* slip: ref_v = psi_O - (psi_O . n) n on selected faces (boundaries/basic/slip.py:46-101) */
int gf_bc_slip(const gf_mesh *m, const double *psi, const uint8_t *slip_mask, double *ref_v,
               gf_stream_t stream);

/* ---- explicit operators (fv/kernels/*.py, fv/fvc/*.py) ------------------ */
/* least_squares.py:29-101; hang_only evaluates only cells next to hanging faces */
int gf_lsq_grad(const gf_mesh *m, int32_t nc, const double *psi, const double *psi_b,
                const uint8_t *covered, int32_t hang_only, double *grad, gf_stream_t stream);
/* Green-Gauss gradient (schemes/grad.py:21-44, gauss_gradient.py:15-73) from
 * face values psi_s [F,nc] and psi_b [B,nc] (gf_interpolate + boundary state):
 * grad [C,nc,3] = sum psi_f Sf / V */
int gf_gauss_grad(const gf_mesh *m, int32_t nc, const double *psi_s, const double *psi_b,
                  double *grad, gf_stream_t stream);
/* Reverse mode of gf_gauss_grad (autograd through fv/schemes/grad.py:21-44):
 * g_grad [C,nc,3] -> g_psi_s [F,nc], g_psi_b [B,nc] (overwritten); chained
 * with gf_interpolate_vjp / gf_bc_state_vjp for the face values' inputs */
int gf_gauss_grad_vjp(const gf_mesh *m, int32_t nc, const double *g_grad, double *g_psi_s,
                      double *g_psi_b, gf_stream_t stream);
/* fvc/reconstruct.py:27-107: U [C,3] from the face flux phi_s [F], phi_b [B]:
 * per axis, the area-weighted average of the face-normal velocities */
int gf_reconstruct(const gf_mesh *m, const double *phi_s, const double *phi_b, double *U,
                   gf_stream_t stream);
/* Reverse mode of gf_reconstruct (autograd through fvc/reconstruct.py:27-107):
 * g_U [C,3] -> g_phi_s [F], g_phi_b [B] (overwritten); work [C,3] scratch */
int gf_reconstruct_vjp(const gf_mesh *m, const double *g_U, double *work, double *g_phi_s,
                       double *g_phi_b, gf_stream_t stream);
/* interpolate.py:20-63 + face_interpolation.py:114-197; grad valid on hang cells */
int gf_interpolate(const gf_mesh *m, int32_t nc, const double *psi, const double *grad,
                   double *psi_f, gf_stream_t stream);
/* fvc/div.py:15-64 */
int gf_fvc_div(const gf_mesh *m, int32_t nc, const double *phi_s, const double *phi_b,
               double *out, gf_stream_t stream);
/* Reverse mode of gf_fvc_div (autograd through fv/fvc/div.py:15-64):
 * g_out [C,nc] -> g_phi_s [F,nc], g_phi_b [B,nc] (overwritten) */
int gf_fvc_div_vjp(const gf_mesh *m, int32_t nc, const double *g_out, double *g_phi_s,
                   double *g_phi_b, gf_stream_t stream);
/* Reverse mode of gf_interpolate (autograd through interpolate.py:20-63 and
 * face_interpolation.py:114-197): g_psi_f [F,nc] -> g_psi [C,nc] and, on
 * grids with hanging faces, g_grad [C,nc,3] on the cells next to them (both
 * added to); chained with gf_lsq_grad_vjp for the gradient's own inputs */
int gf_interpolate_vjp(const gf_mesh *m, int32_t nc, const double *g_psi_f, double *g_psi,
                       double *g_grad, gf_stream_t stream);
/* FP64 mass-conservation check of a face flux (post/diagnostics/continuity.py:12-49):
 * out[0] = dt sum(|div phi| V)/sum V (local), out[1] = dt sum(div phi V)/sum V (global);
 * div_scratch [C] receives div(phi) */
int gf_continuity_error(const gf_mesh *m, const double *phi_s, const double *phi_b, double dt,
                        double *div_scratch, double *out, gf_stream_t stream);
/* fvc/sn_grad + face_interpolation.py:200-246: hanging snGrad correction [H,nc] */
int gf_sn_grad_hang(const gf_mesh *m, int32_t nc, const double *grad, double *corr,
                    gf_stream_t stream);

/* ---- implicit assembly (fvm/laplacian.py:101-201, fvm/div.py:15-73,
 * fvm/ddt.py:7-40, fvmatrix.py:181-211) -----------------------------------
 * One fused kernel family: result = s_ddt*ddt + s_div*div(phi,psi)
 *                                   + s_lap*laplacian(gamma,psi).
 * Terms with a zero sign are skipped.  gamma is a scalar (gamma_cell NULL)
 * or a cell field. Hanging-face explicit parts come from grad (hang cells). */
typedef struct gf_assembly {
    int32_t nc;
    double s_ddt, s_div, s_lap;
    /* ddt: a, b, c weights, dt; psi_old/psi_older [C,nc] (older may be NULL) */
    double ddt_a, ddt_b, ddt_c, dt;
    const double *psi_old, *psi_older;
    /* div: scheme 0 upwind, 1 linear, 2..7 TVD deferred correction (vanLeer, minmod,
     * superbee, monotonized central, limitedLinear, Koren; schemes/div.py:145-461);
     * phi_s [F], phi_b [B] */
    int32_t div_scheme;
    const double *phi_s, *phi_b;
    /* laplacian: face diffusivity from gamma_cell by linear interpolation, or
     * harmonic (series resistances, laplacian.py:38-73) with lap_harmonic */
    double gamma;
    const double *gamma_cell;
    int32_t lap_corrected;
    int32_t lap_harmonic;
    /* field + boundary state */
    const double *psi, *grad;          /* [C,nc], [C,nc,3] (hang cells; every cell for TVD) */
    const double *frac, *ref_v, *ref_g, *psi_b; /* boundary state */
    const uint8_t *covered;
} gf_assembly;
int gf_assemble(const gf_mesh *m, const gf_assembly *a, double *diag, double *upper,
                double *lower, double *source, double *ffc, gf_stream_t stream);
/* Reverse mode of gf_assemble, for the assembly inputs that autograd
 * differentiates in the reference's torch assembly (fv/fvm/laplacian.py:38-201,
 * fv/fvm/div.py:15-73, fv/schemes/div.py:36-142, fv/fvm/ddt.py:7-40): given
 * the gradients of a scalar with respect to diag [C], upper [F], lower [F],
 * source [C,nc] and ffc [F,nc] (g_ffc may be NULL: zero), writes
 *   g_gamma [C]        d/d gamma_cell; with a uniform gamma, per-cell shares
 *                      whose sum is d/d gamma (needs work [F] scratch),
 *   g_phi_s [F], g_phi_b [B], g_psi_old [C,nc], g_psi_older [C,nc],
 *   g_ref_v [B,nc], g_ref_g [B,nc]   (boundary faces not covered: 0).
 * Any output may be NULL.  The field values psi (hanging-face corrections)
 * and the boundary fractions are held fixed.  Upwind follows torch.clamp's
 * backward at phi = 0; the harmonic mean follows torch.minimum's tie rule
 * where both diffusivities are zero. */
int gf_assemble_vjp(const gf_mesh *m, const gf_assembly *a, const double *g_diag,
                    const double *g_upper, const double *g_lower, const double *g_source,
                    const double *g_ffc, double *work, double *g_gamma, double *g_phi_s,
                    double *g_phi_b, double *g_psi_old, double *g_psi_older, double *g_ref_v,
                    double *g_ref_g, double *g_grad, gf_stream_t stream);
/* g_grad [C,nc,3] (may be NULL): d/d of the hang-cell least-squares gradient
 * that the hanging-face corrections read (written on hang cells only; zero
 * it first); chained with gf_lsq_grad_vjp and gf_bc_state_vjp it carries the
 * field values and boundary values through those corrections.  Not for TVD
 * sources (GF_EINVAL). */
/* Reverse mode of gf_lsq_grad (least_squares.py:29-101): g_grad [C,nc,3] ->
 * g_psi [C,nc], g_psi_b [B,nc] (covered faces; 0 elsewhere); hang_only as in
 * the forward; work [C,nc,3] must be zero on entry. */
int gf_lsq_grad_vjp(const gf_mesh *m, int32_t nc, const double *g_grad, const uint8_t *covered,
                    int32_t hang_only, double *work, double *g_psi, double *g_psi_b,
                    gf_stream_t stream);
/* Reverse mode of gf_bc_state (+ gf_bc_slip on slip_mask faces, may be NULL):
 * g_psi_b [B,nc] -> g_psi [C,nc], g_ref_v, g_ref_g [B,nc] (all added to; the
 * reference arrays may be NULL); on slip faces the ref_v gradient is carried
 * on to the cell values and left 0 */
int gf_bc_state_vjp(const gf_mesh *m, int32_t nc, const double *frac, const uint8_t *covered,
                    const uint8_t *slip_mask, const double *g_psi_b, double *g_psi,
                    double *g_ref_v, double *g_ref_g, gf_stream_t stream);

/* ---- LDU algebra (fvmatrix.py:242-323, :424-468) ------------------------ */
int gf_ldu_multiply(const gf_mesh *m, int32_t nc, const double *diag, const double *upper,
                    const double *lower, const double *x, double *y, gf_stream_t stream);
/* rAU = 1/(diag/V); HbyA = rAU * (b - A x + diag x)/V   (piso.py:264-267) */
int gf_rau_hbya(const gf_mesh *m, const double *diag, const double *upper, const double *lower,
                const double *source, const double *U, double *rAU, double *HbyA,
                gf_stream_t stream);
/* flux = upper psi_N - lower psi_O + ffc on internal faces */
int gf_ldu_flux(const gf_mesh *m, int32_t nc, double sign, const double *upper,
                const double *lower, const double *ffc, const double *psi, double *flux,
                gf_stream_t stream);

/* ---- flux helpers (flux.py:32-162, ddt_corr.py:13-78, adjust_phi.py:44-127,
 * pressure_correction.py:197-363) ---------------------------------------- */
int gf_face_flux(const gf_mesh *m, const double *Uf_s, const double *Uf_b,
                 const uint8_t *covered, double *phi_s, double *phi_b, gf_stream_t stream);
int gf_ddt_corr_add(const gf_mesh *m, const double *U_old, const double *U_older,
                    const double *phi_old, const double *phi_older, const double *rAU, double b,
                    double c, double dt, double *phi_s, gf_stream_t stream);
int gf_constrain_bnd_flux(const gf_mesh *m, const double *psi_b, const uint8_t *mask,
                          double *phi_b, gf_stream_t stream);
/* kind per boundary face: 0 none, 1 fixed (dirichlet), 2 adjustable */
int gf_adjust_phi(const gf_mesh *m, const int8_t *kind, double *phi_b, double *scratch,
                  gf_stream_t stream);
/* phi_s = phi_hbya_s + sign_p * flux; boundary: phi_b -= rAtU_P |Sf| snGrad_b */
int gf_correct_phi(const gf_mesh *m, const double *phi_hbya_s, const double *phi_hbya_b,
                   const double *upper, const double *lower, const double *ffc, const double *p,
                   const double *rAtU, const double *frac, const double *ref_v,
                   const double *ref_g, const uint8_t *covered, double *phi_s, double *phi_b,
                   gf_stream_t stream);
/* U = HbyA - rAtU grad p */
int gf_correct_velocity(int32_t n_cells, const double *HbyA, const double *rAtU,
                        const double *grad_p, double *U, gf_stream_t stream);
/* Symmetric elimination of prescribed cell values (fvmatrix.py:325-394):
 * couplings of fixed cells zeroed, their neighbours' contributions moved to the
 * right-hand side, fixed rows diag*x = diag*value (zero diagonal -> +-1 by the
 * sign of sum(diag)).  cells [K] unique, values [K,nc]; fixed [C] (uint8),
 * presc [C,nc] and dsum [1] are caller scratch. */
int gf_fixed_values(const gf_mesh *m, int32_t nc, int32_t n_fixed, const int32_t *cells,
                    const double *values, const double *diag, const double *upper,
                    const double *lower, const double *source, uint8_t *fixed, double *presc,
                    double *dsum, double *out_diag, double *out_upper, double *out_lower,
                    double *out_source, gf_stream_t stream);
/* H1 = -(sum of off-diagonals of the row)/V  (fvmatrix.py:250-264) */
int gf_ldu_h1(const gf_mesh *m, const double *upper, const double *lower, double *h1, gf_stream_t stream);
/* SIMPLEC rAtU = 1/(1/rAU - H1), bounded by 0.1/rAU when bounded != 0
 * (pressure_correction.py:163-193) */
int gf_simplec_rautu(int32_t n_cells, const double *rAU, const double *h1, int32_t bounded, double *rAtU,
                     gf_stream_t stream);
/* diag[ref] += diag[ref]; source[ref] += diag_old*value (reference_value.py:33-51) */
int gf_set_reference(double *diag, double *source, int32_t cell, double value,
                     gf_stream_t stream);
/* SIMPLE momentum under-relaxation: diag /= alpha; source += ((1-alpha)/alpha) diag_old U
 * (algorithms/simple.py:302-311) */
int gf_relax_momentum(int32_t n_cells, double alpha, double *diag, double *source, const double *U,
                      gf_stream_t stream);
/* out = source - x V: momentum predictor source - V grad(p) (piso.py:247-250,
 * simple.py:315-318) and pressure source - div(phiHbyA) V (pressure_correction.py:121-124) */
int gf_predictor_source(int32_t n_cells, int32_t nc, const double *source, const double *volumes,
                        const double *x, double *out, gf_stream_t stream);
/* x = old + alpha (x - old)  (simple.py:386) */
int gf_relax_field(int64_t n, double alpha, const double *old, double *x, gf_stream_t stream);
/* dst = src, 16 bytes per lane (no counterpart in the reference: the HBM
 * reference stream the benchmark times and scales its HBM counters with) */
int gf_stream_copy(const void *src, void *dst, int64_t bytes, gf_stream_t stream);
/* out = a*x + b*y elementwise (FvMatrix +/-) */
int gf_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out,
             gf_stream_t stream);
/* ||A psi - b||_2 / (||b||_2 + eps) written to out[0] (residual.py:12-36) */
int gf_initial_residual(const gf_mesh *m, int32_t nc, const double *diag, const double *upper,
                        const double *lower, const double *source, const double *psi,
                        double *scratch, double *out, gf_stream_t stream);
/* gf_initial_residual without the scratch argument and with no allocation of
 * any kind: its per-workgroup partials live in a static device array of the
 * library (2 MiB per GPU, held until the process ends), so n_cells is limited
 * to 33,554,432 (GF_EINVAL beyond).  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  Calls that overlap on one GPU share the array and race. */
int gf_initial_residual_noalloc(const gf_mesh *m, int32_t nc, const double *diag, const double *upper,
                                const double *lower, const double *source, const double *psi,
                                double *out, gf_stream_t stream);

/* ---- Krylov solves (solvers/cg.py:75-174, bicgstab.py:76-272,
 * preconditioners.py:76-96, krylov.py:34-82, base.py:140-162) -------------
 * Solves nrhs independent components that share one LDU matrix in a single
 * batched launch sequence; results equal per-component solves.
 * stats_out[nrhs][4] = {initial_residual, final_residual, iterations, converged}. */
typedef struct gf_krylov_cfg {
    int32_t method;        /* 0 cg, 1 bicgstab */
    int32_t precond;       /* 0 none, 1 jacobi */
    int32_t norm;          /* 0 L2, 1 Linf */
    int32_t max_iter, max_restart;
    double atol, rtol;
    int32_t mixed;         /* cg only: FP32 inner solve inside FP64 refinement */
    double inner_rtol;     /* mixed: the FP64 refinement runs once the FP32 recursive residual
                              is below inner_rtol x the last true residual (0 < inner_rtol < 1) */
    int32_t check_every;   /* host polls the device state every k iterations (PCG: k rounded up to even) */
    int32_t symmetric;     /* caller guarantees upper == lower: a one-rhs PCG then runs on the brick layout */
} gf_krylov_cfg;
/* workspace bytes needed for (C, nrhs, cfg); caller allocates once */
int64_t gf_krylov_workspace_bytes(const gf_mesh *m, int32_t nrhs, const gf_krylov_cfg *cfg);
int gf_krylov_solve(const gf_mesh *m, const gf_krylov_cfg *cfg, int32_t nrhs, int32_t transpose,
                    const double *diag, const double *upper, const double *lower,
                    const double *b /* [C,nrhs] */, double *x /* [C,nrhs], in/out */,
                    void *workspace, double *stats_out, gf_stream_t stream);
/* PCG chunks of check_every iterations are replayed through HIP graphs cached
 * per workspace (no counterpart in the reference: launch-overhead control).
 * gf_krylov_use_graphs(0) launches them one by one; returns the previous
 * setting.  gf_krylov_release drops a workspace's graphs before it is freed. */
int gf_krylov_use_graphs(int32_t on);
int gf_krylov_release(const void *workspace);
/* ldu_grads.py:11-53: gd = -sum lambda x, gu = -sum lambda_O x_N, gl = -sum lambda_N x_O */
int gf_ldu_grads(const gf_mesh *m, int32_t nc, const double *lambda, const double *x,
                 double *g_diag, double *g_upper, double *g_lower, gf_stream_t stream);
/* Benchmark of the PCG direction/SpMV kernel on its own (beta = 0, so p = z =
 * x and Ap = A x), on the layout a symmetric one-rhs PCG solve uses (bricks
 * when the mesh has them); no counterpart in the reference (measurement).
 * phase 0 packs the operator from diag/upper/lower and copies x [C] (FP64,
 * cell order) in; phase 1 launches the kernel reps times; phase 2 writes
 * y = A x [C] (FP64, cell order).  fp32 = 1: the mixed-precision operator
 * (FP32 values and vectors, FP64 diagonal and accumulation).  Scratch of
 * gf_pcg_dir_bench_bytes(m, fp32) bytes, kept between the phases. */
int64_t gf_pcg_dir_bench_bytes(const gf_mesh *m, int32_t fp32);
int gf_pcg_dir_bench(const gf_mesh *m, int32_t phase, int32_t fp32, int32_t reps, const double *diag,
                     const double *upper, const double *lower, const double *x, double *y, void *scratch,
                     gf_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* GRIDFOAM_HIP_H */
