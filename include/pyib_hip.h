/*
 * pyib_hip.h -- C ABI of the MI355X (gfx950) time-step hot path of PyIB.
 *
 * The reference solver has no FFI: its seam is the array backend chosen at
 * import (src/pyib/solver/backend.py:10-34) plus duck-typed operator objects
 * (IBCore, Driver, ImmersedBoundary, ClosedSphereLoadCorrection).  Each entry
 * point below replaces one device-side operation those objects perform through
 * NumPy/CuPy; the comment on each names the reference call it replaces.
 *
 * Conventions
 *  - All array arguments are device pointers (hipMalloc'd or torch-owned); no
 *    entry point allocates, frees or synchronises.  Workspaces are passed in.
 *  - `dtype` selects the floating type of coefficients and vectors:
 *    PYIB_F64 (double) or PYIB_F32 (float).  Reductions always accumulate and
 *    return in double.
 *  - `stream` is a hipStream_t passed as void*; NULL means the default stream.
 *  - Signed link indices encode the circulation sign: v >= 0 means +x[v],
 *    v < 0 means -x[-v-1].
 *  - Every function returns 0 on success and a nonzero code on failure; the
 *    message is then available from pyib_last_error().
 */
#ifndef PYIB_HIP_H
#define PYIB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PYIB_F64 = 0, PYIB_F32 = 1 };

/* Incidence/geometry tables of one mesh, in PyIB's global numbering.
 * Built once on the host from the reference geometry (ib_core_np.py:24-126). */
typedef struct pyib_mesh {
  int32_t dtype;            /* PYIB_F64 or PYIB_F32: type of every float array below */
  int32_t reserved;
  int64_t ncell, nface, nedge, nghost;
  /* curl: face <- edges over the flat face/edge links (ib_core_np.py:133-134) */
  const int32_t *curl_ptr;  /* nface+1 */
  const int32_t *curl_idx;  /* signed edge index per link, link order */
  /* rot: edge <- faces, the transpose of curl (ib_core_np.py:129-131) */
  const int32_t *rot_ptr;   /* nedge+1 */
  const int32_t *rot_idx;   /* signed face index */
  /* cell <- face incidence: for cell c, entries [cf_ptr[c], cf_mid[c]) are the
   * faces whose first owner is c, [cf_mid[c], cf_ptr[c+1]) those whose second
   * owner is c, each in face order; cf_other is the opposite owner, or the
   * ghost index ncell+k for boundary face k (ib_driver_np.py:112-118). */
  const int32_t *cf_ptr;    /* ncell+1 */
  const int32_t *cf_mid;    /* ncell */
  const int32_t *cf_face;
  const int32_t *cf_other;
  const int32_t *f2c;       /* nface x 2, second owner -1 on boundary faces */
  const void *rf1, *rf2;    /* nface x 3 */
  const void *aol;          /* nface: (rf.nf)/|rf|^2 */
  const void *vol_ci;       /* ncell: inverse cell volume */
  /* boundary ghosts, one per boundary face k (ib_driver_np.py:275-291) */
  const int32_t *gh_face;   /* nghost */
  const int32_t *gh_cell;   /* first owner */
  const int32_t *gh_kind;   /* 0 inlet zero-tangential, 1 zero-gradient, 2 wall, 3 lid, 4 spanwise copy */
  const void *gh_nf;        /* nghost x 3: face area vector */
  const void *gh_af2;       /* nghost: Af^2 */
  /* inlet stream-function values (ib_driver_np.py:271) */
  int64_t ninlet;
  const int32_t *inlet_edge;
  const void *inlet_val;
  const uint8_t *dir_mask;  /* nedge: Dirichlet edges, residual forced to zero */
} pyib_mesh;

/* Workspaces for one operator application (owned by the caller). */
typedef struct pyib_work {
  void *face;               /* nface */
  void *cell_a;             /* ncell x 3 */
  void *cell_b;             /* ncell x 3 */
  double *partials;         /* >= PYIB_MAX_PARTIALS doubles */
} pyib_work;

#define PYIB_MAX_PARTIALS 8192

/* CG state; vectors have the mesh dtype.  scal is a device array of
 * PYIB_NSCAL doubles: [eta, eta_new, den, alpha, beta, rr, err_tot, l2_init]. */
typedef struct pyib_cg {
  void *x, *r, *p, *w;
  const void *prec;
  double *scal;
} pyib_cg;
#define PYIB_NSCAL 8

/* Structured per-level layout of the elliptic operator (one box per 2:1 level).
 * Edge blocks per level: x (nx, ny+1, nz+1), y (nx+1, ny, nz+1), z (nx+1, ny+1, nz),
 * +axis oriented; cell block (nx, ny, nz); cell vectors interleaved (x, y, z of a cell
 * adjacent: element 3*c + d). */
#define PYIB_MAX_LEVELS 8
typedef struct pyib_level {
  int32_t nx, ny, nz, reserved;
  int64_t cell_off;         /* first internal cell of the level */
  int64_t e_off[3];         /* first internal x/y/z edge of the level */
  double h;                 /* spacing */
  int32_t hole_lo[3], hole_hi[3];   /* cells [hole_lo, hole_hi) of the box belong to the finer
                                     * level (no real cell, no real edge; empty: lo == hi) */
} pyib_level;

typedef struct pyib_grid {
  int32_t dtype;            /* PYIB_F64 or PYIB_F32: composite coefficients and vectors */
  int32_t nlevels;
  int64_t ncell, nedge;     /* internal (padded) sizes */
  pyib_level lev[PYIB_MAX_LEVELS];
  const uint8_t *cell_reg;  /* ncell: V row computed by the stencil (bit 0); bit 1: a box-face row of
                             * the coarsest level, whose faces on the box take twice the coefficient
                             * (zero ghost, AoL = 2h) */
  const uint8_t *cell_reg_g1; /* ncell: G1 row computed by the stencil */
  const uint8_t *edge_reg;  /* nedge: G3 row computed by the stencil (w must be zero-initialised:
                             * empty slots, and Dirichlet rows left out of the G3 composite, are
                             * never written) */
  const uint8_t *dir;       /* nedge: Dirichlet edge or empty slot (residual forced to 0) */
  const uint8_t *cell_g1v;  /* ncell, flags of the fused G1+V pass: bit 0 = cell_reg_g1, bit 1 = its
                             * x = interp(curl p) is stored for the composite V rows (the pass keeps
                             * the rest in LDS), bit 2 = cell_reg bit 0, bit 3 = x is read from memory
                             * (a real cell with a composite G1 row), bit 4 = cell_reg bit 1; x = 0 on
                             * the other cells */
  /* composite rows of the other (interface/boundary) rows in sliced ELL: slices of 64
   * rows (one wavefront), entry j of lane l of slice s at sptr[s] + 64*j + l, width[s]
   * entries per lane, row[64*s + l] = internal row id or -1 (padding, col 0, coef 0) */
  int64_t g1_nslice; const int32_t *g1_row, *g1_width; const int64_t *g1_sptr;
  const int32_t *g1_col; const void *g1_coef;   /* coef entries x 3 */
  int64_t v_nslice;  const int32_t *v_row, *v_width;  const int64_t *v_sptr;
  const int32_t *v_col;  const void *v_coef;    /* coef entries */
  const double *v_rw;   /* 64*v_nslice: volume (1/vol_ci) of the row's cell, 0 on padding */
  int64_t g3_nslice; const int32_t *g3_row, *g3_width; const int64_t *g3_sptr;
  const int32_t *g3_col; const void *g3_coef;   /* coef entries x 3 */
  /* Classes of the Jacobi preconditioner over the edge slots.  Used by the PCG iteration when
   * edge_cls is non-NULL and the pyib_cg's prec is cls_prec; any other prec array is read as is.
   * On a regular edge of a level the inverse diagonal is the same number at every edge, so the
   * iteration takes it from prec_level instead of reading prec, and it skips the slots whose
   * direction stays zero (empty slots and Dirichlet rows: r = 0 and p = 0 there). */
  const uint8_t *edge_cls;  /* nedge: bit 0 = edge_reg, bit 1 live (not empty, not Dirichlet),
                             * bit 2 prec == prec_level[level] bitwise, bits 4-6 the level */
  const void *cls_prec;     /* the prec array (of this dtype) the classes were derived from */
  double prec_level[PYIB_MAX_LEVELS];     /* prec on the bit-2 slots of each level */
  double prec_level_g3[PYIB_MAX_LEVELS];  /* prec_level when every regular G3 row of the level is
                                           * bit 2, else 0 (then the G3 pass reads prec) */
  const uint8_t *g3_mask;   /* Dirichlet mask the G3 composite applies to its rows, or NULL when its
                             * rows hold no Dirichlet row (after the mask is known) */
  const void *g3_rprec;     /* 64*g3_nslice, this dtype: cls_prec at each lane's G3 composite row (0 on
                             * padding), read instead of gathering prec[row] when the PCG's prec is
                             * cls_prec; or NULL */
} pyib_grid;

const char *pyib_last_error(void);
int pyib_device_count(int *count);

/* ---- incidence operators (ib_core_np.py:129-183) ---- */
int pyib_curl(const pyib_mesh *m, const void *edge, void *face, void *stream);          /* IBCore.curl */
int pyib_rot(const pyib_mesh *m, const void *face, void *edge, void *stream);           /* IBCore.rot */
int pyib_interp_f2c(const pyib_mesh *m, const void *face, void *cellvec, void *stream); /* IBCore.interp_f2c */
int pyib_integr_c2f(const pyib_mesh *m, const void *cellvec, void *face, void *stream); /* IBCore.integr_c2f */
int pyib_cdiv(const pyib_mesh *m, const void *face, void *cell, void *stream);          /* IBCore.cdiv */
/* IBCore.flow_mat + DirBC (ib_core_np.py:168-175, ib_driver_np.py:387-388):
 * w = rot(integr(V(interp(curl p)))), V x = rou x - hdt*vol_ci*cdiv(AoL*visc*cgrad0 x),
 * hdt = 0.5*dt; w is zeroed on dir_mask when mask_dir != 0.  If dot_out is
 * non-NULL, *dot_out (device) receives p.w in double (deterministic). */
int pyib_flow_mat(const pyib_mesh *m, const pyib_work *ws, double rou, double visc, double hdt,
                  const void *p, void *w, int mask_dir, double *dot_out, void *stream);

/* ---- time-step pieces (ib_driver_np.py:267-303, 458-496) ---- */
/* Driver.calc_vel: s_out = s with inlet values; U = curl(s_out); vc (ncell+nghost) x 3 */
int pyib_calc_vel(const pyib_mesh *m, const void *s_in, void *s_out, void *U, void *vc,
                  const double *lid_vec, void *stream);
/* Driver.conv_diff + momentum residual:
 * out = [base - a_vc*vc | a_vc*vc] - k_conv*conv + k_diff*diff (+ k_force*force)
 * base may be NULL (then the first term is +a_vc*vc); force may be NULL. */
int pyib_momentum_residual(const pyib_mesh *m, const void *U, const void *vc, double visc,
                           const void *base, double a_vc, double k_conv, double k_diff,
                           const void *force, double k_force, void *out, void *stream);
/* rot(integr_c2f(res_c)) (ib_driver_np.py:495-496) */
int pyib_curl_rhs(const pyib_mesh *m, const pyib_work *ws, const void *res_c, void *res_e, void *stream);

/* ---- Jacobi-PCG of Driver.helm_solve (ib_driver_np.py:311-410) ---- */
/* r = res_e (0 on dir), x = s, p = r*prec; scal[eta]=p.r, scal[l2_init]=r.r,
 * scal[err_tot] = (x.x + 1e-16) tol1^2 + tol2^2 eta */
int pyib_cg_begin(const pyib_mesh *m, const pyib_work *ws, const pyib_cg *cg, const void *res_e,
                  const void *s, double tol1, double tol2, void *stream);
/* n_iter PCG iterations.  The last iteration stops before the direction
 * update when skip_last_direction != 0, so the host can test scal[eta_new]. */
int pyib_cg_iterate(const pyib_mesh *m, const pyib_work *ws, const pyib_cg *cg, double rou,
                    double visc, double hdt, int32_t n_iter, int32_t skip_last_direction, void *stream);
/* p = r*prec + beta*p; eta = eta_new */
int pyib_cg_direction(const pyib_mesh *m, const pyib_work *ws, const pyib_cg *cg, void *stream);
/* true residual tr = res_e - A(x - s) (0 on dir); out2[0] = (tr*prec).tr, out2[1] = tr.tr */
int pyib_cg_true_residual(const pyib_mesh *m, const pyib_work *ws, const pyib_cg *cg,
                          const void *res_e, const void *s, void *tmp_edge, void *tr,
                          double rou, double visc, double hdt, double *out2, void *stream);

/* Driver.step guess: out = s + (s - s_old)*dt/dt_old (ib_driver_np.py:461), float64 */
int pyib_extrapolate(int64_t n, const double *s, const double *s_old, double dt, double dt_old,
                     double *out, void *stream);

/* ---- mesh geometry on the host (setup; IBCore.__init__, ib_core_np.py:24-126) ----
 * Host pointers, float64 / int64, row-major.  Inputs: nodes (nnode,3), e2n (nedge,2),
 * f2c (nface,2; -1 = no neighbour), f2e (nface,4), ins (nins,2) = link_fe_insert (face, edge).
 * Outputs (preallocated): lf_face, lf_edge (nlink = 4 nface + nins), epos, te (nedge,3),
 * fpos (nface,3), cpos (ncell,3), sgn (nlink), nf (nface,3), af (nface), rf1, rf2, rf
 * (nface,3), vol_ci (ncell), aol (nface).  Sums follow the reference's link order, so the
 * values are bitwise those of its NumPy code. */
int pyib_geometry(int64_t nnode, const double *nodes, int64_t nedge, const int64_t *e2n, int64_t nface,
                  const int64_t *f2c, const int64_t *f2e, int64_t nins, const int64_t *ins, int64_t ncell,
                  int64_t *lf_face, int64_t *lf_edge, double *epos, double *te, double *fpos, double *cpos,
                  double *sgn, double *nf, double *af, double *rf1, double *rf2, double *rf, double *vol_ci,
                  double *aol);

/* Gather tables of the incidence operators (host, setup; the reference's bincount order over
 * the link list, ib_core_np.py:129-157): curl (face <- its links in link order), rot (edge <-
 * its links in link order), cell <- faces (first-owner faces, then second-owner faces, each in
 * face order; cf_mid = start of the second group; boundary face k's opposite is ncell + k).
 * Signed int32 indices (v >= 0 -> +x[v], v < 0 -> -x[-v-1]).  Fails (non-zero, message in
 * pyib_last_error) if a size does not fit int32. */
int pyib_tables(int64_t nface, int64_t nedge, int64_t ncell, int64_t nlink, const int64_t *lf_face,
                const int64_t *lf_edge, const double *sgn, const int64_t *f2c, int32_t *curl_ptr,
                int32_t *curl_idx, int32_t *rot_ptr, int32_t *rot_idx, int32_t *cf_ptr, int32_t *cf_mid,
                int32_t *cf_face, int32_t *cf_other);

/* ---- offline pressure recovery of a saved field (recover_pressure, ib_postprocess.py:49-105) ----
 * The cell pressure whose face jumps best fit the saved gradient: L p = I^T W j on the cell graph
 * (w_f = area / |x2 - x1|, j_f = (g1 + g2)/2 . (x2 - x1) on internal faces), cell 0 pinned,
 * Jacobi-preconditioned CG, then the mean over the gauge cells removed.
 * pyib_pressure_graph (host) builds the cell-major adjacency: for cell c, entries [ptr[c],
 * ptr[c+1]) are its internal faces in face order, nbr the other cell, sface the face index
 * (f if c is the face's second cell, -f-1 if its first); arrays of 2 x (internal faces).
 * pyib_pressure_begin forms w, j, the right-hand side, D^-1 and the CG start; iterate runs
 * n_iter CG iterations with the scalars on the device (scal[4] = r.r, scal[5] = b.b, the host
 * stops when r.r < tol^2 b.b); finish removes the gauge mean from x and writes out4 = [sum of x
 * over the gauge cells, their count, sum w (p2 - p1 - j)^2, sum w]. */
typedef struct pyib_pgraph {
  int64_t ncell, nface;
  const int32_t *fc;                       /* nface x 2, second cell -1 on boundary faces */
  const double *pos, *grad, *area;         /* ncell x 3, ncell x 3, nface */
  const int32_t *ptr, *nbr, *sface;        /* adjacency of pyib_pressure_graph */
  const uint8_t *gauge;                    /* ncell: 1 on the gauge cells */
  double *w, *jmp;                         /* nface */
  double *dinv, *x, *r, *z, *p, *q;        /* ncell; x is the pressure */
  double *scal;                            /* 8 */
  double *partials;                        /* >= 2 * PYIB_MAX_PARTIALS */
} pyib_pgraph;
int pyib_pressure_graph(int64_t ncell, int64_t nface, const int64_t *face_cells, int32_t *ptr, int32_t *nbr,
                        int32_t *sface);
int pyib_pressure_begin(const pyib_pgraph *g, void *stream);
int pyib_pressure_iterate(const pyib_pgraph *g, int32_t n_iter, void *stream);
int pyib_pressure_finish(const pyib_pgraph *g, double *out4, void *stream);

/* ---- nested 2:1 mesh on the host (setup; make_nested_mesh, src/pyib/make_euler_mesh.py:330-463,
 * with the hanging-node finalisation of make_euler_mesh.py:787-950) ----
 * Inputs per level k (finest first): shape (nlevels,3) cells, lo (nlevels,3) box corner, dx
 * (nlevels,3) spacing, hole (nlevels,3) lower corner of level k-1's box in level k's cells (row 0
 * unused).  pyib_mesh_build keeps the arrays in an opaque handle and returns their counts:
 * counts[0..6] = nodes, cells, outer boundary quads, faces, edges, interface pairs,
 * link_fe_insert rows.  pyib_mesh_fetch copies them into host arrays of those sizes (int64
 * row-major, nodes float64): nodes (n,3), cells (n,8), boundary quads (n,4) and families (n),
 * faces_to_nodes (n,4), faces_to_cells (n,2), edges_to_nodes (n,2), faces_to_edges (n,4),
 * bc_node, bc_edge, bc_face, interface_pairs (n,5), link_fe_insert (n,2), cells and new nodes
 * per level (nlevels).  A NULL output is skipped.  pyib_mesh_free releases the handle. */
int pyib_mesh_build(int32_t nlevels, const int64_t *shape, const double *lo, const double *dx,
                    const int64_t *hole, void **handle, int64_t *counts);
int pyib_mesh_fetch(void *handle, double *nodes, int64_t *cells, int64_t *bfaces, int64_t *bfam, int64_t *f2n,
                    int64_t *f2c, int64_t *e2n, int64_t *f2e, int64_t *bc_node, int64_t *bc_edge,
                    int64_t *bc_face, int64_t *pairs, int64_t *insert, int64_t *layer_cells,
                    int64_t *layer_nodes);
int pyib_mesh_free(void *handle);

/* ---- structured layout of the elliptic operator on the host (setup; our design, no
 * counterpart in the reference: it renumbers the mesh of make_euler_mesh.py:330-463 into
 * per-level dense boxes and assembles, from the IBCore geometry of ib_core_np.py:24-126, the
 * rows of A = rot.integr.V.interp.curl (ib_core_np.py:168-175) that are not constant stencils).
 * lev_i64 (nlevels,13): nx, ny, nz, hole lo (3), hole hi (3), cell offset, x/y/z edge-block
 * offsets; lev_f64 (nlevels,4): h, box corner (3); layer_cells (nlevels).  The mesh arrays are
 * nodes (nnode,3), edges_to_nodes (nedge,2), tangents te (nedge,3); the geometry vol_ci (ncell),
 * faces_to_cells (nface,2), nf/rf1/rf2 (nface,3), AoL/Af (nface); the gather tables those of
 * pyib_tables.  counts[0..8] = G1 rows, G1 entries, V rows, V entries, G3 rows, G3 entries,
 * regular V cells, regular G1 cells, regular G3 edges.  pyib_layout_fetch copies out cell_map
 * (ncell), edge_map/edge_sigma (nedge), edge_inv/sigma_int/reg_edge (internal edges),
 * reg_cell/reg_g1 (internal cells), then rows/ptr/col/coef of G1 (coef x3), V (x1), G3 (x3);
 * a NULL output is skipped.  pyib_layout_free releases the handle. */
int pyib_layout_build(int32_t nlevels, const int64_t *lev_i64, const double *lev_f64, const int64_t *layer_cells,
                      int64_t ncell_i, int64_t nedge_i, int64_t nnode, const double *nodes, int64_t nedge,
                      const int64_t *e2n, const double *te, int64_t ncell, int64_t nface, const double *vol_ci,
                      const int64_t *f2c, const double *nf, const double *rf1, const double *rf2, const double *aol,
                      const double *af, const int32_t *curl_ptr, const int32_t *curl_idx, const int32_t *rot_ptr,
                      const int32_t *rot_idx, const int32_t *cf_ptr, const int32_t *cf_mid, const int32_t *cf_face,
                      const int32_t *cf_other, int32_t margin, double tol, int32_t demote_neighbours,
                      void **handle, int64_t *counts);
int pyib_layout_fetch(void *handle, int64_t *cell_map, int64_t *edge_map, int8_t *edge_sigma, int64_t *edge_inv,
                      int8_t *sigma_int, uint8_t *reg_cell, uint8_t *reg_edge, uint8_t *reg_g1, int32_t *g1_rows,
                      int32_t *g1_ptr, int32_t *g1_col, double *g1_coef, int32_t *v_rows, int32_t *v_ptr,
                      int32_t *v_col, double *v_coef, int32_t *g3_rows, int32_t *g3_ptr, int32_t *g3_col,
                      double *g3_coef);
int pyib_layout_free(void *handle);

/* ---- the same CG on the structured layout (ib_driver_np.py:311-410) ----
 * Workspace cell_a/cell_b hold 3*ncell (internal) values; partials >= 4*PYIB_MAX_PARTIALS.
 * flow_mat: w = A p (0 on dir; a caller that drops the Dirichlet rows from the G3 composite
 * zero-initialises w).  If dot_out is non-NULL, *dot_out (device) receives
 * p.A p = sum_cells vol x.(V x), x = interp(curl p), in double (deterministic): p.w when
 * p is zero on the dir rows. */
int pyib_grid_flow_mat(const pyib_grid *g, const pyib_work *ws, double rou, double visc, double hdt,
                       const void *p, void *w, double *dot_out, void *stream);
int pyib_grid_cg_begin(const pyib_grid *g, const pyib_work *ws, const pyib_cg *cg, const void *res,
                       const void *s, double tol1, double tol2, void *stream);
/* n_iter iterations as pyib_cg_iterate.  cg->w is not used: r -= alpha A p is fused into the
 * last operator pass, and x += alpha p into the direction update (done alone when the last
 * iteration skips the direction, so x is current for the host's stop test). */
int pyib_grid_cg_iterate(const pyib_grid *g, const pyib_work *ws, const pyib_cg *cg, double rou,
                         double visc, double hdt, int32_t n_iter, int32_t skip_last_direction,
                         void *stream);
/* p = r*prec + beta*p; eta = eta_new */
int pyib_grid_cg_direction(const pyib_grid *g, const pyib_cg *cg, void *stream);
int pyib_grid_cg_true_residual(const pyib_grid *g, const pyib_work *ws, const pyib_cg *cg,
                               const void *res, const void *s, void *tmp, void *tr, double rou,
                               double visc, double hdt, double *out2, void *stream);
/* out[i] = idx[i] >= 0 ? sigma[i]*in[idx[i]] : 0 (renumbering between PyIB's and the layout's order) */
int pyib_permute(int dtype_in, const void *in, int dtype_out, void *out, int64_t n_out,
                 const int64_t *idx, const int8_t *sigma, void *stream);

/* ---- adjoints (transposes) of the time-step operators, float64 ----
 * face_ghost: nface, ghost index k of boundary face f or -1. */
/* Driver.calc_vel^T: folds the ghost rows of g_vc into their owners (g_vc is
 * modified), then g_U = interp^T g_vc + ghost terms + g_U_in, g_s = curl^T g_U
 * with the inlet edges zeroed. */
int pyib_calc_vel_adjoint(const pyib_mesh *m, const int32_t *face_ghost, double *g_vc,
                          const double *g_U_in, double *g_U, double *g_s, void *stream);
/* pyib_curl_rhs^T: g_res_c (ncell,3) = integr_c2f^T curl g_edge  (uses ws->face) */
int pyib_curl_rhs_adjoint(const pyib_mesh *m, const pyib_work *ws, const double *g_edge, double *g_res_c,
                          void *stream);
/* momentum residual^T: g_vc (ncell+nghost, 3) += d out/d vc^T g_out, g_U += d out/d U^T g_out */
int pyib_momentum_residual_adjoint(const pyib_mesh *m, const int32_t *face_ghost, const double *U,
                                   const double *vc, double visc, int has_base, double a_vc,
                                   double k_conv, double k_diff, const double *g_out, double *g_vc,
                                   double *g_U, void *stream);
/* d w / d marker for every (marker, tap) of pyib_ib_stencil: dw (nmark*27 x 3) */
int pyib_ib_dweights(int64_t nmark, const double *markers, double pos_ref_x, double pos_ref_y,
                     double pos_ref_z, double dx, double hh, int32_t nx, int32_t ny, int32_t nz,
                     double *dw, void *stream);
/* out[m] += scale[m]*scale_all*sum_t (field[cell].vec[m]) dw[m,t]  (scale may be NULL) */
int pyib_ib_marker_grad(int64_t nmark, const int32_t *cell, const double *dw, const double *field,
                        const double *vec, const double *scale, double scale_all, double *out,
                        void *stream);
/* out[m] = scale[m] * sum_t field[cell] w[m,t]  (spread^T) */
int pyib_ib_gather(int64_t nmark, const int32_t *cell, const double *w, const double *field,
                   const double *scale, double *out, void *stream);
/* g_F[idx] += g * vol  (force_fluid^T) */
int pyib_fluid_load_adjoint(int64_t n, const int32_t *idx, const double *vol, double g0, double g1,
                            double g2, double *gF, void *stream);
/* linear interior momentum^T: g_vc[idx] += rho g wv; gcenter3 = d(P.g)/d center */
int pyib_interior_momentum_adjoint(int64_t n, const int32_t *idx, const double *pos, const double *vol,
                                   const double *vc, double cx, double cy, double cz, double radius,
                                   double width, double rho, double g0, double g1, double g2,
                                   double *gvc, double *partials, double *gcenter3, void *stream);

/* ---- mixed precision helpers ---- */
/* Residual replacement in a CG state stopped at a convergence check (direction not yet
 * updated): scal[eta_new] <- *eta (device), scal[beta] <- *eta / scal[eta].  The caller has
 * written the replaced residual into cg->r; the next pyib_*cg_direction continues the CG
 * with the kept search direction.  No reference counterpart (mixed precision is additive). */
int pyib_cg_replace_residual(double *scal, const double *eta, void *stream);
int pyib_convert(int dtype_in, const void *in, int dtype_out, void *out, int64_t n, void *stream);
/* y[i] += a[i] (a in dtype_a, y double) */
int pyib_add_into_f64(int dtype_a, const void *a, double *y, int64_t n, void *stream);

/* ---- generic vector helpers (double reductions) ---- */
int pyib_dot(int dtype, const void *a, const void *b, const void *c, int64_t n, double *partials,
             double *out, void *stream); /* sum a*b (*c if c) */
int pyib_axpby(int dtype, double a, const void *x, double b, const void *y, void *out, int64_t n,
               void *stream);

/* ---- immersed boundary (ib_immersed.py:12-120) ---- */
/* Roma-Peskin 3-point stencil of every marker: 27 taps (the 125-tap window of
 * the reference has zero weight outside the 3^3 core).  cell[m*27+t] = finest
 * cell id or -1 when outside the finest box; w[...] = delta product (double). */
int pyib_ib_stencil(int64_t nmark, const double *markers, double pos_ref_x, double pos_ref_y,
                    double pos_ref_z, double dx, double hh, int32_t nx, int32_t ny, int32_t nz,
                    int32_t *cell, double *w, void *stream);
/* vel[m] = sum_t vc[cell]*w*hh^3 (ImmersedBoundary._interp_stencil) */
int pyib_ib_interp(int dtype, int64_t nmark, const int32_t *cell, const double *w, double hh3,
                   const void *vc, double *vel, void *stream);
/* Deterministic spreading (ImmersedBoundary._spread_stencil): the (cell, tap)
 * pairs are radix-sorted by cell (stable), then every cell's contributions are
 * summed in marker order, as NumPy's bincount does.  keys/vals are caller
 * workspaces of nmark*27 entries; temp is the radix-sort scratch of
 * pyib_ib_sort_temp_bytes bytes.  pyib_ib_spread writes scale*sum into the
 * touched cells of out (ncell x 3), or adds it when accumulate != 0; cells no
 * marker touches are left as they are (the caller clears out first when it
 * wants the reference's fresh array). */
int pyib_ib_sort_temp_bytes(int64_t n, int64_t *bytes);
int pyib_ib_sort(int64_t nmark, const int32_t *cell, int32_t *keys_tmp, int32_t *vals_tmp,
                 int32_t *keys_sorted, int32_t *vals_sorted, void *temp, int64_t temp_bytes,
                 void *stream);
/* Allocation-free variant of pyib_ib_sort.  It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  The key/value staging and the radix-sort storage live in a static device array of
 * PYIB_IB_NOALLOC_SCRATCH_BYTES bytes per GPU (held until the process ends); nmark may be at
 * most PYIB_IB_NOALLOC_MAX_MARKERS (larger counts fail with a message and launch nothing).
 * The caller passes only the inputs and the sorted outputs, which equal pyib_ib_sort's.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two threads) race and must be serialised by the caller. */
#define PYIB_IB_NOALLOC_MAX_MARKERS 16384
#define PYIB_IB_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)
int pyib_ib_sort_noalloc(int64_t nmark, const int32_t *cell, int32_t *keys_sorted, int32_t *vals_sorted,
                         void *stream);
int pyib_ib_spread(int dtype, int64_t nmark, const int32_t *keys_sorted, const int32_t *vals_sorted,
                   const double *w, const double *flag, const double *ds3, double scale,
                   int accumulate, void *out, void *stream);
/* Implicit forcing's marker coupling matrix (ImmersedBoundary.prepare_implicit,
 * ib_immersed.py:122-145): M (nmark x nmark, row-major, overwritten) with
 * M[i][j] = hh3 * sum over shared cells of w_i w_j ds3[j], from the stencil (cell, w) of
 * pyib_ib_stencil and the cell-sorted taps of pyib_ib_sort.  Without the diagonal jitter. */
int pyib_ib_coupling(int64_t nmark, const int32_t *cell, const double *w, const int32_t *keys_sorted,
                     const int32_t *vals_sorted, const double *ds3, double hh3, double *M, void *stream);
/* f_L = (vel_desired - interp)/dt (ImmersedBoundary.direct_force) */
int pyib_ib_force(int64_t nmark, const double *vel_desired, const double *vel_ib, double dt,
                  double *f_lag, void *stream);

/* ---- FP64 reductions of the closed-body load correction (ib_interior.py:57-70, 291-296) ---- */
/* over candidate cells idx (pos/vol are per candidate): sums of v*wv, (pos x v)*wv
 * and wv with wv the smoothed-sphere indicator times volume (or fixed_wv when
 * fixed != 0); out7 = [lin(3), ang(3), vol], density applied by the caller */
int pyib_interior_momenta(int dtype, int64_t n, const int32_t *idx, const double *pos, const double *vol,
                          const void *vc, double cx, double cy, double cz, double radius, double width,
                          double rho, int fixed, const double *fixed_wv, double *partials, double *out7,
                          void *stream);
/* out6 = [sum F*vol (3), sum (pos x F)*vol (3)] over force-support cells idx */
int pyib_fluid_load(int dtype, int64_t n, const int32_t *idx, const double *pos, const double *vol,
                    const void *F, double *partials, double *out6, void *stream);
/* diagnostics (ib_case.py:1147-1151): out2[0] = max |vol_ci*cdiv(U)|, out2[1] = max |vc|
 * over volume cells; tmp_cell holds >= ncell + 2*PYIB_MAX_PARTIALS doubles */
int pyib_diag_max(const pyib_mesh *m, const void *U, const void *vc, void *tmp_cell, double *out2,
                  void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PYIB_HIP_H */
