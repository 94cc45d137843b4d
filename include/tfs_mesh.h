/*
 * tfs_mesh.h -- host-side C ABI of libtfs_mesh.so: CRS topology from a face
 * list, the synthetic hex-mesh cases the solver is measured on, and the
 * reference's own mesh files (.cns / .geo) with the pressure matrix on them.
 *
 * The topology builder is the counterpart of the reference's
 * Matrix_Mod_Topology (Sources/Process/Matrix_Mod/Topology.f90:2-127): same
 * layout (columns ascending within a row, diagonal position in dia, face ->
 * entry map in pos), selectable 0- or 1-based output.
 *
 * The case generator builds a structured hex box (tanh clustering, node
 * jitter, random cell numbering) and assembles the two matrix families the
 * solver serves:
 *   TFS_CASE_POISSON  -- the fractional-step pressure matrix,
 *                        Sources/Process/Pressure_Matrix_Fractional.f90:26-48
 *                        with Scoef from Sources/Process/Compute_Geometry.f90:22-49;
 *   TFS_CASE_MOMENTUM -- fully implicit upwind convection-diffusion with an
 *                        inertia diagonal, Sources/Process/NewUVW.f90:349-384,500-506.
 * All random draws come from SplitMix64 (counter form), so every process and
 * every language sees identical inputs for a given seed.
 */
#ifndef TFS_MESH_H
#define TFS_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers ------------------------------------------------------ */

/* Uniform in [0,1) from 53 bits of SplitMix64: the i-th draw (i >= 0) of the
 * stream keyed by (seed, tag). */
double tfs_rng_u01(uint64_t seed, uint64_t tag, uint64_t i);

/* ---- topology (Matrix_Mod_Topology) -------------------------------------- */

/* Number of stored entries for faces_c (2 x n_faces, column-major, 1-based
 * cells, c2 <= 0 marks a boundary face).  Topology.f90:33-48. */
int64_t tfs_topology_nnz(int n_cells, int64_t n_faces, const int *faces_c);

/* Fills row[n+1], col[nnz], dia[n] and (if non-NULL) pos[2*n_faces] with
 * index base `base` (0 or 1).  pos entries of boundary faces are base-1.
 * Returns 0, or a negative value on invalid input (bad cell number, nnz
 * beyond the int32 range the reference's default integers hold).
 * Topology.f90:54-121. */
int tfs_topology_build(int n_cells, int64_t n_faces, const int *faces_c,
                       int base, int *row, int *col, int *dia, int *pos);

/* ---- synthetic cases ----------------------------------------------------- */

enum { TFS_CASE_POISSON = 0, TFS_CASE_MOMENTUM = 1 };

typedef struct {
  int kind;            /* TFS_CASE_POISSON or TFS_CASE_MOMENTUM              */
  int nx, ny, nz;      /* cells per axis                                      */
  double beta[3];      /* tanh clustering per axis, 0 = uniform               */
  double jitter;       /* node jitter, fraction of local spacing (0.1 = 10%)  */
  int permute;         /* 1 = random cell numbering                           */
  int pin_first;       /* Poisson: double a_11 to fix the null space          */
  int rhs_zero_mean;   /* remove the mean of the random RHS                   */
  double nu;           /* momentum: kinematic viscosity                       */
  double cfl;          /* momentum: dt = cfl * h_min / |u|max                 */
  int n_modes;         /* momentum: Fourier modes of the vector potential     */
  uint64_t seed;
} tfs_case_params;

typedef struct tfs_case tfs_case;

int tfs_case_create(const tfs_case_params *p, tfs_case **out);
void tfs_case_destroy(tfs_case *c);
/* n_cells, n_faces (interior + boundary), nnz */
void tfs_case_sizes(const tfs_case *c, int *n_cells, int64_t *n_faces, int64_t *nnz);
/* CRS arrays with index base `base`, RHS b[n]. */
int tfs_case_crs(const tfs_case *c, int base, int *row, int *col, int *dia,
                 double *val, double *b);
/* Face list as the reference Grid_Type holds it: faces_c[2*n_faces],
 * 1-based cells, negative c2 for boundary faces. */
int tfs_case_faces(const tfs_case *c, int *faces_c);
/* The assembly inputs: per face the coefficient of c2 in row c1 (a12; for a
 * boundary face the diagonal addition of c1) and of c1 in row c2 (a21); per
 * cell the extra diagonal (inertia).  val = scatter through pos/dia in face
 * order, then the extra diagonal, then (Poisson, pin_first) a_11 doubled. */
int tfs_case_face_coefs(const tfs_case *c, double *a12, double *a21, double *dextra);
/* Time step used for the momentum inertia term (0 for Poisson). */
double tfs_case_dt(const tfs_case *c);
/* Largest cell aspect ratio (max edge / min edge over the cell's 12 edges). */
double tfs_case_max_aspect(const tfs_case *c);

/* ---- the solver's host-side layouts (no device; for inspection and tests) ----
 * Greedy colouring of the symmetrised pattern in index order: colour[i] per
 * row, perm = rows sorted by colour (index order within a colour); returns
 * the number of colours.  The FP32 inner solves of the refinement run in
 * this order (S-order rows as input there). */
int tfs_layout_colour(int n, const int *row, const int *col, int *perm, int *colour);
/* The same with runs of `block` consecutive rows coloured as units. */
int tfs_layout_block_colour(int n, const int *row, const int *col, int block, int *perm,
                            int *colour);
/* The block multi-colour order of the FP32 inner solves: blocks of at most
 * `block` rows grown breadth-first (seeds in index order, neighbours in
 * ascending index), the block graph coloured greedily in growth order, perm =
 * rows by colour, then block, then discovery order; colour[i] and block_of[i]
 * per row.  Returns the number of colours. */
int tfs_layout_block_colour_bfs(int n, const int *row, const int *col, int block, int *perm,
                                int *colour, int *block_of);
/* The brick order of the FP32 inner solves (TFS_INNER_BRICKS; 0-based CRS in
 * the caller's order, columns ascending, dia the diagonal's entry): rows
 * grouped by their monotone potentials divided by `edge` (on a
 * lexicographically numbered grid, edge^3 bricks), groups over `cap` rows
 * split, blocks coloured greedily; perm = rows by colour, then block, then
 * caller order; colour[i] and block_of[i] per row (blocks numbered in perm
 * order).  Returns the number of colours. */
int tfs_layout_brick_colour(int n, const int *row, const int *col, const int *dia, int edge,
                            int cap, int *perm, int *colour, int *block_of);
/* The layouts the solver builds for a matrix (0-based CRS): s_orig = caller
 * row at each S (reverse Cuthill-McKee) position; level of each caller row in
 * the lower dependency DAG of the caller's numbering (forward sweep) and in
 * the backward sweep's order (the lower levels counted from the deepest, so
 * every upper dependency of a row sits at a smaller level); bandwidth of the
 * pattern in S order. */
int tfs_layout_summary(int n, const int *row, const int *col, const int *dia, int *s_orig,
                       int *level_lower, int *level_upper, int *bandwidth);

/* ---- T-FlowS mesh files -------------------------------------------------
 * The .cns (connectivity) and .geo (geometry) files the reference's mesh
 * programs write (Sources/Shared/Save_Cns_Geo.f90) and its Process reads:
 * Fortran unformatted sequential records, 4-byte integers, 8-byte reals.
 * Cells are 1..n_cells; boundary cells -1..-n_bnd_cells.  Status: */
enum {
  TFS_MESH_OK = 0,
  TFS_MESH_ERR_INVALID = -1, /* bad argument or inconsistent indices        */
  TFS_MESH_ERR_IO = -2,      /* the file cannot be opened                   */
  TFS_MESH_ERR_FORMAT = -3   /* a record is missing or of the wrong length  */
};
/* Reads a .cns file (replaces Sources/Process/Load_Cns.f90:25-69).  Always
 * returns the sizes; with faces_c non-NULL also the face list
 * faces_c[2*n_faces] as (c1, c2) pairs -- the reference's faces_c(2,n_faces)
 * in column-major order, c2 negative for boundary faces -- and with material
 * non-NULL the material of the cells, then of the boundary cells
 * (n_cells + n_bnd_cells entries). */
int tfs_mesh_read_cns(const char *path, int *n_cells, int *n_bnd_cells, int *n_faces,
                      int *faces_c, int *material);
/* The records after the faces (Sources/Process/Load_Cns.f90:62-69): bcmark[b]
 * = the boundary-condition number of boundary cell -(b+1) (the index of its
 * name in the file; the caller maps it to a kind), copy_c[b] = its copy cell
 * (0 where it has none).  n_bnd_cells as tfs_mesh_read_cns returned it; either
 * output may be NULL. */
int tfs_mesh_read_cns_marks(const char *path, int n_bnd_cells, int *bcmark, int *copy_c);

/* Reads a .geo file (replaces Sources/Shared/Load_Geo.f90:23-50) for the sizes
 * of its .cns.  Outputs (each may be NULL), x, y and z stored one block after
 * the other:
 *   xc    3*(n_cells+n_bnd_cells)  cell centres: cells, then boundary cells
 *                                  -1..-n_bnd (the boundary face centres)
 *   vol   n_cells                  cell volumes
 *   s     3*n_faces                face surface vectors (pointing c1 -> c2)
 *   shift 3*n_faces                periodic shift of c2 (zero off periodic faces)
 *   f     n_faces                  interpolation weights
 *   xf    3*n_faces                face centres */
int tfs_mesh_read_geo(const char *path, int n_cells, int n_bnd_cells, int n_faces, double *xc,
                      double *vol, double *s, double *shift, double *f, double *xf);
/* Face coefficients of Sources/Process/Compute_Geometry.f90:22-49:
 * scoef[s] = |S|^2 / (d . S), d = centre of c2 (+ shift) - centre of c1;
 * d (3*n_faces, may be NULL) receives the distance vectors. */
int tfs_mesh_face_coef(int n_cells, int n_bnd_cells, int n_faces, const int *faces_c,
                       const double *xc, const double *s, const double *shift, double *d,
                       double *scoef);
/* The fractional-step pressure matrix values of
 * Sources/Process/Pressure_Matrix_Fractional.f90:19-48 on a topology from
 * tfs_topology_build (pos, dia with index base `base`): val[nnz] := 0; each
 * interior face sets its two off-diagonals to -dt scoef and adds dt scoef to
 * both diagonals, in face order; a boundary face adds dt scoef to its cell's
 * diagonal where bnd_on[s] is set (NULL: nowhere -- walls and symmetry, the
 * sequential case). */
int tfs_mesh_pressure_values(int n_cells, int64_t n_faces, const int *faces_c, const int *pos,
                             const int *dia, int base, int64_t nnz, const double *scoef,
                             double dt, const unsigned char *bnd_on, double *val);

#ifdef __cplusplus
}
#endif

#endif
