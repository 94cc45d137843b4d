/*
 * tfs_oracle.h -- CPU restatement of the T-FlowS linear-solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: the GPU product is checked
 * against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; nothing in the product links or calls it.
 *
 * Indexing: all arrays are 0-based here.  row[n+1] holds row starts
 * (row[n] == nnz), col[nnz] sorted ascending within each row, dia[n] the
 * position of the diagonal entry, val[nnz] the FP64 coefficients.  This is the
 * reference Matrix_Type (Sources/Process/Matrix_Mod.f90:36-49) shifted by one.
 *
 * Arithmetic follows the reference loop by loop (left-to-right sums in row
 * order, no fused multiply-add), so on x86-64 the results are bitwise equal to
 * the reference compiled with default floating-point semantics.
 */
#ifndef TFS_ORACLE_H
#define TFS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n;
  const int *row;
  const int *col;
  const int *dia;
  const double *val;
} or_matrix;

/* Sources/Process/Solvers_Mod/Prec_Form.f90:2-51 -> d[n] (the used diagonal of D). */
void or_prec_form(const or_matrix *A, int prec, double *d);

/* Sources/Process/Solvers_Mod/Prec_Solve.f90:2-78: x := M^-1 b. */
void or_prec_solve(const or_matrix *A, const double *d, int prec, double *x,
                   const double *b);

/* Sources/Process/Solvers_Mod/Residual.f90:22-27: r := r - A x (in place). */
void or_residual(const or_matrix *A, const double *x, double *r);

/* CRS product y := A x, summed as in Sources/Process/Solvers_Mod/Cg.f90:128-134. */
void or_spmv(const or_matrix *A, const double *x, double *y);

/* Krylov solvers with the reference signature (Cg.f90:2-4, Bicg.f90:2-4,
 * Cgs.f90:2-4).  x in/out, r1 in = b / out = final residual, niter in = cap /
 * out = count.  ini_res is only written where the reference writes it.
 * On the ||b|| < tol early exit the reference returns an uninitialised
 * `error` as fin_res; this restatement returns ||b|| there.
 * Returns 0. */
int or_cg(const or_matrix *A, double *x, double *r1, int prec, int *niter,
          double tol, double *ini_res, double *fin_res);
int or_bicg(const or_matrix *A, double *x, double *r1, int prec, int *niter,
            double tol, double *ini_res, double *fin_res);
int or_cgs(const or_matrix *A, double *x, double *r1, int prec, int *niter,
           double tol, double *ini_res, double *fin_res);

/* Sources/Process/Matrix_Mod/Topology.f90:33-121.  faces_c is (2, n_faces)
 * column-major with 1-based cell numbers (c2 <= 0 marks a boundary face), as
 * the reference Grid_Type holds it.  Outputs are 1-based as in the reference:
 * row[n+1], col[nnz], dia[n], mir[nnz], pos[2*n_faces] (0 where unset).
 * Returns nnz (without the reference's one spare entry). */
int or_topology_nnz(int n_cells, int n_faces, const int *faces_c);
void or_topology(int n_cells, int n_faces, const int *faces_c, int *row,
                 int *col, int *dia, int *mir, int *pos);

/* The face-loop assembly of A % val (Sources/Process/Pressure_Matrix_Fractional.f90:19-48;
 * Sources/Process/NewUVW.f90:372-384 faces, :500-506 inertia, :586-590
 * under-relaxation).  faces_c / pos / dia 1-based as the reference holds them
 * (pos = Matrix_Mod_Topology's pos(1:2, s)); a21 NULL means a21 = a12; a
 * boundary face adds a12 to its cell's diagonal when bnd_on is NULL or
 * bnd_on[s] != 0; cell NULL adds nothing; sav NULL is not written.
 * val[nnz] and sav[n] are 0-based outputs. */
void or_assemble(int n_cells, int n_faces, const int *faces_c, const int *pos, const int *dia,
                 int nnz, const double *a12, const double *a21, const unsigned char *bnd_on,
                 const double *cell, double urf, double *val, double *sav);

/* Least-squares gradients (Sources/Process/Compute_Gradient_Matrix.f90,
 * GraPhi.f90, GradP.f90), Boundary = .TRUE. as every call site passes it and
 * every material a fluid.  Cells 0..n_cells-1, boundary cells -1..-n_bnd at
 * n_cells + (-c2 - 1) in phi / xc; d = the face vectors after
 * Compute_Geometry in three blocks of n_faces (dx, dy, dz); g = G(1:6, c)
 * column-major, 6 per cell.  bc[b] is the kind of boundary cell -(b+1):
 * OR_BC_* below (NULL: every boundary a wall). */
enum { OR_BC_WALL = 0, OR_BC_SYMMETRY = 1, OR_BC_PRESSURE = 2, OR_BC_BUFFER = 3 };
void or_gradient_matrix(int n_cells, int n_faces, const int *faces_c, const double *d,
                        double *g);
void or_graphi(int n_cells, int n_faces, const int *faces_c, const double *d, const double *g,
               const double *phi, int i, const unsigned char *bc, double *phii);
void or_gradp(int n_cells, int n_faces, const int *faces_c, const double *d, const double *g,
              const double *xc, int n_bnd, const unsigned char *bc, double *phi, double *px,
              double *py, double *pz);

/* Correct_Velocity (Sources/Process/Correct_Velocity.f90:34-135) for one
 * material: vol, sav, px/py/pz, den per cell (den = DENc(material(c)));
 * u, v, w, pp on cells and boundary cells; coef per face = A % val(A %
 * pos(1,s)); bou per boundary cell (A % bou; NULL = zeros); in_m per face
 * (material(c1) or material(c2) is the material reported; NULL = every
 * face).  Updates u, v, w, flux; writes b (the mass error per cell);
 * returns errmax, with the CFL and Peclet maxima in *cfl, *pe. */
double or_correct_velocity(int n_cells, int n_faces, const int *faces_c, const double *vol,
                           const double *sav, const double *px, const double *py,
                           const double *pz, double *u, double *v, double *w, const double *pp,
                           const double *coef, const double *bou, const unsigned char *bc,
                           const double *den, const unsigned char *in_m, const double *scoef,
                           const double *d, double dt, double visc, double *flux, double *b,
                           double *cfl, double *pe);

/* The face loop of Compute_Pressure_Fractional (Sources/Process/
 * Compute_Pressure_Fractional.f90:47-131), every material a fluid, no BUFFER
 * faces: per face the density (f-weighted), the face velocity rebuilt from
 * the cells' velocities and pressure gradients, the face flux and the
 * pressure-matrix coefficient A12 (0 on boundary faces, which add nothing to
 * the matrix); per cell b = inflow - outflow in face order.  f, s (3 blocks)
 * per face; vol, sav, px/py/pz per cell; u/v/w, p, den on cells then
 * boundary cells.  Outputs flux, a12 (per face), b (per cell). */
void or_pressure_fractional(int n_cells, int n_faces, const int *faces_c, const double *f,
                            const double *s, const double *vol, const double *sav,
                            const double *u, const double *v, const double *w, const double *px,
                            const double *py, const double *pz, const double *p,
                            const double *den, double *flux, double *a12, double *b);

/* The face loop of Compute_Pressure_Simple (Sources/Process/
 * Compute_Pressure_Simple.f90:45-180), every material a fluid, no BUFFER
 * faces: A % val and b cleared; per interior face the f-weighted density,
 * SMDPN = |S|^2 / (S . d), the interpolated velocity, A12 = 0.5 DENs SMDPN
 * (vol1/sav1 + vol2/sav2) with the off-diagonals set to -A12 and both
 * diagonals increased, the interpolated pressure gradient and the face flux;
 * per boundary face by the kind of its boundary cell (bc: 0 wall,
 * 1 symmetry, 2 pressure, 4 inflow, 5 outflow, 6 convective outflow): the
 * flux from the boundary cell's velocity (inflow, outflow, convective) or the
 * inside cell's (pressure), with A12 = DENs SMDPN vol1/sav1 on the diagonal
 * for outflow, convective and pressure boundaries, and no flux on walls and
 * symmetry planes; b(c1) -= flux, b(c2) += flux in face order.  f, s and d
 * (3 blocks each) per face; vol, sav, px/py/pz per cell; u/v/w, p, den on
 * cells then boundary cells; pos (2 per face), dia (1-based), nnz.  Outputs
 * flux, a12 (the coefficient each face adds to the matrix, 0 for the faces
 * that add none), bnd_on (1 for a boundary face that adds to its diagonal),
 * val (nnz), b (per cell). */
void or_pressure_simple(int n_cells, int n_faces, const int *faces_c, const double *f,
                        const double *s, const double *d, const double *vol, const double *sav,
                        const double *u, const double *v, const double *w, const double *p,
                        const double *px, const double *py, const double *pz, const double *den,
                        const unsigned char *bc, const int *pos, const int *dia, int nnz,
                        double *flux, double *a12, unsigned char *bnd_on, double *val, double *b);

#ifdef __cplusplus
}
#endif

#endif
