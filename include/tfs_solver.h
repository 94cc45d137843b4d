/*
 * tfs_solver.h -- C ABI of libtfs_solver.so, the MI355X (gfx950) drop-in for
 * T-FlowS's preconditioned Krylov solvers.
 *
 * Reference interface replaced (all paths relative to the T-FlowS tree):
 *   subroutine Cg  (A, x, r1, prec, niter, tol, ini_res, fin_res)
 *                                  Sources/Process/Solvers_Mod/Cg.f90:2-4,24-29
 *   subroutine Bicg(same)          Sources/Process/Solvers_Mod/Bicg.f90:2-4
 *   subroutine Cgs (same)          Sources/Process/Solvers_Mod/Cgs.f90:2-4
 *   type(Matrix_Type) A            Sources/Process/Matrix_Mod.f90:36-49
 *   Prec_Form / Prec_Solve         Sources/Process/Solvers_Mod/Prec_Form.f90:2-51,
 *                                  Sources/Process/Solvers_Mod/Prec_Solve.f90:2-78
 *   Residual                       Sources/Process/Solvers_Mod/Residual.f90:2-40
 *
 * Semantics kept from the reference:
 *   - tol is ABSOLUTE on ||r||_2 (Cg.f90:174);
 *   - niter: in = cap; out = k on convergence at iteration k, 0 on an early
 *     exit, cap+1 without convergence (Cg.f90:55-58,87-90,97,181);
 *   - x is updated in place from the caller's initial guess; r1 comes in as b
 *     and goes out as the final recurrence residual (Cg.f90:63,161);
 *   - ini_res is written only after the residual is formed (Cg.f90:85);
 *     on the ||b|| < tol exit fin_res = ||b|| (the reference leaves an
 *     uninitialised value there);
 *   - prec 0 = none, 1 = diagonal, 2 = the reference's diagonal incomplete
 *     Cholesky (Prec_Form.f90:32-40 with Prec_Solve.f90:46-67), applied
 *     exactly (same factor, same row-internal order) by level-scheduled sweeps;
 *   - Bicg equals Cg: the reference's shadow sequence (Bicg.f90:91-175) is
 *     bitwise equal to the primary one, so it is computed once.
 * Differences: dot products are reduced in a fixed tree instead of left to
 * right, so iterates agree with the reference to rounding, not bitwise.
 *
 * All entry points are synchronous (they return when the device work is
 * done) and are not reentrant for the same matrix handle.  Indices cross the
 * ABI as int32 (the reference's default integer); `base` selects 0- or
 * 1-based numbering.  Device pointers are HIP device addresses on the device
 * the library was initialised on.  Each matrix does its work on a stream of
 * its own that does not wait for other streams: device arrays passed in must
 * be complete when the call is made (the caller synchronises whatever wrote
 * them), and are complete when it returns.
 */
#ifndef TFS_SOLVER_H
#define TFS_SOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ---------------------------------------------------------------- */
enum {
  TFS_OK = 0,
  TFS_NOT_CONVERGED = 1,    /* niter came back as cap+1 (the reference's only signal) */
  TFS_CANCELLED = 2,        /* serving: the request was withdrawn before it ran       */
  TFS_ERR_INVALID = -1,     /* bad argument or inconsistent CRS structure            */
  TFS_ERR_NOMEM = -2,       /* device allocation failed                              */
  TFS_ERR_BREAKDOWN = -3,   /* rho, (p,q) or ||r|| became 0, Inf or NaN              */
  TFS_ERR_DEVICE = -4,      /* HIP runtime error                                      */
  TFS_ERR_NO_DEVICE = -5    /* no GPU visible                                         */
};
enum { TFS_CG = 0, TFS_BICG = 1, TFS_CGS = 2 };

const char *tfs_status_string(int status);
/* Library and device set-up; called implicitly by tfs_matrix_create. */
int tfs_init(int device);

/* ---- matrix: the Matrix_Type counterpart ------------------------------------
 * Structure (row/col/dia) is uploaded once, as it is fixed after
 * Matrix_Mod_Topology (Sources/Process/Matrix_Mod/Topology.f90:2-127); the
 * level schedules of the triangular sweeps are built from it here.  Values
 * change before every solve and are uploaded with tfs_matrix_set_values,
 * which invalidates the cached preconditioner factor. */
typedef struct tfs_matrix tfs_matrix;

int tfs_matrix_create(int n, long long nnz, const int *row, const int *col,
                      const int *dia, int base, tfs_matrix **out);
int tfs_matrix_destroy(tfs_matrix *A);
/* val[nnz] from host memory (the reference's A % val). */
int tfs_matrix_set_values(tfs_matrix *A, const double *val);
/* val[nnz] already on the device (copied device to device). */
int tfs_matrix_set_values_device(tfs_matrix *A, const double *d_val);
/* The caller rewrote the matrix's own device values in place
 * (tfs_matrix_values_device); the cached factor is dropped. */
/* Copies the caller-order values (nnz doubles) back to host memory. */
int tfs_matrix_get_values(tfs_matrix *A, double *val);
int tfs_matrix_values_changed(tfs_matrix *A);
/* Device address of the matrix's FP64 values (nnz entries). */
double *tfs_matrix_values_device(tfs_matrix *A);

/* Schedule of the preconditioner sweeps (bitwise the same results):
 *   TFS_SWEEP_LEVELS       one launch per level of the dependency DAG;
 *   TFS_SWEEP_STATIC       one persistent launch per sweep, workgroup w owning
 *                          row chunks w, w+G, ... (default where it times
 *                          faster than the levels); needs all G workgroups
 *                          resident, which an ordinary launch gets unless a
 *                          never-ending kernel holds CUs -- then the sweep's
 *                          bounded waits end it with TFS_ERR_DEVICE;
 *   TFS_SWEEP_DYNAMIC      one persistent launch taking chunks from a counter --
 *                          safe when several solves share the GPU concurrently;
 *   TFS_SWEEP_COOPERATIVE  the static schedule under a cooperative launch (the
 *                          runtime guarantees co-residency; slower to launch). */
enum {
  TFS_SWEEP_LEVELS = 0,
  TFS_SWEEP_STATIC = 1,
  TFS_SWEEP_DYNAMIC = 2,
  TFS_SWEEP_COOPERATIVE = 3
};
int tfs_matrix_set_sweep_schedule(tfs_matrix *A, int schedule);

typedef struct {
  int n;
  long long nnz;
  int levels_lower;     /* depth of the forward-sweep DAG                    */
  int levels_upper;     /* depth of the backward-sweep DAG                   */
  int row_blocks;       /* SpMV row blocks (wavefront work units)            */
  long long device_bytes;
  int inner_colours;    /* colours of the inner order (0 until first built)  */
  int spmv_rows;        /* SpMV shape: rows per row block (one wavefront)    */
  int spmv_waves;       /*             wavefronts per workgroup              */
  int inner_block;      /* rows per block of the inner order (1 = points)    */
  int inner_levels;     /* depth of the inner forward sweep (0 until built)  */
  int ell64_slices;     /* FP64 SpMV: sliced-ELL slices (0: CSR kernel)      */
  int inner_order;      /* the inner order built (TFS_INNER_*; -1 until then) */
  int sweep_groups;     /* DIC sweeps: workgroups of the persistent launch,
                           0 for one launch per level                        */
  int inner_sweep_groups; /* the same for the inner copy (-1 until built)    */
  int inner_brick_edge; /* TFS_INNER_BRICKS built: the brick edge, else 0    */
} tfs_matrix_info;
int tfs_matrix_get_info(const tfs_matrix *A, tfs_matrix_info *info);

/* ---- the solvers (Cg / Bicg / Cgs) -------------------------------------------
 * Host-memory variant: x1 = c_loc(x(1)) of x(-NB:N), r1 = b(1:N). */
int tfs_solve(tfs_matrix *A, int method, int prec, double *x1, double *r1,
              int *niter, double tol, double *ini_res, double *fin_res);
/* Device-memory variant: x and r1 are device arrays of n doubles. */
int tfs_solve_device(tfs_matrix *A, int method, int prec, double *d_x, double *d_r1,
                     int *niter, double tol, double *ini_res, double *fin_res);

/* ---- FP32-inner solve inside FP64 iterative refinement -------------------------
 * Not in the reference: values and inner Krylov vectors stream in FP32 while
 * the residual b - A x, the solution and every dot product stay FP64 on the
 * device.  The outer loop is set by tfs_matrix_set_refinement; by default
 * (TFS_REFINE_RELIABLE) one FP32 PCG recurrence (the prec factor recomputed
 * in FP32 from FP32 values) runs until its residual has fallen by
 * inner_rtol (0.1 if <= 0) against the last true residual, x += the FP64 sum
 * of its steps, r = b - A x in FP64, and the recurrence continues from that
 * r -- until ||r|| < tol (a segment that has not reduced its residual by
 * inner_rtol within 20000 FP32 iterations ends the solve, not converged).
 * niter: in = cap on refinement steps (residual replacements); out = steps
 * taken (cap+1 if tol was not reached, which also happens when FP32 stops
 * reducing the true residual); inner_iters = FP32 PCG iterations over all
 * steps; ini_res = ||b - A x0||; fin_res = ||b - A x|| (the true FP64
 * residual, left in r1).
 * x_rtol > 0 adds a stopping condition on the solution: the last update of
 * x must also satisfy ||dx|| <= x_rtol ||x||.  With reliable updates the FP32
 * recurrence cannot see residual components below FP32 resolution (the error
 * along the smallest eigenvectors), so once it has reached tol the
 * flexible-CG loop continues from its x to meet the criterion.  With the flexible-CG outer
 * loop (the default) this gives the FP64 solver's accuracy at the same tol
 * (x_rtol = tol / ||b||); defect correction's update omits the error it
 * cannot resolve in FP32 (along the smallest eigenvectors when
 * cond(A) * 2^-24 is not small), so there it is a weak check.  0 stops on the
 * residual alone. */
int tfs_solve_mixed_device(tfs_matrix *A, int prec, double *d_x, double *d_r1, int *niter,
                           int *inner_iters, double tol, double inner_rtol, double x_rtol,
                           double *ini_res, double *fin_res);
/* ---- SpMV shape ------------------------------------------------------------------
 * Not in the reference.  The SpMV works on row blocks of at most 64 rows and
 * 512 entries, one wavefront each, 2, 4 or 8 wavefronts per workgroup
 * (default 64 rows, 8 wavefronts).  Row sums are the same for every shape
 * (each row is summed left to right); the dot products fused into the SpMV
 * of the Krylov iterations are summed per workgroup, so their last-bit
 * rounding follows the shape.  tfs_matrix_tune_spmv times 16/32/64 rows x
 * 2/4/8 wavefronts and keeps the fastest (the batch does this at creation,
 * where no result depends on it); tfs_matrix_set_spmv_shape fixes a shape. */
int tfs_matrix_tune_spmv(tfs_matrix *A, int *rows_per_block, int *waves_per_group);
int tfs_matrix_set_spmv_shape(tfs_matrix *A, int rows_per_block, int waves_per_group);

/* Row order of the FP32 inner solves (the outer FP64 loop is unaffected):
 *   TFS_INNER_REFERENCE  the caller's order: the inner factor is the
 *                        reference's, its sweeps have the caller DAG's depth;
 *   TFS_INNER_COLOURED   (default) a greedy colouring of the pattern: rows of
 *                        one colour are independent, so each sweep has one
 *                        level per colour (2 for a hex mesh in lexicographic
 *                        order).  The inner preconditioner is then the same
 *                        incomplete Cholesky formula on the permuted matrix;
 *                        the refinement still converges to the FP64 tol;
 *   TFS_INNER_BRICKS     bricks of E^3 rows (on a lexicographically numbered
 *                        grid; in general blocks from the monotone potentials
 *                        of the caller's lower DAG) coloured as units, each
 *                        block in the caller's order inside: a stronger factor
 *                        than point colouring (fewer inner iterations) whose
 *                        sweeps are a few levels deep (4E - 2 for bricks in
 *                        two colours).  E = 16 from 3 * 2^22 rows, 8 from
 *                        3 * 2^21, 4 below;
 *   TFS_INNER_AUTO       (default) the bricks for systems of at least 2^22
 *                        rows whose bricks take two colours (measured, bricks
 *                        against point colouring per solve: 256^3 2.48
 *                        against 3.00 s, 192^3 0.82 against 0.99 s, 128^3 a
 *                        tie at 0.22 s), point colouring otherwise.
 * tfs_matrix_info.inner_order reports the order built (0, 1 or 2). */
enum {
  TFS_INNER_REFERENCE = 0,
  TFS_INNER_COLOURED = 1,
  TFS_INNER_BRICKS = 2,
  TFS_INNER_AUTO = 3
};
int tfs_matrix_set_inner_order(tfs_matrix *A, int order);
/* Block multi-colouring of the inner order: blocks of at most `rows` rows
 * grown breadth-first over the pattern are coloured as units, keeping their
 * couplings inside the factor (a stronger preconditioner than 1-row
 * colouring, with a block's internal depth of sweep levels per colour).
 * 1 = plain colouring. */
int tfs_matrix_set_inner_block(tfs_matrix *A, int rows);
/* Outer loop of the refinement:
 *   TFS_REFINE_RELIABLE (default) one FP32 CG recurrence with reliable FP64
 *                      updates: its directions are summed into x in FP64
 *                      every iteration, and whenever its residual has fallen
 *                      by inner_rtol the FP64 residual b - A x replaces it
 *                      (direction and rho kept, so nothing restarts).  The
 *                      FP32 operator is applied in difference form
 *                      (sum a_ij (x_j - x_i) + rowsum_i x_i), whose rounding
 *                      does not grow with cond(A): the iteration count stays
 *                      that of FP64 CG with the same preconditioner;
 *   TFS_REFINE_FCG     flexible CG in FP64 with each FP32 inner solve (to
 *                      inner_rtol, restarted per step) as its variable
 *                      preconditioner;
 *   TFS_REFINE_DEFECT  classical defect correction x += z, r = b - A x, which
 *                      contracts by about cond(A) * 2^-24 per step.
 * All stop on the true FP64 residual. */
enum { TFS_REFINE_FCG = 0, TFS_REFINE_DEFECT = 1, TFS_REFINE_RELIABLE = 2 };
int tfs_matrix_set_refinement(tfs_matrix *A, int outer);

/* ---- batched solves ----------------------------------------------------------------
 * Not in the reference (its callers solve one system at a time): many
 * independent systems of uneven size, e.g. ensemble members or Divide
 * subdomain blocks, solved together on one GPU.  Internally one
 * block-diagonal matrix; each system keeps its own scalars, tolerance, cap
 * and convergence, and one iteration of all systems is replayed from a
 * hipGraph.  System s: n[s] rows, nnz[s] entries, its own CRS arrays. */
typedef struct tfs_batch tfs_batch;
int tfs_batch_create(int nsys, const int *n, const long long *nnz, const int *const *row,
                     const int *const *col, const int *const *dia, int base, tfs_batch **out);
int tfs_batch_destroy(tfs_batch *B);
/* val of system `sys` (host memory, nnz[sys] doubles). */
int tfs_batch_set_values(tfs_batch *B, int sys, const double *val);
/* Cg semantics per system (x[s], r1[s]: device arrays of n[s] doubles); the
 * per-system outputs follow tfs_solve; returns the worst status. */
int tfs_batch_solve_device(tfs_batch *B, int prec, double *const *d_x, double *const *d_r1,
                           const int *cap, const double *tol, int *niter, double *ini_res,
                           double *fin_res, int *status);
/* FP32-inner refinement per system (tfs_solve_mixed_device semantics). */
int tfs_batch_solve_mixed_device(tfs_batch *B, int prec, double *const *d_x,
                                 double *const *d_r1, int cap, const double *tol,
                                 double inner_rtol, int *steps, int *inner_iters,
                                 double *ini_res, double *fin_res, int *status);
/* Row order of the batch's FP32 inner solves, as tfs_matrix_set_inner_order:
 * TFS_INNER_BRICKS (the default: each system in bricks from its own
 * numbering, the edge chosen by the system's size), TFS_INNER_COLOURED
 * (every system's rows grouped by colour) or TFS_INNER_REFERENCE; the systems
 * stay in their order, so a system's results do not depend on the others.
 * Changing it drops the inner copy (rebuilt on the next refinement solve). */
int tfs_batch_set_inner_order(tfs_batch *B, int order);
int tfs_batch_set_inner_block(tfs_batch *B, int rows);
/* Outer loop of the batch's refinement, as tfs_matrix_set_refinement
 * (default TFS_REFINE_RELIABLE: per system one FP32 CG recurrence on the
 * colour-ordered copy with reliable FP64 updates, in rounds in which every
 * unconverged system reduces its residual by inner_rtol; with
 * TFS_INNER_REFERENCE the batch runs flexible CG instead). */
int tfs_batch_set_refinement(tfs_batch *B, int outer);
/* spmv_rows / spmv_waves: the SpMV shape the batch's tuner picked at creation. */
int tfs_batch_get_info(const tfs_batch *B, int *nsys, int *n_total, long long *nnz_total,
                       int *graph_launches, int *spmv_rows, int *spmv_waves);

/* ---- device-side assembly of A % val from face coefficients ------------------
 * Replaces the face loops of Pressure_Matrix_Fractional.f90:19-48 and
 * NewUVW.f90:349-384,500-506,586-590 (scatter through Matrix_Mod/Topology.f90
 * pos/dia), so values built on the device need no upload:
 *   val := 0;  for each face s, in order: interior (c2 > 0):
 *     val(pos(1,s)) -= a12(s);  val(dia(c1)) += a12(s);
 *     val(pos(2,s)) -= a21(s);  val(dia(c2)) += a21(s);
 *   boundary (c2 <= 0, and bnd_on(s) != 0 if bnd_on is given):
 *     val(dia(c1)) += a12(s);
 *   for each cell: val(dia(c)) += cell(c);  sav(c) = val(dia(c));
 *                  val(dia(c)) /= urf.
 * The result is bitwise that of the sequential loop (each diagonal is summed
 * over its faces in face order).  faces_c = the reference's faces_c(1:2,s),
 * column-major (c1, c2 per face), pos = pos(1:2,s), dia as the topology built
 * them, all in `base`.  a21 NULL means a21 = a12; bnd_on, cell and sav may be
 * NULL; urf = 1 leaves the diagonal unscaled.  Device arrays: a12/a21 per
 * face, bnd_on (bytes) per face, cell/sav per cell.  The matrix's values are
 * replaced and marked changed, as after tfs_matrix_set_values_device. */
typedef struct tfs_assembly tfs_assembly;
int tfs_assembly_create(tfs_matrix *A, long long n_faces, const int *faces_c, const int *pos,
                        const int *dia, int base, tfs_assembly **out);
int tfs_assembly_destroy(tfs_assembly *S);
int tfs_assemble_device(tfs_assembly *S, const double *d_a12, const double *d_a21,
                        const unsigned char *d_bnd_on, const double *d_cell, double urf,
                        double *d_sav);

/* ---- topology on the device (replaces Matrix_Mod_Topology) -------------------
 * Sources/Process/Matrix_Mod/Topology.f90:2-127 (its row/col/dia/pos; mir is
 * not formed): the CRS structure of a face list built from device arrays.
 * d_faces_c = faces_c(1:2, n_faces) column-major, 1-based cells, c2 <= 0 a
 * boundary face (as tfs_mesh_read_cns returns it).  Output in index base
 * `base`: d_row[n_cells + 1], d_col[<= col_cap] (each row ascending),
 * d_dia[n_cells] (the last entry of row c equal to c), and, if d_pos is not
 * NULL, d_pos[2 n_faces] (the last position of c2 in row c1 and of c1 in row
 * c2; base - 1 for boundary faces).  Bitwise equal to tfs_topology_build.
 * d_faces_c and d_pos must be 8-byte aligned.  *nnz receives the entry count
 * (n_cells + 2 x interior faces); TFS_ERR_INVALID for a face with c1 outside
 * 1..n_cells or c2 above n_cells, or for nnz above col_cap or 2^31 - 2
 * (*nnz then says how many are needed).  Synchronous: runs on the calling
 * thread's default stream and returns when the arrays are complete. */
int tfs_topology_build_device(int n_cells, long long n_faces, const int *d_faces_c, int base,
                              long long col_cap, int *d_row, int *d_col, int *d_dia, int *d_pos,
                              long long *nnz);

/* ---- least-squares cell gradients ----------------------------------------------
 * Sources/Process/Compute_Gradient_Matrix.f90:21-116 (G = inverse of the sum
 * of d d^T over a cell's faces), GraPhi.f90:25-173 (the gradient of phi, all
 * three components in one call) and GradP.f90:24-64 (the pressure gradient
 * with the boundary values copied, extrapolated by 1.2 x the gradient, and the
 * gradient formed again), with Boundary = .TRUE. as every call site passes it
 * and every material a fluid.  Bitwise equal to the reference (each cell sums
 * its faces in face order).
 * tfs_gradient_create (host arrays, once per mesh): faces_c as
 * tfs_mesh_read_cns returns it (c1, c2 per face, 1-based, c2 = -1..-n_bnd for
 * boundary cells; each boundary cell on one face); d = the face vectors after
 * Compute_Geometry (tfs_mesh_face_coef), three blocks of n_faces (dx, dy,
 * dz); xc = cell centres, three blocks of n_cells + n_bnd (cells, then
 * boundary cells; needed by tfs_grad_p_device only, may be NULL otherwise);
 * bc = the kind of each boundary cell (TFS_BC_*; NULL = walls).  G is formed
 * on the device at creation.
 * Device arrays: phi has n_cells + n_bnd entries (cells, then boundary cells);
 * px, py, pz n_cells each.  tfs_grad_p_device rewrites phi's boundary part,
 * as GradP does.  Synchronous, on the calling thread's default stream. */
enum {
  TFS_BC_WALL = 0, TFS_BC_SYMMETRY = 1, TFS_BC_PRESSURE = 2, TFS_BC_BUFFER = 3,
  /* as walls for the gradients and the velocity correction (which single out
     symmetry, pressure and buffer cells only); told apart by the SIMPLE face
     loop */
  TFS_BC_INFLOW = 4, TFS_BC_OUTFLOW = 5, TFS_BC_CONVECT = 6
};
typedef struct tfs_gradient tfs_gradient;
int tfs_gradient_create(int n_cells, int n_bnd_cells, long long n_faces, const int *faces_c,
                        const double *d, const double *xc, const unsigned char *bc,
                        tfs_gradient **out);
int tfs_gradient_destroy(tfs_gradient *S);
/* host copy of G(1:6, c), six per cell (the reference's column-major G) */
int tfs_gradient_matrix_get(tfs_gradient *S, double *g);
int tfs_grad_phi_device(tfs_gradient *S, const double *d_phi, double *d_px, double *d_py,
                        double *d_pz);
int tfs_grad_p_device(tfs_gradient *S, double *d_phi, double *d_px, double *d_py, double *d_pz);

/* Correct_Velocity (Sources/Process/Correct_Velocity.f90:34-135) on the mesh of
 * a tfs_gradient object (its faces, face vectors and boundary kinds), for one
 * material: u, v, w (cells then boundary cells) -= (px, py, pz) vol / sav;
 * PRESSURE boundary cells take the inside velocity; flux(s) += (pp(c2) -
 * pp(c1)) coef(s) on interior faces (bou of the boundary cell on BUFFER
 * faces; d_bou NULL = 0); b(c) = (incoming - outgoing fluxes, summed in face
 * order) / (vol den); *errmax = max |b|; *cfl_max, *pe_max = the maxima of
 * |dt flux / (Scoef |d|^2)| and |flux / Scoef / (visc + TINY)| over the
 * interior and BUFFER faces with d_in_m[s] != 0 (d_in_m NULL = all faces).
 * coef(s) = A % val(A % pos(1,s)); den(c) = DENc(material(c)).  Per cell:
 * vol, sav, px, py, pz, den, b; per face: coef, scoef, flux, in_m (bytes).
 * Bitwise equal to the reference for the arrays and errmax. */
int tfs_correct_velocity_device(tfs_gradient *S, const double *d_vol, const double *d_sav,
                                const double *d_px, const double *d_py, const double *d_pz,
                                double *d_u, double *d_v, double *d_w, const double *d_pp,
                                const double *d_coef, const double *d_bou, const double *d_den,
                                const unsigned char *d_in_m, const double *d_scoef, double dt,
                                double visc, double *d_flux, double *d_b, double *errmax,
                                double *cfl_max, double *pe_max);

/* The face loop of Compute_Pressure_Fractional
 * (Sources/Process/Compute_Pressure_Fractional.f90:47-131) on the mesh of a
 * tfs_gradient object, every material a fluid and no BUFFER boundary
 * (TFS_ERR_INVALID otherwise): per face the density f den(c1) + (1-f)
 * den(c2), the face velocity rebuilt from the cells' velocities and pressure
 * gradients, flux = den (u_f . S) and a12 = den |S|^2 (f/sav(c1) +
 * (1-f)/sav(c2)) (0 on boundary faces, where the boundary cell's velocity
 * gives the flux); per cell b = inflow - outflow in face order.  Bitwise the
 * reference's.  The matrix then comes from tfs_assemble_device with a12 (the
 * boundary faces add 0).  Per face: f, flux, a12, and s in three blocks (sx,
 * sy, sz); per cell: vol, sav, px, py, pz, b; on cells then boundary cells:
 * u, v, w, p, den. */
int tfs_pressure_fractional_device(tfs_gradient *S, const double *d_f, const double *d_s,
                                   const double *d_vol, const double *d_sav, const double *d_u,
                                   const double *d_v, const double *d_w, const double *d_px,
                                   const double *d_py, const double *d_pz, const double *d_p,
                                   const double *d_den, double *d_flux, double *d_a12,
                                   double *d_b);

/* The face loop of Compute_Pressure_Simple
 * (Sources/Process/Compute_Pressure_Simple.f90:45-180) on the mesh of a
 * tfs_gradient object, every material a fluid and no BUFFER boundary
 * (TFS_ERR_INVALID otherwise): per interior face the density f den(c1) +
 * (1-f) den(c2), SMDPN = |S|^2 / (S . d) (d: the object's face vectors), the
 * interpolated velocity and pressure gradient, flux = DENs (u_f . S) +
 * A12 (p1 - p2) + A12 (grad p_f . d), A12 = 0.5 DENs SMDPN (vol1/sav1 +
 * vol2/sav2); per boundary face by the kind of its boundary cell (TFS_BC_*):
 * inflow / outflow / convective take the boundary cell's velocity, pressure
 * the inside cell's, walls and symmetry planes have no flux; outflow,
 * convective and pressure faces add A12 = DENs SMDPN vol1/sav1 to their
 * diagonal (bnd_on = 1, else 0; a12 = 0 where a face adds nothing); b per
 * cell = inflow - outflow in face order.  Bitwise the reference's.  The
 * matrix then comes from tfs_assemble_device(a12, NULL, bnd_on, NULL, 1, NULL)
 * (the reference sets the off-diagonals to -A12, which 0 - A12 equals for
 * every A12 != 0).  Device arrays: f, s (sx, sy, sz blocks), flux, a12,
 * bnd_on per face; vol, sav, px, py, pz, b per cell; u, v, w, p, den per cell
 * then boundary cell. */
int tfs_pressure_simple_device(tfs_gradient *S, const double *d_f, const double *d_s,
                               const double *d_vol, const double *d_sav, const double *d_u,
                               const double *d_v, const double *d_w, const double *d_px,
                               const double *d_py, const double *d_pz, const double *d_p,
                               const double *d_den, double *d_flux, double *d_a12,
                               unsigned char *d_bnd_on, double *d_b);

/* ---- serving: continuous batching of solves on one GPU ------------------------
 * Not in the reference.  A case (one system) is registered with its structure
 * and values; its GPU copy is made when its first solve is admitted and stays
 * resident across visits until a solve is marked as the case's last.
 * Requests queue in arrival order; `workers` host threads each take the first
 * request whose case is idle and resident or fits in the HBM budget, and solve
 * it on the case's own stream, so solves of different cases overlap on the
 * device.  A result does not depend on what else runs (per-case state,
 * fixed-order reductions).  Calls on one server may come from any thread;
 * times are seconds since the server was created. */
typedef struct tfs_server tfs_server;
int tfs_server_create(long long hbm_budget_bytes, int workers, tfs_server **out);
/* Waits for all queued requests, then frees every resident case. */
int tfs_server_destroy(tfs_server *S);
/* Copies the structure and values; TFS_ERR_NOMEM if it could never fit.
 * When a case does not fit, idle resident cases with no queued request are
 * evicted, least recently used first (their next visit uploads them again). */
int tfs_server_add_case(tfs_server *S, int n, long long nnz, const int *row, const int *col,
                        const int *dia, const double *val, int base, int *case_id);
/* Queues one solve (tfs_solve semantics, x0 = 0, b copied); release != 0
 * frees the case's GPU copy after this solve. */
int tfs_server_submit(tfs_server *S, int case_id, int method, int prec, const double *b,
                      double tol, int cap, int release, long long *req);
/* Blocks until `req` is done; x (n doubles) may be NULL.  Each request can be
 * waited for once. */
int tfs_server_wait(tfs_server *S, long long req, double *x, int *niter, double *ini_res,
                    double *fin_res, int *status, double *t_submit, double *t_start,
                    double *t_end);
/* Withdraws a request that has not started (its wait then reports
 * TFS_CANCELLED); TFS_ERR_INVALID once it runs or is done. */
int tfs_server_cancel(tfs_server *S, long long req);
int tfs_server_stats(tfs_server *S, long long *hbm_in_use, long long *hbm_peak,
                     int *resident_cases, long long *completed);

/* ---- building blocks on device vectors ---------------------------------------- */
/* y := A x  (row sums in the reference order, Cg.f90:128-134). */
int tfs_spmv_device(tfs_matrix *A, const double *d_x, double *d_y);
/* r := r - A x  (Residual.f90:22-27, term by term in row order). */
int tfs_residual_device(tfs_matrix *A, const double *d_x, double *d_r);
/* The same without allocating device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its two
 * work vectors live in a library scratch array of 64 MiB held on each GPU the
 * library runs on: matrices of up to 4,194,304 rows (more: TFS_ERR_INVALID).
 * The matrix must exist (its device layouts are allocated at creation); calls
 * that overlap on one GPU share the scratch array and race. */
int tfs_residual_device_noalloc(tfs_matrix *A, const double *d_x, double *d_r);
/* Prec_Form: builds (or reuses) the factor for prec. */
int tfs_prec_form(tfs_matrix *A, int prec);
/* Prec_Solve: x := M^-1 b with the current factor (tfs_prec_form first). */
int tfs_prec_solve_device(tfs_matrix *A, int prec, const double *d_b, double *d_x);
/* Copies the n used diagonal entries of the preconditioner D to d_out. */
int tfs_prec_diag_device(tfs_matrix *A, double *d_out);
/* The same for the FP32 factor of the refinement solver (built if needed),
 * widened to FP64. */
int tfs_prec_diag32_device(tfs_matrix *A, int prec, double *d_out);

/* ---- measurement ----------------------------------------------------------------
 * Runs `reps` back-to-back SpMV launches y := A x on the library's stream and
 * returns the mean launch duration in ms from HIP events on that stream. */
int tfs_spmv_time(tfs_matrix *A, const double *d_x, double *d_y, int reps, double *ms);
/* The same with the 256 MB memory-side cache flushed before every launch
 * (640 MiB copied), each launch timed alone: the HBM-bound figure. */
int tfs_spmv_time_cold(tfs_matrix *A, const double *d_x, double *d_y, int reps, double *ms);
/* Average time of one FP64 incomplete-Cholesky sweep (forward and backward
 * alternating, `reps` pairs back to back, factor formed first). */
int tfs_sweep_time(tfs_matrix *A, int reps, double *ms);

/* Kernel classes for event timing during real solves. */
enum {
  TFS_KC_SPMV = 0,    /* k_spmv in every mode                                  */
  TFS_KC_SWEEP = 1,   /* forward / backward preconditioner sweeps (one bracket
                         per sweep, also where a sweep is a launch per level) */
  TFS_KC_FACTOR = 2,  /* Prec_Form level kernels                               */
  TFS_KC_VECTOR = 3,  /* reserved: vector update / reduction kernels           */
  TFS_KC_SPMV32 = 4   /* k_ell_diff, the refinement's FP32 operator            */
};
/* From now on every launch of `kclass` on the library stream is bracketed by
 * a HIP event pair (up to 4096 pairs between stream synchronisations). */
int tfs_timing_start(int kclass);
/* Stops recording; total bracketed milliseconds and number of launches. */
int tfs_timing_stop(double *ms, long long *launches);

#ifdef __cplusplus
}
#endif

#endif
