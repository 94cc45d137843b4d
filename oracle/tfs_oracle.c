/*
 * tfs_oracle.c -- CPU restatement of the T-FlowS linear-solver path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The GPU library never links or calls
 * this file; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it as the checker.
 *
 * Every function follows one reference routine loop by loop and cites it.
 * Arrays are 0-based; the reference is 1-based Fortran.  Build with
 * -ffp-contract=off so that a*b+c is never fused, as in the reference FAST
 * build on x86-64 (no FMA in the baseline ISA).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   or_spmv / or_residual : dense brute force on tiny matrices, the 4x4 CRS
 *                           example of Matrix_Mod.f90:15-34, Laplacian times a
 *                           linear field, bitwise against the compiled reference.
 *   or_prec_form          : closed form d_i = i/(i+1) on tridiag(-1,2,-1);
 *                           diag(M) == diag(A) for symmetric A; bitwise against
 *                           the compiled reference (through Cg's iterates).
 *   or_prec_solve         : M z == r with M = (Dt+L) Dt^-1 (Dt+U) formed densely.
 *   or_cg                 : textbook PCG (scipy cg with the same M) on small SPD
 *                           systems; exact termination in <= n steps; bitwise
 *                           against the compiled reference.
 *   or_bicg               : bitwise identical to or_cg (the shadow sequence of the
 *                           reference equals the primary one); bitwise against the
 *                           compiled reference.
 *   or_cgs                : dense solve on tiny nonsymmetric systems; bitwise
 *                           against the compiled reference.
 *   or_topology           : bitwise against the reference Matrix_Mod_Topology.
 *   or_assemble           : bitwise against the reference Pressure_Matrix_Fractional;
 *                           exact (integer-valued) coefficients against a dense
 *                           sum of per-face 2x2 element matrices.
 *   or_pressure_fractional: fluxes and matrix bitwise against the reference's
 *                           Compute_Pressure_Fractional, its right-hand side
 *                           through the reference's pressure-correction solution.
 *   or_correct_velocity   : bitwise against the reference's Correct_Velocity
 *                           (velocities, fluxes, mass errors, errmax) on real
 *                           meshes, CFL / Peclet to the 4 digits it reports;
 *                           a divergence-free flux field has zero mass error.
 *   or_gradient_matrix / or_graphi / or_gradp : bitwise against the reference's
 *                           Compute_Gradient_Matrix / GraPhi / GradP on real
 *                           meshes; G times the normal matrix is the identity;
 *                           the gradient of a linear field is exact away from
 *                           periodic faces; symmetric boundaries drop their faces.
 */
#include "tfs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Sources/Shared/allp_mod.f90:40 (real promoted to FP64 by -fdefault-real-8). */
#define OR_TINY 1.e-30

/* Sources/Process/Solvers_Mod/Prec_Form.f90:20-48 */
void or_prec_form(const or_matrix *A, int prec, double *d) {
  const int n = A->n;
  if (prec == 1) { /* Prec_Form.f90:24-27 */
    for (int i = 0; i < n; ++i) d[i] = A->val[A->dia[i]];
  } else if (prec == 2) { /* Prec_Form.f90:32-40 */
    for (int i = 0; i < n; ++i) {
      double sum1 = A->val[A->dia[i]];
      for (int j = A->row[i]; j < A->dia[i]; ++j) {
        const int k = A->col[j];
        sum1 = sum1 - d[k] * A->val[j] * A->val[j];
      }
      d[i] = 1.0 / sum1;
    }
  } else { /* Prec_Form.f90:45-48 */
    for (int i = 0; i < n; ++i) d[i] = 1.0;
  }
}

/* Sources/Process/Solvers_Mod/Prec_Solve.f90:31-75 */
void or_prec_solve(const or_matrix *A, const double *d, int prec, double *x,
                   const double *b) {
  const int n = A->n;
  if (prec == 1) { /* Prec_Solve.f90:35-38 */
    for (int i = 0; i < n; ++i) x[i] = b[i] / d[i];
  } else if (prec == 2) {
    /* forward substitution, Prec_Solve.f90:46-53 */
    for (int i = 0; i < n; ++i) {
      double sum1 = b[i];
      for (int j = A->row[i]; j < A->dia[i]; ++j)
        sum1 = sum1 - A->val[j] * x[A->col[j]];
      x[i] = sum1 * d[i];
    }
    /* scale pass, Prec_Solve.f90:55-57 */
    for (int i = 0; i < n; ++i) x[i] = x[i] / (d[i] + OR_TINY);
    /* backward substitution, Prec_Solve.f90:60-67 */
    for (int i = n - 1; i >= 0; --i) {
      double sum1 = x[i];
      for (int j = A->dia[i] + 1; j < A->row[i + 1]; ++j)
        sum1 = sum1 - A->val[j] * x[A->col[j]];
      x[i] = sum1 * d[i];
    }
  } else { /* Prec_Solve.f90:72-75 */
    for (int i = 0; i < n; ++i) x[i] = b[i];
  }
}

/* Sources/Process/Solvers_Mod/Residual.f90:22-27 (buffer loop empty, n_proc=0) */
void or_residual(const or_matrix *A, const double *x, double *r) {
  for (int i = 0; i < A->n; ++i)
    for (int j = A->row[i]; j < A->row[i + 1]; ++j)
      r[i] = r[i] - A->val[j] * x[A->col[j]];
}

/* Sources/Process/Solvers_Mod/Cg.f90:128-134 */
void or_spmv(const or_matrix *A, const double *x, double *y) {
  for (int i = 0; i < A->n; ++i) {
    double s = 0.0;
    for (int j = A->row[i]; j < A->row[i + 1]; ++j) s = s + A->val[j] * x[A->col[j]];
    y[i] = s;
  }
}

static double or_sumsq(int n, const double *v) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + v[i] * v[i];
  return s;
}

static double or_dot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + a[i] * b[i];
  return s;
}

/* Sources/Process/Solvers_Mod/Cg.f90:36-181 */
int or_cg(const or_matrix *A, double *x, double *r1, int prec, int *niter,
          double tol, double *ini_res, double *fin_res) {
  const int n = A->n;
  const int cap = *niter;
  double *d = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *p1 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *q1 = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double rho = 0.0, rhoold = 0.0, alfa, beta, error;
  int iter;

  or_prec_form(A, prec, d);                                  /* Cg.f90:42 */
  const double bnrm2 = sqrt(or_sumsq(n, r1));                /* Cg.f90:48-53 */
  if (bnrm2 < tol) { iter = 0; error = bnrm2; goto done; }   /* Cg.f90:55-58 */

  or_residual(A, x, r1);                                     /* Cg.f90:63 */
  memcpy(p1, r1, sizeof(double) * (size_t)n);                /* Cg.f90:68-70 */
  error = sqrt(or_sumsq(n, r1));                             /* Cg.f90:75-80 */
  *ini_res = error;                                          /* Cg.f90:85 */
  if (error < tol) { iter = 0; goto done; }                  /* Cg.f90:87-90 */

  for (iter = 1; iter <= cap; ++iter) {                      /* Cg.f90:97 */
    or_prec_solve(A, d, prec, q1, r1);                       /* Cg.f90:103 */
    rho = or_dot(n, r1, q1);                                 /* Cg.f90:108-112 */
    if (iter == 1) {
      memcpy(p1, q1, sizeof(double) * (size_t)n);
    } else {
      beta = rho / rhoold;                                   /* Cg.f90:119-122 */
      for (int i = 0; i < n; ++i) p1[i] = q1[i] + beta * p1[i];
    }
    or_spmv(A, p1, q1);                                      /* Cg.f90:128-134 */
    alfa = or_dot(n, p1, q1);                                /* Cg.f90:148-153 */
    alfa = rho / alfa;
    for (int i = 0; i < n; ++i) {                            /* Cg.f90:159-162 */
      x[i] = x[i] + alfa * p1[i];
      r1[i] = r1[i] - alfa * q1[i];
    }
    error = sqrt(or_sumsq(n, r1));                           /* Cg.f90:167-172 */
    if (error < tol) goto done;                              /* Cg.f90:174 */
    rhoold = rho;                                            /* Cg.f90:176 */
  }
  /* Fortran DO leaves iter == cap+1 on normal completion (Cg.f90:178,181). */
done:
  *fin_res = error;
  *niter = iter;
  free(d);
  free(p1);
  free(q1);
  return 0;
}

/* Sources/Process/Solvers_Mod/Bicg.f90:39-194 */
int or_bicg(const or_matrix *A, double *x, double *r1, int prec, int *niter,
            double tol, double *ini_res, double *fin_res) {
  const int n = A->n;
  const int cap = *niter;
  const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
  double *d = (double *)malloc(nb);
  double *p1 = (double *)malloc(nb), *p2 = (double *)malloc(nb);
  double *q1 = (double *)malloc(nb), *q2 = (double *)malloc(nb);
  double *r2 = (double *)malloc(nb);
  double rho = 0.0, rhoold = 0.0, alfa, beta, error;
  int iter;

  or_prec_form(A, prec, d);                                  /* Bicg.f90:45 */
  const double bnrm2 = sqrt(or_sumsq(n, r1));                /* Bicg.f90:51-56 */
  if (bnrm2 < tol) { iter = 0; error = bnrm2; goto done; }   /* Bicg.f90:58-61 */

  or_residual(A, x, r1);                                     /* Bicg.f90:66 */
  error = sqrt(or_sumsq(n, r1));                             /* Bicg.f90:71-76 */
  *ini_res = error;                                          /* Bicg.f90:81 */
  if (error < tol) { iter = 0; goto done; }                  /* Bicg.f90:83-86 */
  memcpy(r2, r1, nb);                                        /* Bicg.f90:91-93 */

  for (iter = 1; iter <= cap; ++iter) {                      /* Bicg.f90:100 */
    or_prec_solve(A, d, prec, q1, r1);                       /* Bicg.f90:107 */
    or_prec_solve(A, d, prec, q2, r2);                       /* Bicg.f90:108 */
    rho = or_dot(n, q1, r2);                                 /* Bicg.f90:113-117 */
    if (iter == 1) {                                         /* Bicg.f90:119-130 */
      memcpy(p1, q1, nb);
      memcpy(p2, q2, nb);
    } else {
      beta = rho / rhoold;
      for (int i = 0; i < n; ++i) {
        p1[i] = q1[i] + beta * p1[i];
        p2[i] = q2[i] + beta * p2[i];
      }
    }
    for (int i = 0; i < n; ++i) {                            /* Bicg.f90:136-144 */
      double s1 = 0.0, s2 = 0.0;
      for (int j = A->row[i]; j < A->row[i + 1]; ++j) {
        s1 = s1 + A->val[j] * p1[A->col[j]];
        s2 = s2 + A->val[j] * p2[A->col[j]];
      }
      q1[i] = s1;
      q2[i] = s2;
    }
    alfa = or_dot(n, p2, q1);                                /* Bicg.f90:160-165 */
    alfa = rho / alfa;
    for (int i = 0; i < n; ++i) {                            /* Bicg.f90:171-175 */
      x[i] = x[i] + alfa * p1[i];
      r1[i] = r1[i] - alfa * q1[i];
      r2[i] = r2[i] - alfa * q2[i];
    }
    error = sqrt(or_sumsq(n, r1));                           /* Bicg.f90:180-185 */
    if (error < tol) goto done;                              /* Bicg.f90:187 */
    rhoold = rho;                                            /* Bicg.f90:189 */
  }
done:
  *fin_res = error;
  *niter = iter;
  free(d);
  free(p1); free(p2); free(q1); free(q2); free(r2);
  return 0;
}

/* Sources/Process/Solvers_Mod/Cgs.f90:43-230 */
int or_cgs(const or_matrix *A, double *x, double *r1, int prec, int *niter,
           double tol, double *ini_res, double *fin_res) {
  const int n = A->n;
  const int cap = *niter;
  const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
  double *d = (double *)malloc(nb);
  double *p1 = (double *)malloc(nb), *p2 = (double *)malloc(nb);
  double *q1 = (double *)malloc(nb), *q2 = (double *)malloc(nb);
  double *r2 = (double *)malloc(nb), *u1 = (double *)malloc(nb);
  double *u2 = (double *)malloc(nb), *v2 = (double *)malloc(nb);
  double *u1q1 = (double *)malloc(nb);
  double rho = 0.0, rhoold = 0.0, alfa, beta, error;
  int iter;

  or_prec_form(A, prec, d);                                  /* Cgs.f90:49 */
  const double bnrm2 = sqrt(or_sumsq(n, r1));                /* Cgs.f90:55-60 */
  if (bnrm2 < tol) { iter = 0; error = bnrm2; goto done; }   /* Cgs.f90:62-65 */

  or_residual(A, x, r1);                                     /* Cgs.f90:70 */
  memcpy(r2, r1, nb);                                        /* Cgs.f90:75-77 */
  error = sqrt(or_sumsq(n, r1));                             /* Cgs.f90:82-87 */
  *ini_res = error;                                          /* Cgs.f90:92 */
  if (error < tol) { iter = 0; goto done; }                  /* Cgs.f90:94-97 */

  for (iter = 1; iter <= cap; ++iter) {                      /* Cgs.f90:104 */
    rho = or_dot(n, r1, r2);                                 /* Cgs.f90:109-113 */
    if (iter == 1) {                                         /* Cgs.f90:115-126 */
      for (int i = 0; i < n; ++i) {
        u1[i] = r1[i];
        u2[i] = u1[i];
      }
    } else {
      beta = rho / rhoold;
      for (int i = 0; i < n; ++i) {
        u1[i] = r1[i] + beta * q1[i];
        u2[i] = u1[i] + beta * (q1[i] + beta * u2[i]);
      }
    }
    or_prec_solve(A, d, prec, p2, u2);                       /* Cgs.f90:131 */
    /* v2 = A p2, Cgs.f90:136-143.  The `alfa` accumulated inside that loop is
     * overwritten at Cgs.f90:157 before any use, so it is not restated. */
    or_spmv(A, p2, v2);
    alfa = or_dot(n, r2, v2);                                /* Cgs.f90:157-162 */
    alfa = rho / alfa;
    for (int i = 0; i < n; ++i) q1[i] = u1[i] - alfa * v2[i]; /* Cgs.f90:167-169 */
    for (int i = 0; i < n; ++i) u1q1[i] = u1[i] + q1[i];      /* Cgs.f90:174-176 */
    or_prec_solve(A, d, prec, p1, u1q1);                     /* Cgs.f90:177 */
    for (int i = 0; i < n; ++i) x[i] = x[i] + alfa * p1[i];  /* Cgs.f90:182-184 */
    or_spmv(A, p1, q2);                                      /* Cgs.f90:189-195 */
    for (int i = 0; i < n; ++i) r1[i] = r1[i] - alfa * q2[i]; /* Cgs.f90:209-211 */
    error = sqrt(or_sumsq(n, r1));                           /* Cgs.f90:216-221 */
    if (error < tol) goto done;                              /* Cgs.f90:223 */
    rhoold = rho;                                            /* Cgs.f90:225 */
  }
done:
  *fin_res = error;
  *niter = iter;
  free(d);
  free(p1); free(p2); free(q1); free(q2); free(r2);
  free(u1); free(u2); free(v2); free(u1q1);
  return 0;
}

/* Sources/Process/Matrix_Mod/Topology.f90:33-47: nnz = sum of stencil widths. */
int or_topology_nnz(int n_cells, int n_faces, const int *faces_c) {
  long n = n_cells;
  for (int s = 0; s < n_faces; ++s)
    if (faces_c[2 * s + 1] > 0) n += 2;
  return (int)n;
}

static int or_cmp_int(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* Sources/Process/Matrix_Mod/Topology.f90:26-121 (1-based outputs). */
void or_topology(int n_cells, int n_faces, const int *faces_c, int *row,
                 int *col, int *dia, int *mir, int *pos) {
  int *stencw = (int *)malloc(sizeof(int) * (size_t)(n_cells + 1));
  const int nnz = or_topology_nnz(n_cells, n_faces, faces_c);
  for (int c = 1; c <= n_cells; ++c) stencw[c] = 1;
  for (int s = 0; s < n_faces; ++s) {                        /* :34-41 */
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    if (c2 > 0) { stencw[c1]++; stencw[c2]++; }
  }
  for (int k = 0; k < nnz; ++k) { col[k] = 0; mir[k] = 0; }
  for (int s = 0; s < 2 * n_faces; ++s) pos[s] = 0;
  row[0] = 1;                                                /* :54-59 */
  for (int c = 1; c <= n_cells; ++c) {
    row[c] = row[c - 1] + stencw[c];
    col[row[c - 1] - 1] = c;
    stencw[c] = 1;
  }
  for (int s = 0; s < n_faces; ++s) {                        /* :62-71 */
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    if (c2 > 0) {
      col[row[c1 - 1] + stencw[c1] - 1] = c2;
      col[row[c2 - 1] + stencw[c2] - 1] = c1;
      stencw[c1]++;
      stencw[c2]++;
    }
  }
  for (int c = 1; c <= n_cells; ++c) {                       /* :75-82 */
    qsort(&col[row[c - 1] - 1], (size_t)stencw[c], sizeof(int), or_cmp_int);
    for (int p = row[c - 1]; p <= row[c] - 1; ++p)
      if (col[p - 1] == c) dia[c - 1] = p;
  }
  for (int c1 = 1; c1 <= n_cells; ++c1) {                    /* :85-103 */
    for (int p1 = row[c1 - 1]; p1 <= row[c1] - 1; ++p1) {
      const int c2 = col[p1 - 1];
      for (int p2 = row[c2 - 1]; p2 <= row[c2] - 1; ++p2) {
        if (col[p2 - 1] == c1) {
          mir[p1 - 1] = p2;
          mir[p2 - 1] = p1;
          break;
        }
      }
    }
  }
  for (int s = 0; s < n_faces; ++s) {                        /* :106-121 */
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    if (c2 > 0) {
      for (int p = row[c1 - 1]; p <= row[c1] - 1; ++p)
        if (col[p - 1] == c2) pos[2 * s] = p;
      for (int p = row[c2 - 1]; p <= row[c2] - 1; ++p)
        if (col[p - 1] == c1) pos[2 * s + 1] = p;
    }
  }
  free(stencw);
}

/* Pressure_Matrix_Fractional.f90:19-21 (val := 0), :26-48 (face loop; the
 * BUFFER branch is the boundary add here), NewUVW.f90:372-384 (the same loop
 * with a12 != a21 and wall/inflow boundaries adding to the diagonal),
 * :500-506 (inertia cell term), :586-590 (sav, then division by URF). */
void or_assemble(int n_cells, int n_faces, const int *faces_c, const int *pos, const int *dia,
                 int nnz, const double *a12, const double *a21, const unsigned char *bnd_on,
                 const double *cell, double urf, double *val, double *sav) {
  for (int k = 0; k < nnz; ++k) val[k] = 0.0;
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    const double A12 = a12[s], A21 = a21 ? a21[s] : a12[s];
    if (c2 > 0) {
      val[pos[2 * s] - 1] = val[pos[2 * s] - 1] - A12;
      val[dia[c1 - 1] - 1] = val[dia[c1 - 1] - 1] + A12;
      val[pos[2 * s + 1] - 1] = val[pos[2 * s + 1] - 1] - A21;
      val[dia[c2 - 1] - 1] = val[dia[c2 - 1] - 1] + A21;
    } else if (!bnd_on || bnd_on[s]) {
      val[dia[c1 - 1] - 1] = val[dia[c1 - 1] - 1] + A12;
    }
  }
  for (int c = 0; c < n_cells; ++c) {
    double *d = &val[dia[c] - 1];
    if (cell) *d = *d + cell[c];
    if (sav) sav[c] = *d;
    if (urf != 1.0) *d = *d / urf;
  }
}

/* Compute_Gradient_Matrix.f90:21-28 (G := 0), :30-70 (Boundary = .TRUE.: both
 * cells of an interior face, the inside cell of a boundary face, gain d d^T;
 * the fluid-solid branch :41-53 is not taken), :96-116 (inverse of the
 * symmetric 3x3 G by cofactors over Jac + TINY). */
void or_gradient_matrix(int n_cells, int n_faces, const int *faces_c, const double *d,
                        double *g) {
  const double *dx = d, *dy = d + n_faces, *dz = d + 2 * (size_t)n_faces;
  for (int c = 0; c < 6 * n_cells; ++c) g[c] = 0.0;
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    const double x = dx[s], y = dy[s], z = dz[s];
    double *g1 = g + 6 * (size_t)(c1 - 1);
    g1[0] = g1[0] + x * x;
    g1[1] = g1[1] + y * y;
    g1[2] = g1[2] + z * z;
    g1[3] = g1[3] + x * y;
    g1[4] = g1[4] + x * z;
    g1[5] = g1[5] + y * z;
    if (c2 > 0) {
      double *g2 = g + 6 * (size_t)(c2 - 1);
      g2[0] = g2[0] + x * x;
      g2[1] = g2[1] + y * y;
      g2[2] = g2[2] + z * z;
      g2[3] = g2[3] + x * y;
      g2[4] = g2[4] + x * z;
      g2[5] = g2[5] + y * z;
    }
  }
  for (int c = 0; c < n_cells; ++c) {
    double *G = g + 6 * (size_t)c, inv[6];
    const double jac = G[0] * G[1] * G[2] - G[0] * G[5] * G[5] - G[3] * G[3] * G[2] +
                       2.0 * G[3] * G[4] * G[5] - G[4] * G[4] * G[1];
    inv[0] = +(G[1] * G[2] - G[5] * G[5]) / (jac + OR_TINY);
    inv[1] = +(G[0] * G[2] - G[4] * G[4]) / (jac + OR_TINY);
    inv[2] = +(G[0] * G[1] - G[3] * G[3]) / (jac + OR_TINY);
    inv[3] = -(G[3] * G[2] - G[4] * G[5]) / (jac + OR_TINY);
    inv[4] = +(G[3] * G[5] - G[4] * G[1]) / (jac + OR_TINY);
    inv[5] = -(G[0] * G[5] - G[3] * G[4]) / (jac + OR_TINY);
    for (int k = 0; k < 6; ++k) G[k] = inv[k];
  }
}

/* GraPhi.f90:25-27 (phii := 0) and the face loop of component i (:29-73 for
 * x, :75-123 for y, :125-173 for z) with Boundary = .TRUE. (:62-66): both
 * cells use the same d and the same difference phi(c2) - phi(c1), zero on a
 * SYMMETRY boundary (:41-44), each weighted by its own row of G. */
void or_graphi(int n_cells, int n_faces, const int *faces_c, const double *d, const double *g,
               const double *phi, int i, const unsigned char *bc, double *phii) {
  const double *dx = d, *dy = d + n_faces, *dz = d + 2 * (size_t)n_faces;
  /* the G entries of row i: (1,4,5), (4,2,6), (5,6,3) */
  static const int row[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  const int *r = row[i - 1];
  for (int c = 0; c < n_cells; ++c) phii[c] = 0.0;
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    const size_t k1 = (size_t)(c1 - 1);
    const size_t k2 = c2 > 0 ? (size_t)(c2 - 1) : (size_t)n_cells + (size_t)(-c2 - 1);
    double dphi = phi[k2] - phi[k1];
    if (c2 < 0 && bc && bc[-c2 - 1] == OR_BC_SYMMETRY) dphi = 0.0;
    const double x = dx[s], y = dy[s], z = dz[s];
    const double *g1 = g + 6 * k1;
    phii[k1] = phii[k1] + dphi * (g1[r[0]] * x + g1[r[1]] * y + g1[r[2]] * z);
    if (c2 > 0) {
      const double *g2 = g + 6 * k2;
      phii[k2] = phii[k2] + dphi * (g2[r[0]] * x + g2[r[1]] * y + g2[r[2]] * z);
    }
  }
}

/* GradP.f90:26-44: boundary cells other than BUFFER and PRESSURE take the
 * inside value, the gradient follows; :46-64: one pass extrapolating those
 * boundary values along the gradient with factor 1.2, then the gradient
 * again.  phi's boundary part is rewritten, as in the reference. */
void or_gradp(int n_cells, int n_faces, const int *faces_c, const double *d, const double *g,
              const double *xc, int n_bnd, const unsigned char *bc, double *phi, double *px,
              double *py, double *pz) {
  const size_t na = (size_t)n_cells + (size_t)n_bnd;
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    const int kind = bc && c2 < 0 ? bc[-c2 - 1] : OR_BC_WALL;
    if (c2 < 0 && kind != OR_BC_BUFFER && kind != OR_BC_PRESSURE)
      phi[(size_t)n_cells + (size_t)(-c2 - 1)] = phi[c1 - 1];
  }
  or_graphi(n_cells, n_faces, faces_c, d, g, phi, 1, bc, px);
  or_graphi(n_cells, n_faces, faces_c, d, g, phi, 2, bc, py);
  or_graphi(n_cells, n_faces, faces_c, d, g, phi, 3, bc, pz);
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    const int kind = bc && c2 < 0 ? bc[-c2 - 1] : OR_BC_WALL;
    if (c2 < 0 && kind != OR_BC_BUFFER && kind != OR_BC_PRESSURE) {
      const size_t k1 = (size_t)(c1 - 1), k2 = (size_t)n_cells + (size_t)(-c2 - 1);
      phi[k2] = phi[k1] + 1.2 * (px[k1] * (xc[k2] - xc[k1]) + py[k1] * (xc[na + k2] - xc[na + k1]) +
                                 pz[k1] * (xc[2 * na + k2] - xc[2 * na + k1]));
    }
  }
  or_graphi(n_cells, n_faces, faces_c, d, g, phi, 1, bc, px);
  or_graphi(n_cells, n_faces, faces_c, d, g, phi, 2, bc, py);
  or_graphi(n_cells, n_faces, faces_c, d, g, phi, 3, bc, pz);
}

/* Correct_Velocity.f90:34-46 (velocities minus the pressure gradient times
 * vol / sav), :48-59 (PRESSURE boundary cells take the inside velocity),
 * :72-82 (flux correction on interior and BUFFER faces), :88-111 (mass error
 * per cell and its maximum), :117-135 (CFL and Peclet maxima over the faces
 * of the material).  Boundary cell -(b+1) sits at n_cells + b. */
double or_correct_velocity(int n_cells, int n_faces, const int *faces_c, const double *vol,
                           const double *sav, const double *px, const double *py,
                           const double *pz, double *u, double *v, double *w, const double *pp,
                           const double *coef, const double *bou, const unsigned char *bc,
                           const double *den, const unsigned char *in_m, const double *scoef,
                           const double *d, double dt, double visc, double *flux, double *b,
                           double *cfl, double *pe) {
  const double *dx = d, *dy = d + n_faces, *dz = d + 2 * (size_t)n_faces;
  for (int c = 0; c < n_cells; ++c) {
    u[c] = u[c] - px[c] * vol[c] / sav[c];
    v[c] = v[c] - py[c] * vol[c] / sav[c];
    w[c] = w[c] - pz[c] * vol[c] / sav[c];
  }
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    if (c2 < 0 && bc && bc[-c2 - 1] == OR_BC_PRESSURE) {
      const size_t k1 = (size_t)(c1 - 1), k2 = (size_t)n_cells + (size_t)(-c2 - 1);
      u[k2] = u[k1];
      v[k2] = v[k1];
      w[k2] = w[k1];
    }
  }
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    const int buffer = c2 < 0 && bc && bc[-c2 - 1] == OR_BC_BUFFER;
    if (c2 > 0 || buffer) {
      const size_t k1 = (size_t)(c1 - 1);
      const size_t k2 = c2 > 0 ? (size_t)(c2 - 1) : (size_t)n_cells + (size_t)(-c2 - 1);
      if (c2 > 0)
        flux[s] = flux[s] + (pp[k2] - pp[k1]) * coef[s];
      else
        flux[s] = flux[s] + (pp[k2] - pp[k1]) * (bou ? bou[-c2 - 1] : 0.0);
    }
  }
  for (int c = 0; c < n_cells; ++c) b[c] = 0.0;
  for (int s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    b[c1 - 1] = b[c1 - 1] - flux[s];
    if (c2 > 0) b[c2 - 1] = b[c2 - 1] + flux[s];
  }
  double errmax = 0.0;
  for (int c = 0; c < n_cells; ++c) {
    b[c] = b[c] / (vol[c] * den[c]);
    errmax = fmax(errmax, fabs(b[c]));
  }
  double cm = 0.0, pm = 0.0;
  for (int s = 0; s < n_faces; ++s) {
    const int c2 = faces_c[2 * s + 1];
    if (in_m && !in_m[s]) continue;
    const int buffer = c2 < 0 && bc && bc[-c2 - 1] == OR_BC_BUFFER;
    if (c2 > 0 || buffer) {
      const double ct =
          fabs(dt * flux[s] / (scoef[s] * (dx[s] * dx[s] + dy[s] * dy[s] + dz[s] * dz[s])));
      const double pt = fabs(flux[s] / scoef[s] / (visc + OR_TINY));
      cm = fmax(cm, ct);
      pm = fmax(pm, pt);
    }
  }
  *cfl = cm;
  *pe = pm;
  return errmax;
}

/* Compute_Pressure_Fractional.f90:47-48 (b := 0), :50-131 (the face loop; the
 * fluid-fluid density branch :55-57; interior faces :69-115 -- the face
 * velocity, flux and A12; boundary faces :117-127 -- the boundary cell's
 * velocity, flux only). */
void or_pressure_fractional(int n_cells, int n_faces, const int *faces_c, const double *f,
                            const double *s, const double *vol, const double *sav,
                            const double *u, const double *v, const double *w, const double *px,
                            const double *py, const double *pz, const double *p,
                            const double *den, double *flux, double *a12, double *b) {
  const double *sx = s, *sy = s + n_faces, *sz = s + 2 * (size_t)n_faces;
  for (int c = 0; c < n_cells; ++c) b[c] = 0.0;
  for (int k = 0; k < n_faces; ++k) {
    const int c1 = faces_c[2 * k], c2 = faces_c[2 * k + 1];
    const size_t k1 = (size_t)(c1 - 1);
    const size_t k2 = c2 > 0 ? (size_t)(c2 - 1) : (size_t)n_cells + (size_t)(-c2 - 1);
    const double fs = f[k];
    const double dens = fs * den[k1] + (1.0 - fs) * den[k2];
    if (c2 > 0) {
      double us = fs * (u[k1] + px[k1] * vol[k1] / sav[k1]) +
                  (1.0 - fs) * (u[k2] + px[k2] * vol[k2] / sav[k2]);
      double vs = fs * (v[k1] + py[k1] * vol[k1] / sav[k1]) +
                  (1.0 - fs) * (v[k2] + py[k2] * vol[k2] / sav[k2]);
      double ws = fs * (w[k1] + pz[k1] * vol[k1] / sav[k1]) +
                  (1.0 - fs) * (w[k2] + pz[k2] * vol[k2] / sav[k2]);
      us = us + (p[k1] - p[k2]) * sx[k] * (fs / sav[k1] + (1.0 - fs) / sav[k2]);
      vs = vs + (p[k1] - p[k2]) * sy[k] * (fs / sav[k1] + (1.0 - fs) / sav[k2]);
      ws = ws + (p[k1] - p[k2]) * sz[k] * (fs / sav[k1] + (1.0 - fs) / sav[k2]);
      flux[k] = dens * (us * sx[k] + vs * sy[k] + ws * sz[k]);
      double a = dens * (sx[k] * sx[k] + sy[k] * sy[k] + sz[k] * sz[k]);
      a12[k] = a * (fs / sav[k1] + (1.0 - fs) / sav[k2]);
      b[k1] = b[k1] - flux[k];
      b[k2] = b[k2] + flux[k];
    } else {
      flux[k] = dens * (u[k2] * sx[k] + v[k2] * sy[k] + w[k2] * sz[k]);
      a12[k] = 0.0;
      b[k1] = b[k1] - flux[k];
    }
  }
}

/* Sources/Process/Compute_Pressure_Simple.f90:45-180 (boundary-cell kinds as
 * in or_correct_velocity's bc; no BUFFER faces). */
void or_pressure_simple(int n_cells, int n_faces, const int *faces_c, const double *f,
                        const double *s, const double *d, const double *vol, const double *sav,
                        const double *u, const double *v, const double *w, const double *p,
                        const double *px, const double *py, const double *pz, const double *den,
                        const unsigned char *bc, const int *pos, const int *dia, int nnz,
                        double *flux, double *a12, unsigned char *bnd_on, double *val, double *b) {
  const double *sx = s, *sy = s + n_faces, *sz = s + 2 * (size_t)n_faces;
  const double *dx = d, *dy = d + n_faces, *dz = d + 2 * (size_t)n_faces;
  for (int k = 0; k < nnz; ++k) val[k] = 0.0; /* :46 */
  for (int c = 0; c < n_cells; ++c) b[c] = 0.0; /* :47 */
  for (int k = 0; k < n_faces; ++k) {
    const int c1 = faces_c[2 * k], c2 = faces_c[2 * k + 1];
    const size_t k1 = (size_t)(c1 - 1);
    const size_t k2 = c2 > 0 ? (size_t)(c2 - 1) : (size_t)n_cells + (size_t)(-c2 - 1);
    const double fs = f[k];
    const double dens = fs * den[k1] + (1.0 - fs) * den[k2]; /* :61-62 */
    a12[k] = 0.0;
    bnd_on[k] = 0;
    if (c2 > 0) { /* :82-127 */
      const double smdpn = (sx[k] * sx[k] + sy[k] * sy[k] + sz[k] * sz[k]) /
                           (sx[k] * dx[k] + sy[k] * dy[k] + sz[k] * dz[k]);
      const double us = fs * u[k1] + (1.0 - fs) * u[k2];
      const double vs = fs * v[k1] + (1.0 - fs) * v[k2];
      const double ws = fs * w[k1] + (1.0 - fs) * w[k2];
      const double A12 = 0.5 * dens * smdpn * (vol[k1] / sav[k1] + vol[k2] / sav[k2]);
      val[pos[2 * k] - 1] = -A12;
      val[pos[2 * k + 1] - 1] = -A12;
      val[dia[c1 - 1] - 1] = val[dia[c1 - 1] - 1] + A12;
      val[dia[c2 - 1] - 1] = val[dia[c2 - 1] - 1] + A12;
      const double dpxi = .5 * (px[k1] + px[k2]) * dx[k];
      const double dpyi = .5 * (py[k1] + py[k2]) * dy[k];
      const double dpzi = .5 * (pz[k1] + pz[k2]) * dz[k];
      flux[k] = dens * (us * sx[k] + vs * sy[k] + ws * sz[k]) + A12 * (p[k1] - p[k2]) +
                A12 * (dpxi + dpyi + dpzi);
      b[k1] = b[k1] - flux[k];
      b[k2] = b[k2] + flux[k];
      a12[k] = A12;
    } else { /* :130-177 */
      const int kind = bc ? bc[-c2 - 1] : 0;
      if (kind == 4 || kind == 5 || kind == 6 || kind == 2) {
        const size_t kv = kind == 2 ? k1 : k2; /* pressure: the inside velocity */
        flux[k] = dens * (u[kv] * sx[k] + v[kv] * sy[k] + w[kv] * sz[k]);
        b[k1] = b[k1] - flux[k];
        if (kind != 4) {
          const double smdpn = (sx[k] * sx[k] + sy[k] * sy[k] + sz[k] * sz[k]) /
                               (sx[k] * dx[k] + sy[k] * dy[k] + sz[k] * dz[k]);
          const double A12 = dens * smdpn * vol[k1] / sav[k1];
          val[dia[c1 - 1] - 1] = val[dia[c1 - 1] - 1] + A12;
          a12[k] = A12;
          bnd_on[k] = 1;
        }
      } else { /* walls and symmetry planes */
        flux[k] = 0.0;
      }
    }
  }
}
