!==============================================================================!
!   ref_driver.f90 -- C-callable driver around the REFERENCE solver modules.  !
!                                                                              !
!   TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together    !
!   with the reference's own Solvers_Mod / Matrix_Mod / Grid_Mod / Work_Mod   !
!   sources where they lie, into oracle/_ref/libtfs_ref.so.  Nothing of the    !
!   reference is copied into this repository.  Tests use the library to pin   !
!   the C restatement (oracle/tfs_oracle.c) bitwise, and bench.py times it as  !
!   the CPU baseline.  The product never loads it.                            !
!                                                                              !
!   It fills a Grid_Type with n_cells (and, for the topology entry, faces_c), !
!   fills the Matrix_Type components directly, sets the sequential-build      !
!   processor count (n_proc = 0, Communication/Sequential/StaPar.f90:12-13),  !
!   and calls the reference subroutines unchanged.                            !
!==============================================================================!
  module tfs_ref_driver
  use iso_c_binding
  use par_mod
  use Grid_Mod
  use Matrix_Mod
  use Work_Mod
  use Solvers_Mod
  implicit none

  type(Grid_Type),   target, save :: grid
  type(Matrix_Type),         save :: A

  contains

  !---------------------------------------------------------------------------!
  subroutine ref_reset_work(n)
  integer :: n
  if(allocated(r_cell_01)) then
    if(size(r_cell_01) == n + 1) return
    deallocate(r_cell_01, r_cell_02, r_cell_03, r_cell_04, r_cell_05,  &
               r_cell_06, r_cell_07, r_cell_08, r_cell_09)
  end if
  call Work_Mod_Allocate_Real_Cells(grid, 9)
  end subroutine

  !---------------------------------------------------------------------------!
  !   Load a CRS matrix (1-based row/col/dia as the reference stores them).   !
  !---------------------------------------------------------------------------!
  subroutine tfs_ref_set_matrix(n, nnz, row, col, dia, val)  &
             bind(C, name='tfs_ref_set_matrix')
  integer(c_int), value      :: n, nnz
  integer(c_int), intent(in) :: row(n+1), col(nnz), dia(n)
  real(c_double), intent(in) :: val(nnz)

  n_proc = 0
  this_proc = 0
  grid % n_cells     = n
  grid % n_bnd_cells = 0
  grid % n_faces     = 0

  if(allocated(A % row)) deallocate(A % row, A % col, A % dia, A % val)
  if(allocated(D % row)) deallocate(D % row, D % col, D % dia, D % val)
  allocate(A % row(n+1), A % col(nnz+1), A % dia(n), A % val(nnz+1))
  allocate(D % row(n+1), D % col(nnz+1), D % dia(n), D % val(nnz+1))
  A % pnt_grid => grid
  D % pnt_grid => grid
  A % nonzeros = nnz + 1
  D % nonzeros = nnz + 1
  A % row = row;  A % col(1:nnz) = col;  A % col(nnz+1) = 0
  A % dia = dia;  A % val(1:nnz) = val;  A % val(nnz+1) = 0.0
  D % row = row;  D % col = A % col;     D % dia = dia
  D % val = 0.0

  call ref_reset_work(n)
  end subroutine

  !---------------------------------------------------------------------------!
  !   One solve: method 0 = Cg, 1 = Bicg, 2 = Cgs.  x(1:n) in/out,            !
  !   b(1:n) in = RHS, out = final residual.  Returns wall seconds of the     !
  !   reference call (Prec_Form included, as the reference does per call).    !
  !---------------------------------------------------------------------------!
  real(c_double) function tfs_ref_solve(method, n, x, b, prec, niter, tol,  &
                                        ini_res, fin_res)                   &
                 bind(C, name='tfs_ref_solve')
  integer(c_int), value         :: method, n, prec
  real(c_double), intent(inout) :: x(n), b(n)
  integer(c_int), intent(inout) :: niter
  real(c_double), value         :: tol
  real(c_double), intent(inout) :: ini_res, fin_res
  real,    allocatable :: xf(:), bf(:)
  integer :: it
  real    :: t, ir, fr
  integer(8) :: c0, c1, rate

  allocate(xf(0:n), bf(n))
  xf(0) = 0.0
  xf(1:n) = x
  bf = b
  it = niter;  t = tol;  ir = ini_res;  fr = fin_res

  call system_clock(c0, rate)
  if(method == 0) call Cg  (A, xf, bf, prec, it, t, ir, fr)
  if(method == 1) call Bicg(A, xf, bf, prec, it, t, ir, fr)
  if(method == 2) call Cgs (A, xf, bf, prec, it, t, ir, fr)
  call system_clock(c1)

  x = xf(1:n);  b = bf
  niter = it;  ini_res = ir;  fin_res = fr
  tfs_ref_solve = real(c1 - c0, c_double) / real(rate, c_double)
  deallocate(xf, bf)
  end function

  !---------------------------------------------------------------------------!
  !   Prec_Form followed by one Prec_Solve: z = M^-1 b.  dg(1:n) returns the !
  !   used diagonal of the preconditioning matrix D.                          !
  !---------------------------------------------------------------------------!
  subroutine tfs_ref_prec(n, prec, b, z, dg) bind(C, name='tfs_ref_prec')
  integer(c_int), value       :: n, prec
  real(c_double), intent(in)  :: b(n)
  real(c_double), intent(out) :: z(n), dg(n)
  real, allocatable :: zf(:), bf(:)
  integer :: i

  allocate(zf(0:n), bf(n))
  zf = 0.0
  bf = b
  call Prec_Form(A, prec)
  call Prec_Solve(A, zf, bf, prec)
  z = zf(1:n)
  do i = 1, n
    dg(i) = D % val(D % dia(i))
  end do
  deallocate(zf, bf)
  end subroutine

  !---------------------------------------------------------------------------!
  !   r := r - A x through the reference Residual.                            !
  !---------------------------------------------------------------------------!
  subroutine tfs_ref_residual(n, x, r) bind(C, name='tfs_ref_residual')
  integer(c_int), value         :: n
  real(c_double), intent(in)    :: x(n)
  real(c_double), intent(inout) :: r(n)
  real, allocatable :: xf(:), rf(:)

  allocate(xf(0:n), rf(n))
  xf(0) = 0.0
  xf(1:n) = x
  rf = r
  call Residual(n, 0, A, xf, rf)
  r = rf
  deallocate(xf, rf)
  end subroutine

  !---------------------------------------------------------------------------!
  !   Matrix_Mod_Allocate + Matrix_Mod_Topology on a face list.               !
  !   faces_c(2, n_faces), 1-based cells, c2 <= 0 for boundary faces.         !
  !   Outputs are 1-based and sized by the caller: row(n+1), col(nnz),       !
  !   dia(n), mir(nnz), pos(2, n_faces).  Returns nnz.                        !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_topology(n_cells, n_bnd, n_faces, faces_c,  &
                                           nnz_cap, row, col, dia, mir, pos)  &
                 bind(C, name='tfs_ref_topology')
  integer(c_int), value       :: n_cells, n_bnd, n_faces, nnz_cap
  integer(c_int), intent(in)  :: faces_c(2, n_faces)
  integer(c_int), intent(out) :: row(n_cells+1), col(nnz_cap), dia(n_cells)
  integer(c_int), intent(out) :: mir(nnz_cap), pos(2, n_faces)
  type(Grid_Type), target     :: g
  type(Matrix_Type)           :: M
  integer :: nnz

  this_proc = 0
  g % n_cells     = n_cells
  g % n_bnd_cells = n_bnd
  g % n_faces     = n_faces
  allocate(g % faces_c(2, n_faces))
  g % faces_c = faces_c

  call Matrix_Mod_Allocate(g, M)
  call Matrix_Mod_Topology(g, M)

  nnz = M % nonzeros - 1
  tfs_ref_topology = nnz
  if(nnz > nnz_cap) return
  row = M % row
  col(1:nnz) = M % col(1:nnz)
  dia = M % dia
  mir(1:nnz) = M % mir(1:nnz)
  pos = M % pos
  end function

  end module
