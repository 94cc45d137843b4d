!==============================================================================!
  module Solvers_Mod
!------------------------------------------------------------------------------!
!   Drop-in for T-FlowS's Solvers_Mod (Sources/Process/Solvers_Mod.f90:2-23): !
!   the same public names -- Cg, Bicg, Cgs with the reference dummy arguments !
!   (Sources/Process/Solvers_Mod/Cg.f90:2-4,24-29) and the module variable D -- !
!   forwarding each solve to libtfs_solver.so (MI355X, HIP) through           !
!   ISO_C_BINDING.  Compile with the reference's own flags                    !
!   (-fdefault-real-8): `real` here is the reference's FP64 real.             !
!                                                                              !
!   The structure (row/col/dia) is uploaded once per matrix shape, as it is   !
!   fixed after Matrix_Mod_Topology; the values are uploaded on every call,   !
!   since the callers reassemble A before each solve -- unless the caller     !
!   assembled them on the device (Solvers_Assemble_On_Device) or handed a     !
!   device copy (Solvers_Values_From_Device) since the last solve: then the   !
!   next solve uses those and uploads nothing.  A failed structure or value   !
!   transfer stops the run; another non-zero library status other than       !
!   non-convergence is reported on stdout; as in the reference,               !
!   non-convergence shows only as niter = cap+1.                              !
!------------------------------------------------------------------------------!
  use iso_c_binding
  use Matrix_Mod
!------------------------------------------------------------------------------!
  implicit none
!==============================================================================!

  ! Preconditioning "matrix" (D) -- kept public because Main.f90 builds its
  ! topology (Sources/Process/Main.f90:142); the factor itself lives on the GPU.
  type(Matrix_Type) :: D

  type(c_ptr), private, save :: tfs_handle = c_null_ptr
  integer,     private, save :: tfs_n = -1, tfs_nnz = -1
  type(c_ptr), private, save :: tfs_asm = c_null_ptr     ! device assembly of tfs_handle
  logical,     private, save :: tfs_values_on_device = .false.

  interface
    integer(c_int) function tfs_matrix_create(n, nnz, row, col, dia, base, h)  &
                   bind(C, name='tfs_matrix_create')
      import :: c_int, c_long_long, c_ptr
      integer(c_int),       value :: n
      integer(c_long_long), value :: nnz
      integer(c_int)              :: row(*), col(*), dia(*)
      integer(c_int),       value :: base
      type(c_ptr)                 :: h
    end function
    integer(c_int) function tfs_matrix_destroy(h) bind(C, name='tfs_matrix_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function tfs_matrix_set_values(h, val)  &
                   bind(C, name='tfs_matrix_set_values')
      import :: c_int, c_ptr, c_double
      type(c_ptr),    value :: h
      real(c_double)        :: val(*)
    end function
    integer(c_int) function tfs_matrix_set_values_device(h, d_val)  &
                   bind(C, name='tfs_matrix_set_values_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, d_val
    end function
    integer(c_int) function tfs_assembly_create(h, nf, faces_c, pos, dia, base, s)  &
                   bind(C, name='tfs_assembly_create')
      import :: c_int, c_long_long, c_ptr
      type(c_ptr),          value :: h
      integer(c_long_long), value :: nf
      integer(c_int)              :: faces_c(*), pos(*), dia(*)
      integer(c_int),       value :: base
      type(c_ptr)                 :: s
    end function
    integer(c_int) function tfs_assembly_destroy(s) bind(C, name='tfs_assembly_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function tfs_assemble_device(s, a12, a21, bnd_on, cell, urf, sav)  &
                   bind(C, name='tfs_assemble_device')
      import :: c_int, c_ptr, c_double
      type(c_ptr),    value :: s, a12, a21, bnd_on, cell
      real(c_double), value :: urf
      type(c_ptr),    value :: sav
    end function
    integer(c_int) function tfs_solve(h, method, prec, x1, r1, niter, tol,  &
                                      ini_res, fin_res) bind(C, name='tfs_solve')
      import :: c_int, c_ptr, c_double
      type(c_ptr),    value :: h
      integer(c_int), value :: method, prec
      real(c_double)        :: x1(*), r1(*)
      integer(c_int)        :: niter
      real(c_double), value :: tol
      real(c_double)        :: ini_res, fin_res
    end function
  end interface

  contains

!==============================================================================!
  subroutine Cg(A, x, r1, prec, niter, tol, ini_res, fin_res)
!------------------------------------------------------------------------------!
!   Preconditioned CG (Sources/Process/Solvers_Mod/Cg.f90).                    !
!------------------------------------------------------------------------------!
  type(Matrix_Type) :: A
  real              :: x(-A % pnt_grid % n_bnd_cells : A % pnt_grid % n_cells)
  real              :: r1(A % pnt_grid % n_cells)
  integer           :: prec, niter
  real              :: tol, ini_res, fin_res
!==============================================================================!
  call Tfs_Solve_On_Gpu(0, A, x, r1, prec, niter, tol, ini_res, fin_res)
  end subroutine

!==============================================================================!
  subroutine Bicg(A, x, r1, prec, niter, tol, ini_res, fin_res)
!------------------------------------------------------------------------------!
!   "BiCG" (Sources/Process/Solvers_Mod/Bicg.f90); its shadow sequence equals  !
!   the primary one bit for bit, so the library runs the CG recurrence once.   !
!------------------------------------------------------------------------------!
  type(Matrix_Type) :: A
  real              :: x(-A % pnt_grid % n_bnd_cells : A % pnt_grid % n_cells)
  real              :: r1(A % pnt_grid % n_cells)
  integer           :: prec, niter
  real              :: tol, ini_res, fin_res
!==============================================================================!
  call Tfs_Solve_On_Gpu(1, A, x, r1, prec, niter, tol, ini_res, fin_res)
  end subroutine

!==============================================================================!
  subroutine Cgs(A, x, r1, prec, niter, tol, ini_res, fin_res)
!------------------------------------------------------------------------------!
!   Preconditioned CGS (Sources/Process/Solvers_Mod/Cgs.f90).                  !
!------------------------------------------------------------------------------!
  type(Matrix_Type) :: A
  real              :: x(-A % pnt_grid % n_bnd_cells : A % pnt_grid % n_cells)
  real              :: r1(A % pnt_grid % n_cells)
  integer           :: prec, niter
  real              :: tol, ini_res, fin_res
!==============================================================================!
  call Tfs_Solve_On_Gpu(2, A, x, r1, prec, niter, tol, ini_res, fin_res)
  end subroutine

!==============================================================================!
  subroutine Solvers_Assemble_On_Device(A, n_faces, faces_c, a12, a21, bnd_on,  &
                                        cell, urf, sav)
!------------------------------------------------------------------------------!
!   Forms A % val on the device from per-face coefficients already in device  !
!   memory, as the face loops of Pressure_Matrix_Fractional.f90:19-48 and     !
!   NewUVW.f90:349-384,500-506,586-590 form it on the host (bitwise the same  !
!   values): the next Cg / Bicg / Cgs uses them without a host upload.        !
!   a12, a21 (or c_null_ptr: a21 = a12), bnd_on (bytes, or c_null_ptr) per    !
!   face and cell, sav (or c_null_ptr) per cell are device addresses;         !
!   faces_c is the grid's faces_c(2, n_faces).  A % val on the host is not    !
!   touched.                                                                  !
!------------------------------------------------------------------------------!
  type(Matrix_Type) :: A
  integer           :: n_faces
  integer           :: faces_c(2, n_faces)
  type(c_ptr)       :: a12, a21, bnd_on, cell, sav
  real              :: urf
!-----------------------------------[Locals]-----------------------------------!
  integer :: rc
!==============================================================================!
  call Tfs_Structure(A)
  if(.not. c_associated(tfs_asm)) then
    rc = tfs_assembly_create(tfs_handle, int(n_faces, c_long_long), faces_c,  &
                             A % pos, A % dia, 1_c_int, tfs_asm)
    if(rc /= 0) then
      write(*,*) '# Solvers_Mod: tfs_assembly_create failed, status ', rc
      stop
    end if
  end if
  rc = tfs_assemble_device(tfs_asm, a12, a21, bnd_on, cell, real(urf, c_double), sav)
  if(rc /= 0) then
    write(*,*) '# Solvers_Mod: tfs_assemble_device failed, status ', rc
    stop
  end if
  tfs_values_on_device = .true.

  end subroutine

!==============================================================================!
  subroutine Solvers_Values_From_Device(A, d_val)
!------------------------------------------------------------------------------!
!   The next solve takes A % val from device memory (d_val: nnz doubles in   !
!   A % col order, e.g. assembled there by the caller) instead of uploading  !
!   the host array.                                                           !
!------------------------------------------------------------------------------!
  type(Matrix_Type) :: A
  type(c_ptr)       :: d_val
!-----------------------------------[Locals]-----------------------------------!
  integer :: rc
!==============================================================================!
  call Tfs_Structure(A)
  rc = tfs_matrix_set_values_device(tfs_handle, d_val)
  if(rc /= 0) then
    write(*,*) '# Solvers_Mod: tfs_matrix_set_values_device failed, status ', rc
    stop
  end if
  tfs_values_on_device = .true.

  end subroutine

!==============================================================================!
  subroutine Tfs_Structure(A)
!------------------------------------------------------------------------------!
!   The device structure of A: once per shape (1-based, as                   !
!   Matrix_Mod_Topology builds it).                                           !
!------------------------------------------------------------------------------!
  type(Matrix_Type) :: A
!-----------------------------------[Locals]-----------------------------------!
  integer :: n, nnz, rc
!==============================================================================!
  n   = A % pnt_grid % n_cells
  nnz = A % row(n+1) - 1
  if(.not. c_associated(tfs_handle) .or. n /= tfs_n .or. nnz /= tfs_nnz) then
    if(c_associated(tfs_asm)) rc = tfs_assembly_destroy(tfs_asm)
    tfs_asm = c_null_ptr
    if(c_associated(tfs_handle)) rc = tfs_matrix_destroy(tfs_handle)
    rc = tfs_matrix_create(int(n, c_int), int(nnz, c_long_long),  &
                           A % row, A % col, A % dia, 1_c_int, tfs_handle)
    if(rc /= 0) then
      write(*,*) '# Solvers_Mod: tfs_matrix_create failed, status ', rc
      stop
    end if
    tfs_n = n
    tfs_nnz = nnz
    tfs_values_on_device = .false.
  end if

  end subroutine

!==============================================================================!
  subroutine Tfs_Solve_On_Gpu(method, A, x, r1, prec, niter, tol, ini_res, fin_res)
!------------------------------------------------------------------------------!
  integer           :: method
  type(Matrix_Type) :: A
  real              :: x(-A % pnt_grid % n_bnd_cells : A % pnt_grid % n_cells)
  real              :: r1(A % pnt_grid % n_cells)
  integer           :: prec, niter
  real              :: tol, ini_res, fin_res
!-----------------------------------[Locals]-----------------------------------!
  integer        :: rc
  integer(c_int) :: it
!==============================================================================!
  call Tfs_Structure(A)

  ! values: every call (A was reassembled by the caller), unless they were
  ! formed on or handed over from the device since the last solve
  if(.not. tfs_values_on_device) then
    rc = tfs_matrix_set_values(tfs_handle, A % val)
    if(rc /= 0) then
      write(*,*) '# Solvers_Mod: tfs_matrix_set_values failed, status ', rc
      stop
    end if
  end if
  tfs_values_on_device = .false.

  it = niter
  rc = tfs_solve(tfs_handle, int(method, c_int), int(prec, c_int), x(1), r1, it,  &
                 tol, ini_res, fin_res)
  niter = it
  if(rc < 0) write(*,*) '# Solvers_Mod: GPU solve status ', rc

  end subroutine

  end module
