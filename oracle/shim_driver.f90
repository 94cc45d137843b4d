!==============================================================================!
!   shim_driver.f90 -- TEST DRIVER for the Fortran drop-in module            !
!   (t-flows_amd/fortran/solvers_mod.f90).                                    !
!                                                                              !
!   Built by `make -C oracle shim` against the REFERENCE's own Grid_Mod and   !
!   Matrix_Mod objects (compiled from its sources by the `ref` target), so    !
!   the caller side is exactly T-FlowS's: the matrix comes out of the         !
!   reference Matrix_Mod_Topology and is passed as its Matrix_Type to        !
!   Cg / Bicg / Cgs of the drop-in, which solves on the GPU.                  !
!                                                                              !
!   Usage: shim_driver <faces.bin> <case.bin> <out.bin>                       !
!     faces.bin: n_cells, n_faces (int32), faces_c(2, n_faces) (int32)       !
!     case.bin : per face a12, a21 (float64), per cell extra diagonal and b   !
!     out.bin  : for method = Cg, Bicg, Cgs and prec = 0..2, then two more    !
!                Cg prec 2 solves -- values assembled on the device through   !
!                Solvers_Assemble_On_Device, and a device copy of A % val     !
!                handed over through Solvers_Values_From_Device:              !
!                niter (int32), ini_res, fin_res, x(1:n) (float64)            !
!==============================================================================!
  program Shim_Driver
  use iso_c_binding
  use Grid_Mod
  use Matrix_Mod
  use Solvers_Mod
  implicit none
  interface  ! the HIP runtime, for the device copies a caller would hold
    integer(c_int) function hipMalloc(p, bytes) bind(C, name='hipMalloc')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr) :: p
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name='hipMemcpy')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
    end function
    integer(c_int) function hipFree(p) bind(C, name='hipFree')
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
  end interface
  type(Grid_Type), target :: grid
  type(Matrix_Type), target :: A
  character(len=512)     :: f_faces, f_case, f_out
  integer(c_int)         :: nc, nf
  integer(c_int), allocatable :: fc(:,:)
  real(c_double), allocatable, target :: a12(:), a21(:), dex(:), b(:)
  real, allocatable      :: x(:), r(:)
  integer :: s, c1, c2, method, prec, niter, c, k, nnz
  real    :: tol, ini_res, fin_res, bnorm
  type(c_ptr) :: d_a12, d_a21, d_dex, d_val

  call get_command_argument(1, f_faces)
  call get_command_argument(2, f_case)
  call get_command_argument(3, f_out)

  open(11, file=trim(f_faces), access='stream', form='unformatted', status='old')
  read(11) nc, nf
  allocate(fc(2, nf))
  read(11) fc
  close(11)
  allocate(a12(nf), a21(nf), dex(nc), b(nc))
  open(12, file=trim(f_case), access='stream', form='unformatted', status='old')
  read(12) a12, a21, dex, b
  close(12)

  grid % n_cells     = nc
  grid % n_bnd_cells = max(0, -minval(fc(2,:)))
  grid % n_faces     = nf
  allocate(grid % faces_c(2, nf))
  grid % faces_c = fc
  call Matrix_Mod_Allocate(grid, A)
  call Matrix_Mod_Topology(grid, A)

  ! assembly through pos / dia, the reference's scatter pattern
  A % val = 0.0
  do s = 1, nf
    c1 = fc(1, s)
    c2 = fc(2, s)
    if(c2 > 0) then
      A % val(A % pos(1,s)) = A % val(A % pos(1,s)) - a12(s)
      A % val(A % dia(c1))  = A % val(A % dia(c1))  + a12(s)
      A % val(A % pos(2,s)) = A % val(A % pos(2,s)) - a21(s)
      A % val(A % dia(c2))  = A % val(A % dia(c2))  + a21(s)
    else
      A % val(A % dia(c1))  = A % val(A % dia(c1))  + a12(s)
    end if
  end do
  do c = 1, nc
    A % val(A % dia(c)) = A % val(A % dia(c)) + dex(c)
  end do

  bnorm = sqrt(sum(b * b))
  tol = 1.0e-10 * bnorm
  allocate(x(-grid % n_bnd_cells : nc), r(nc))
  open(13, file=trim(f_out), access='stream', form='unformatted', status='replace')
  do method = 0, 2
    do prec = 0, 2
      x = 0.0
      r = b
      niter = 3000
      ini_res = 0.0
      fin_res = 0.0
      if(method == 0) call Cg  (A, x, r, prec, niter, tol, ini_res, fin_res)
      if(method == 1) call Bicg(A, x, r, prec, niter, tol, ini_res, fin_res)
      if(method == 2) call Cgs (A, x, r, prec, niter, tol, ini_res, fin_res)
      write(*,'(a,i2,a,i2,a,i6,a,es12.4)') ' method', method, ' prec', prec,  &
            ' niter', niter, ' fin_res', fin_res
      write(13) int(niter, c_int), real(ini_res, c_double), real(fin_res, c_double),  &
                real(x(1:nc), c_double)
    end do
  end do

  ! the same system with its values formed on the device, then handed over
  ! from a device copy: no host upload of A % val
  nnz = A % row(nc+1) - 1
  if(hipMalloc(d_a12, int(8*nf, c_size_t)) /= 0 .or.  &
     hipMalloc(d_a21, int(8*nf, c_size_t)) /= 0 .or.  &
     hipMalloc(d_dex, int(8*nc, c_size_t)) /= 0 .or.  &
     hipMalloc(d_val, int(8*nnz, c_size_t)) /= 0) stop 'hipMalloc failed'
  if(hipMemcpy(d_a12, c_loc(a12), int(8*nf, c_size_t), 1_c_int) /= 0 .or.  &
     hipMemcpy(d_a21, c_loc(a21), int(8*nf, c_size_t), 1_c_int) /= 0 .or.  &
     hipMemcpy(d_dex, c_loc(dex), int(8*nc, c_size_t), 1_c_int) /= 0 .or.  &
     hipMemcpy(d_val, c_loc(A % val(1)), int(8*nnz, c_size_t), 1_c_int) /= 0)  &
    stop 'hipMemcpy failed'
  do k = 1, 2
    if(k == 1) call Solvers_Assemble_On_Device(A, nf, fc, d_a12, d_a21, c_null_ptr, d_dex,  &
                                               1.0, c_null_ptr)
    if(k == 2) call Solvers_Values_From_Device(A, d_val)
    x = 0.0
    r = b
    niter = 3000
    call Cg(A, x, r, 2, niter, tol, ini_res, fin_res)
    write(*,'(a,i2,a,i6,a,es12.4)') ' device values', k, ' niter', niter, ' fin_res', fin_res
    write(13) int(niter, c_int), real(ini_res, c_double), real(fin_res, c_double),  &
              real(x(1:nc), c_double)
  end do
  if(hipFree(d_a12) + hipFree(d_a21) + hipFree(d_dex) + hipFree(d_val) /= 0) stop 'hipFree'
  close(13)
  end program
