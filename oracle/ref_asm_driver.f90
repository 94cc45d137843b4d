!==============================================================================!
!   ref_asm_driver.f90 -- C-callable driver around the REFERENCE pressure     !
!   matrix assembly (Sources/Process/Pressure_Matrix_Fractional.f90).         !
!                                                                              !
!   TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles it with the           !
!   reference's own pro_mod / all_mod / Var_Mod / Bulk_Mod / Grid_Mod /       !
!   Matrix_Mod sources where they lie, into oracle/_ref/libtfs_ref.so; tests  !
!   use it to pin the C restatement of the face-loop assembly bitwise.  The   !
!   product never loads it.                                                   !
!                                                                              !
!   It builds the global pro_mod matrix A with the reference's own            !
!   Matrix_Mod_Allocate / Matrix_Mod_Topology from a face list, sets dt and   !
!   Scoef, marks every boundary as a wall (no BUFFER faces: the sequential    !
!   build) and calls Pressure_Matrix_Fractional unchanged.                    !
!==============================================================================!
  module tfs_ref_asm_driver
  use iso_c_binding
  use all_mod
  use pro_mod
  use Grid_Mod
  use Matrix_Mod
  implicit none

  contains

  !---------------------------------------------------------------------------!
  !   faces_c(2, n_faces), 1-based cells, c2 <= 0 for boundary faces.         !
  !   val(nnz_cap) receives A % val(1:nnz); returns nnz.                      !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_pressure_matrix(n_cells, n_bnd, n_faces,    &
                                                  faces_c, dt_in, scoef_in,   &
                                                  nnz_cap, val)               &
                 bind(C, name='tfs_ref_pressure_matrix')
  integer(c_int), value       :: n_cells, n_bnd, n_faces, nnz_cap
  integer(c_int), intent(in)  :: faces_c(2, n_faces)
  real(c_double), value       :: dt_in
  real(c_double), intent(in)  :: scoef_in(n_faces)
  real(c_double), intent(out) :: val(nnz_cap)
  type(Grid_Type), target     :: g
  integer :: nnz

  g % n_cells     = n_cells
  g % n_bnd_cells = n_bnd
  g % n_faces     = n_faces
  allocate(g % faces_c(2, n_faces))
  g % faces_c = faces_c

  if(allocated(A % row)) deallocate(A % row)
  if(allocated(A % col)) deallocate(A % col)
  if(allocated(A % dia)) deallocate(A % dia)
  if(allocated(A % val)) deallocate(A % val)
  if(allocated(A % mir)) deallocate(A % mir)
  if(allocated(A % pos)) deallocate(A % pos)
  if(allocated(A % sav)) deallocate(A % sav)
  if(allocated(A % bou)) deallocate(A % bou)
  call Matrix_Mod_Allocate(g, A)
  call Matrix_Mod_Topology(g, A)

  if(allocated(Scoef))  deallocate(Scoef)
  if(allocated(TypeBC)) deallocate(TypeBC)
  allocate(Scoef(n_faces));  Scoef = scoef_in
  allocate(TypeBC(-n_bnd:n_cells));  TypeBC = WALL
  dt = dt_in

  call Pressure_Matrix_Fractional(g)

  nnz = A % nonzeros - 1
  tfs_ref_pressure_matrix = nnz
  if(nnz > nnz_cap) return
  val(1:nnz) = A % val(1:nnz)
  end function

  end module
