!==============================================================================!
!   ref_mesh_driver.f90 -- C-callable driver around the REFERENCE mesh        !
!   readers and the geometry / pressure-matrix routines that use them:        !
!     Load_Cns          Sources/Process/Load_Cns.f90:25-69                    !
!     Load_Geo          Sources/Shared/Load_Geo.f90:23-50                     !
!     Compute_Geometry  Sources/Process/Compute_Geometry.f90:22-49 (Scoef)    !
!     Matrix_Mod_Topology, Pressure_Matrix_Fractional                         !
!                                                                              !
!   TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles it with the           !
!   reference's own sources where they lie, into oracle/_ref/libtfs_ref.so;   !
!   tests use it to pin the product's .cns/.geo readers, its face            !
!   coefficients and its pressure-matrix values bitwise on a real T-FlowS     !
!   mesh.  The product never loads it.                                        !
!                                                                              !
!   tfs_ref_mesh_load reads <name>.cns and <name>.geo from the working        !
!   directory, keeps copies of what was read, runs Compute_Geometry, builds   !
!   A with Matrix_Mod_Topology and assembles the fractional-step pressure     !
!   matrix with every boundary a wall (the sequential build, no BUFFER        !
!   faces).  tfs_ref_mesh_get copies one of the kept arrays out.              !
!==============================================================================!
  module tfs_ref_mesh_driver
  use iso_c_binding
  use all_mod
  use pro_mod
  use Grid_Mod
  use Matrix_Mod
  use Info_Mod
  use Parameters_Mod, only: FRACT, SIMPLE
  use Work_Mod
  use Solvers_Mod, only: D
  implicit none

  integer :: m_nc = 0, m_nb = 0, m_nf = 0, m_nnz = 0
  type(Grid_Type), allocatable, target :: mg     ! the grid last loaded
  integer, allocatable :: m_faces(:), m_row(:), m_col(:), m_dia(:)
  real,    allocatable :: m_xc(:), m_vol(:), m_s(:), m_shift(:), m_f(:), m_xf(:)
  real,    allocatable :: m_d(:), m_scoef(:), m_val(:)

  contains

  !---------------------------------------------------------------------------!
  !   sizes(4) <- n_cells, n_bnd_cells, n_faces, nnz;  returns 0 on success.  !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_mesh_load(cname, nlen, dt_in, sizes)      &
                 bind(C, name='tfs_ref_mesh_load')
  character(kind=c_char), intent(in) :: cname(*)
  integer(c_int), value              :: nlen
  real(c_double), value              :: dt_in
  integer(c_int), intent(out)        :: sizes(4)
  integer :: k, c, s, nc, nb, nf

  name = ' '
  do k = 1, nlen
    name(k:k) = cname(k)
  end do

  if(allocated(material)) deallocate(material)
  if(allocated(TypeBC))   deallocate(TypeBC)
  if(allocated(bcmark))   deallocate(bcmark)
  if(allocated(CopyC))    deallocate(CopyC)
  if(allocated(mg)) deallocate(mg)
  allocate(mg)
  call Load_Cns(mg, 0)
  nc = mg % n_cells;  nb = mg % n_bnd_cells;  nf = mg % n_faces

  if(allocated(delta))  deallocate(delta)
  if(allocated(WallDs)) deallocate(WallDs)
  allocate(delta(-nb:nc));  delta = 0.0
  allocate(WallDs(-nb:nc)); WallDs = 0.0
  call Load_Geo(mg, 0)

  ! what was read: faces, cell centres (cells, then boundary cells -1..-nb),
  ! volumes, face surfaces, the periodic shifts, weights, face centres
  if(allocated(m_faces)) deallocate(m_faces, m_xc, m_vol, m_s, m_shift, m_f, m_xf)
  allocate(m_faces(2*nf), m_xc(3*(nc+nb)), m_vol(nc), m_s(3*nf), m_shift(3*nf),  &
           m_f(nf), m_xf(3*nf))
  do s = 1, nf
    m_faces(2*s-1) = mg % faces_c(1,s)
    m_faces(2*s)   = mg % faces_c(2,s)
  end do
  do c = 1, nc
    m_xc(c)            = mg % xc(c)
    m_xc(nc+nb+c)      = mg % yc(c)
    m_xc(2*(nc+nb)+c)  = mg % zc(c)
    m_vol(c)           = mg % vol(c)
  end do
  do c = 1, nb
    m_xc(nc+c)            = mg % xc(-c)
    m_xc(nc+nb+nc+c)      = mg % yc(-c)
    m_xc(2*(nc+nb)+nc+c)  = mg % zc(-c)
  end do
  do s = 1, nf
    m_s(s)        = mg % sx(s);  m_s(nf+s)      = mg % sy(s);  m_s(2*nf+s)      = mg % sz(s)
    m_shift(s)    = mg % dx(s);  m_shift(nf+s)  = mg % dy(s);  m_shift(2*nf+s)  = mg % dz(s)
    m_f(s)        = mg % f(s)
    m_xf(s)       = mg % xf(s);  m_xf(nf+s)     = mg % yf(s);  m_xf(2*nf+s)     = mg % zf(s)
  end do

  ! Compute_Geometry's inputs beyond the grid: every material a fluid,
  ! every boundary a wall
  if(allocated(Scoef))      deallocate(Scoef)
  if(allocated(fF))         deallocate(fF)
  if(allocated(IsNearWall)) deallocate(IsNearWall)
  if(allocated(xp))         deallocate(xp, yp, zp)
  allocate(Scoef(nf), fF(nf), IsNearWall(-nb:nc))
  allocate(xp(mg % n_materials), yp(mg % n_materials), zp(mg % n_materials))
  xp = 0.0;  yp = 0.0;  zp = 0.0
  StateMat = FLUID
  TypeBC = WALL
  call Compute_Geometry(mg)

  if(allocated(m_d)) deallocate(m_d, m_scoef)
  allocate(m_d(3*nf), m_scoef(nf))
  do s = 1, nf
    m_d(s) = mg % dx(s);  m_d(nf+s) = mg % dy(s);  m_d(2*nf+s) = mg % dz(s)
    m_scoef(s) = Scoef(s)
  end do

  if(allocated(A % row)) deallocate(A % row)
  if(allocated(A % col)) deallocate(A % col)
  if(allocated(A % dia)) deallocate(A % dia)
  if(allocated(A % val)) deallocate(A % val)
  if(allocated(A % mir)) deallocate(A % mir)
  if(allocated(A % pos)) deallocate(A % pos)
  if(allocated(A % sav)) deallocate(A % sav)
  if(allocated(A % bou)) deallocate(A % bou)
  call Matrix_Mod_Allocate(mg, A)
  call Matrix_Mod_Topology(mg, A)
  dt = dt_in
  call Pressure_Matrix_Fractional(mg)

  m_nnz = A % nonzeros - 1
  if(allocated(m_row)) deallocate(m_row, m_col, m_dia, m_val)
  allocate(m_row(nc+1), m_col(m_nnz), m_dia(nc), m_val(m_nnz))
  m_row = A % row(1:nc+1)
  m_col = A % col(1:m_nnz)
  m_dia = A % dia(1:nc)
  m_val = A % val(1:m_nnz)

  m_nc = nc;  m_nb = nb;  m_nf = nf
  sizes(1) = nc;  sizes(2) = nb;  sizes(3) = nf;  sizes(4) = m_nnz
  tfs_ref_mesh_load = 0
  end function

  !---------------------------------------------------------------------------!
  !   The least-squares gradients on the grid last loaded, every boundary      !
  !   cell -b of kind bc(b): 0 WALL, 1 SYMMETRY, 2 PRESSURE, 3 BUFFER.          !
  !   mode 0: Compute_Gradient_Matrix(grid, .TRUE.) -> g(6*nc), G(1:6,c)       !
  !   mode 1: GraPhi(grid, phi, i, phi_i, .TRUE.) for i = 1, 2, 3 -> px, py, pz!
  !   mode 2: GradP(grid, phi, px, py, pz); phi's boundary part rewritten      !
  !   phi holds cells 1..nc, then boundary cells -1..-nb (nc+nb entries).      !
  !   Modes 1 and 2 use the G of the last mode-0 call.                         !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_mesh_grad(mode, bc, phi, px, py, pz, gout) &
                 bind(C, name='tfs_ref_mesh_grad')
  integer(c_int), value          :: mode
  integer(c_signed_char), intent(in) :: bc(*)
  real(c_double), intent(inout)  :: phi(*)
  real(c_double), intent(out)    :: px(*), py(*), pz(*), gout(*)
  real, allocatable :: f(:), fx(:), fy(:), fz(:)
  integer :: c, nc, nb
  integer, parameter :: kinds(0:3) = (/WALL, SYMMETRY, PRESSURE, BUFFER/)

  tfs_ref_mesh_grad = -1
  if(.not. allocated(mg)) return
  nc = m_nc;  nb = m_nb
  do c = 1, nb
    TypeBC(-c) = kinds(bc(c))
  end do
  if(mode == 0) then
    if(allocated(G)) deallocate(G)
    allocate(G(6, nc))
    call Compute_Gradient_Matrix(mg, .TRUE.)
    do c = 1, nc
      gout(6*(c-1)+1 : 6*c) = G(1:6, c)
    end do
    tfs_ref_mesh_grad = 0
    return
  end if
  if(.not. allocated(G)) return
  allocate(f(-nb:nc), fx(-nb:nc), fy(-nb:nc), fz(-nb:nc))
  fx = 0.0;  fy = 0.0;  fz = 0.0
  f(1:nc) = phi(1:nc)
  do c = 1, nb
    f(-c) = phi(nc + c)
  end do
  if(mode == 1) then
    call GraPhi(mg, f, 1, fx, .TRUE.)
    call GraPhi(mg, f, 2, fy, .TRUE.)
    call GraPhi(mg, f, 3, fz, .TRUE.)
  else if(mode == 2) then
    call GradP(mg, f, fx, fy, fz)
  else
    return
  end if
  px(1:nc) = fx(1:nc);  py(1:nc) = fy(1:nc);  pz(1:nc) = fz(1:nc)
  phi(1:nc) = f(1:nc)
  do c = 1, nb
    phi(nc + c) = f(-c)
  end do
  tfs_ref_mesh_grad = 0
  end function

  !---------------------------------------------------------------------------!
  !   Correct_Velocity on the grid and pressure matrix last loaded (material  !
  !   1 everywhere, density den, viscosity visc, time step as loaded).        !
  !   u, v, w, pp: cells then boundary cells; px, py, pz, sav: cells; flux:   !
  !   faces; bc as for tfs_ref_mesh_grad.  u, v, w, flux updated in place;    !
  !   b(nc) <- the mass error; res <- errmax, and the CFL and Peclet maxima    !
  !   as the reference reports them (read back from its 1pe9.3 text).         !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_mesh_correct(bc, cu, cv, cw, gx, gy, gz, cpp, csav, &
                                               cflux, den, cvisc, bout, res)     &
                 bind(C, name='tfs_ref_mesh_correct')
  integer(c_signed_char), intent(in) :: bc(*)
  real(c_double), intent(inout) :: cu(*), cv(*), cw(*), cflux(*)
  real(c_double), intent(in)    :: gx(*), gy(*), gz(*), cpp(*), csav(*)
  real(c_double), value         :: den, cvisc
  real(c_double), intent(out)   :: bout(*), res(3)
  real, external :: Correct_Velocity
  integer :: c, nc, nb, nf
  integer, parameter :: kinds(0:3) = (/WALL, SYMMETRY, PRESSURE, BUFFER/)

  tfs_ref_mesh_correct = -1
  if(.not. allocated(mg)) return
  nc = m_nc;  nb = m_nb;  nf = m_nf
  do c = 1, nb
    TypeBC(-c) = kinds(bc(c))
  end do
  material = 1
  DENc(1) = den
  VISc    = cvisc
  ALGOR   = FRACT
  if(allocated(U % n)) deallocate(U % n, V % n, W % n, PP % n, P % x, P % y, P % z)
  if(allocated(Flux))  deallocate(Flux)
  if(allocated(b))     deallocate(b)
  allocate(U % n(-nb:nc), V % n(-nb:nc), W % n(-nb:nc), PP % n(-nb:nc))
  allocate(P % x(-nb:nc), P % y(-nb:nc), P % z(-nb:nc), Flux(nf), b(-nb:nc))
  P % x = 0.0;  P % y = 0.0;  P % z = 0.0;  b = 0.0
  U % n(1:nc) = cu(1:nc);  V % n(1:nc) = cv(1:nc);  W % n(1:nc) = cw(1:nc)
  PP % n(1:nc) = cpp(1:nc)
  do c = 1, nb
    U % n(-c) = cu(nc+c);  V % n(-c) = cv(nc+c);  W % n(-c) = cw(nc+c)
    PP % n(-c) = cpp(nc+c)
  end do
  P % x(1:nc) = gx(1:nc);  P % y(1:nc) = gy(1:nc);  P % z(1:nc) = gz(1:nc)
  A % sav(1:nc) = csav(1:nc)
  A % bou = 0.0
  Flux(1:nf) = cflux(1:nf)

  res(1) = Correct_Velocity(mg)
  read(bulk_info % lines(1)(29:37), *) res(2)
  read(bulk_info % lines(1)(68:76), *) res(3)

  cu(1:nc) = U % n(1:nc);  cv(1:nc) = V % n(1:nc);  cw(1:nc) = W % n(1:nc)
  do c = 1, nb
    cu(nc+c) = U % n(-c);  cv(nc+c) = V % n(-c);  cw(nc+c) = W % n(-c)
  end do
  cflux(1:nf) = Flux(1:nf)
  bout(1:nc) = b(1:nc)
  tfs_ref_mesh_correct = 0
  end function

  !---------------------------------------------------------------------------!
  !   Compute_Pressure_Fractional on the grid and matrix last loaded (every    !
  !   material a fluid of density den, every boundary a wall): the face       !
  !   fluxes, the pressure-correction matrix, its Cg solve (prec, tolerance   !
  !   stol, cap 200 as the routine sets it) and the pressure update by urf.    !
  !   cu, cv, cw, cp: cells then boundary cells; gx, gy, gz, csav: cells.     !
  !   Out: flux(nf), val(nnz) = A % val, pp(nc) = PP % n, pn(nc) = P % n.     !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_mesh_press_frac(cu, cv, cw, cp, gx, gy, gz, csav, den, &
                          cprec, stol, urf, cflux, cval, cpp, cpn)                       &
                 bind(C, name='tfs_ref_mesh_press_frac')
  real(c_double), intent(in)  :: cu(*), cv(*), cw(*), cp(*), gx(*), gy(*), gz(*), csav(*)
  real(c_double), value       :: den, stol, urf
  integer(c_int), value       :: cprec
  real(c_double), intent(out) :: cflux(*), cval(*), cpp(*), cpn(*)
  integer :: c, nc, nb, nf

  tfs_ref_mesh_press_frac = -1
  if(.not. allocated(mg)) return
  nc = m_nc;  nb = m_nb;  nf = m_nf
  TypeBC(-nb:-1) = WALL
  material = 1
  StateMat = FLUID
  DENc(1) = den
  ALGOR   = FRACT
  PREC    = cprec
  PP % STol = stol
  P % URF   = urf
  if(allocated(U % n)) deallocate(U % n, V % n, W % n, PP % n, P % x, P % y, P % z)
  if(allocated(P % n)) deallocate(P % n)
  if(allocated(Flux))  deallocate(Flux)
  if(allocated(b))     deallocate(b)
  allocate(U % n(-nb:nc), V % n(-nb:nc), W % n(-nb:nc), PP % n(-nb:nc), P % n(-nb:nc))
  allocate(P % x(-nb:nc), P % y(-nb:nc), P % z(-nb:nc), Flux(nf), b(nc))
  do c = 1, nc
    U % n(c) = cu(c);  V % n(c) = cv(c);  W % n(c) = cw(c);  P % n(c) = cp(c)
    P % x(c) = gx(c);  P % y(c) = gy(c);  P % z(c) = gz(c)
  end do
  do c = 1, nb
    U % n(-c) = cu(nc+c);  V % n(-c) = cv(nc+c);  W % n(-c) = cw(nc+c);  P % n(-c) = cp(nc+c)
    P % x(-c) = 0.0;  P % y(-c) = 0.0;  P % z(-c) = 0.0
  end do
  PP % n = 0.0
  A % sav(1:nc) = csav(1:nc)
  ! the solver's work arrays and preconditioner matrix on this grid
  if(allocated(r_cell_01)) deallocate(r_cell_01, r_cell_02, r_cell_03, r_cell_04,      &
                                      r_cell_05, r_cell_06, r_cell_07, r_cell_08, r_cell_09)
  call Work_Mod_Allocate_Real_Cells(mg, 9)
  if(allocated(D % row)) deallocate(D % row, D % col, D % dia, D % val)
  allocate(D % row(size(A % row)), D % col(size(A % col)), D % dia(size(A % dia)),      &
           D % val(size(A % val)))
  D % row = A % row;  D % col = A % col;  D % dia = A % dia;  D % val = 0.0
  D % nonzeros = A % nonzeros
  D % pnt_grid => mg

  call Compute_Pressure_Fractional(mg)

  cflux(1:nf)    = Flux(1:nf)
  cval(1:m_nnz)  = A % val(1:m_nnz)
  cpp(1:nc)      = PP % n(1:nc)
  cpn(1:nc)      = P % n(1:nc)
  tfs_ref_mesh_press_frac = 0
  end function

  !---------------------------------------------------------------------------!
  !   Compute_Pressure_Simple on the grid and matrix last loaded (every        !
  !   material a fluid of density den): boundary cell -b of kind bc(b) (0 WALL, !
  !   1 SYMMETRY, 2 PRESSURE, 4 INFLOW, 5 OUTFLOW, 6 CONVECT), the face fluxes, !
  !   the pressure-correction matrix, its Bicg solve (prec, tolerance stol,    !
  !   cap 40 as the routine sets it) and the pressure update by urf.           !
  !   cu, cv, cw, cp: cells then boundary cells; gx, gy, gz, csav: cells.     !
  !   Out: flux(nf), val(nnz) = A % val, pp(nc) = PP % n, pn(nc) = P % n.     !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_mesh_press_simple(cu, cv, cw, cp, gx, gy, gz, csav, den, &
                          bc, cprec, stol, urf, cflux, cval, cpp, cpn)                     &
                 bind(C, name='tfs_ref_mesh_press_simple')
  real(c_double), intent(in)  :: cu(*), cv(*), cw(*), cp(*), gx(*), gy(*), gz(*), csav(*)
  real(c_double), value       :: den, stol, urf
  integer(c_int), intent(in)  :: bc(*)
  integer(c_int), value       :: cprec
  real(c_double), intent(out) :: cflux(*), cval(*), cpp(*), cpn(*)
  integer, parameter :: kinds(0:6) = (/WALL, SYMMETRY, PRESSURE, BUFFER, INFLOW, OUTFLOW,  &
                                       CONVECT/)
  integer :: c, nc, nb, nf

  tfs_ref_mesh_press_simple = -1
  if(.not. allocated(mg)) return
  nc = m_nc;  nb = m_nb;  nf = m_nf
  do c = 1, nb
    TypeBC(-c) = kinds(bc(c))
  end do
  material = 1
  StateMat = FLUID
  DENc(1) = den
  ALGOR   = SIMPLE
  PREC    = cprec
  PP % STol = stol
  P % URF   = urf
  if(allocated(U % n)) deallocate(U % n, V % n, W % n, PP % n, P % x, P % y, P % z)
  if(allocated(P % n)) deallocate(P % n)
  if(allocated(Flux))  deallocate(Flux)
  if(allocated(b))     deallocate(b)
  allocate(U % n(-nb:nc), V % n(-nb:nc), W % n(-nb:nc), PP % n(-nb:nc), P % n(-nb:nc))
  allocate(P % x(-nb:nc), P % y(-nb:nc), P % z(-nb:nc), Flux(nf), b(nc))
  do c = 1, nc
    U % n(c) = cu(c);  V % n(c) = cv(c);  W % n(c) = cw(c);  P % n(c) = cp(c)
    P % x(c) = gx(c);  P % y(c) = gy(c);  P % z(c) = gz(c)
  end do
  do c = 1, nb
    U % n(-c) = cu(nc+c);  V % n(-c) = cv(nc+c);  W % n(-c) = cw(nc+c);  P % n(-c) = cp(nc+c)
    P % x(-c) = 0.0;  P % y(-c) = 0.0;  P % z(-c) = 0.0
  end do
  PP % n = 0.0
  A % sav(1:nc) = csav(1:nc)
  if(allocated(r_cell_01)) deallocate(r_cell_01, r_cell_02, r_cell_03, r_cell_04,      &
                                      r_cell_05, r_cell_06, r_cell_07, r_cell_08, r_cell_09)
  call Work_Mod_Allocate_Real_Cells(mg, 9)
  if(allocated(D % row)) deallocate(D % row, D % col, D % dia, D % val)
  allocate(D % row(size(A % row)), D % col(size(A % col)), D % dia(size(A % dia)),      &
           D % val(size(A % val)))
  D % row = A % row;  D % col = A % col;  D % dia = A % dia;  D % val = 0.0
  D % nonzeros = A % nonzeros
  D % pnt_grid => mg

  call Compute_Pressure_Simple(mg)

  cflux(1:nf)    = Flux(1:nf)
  cval(1:m_nnz)  = A % val(1:m_nnz)
  cpp(1:nc)      = PP % n(1:nc)
  cpn(1:nc)      = P % n(1:nc)
  TypeBC(-nb:-1) = WALL
  tfs_ref_mesh_press_simple = 0
  end function

  !---------------------------------------------------------------------------!
  !   which: 1 faces_c (2*nf, int)   2 row (nc+1, int)   3 col (nnz, int)      !
  !          4 dia (nc, int)         5 cell centres x..,y..,z.. (3*(nc+nb))    !
  !          6 vol (nc)              7 surfaces (3*nf)   8 shifts as read      !
  !          9 f (nf)               10 face centres (3*nf)                     !
  !         11 distances after Compute_Geometry (3*nf)  12 Scoef (nf)          !
  !         13 val (nnz)  14 bcmark (nb, int)  15 CopyC (nb, int), both of    !
  !            boundary cells -1..-nb                                           !
  !   Integer arrays go to iout, real ones to rout; returns the count copied,  !
  !   -1 if cap is too small or which is unknown.                              !
  !---------------------------------------------------------------------------!
  integer(c_int) function tfs_ref_mesh_get(which, cap, iout, rout)          &
                 bind(C, name='tfs_ref_mesh_get')
  integer(c_int), value        :: which, cap
  integer(c_int), intent(out)  :: iout(*)
  real(c_double), intent(out)  :: rout(*)
  integer :: n

  tfs_ref_mesh_get = -1
  select case(which)
    case(1);  n = size(m_faces); if(n > cap) return; iout(1:n) = m_faces
    case(2);  n = size(m_row);   if(n > cap) return; iout(1:n) = m_row
    case(3);  n = size(m_col);   if(n > cap) return; iout(1:n) = m_col
    case(4);  n = size(m_dia);   if(n > cap) return; iout(1:n) = m_dia
    case(5);  n = size(m_xc);    if(n > cap) return; rout(1:n) = m_xc
    case(6);  n = size(m_vol);   if(n > cap) return; rout(1:n) = m_vol
    case(7);  n = size(m_s);     if(n > cap) return; rout(1:n) = m_s
    case(8);  n = size(m_shift); if(n > cap) return; rout(1:n) = m_shift
    case(9);  n = size(m_f);     if(n > cap) return; rout(1:n) = m_f
    case(10); n = size(m_xf);    if(n > cap) return; rout(1:n) = m_xf
    case(11); n = size(m_d);     if(n > cap) return; rout(1:n) = m_d
    case(12); n = size(m_scoef); if(n > cap) return; rout(1:n) = m_scoef
    case(13); n = size(m_val);   if(n > cap) return; rout(1:n) = m_val
    case(14); n = m_nb; if(n > cap) return; iout(1:n) = bcmark(-1:-m_nb:-1)
    case(15); n = m_nb; if(n > cap) return; iout(1:n) = CopyC(-1:-m_nb:-1)
    case default; return
  end select
  tfs_ref_mesh_get = n
  end function

  end module
