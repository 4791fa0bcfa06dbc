!-----------------------------------------------------------------------------
! glissade_velo_higher_pcg.F90 (drop-in replacement)
!
! Drop-in Fortran entry points for the structured 3-D and 2-D velocity
! solvers of CISM's Glissade dycore, backed by the MI355X library
! libcismgpu.so through its C ABI (include/cismgpu.h).  The module keeps
! the reference module's name, glissade_velo_higher_pcg, and its public
! subroutines keep the names and argument lists of the reference's
! (glissade_velo_higher_pcg.F90:56-58, :67-78, :827-838, :1541-1552,
! :2632-2643, :5528-5540, :5656-5663) and the
! Auu/Auv/Avu/Avv(27,nz,nx-1,ny-1) and (nx-1,ny-1,9) block layouts, so a
! build switches by compiling this file in place of
! libglissade/glissade_velo_higher_pcg.F90; the caller's use statement
! (glissade_velo_higher.F90:77-80) stays as it is.
!
! Arrays are passed by address (no copies on the host side).  A nonzero
! library status is reported through write_log(..., GM_FATAL), as the
! reference does for a NaN in the PCG scalars; non-convergence is silent
! (niters = maxiters), as in the reference.
!-----------------------------------------------------------------------------
module glissade_velo_higher_pcg

  use iso_c_binding, only: c_int32_t, c_double, c_char, c_ptr, c_null_char, c_associated, c_f_pointer
  use ieee_arithmetic, only: ieee_value, ieee_quiet_nan
  use glimmer_global, only: dp
  use glimmer_log, only: write_log, GM_FATAL
  use cism_parallel, only: parallel_type

  implicit none
  private
  public :: pcg_solver_standard_3d, pcg_solver_chrongear_3d, matvec_multiply_structured_3d
  public :: pcg_solver_standard_2d, pcg_solver_chrongear_2d, matvec_multiply_structured_2d

  ! mirror of cismgpu_grid
  type, bind(c) :: cismgpu_grid
     integer(c_int32_t) :: nx, ny, nz
     integer(c_int32_t) :: ilo, ihi, jlo, jhi
     integer(c_int32_t) :: global_bc
     integer(c_int32_t) :: reprosum
  end type cismgpu_grid

  interface
     integer(c_int32_t) function c_pcg_standard(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
          precond, ncheck, tolerance, maxiters, err, niters) bind(c, name='cismgpu_pcg_solver_standard_3d')
       import :: c_int32_t, c_double, cismgpu_grid
       type(cismgpu_grid), intent(in) :: grid
       integer(c_int32_t), intent(in) :: active(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), bu(*), bv(*)
       real(c_double), intent(inout) :: xu(*), xv(*)
       integer(c_int32_t), value :: precond, ncheck
       real(c_double), value :: tolerance
       integer(c_int32_t), value :: maxiters
       real(c_double), intent(inout) :: err
       integer(c_int32_t), intent(inout) :: niters
     end function c_pcg_standard

     integer(c_int32_t) function c_pcg_chrongear(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
          precond, ncheck, tolerance, maxiters, err, niters) bind(c, name='cismgpu_pcg_solver_chrongear_3d')
       import :: c_int32_t, c_double, cismgpu_grid
       type(cismgpu_grid), intent(in) :: grid
       integer(c_int32_t), intent(in) :: active(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), bu(*), bv(*)
       real(c_double), intent(inout) :: xu(*), xv(*)
       integer(c_int32_t), value :: precond, ncheck
       real(c_double), value :: tolerance
       integer(c_int32_t), value :: maxiters
       real(c_double), intent(inout) :: err
       integer(c_int32_t), intent(inout) :: niters
     end function c_pcg_chrongear

     integer(c_int32_t) function c_matvec(grid, active, Auu, Auv, Avu, Avv, xu, xv, yu, yv) &
          bind(c, name='cismgpu_matvec_multiply_structured_3d')
       import :: c_int32_t, c_double, cismgpu_grid
       type(cismgpu_grid), intent(in) :: grid
       integer(c_int32_t), intent(in) :: active(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), xu(*), xv(*)
       real(c_double), intent(inout) :: yu(*), yv(*)
     end function c_matvec

     integer(c_int32_t) function c_pcg_standard_2d(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
          precond, ncheck, tolerance, maxiters, err, niters) bind(c, name='cismgpu_pcg_solver_standard_2d')
       import :: c_int32_t, c_double, cismgpu_grid
       type(cismgpu_grid), intent(in) :: grid
       integer(c_int32_t), intent(in) :: active(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), bu(*), bv(*)
       real(c_double), intent(inout) :: xu(*), xv(*)
       integer(c_int32_t), value :: precond, ncheck
       real(c_double), value :: tolerance
       integer(c_int32_t), value :: maxiters
       real(c_double), intent(inout) :: err
       integer(c_int32_t), intent(inout) :: niters
     end function c_pcg_standard_2d

     integer(c_int32_t) function c_pcg_chrongear_2d(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
          precond, ncheck, tolerance, maxiters, err, niters) bind(c, name='cismgpu_pcg_solver_chrongear_2d')
       import :: c_int32_t, c_double, cismgpu_grid
       type(cismgpu_grid), intent(in) :: grid
       integer(c_int32_t), intent(in) :: active(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), bu(*), bv(*)
       real(c_double), intent(inout) :: xu(*), xv(*)
       integer(c_int32_t), value :: precond, ncheck
       real(c_double), value :: tolerance
       integer(c_int32_t), value :: maxiters
       real(c_double), intent(inout) :: err
       integer(c_int32_t), intent(inout) :: niters
     end function c_pcg_chrongear_2d

     integer(c_int32_t) function c_matvec_2d(grid, active, Auu, Auv, Avu, Avv, xu, xv, yu, yv) &
          bind(c, name='cismgpu_matvec_multiply_structured_2d')
       import :: c_int32_t, c_double, cismgpu_grid
       type(cismgpu_grid), intent(in) :: grid
       integer(c_int32_t), intent(in) :: active(*)
       real(c_double), intent(in) :: Auu(*), Auv(*), Avu(*), Avv(*), xu(*), xv(*)
       real(c_double), intent(inout) :: yu(*), yv(*)
     end function c_matvec_2d

     type(c_ptr) function c_last_error() bind(c, name='cismgpu_last_error')
       import :: c_ptr
     end function c_last_error
  end interface

contains

  ! The kernels use the neighbour numbering of glissade_velo_higher_init
  ! (glissade_velo_higher.F90:607-640).
  subroutine check_indxA_3d(indxA_3d)
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA_3d
    integer :: iA, jA, kA
    do kA = -1, 1
       do jA = -1, 1
          do iA = -1, 1
             if (indxA_3d(iA,jA,kA) /= 1 + (iA+1) + 3*(jA+1) + 9*(kA+1)) then
                call write_log('GPU PCG solver: unexpected indxA_3d ordering', GM_FATAL)
             endif
          enddo
       enddo
    enddo
  end subroutine check_indxA_3d

  subroutine check_indxA_2d(indxA_2d)
    integer, dimension(-1:1,-1:1), intent(in) :: indxA_2d
    integer :: iA, jA
    do jA = -1, 1
       do iA = -1, 1
          if (indxA_2d(iA,jA) /= 1 + (iA+1) + 3*(jA+1)) then
             call write_log('GPU PCG solver: unexpected indxA_2d ordering', GM_FATAL)
          endif
       enddo
    enddo
  end subroutine check_indxA_2d

  ! Grid bounds and halo rule of the one task this library supports.
  subroutine make_grid(nx, ny, nz, parallel, grid)
    integer, intent(in) :: nx, ny, nz
    type(parallel_type), intent(in) :: parallel
    type(cismgpu_grid), intent(out) :: grid

    if (parallel%tasks_row /= 1 .or. parallel%tasks_col /= 1) then
       call write_log('GPU PCG solver: one task per row and column is supported', GM_FATAL)
    endif
    grid%nx = nx
    grid%ny = ny
    grid%nz = nz
    grid%ilo = parallel%staggered_ilo
    grid%ihi = parallel%staggered_ihi
    grid%jlo = parallel%staggered_jlo
    grid%jhi = parallel%staggered_jhi
    if (parallel%no_ice_bc) then
       grid%global_bc = 3
    elseif (parallel%outflow_bc) then
       grid%global_bc = 1
    else
       grid%global_bc = 0          ! periodic halo logic (also used for no_penetration)
    endif
    if (parallel%reprosum) then
       grid%reprosum = 1
    else
       grid%reprosum = 0
    endif
  end subroutine make_grid

  subroutine report(status, what)
    integer(c_int32_t), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    type(c_ptr) :: p
    text = ''
    p = c_last_error()
    if (c_associated(p)) then
       call c_f_pointer(p, msg, [512])
       n = 0
       do while (n < 512)
          if (msg(n+1) == c_null_char) exit
          n = n + 1
          text(n:n) = msg(n)
       enddo
    endif
    if (status == 1) then
       call write_log(what//' has failed, alpha = NaN', GM_FATAL)
    else
       call write_log(what//': '//trim(text), GM_FATAL)
    endif
  end subroutine report

  subroutine pcg_solver_standard_3d(nx,        ny,            &
                                    nz,        parallel,      &
                                    indxA_2d,  indxA_3d,      &
                                    active_vertex,            &
                                    Auu,       Auv,           &
                                    Avu,       Avv,           &
                                    bu,        bv,            &
                                    xu,        xv,            &
                                    precond,   linear_solve_ncheck,  &
                                    tolerance, maxiters,      &
                                    err,       niters,        &
                                    itest, jtest, rtest)
    integer, intent(in) :: nx, ny, nz
    type(parallel_type), intent(in) :: parallel
    integer, dimension(-1:1,-1:1), intent(in) :: indxA_2d
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA_3d
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    integer, intent(in) :: linear_solve_ncheck
    integer, intent(in) :: maxiters
    real(dp), intent(in) :: tolerance
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer, intent(in) :: itest, jtest, rtest

    type(cismgpu_grid) :: grid
    integer(c_int32_t), dimension(nx-1,ny-1) :: active
    integer(c_int32_t) :: status, nit
    real(c_double) :: e

    call check_indxA_3d(indxA_3d)
    call make_grid(nx, ny, nz, parallel, grid)
    active = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    e = ieee_value(e, ieee_quiet_nan)   ! left NaN when no convergence check ran
    nit = 0
    status = c_pcg_standard(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
                            int(precond, c_int32_t), int(linear_solve_ncheck, c_int32_t), &
                            real(tolerance, c_double), int(maxiters, c_int32_t), e, nit)
    if (status /= 0) call report(status, 'PCG solver')
    err = e
    niters = nit
  end subroutine pcg_solver_standard_3d

  subroutine pcg_solver_chrongear_3d(nx,        ny,            &
                                     nz,        parallel,      &
                                     indxA_2d,  indxA_3d,      &
                                     active_vertex,            &
                                     Auu,       Auv,           &
                                     Avu,       Avv,           &
                                     bu,        bv,            &
                                     xu,        xv,            &
                                     precond,   linear_solve_ncheck,  &
                                     tolerance, maxiters,      &
                                     err,       niters,        &
                                     itest, jtest, rtest)
    integer, intent(in) :: nx, ny, nz
    type(parallel_type), intent(in) :: parallel
    integer, dimension(-1:1,-1:1), intent(in) :: indxA_2d
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA_3d
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nz,nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    integer, intent(in) :: linear_solve_ncheck
    integer, intent(in) :: maxiters
    real(dp), intent(in) :: tolerance
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer, intent(in) :: itest, jtest, rtest

    type(cismgpu_grid) :: grid
    integer(c_int32_t), dimension(nx-1,ny-1) :: active
    integer(c_int32_t) :: status, nit
    real(c_double) :: e

    call check_indxA_3d(indxA_3d)
    call make_grid(nx, ny, nz, parallel, grid)
    active = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    e = ieee_value(e, ieee_quiet_nan)   ! left NaN when no convergence check ran
    nit = 0
    status = c_pcg_chrongear(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
                             int(precond, c_int32_t), int(linear_solve_ncheck, c_int32_t), &
                             real(tolerance, c_double), int(maxiters, c_int32_t), e, nit)
    if (status /= 0) call report(status, 'Chron-Gear PCG solver')
    err = e
    niters = nit
  end subroutine pcg_solver_chrongear_3d

  subroutine matvec_multiply_structured_3d(nx,        ny,              &
                                           nz,        parallel,        &
                                           indxA_3d,  active_vertex,   &
                                           Auu,       Auv,             &
                                           Avu,       Avv,             &
                                           xu,        xv,              &
                                           yu,        yv)
    integer, intent(in) :: nx, ny, nz
    type(parallel_type), intent(in) :: parallel
    integer, dimension(-1:1,-1:1,-1:1), intent(in) :: indxA_3d
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(27,nz,nx-1,ny-1), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nz,nx-1,ny-1), intent(in) :: xu, xv
    real(dp), dimension(nz,nx-1,ny-1), intent(out) :: yu, yv

    type(cismgpu_grid) :: grid
    integer(c_int32_t), dimension(nx-1,ny-1) :: active
    integer(c_int32_t) :: status

    call check_indxA_3d(indxA_3d)
    call make_grid(nx, ny, nz, parallel, grid)
    active = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    status = c_matvec(grid, active, Auu, Auv, Avu, Avv, xu, xv, yu, yv)
    if (status /= 0) call report(status, 'matvec_multiply_structured_3d')
  end subroutine matvec_multiply_structured_3d

  subroutine pcg_solver_standard_2d(nx,        ny,            &
                                    parallel,                 &
                                    indxA_2d,  active_vertex, &
                                    Auu,       Auv,           &
                                    Avu,       Avv,           &
                                    bu,        bv,            &
                                    xu,        xv,            &
                                    precond,   linear_solve_ncheck,  &
                                    tolerance, maxiters,      &
                                    err,       niters,        &
                                    itest, jtest, rtest)
    integer, intent(in) :: nx, ny
    type(parallel_type), intent(in) :: parallel
    integer, dimension(-1:1,-1:1), intent(in) :: indxA_2d
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(nx-1,ny-1,9), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    integer, intent(in) :: linear_solve_ncheck
    integer, intent(in) :: maxiters
    real(dp), intent(in) :: tolerance
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer, intent(in) :: itest, jtest, rtest

    type(cismgpu_grid) :: grid
    integer(c_int32_t), dimension(nx-1,ny-1) :: active
    integer(c_int32_t) :: status, nit
    real(c_double) :: e

    call check_indxA_2d(indxA_2d)
    call make_grid(nx, ny, 1, parallel, grid)
    active = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    e = ieee_value(e, ieee_quiet_nan)   ! left NaN when no convergence check ran
    nit = 0
    status = c_pcg_standard_2d(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
                               int(precond, c_int32_t), int(linear_solve_ncheck, c_int32_t), &
                               real(tolerance, c_double), int(maxiters, c_int32_t), e, nit)
    if (status /= 0) call report(status, 'PCG solver')
    err = e
    niters = nit
  end subroutine pcg_solver_standard_2d

  subroutine pcg_solver_chrongear_2d(nx,        ny,            &
                                     parallel,                 &
                                     indxA_2d,  active_vertex, &
                                     Auu,       Auv,           &
                                     Avu,       Avv,           &
                                     bu,        bv,            &
                                     xu,        xv,            &
                                     precond,   linear_solve_ncheck,  &
                                     tolerance, maxiters,      &
                                     err,       niters,        &
                                     itest, jtest, rtest)
    integer, intent(in) :: nx, ny
    type(parallel_type), intent(in) :: parallel
    integer, dimension(-1:1,-1:1), intent(in) :: indxA_2d
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(nx-1,ny-1,9), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nx-1,ny-1), intent(in) :: bu, bv
    real(dp), dimension(nx-1,ny-1), intent(inout) :: xu, xv
    integer, intent(in) :: precond
    integer, intent(in) :: linear_solve_ncheck
    integer, intent(in) :: maxiters
    real(dp), intent(in) :: tolerance
    real(dp), intent(out) :: err
    integer, intent(out) :: niters
    integer, intent(in) :: itest, jtest, rtest

    type(cismgpu_grid) :: grid
    integer(c_int32_t), dimension(nx-1,ny-1) :: active
    integer(c_int32_t) :: status, nit
    real(c_double) :: e

    call check_indxA_2d(indxA_2d)
    call make_grid(nx, ny, 1, parallel, grid)
    active = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    e = ieee_value(e, ieee_quiet_nan)   ! left NaN when no convergence check ran
    nit = 0
    status = c_pcg_chrongear_2d(grid, active, Auu, Auv, Avu, Avv, bu, bv, xu, xv, &
                                int(precond, c_int32_t), int(linear_solve_ncheck, c_int32_t), &
                                real(tolerance, c_double), int(maxiters, c_int32_t), e, nit)
    if (status /= 0) call report(status, 'Chron-Gear PCG solver')
    err = e
    niters = nit
  end subroutine pcg_solver_chrongear_2d

  subroutine matvec_multiply_structured_2d(nx,            ny,            &
                                           parallel,                     &
                                           indxA_2d,      active_vertex, &
                                           Auu,           Auv,           &
                                           Avu,           Avv,           &
                                           xu,            xv,            &
                                           yu,            yv)
    integer, intent(in) :: nx, ny
    type(parallel_type), intent(in) :: parallel
    integer, dimension(-1:1,-1:1), intent(in) :: indxA_2d
    logical, dimension(nx-1,ny-1), intent(in) :: active_vertex
    real(dp), dimension(nx-1,ny-1,9), intent(in) :: Auu, Auv, Avu, Avv
    real(dp), dimension(nx-1,ny-1), intent(in) :: xu, xv
    real(dp), dimension(nx-1,ny-1), intent(out) :: yu, yv

    type(cismgpu_grid) :: grid
    integer(c_int32_t), dimension(nx-1,ny-1) :: active
    integer(c_int32_t) :: status

    call check_indxA_2d(indxA_2d)
    call make_grid(nx, ny, 1, parallel, grid)
    active = merge(1_c_int32_t, 0_c_int32_t, active_vertex)
    status = c_matvec_2d(grid, active, Auu, Auv, Avu, Avv, xu, xv, yu, yv)
    if (status /= 0) call report(status, 'matvec_multiply_structured_2d')
  end subroutine matvec_multiply_structured_2d

end module glissade_velo_higher_pcg
