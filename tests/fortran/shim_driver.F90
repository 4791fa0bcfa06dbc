! Reads a system written by tests/test_fortran_shim.py, calls the drop-in
! pcg_solver_{standard,chrongear}_3d (nz > 0) or _2d (nz = 0 in the header)
! exactly as glissade_velo_higher.F90 does (:3245-3271, :3288-3316), and
! writes the solution, err and niters back.
program shim_driver
  use glimmer_global, only: dp
  use cism_parallel, only: parallel_type
  use glissade_velo_higher_pcg, only: pcg_solver_standard_3d, pcg_solver_chrongear_3d, &
                                          pcg_solver_standard_2d, pcg_solver_chrongear_2d
  implicit none
  integer :: hdr(12), nx, ny, nz, bc, algo, precond, ncheck, maxiters, niters, i, j, k, m, u
  real(dp) :: tol, err
  type(parallel_type) :: parallel
  integer :: indxA_2d(-1:1,-1:1), indxA_3d(-1:1,-1:1,-1:1)
  integer, allocatable :: act(:,:)
  logical, allocatable :: active_vertex(:,:)
  real(dp), allocatable :: Auu(:,:,:,:), Auv(:,:,:,:), Avu(:,:,:,:), Avv(:,:,:,:)
  real(dp), allocatable :: bu(:,:,:), bv(:,:,:), xu(:,:,:), xv(:,:,:)
  real(dp), allocatable :: A2uu(:,:,:), A2uv(:,:,:), A2vu(:,:,:), A2vv(:,:,:)
  real(dp), allocatable :: b2u(:,:), b2v(:,:), x2u(:,:), x2v(:,:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) tol
  nx = hdr(1); ny = hdr(2); nz = hdr(3)
  parallel%staggered_ilo = hdr(4); parallel%staggered_ihi = hdr(5)
  parallel%staggered_jlo = hdr(6); parallel%staggered_jhi = hdr(7)
  bc = hdr(8); algo = hdr(9); precond = hdr(10); ncheck = hdr(11); maxiters = hdr(12)
  parallel%periodic_bc = (bc == 0 .or. bc == 2)
  parallel%outflow_bc = (bc == 1)
  parallel%no_ice_bc = (bc == 3)
  allocate(act(nx-1,ny-1), active_vertex(nx-1,ny-1))
  read(u) act
  if (nz > 0) then
     allocate(Auu(27,nz,nx-1,ny-1), Auv(27,nz,nx-1,ny-1), Avu(27,nz,nx-1,ny-1), Avv(27,nz,nx-1,ny-1))
     allocate(bu(nz,nx-1,ny-1), bv(nz,nx-1,ny-1), xu(nz,nx-1,ny-1), xv(nz,nx-1,ny-1))
     read(u) Auu, Auv, Avu, Avv, bu, bv, xu, xv
  else
     allocate(A2uu(nx-1,ny-1,9), A2uv(nx-1,ny-1,9), A2vu(nx-1,ny-1,9), A2vv(nx-1,ny-1,9))
     allocate(b2u(nx-1,ny-1), b2v(nx-1,ny-1), x2u(nx-1,ny-1), x2v(nx-1,ny-1))
     read(u) A2uu, A2uv, A2vu, A2vv, b2u, b2v, x2u, x2v
  endif
  close(u)
  active_vertex = act /= 0
  ! neighbour numbering of glissade_velo_higher_init (glissade_velo_higher.F90:619-640)
  m = 0
  do k = -1, 1
     do j = -1, 1
        do i = -1, 1
           m = m + 1
           indxA_3d(i,j,k) = m
        enddo
     enddo
  enddo
  m = 0
  do j = -1, 1
     do i = -1, 1
        m = m + 1
        indxA_2d(i,j) = m
     enddo
  enddo
  if (nz == 0) then
     if (algo == 3) then
        call pcg_solver_chrongear_2d(nx, ny, parallel, indxA_2d, active_vertex, &
                                     A2uu, A2uv, A2vu, A2vv, b2u, b2v, x2u, x2v, precond, ncheck, tol, &
                                     maxiters, err, niters, 1, 1, 0)
     else
        call pcg_solver_standard_2d(nx, ny, parallel, indxA_2d, active_vertex, &
                                    A2uu, A2uv, A2vu, A2vv, b2u, b2v, x2u, x2v, precond, ncheck, tol, &
                                    maxiters, err, niters, 1, 1, 0)
     endif
  elseif (algo == 3) then
     call pcg_solver_chrongear_3d(nx, ny, nz, parallel, indxA_2d, indxA_3d, active_vertex, &
                                  Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, ncheck, tol, maxiters, &
                                  err, niters, 1, 1, 0)
  else
     call pcg_solver_standard_3d(nx, ny, nz, parallel, indxA_2d, indxA_3d, active_vertex, &
                                 Auu, Auv, Avu, Avv, bu, bv, xu, xv, precond, ncheck, tol, maxiters, &
                                 err, niters, 1, 1, 0)
  endif
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) niters
  write(u) err
  if (nz == 0) then
     write(u) x2u, x2v
  else
     write(u) xu, xv
  endif
  close(u)
end program shim_driver
