! Minimal test doubles of the three CISM modules the drop-in shim uses
! (glimmer_global, glimmer_log, cism_parallel), so that the shim can be
! compiled and exercised outside a CISM build.  Field names follow
! parallel_type (parallel_mpi.F90:70-150).
module glimmer_global
  implicit none
  integer, parameter :: dp = kind(1.0d0)
end module glimmer_global

module glimmer_log
  implicit none
  integer, parameter :: GM_FATAL = 6
contains
  subroutine write_log(msg, type)
    character(len=*), intent(in) :: msg
    integer, intent(in), optional :: type
    write(*,'(a,a)') 'write_log: ', trim(msg)
    if (present(type)) then
       if (type == GM_FATAL) stop 3
    endif
  end subroutine write_log
end module glimmer_log

module cism_parallel
  implicit none
  type :: parallel_type
     integer :: staggered_ilo, staggered_ihi, staggered_jlo, staggered_jhi
     integer :: tasks_row = 1, tasks_col = 1
     logical :: periodic_bc = .false., outflow_bc = .false., no_ice_bc = .false.
     logical :: reprosum = .false.
  end type parallel_type
end module cism_parallel
