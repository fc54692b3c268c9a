!-------------------------------------------------------------------------------
! Harness around the reference's own Fortran modules (TEST INFRASTRUCTURE ONLY).
!
! Compiled together with the reference's mod_precision.f90 / mod_misc.f90 and
! one kernel module, straight from the reference tree, into oracle/_ref/.
! It sets the size variables that the reference keeps as module globals
! (kernels/NICAM/*/src/problem_size.inc:28-50) for ONE region at a time, since
! ADM_have_sgp is declared with extent 1 there (problem_size.inc:50).
!-------------------------------------------------------------------------------
module ref_harness_common
  use iso_c_binding
  use mod_precision
  use mod_misc
  implicit none
  private
  public :: harness_set_sizes
  public :: ref_const_setup
  public :: ref_set_eps

  logical, save :: log_opened = .false.
  ! CONST_EPS to install after CONST_setup; negative = keep CONST_setup's
  real(c_double), save :: eps_override = -1.0_c_double

contains

  subroutine harness_set_sizes( gall_1d, kall, gall_pl, have_pl, have_sgp )
    integer, intent(in) :: gall_1d, kall, gall_pl
    logical, intent(in) :: have_pl, have_sgp

    if ( .not. log_opened ) then
       ! CONST_setup and the timers write to ADM_LOG_FID; keep them quiet.
       ADM_LOG_FID = 77
       open( unit=ADM_LOG_FID, file='/dev/null', status='old', action='write' )
       log_opened = .true.
    endif

    ADM_lall        = 1
    ADM_gall_1d     = gall_1d
    ADM_gall        = gall_1d * gall_1d
    ADM_gmin        = 2
    ADM_gmax        = gall_1d - 1
    ADM_gall_pl     = gall_pl
    ADM_gmax_pl     = gall_pl
    ADM_kall        = kall
    ADM_kmin        = 2
    ADM_kmax        = kall - 1
    ADM_vlayer      = kall - 2
    ADM_have_pl     = have_pl
    ADM_have_sgp(1) = have_sgp

    call CONST_setup
    if ( eps_override >= 0.0_c_double ) CONST_EPS = real( eps_override, RP )
  end subroutine harness_set_sizes

  !> Replaces CONST_EPS after every CONST_setup (a negative value restores it)
  subroutine ref_set_eps( eps ) bind(C, name='ref_set_eps')
    real(c_double), value, intent(in) :: eps
    eps_override = eps
  end subroutine ref_set_eps

  !> CONST_EPS / CONST_HUGE as CONST_setup leaves them
  subroutine ref_const_setup( eps, big ) bind(C, name='ref_const_setup')
    real(c_double), intent(out) :: eps, big
    call harness_set_sizes( 4, 4, 6, .false., .false. )
    eps = CONST_EPS
    big = CONST_HUGE
  end subroutine ref_const_setup

end module ref_harness_common
