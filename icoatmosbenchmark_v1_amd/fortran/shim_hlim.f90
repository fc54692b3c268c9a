!-------------------------------------------------------------------------------
!> Drop-in module for kernels/NICAM/dyn_horiz_adv_limiter/src/mod_src_tracer.f90
!> (horizontal_limiter_thuburn, :42-430): same module name, procedure name and
!> dummy arguments, the work done by libnicam_tt on the GPU.  Compile it instead
!> of that file, against the reference's mod_precision and mod_misc, add
!> nicam_tt_f.f90 and link libnicam_tt.so; the driver source is unchanged.
!>
!> The first call creates a context from the mod_misc globals (ADM_* sizes,
!> ADM_have_sgp, ADM_have_pl, CONST_EPS, CONST_HUGE).  Arrays are host arrays
!> (NICAM_PTR_HOST), RP must be DP.
!-------------------------------------------------------------------------------
module mod_src_tracer
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: horizontal_limiter_thuburn

  type(c_ptr), save :: ctx = c_null_ptr

contains

  subroutine setup
    integer(c_int) :: sgp(ADM_lall)
    integer :: l
    if (c_associated(ctx)) return
    if (RP /= c_double) then
       write(*, *) 'xxx the libnicam_tt shim needs RP = DP'
       error stop 1
    end if
    do l = 1, ADM_lall
       sgp(l) = merge(1, 0, ADM_have_sgp(l))
    end do
    call nicam_check(nicam_ctx_create(ctx, ADM_gall_1d, ADM_kall, ADM_lall, ADM_gall_pl, &
                                      merge(1, 0, ADM_have_pl), sgp), 'nicam_ctx_create')
    call nicam_check(nicam_ctx_set_pointer_mode(ctx, NICAM_PTR_HOST), 'nicam_ctx_set_pointer_mode')
    call nicam_check(nicam_set_const(ctx, real(CONST_EPS, c_double), real(CONST_HUGE, c_double)), &
                     'nicam_set_const')
  end subroutine setup

  subroutine horizontal_limiter_thuburn( &
       q_a,    q_a_pl,  &
       q,      q_pl,    &
       d,      d_pl,    &
       ch,     ch_pl,   &
       cmask,  cmask_pl, &
       Qout_prev, Qout_prev_pl, &
       Qout_post, Qout_post_pl  )
    real(RP), intent(inout), target :: q_a     (ADM_gall   ,ADM_kall,ADM_lall   ,6)
    real(RP), intent(inout), target :: q_a_pl  (ADM_gall_pl,ADM_kall,ADM_lall_pl  )
    real(RP), intent(in),    target :: q       (ADM_gall   ,ADM_kall,ADM_lall     )
    real(RP), intent(in),    target :: q_pl    (ADM_gall_pl,ADM_kall,ADM_lall_pl  )
    real(RP), intent(in),    target :: d       (ADM_gall   ,ADM_kall,ADM_lall     )
    real(RP), intent(in),    target :: d_pl    (ADM_gall_pl,ADM_kall,ADM_lall_pl  )
    real(RP), intent(in),    target :: ch      (ADM_gall   ,ADM_kall,ADM_lall   ,6)
    real(RP), intent(in),    target :: ch_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl  )
    real(RP), intent(in),    target :: cmask   (ADM_gall   ,ADM_kall,ADM_lall   ,6)
    real(RP), intent(in),    target :: cmask_pl(ADM_gall_pl,ADM_kall,ADM_lall_pl  )
    real(RP), intent(inout), target :: Qout_prev   (ADM_gall   ,ADM_kall,ADM_lall   ,2)
    real(RP), intent(inout), target :: Qout_prev_pl(ADM_gall_pl,ADM_kall,ADM_lall_pl,2)
    real(RP), intent(in),    target :: Qout_post   (ADM_gall   ,ADM_kall,ADM_lall   ,2)
    real(RP), intent(in),    target :: Qout_post_pl(ADM_gall_pl,ADM_kall,ADM_lall_pl,2)

    call setup
    call nicam_check(nicam_horizontal_limiter_thuburn(ctx, c_loc(q_a), c_loc(q_a_pl), c_loc(q), c_loc(q_pl),    &
                                                      c_loc(d), c_loc(d_pl), c_loc(ch), c_loc(ch_pl),            &
                                                      c_loc(cmask), c_loc(cmask_pl), c_loc(Qout_prev),           &
                                                      c_loc(Qout_prev_pl), c_loc(Qout_post), c_loc(Qout_post_pl)), &
                     'nicam_horizontal_limiter_thuburn')
  end subroutine horizontal_limiter_thuburn

end module mod_src_tracer
