!-------------------------------------------------------------------------------
!> Drop-in module for kernels/NICAM/dyn_diffusion/src/mod_oprt.f90
!> (OPRT_diffusion, :41-296): same module name, procedure name and dummy
!> arguments, the work done by libnicam_tt on the GPU.  Compile it instead of
!> that file, against the reference's mod_precision and mod_misc, add
!> nicam_tt_f.f90 and link libnicam_tt.so; the driver source is unchanged.
!>
!> The first call creates a context from the mod_misc globals (ADM_* sizes,
!> ADM_have_sgp, ADM_have_pl).  Arrays are host arrays (NICAM_PTR_HOST), RP
!> must be DP.
!-------------------------------------------------------------------------------
module mod_oprt
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: OPRT_diffusion

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
  end subroutine setup

  subroutine OPRT_diffusion( &
       dscl,      dscl_pl,      &
       scl,       scl_pl,       &
       kh,        kh_pl,        &
       coef_intp, coef_intp_pl, &
       coef_diff, coef_diff_pl  )
    real(RP), intent(inout), target :: dscl        (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(inout), target :: dscl_pl     (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: scl         (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: scl_pl      (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: kh          (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: kh_pl       (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: coef_intp   (ADM_gall   ,1:3,ADM_nxyz,TI:TJ,ADM_lall   )
    real(RP), intent(in),    target :: coef_intp_pl(ADM_gall_pl,1:3,ADM_nxyz,      ADM_lall_pl)
    real(RP), intent(in),    target :: coef_diff   (ADM_gall   ,1:6,ADM_nxyz,      ADM_lall   )
    real(RP), intent(in),    target :: coef_diff_pl(           1:ADM_vlink,ADM_nxyz,ADM_lall_pl)

    call setup
    call nicam_check(nicam_oprt_diffusion(ctx, c_loc(dscl), c_loc(dscl_pl), c_loc(scl), c_loc(scl_pl),         &
                                          c_loc(kh), c_loc(kh_pl), c_loc(coef_intp), c_loc(coef_intp_pl),      &
                                          c_loc(coef_diff), c_loc(coef_diff_pl)), 'nicam_oprt_diffusion')
  end subroutine OPRT_diffusion

end module mod_oprt
