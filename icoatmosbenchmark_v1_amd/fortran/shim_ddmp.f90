!-------------------------------------------------------------------------------
!> Drop-in module for kernels/NICAM/dyn_divdamp/src/mod_oprt3d.f90
!> (OPRT3D_divdamp, :41-402): same module name, procedure name and dummy
!> arguments, the work done by libnicam_tt on the GPU.  Compile it instead of
!> that file, against the reference's mod_precision and mod_misc, add
!> nicam_tt_f.f90 and link libnicam_tt.so; the driver source is unchanged.
!>
!> The first call creates a context from the mod_misc globals (ADM_* sizes,
!> ADM_have_sgp, ADM_have_pl) and uploads what the reference reads from
!> mod_misc: VMTR_RGSQRTH, VMTR_RGAM, VMTR_RGAMH, VMTR_C2WfactGz and their pole
!> forms, and GRD_rdgz (after GRD_setup).  Arrays are host arrays
!> (NICAM_PTR_HOST), RP must be DP.
!-------------------------------------------------------------------------------
module mod_oprt3d
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: OPRT3D_divdamp

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
    call nicam_check(nicam_set_vmetrics3d(ctx, VMTR_RGSQRTH, VMTR_RGSQRTH_pl, VMTR_RGAM, VMTR_RGAM_pl, &
                                          VMTR_RGAMH, VMTR_RGAMH_pl, VMTR_C2WfactGz, VMTR_C2WfactGz_pl), &
                     'nicam_set_vmetrics3d')
    call nicam_check(nicam_set_rdgz(ctx, GRD_rdgz), 'nicam_set_rdgz')
  end subroutine setup

  subroutine OPRT3D_divdamp( &
       ddivdx,    ddivdx_pl,    &
       ddivdy,    ddivdy_pl,    &
       ddivdz,    ddivdz_pl,    &
       rhogvx,    rhogvx_pl,    &
       rhogvy,    rhogvy_pl,    &
       rhogvz,    rhogvz_pl,    &
       rhogw,     rhogw_pl,     &
       coef_intp, coef_intp_pl, &
       coef_diff, coef_diff_pl  )
    real(RP), intent(inout), target :: ddivdx      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(inout), target :: ddivdx_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(inout), target :: ddivdy      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(inout), target :: ddivdy_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(inout), target :: ddivdz      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(inout), target :: ddivdz_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhogvx      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhogvx_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhogvy      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhogvy_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhogvz      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhogvz_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhogw       (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhogw_pl    (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: coef_intp   (ADM_gall   ,1:3,ADM_nxyz,TI:TJ,ADM_lall   )
    real(RP), intent(in),    target :: coef_intp_pl(ADM_gall_pl,1:3,ADM_nxyz,      ADM_lall_pl)
    real(RP), intent(in),    target :: coef_diff   (ADM_gall   ,1:6,ADM_nxyz,      ADM_lall   )
    real(RP), intent(in),    target :: coef_diff_pl(           1:ADM_vlink,ADM_nxyz,ADM_lall_pl)

    call setup
    call nicam_check(nicam_oprt3d_divdamp(ctx, c_loc(ddivdx), c_loc(ddivdx_pl), c_loc(ddivdy), c_loc(ddivdy_pl), &
                                          c_loc(ddivdz), c_loc(ddivdz_pl), c_loc(rhogvx), c_loc(rhogvx_pl),      &
                                          c_loc(rhogvy), c_loc(rhogvy_pl), c_loc(rhogvz), c_loc(rhogvz_pl),      &
                                          c_loc(rhogw), c_loc(rhogw_pl), c_loc(coef_intp), c_loc(coef_intp_pl),  &
                                          c_loc(coef_diff), c_loc(coef_diff_pl)), 'nicam_oprt3d_divdamp')
  end subroutine OPRT3D_divdamp

end module mod_oprt3d
