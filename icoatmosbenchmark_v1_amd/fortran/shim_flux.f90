!-------------------------------------------------------------------------------
!> Drop-in module for kernels/NICAM/dyn_horiz_adv_flux/src/mod_src_tracer.f90
!> (horizontal_flux, :42-379): same module name, procedure name and dummy
!> arguments, the work done by libnicam_tt on the GPU.  Compile it instead of
!> that file, against the reference's mod_precision and mod_misc, add
!> nicam_tt_f.f90 and link libnicam_tt.so; the driver source is unchanged.
!>
!> The first call creates a context from the mod_misc globals (ADM_* sizes,
!> ADM_have_sgp, ADM_have_pl, CONST_EPS, CONST_HUGE) and uploads the
!> horizontal metrics GRD_xr, GMTR_p/t/a and their pole forms; later calls
!> reuse it.  Arrays are host arrays (NICAM_PTR_HOST), RP must be DP.
!-------------------------------------------------------------------------------
module mod_src_tracer
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: horizontal_flux

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
    call nicam_check(nicam_set_hmetrics(ctx, GRD_xr, GRD_xr_pl, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, &
                                        GMTR_a, GMTR_a_pl), 'nicam_set_hmetrics')
  end subroutine setup

  subroutine horizontal_flux( &
       flx_h,  flx_h_pl,  &
       GRD_xc, GRD_xc_pl, &
       rho,    rho_pl,    &
       rhovx,  rhovx_pl,  &
       rhovy,  rhovy_pl,  &
       rhovz,  rhovz_pl,  &
       dt                 )
    ! flx_h and GRD_xc are inout here: the reference leaves some of their
    ! points unwritten, and the library copies them in and back unchanged
    real(RP), intent(inout), target :: flx_h    (ADM_gall   ,ADM_kall,ADM_lall   ,6)
    real(RP), intent(inout), target :: flx_h_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl  )
    real(RP), intent(inout), target :: GRD_xc   (ADM_gall   ,ADM_kall,ADM_lall   ,3,3)
    real(RP), intent(inout), target :: GRD_xc_pl(ADM_gall_pl,ADM_kall,ADM_lall_pl  ,3)
    real(RP), intent(in),    target :: rho      (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rho_pl   (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhovx    (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhovx_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhovy    (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhovy_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhovz    (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhovz_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in)            :: dt

    call setup
    call nicam_check(nicam_horizontal_flux(ctx, c_loc(flx_h), c_loc(flx_h_pl), c_loc(GRD_xc), c_loc(GRD_xc_pl), &
                                           c_loc(rho), c_loc(rho_pl), c_loc(rhovx), c_loc(rhovx_pl),          &
                                           c_loc(rhovy), c_loc(rhovy_pl), c_loc(rhovz), c_loc(rhovz_pl),      &
                                           real(dt, c_double)), 'nicam_horizontal_flux')
  end subroutine horizontal_flux

end module mod_src_tracer
