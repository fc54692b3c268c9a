!-------------------------------------------------------------------------------
!> Drop-in module for kernels/NICAM/dyn_vi_rhow_solver/src/mod_vi.f90
!> (vi_rhow_solver_putcoef :58-89, vi_rhow_solver :93-306): same module name,
!> procedure names and dummy arguments, the work done by libnicam_tt on the
!> GPU.  Compile it instead of that file, against the reference's
!> mod_precision and mod_misc, add nicam_tt_f.f90 and link libnicam_tt.so;
!> the driver source is unchanged.
!>
!> vi_rhow_solver_putcoef keeps Mc, Ml, Mu on the device (the reference keeps
!> them as module arrays).  The first vi_rhow_solver call uploads what the
!> reference reads from mod_misc: VMTR_RGAM, _RGAMH, _RGSGAM2, _RGSGAM2H,
!> _GSGAM2H and their pole forms, GRD_rdgzh, GRD_afact, GRD_bfact (after
!> GRD_setup), CONST_GRAV, CONST_Rdry, CONST_CVdry and NON_HYDRO_ALPHA.
!> Arrays are host arrays (NICAM_PTR_HOST), RP must be DP.
!-------------------------------------------------------------------------------
module mod_vi
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: vi_rhow_solver_putcoef
  public :: vi_rhow_solver

  type(c_ptr), save :: ctx = c_null_ptr
  logical,     save :: have_metrics = .false.

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

  subroutine vi_rhow_solver_putcoef( &
       Mc_in, Mc_pl_in, &
       Ml_in, Ml_pl_in, &
       Mu_in, Mu_pl_in  )
    real(RP), intent(in) :: Mc_in   (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in) :: Mc_pl_in(ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in) :: Ml_in   (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in) :: Ml_pl_in(ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in) :: Mu_in   (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in) :: Mu_pl_in(ADM_gall_pl,ADM_kall,ADM_lall_pl)

    call setup
    call nicam_check(nicam_vi_rhow_solver_putcoef(ctx, Mc_in, Mc_pl_in, Ml_in, Ml_pl_in, Mu_in, Mu_pl_in), &
                     'nicam_vi_rhow_solver_putcoef')
  end subroutine vi_rhow_solver_putcoef

  subroutine vi_rhow_solver( &
       rhogw,  rhogw_pl,  &
       rhogw0, rhogw0_pl, &
       preg0,  preg0_pl,  &
       rhog0,  rhog0_pl,  &
       Srho,   Srho_pl,   &
       Sw,     Sw_pl,     &
       Spre,   Spre_pl,   &
       dt                 )
    real(RP), intent(inout), target :: rhogw    (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(inout), target :: rhogw_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhogw0   (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhogw0_pl(ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: preg0    (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: preg0_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: rhog0    (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: rhog0_pl (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: Srho     (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: Srho_pl  (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: Sw       (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: Sw_pl    (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in),    target :: Spre     (ADM_gall   ,ADM_kall,ADM_lall   )
    real(RP), intent(in),    target :: Spre_pl  (ADM_gall_pl,ADM_kall,ADM_lall_pl)
    real(RP), intent(in)            :: dt

    call setup
    if (.not. have_metrics) then
       call nicam_check(nicam_set_vmetrics(ctx, VMTR_RGAM, VMTR_RGAM_pl, VMTR_RGAMH, VMTR_RGAMH_pl,        &
                                           VMTR_RGSGAM2, VMTR_RGSGAM2_pl, VMTR_RGSGAM2H, VMTR_RGSGAM2H_pl, &
                                           VMTR_GSGAM2H, VMTR_GSGAM2H_pl), 'nicam_set_vmetrics')
       call nicam_check(nicam_set_vgrid_factors(ctx, GRD_rdgzh, GRD_afact, GRD_bfact), 'nicam_set_vgrid_factors')
       call nicam_check(nicam_set_phys_const(ctx, real(CONST_GRAV, c_double), real(CONST_Rdry, c_double), &
                                             real(CONST_CVdry, c_double), int(NON_HYDRO_ALPHA, c_int)),   &
                        'nicam_set_phys_const')
       have_metrics = .true.
    end if
    call nicam_check(nicam_vi_rhow_solver(ctx, c_loc(rhogw), c_loc(rhogw_pl), c_loc(rhogw0), c_loc(rhogw0_pl), &
                                          c_loc(preg0), c_loc(preg0_pl), c_loc(rhog0), c_loc(rhog0_pl),         &
                                          c_loc(Srho), c_loc(Srho_pl), c_loc(Sw), c_loc(Sw_pl),                 &
                                          c_loc(Spre), c_loc(Spre_pl), real(dt, c_double)), 'nicam_vi_rhow_solver')
  end subroutine vi_rhow_solver

end module mod_vi
