!-------------------------------------------------------------------------------
! bind(C) entry around the reference's horizontal_flux
! (kernels/NICAM/dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1); the
! metrics of region l are copied into the mod_misc globals before each call.
!-------------------------------------------------------------------------------
! The interface kind follows the build: c_double, or c_float when the
! reference is compiled with -DSINGLE (RP = SP, mod_precision.f90:30-34).
#ifdef SINGLE
#define HK c_float
#else
#define HK c_double
#endif
module ref_harness_flux
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_src_tracer, only: horizontal_flux
  implicit none
  private
  public :: ref_horizontal_flux
contains

  subroutine ref_horizontal_flux( gall_1d, kall, lall, gall_pl, have_pl, have_sgp, &
       flx_h, flx_h_pl, xc, xc_pl, rho, rho_pl, rvx, rvx_pl, rvy, rvy_pl, rvz, rvz_pl, dt, &
       xr, xr_pl, gp, gp_pl, gt, gt_pl, ga, ga_pl ) bind(C, name='ref_horizontal_flux')
    integer(c_int), value         :: gall_1d, kall, lall, gall_pl, have_pl
    integer(c_int), intent(in)    :: have_sgp(lall)
    real(HK), intent(inout) :: flx_h (gall_1d*gall_1d,kall,lall,6)
    real(HK), intent(inout) :: flx_h_pl(gall_pl,kall,2)
    real(HK), intent(inout) :: xc    (gall_1d*gall_1d,kall,lall,3,3)
    real(HK), intent(inout) :: xc_pl (gall_pl,kall,2,3)
    real(HK), intent(in)    :: rho   (gall_1d*gall_1d,kall,lall), rho_pl(gall_pl,kall,2)
    real(HK), intent(in)    :: rvx   (gall_1d*gall_1d,kall,lall), rvx_pl(gall_pl,kall,2)
    real(HK), intent(in)    :: rvy   (gall_1d*gall_1d,kall,lall), rvy_pl(gall_pl,kall,2)
    real(HK), intent(in)    :: rvz   (gall_1d*gall_1d,kall,lall), rvz_pl(gall_pl,kall,2)
    real(HK), value         :: dt
    real(HK), intent(in)    :: xr(gall_1d*gall_1d,1,lall,3,3),  xr_pl(gall_pl,1,2,3)
    real(HK), intent(in)    :: gp(gall_1d*gall_1d,1,lall,8),    gp_pl(gall_pl,1,2,8)
    real(HK), intent(in)    :: gt(gall_1d*gall_1d,1,lall,2,5),  gt_pl(gall_pl,1,2,5)
    real(HK), intent(in)    :: ga(gall_1d*gall_1d,1,lall,3,12), ga_pl(gall_pl,1,2,18)

    real(RP), allocatable :: fpl(:,:,:), xcpl(:,:,:,:)
    real(RP) :: dt_rp
    integer  :: gall, l
    logical  :: pl_here

    gall  = gall_1d*gall_1d
    dt_rp = dt
    allocate( fpl(gall_pl,kall,2), xcpl(gall_pl,kall,2,3) )

    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, kall, gall_pl, pl_here, have_sgp(l) /= 0 )

       if ( allocated(GRD_xr) ) deallocate( GRD_xr, GRD_xr_pl, GMTR_p, GMTR_p_pl, &
                                            GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl )
       allocate( GRD_xr(gall,1,1,3,3),  GRD_xr_pl(gall_pl,1,2,3) )
       allocate( GMTR_p(gall,1,1,8),    GMTR_p_pl(gall_pl,1,2,8) )
       allocate( GMTR_t(gall,1,1,2,5),  GMTR_t_pl(gall_pl,1,2,5) )
       allocate( GMTR_a(gall,1,1,3,12), GMTR_a_pl(gall_pl,1,2,18) )
       GRD_xr(:,:,1,:,:) = xr(:,:,l,:,:);  GRD_xr_pl(:,:,:,:) = xr_pl(:,:,:,:)
       GMTR_p(:,:,1,:)   = gp(:,:,l,:);    GMTR_p_pl(:,:,:,:) = gp_pl(:,:,:,:)
       GMTR_t(:,:,1,:,:) = gt(:,:,l,:,:);  GMTR_t_pl(:,:,:,:) = gt_pl(:,:,:,:)
       GMTR_a(:,:,1,:,:) = ga(:,:,l,:,:);  GMTR_a_pl(:,:,:,:) = ga_pl(:,:,:,:)

       if ( pl_here ) then
          fpl(:,:,:) = flx_h_pl(:,:,:)
          xcpl(:,:,:,:) = xc_pl(:,:,:,:)
       endif
       call horizontal_flux( flx_h(:,:,l,:), fpl(:,:,:),           &
                             xc(:,:,l,:,:),  xcpl(:,:,:,:),        &
                             rho(:,:,l),     rho_pl(:,:,:),        &
                             rvx(:,:,l),     rvx_pl(:,:,:),        &
                             rvy(:,:,l),     rvy_pl(:,:,:),        &
                             rvz(:,:,l),     rvz_pl(:,:,:),        &
                             dt_rp                                 )
       if ( pl_here ) then
          ! point 1 of the pole arrays is intent(out) and never written there
          flx_h_pl(2:,:,:) = fpl(2:,:,:)
          xc_pl(2:,:,:,:)  = xcpl(2:,:,:,:)
       endif
    enddo
  end subroutine ref_horizontal_flux

end module ref_harness_flux
