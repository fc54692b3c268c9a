!-------------------------------------------------------------------------------
! bind(C) entry around the reference's horizontal metrics
! (kernels/NICAM/dyn_metrics/src/mod_gmtr.f90: GMTR_p_setup :46-233,
! GMTR_t_setup :237-400, GMTR_a_setup :404-741) and the operator coefficients
! (dyn_metrics/src/mod_oprt.f90).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1); the
! pole parts run with the first region, and each setup zeroes its pole output
! on entry, so the poles are kept from that first call.
!-------------------------------------------------------------------------------
module ref_harness_gmtr
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_gmtr, only: GMTR_p_setup, GMTR_t_setup, GMTR_a_setup
  use mod_oprt, only: OPRT_diffusion_setup, OPRT_divergence_setup, OPRT_rotation_setup, &
                      OPRT_gradient_setup, OPRT_laplacian_setup
  implicit none
  private
  public :: ref_gmtr_setup
  public :: ref_oprt_diffusion_setup
  public :: ref_oprt_vector_setup
  public :: ref_oprt_laplacian_setup
contains

  !> OPRT_divergence_setup (kind 1), OPRT_rotation_setup (2) or
  !> OPRT_gradient_setup (3) (dyn_metrics/src/mod_oprt.f90:45-734), one
  !> region per call, the poles kept from the first region's call.
  subroutine ref_oprt_vector_setup( kind, gall_1d, lall, gall_pl, have_pl, have_sgp, &
       GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef, coef_pl )     &
       bind(C, name='ref_oprt_vector_setup')
    integer(c_int), value       :: kind, gall_1d, lall, gall_pl, have_pl
    integer(c_int), intent(in)  :: have_sgp(lall)
    real(c_double), intent(in)  :: GMTR_p(gall_1d*gall_1d,1,lall,8),    GMTR_p_pl(gall_pl,1,2,8)
    real(c_double), intent(in)  :: GMTR_t(gall_1d*gall_1d,1,lall,2,5),  GMTR_t_pl(gall_pl,1,2,5)
    real(c_double), intent(in)  :: GMTR_a(gall_1d*gall_1d,1,lall,3,12), GMTR_a_pl(gall_pl,1,2,18)
    real(c_double), intent(out) :: coef(gall_1d*gall_1d,0:6,3,lall), coef_pl(0:gall_pl-1,3,2)

    real(RP), allocatable :: c_pl(:,:,:)
    integer :: l
    logical :: pl_here

    allocate( c_pl(0:gall_pl-1,3,2) )
    coef_pl(:,:,:) = 0.0_RP
    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, 4, gall_pl, pl_here, have_sgp(l) /= 0 )
       select case ( kind )
       case ( 1 )
          call OPRT_divergence_setup( GMTR_p(:,:,l:l,:), GMTR_p_pl, GMTR_t(:,:,l:l,:,:), GMTR_t_pl, &
                                      GMTR_a(:,:,l:l,:,:), GMTR_a_pl, coef(:,:,:,l:l), c_pl )
       case ( 2 )
          call OPRT_rotation_setup( GMTR_p(:,:,l:l,:), GMTR_p_pl, GMTR_t(:,:,l:l,:,:), GMTR_t_pl, &
                                    GMTR_a(:,:,l:l,:,:), GMTR_a_pl, coef(:,:,:,l:l), c_pl )
       case default
          call OPRT_gradient_setup( GMTR_p(:,:,l:l,:), GMTR_p_pl, GMTR_t(:,:,l:l,:,:), GMTR_t_pl, &
                                    GMTR_a(:,:,l:l,:,:), GMTR_a_pl, coef(:,:,:,l:l), c_pl )
       end select
       if ( pl_here ) coef_pl(:,:,:) = c_pl(:,:,:)
    enddo
  end subroutine ref_oprt_vector_setup

  !> OPRT_laplacian_setup (dyn_metrics/src/mod_oprt.f90:737-1232), as above.
  subroutine ref_oprt_laplacian_setup( gall_1d, lall, gall_pl, have_pl, have_sgp, &
       GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef, coef_pl ) &
       bind(C, name='ref_oprt_laplacian_setup')
    integer(c_int), value       :: gall_1d, lall, gall_pl, have_pl
    integer(c_int), intent(in)  :: have_sgp(lall)
    real(c_double), intent(in)  :: GMTR_p(gall_1d*gall_1d,1,lall,8),    GMTR_p_pl(gall_pl,1,2,8)
    real(c_double), intent(in)  :: GMTR_t(gall_1d*gall_1d,1,lall,2,5),  GMTR_t_pl(gall_pl,1,2,5)
    real(c_double), intent(in)  :: GMTR_a(gall_1d*gall_1d,1,lall,3,12), GMTR_a_pl(gall_pl,1,2,18)
    real(c_double), intent(out) :: coef(gall_1d*gall_1d,0:6,lall), coef_pl(0:gall_pl-1,2)

    real(RP), allocatable :: c_pl(:,:)
    integer :: l
    logical :: pl_here

    allocate( c_pl(0:gall_pl-1,2) )
    coef_pl(:,:) = 0.0_RP
    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, 4, gall_pl, pl_here, have_sgp(l) /= 0 )
       call OPRT_laplacian_setup( GMTR_p(:,:,l:l,:), GMTR_p_pl, GMTR_t(:,:,l:l,:,:), GMTR_t_pl, &
                                  GMTR_a(:,:,l:l,:,:), GMTR_a_pl, coef(:,:,l:l), c_pl )
       if ( pl_here ) coef_pl(:,:) = c_pl(:,:)
    enddo
  end subroutine ref_oprt_laplacian_setup

  !> OPRT_diffusion_setup (dyn_metrics/src/mod_oprt.f90:1235-1390), one region
  !> per call; it zeroes its pole output on entry, so the poles are kept from
  !> the first region's call.
  subroutine ref_oprt_diffusion_setup( gall_1d, lall, gall_pl, have_pl, have_sgp, &
       GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl,                  &
       coef_intp, coef_intp_pl, coef_diff, coef_diff_pl )                        &
       bind(C, name='ref_oprt_diffusion_setup')
    integer(c_int), value       :: gall_1d, lall, gall_pl, have_pl
    integer(c_int), intent(in)  :: have_sgp(lall)
    real(c_double), intent(in)  :: GMTR_p(gall_1d*gall_1d,1,lall,8),    GMTR_p_pl(gall_pl,1,2,8)
    real(c_double), intent(in)  :: GMTR_t(gall_1d*gall_1d,1,lall,2,5),  GMTR_t_pl(gall_pl,1,2,5)
    real(c_double), intent(in)  :: GMTR_a(gall_1d*gall_1d,1,lall,3,12), GMTR_a_pl(gall_pl,1,2,18)
    real(c_double), intent(out) :: coef_intp(gall_1d*gall_1d,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2)
    real(c_double), intent(out) :: coef_diff(gall_1d*gall_1d,6,3,lall),   coef_diff_pl(gall_pl-1,3,2)

    real(RP), allocatable :: ci_pl(:,:,:,:), cd_pl(:,:,:)
    integer :: l
    logical :: pl_here

    allocate( ci_pl(gall_pl,3,3,2), cd_pl(gall_pl-1,3,2) )
    coef_intp_pl(:,:,:,:) = 0.0_RP
    coef_diff_pl(:,:,:)   = 0.0_RP
    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, 4, gall_pl, pl_here, have_sgp(l) /= 0 )
       call OPRT_diffusion_setup( GMTR_p(:,:,l:l,:),   GMTR_p_pl(:,:,:,:), &
                                  GMTR_t(:,:,l:l,:,:), GMTR_t_pl(:,:,:,:), &
                                  GMTR_a(:,:,l:l,:,:), GMTR_a_pl(:,:,:,:), &
                                  coef_intp(:,:,:,:,l:l), ci_pl(:,:,:,:),  &
                                  coef_diff(:,:,:,l:l),   cd_pl(:,:,:)     )
       if ( pl_here ) then
          coef_intp_pl(:,:,:,:) = ci_pl(:,:,:,:)
          coef_diff_pl(:,:,:)   = cd_pl(:,:,:)
       endif
    enddo
  end subroutine ref_oprt_diffusion_setup

  subroutine ref_gmtr_setup( gall_1d, lall, gall_pl, have_pl, have_sgp, radius, &
       GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, GRD_s, GRD_s_pl,                    &
       GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl )               &
       bind(C, name='ref_gmtr_setup')
    integer(c_int), value         :: gall_1d, lall, gall_pl, have_pl
    integer(c_int), intent(in)    :: have_sgp(lall)
    real(c_double), value         :: radius
    real(c_double), intent(in)    :: GRD_x (gall_1d*gall_1d,1,lall,3),   GRD_x_pl (gall_pl,1,2,3)
    real(c_double), intent(in)    :: GRD_xt(gall_1d*gall_1d,1,lall,2,3), GRD_xt_pl(gall_pl,1,2,3)
    real(c_double), intent(in)    :: GRD_s (gall_1d*gall_1d,1,lall,2),   GRD_s_pl (gall_pl,1,2,2)
    real(c_double), intent(out)   :: GMTR_p(gall_1d*gall_1d,1,lall,8),   GMTR_p_pl(gall_pl,1,2,8)
    real(c_double), intent(out)   :: GMTR_t(gall_1d*gall_1d,1,lall,2,5), GMTR_t_pl(gall_pl,1,2,5)
    real(c_double), intent(out)   :: GMTR_a(gall_1d*gall_1d,1,lall,3,12), GMTR_a_pl(gall_pl,1,2,18)

    real(RP), allocatable :: p_pl(:,:,:,:), t_pl(:,:,:,:), a_pl(:,:,:,:)
    real(RP), allocatable :: p(:,:,:,:), t(:,:,:,:,:), a(:,:,:,:,:)
    integer :: l
    logical :: pl_here

    allocate( p_pl(gall_pl,1,2,8), t_pl(gall_pl,1,2,5), a_pl(gall_pl,1,2,18) )
    allocate( p(gall_1d*gall_1d,1,1,8), t(gall_1d*gall_1d,1,1,2,5), a(gall_1d*gall_1d,1,1,3,12) )
    GMTR_p_pl(:,:,:,:) = 0.0_RP
    GMTR_t_pl(:,:,:,:) = 0.0_RP
    GMTR_a_pl(:,:,:,:) = 0.0_RP
    GRD_rscale = radius
    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, 4, gall_pl, pl_here, have_sgp(l) /= 0 )
       call GMTR_p_setup( GRD_x (:,:,l:l,:),   GRD_x_pl (:,:,:,:), &
                          GRD_xt(:,:,l:l,:,:), GRD_xt_pl(:,:,:,:), &
                          GRD_s (:,:,l:l,:),   GRD_s_pl (:,:,:,:), &
                          p(:,:,:,:),          p_pl(:,:,:,:),      &
                          GRD_rscale                               )
       call GMTR_t_setup( GRD_x (:,:,l:l,:),   GRD_x_pl (:,:,:,:), &
                          GRD_xt(:,:,l:l,:,:), GRD_xt_pl(:,:,:,:), &
                          t(:,:,:,:,:),        t_pl(:,:,:,:),      &
                          GRD_rscale                               )
       call GMTR_a_setup( GRD_x (:,:,l:l,:),   GRD_x_pl (:,:,:,:), &
                          GRD_xt(:,:,l:l,:,:), GRD_xt_pl(:,:,:,:), &
                          a(:,:,:,:,:),        a_pl(:,:,:,:),      &
                          GRD_rscale                               )
       GMTR_p(:,:,l,:)   = p(:,:,1,:)
       GMTR_t(:,:,l,:,:) = t(:,:,1,:,:)
       GMTR_a(:,:,l,:,:) = a(:,:,1,:,:)
       if ( pl_here ) then
          GMTR_p_pl(:,:,:,:) = p_pl(:,:,:,:)
          GMTR_t_pl(:,:,:,:) = t_pl(:,:,:,:)
          GMTR_a_pl(:,:,:,:) = a_pl(:,:,:,:)
       endif
    enddo
  end subroutine ref_gmtr_setup

end module ref_harness_gmtr
