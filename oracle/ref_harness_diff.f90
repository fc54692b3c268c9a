!-------------------------------------------------------------------------------
! bind(C) entry around the reference's OPRT_diffusion
! (kernels/NICAM/dyn_diffusion/src/mod_oprt.f90:41-296).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1); the
! pole part runs with the first region.
!-------------------------------------------------------------------------------
module ref_harness_diff
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_oprt, only: OPRT_diffusion
  implicit none
  private
  public :: ref_oprt_diffusion
contains

  subroutine ref_oprt_diffusion( gall_1d, kall, lall, gall_pl, have_pl, have_sgp, &
       dscl, dscl_pl, scl, scl_pl, kh, kh_pl, coef_intp, coef_intp_pl, coef_diff, coef_diff_pl ) &
       bind(C, name='ref_oprt_diffusion')
    integer(c_int), value         :: gall_1d, kall, lall, gall_pl, have_pl
    integer(c_int), intent(in)    :: have_sgp(lall)
    real(c_double), intent(inout) :: dscl(gall_1d*gall_1d,kall,lall), dscl_pl(gall_pl,kall,2)
    real(c_double), intent(in)    :: scl (gall_1d*gall_1d,kall,lall), scl_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: kh  (gall_1d*gall_1d,kall,lall), kh_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: coef_intp(gall_1d*gall_1d,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2)
    real(c_double), intent(in)    :: coef_diff(gall_1d*gall_1d,6,3,lall),   coef_diff_pl(5,3,2)

    real(RP), allocatable :: dpl(:,:,:)
    integer :: l
    logical :: pl_here

    allocate( dpl(gall_pl,kall,2) )
    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, kall, gall_pl, pl_here, have_sgp(l) /= 0 )
       dpl(:,:,:) = dscl_pl(:,:,:)
       call OPRT_diffusion( dscl(:,:,l),            dpl(:,:,:),          &
                            scl(:,:,l),             scl_pl(:,:,:),       &
                            kh(:,:,l),              kh_pl(:,:,:),        &
                            coef_intp(:,:,:,:,l),   coef_intp_pl(:,:,:,:), &
                            coef_diff(:,:,:,l),     coef_diff_pl(:,:,:)  )
       if ( l == 1 ) dscl_pl(:,:,:) = dpl(:,:,:)
    enddo
  end subroutine ref_oprt_diffusion

end module ref_harness_diff
