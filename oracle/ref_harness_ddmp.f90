!-------------------------------------------------------------------------------
! bind(C) entry around the reference's OPRT3D_divdamp
! (kernels/NICAM/dyn_divdamp/src/mod_oprt3d.f90:41-402).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1); the
! pole part runs with the first region.  The module arrays the kernel reads
! (VMTR_RGSQRTH, VMTR_RGAM, VMTR_RGAMH, VMTR_C2WfactGz and their pole forms,
! GRD_rdgz) are set from the arguments before each call.
!-------------------------------------------------------------------------------
module ref_harness_ddmp
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_oprt3d, only: OPRT3D_divdamp
  implicit none
  private
  public :: ref_oprt3d_divdamp
contains

  subroutine ref_oprt3d_divdamp( gall_1d, kall, lall, gall_pl, have_pl, have_sgp,          &
       ddx, ddx_pl, ddy, ddy_pl, ddz, ddz_pl, vx, vx_pl, vy, vy_pl, vz, vz_pl, w, w_pl,     &
       coef_intp, coef_intp_pl, coef_diff, coef_diff_pl, rgsqrth, rgsqrth_pl, rgam, rgam_pl, &
       rgamh, rgamh_pl, c2w, c2w_pl, rdgz ) bind(C, name='ref_oprt3d_divdamp')
    integer(c_int), value         :: gall_1d, kall, lall, gall_pl, have_pl
    integer(c_int), intent(in)    :: have_sgp(lall)
    real(c_double), intent(inout) :: ddx(gall_1d*gall_1d,kall,lall), ddx_pl(gall_pl,kall,2)
    real(c_double), intent(inout) :: ddy(gall_1d*gall_1d,kall,lall), ddy_pl(gall_pl,kall,2)
    real(c_double), intent(inout) :: ddz(gall_1d*gall_1d,kall,lall), ddz_pl(gall_pl,kall,2)
    real(c_double), intent(in)    :: vx (gall_1d*gall_1d,kall,lall), vx_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: vy (gall_1d*gall_1d,kall,lall), vy_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: vz (gall_1d*gall_1d,kall,lall), vz_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: w  (gall_1d*gall_1d,kall,lall), w_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: coef_intp(gall_1d*gall_1d,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2)
    real(c_double), intent(in)    :: coef_diff(gall_1d*gall_1d,6,3,lall),   coef_diff_pl(5,3,2)
    real(c_double), intent(in)    :: rgsqrth(gall_1d*gall_1d,kall,lall), rgsqrth_pl(gall_pl,kall,2)
    real(c_double), intent(in)    :: rgam   (gall_1d*gall_1d,kall,lall), rgam_pl   (gall_pl,kall,2)
    real(c_double), intent(in)    :: rgamh  (gall_1d*gall_1d,kall,lall), rgamh_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: c2w(gall_1d*gall_1d,kall,6,lall),   c2w_pl(gall_pl,kall,6,2)
    real(c_double), intent(in)    :: rdgz(kall)

    real(RP), allocatable :: xpl(:,:,:), ypl(:,:,:), zpl(:,:,:)
    integer :: l, gall
    logical :: pl_here

    gall = gall_1d * gall_1d
    allocate( xpl(gall_pl,kall,2), ypl(gall_pl,kall,2), zpl(gall_pl,kall,2) )
    if ( allocated(GRD_rdgz) ) deallocate( GRD_rdgz )
    allocate( GRD_rdgz(kall) )
    GRD_rdgz(:) = rdgz(:)
    call realloc3( VMTR_RGSQRTH, gall, kall, 1 )
    call realloc3( VMTR_RGAM,    gall, kall, 1 )
    call realloc3( VMTR_RGAMH,   gall, kall, 1 )
    call realloc3( VMTR_RGSQRTH_pl, gall_pl, kall, 2 )
    call realloc3( VMTR_RGAM_pl,    gall_pl, kall, 2 )
    call realloc3( VMTR_RGAMH_pl,   gall_pl, kall, 2 )
    if ( allocated(VMTR_C2WfactGz) ) deallocate( VMTR_C2WfactGz )
    if ( allocated(VMTR_C2WfactGz_pl) ) deallocate( VMTR_C2WfactGz_pl )
    allocate( VMTR_C2WfactGz(gall,kall,6,1), VMTR_C2WfactGz_pl(gall_pl,kall,6,2) )
    VMTR_RGSQRTH_pl(:,:,:) = rgsqrth_pl(:,:,:)
    VMTR_RGAM_pl   (:,:,:) = rgam_pl   (:,:,:)
    VMTR_RGAMH_pl  (:,:,:) = rgamh_pl  (:,:,:)
    VMTR_C2WfactGz_pl(:,:,:,:) = c2w_pl(:,:,:,:)

    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, kall, gall_pl, pl_here, have_sgp(l) /= 0 )
       VMTR_RGSQRTH(:,:,1) = rgsqrth(:,:,l)
       VMTR_RGAM   (:,:,1) = rgam   (:,:,l)
       VMTR_RGAMH  (:,:,1) = rgamh  (:,:,l)
       VMTR_C2WfactGz(:,:,:,1) = c2w(:,:,:,l)
       xpl(:,:,:) = ddx_pl(:,:,:)
       ypl(:,:,:) = ddy_pl(:,:,:)
       zpl(:,:,:) = ddz_pl(:,:,:)
       call OPRT3D_divdamp( ddx(:,:,l), xpl(:,:,:), ddy(:,:,l), ypl(:,:,:), ddz(:,:,l), zpl(:,:,:), &
                            vx(:,:,l), vx_pl(:,:,:), vy(:,:,l), vy_pl(:,:,:), vz(:,:,l), vz_pl(:,:,:), &
                            w(:,:,l), w_pl(:,:,:), coef_intp(:,:,:,:,l), coef_intp_pl(:,:,:,:),        &
                            coef_diff(:,:,:,l), coef_diff_pl(:,:,:) )
       if ( l == 1 ) then
          ddx_pl(:,:,:) = xpl(:,:,:)
          ddy_pl(:,:,:) = ypl(:,:,:)
          ddz_pl(:,:,:) = zpl(:,:,:)
       endif
    enddo
  end subroutine ref_oprt3d_divdamp

  subroutine realloc3( a, n1, n2, n3 )
    real(RP), allocatable, intent(inout) :: a(:,:,:)
    integer, intent(in) :: n1, n2, n3
    if ( allocated(a) ) deallocate( a )
    allocate( a(n1,n2,n3) )
  end subroutine realloc3

end module ref_harness_ddmp
