!-------------------------------------------------------------------------------
! bind(C) entry around the reference's horizontal_limiter_thuburn
! (kernels/NICAM/dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1).
!-------------------------------------------------------------------------------
! The interface kind follows the build: c_double, or c_float when the
! reference is compiled with -DSINGLE (RP = SP, mod_precision.f90:30-34).
#ifdef SINGLE
#define HK c_float
#else
#define HK c_double
#endif
module ref_harness_hlim
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_src_tracer, only: horizontal_limiter_thuburn
  implicit none
  private
  public :: ref_horizontal_limiter_thuburn
contains

  subroutine ref_horizontal_limiter_thuburn( gall_1d, kall, lall, gall_pl, have_pl, have_sgp, &
       q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl,                              &
       Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl ) bind(C, name='ref_horizontal_limiter_thuburn')
    integer(c_int), value         :: gall_1d, kall, lall, gall_pl, have_pl
    integer(c_int), intent(in)    :: have_sgp(lall)
    real(HK), intent(inout) :: q_a      (gall_1d*gall_1d,kall,lall,6), q_a_pl(gall_pl,kall,2)
    real(HK), intent(in)    :: q        (gall_1d*gall_1d,kall,lall),   q_pl  (gall_pl,kall,2)
    real(HK), intent(in)    :: d        (gall_1d*gall_1d,kall,lall),   d_pl  (gall_pl,kall,2)
    real(HK), intent(in)    :: ch       (gall_1d*gall_1d,kall,lall,6), ch_pl (gall_pl,kall,2)
    real(HK), intent(in)    :: cmask    (gall_1d*gall_1d,kall,lall,6), cmask_pl(gall_pl,kall,2)
    real(HK), intent(inout) :: Qout_prev(gall_1d*gall_1d,kall,lall,2), Qout_prev_pl(gall_pl,kall,2,2)
    real(HK), intent(in)    :: Qout_post(gall_1d*gall_1d,kall,lall,2), Qout_post_pl(gall_pl,kall,2,2)

    real(RP), allocatable :: qapl(:,:,:), qprevpl(:,:,:,:)
    integer :: l
    logical :: pl_here

    allocate( qapl(gall_pl,kall,2), qprevpl(gall_pl,kall,2,2) )

    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, kall, gall_pl, pl_here, have_sgp(l) /= 0 )
       qapl(:,:,:) = q_a_pl(:,:,:)
       call horizontal_limiter_thuburn( q_a(:,:,l,:),       qapl(:,:,:),          &
                                        q(:,:,l),           q_pl(:,:,:),          &
                                        d(:,:,l),           d_pl(:,:,:),          &
                                        ch(:,:,l,:),        ch_pl(:,:,:),         &
                                        cmask(:,:,l,:),     cmask_pl(:,:,:),      &
                                        Qout_prev(:,:,l,:), qprevpl(:,:,:,:),     &
                                        Qout_post(:,:,l,:), Qout_post_pl(:,:,:,:) )
       if ( pl_here ) then
          q_a_pl(:,:,:) = qapl(:,:,:)
          Qout_prev_pl(:,:,:,:) = qprevpl(:,:,:,:)
       endif
    enddo
  end subroutine ref_horizontal_limiter_thuburn

end module ref_harness_hlim
