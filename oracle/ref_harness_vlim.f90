!-------------------------------------------------------------------------------
! bind(C) entry around the reference's vertical_limiter_thuburn
! (kernels/NICAM/dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1).
!-------------------------------------------------------------------------------
! The interface kind follows the build: c_double, or c_float when the
! reference is compiled with -DSINGLE (RP = SP, mod_precision.f90:30-34).
#ifdef SINGLE
#define HK c_float
#else
#define HK c_double
#endif
module ref_harness_vlim
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_src_tracer, only: vertical_limiter_thuburn
  implicit none
  private
  public :: ref_vertical_limiter_thuburn
contains

  subroutine ref_vertical_limiter_thuburn( gall_1d, kall, lall, gall_pl, have_pl, &
       q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl ) bind(C, name='ref_vertical_limiter_thuburn')
    integer(c_int), value         :: gall_1d, kall, lall, gall_pl, have_pl
    real(HK), intent(inout) :: q_h(gall_1d*gall_1d,kall,lall),   q_h_pl(gall_pl,kall,2)
    real(HK), intent(in)    :: q  (gall_1d*gall_1d,kall,lall),   q_pl  (gall_pl,kall,2)
    real(HK), intent(in)    :: d  (gall_1d*gall_1d,kall,lall),   d_pl  (gall_pl,kall,2)
    real(HK), intent(in)    :: ck (gall_1d*gall_1d,kall,lall,2), ck_pl (gall_pl,kall,2,2)

    real(RP), allocatable :: qhpl(:,:,:)
    integer :: l
    logical :: pl_here

    allocate( qhpl(gall_pl,kall,2) )
    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, kall, gall_pl, pl_here, .false. )
       qhpl(:,:,:) = q_h_pl(:,:,:)
       call vertical_limiter_thuburn( q_h(:,:,l),   qhpl(:,:,:),    &
                                      q(:,:,l),     q_pl(:,:,:),    &
                                      d(:,:,l),     d_pl(:,:,:),    &
                                      ck(:,:,l,:),  ck_pl(:,:,:,:)  )
       if ( pl_here ) q_h_pl(:,:,:) = qhpl(:,:,:)
    enddo
  end subroutine ref_vertical_limiter_thuburn

end module ref_harness_vlim
