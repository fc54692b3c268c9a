!-------------------------------------------------------------------------------
!> Drop-in modules for kernels/NICAM/dyn_metrics/src/mod_gmtr.f90
!> (GMTR_p_setup :46-233, GMTR_t_setup :237-400, GMTR_a_setup :404-741) and
!> dyn_metrics/src/mod_oprt.f90 (OPRT_divergence/rotation/gradient/
!> laplacian/diffusion_setup :45-1390): same module names, procedure names
!> and dummy arguments, the work done by libnicam_tt's host routines.
!> Compile this file instead of mod_vector.f90, mod_gmtr.f90 and mod_oprt.f90,
!> against the reference's mod_precision and mod_misc, add nicam_tt_f.f90 and
!> link libnicam_tt.so; the driver source is unchanged.  No GPU is used.
!>
!> Sizes and flags come from the mod_misc globals (ADM_* sizes, ADM_have_sgp,
!> ADM_have_pl); RP must be DP.
!-------------------------------------------------------------------------------
module mod_gmtr
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: GMTR_p_setup
  public :: GMTR_t_setup
  public :: GMTR_a_setup

contains

  !> nicam_gmtr_setup forms all three metric sets in one call; each setup
  !> keeps its own and discards the other two.
  subroutine run( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, GRD_s, GRD_s_pl, p, p_pl, t, t_pl, a, a_pl, radius )
    real(RP), intent(in)  :: GRD_x(*), GRD_x_pl(*), GRD_xt(*), GRD_xt_pl(*), GRD_s(*), GRD_s_pl(*)
    real(RP), intent(out) :: p(*), p_pl(*), t(*), t_pl(*), a(*), a_pl(*)
    real(RP), intent(in)  :: radius
    integer(c_int) :: sgp(ADM_lall), have_pl
    integer :: l
    if (RP /= c_double) then
       write(*,*) 'mod_gmtr (libnicam_tt shim): RP must be double precision'
       error stop 1
    endif
    do l = 1, ADM_lall
       sgp(l) = merge(1, 0, ADM_have_sgp(l))
    enddo
    have_pl = merge(1, 0, ADM_have_pl)
    call nicam_check( nicam_gmtr_setup( int(ADM_gall_1d, c_int), int(ADM_lall, c_int), int(ADM_gall_pl, c_int), &
                                        have_pl, sgp, real(radius, c_double), GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, &
                                        GRD_s, GRD_s_pl, p, p_pl, t, t_pl, a, a_pl ), 'nicam_gmtr_setup' )
  end subroutine run

  subroutine GMTR_p_setup( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, GRD_s, GRD_s_pl, GMTR_p, GMTR_p_pl, GRD_rscale )
    real(RP), intent(in)  :: GRD_x    (ADM_gall   ,k0,ADM_lall   ,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_x_pl (ADM_gall_pl,k0,ADM_lall_pl,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_xt   (ADM_gall   ,k0,ADM_lall   ,TI:TJ,ADM_nxyz)
    real(RP), intent(in)  :: GRD_xt_pl(ADM_gall_pl,k0,ADM_lall_pl,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_s    (ADM_gall   ,k0,ADM_lall   ,      2)
    real(RP), intent(in)  :: GRD_s_pl (ADM_gall_pl,k0,ADM_lall_pl,      2)
    real(RP), intent(out) :: GMTR_p   (ADM_gall   ,k0,ADM_lall   ,GMTR_p_nmax)
    real(RP), intent(out) :: GMTR_p_pl(ADM_gall_pl,k0,ADM_lall_pl,GMTR_p_nmax)
    real(RP), intent(in)  :: GRD_rscale
    real(RP), allocatable :: t(:), t_pl(:), a(:), a_pl(:)
    allocate( t(ADM_gall*ADM_lall*2*GMTR_t_nmax), t_pl(ADM_gall_pl*ADM_lall_pl*GMTR_t_nmax) )
    allocate( a(ADM_gall*ADM_lall*3*GMTR_a_nmax), a_pl(ADM_gall_pl*ADM_lall_pl*GMTR_a_nmax_pl) )
    call run( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, GRD_s, GRD_s_pl, GMTR_p, GMTR_p_pl, t, t_pl, a, a_pl, GRD_rscale )
  end subroutine GMTR_p_setup

  subroutine GMTR_t_setup( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, GMTR_t, GMTR_t_pl, GRD_rscale )
    real(RP), intent(in)  :: GRD_x    (ADM_gall   ,k0,ADM_lall   ,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_x_pl (ADM_gall_pl,k0,ADM_lall_pl,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_xt   (ADM_gall   ,k0,ADM_lall   ,TI:TJ,ADM_nxyz)
    real(RP), intent(in)  :: GRD_xt_pl(ADM_gall_pl,k0,ADM_lall_pl,      ADM_nxyz)
    real(RP), intent(out) :: GMTR_t   (ADM_gall   ,k0,ADM_lall   ,TI:TJ,GMTR_t_nmax)
    real(RP), intent(out) :: GMTR_t_pl(ADM_gall_pl,k0,ADM_lall_pl,      GMTR_t_nmax)
    real(RP), intent(in)  :: GRD_rscale
    real(RP), allocatable :: s(:), s_pl(:), p(:), p_pl(:), a(:), a_pl(:)
    allocate( s(ADM_gall*ADM_lall*2), s_pl(ADM_gall_pl*ADM_lall_pl*2) )
    allocate( p(ADM_gall*ADM_lall*GMTR_p_nmax), p_pl(ADM_gall_pl*ADM_lall_pl*GMTR_p_nmax) )
    allocate( a(ADM_gall*ADM_lall*3*GMTR_a_nmax), a_pl(ADM_gall_pl*ADM_lall_pl*GMTR_a_nmax_pl) )
    s(:) = 0.0_RP  ! only GMTR_p reads the longitudes
    s_pl(:) = 0.0_RP
    call run( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, s, s_pl, p, p_pl, GMTR_t, GMTR_t_pl, a, a_pl, GRD_rscale )
  end subroutine GMTR_t_setup

  subroutine GMTR_a_setup( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, GMTR_a, GMTR_a_pl, GRD_rscale )
    real(RP), intent(in)  :: GRD_x    (ADM_gall   ,k0,ADM_lall   ,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_x_pl (ADM_gall_pl,k0,ADM_lall_pl,      ADM_nxyz)
    real(RP), intent(in)  :: GRD_xt   (ADM_gall   ,k0,ADM_lall   ,TI:TJ,ADM_nxyz)
    real(RP), intent(in)  :: GRD_xt_pl(ADM_gall_pl,k0,ADM_lall_pl,      ADM_nxyz)
    real(RP), intent(out) :: GMTR_a   (ADM_gall   ,k0,ADM_lall   ,AI:AJ,GMTR_a_nmax   )
    real(RP), intent(out) :: GMTR_a_pl(ADM_gall_pl,k0,ADM_lall_pl,      GMTR_a_nmax_pl)
    real(RP), intent(in)  :: GRD_rscale
    real(RP), allocatable :: s(:), s_pl(:), p(:), p_pl(:), t(:), t_pl(:)
    allocate( s(ADM_gall*ADM_lall*2), s_pl(ADM_gall_pl*ADM_lall_pl*2) )
    allocate( p(ADM_gall*ADM_lall*GMTR_p_nmax), p_pl(ADM_gall_pl*ADM_lall_pl*GMTR_p_nmax) )
    allocate( t(ADM_gall*ADM_lall*2*GMTR_t_nmax), t_pl(ADM_gall_pl*ADM_lall_pl*GMTR_t_nmax) )
    s(:) = 0.0_RP
    s_pl(:) = 0.0_RP
    call run( GRD_x, GRD_x_pl, GRD_xt, GRD_xt_pl, s, s_pl, p, p_pl, t, t_pl, GMTR_a, GMTR_a_pl, GRD_rscale )
  end subroutine GMTR_a_setup

end module mod_gmtr

module mod_oprt
  use iso_c_binding
  use mod_precision
  use mod_misc
  use nicam_tt_f
  implicit none
  private

  public :: OPRT_divergence_setup
  public :: OPRT_rotation_setup
  public :: OPRT_gradient_setup
  public :: OPRT_laplacian_setup
  public :: OPRT_diffusion_setup

contains

  subroutine flags( sgp, have_pl )
    integer(c_int), intent(out) :: sgp(ADM_lall), have_pl
    integer :: l
    if (RP /= c_double) then
       write(*,*) 'mod_oprt (libnicam_tt shim): RP must be double precision'
       error stop 1
    endif
    do l = 1, ADM_lall
       sgp(l) = merge(1, 0, ADM_have_sgp(l))
    enddo
    have_pl = merge(1, 0, ADM_have_pl)
  end subroutine flags

  subroutine OPRT_divergence_setup( GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef_div, coef_div_pl )
    real(RP), intent(in)  :: GMTR_p     (ADM_gall   ,K0,ADM_lall   ,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_p_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_t     (ADM_gall   ,K0,ADM_lall   ,TI:TJ,GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_t_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_a     (ADM_gall   ,K0,ADM_lall   ,AI:AJ,GMTR_a_nmax   )
    real(RP), intent(in)  :: GMTR_a_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_a_nmax_pl)
    real(RP), intent(out) :: coef_div   (ADM_gall,0:6        ,ADM_nxyz,ADM_lall   )
    real(RP), intent(out) :: coef_div_pl(         0:ADM_vlink,ADM_nxyz,ADM_lall_pl)
    integer(c_int) :: sgp(ADM_lall), have_pl
    call flags( sgp, have_pl )
    call nicam_check( nicam_oprt_divergence_setup( int(ADM_gall_1d, c_int), int(ADM_lall, c_int), &
                      int(ADM_gall_pl, c_int), have_pl, sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, &
                      GMTR_a_pl, coef_div, coef_div_pl ), 'nicam_oprt_divergence_setup' )
  end subroutine OPRT_divergence_setup

  subroutine OPRT_rotation_setup( GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef_rot, coef_rot_pl )
    real(RP), intent(in)  :: GMTR_p     (ADM_gall   ,K0,ADM_lall   ,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_p_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_t     (ADM_gall   ,K0,ADM_lall   ,TI:TJ,GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_t_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_a     (ADM_gall   ,K0,ADM_lall   ,AI:AJ,GMTR_a_nmax   )
    real(RP), intent(in)  :: GMTR_a_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_a_nmax_pl)
    real(RP), intent(out) :: coef_rot   (ADM_gall,0:6        ,ADM_nxyz,ADM_lall   )
    real(RP), intent(out) :: coef_rot_pl(         0:ADM_vlink,ADM_nxyz,ADM_lall_pl)
    integer(c_int) :: sgp(ADM_lall), have_pl
    call flags( sgp, have_pl )
    call nicam_check( nicam_oprt_rotation_setup( int(ADM_gall_1d, c_int), int(ADM_lall, c_int), &
                      int(ADM_gall_pl, c_int), have_pl, sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, &
                      GMTR_a_pl, coef_rot, coef_rot_pl ), 'nicam_oprt_rotation_setup' )
  end subroutine OPRT_rotation_setup

  subroutine OPRT_gradient_setup( GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef_grad, coef_grad_pl )
    real(RP), intent(in)  :: GMTR_p      (ADM_gall   ,K0,ADM_lall   ,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_p_pl   (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_t      (ADM_gall   ,K0,ADM_lall   ,TI:TJ,GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_t_pl   (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_a      (ADM_gall   ,K0,ADM_lall   ,AI:AJ,GMTR_a_nmax   )
    real(RP), intent(in)  :: GMTR_a_pl   (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_a_nmax_pl)
    real(RP), intent(out) :: coef_grad   (ADM_gall,0:6        ,ADM_nxyz,ADM_lall   )
    real(RP), intent(out) :: coef_grad_pl(         0:ADM_vlink,ADM_nxyz,ADM_lall_pl)
    integer(c_int) :: sgp(ADM_lall), have_pl
    call flags( sgp, have_pl )
    call nicam_check( nicam_oprt_gradient_setup( int(ADM_gall_1d, c_int), int(ADM_lall, c_int), &
                      int(ADM_gall_pl, c_int), have_pl, sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, &
                      GMTR_a_pl, coef_grad, coef_grad_pl ), 'nicam_oprt_gradient_setup' )
  end subroutine OPRT_gradient_setup

  subroutine OPRT_laplacian_setup( GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef_lap, coef_lap_pl )
    real(RP), intent(in)  :: GMTR_p     (ADM_gall   ,K0,ADM_lall   ,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_p_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_t     (ADM_gall   ,K0,ADM_lall   ,TI:TJ,GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_t_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_a     (ADM_gall   ,K0,ADM_lall   ,AI:AJ,GMTR_a_nmax   )
    real(RP), intent(in)  :: GMTR_a_pl  (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_a_nmax_pl)
    real(RP), intent(out) :: coef_lap   (ADM_gall,0:6        ,ADM_lall   )
    real(RP), intent(out) :: coef_lap_pl(         0:ADM_vlink,ADM_lall_pl)
    integer(c_int) :: sgp(ADM_lall), have_pl
    call flags( sgp, have_pl )
    call nicam_check( nicam_oprt_laplacian_setup( int(ADM_gall_1d, c_int), int(ADM_lall, c_int), &
                      int(ADM_gall_pl, c_int), have_pl, sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, &
                      GMTR_a_pl, coef_lap, coef_lap_pl ), 'nicam_oprt_laplacian_setup' )
  end subroutine OPRT_laplacian_setup

  subroutine OPRT_diffusion_setup( GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, &
                                   coef_intp, coef_intp_pl, coef_diff, coef_diff_pl )
    real(RP), intent(in)  :: GMTR_p      (ADM_gall   ,K0,ADM_lall   ,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_p_pl   (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_p_nmax   )
    real(RP), intent(in)  :: GMTR_t      (ADM_gall   ,K0,ADM_lall   ,TI:TJ,GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_t_pl   (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_t_nmax   )
    real(RP), intent(in)  :: GMTR_a      (ADM_gall   ,K0,ADM_lall   ,AI:AJ,GMTR_a_nmax   )
    real(RP), intent(in)  :: GMTR_a_pl   (ADM_gall_pl,K0,ADM_lall_pl,      GMTR_a_nmax_pl)
    real(RP), intent(out) :: coef_intp   (ADM_gall   ,1:3,ADM_nxyz,TI:TJ,ADM_lall   )
    real(RP), intent(out) :: coef_intp_pl(ADM_gall_pl,1:3,ADM_nxyz,      ADM_lall_pl)
    real(RP), intent(out) :: coef_diff   (ADM_gall,1:6        ,ADM_nxyz,ADM_lall   )
    real(RP), intent(out) :: coef_diff_pl(         1:ADM_vlink,ADM_nxyz,ADM_lall_pl)
    integer(c_int) :: sgp(ADM_lall), have_pl
    call flags( sgp, have_pl )
    call nicam_check( nicam_oprt_diffusion_setup( int(ADM_gall_1d, c_int), int(ADM_lall, c_int), &
                      int(ADM_gall_pl, c_int), have_pl, sgp, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, &
                      GMTR_a_pl, coef_intp, coef_intp_pl, coef_diff, coef_diff_pl ), 'nicam_oprt_diffusion_setup' )
  end subroutine OPRT_diffusion_setup

end module mod_oprt
