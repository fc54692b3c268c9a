!-------------------------------------------------------------------------------
! bind(C) entries around the reference's vi_rhow_solver_putcoef/vi_rhow_solver
! (kernels/NICAM/dyn_vi_rhow_solver/src/mod_vi.f90:58-306) and GRD_setup
! (kernels/NICAM/dyn_vi_rhow_solver/src/mod_misc.f90:464-570).
! TEST INFRASTRUCTURE ONLY.  Regions are run one at a time (lall = 1).
!-------------------------------------------------------------------------------
module ref_harness_vi
  use iso_c_binding
  use mod_precision
  use mod_misc
  use ref_harness_common
  use mod_vi, only: vi_rhow_solver_putcoef, vi_rhow_solver, Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl
  implicit none
  private
  public :: ref_vi_rhow_solver
  public :: ref_grd_setup
contains

  subroutine ref_vi_rhow_solver( gall_1d, kall, lall, gall_pl, have_pl,                   &
       rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl,               &
       Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl, dt,                                         &
       c_Mc, c_Mc_pl, c_Ml, c_Ml_pl, c_Mu, c_Mu_pl,                                          &
       rgam, rgam_pl, rgamh, rgamh_pl, rgsgam2, rgsgam2_pl, rgsgam2h, rgsgam2h_pl,          &
       gsgam2h, gsgam2h_pl, rdgzh, afact, bfact ) bind(C, name='ref_vi_rhow_solver')
    integer(c_int), value         :: gall_1d, kall, lall, gall_pl, have_pl
    real(c_double), intent(inout) :: rhogw (gall_1d*gall_1d,kall,lall), rhogw_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: rhogw0(gall_1d*gall_1d,kall,lall), rhogw0_pl(gall_pl,kall,2)
    real(c_double), intent(in)    :: preg0 (gall_1d*gall_1d,kall,lall), preg0_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: rhog0 (gall_1d*gall_1d,kall,lall), rhog0_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: Srho  (gall_1d*gall_1d,kall,lall), Srho_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: Sw    (gall_1d*gall_1d,kall,lall), Sw_pl    (gall_pl,kall,2)
    real(c_double), intent(in)    :: Spre  (gall_1d*gall_1d,kall,lall), Spre_pl  (gall_pl,kall,2)
    real(c_double), value         :: dt
    real(c_double), intent(in)    :: c_Mc  (gall_1d*gall_1d,kall,lall), c_Mc_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: c_Ml  (gall_1d*gall_1d,kall,lall), c_Ml_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: c_Mu  (gall_1d*gall_1d,kall,lall), c_Mu_pl  (gall_pl,kall,2)
    real(c_double), intent(in)    :: rgam    (gall_1d*gall_1d,kall,lall), rgam_pl    (gall_pl,kall,2)
    real(c_double), intent(in)    :: rgamh   (gall_1d*gall_1d,kall,lall), rgamh_pl   (gall_pl,kall,2)
    real(c_double), intent(in)    :: rgsgam2 (gall_1d*gall_1d,kall,lall), rgsgam2_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: rgsgam2h(gall_1d*gall_1d,kall,lall), rgsgam2h_pl(gall_pl,kall,2)
    real(c_double), intent(in)    :: gsgam2h (gall_1d*gall_1d,kall,lall), gsgam2h_pl (gall_pl,kall,2)
    real(c_double), intent(in)    :: rdgzh(kall), afact(kall), bfact(kall)

    real(RP), allocatable :: wpl(:,:,:)
    real(RP) :: dt_rp
    integer  :: gall, l
    logical  :: pl_here

    gall  = gall_1d*gall_1d
    dt_rp = dt
    allocate( wpl(gall_pl,kall,2) )

    do l = 1, lall
       pl_here = ( have_pl /= 0 .and. l == 1 )
       call harness_set_sizes( gall_1d, kall, gall_pl, pl_here, .false. )

       if ( allocated(GRD_rdgzh) ) deallocate( GRD_rdgzh )
       if ( allocated(GRD_afact) ) deallocate( GRD_afact )
       if ( allocated(GRD_bfact) ) deallocate( GRD_bfact )
       allocate( GRD_rdgzh(kall), GRD_afact(kall), GRD_bfact(kall) )
       GRD_rdgzh(:) = rdgzh(:)
       GRD_afact(:) = afact(:)
       GRD_bfact(:) = bfact(:)

       if ( allocated(VMTR_RGAM) ) deallocate( VMTR_RGAM, VMTR_RGAM_pl, VMTR_RGAMH, VMTR_RGAMH_pl,       &
                                               VMTR_RGSGAM2, VMTR_RGSGAM2_pl, VMTR_RGSGAM2H,              &
                                               VMTR_RGSGAM2H_pl, VMTR_GSGAM2H, VMTR_GSGAM2H_pl )
       allocate( VMTR_RGAM    (gall,kall,1), VMTR_RGAM_pl    (gall_pl,kall,2) )
       allocate( VMTR_RGAMH   (gall,kall,1), VMTR_RGAMH_pl   (gall_pl,kall,2) )
       allocate( VMTR_RGSGAM2 (gall,kall,1), VMTR_RGSGAM2_pl (gall_pl,kall,2) )
       allocate( VMTR_RGSGAM2H(gall,kall,1), VMTR_RGSGAM2H_pl(gall_pl,kall,2) )
       allocate( VMTR_GSGAM2H (gall,kall,1), VMTR_GSGAM2H_pl (gall_pl,kall,2) )
       VMTR_RGAM    (:,:,1) = rgam    (:,:,l); VMTR_RGAM_pl    (:,:,:) = rgam_pl    (:,:,:)
       VMTR_RGAMH   (:,:,1) = rgamh   (:,:,l); VMTR_RGAMH_pl   (:,:,:) = rgamh_pl   (:,:,:)
       VMTR_RGSGAM2 (:,:,1) = rgsgam2 (:,:,l); VMTR_RGSGAM2_pl (:,:,:) = rgsgam2_pl (:,:,:)
       VMTR_RGSGAM2H(:,:,1) = rgsgam2h(:,:,l); VMTR_RGSGAM2H_pl(:,:,:) = rgsgam2h_pl(:,:,:)
       VMTR_GSGAM2H (:,:,1) = gsgam2h (:,:,l); VMTR_GSGAM2H_pl (:,:,:) = gsgam2h_pl (:,:,:)

       ! vi_rhow_solver_putcoef allocates its module copies on every call
       if ( allocated(Mc) ) deallocate( Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl )
       call vi_rhow_solver_putcoef( c_Mc(:,:,l), c_Mc_pl(:,:,:), &
                                    c_Ml(:,:,l), c_Ml_pl(:,:,:), &
                                    c_Mu(:,:,l), c_Mu_pl(:,:,:)  )

       wpl(:,:,:) = rhogw_pl(:,:,:)
       call vi_rhow_solver( rhogw (:,:,l), wpl      (:,:,:), &
                            rhogw0(:,:,l), rhogw0_pl(:,:,:), &
                            preg0 (:,:,l), preg0_pl (:,:,:), &
                            rhog0 (:,:,l), rhog0_pl (:,:,:), &
                            Srho  (:,:,l), Srho_pl  (:,:,:), &
                            Sw    (:,:,l), Sw_pl    (:,:,:), &
                            Spre  (:,:,l), Spre_pl  (:,:,:), &
                            dt_rp                            )
       if ( pl_here ) rhogw_pl(:,:,:) = wpl(:,:,:)
    enddo
  end subroutine ref_vi_rhow_solver

  !> GRD_setup on a vertical-grid file; returns the derived factors
  subroutine ref_grd_setup( kall, fname, nlen, gz, gzh, dgz, dgzh, rdgz, rdgzh,   &
       afact, bfact, cfact, dfact ) bind(C, name='ref_grd_setup')
    integer(c_int), value          :: kall, nlen
    character(kind=c_char), intent(in) :: fname(nlen)
    real(c_double), intent(out)    :: gz(kall), gzh(kall), dgz(kall), dgzh(kall)
    real(c_double), intent(out)    :: rdgz(kall), rdgzh(kall)
    real(c_double), intent(out)    :: afact(kall), bfact(kall), cfact(kall), dfact(kall)
    integer :: n

    call harness_set_sizes( 4, kall, 6, .false., .false. )
    vgrid_fname = ''
    do n = 1, nlen
       vgrid_fname(n:n) = fname(n)
    enddo
    if ( allocated(GRD_gz) ) deallocate( GRD_gz, GRD_gzh, GRD_dgz, GRD_dgzh, GRD_rdgz, GRD_rdgzh )
    if ( allocated(GRD_afact) ) deallocate( GRD_afact, GRD_bfact, GRD_cfact, GRD_dfact )
    call GRD_setup
    gz(:) = GRD_gz(:);     gzh(:) = GRD_gzh(:)
    dgz(:) = GRD_dgz(:);   dgzh(:) = GRD_dgzh(:)
    rdgz(:) = GRD_rdgz(:); rdgzh(:) = GRD_rdgzh(:)
    afact(:) = GRD_afact(:); bfact(:) = GRD_bfact(:)
    cfact(:) = GRD_cfact(:); dfact(:) = GRD_dfact(:)
  end subroutine ref_grd_setup

end module ref_harness_vi
