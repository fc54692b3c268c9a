!-------------------------------------------------------------------------------
!> Fortran 2003 binding of libnicam_tt (include/nicam_tt.h), one interface per
!> exported C function.  A context handle is a type(c_ptr).  Field arguments
!> are passed as type(c_ptr) by value: c_loc() of a host array in
!> NICAM_PTR_HOST mode, or a device address in NICAM_PTR_DEVICE mode; the
!> set-up calls take host arrays directly.  nicam_check() turns a nonzero
!> status into an error stop that prints the library's message.
!-------------------------------------------------------------------------------
module nicam_tt_f
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: NICAM_OK        = 0
  integer(c_int), parameter, public :: NICAM_ERR_ARG   = 1
  integer(c_int), parameter, public :: NICAM_ERR_STATE = 2
  integer(c_int), parameter, public :: NICAM_ERR_HIP   = 3
  integer(c_int), parameter, public :: NICAM_ERR_IO    = 4
  integer(c_int), parameter, public :: NICAM_ERR_NOMEM = 5
  integer(c_int), parameter, public :: NICAM_PTR_DEVICE = 0
  integer(c_int), parameter, public :: NICAM_PTR_HOST   = 1
  integer(c_int), parameter, public :: NICAM_TUNE_HFLUX_KPER  = 1
  integer(c_int), parameter, public :: NICAM_TUNE_VLIM_KCHUNK = 3
  integer(c_int), parameter, public :: NICAM_TUNE_VI_VARIANT  = 4
  integer(c_int), parameter, public :: NICAM_TUNE_DIFF_KPER   = 5
  integer(c_int), parameter, public :: NICAM_TUNE_DDMP_KPER   = 6

  public :: nicam_check

  interface
     function nicam_version() bind(C, name='nicam_version')
       import :: c_ptr
       type(c_ptr) :: nicam_version
     end function nicam_version

     function nicam_last_error() bind(C, name='nicam_last_error')
       import :: c_ptr
       type(c_ptr) :: nicam_last_error
     end function nicam_last_error

     function nicam_ctx_create(ctx, gall_1d, kall, lall, gall_pl, have_pl, have_sgp) &
          bind(C, name='nicam_ctx_create')
       import :: c_ptr, c_int
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: gall_1d, kall, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       integer(c_int) :: nicam_ctx_create
     end function nicam_ctx_create

     function nicam_ctx_destroy(ctx) bind(C, name='nicam_ctx_destroy')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: nicam_ctx_destroy
     end function nicam_ctx_destroy

     function nicam_ctx_set_stream(ctx, stream) bind(C, name='nicam_ctx_set_stream')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, stream
       integer(c_int) :: nicam_ctx_set_stream
     end function nicam_ctx_set_stream

     function nicam_ctx_set_pointer_mode(ctx, mode) bind(C, name='nicam_ctx_set_pointer_mode')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: mode
       integer(c_int) :: nicam_ctx_set_pointer_mode
     end function nicam_ctx_set_pointer_mode

     function nicam_ctx_set_tuning(ctx, param, val) bind(C, name='nicam_ctx_set_tuning')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: param, val
       integer(c_int) :: nicam_ctx_set_tuning
     end function nicam_ctx_set_tuning

     function nicam_ctx_device_bytes(ctx) bind(C, name='nicam_ctx_device_bytes')
       import :: c_ptr, c_size_t
       type(c_ptr), value :: ctx
       integer(c_size_t) :: nicam_ctx_device_bytes
     end function nicam_ctx_device_bytes

     function nicam_const_setup(ctx) bind(C, name='nicam_const_setup')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: nicam_const_setup
     end function nicam_const_setup

     function nicam_set_const(ctx, eps, huge_val) bind(C, name='nicam_set_const')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: eps, huge_val
       integer(c_int) :: nicam_set_const
     end function nicam_set_const

     function nicam_set_phys_const(ctx, grav, rdry, cvdry, non_hydro_alpha) bind(C, name='nicam_set_phys_const')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: grav, rdry, cvdry
       integer(c_int), value :: non_hydro_alpha
       integer(c_int) :: nicam_set_phys_const
     end function nicam_set_phys_const

     function nicam_grd_input_vgrid(fname, kall, gz, gzh) bind(C, name='nicam_grd_input_vgrid')
       import :: c_char, c_int, c_double
       character(kind=c_char), intent(in) :: fname(*)
       integer(c_int), value :: kall
       real(c_double), intent(out) :: gz(*), gzh(*)
       integer(c_int) :: nicam_grd_input_vgrid
     end function nicam_grd_input_vgrid

     function nicam_grd_setup(ctx, gz, gzh, rdgzh, afact, bfact) bind(C, name='nicam_grd_setup')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, gz, gzh, rdgzh, afact, bfact
       integer(c_int) :: nicam_grd_setup
     end function nicam_grd_setup

     function nicam_set_vgrid_factors(ctx, rdgzh, afact, bfact) bind(C, name='nicam_set_vgrid_factors')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rdgzh(*), afact(*), bfact(*)
       integer(c_int) :: nicam_set_vgrid_factors
     end function nicam_set_vgrid_factors

     function nicam_set_hmetrics(ctx, GRD_xr, GRD_xr_pl, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, &
          GMTR_a, GMTR_a_pl) bind(C, name='nicam_set_hmetrics')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: GRD_xr(*), GRD_xr_pl(*), GMTR_p(*), GMTR_p_pl(*)
       real(c_double), intent(in) :: GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       integer(c_int) :: nicam_set_hmetrics
     end function nicam_set_hmetrics

     function nicam_set_vmetrics(ctx, RGAM, RGAM_pl, RGAMH, RGAMH_pl, RGSGAM2, RGSGAM2_pl, &
          RGSGAM2H, RGSGAM2H_pl, GSGAM2H, GSGAM2H_pl) bind(C, name='nicam_set_vmetrics')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: RGAM(*), RGAM_pl(*), RGAMH(*), RGAMH_pl(*), RGSGAM2(*), RGSGAM2_pl(*)
       real(c_double), intent(in) :: RGSGAM2H(*), RGSGAM2H_pl(*), GSGAM2H(*), GSGAM2H_pl(*)
       integer(c_int) :: nicam_set_vmetrics
     end function nicam_set_vmetrics

     function nicam_horizontal_flux(ctx, flx_h, flx_h_pl, GRD_xc, GRD_xc_pl, rho, rho_pl, rhovx, rhovx_pl, &
          rhovy, rhovy_pl, rhovz, rhovz_pl, dt) bind(C, name='nicam_horizontal_flux')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, flx_h, flx_h_pl, GRD_xc, GRD_xc_pl, rho, rho_pl
       type(c_ptr), value :: rhovx, rhovx_pl, rhovy, rhovy_pl, rhovz, rhovz_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_horizontal_flux
     end function nicam_horizontal_flux

     function nicam_horizontal_limiter_thuburn(ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, &
          Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl) bind(C, name='nicam_horizontal_limiter_thuburn')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl
       type(c_ptr), value :: Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl
       integer(c_int) :: nicam_horizontal_limiter_thuburn
     end function nicam_horizontal_limiter_thuburn

     function nicam_vertical_limiter_thuburn(ctx, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl) &
          bind(C, name='nicam_vertical_limiter_thuburn')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl
       integer(c_int) :: nicam_vertical_limiter_thuburn
     end function nicam_vertical_limiter_thuburn

     function nicam_vi_rhow_solver_putcoef(ctx, Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl) &
          bind(C, name='nicam_vi_rhow_solver_putcoef')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: Mc(*), Mc_pl(*), Ml(*), Ml_pl(*), Mu(*), Mu_pl(*)
       integer(c_int) :: nicam_vi_rhow_solver_putcoef
     end function nicam_vi_rhow_solver_putcoef

     function nicam_vi_rhow_solver(ctx, rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl, &
          Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl, dt) bind(C, name='nicam_vi_rhow_solver')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl
       type(c_ptr), value :: Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_vi_rhow_solver
     end function nicam_vi_rhow_solver

     function nicam_vi_rhow_solver_mixed(ctx, rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, &
          rhog0_pl, Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl, dt) bind(C, name='nicam_vi_rhow_solver_mixed')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, rhogw, rhogw_pl, rhogw0, rhogw0_pl, preg0, preg0_pl, rhog0, rhog0_pl
       type(c_ptr), value :: Srho, Srho_pl, Sw, Sw_pl, Spre, Spre_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_vi_rhow_solver_mixed
     end function nicam_vi_rhow_solver_mixed

     function nicam_vi_rhow_solver_adjoint(ctx, rhogw_b, rhogw_b_pl, rhogw0_b, rhogw0_b_pl, preg0_b, preg0_b_pl, &
          rhog0_b, rhog0_b_pl, Srho_b, Srho_b_pl, Sw_b, Sw_b_pl, Spre_b, Spre_b_pl, dt) &
          bind(C, name='nicam_vi_rhow_solver_adjoint')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, rhogw_b, rhogw_b_pl, rhogw0_b, rhogw0_b_pl, preg0_b, preg0_b_pl
       type(c_ptr), value :: rhog0_b, rhog0_b_pl, Srho_b, Srho_b_pl, Sw_b, Sw_b_pl, Spre_b, Spre_b_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_vi_rhow_solver_adjoint
     end function nicam_vi_rhow_solver_adjoint

     function nicam_horizontal_flux_adjoint(ctx, flx_h_b, flx_h_b_pl, GRD_xc_b, GRD_xc_b_pl, rho_b, rho_b_pl, &
          rhovx_b, rhovx_b_pl, rhovy_b, rhovy_b_pl, rhovz_b, rhovz_b_pl, rho, rho_pl, rhovx, rhovx_pl, rhovy, &
          rhovy_pl, rhovz, rhovz_pl, dt) bind(C, name='nicam_horizontal_flux_adjoint')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, flx_h_b, flx_h_b_pl, GRD_xc_b, GRD_xc_b_pl, rho_b, rho_b_pl, rhovx_b, rhovx_b_pl
       type(c_ptr), value :: rhovy_b, rhovy_b_pl, rhovz_b, rhovz_b_pl, rho, rho_pl, rhovx, rhovx_pl, rhovy
       type(c_ptr), value :: rhovy_pl, rhovz, rhovz_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_horizontal_flux_adjoint
     end function nicam_horizontal_flux_adjoint

     function nicam_horizontal_limiter_thuburn_mask(ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, &
          cmask_pl, Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl, mask, mask_pl) &
          bind(C, name='nicam_horizontal_limiter_thuburn_mask')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl
       type(c_ptr), value :: Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl, mask, mask_pl
       integer(c_int) :: nicam_horizontal_limiter_thuburn_mask
     end function nicam_horizontal_limiter_thuburn_mask

     function nicam_horizontal_limiter_thuburn_adjoint(ctx, q_a_b, q_a_b_pl, q_b, q_b_pl, Qout_prev_b, &
          Qout_prev_b_pl, Qout_post_b, Qout_post_b_pl, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, &
          cmask_pl, Qout_post, Qout_post_pl, mask, mask_pl) bind(C, name='nicam_horizontal_limiter_thuburn_adjoint')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_a_b, q_a_b_pl, q_b, q_b_pl, Qout_prev_b, Qout_prev_b_pl, Qout_post_b
       type(c_ptr), value :: Qout_post_b_pl, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl
       type(c_ptr), value :: Qout_post, Qout_post_pl, mask, mask_pl
       integer(c_int) :: nicam_horizontal_limiter_thuburn_adjoint
     end function nicam_horizontal_limiter_thuburn_adjoint

     function nicam_vertical_limiter_thuburn_mask(ctx, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl) &
          bind(C, name='nicam_vertical_limiter_thuburn_mask')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl
       integer(c_int) :: nicam_vertical_limiter_thuburn_mask
     end function nicam_vertical_limiter_thuburn_mask

     function nicam_vertical_limiter_thuburn_adjoint(ctx, q_h_b, q_h_b_pl, q_b, q_b_pl, q_h, q_h_pl, q, q_pl, &
          d, d_pl, ck, ck_pl, mask, mask_pl) bind(C, name='nicam_vertical_limiter_thuburn_adjoint')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_h_b, q_h_b_pl, q_b, q_b_pl, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl
       type(c_ptr), value :: mask, mask_pl
       integer(c_int) :: nicam_vertical_limiter_thuburn_adjoint
     end function nicam_vertical_limiter_thuburn_adjoint

     function nicam_ctx_synchronize(ctx) bind(C, name='nicam_ctx_synchronize')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: nicam_ctx_synchronize
     end function nicam_ctx_synchronize

     function nicam_halo_map(rlevel, gall_1d, gall_pl, capacity, n_out, dst_kind, dst_l, dst_g, &
          src_kind, src_l, src_g, have_sgp) bind(C, name='nicam_halo_map')
       import :: c_ptr, c_int
       integer(c_int), value :: rlevel, gall_1d, gall_pl, capacity
       integer(c_int), intent(out) :: n_out
       type(c_ptr), value :: dst_kind, dst_l, dst_g, src_kind, src_l, src_g, have_sgp
       integer(c_int) :: nicam_halo_map
     end function nicam_halo_map

     function nicam_pole_fill_map(rlevel, gall_1d, src_l, src_g) bind(C, name='nicam_pole_fill_map')
       import :: c_int
       integer(c_int), value :: rlevel, gall_1d
       integer(c_int), intent(out) :: src_l(2), src_g(2)
       integer(c_int) :: nicam_pole_fill_map
     end function nicam_pole_fill_map

     function nicam_comm_setup(ctx, rlevel) bind(C, name='nicam_comm_setup')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: rlevel
       integer(c_int) :: nicam_comm_setup
     end function nicam_comm_setup

     function nicam_comm_data_transfer(ctx, var, var_pl, kmax, vmax) bind(C, name='nicam_comm_data_transfer')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, var, var_pl
       integer(c_int), value :: kmax, vmax
       integer(c_int) :: nicam_comm_data_transfer
     end function nicam_comm_data_transfer

     function nicam_comm_var(ctx, var, var_pl, kmax, vmax) bind(C, name='nicam_comm_var')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, var, var_pl
       integer(c_int), value :: kmax, vmax
       integer(c_int) :: nicam_comm_var
     end function nicam_comm_var

     function nicam_oprt_diffusion(ctx, dscl, dscl_pl, scl, scl_pl, kh, kh_pl, coef_intp, coef_intp_pl, &
          coef_diff, coef_diff_pl) bind(C, name='nicam_oprt_diffusion')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, dscl, dscl_pl, scl, scl_pl, kh, kh_pl
       type(c_ptr), value :: coef_intp, coef_intp_pl, coef_diff, coef_diff_pl
       integer(c_int) :: nicam_oprt_diffusion
     end function nicam_oprt_diffusion

     function nicam_oprt_gradient(ctx, grad, grad_pl, scl, scl_pl, coef_grad, coef_grad_pl) &
          bind(C, name='nicam_oprt_gradient')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, grad, grad_pl, scl, scl_pl, coef_grad, coef_grad_pl
       integer(c_int) :: nicam_oprt_gradient
     end function nicam_oprt_gradient

     function nicam_tracer_courant(ctx, ch, ch_pl, cmask, cmask_pl, flx_h, flx_h_pl, rhog, rhog_pl) &
          bind(C, name='nicam_tracer_courant')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, ch, ch_pl, cmask, cmask_pl, flx_h, flx_h_pl, rhog, rhog_pl
       integer(c_int) :: nicam_tracer_courant
     end function nicam_tracer_courant

     function nicam_horizontal_remap(ctx, q_a, q_a_pl, q, q_pl, grad, grad_pl, cmask, cmask_pl, &
          GRD_xc, GRD_xc_pl, GRD_x, GRD_x_pl) bind(C, name='nicam_horizontal_remap')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_a, q_a_pl, q, q_pl, grad, grad_pl, cmask, cmask_pl
       type(c_ptr), value :: GRD_xc, GRD_xc_pl, GRD_x, GRD_x_pl
       integer(c_int) :: nicam_horizontal_remap
     end function nicam_horizontal_remap

     function nicam_tracer_update(ctx, rhogq, rhogq_pl, flx_h, flx_h_pl, q_a, q_a_pl) &
          bind(C, name='nicam_tracer_update')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, rhogq, rhogq_pl, flx_h, flx_h_pl, q_a, q_a_pl
       integer(c_int) :: nicam_tracer_update
     end function nicam_tracer_update

     function nicam_vertical_mass_flux(ctx, flx_v, flx_v_pl, rhogvx, rhogvx_pl, rhogvy, rhogvy_pl, &
          rhogvz, rhogvz_pl, rhogw, rhogw_pl, dt) bind(C, name='nicam_vertical_mass_flux')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, flx_v, flx_v_pl, rhogvx, rhogvx_pl, rhogvy, rhogvy_pl
       type(c_ptr), value :: rhogvz, rhogvz_pl, rhogw, rhogw_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_vertical_mass_flux
     end function nicam_vertical_mass_flux

     function nicam_vertical_courant(ctx, ck, ck_pl, flx_v, flx_v_pl, rhog, rhog_pl) &
          bind(C, name='nicam_vertical_courant')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, ck, ck_pl, flx_v, flx_v_pl, rhog, rhog_pl
       integer(c_int) :: nicam_vertical_courant
     end function nicam_vertical_courant

     function nicam_vertical_face(ctx, q_h, q_h_pl, q, q_pl) bind(C, name='nicam_vertical_face')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_h, q_h_pl, q, q_pl
       integer(c_int) :: nicam_vertical_face
     end function nicam_vertical_face

     function nicam_tracer_update_vertical(ctx, rhogq, rhogq_pl, flx_v, flx_v_pl, q_h, q_h_pl) &
          bind(C, name='nicam_tracer_update_vertical')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, rhogq, rhogq_pl, flx_v, flx_v_pl, q_h, q_h_pl
       integer(c_int) :: nicam_tracer_update_vertical
     end function nicam_tracer_update_vertical

     function nicam_vi_rhow_update_matrix(ctx, eth, eth_pl, g_tilde, g_tilde_pl, dt) &
          bind(C, name='nicam_vi_rhow_update_matrix')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, eth, eth_pl, g_tilde, g_tilde_pl
       real(c_double), value :: dt
       integer(c_int) :: nicam_vi_rhow_update_matrix
     end function nicam_vi_rhow_update_matrix

     function nicam_vi_rhow_getcoef(ctx, Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl) bind(C, name='nicam_vi_rhow_getcoef')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl
       integer(c_int) :: nicam_vi_rhow_getcoef
     end function nicam_vi_rhow_getcoef

     function nicam_horizontal_flux_f32(ctx, flx_h, flx_h_pl, GRD_xc, GRD_xc_pl, rho, rho_pl, rhovx, rhovx_pl, &
          rhovy, rhovy_pl, rhovz, rhovz_pl, dt) bind(C, name='nicam_horizontal_flux_f32')
       import :: c_ptr, c_int, c_float
       type(c_ptr), value :: ctx, flx_h, flx_h_pl, GRD_xc, GRD_xc_pl, rho, rho_pl
       type(c_ptr), value :: rhovx, rhovx_pl, rhovy, rhovy_pl, rhovz, rhovz_pl
       real(c_float), value :: dt
       integer(c_int) :: nicam_horizontal_flux_f32
     end function nicam_horizontal_flux_f32

     function nicam_horizontal_limiter_thuburn_f32(ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, &
          cmask_pl, Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl) &
          bind(C, name='nicam_horizontal_limiter_thuburn_f32')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl
       type(c_ptr), value :: Qout_prev, Qout_prev_pl, Qout_post, Qout_post_pl
       integer(c_int) :: nicam_horizontal_limiter_thuburn_f32
     end function nicam_horizontal_limiter_thuburn_f32

     function nicam_vertical_limiter_thuburn_f32(ctx, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl) &
          bind(C, name='nicam_vertical_limiter_thuburn_f32')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl
       integer(c_int) :: nicam_vertical_limiter_thuburn_f32
     end function nicam_vertical_limiter_thuburn_f32

     function nicam_grid_icosahedral(glevel, rlevel, radius, GRD_x, GRD_x_pl) bind(C, name='nicam_grid_icosahedral')
       import :: c_int, c_double
       integer(c_int), value :: glevel, rlevel
       real(c_double), value :: radius
       real(c_double), intent(out) :: GRD_x(*), GRD_x_pl(*)
       integer(c_int) :: nicam_grid_icosahedral
     end function nicam_grid_icosahedral

     function nicam_set_rdgz(ctx, rdgz) bind(C, name='nicam_set_rdgz')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rdgz(*)
       integer(c_int) :: nicam_set_rdgz
     end function nicam_set_rdgz

     function nicam_set_vmetrics3d(ctx, RGSQRTH, RGSQRTH_pl, RGAM, RGAM_pl, RGAMH, RGAMH_pl, &
          C2WfactGz, C2WfactGz_pl) bind(C, name='nicam_set_vmetrics3d')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: RGSQRTH(*), RGSQRTH_pl(*), RGAM(*), RGAM_pl(*), RGAMH(*), RGAMH_pl(*)
       real(c_double), intent(in) :: C2WfactGz(*), C2WfactGz_pl(*)
       integer(c_int) :: nicam_set_vmetrics3d
     end function nicam_set_vmetrics3d

     function nicam_oprt3d_divdamp(ctx, ddivdx, ddivdx_pl, ddivdy, ddivdy_pl, ddivdz, ddivdz_pl, &
          rhogvx, rhogvx_pl, rhogvy, rhogvy_pl, rhogvz, rhogvz_pl, rhogw, rhogw_pl, coef_intp, coef_intp_pl, &
          coef_diff, coef_diff_pl) bind(C, name='nicam_oprt3d_divdamp')
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, ddivdx, ddivdx_pl, ddivdy, ddivdy_pl, ddivdz, ddivdz_pl
       type(c_ptr), value :: rhogvx, rhogvx_pl, rhogvy, rhogvy_pl, rhogvz, rhogvz_pl, rhogw, rhogw_pl
       type(c_ptr), value :: coef_intp, coef_intp_pl, coef_diff, coef_diff_pl
       integer(c_int) :: nicam_oprt3d_divdamp
     end function nicam_oprt3d_divdamp

     function c_strlen(s) bind(C, name='strlen')
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: c_strlen
     end function c_strlen
     function nicam_gmtr_setup(gall_1d, lall, gall_pl, have_pl, have_sgp, radius, GRD_x, GRD_x_pl, GRD_xt, &
          GRD_xt_pl, GRD_s, GRD_s_pl, GMTR_p, GMTR_p_pl, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl) &
          bind(C, name='nicam_gmtr_setup')
       import :: c_int, c_double
       integer(c_int), value :: gall_1d, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       real(c_double), value :: radius
       real(c_double), intent(in) :: GRD_x(*), GRD_x_pl(*), GRD_xt(*), GRD_xt_pl(*), GRD_s(*), GRD_s_pl(*)
       real(c_double), intent(out) :: GMTR_p(*), GMTR_p_pl(*), GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       integer(c_int) :: nicam_gmtr_setup
     end function nicam_gmtr_setup

     function nicam_oprt_diffusion_setup(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, &
          GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef_intp, coef_intp_pl, coef_diff, coef_diff_pl) &
          bind(C, name='nicam_oprt_diffusion_setup')
       import :: c_int, c_double
       integer(c_int), value :: gall_1d, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       real(c_double), intent(in) :: GMTR_p(*), GMTR_p_pl(*), GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       real(c_double), intent(out) :: coef_intp(*), coef_intp_pl(*), coef_diff(*), coef_diff_pl(*)
       integer(c_int) :: nicam_oprt_diffusion_setup
     end function nicam_oprt_diffusion_setup

     function nicam_oprt_divergence_setup(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, &
          GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef, coef_pl) bind(C, name='nicam_oprt_divergence_setup')
       import :: c_int, c_double
       integer(c_int), value :: gall_1d, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       real(c_double), intent(in) :: GMTR_p(*), GMTR_p_pl(*), GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       real(c_double), intent(out) :: coef(*), coef_pl(*)
       integer(c_int) :: nicam_oprt_divergence_setup
     end function nicam_oprt_divergence_setup

     function nicam_oprt_rotation_setup(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, &
          GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef, coef_pl) bind(C, name='nicam_oprt_rotation_setup')
       import :: c_int, c_double
       integer(c_int), value :: gall_1d, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       real(c_double), intent(in) :: GMTR_p(*), GMTR_p_pl(*), GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       real(c_double), intent(out) :: coef(*), coef_pl(*)
       integer(c_int) :: nicam_oprt_rotation_setup
     end function nicam_oprt_rotation_setup

     function nicam_oprt_gradient_setup(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, &
          GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef, coef_pl) bind(C, name='nicam_oprt_gradient_setup')
       import :: c_int, c_double
       integer(c_int), value :: gall_1d, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       real(c_double), intent(in) :: GMTR_p(*), GMTR_p_pl(*), GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       real(c_double), intent(out) :: coef(*), coef_pl(*)
       integer(c_int) :: nicam_oprt_gradient_setup
     end function nicam_oprt_gradient_setup

     function nicam_oprt_laplacian_setup(gall_1d, lall, gall_pl, have_pl, have_sgp, GMTR_p, GMTR_p_pl, &
          GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, coef, coef_pl) bind(C, name='nicam_oprt_laplacian_setup')
       import :: c_int, c_double
       integer(c_int), value :: gall_1d, lall, gall_pl, have_pl
       integer(c_int), intent(in) :: have_sgp(*)
       real(c_double), intent(in) :: GMTR_p(*), GMTR_p_pl(*), GMTR_t(*), GMTR_t_pl(*), GMTR_a(*), GMTR_a_pl(*)
       real(c_double), intent(out) :: coef(*), coef_pl(*)
       integer(c_int) :: nicam_oprt_laplacian_setup
     end function nicam_oprt_laplacian_setup

  end interface

  public :: nicam_version, nicam_last_error, nicam_ctx_create, nicam_ctx_destroy, nicam_ctx_set_stream
  public :: nicam_ctx_set_pointer_mode, nicam_ctx_set_tuning, nicam_ctx_device_bytes, nicam_const_setup
  public :: nicam_set_const, nicam_set_phys_const, nicam_grd_input_vgrid, nicam_grd_setup
  public :: nicam_set_vgrid_factors, nicam_set_hmetrics, nicam_set_vmetrics, nicam_horizontal_flux
  public :: nicam_horizontal_limiter_thuburn, nicam_vertical_limiter_thuburn, nicam_vi_rhow_solver_putcoef
  public :: nicam_vi_rhow_solver, nicam_ctx_synchronize, nicam_halo_map, nicam_pole_fill_map
  public :: nicam_comm_setup, nicam_comm_data_transfer, nicam_comm_var, nicam_oprt_diffusion
  public :: nicam_grid_icosahedral, nicam_set_rdgz, nicam_set_vmetrics3d, nicam_oprt3d_divdamp
  public :: nicam_gmtr_setup, nicam_oprt_diffusion_setup, nicam_oprt_divergence_setup
  public :: nicam_oprt_rotation_setup, nicam_oprt_gradient_setup, nicam_oprt_laplacian_setup

contains

  !> Error stop with the library's message when rc is not NICAM_OK (the
  !> reference's own fatal paths stop the program as well).
  subroutine nicam_check(rc, what)
    integer(c_int),   intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == NICAM_OK) return
    p = nicam_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*, '(A,A,I0,A,*(A))') trim(what), ' failed with status ', rc, ': ', msg(1:n)
    error stop 1
  end subroutine nicam_check

end module nicam_tt_f
