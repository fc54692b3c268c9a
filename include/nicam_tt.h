/*
 * nicam_tt.h -- C ABI of the MI355X (gfx950) NICAM tracer-transport and
 * vertical-implicit kernels.
 *
 * Drop-in for four Fortran module procedures of IcoAtmosBenchmark v1 (paths
 * relative to kernels/NICAM/ in the reference tree):
 *   horizontal_flux             dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379
 *   horizontal_limiter_thuburn  dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430
 *   vertical_limiter_thuburn    dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243
 *   vi_rhow_solver_putcoef      dyn_vi_rhow_solver/src/mod_vi.f90:58-89
 *   vi_rhow_solver              dyn_vi_rhow_solver/src/mod_vi.f90:93-306
 * and their set-up routines CONST_setup / GRD_setup / GRD_input_vgrid
 * (dyn_vi_rhow_solver/src/mod_misc.f90:447-570).
 *
 * Arrays keep the reference's Fortran column-major layout:
 *   regular x(gall, kall, lall[, n])   with gall = gall_1d*gall_1d and
 *                                       g = (j-1)*gall_1d + i fastest,
 *   pole    x_pl(gall_pl, kall, 2[, n]).
 * Argument order is the reference's, always in (regular, _pl) pairs.  What the
 * reference reads from module globals (sizes, singular-point flags, constants,
 * GMTR_*, GRD_*, VMTR_* metrics, the Mc/Ml/Mu module copies) lives in a
 * context created once per grid.
 *
 * Pointer mode: NICAM_PTR_DEVICE (default) -- every field argument is a
 * device pointer and the call is asynchronous on the context's stream;
 * NICAM_PTR_HOST -- fields are host pointers, the library stages them through
 * device buffers it owns, and the call returns after the results are back.
 * Set-up calls (metrics, coefficients) accept host or device pointers in
 * either mode and copy the data.
 *
 * Every function returns NICAM_OK (0) or a nonzero status and never aborts;
 * nicam_last_error() gives the message of the last failure on this thread.
 * A context is not thread-safe: use one per host thread, or serialise.
 */
#ifndef NICAM_TT_H
#define NICAM_TT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nicam_ctx_s *nicam_ctx;

enum nicam_status {
    NICAM_OK = 0,
    NICAM_ERR_ARG = 1,     /* bad size, null pointer, inconsistent argument */
    NICAM_ERR_STATE = 2,   /* a required set-up call has not been made       */
    NICAM_ERR_HIP = 3,     /* HIP runtime error (message has the details)    */
    NICAM_ERR_IO = 4,      /* file missing, short or malformed               */
    NICAM_ERR_NOMEM = 5    /* device or host allocation failed               */
};

enum nicam_ptr_mode { NICAM_PTR_DEVICE = 0, NICAM_PTR_HOST = 1 };

/* Library version string, e.g. "nicam_tt 0.1 gfx950". */
const char *nicam_version(void);
/* Message of the last failing call on the calling thread ("" if none). */
const char *nicam_last_error(void);

/* ---- context: replaces the ADM_* size globals
 *      (dyn_vi_rhow_solver/src/problem_size.inc:28-50) ----
 * gall_1d >= 4, kall >= 4, lall >= 1, gall_pl >= 3 (6 on the icosahedral
 * grid); have_sgp[lall] is ADM_have_sgp (nonzero = region holds a singular
 * point); have_pl is ADM_have_pl.  The context starts with CONST_setup's
 * values (EPS = DBL_EPSILON, HUGE = DBL_MAX) and the physical constants of
 * problem_size.inc:56-59 (GRAV 9.80616, Rdry 287, CVdry 717.5, alpha 1). */
int nicam_ctx_create(nicam_ctx *ctx, int gall_1d, int kall, int lall, int gall_pl,
                     int have_pl, const int *have_sgp);
int nicam_ctx_destroy(nicam_ctx ctx);
/* Stream for all work of this context (a hipStream_t; NULL = null stream). */
int nicam_ctx_set_stream(nicam_ctx ctx, void *hip_stream);
int nicam_ctx_set_pointer_mode(nicam_ctx ctx, int mode);
/* Launch-shape parameters (all default to 0 = library's choice):
 *   NICAM_TUNE_HFLUX_KPER   levels per work-group of horizontal_flux
 *   NICAM_TUNE_VLIM_KCHUNK  levels per work-item of the vertical limiter
 *   NICAM_TUNE_VI_VARIANT   the Thomas elimination of vi_rhow_solver:
 *                           1 = factors formed in the solve, kept in LDS;
 *                           2 = factors read from the factorisation that
 *                           putcoef stores, forward solution through rhogw;
 *                           3 = factors formed in the solve, column in
 *                           registers; 4 = stored factors, forward solution
 *                           in registers (3 and 4: kall = 42 or 98, else
 *                           automatic); 0 = automatic.  The adjoint
 *                           follows it: 1 and 3 form the transposed
 *                           system's factors in the solve, 2, 4 and 0 read
 *                           the ones putcoef stores (0 at kall = 42: the
 *                           column in registers)
 *   NICAM_TUNE_DIFF_KPER    levels per work-group of nicam_oprt_diffusion
 *   NICAM_TUNE_DDMP_KPER    levels per work-group of nicam_oprt3d_divdamp
 *   NICAM_TUNE_HLIM_VARIANT horizontal limiter, one tracer: 1 = the streamed
 *                           inputs loaded after the tracer tile's barrier,
 *                           2 = before it; 0 = automatic (2)
 * Results do not depend on these settings (every variant computes the same
 * floating-point operations in the same order). */
enum nicam_tuning {
    NICAM_TUNE_HFLUX_KPER = 1,
    NICAM_TUNE_VLIM_KCHUNK = 3,
    NICAM_TUNE_VI_VARIANT = 4,
    NICAM_TUNE_DIFF_KPER = 5,
    NICAM_TUNE_DDMP_KPER = 6,
    NICAM_TUNE_HLIM_VARIANT = 7
};
int nicam_ctx_set_tuning(nicam_ctx ctx, int param, int value);

/* Device memory currently held by the context, in bytes. */
size_t nicam_ctx_device_bytes(nicam_ctx ctx);

/* CONST_setup (mod_misc.f90:447-461): EPS = epsilon(1d0), HUGE = huge(1d0). */
int nicam_const_setup(nicam_ctx ctx);
/* Direct assignment of CONST_EPS / CONST_HUGE (problem_size.inc:54-55). */
int nicam_set_const(nicam_ctx ctx, double eps, double huge_val);
/* CONST_EPS / CONST_HUGE of the _f32 entry points.  The default is what
 * CONST_setup (mod_misc.f90:447-461) gives the reference built with -DSINGLE:
 * epsilon(0._SP) = 2^-23 and huge(0._SP).  Setting eps to FP64's epsilon
 * (2^-52, representable in FP32) makes the limiters' zerosw switches
 * (dyn_horiz_adv_limiter/src/mod_src_tracer.f90:274,
 * dyn_vert_adv_limiter/src/mod_src_tracer.f90:131,170,217) decide as the FP64
 * path does, which the mixed-precision step uses. */
int nicam_set_const_f32(nicam_ctx ctx, float eps, float huge_val);
/* CONST_GRAV, CONST_Rdry, CONST_CVdry, NON_HYDRO_ALPHA (problem_size.inc:56-59). */
int nicam_set_phys_const(nicam_ctx ctx, double grav, double rdry, double cvdry,
                         int non_hydro_alpha);

/* GRD_input_vgrid (mod_misc.f90:527-570): reads a Fortran unformatted
 * sequential big-endian file with 4-byte record markers: nlayer (int32),
 * gz(kall), gzh(kall) (float64).  nlayer must equal kall-2. */
int nicam_grd_input_vgrid(const char *fname, int kall, double *gz, double *gzh);
/* GRD_setup (mod_misc.f90:464-524) from gz/gzh(kall): computes and stores
 * GRD_rdgzh, GRD_afact, GRD_bfact in the context.  Any of the optional
 * outputs (kall doubles each) may be NULL. */
int nicam_grd_setup(nicam_ctx ctx, const double *gz, const double *gzh,
                    double *rdgzh_out, double *afact_out, double *bfact_out);
/* Direct assignment of GRD_rdgz (kall), which OPRT3D_divdamp reads;
 * nicam_grd_setup also sets it (dyn_divdamp/src/mod_misc.f90:480-496). */
int nicam_set_rdgz(nicam_ctx ctx, const double *rdgz);
/* Direct assignment of GRD_rdgzh, GRD_afact, GRD_bfact (kall each). */
int nicam_set_vgrid_factors(nicam_ctx ctx, const double *rdgzh, const double *afact,
                            const double *bfact);

/* Horizontal metrics, mod_misc.f90 module arrays (shapes as in
 * dyn_horiz_adv_flux/src/main.f90:93-100):
 *   GRD_xr(gall,1,lall,3,3)  GRD_xr_pl(gall_pl,1,2,3)
 *   GMTR_p(gall,1,lall,8)    GMTR_p_pl(gall_pl,1,2,8)
 *   GMTR_t(gall,1,lall,2,5)  GMTR_t_pl(gall_pl,1,2,5)
 *   GMTR_a(gall,1,lall,3,12) GMTR_a_pl(gall_pl,1,2,18)
 * Pole pointers may be NULL when have_pl == 0. */
int nicam_set_hmetrics(nicam_ctx ctx,
                       const double *GRD_xr, const double *GRD_xr_pl,
                       const double *GMTR_p, const double *GMTR_p_pl,
                       const double *GMTR_t, const double *GMTR_t_pl,
                       const double *GMTR_a, const double *GMTR_a_pl);

/* Vertical metrics VMTR_RGAM, _RGAMH, _RGSGAM2, _RGSGAM2H, _GSGAM2H
 * (gall,kall,lall) and (gall_pl,kall,2) (mod_misc.f90:101-112). */
int nicam_set_vmetrics(nicam_ctx ctx,
                       const double *RGAM, const double *RGAM_pl,
                       const double *RGAMH, const double *RGAMH_pl,
                       const double *RGSGAM2, const double *RGSGAM2_pl,
                       const double *RGSGAM2H, const double *RGSGAM2H_pl,
                       const double *GSGAM2H, const double *GSGAM2H_pl);

/* Vertical metrics that OPRT3D_divdamp reads (dyn_divdamp/src/mod_misc.f90,
 * module arrays VMTR_*): RGSQRTH, RGAM, RGAMH (gall,kall,lall) and _pl,
 * C2WfactGz(gall,kall,6,lall) and C2WfactGz_pl(gall_pl,kall,6,2).  RGAM and
 * RGAMH are the same context copies nicam_set_vmetrics sets. */
int nicam_set_vmetrics3d(nicam_ctx ctx,
                         const double *RGSQRTH, const double *RGSQRTH_pl,
                         const double *RGAM, const double *RGAM_pl,
                         const double *RGAMH, const double *RGAMH_pl,
                         const double *C2WfactGz, const double *C2WfactGz_pl);

/* ---- hot path ---------------------------------------------------------- */

/* Region widths: the flux and the h-limiter (and their _f32 and batched
 * forms) take gall_1d up to 514 (gl9/rl0, gl10/rl1); the v-limiter and the
 * rho*w solve have no limit; the step glue's remap, the grouped limiter and
 * the flux and h-limiter adjoints take gall_1d up to 160, the operators up to
 * 255.  Wider regions fail with NICAM_ERR_ARG before anything is launched. */

/* horizontal_flux (dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379).
 *   out flx_h(gall,kall,lall,6), flx_h_pl(gall_pl,kall,2)
 *   out GRD_xc(gall,kall,lall,3,3), GRD_xc_pl(gall_pl,kall,2,3)
 *   in  rho, rhovx, rhovy, rhovz (gall,kall,lall) and _pl
 *   dt  (the driver passes SET_dt_large = 360 s)
 * Point 1 of flx_h_pl and GRD_xc_pl is not written (as in the reference). */
int nicam_horizontal_flux(nicam_ctx ctx,
                          double *flx_h, double *flx_h_pl,
                          double *GRD_xc, double *GRD_xc_pl,
                          const double *rho, const double *rho_pl,
                          const double *rhovx, const double *rhovx_pl,
                          const double *rhovy, const double *rhovy_pl,
                          const double *rhovz, const double *rhovz_pl,
                          double dt);

/* horizontal_limiter_thuburn (dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430).
 *   inout q_a(gall,kall,lall,6), q_a_pl(gall_pl,kall,2)
 *   in    q, d (gall,kall,lall); ch, cmask (gall,kall,lall,6); and _pl
 *   out   Qout_prev(gall,kall,lall,2), Qout_prev_pl(gall_pl,kall,2,2)
 *   in    Qout_post(gall,kall,lall,2), Qout_post_pl(gall_pl,kall,2,2)
 * Qout_prev is the allowable range before the halo exchange; the clamp uses
 * Qout_post, the range after it, exactly as the reference kernel does
 * (:365-375).  Qout_post = NULL asks the library to do that exchange itself
 * (the COMM_data_transfer of Qout the kernel cut out), in place on Qout_prev
 * as NICAM does: Qout_prev then returns the range after the exchange.  It
 * needs nicam_comm_setup; Qout_post_pl is then ignored. */
int nicam_horizontal_limiter_thuburn(nicam_ctx ctx,
                                     double *q_a, double *q_a_pl,
                                     const double *q, const double *q_pl,
                                     const double *d, const double *d_pl,
                                     const double *ch, const double *ch_pl,
                                     const double *cmask, const double *cmask_pl,
                                     double *Qout_prev, double *Qout_prev_pl,
                                     const double *Qout_post, const double *Qout_post_pl);

/* vertical_limiter_thuburn (dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243).
 *   inout q_h(gall,kall,lall), q_h_pl
 *   in    q, d (gall,kall,lall), ck(gall,kall,lall,2), and _pl */
int nicam_vertical_limiter_thuburn(nicam_ctx ctx,
                                   double *q_h, double *q_h_pl,
                                   const double *q, const double *q_pl,
                                   const double *d, const double *d_pl,
                                   const double *ck, const double *ck_pl);

/* vi_rhow_solver_putcoef (dyn_vi_rhow_solver/src/mod_vi.f90:58-89): keeps a
 * device copy of Mc, Ml, Mu (gall,kall,lall) and _pl in the context. */
int nicam_vi_rhow_solver_putcoef(nicam_ctx ctx,
                                 const double *Mc, const double *Mc_pl,
                                 const double *Ml, const double *Ml_pl,
                                 const double *Mu, const double *Mu_pl);

/* vi_rhow_solver (dyn_vi_rhow_solver/src/mod_vi.f90:93-306).
 *   inout rhogw(gall,kall,lall), rhogw_pl (levels kmin and kmax+1 are read)
 *   in    rhogw0, preg0, rhog0, Srho, Sw, Spre (gall,kall,lall) and _pl
 *   dt    (the driver passes SET_dt_small = 60 s)
 * Needs nicam_vi_rhow_solver_putcoef, nicam_set_vmetrics and
 * nicam_grd_setup (or nicam_set_vgrid_factors) first. */
int nicam_vi_rhow_solver(nicam_ctx ctx,
                         double *rhogw, double *rhogw_pl,
                         const double *rhogw0, const double *rhogw0_pl,
                         const double *preg0, const double *preg0_pl,
                         const double *rhog0, const double *rhog0_pl,
                         const double *Srho, const double *Srho_pl,
                         const double *Sw, const double *Sw_pl,
                         const double *Spre, const double *Spre_pl,
                         double dt);

/* ---- neighbouring operators on the same grid --------------------------- */

/* OPRT_diffusion (dyn_diffusion/src/mod_oprt.f90:41-296).
 *   out dscl(gall,kall,lall), dscl_pl(gall_pl,kall,2)
 *   in  scl, kh (gall,kall,lall) and _pl
 *   in  coef_intp(gall,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2)
 *   in  coef_diff(gall,6,3,lall),   coef_diff_pl(5,3,2)
 * As in the reference, the Y and Z passes read the X component of vt as the
 * first point of each pair (:170-214), and dscl_pl is zeroed when the context
 * has no pole regions (dscl_pl may then be NULL). */
int nicam_oprt_diffusion(nicam_ctx ctx,
                         double *dscl, double *dscl_pl,
                         const double *scl, const double *scl_pl,
                         const double *kh, const double *kh_pl,
                         const double *coef_intp, const double *coef_intp_pl,
                         const double *coef_diff, const double *coef_diff_pl);

/* OPRT3D_divdamp (dyn_divdamp/src/mod_oprt3d.f90:41-402).
 *   out ddivdx, ddivdy, ddivdz (gall,kall,lall) and _pl (gall_pl,kall,2)
 *   in  rhogvx, rhogvy, rhogvz, rhogw (gall,kall,lall) and _pl
 *   in  coef_intp, coef_diff and _pl as for nicam_oprt_diffusion
 * Needs nicam_set_vmetrics3d and GRD_rdgz (nicam_grd_setup or
 * nicam_set_rdgz).  The outputs' pole arrays are zeroed when the context has
 * no pole regions (they may then be NULL, all three together). */
int nicam_oprt3d_divdamp(nicam_ctx ctx,
                         double *ddivdx, double *ddivdx_pl,
                         double *ddivdy, double *ddivdy_pl,
                         double *ddivdz, double *ddivdz_pl,
                         const double *rhogvx, const double *rhogvx_pl,
                         const double *rhogvy, const double *rhogvy_pl,
                         const double *rhogvz, const double *rhogvz_pl,
                         const double *rhogw, const double *rhogw_pl,
                         const double *coef_intp, const double *coef_intp_pl,
                         const double *coef_diff, const double *coef_diff_pl);

/* Blocks until all work queued on the context's stream has finished. */
int nicam_ctx_synchronize(nicam_ctx ctx);

/* ---- halo exchange on one GPU ------------------------------------------ */

/* The halo map of the icosahedral grid at region level rlevel, as the
 * reference builds it (communication/src/mod_adm.f90:629-894 region topology,
 * communication/src/mod_comm.f90:213-609 region-to-region lists, :924-1380
 * pole lists, :1384-1502 singular points).  Host-only; needs no context.
 * Entry e copies point (src_kind, src_l, src_g) into (dst_kind, dst_l, dst_g):
 * kind 0 is a regular point, l the region (0..10*4^rlevel-1) and g the 0-based
 * point (j*gall_1d + i); kind 1 is a pole point, l 0 = north, 1 = south, g
 * 0..gall_pl-1.  With capacity < 1 only *n_out is set.  have_sgp (may be
 * NULL) receives the 10*4^rlevel singular-point flags ADM_have_sgp: 1 for the
 * regions whose west vertex joins only three regions (mod_adm.f90:257-264). */
int nicam_halo_map(int rlevel, int gall_1d, int gall_pl, int capacity, int *n_out,
                   int *dst_kind, int *dst_l, int *dst_g,
                   int *src_kind, int *src_l, int *src_g, int *have_sgp);

/* COMM_setup (communication/src/mod_comm.f90:147-209) for a context holding
 * every region of the level-rlevel grid (lall = 10*4^rlevel, have_pl = 1):
 * builds the halo map and keeps it on the device.  The context's have_sgp
 * must agree with the topology. */
int nicam_comm_setup(nicam_ctx ctx, int rlevel);

/* The point whose value COMM_var copies into each pole centre before its
 * exchange (mod_comm.f90:1895-2101 with RGNMNG_rgn4pl, mod_adm.f90:393-405):
 * pole p (0 north, 1 south) takes regular point src_g[p] of region src_l[p]
 * (0-based).  Host-only. */
int nicam_pole_fill_map(int rlevel, int gall_1d, int *src_l, int *src_g);

/* COMM_data_transfer (communication/src/mod_comm.f90:1506-1891) on the
 * device: fills the halo of var(gall,kmax,lall,vmax) and the pole rings of
 * var_pl(gall_pl,kmax,2,vmax) from their owners, then applies the singular
 * point copies, all in one gather launch.  kmax and vmax are the extents
 * the reference takes from the assumed-shape arguments (kmax need not be
 * kall).  Needs nicam_comm_setup. */
int nicam_comm_data_transfer(nicam_ctx ctx, double *var, double *var_pl, int kmax, int vmax);

/* An icosahedral grid (host-only): the positions GRD_x(gall,1,lall,3) and
 * GRD_x_pl(6,1,2,3) that the metrics kernel reads from its snapshot
 * (dyn_metrics/src/main.f90:217-218), for gall_1d = 2^(glevel-rlevel)+2 and
 * lall = 10*4^rlevel, on a sphere of the given radius.  Inner points come from
 * the 10 diamonds of the icosahedron refined glevel times by great-circle
 * midpoints; halo points and pole rings are then filled with the halo map, as
 * COMM_data_transfer(GRD_x, GRD_x_pl) does.  Halo points outside the map are 0. */
int nicam_grid_icosahedral(int glevel, int rlevel, double radius, double *GRD_x, double *GRD_x_pl);

/* COMM_var (communication/src/mod_comm.f90:1895-2101): each pole centre takes
 * its value from the point nicam_pole_fill_map names, then the exchange of
 * nicam_comm_data_transfer runs.  Same arguments. */
int nicam_comm_var(nicam_ctx ctx, double *var, double *var_pl, int kmax, int vmax);

/* The triangle vertices, latitude/longitude and arc midpoints of a grid whose
 * GRD_x halo is filled (host-only): GRD_xt(gall,1,lall,2,3) (TI, TJ),
 * GRD_xt_pl(gall_pl,1,2,3), GRD_s(gall,1,lall,2) (lat, lon),
 * GRD_s_pl(gall_pl,1,2,2), GRD_xr(gall,1,lall,3,3) (AI, AIJ, AJ) and
 * GRD_xr_pl(gall_pl,1,2,3).  The metrics kernel reads GRD_xt and GRD_s from
 * its snapshot (dyn_metrics/src/main.f90:219-222) and the flux kernel reads
 * GRD_xr (dyn_horiz_adv_flux/src/main.f90:157-158); the reference has no
 * generator for them.  TI/TJ are the normalised centroids of the triangles
 * (ij, ip1j, ip1jp1) and (ij, ip1jp1, ijp1), the pole ring's of (pole, v, v+1);
 * lat/lon as VECTR_xyz2latlon (dyn_metrics/src/mod_vector.f90:56-113); the
 * arc midpoints are the normalised midpoints of the hexagon edges whose ends
 * GMTR_a_setup uses for HN/HT (dyn_metrics/src/mod_gmtr.f90:561-650).  Every
 * point lies on the sphere of the given radius.  have_sgp has lall entries. */
int nicam_grid_vertices(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp, double radius,
                        const double *GRD_x, const double *GRD_x_pl, double *GRD_xt, double *GRD_xt_pl,
                        double *GRD_s, double *GRD_s_pl, double *GRD_xr, double *GRD_xr_pl);

/* GMTR_p_setup, GMTR_t_setup and GMTR_a_setup
 * (dyn_metrics/src/mod_gmtr.f90:46-741) on the sphere with spherical
 * polygons, the reference's configuration (dyn_metrics/src/problem_size.inc:69-78),
 * host-only.  Inputs as the metrics kernel reads them
 * (dyn_metrics/src/main.f90:217-222), outputs GMTR_p(gall,1,lall,8),
 * GMTR_p_pl(gall_pl,1,2,8), GMTR_t(gall,1,lall,2,5), GMTR_t_pl(gall_pl,1,2,5),
 * GMTR_a(gall,1,lall,3,12) and GMTR_a_pl(gall_pl,1,2,18), in the operand order
 * of the reference, so they match its compiled kernel.  The _pl arguments are
 * ignored when have_pl is 0. */
int nicam_gmtr_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp, double radius,
                     const double *GRD_x, const double *GRD_x_pl, const double *GRD_xt, const double *GRD_xt_pl,
                     const double *GRD_s, const double *GRD_s_pl, double *GMTR_p, double *GMTR_p_pl,
                     double *GMTR_t, double *GMTR_t_pl, double *GMTR_a, double *GMTR_a_pl);

/* OPRT_diffusion_setup (dyn_metrics/src/mod_oprt.f90:1235-1390), host-only:
 * from the metrics of nicam_gmtr_setup, the coefficients that
 * nicam_oprt_diffusion and nicam_oprt3d_divdamp read,
 * coef_intp(gall,3,3,2,lall), coef_intp_pl(gall_pl,3,3,2),
 * coef_diff(gall,6,3,lall) and coef_diff_pl(gall_pl-1,3,2), bit for bit as the
 * reference computes them (halo entries included).  The _pl arguments are
 * ignored when have_pl is 0. */
int nicam_oprt_diffusion_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                               const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                               const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                               double *coef_intp, double *coef_intp_pl, double *coef_diff, double *coef_diff_pl);

/* OPRT_divergence_setup, OPRT_rotation_setup, OPRT_gradient_setup
 * (dyn_metrics/src/mod_oprt.f90:45-269, 272-496, 499-734), host-only: the
 * 7-point stencil coefficients coef(gall,0:6,3,lall) and
 * coef_pl(0:gall_pl-1,3,2) of the divergence (arc normals HN), rotation (arc
 * tangents HT) and gradient operators, from the metrics of nicam_gmtr_setup,
 * bit for bit as the reference computes them. */
int nicam_oprt_divergence_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                                const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                                const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                                double *coef_div, double *coef_div_pl);
int nicam_oprt_rotation_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                              const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                              const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                              double *coef_rot, double *coef_rot_pl);
int nicam_oprt_gradient_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                              const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                              const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                              double *coef_grad, double *coef_grad_pl);
/* OPRT_laplacian_setup (dyn_metrics/src/mod_oprt.f90:737-1232), host-only:
 * coef_lap(gall,0:6,lall) and coef_lap_pl(0:gall_pl-1,2). */
int nicam_oprt_laplacian_setup(int gall_1d, int lall, int gall_pl, int have_pl, const int *have_sgp,
                               const double *GMTR_p, const double *GMTR_p_pl, const double *GMTR_t,
                               const double *GMTR_t_pl, const double *GMTR_a, const double *GMTR_a_pl,
                               double *coef_lap, double *coef_lap_pl);

/* ---- step glue: the inputs the reference's kernels take "computed in
 * advance by other subroutine" (its kernels have no such code).  The NICAM
 * forms of the Miura upwind scheme (docs/NICAM/src/20_intro_nicam.tex:368-438);
 * no reference interface is replaced, the argument conventions follow the
 * kernels' own (regular, _pl) pairs and layouts.  Pole arguments may be NULL
 * when have_pl == 0. ---- */

/* The gradient of a scalar with the coefficients of OPRT_gradient_setup
 * (dyn_metrics/src/mod_oprt.f90:499-734; nicam_oprt_gradient_setup):
 * grad(gall,kall,lall,3) = sum over ij, ip1j, ip1jp1, ijp1, im1j, im1jm1, ijm1
 * of coef_grad(gall,0:6,3,lall) * scl, on the flat range [gmin, gmax], zero
 * elsewhere; grad_pl(gall_pl,kall,2,3) at the pole centre from
 * coef_grad_pl(0:gall_pl-1,3,2), zero on the ring.  The halo and pole-ring
 * values come from nicam_comm_data_transfer(ctx, grad, grad_pl, kall, 3). */
int nicam_oprt_gradient(nicam_ctx ctx, double *grad, double *grad_pl, const double *scl, const double *scl_pl,
                        const double *coef_grad, const double *coef_grad_pl);

/* Courant numbers and upwind masks of the six faces for the horizontal
 * limiter: ch(gall,kall,lall,6) = flx_h / rhog, cmask = 0.5 - sign(0.5, ch - EPS)
 * (1 on inflow faces); the pole ring against the centre's rhog_pl, the centre
 * entry 0. */
int nicam_tracer_courant(nicam_ctx ctx, double *ch, double *ch_pl, double *cmask, double *cmask_pl,
                         const double *flx_h, const double *flx_h_pl, const double *rhog, const double *rhog_pl);

/* Face values q_a(gall,kall,lall,6) for the horizontal limiter: for each arc
 * AI, AIJ, AJ with origin in [gmin-1, gmax]^2, the upwind node's q plus its
 * gradient dotted with GRD_xc(gall,kall,lall,3,3) - GRD_x(gall,1,lall,3) of
 * that node.  Face a of the origin takes the side cmask(origin,a) picks and
 * face a+3 of the neighbour the side cmask(neighbour,a+3) picks (each face
 * by its owner's mask, as the limiter reads them); other faces are kept.
 * The pole ring faces likewise (GRD_x_pl(gall_pl,1,2,3)). */
int nicam_horizontal_remap(nicam_ctx ctx, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                           const double *grad, const double *grad_pl, const double *cmask, const double *cmask_pl,
                           const double *GRD_xc, const double *GRD_xc_pl, const double *GRD_x,
                           const double *GRD_x_pl);

/* The flux-form update of the tracer mass: rhogq -= sum over the six faces
 * of flx_h * q_a, on the inner cells and the pole centres. */
int nicam_tracer_update(nicam_ctx ctx, double *rhogq, double *rhogq_pl, const double *flx_h, const double *flx_h_pl,
                        const double *q_a, const double *q_a_pl);

/* Adjoints of the four above and of the halo fill, for the differentiable
 * tracer step.  No reference counterpart (the reference has no glue).  Each
 * overwrites its outputs (the *_b arrays it writes) with the transpose of its
 * forward's derivative applied to the given cotangents, cmask held fixed;
 * every output is summed by one work-item in a fixed order, so the results are
 * deterministic.  nicam_oprt_gradient_adjoint: scl_b from grad_b.
 * nicam_tracer_courant_adjoint: flx_h_b = ch_b / rhog and rhog_b =
 * -(sum of ch_b * ch) / rhog.  nicam_horizontal_remap_adjoint: q_b, grad_b and
 * GRD_xc_b from q_a_b (entries the forward does not write get none).
 * nicam_tracer_update_adjoint: flx_h_b and q_a_b from rhogq_b (rhogq's own
 * cotangent passes through unchanged).  nicam_comm_data_transfer_adjoint, in
 * place: every halo destination's cotangent is added to its source's and then
 * cleared (nicam_comm_setup first). */
int nicam_oprt_gradient_adjoint(nicam_ctx ctx, double *scl_b, double *scl_b_pl, const double *grad_b,
                                const double *grad_b_pl, const double *coef_grad, const double *coef_grad_pl);
int nicam_tracer_courant_adjoint(nicam_ctx ctx, double *flx_h_b, double *flx_h_b_pl, double *rhog_b,
                                 double *rhog_b_pl, const double *ch_b, const double *ch_b_pl, const double *flx_h,
                                 const double *flx_h_pl, const double *rhog, const double *rhog_pl);
int nicam_horizontal_remap_adjoint(nicam_ctx ctx, double *q_b, double *q_b_pl, double *grad_b, double *grad_b_pl,
                                   double *GRD_xc_b, double *GRD_xc_b_pl, const double *q_a_b, const double *q_a_b_pl,
                                   const double *grad, const double *grad_pl, const double *cmask,
                                   const double *cmask_pl, const double *GRD_xc, const double *GRD_xc_pl,
                                   const double *GRD_x, const double *GRD_x_pl);
int nicam_tracer_update_adjoint(nicam_ctx ctx, double *flx_h_b, double *flx_h_b_pl, double *q_a_b, double *q_a_b_pl,
                                const double *rhogq_b, const double *rhogq_b_pl, const double *flx_h,
                                const double *flx_h_pl, const double *q_a, const double *q_a_pl);
int nicam_comm_data_transfer_adjoint(nicam_ctx ctx, double *var_b, double *var_b_pl, int kmax, int vmax);

/* FP32 variants of the flux and the two limiters: the same arguments as the
 * FP64 entry points above, every array and dt in single precision.  They
 * reproduce, bit for bit, the reference's own kernels built with -DSINGLE
 * (mod_precision.f90: RP = SP): the same operations in the same order in
 * IEEE single precision, CONST_EPS = epsilon(0._SP), CONST_HUGE = huge(0._SP),
 * and the horizontal metrics of nicam_set_hmetrics rounded to single
 * precision, as real(x, kind=SP) rounds them.  The FP64 VI solve completes the
 * mixed-precision step. */
int nicam_horizontal_flux_f32(nicam_ctx ctx, float *flx_h, float *flx_h_pl, float *GRD_xc, float *GRD_xc_pl,
                              const float *rho, const float *rho_pl, const float *rhovx, const float *rhovx_pl,
                              const float *rhovy, const float *rhovy_pl, const float *rhovz, const float *rhovz_pl,
                              float dt);
int nicam_horizontal_limiter_thuburn_f32(nicam_ctx ctx, float *q_a, float *q_a_pl, const float *q,
                                         const float *q_pl, const float *d, const float *d_pl, const float *ch,
                                         const float *ch_pl, const float *cmask, const float *cmask_pl,
                                         float *Qout_prev, float *Qout_prev_pl, const float *Qout_post,
                                         const float *Qout_post_pl);
int nicam_vertical_limiter_thuburn_f32(nicam_ctx ctx, float *q_h, float *q_h_pl, const float *q, const float *q_pl,
                                       const float *d, const float *d_pl, const float *ck, const float *ck_pl);

/* The vertical half of the step glue.  No reference counterpart: the
 * reference's vertical limiter takes ck and q_h as inputs
 * (dyn_vert_adv_limiter/src/main.f90), computed by code it does not hold.
 * Arrays (gall,kall,lall), poles (gall_pl,kall,2); kmin = 2, kmax = kall-1. */

/* Vertical mass flux through the half levels over dt:
 * flx_v(k) = rhogw_vm(k) * dt on kmin+1..kmax, 0 on every other level, where
 * rhogw_vm is the contravariant vertical mass flux OPRT3D_divdamp forms
 * (dyn_divdamp/src/mod_oprt3d.f90:148-155) from rhogvx/vy/vz at k and k-1,
 * rhogw and the metrics of nicam_set_vmetrics3d. */
int nicam_vertical_mass_flux(nicam_ctx ctx, double *flx_v, double *flx_v_pl, const double *rhogvx,
                             const double *rhogvx_pl, const double *rhogvy, const double *rhogvy_pl,
                             const double *rhogvz, const double *rhogvz_pl, const double *rhogw,
                             const double *rhogw_pl, double dt);

/* Vertical Courant numbers for nicam_vertical_limiter_thuburn:
 * ck(gall,kall,lall,2) with ck(k,1) = -flx_v(k)/rhog(k)*GRD_rdgz(k) (lower
 * face) and ck(k,2) = flx_v(k+1)/rhog(k)*GRD_rdgz(k) (upper face) on the
 * interior faces of kmin..kmax, 0 elsewhere; negative where the face is an
 * inflow.  The closed faces hold +0 in ck(kmin,1) and -0 in ck(kmax+1,1):
 * the limiter's flags (dyn_vert_adv_limiter/src/mod_src_tracer.f90:113-114)
 * read both as outflow, so its bounds never take q at kmin-1 or kmax+1. */
int nicam_vertical_courant(nicam_ctx ctx, double *ck, double *ck_pl, const double *flx_v, const double *flx_v_pl,
                           const double *rhog, const double *rhog_pl);

/* Half-level face values q_h(k) = GRD_afact(k) q(k) + GRD_bfact(k) q(k-1) on
 * kmin+1..kmax (the factors of nicam_grd_setup); other levels kept. */
int nicam_vertical_face(nicam_ctx ctx, double *q_h, double *q_h_pl, const double *q, const double *q_pl);

/* Flux-form vertical update of the tracer mass on kmin..kmax:
 * rhogq(k) -= (flx_v(k+1) q_h(k+1) - flx_v(k) q_h(k)) * GRD_rdgz(k), the
 * boundary faces kmin and kmax+1 carrying no flux. */
int nicam_tracer_update_vertical(nicam_ctx ctx, double *rhogq, double *rhogq_pl, const double *flx_v,
                                 const double *flx_v_pl, const double *q_h, const double *q_h_pl);

/* Adjoints of the vertical glue above, for the differentiable tracer step
 * (no reference counterpart; outputs overwritten, each a fixed-order sum per
 * cell-level).  nicam_vertical_courant_adjoint: flx_v_b and rhog_b from ck_b
 * over the interior faces the forward forms; nicam_vertical_face_adjoint: q_b
 * from q_h_b; nicam_tracer_update_vertical_adjoint: flx_v_b and q_h_b from
 * rhogq_b (rhogq's own cotangent passes through unchanged). */
int nicam_vertical_courant_adjoint(nicam_ctx ctx, double *flx_v_b, double *flx_v_b_pl, double *rhog_b,
                                   double *rhog_b_pl, const double *ck_b, const double *ck_b_pl, const double *flx_v,
                                   const double *flx_v_pl, const double *rhog, const double *rhog_pl);
int nicam_vertical_face_adjoint(nicam_ctx ctx, double *q_b, double *q_b_pl, const double *q_h_b,
                                const double *q_h_b_pl);
int nicam_tracer_update_vertical_adjoint(nicam_ctx ctx, double *flx_v_b, double *flx_v_b_pl, double *q_h_b,
                                         double *q_h_b_pl, const double *rhogq_b, const double *rhogq_b_pl,
                                         const double *flx_v, const double *flx_v_pl, const double *q_h,
                                         const double *q_h_pl);

/* Coefficients of the tridiagonal rho*w system that nicam_vi_rhow_solver
 * reads, computed on the device and installed in the context (what
 * nicam_vi_rhow_solver_putcoef installs from arrays).  No reference
 * counterpart: the reference reads Mc/Ml/Mu from its input file ("calculated
 * in advance by other subroutine", docs/NICAM/src/33_vi_rhow_solver.tex:188-190).
 * Collected from the discretised Helmholtz equation of
 * docs/NICAM/src/20_intro_nicam.tex:480-600 at the half levels kmin+1..kmax
 * (0 on the other levels), with A^o = VMTR_RGSGAM2, gam^2 = 1/VMTR_RGAMH^2,
 * G^1/2 = VMTR_GSGAM2H/gam^2 (nicam_set_vmetrics), GRD_rdgz/rdgzh/cfact/dfact
 * (nicam_grd_setup) and the constants of nicam_set_phys_const:
 *   eth(gall,kall,lall)     h, the enthalpy of the large step at the half levels
 *   g_tilde(gall,kall,lall) the effective gravity of the large step at the half levels
 * The equation's scaling is the one vi_rhow_solver's Sall carries (the factor
 * Cv/(R dt^2) of its source, dyn_vi_rhow_solver/src/mod_vi.f90:192-205). */
int nicam_vi_rhow_update_matrix(nicam_ctx ctx, const double *eth, const double *eth_pl, const double *g_tilde,
                                const double *g_tilde_pl, double dt);

/* Copies the context's Mc/Ml/Mu (gall,kall,lall), _pl (gall_pl,kall,2) out. */
int nicam_vi_rhow_getcoef(nicam_ctx ctx, double *Mc, double *Mc_pl, double *Ml, double *Ml_pl, double *Mu,
                          double *Mu_pl);

/* vi_rhow_solver with FP32 prognostic and source inputs (the mixed-precision
 * step): rhogw0, preg0, rhog0, Srho, Sw, Spre are float arrays, rhogw stays
 * double, and the tridiagonal elimination and every operation run in FP64
 * on the exactly widened inputs.  The source term reads FP32 copies of
 * VMTR_RGAM, VMTR_RGAMH, VMTR_RGSGAM2 kept by nicam_set_vmetrics; Mc/Ml/Mu,
 * RGSGAM2H and GSGAM2H stay FP64.  The result equals nicam_vi_rhow_solver run
 * on the same inputs and those three metrics rounded to FP32. */
int nicam_vi_rhow_solver_mixed(nicam_ctx ctx,
                               double *rhogw, double *rhogw_pl,
                               const float *rhogw0, const float *rhogw0_pl,
                               const float *preg0, const float *preg0_pl,
                               const float *rhog0, const float *rhog0_pl,
                               const float *Srho, const float *Srho_pl,
                               const float *Sw, const float *Sw_pl,
                               const float *Spre, const float *Spre_pl,
                               double dt);

/* ---- Adjoints (hand-written reverse of the kernels above) ----------------
 * Each computes the transpose of its forward kernel's derivative at fixed
 * coefficients.  No counterpart exists in the reference: these replace no
 * reference interface; they serve gradient computations through the
 * forward kernels (data assimilation, parameter estimation).
 *
 * vi_rhow_solver adjoint (transpose of mod_vi.f90:93-306 as a linear map of
 * rhogw(kmin), rhogw(kmax+1), rhogw0, preg0, rhog0, Srho, Sw, Spre, at the
 * context's Mc/Ml/Mu, metrics and the given dt):
 *   inout rhogw_b   in: cotangent of the solver's output rhogw;
 *                   out: cotangent of its input rhogw (zero on the levels the
 *                   solver overwrites without reading; level 1 passes through)
 *   out   rhogw0_b, preg0_b, rhog0_b, Srho_b, Sw_b, Spre_b (every level written)
 * Same set-up as nicam_vi_rhow_solver. */
int nicam_vi_rhow_solver_adjoint(nicam_ctx ctx,
                                 double *rhogw_b, double *rhogw_b_pl,
                                 double *rhogw0_b, double *rhogw0_b_pl,
                                 double *preg0_b, double *preg0_b_pl,
                                 double *rhog0_b, double *rhog0_b_pl,
                                 double *Srho_b, double *Srho_b_pl,
                                 double *Sw_b, double *Sw_b_pl,
                                 double *Spre_b, double *Spre_b_pl,
                                 double dt);

/* Adjoint of horizontal_flux in (rho, rhovx, rhovy, rhovz) at the context's
 * metrics and the given dt (dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379):
 *   in  flx_h_b (gall,kall,lall,6), GRD_xc_b (gall,kall,lall,3,3) and _pl:
 *       cotangents of the outputs (entries the forward leaves at zero are
 *       not read)
 *   out rho_b, rhovx_b, rhovy_b, rhovz_b and _pl (every point written)
 *   in  rho, rhovx, rhovy, rhovz and _pl: the forward's inputs
 * The derivative of GRD_xc in rho is taken where max(rho_arc, EPS) selects
 * rho_arc. */
int nicam_horizontal_flux_adjoint(nicam_ctx ctx,
                                  const double *flx_h_b, const double *flx_h_b_pl,
                                  const double *GRD_xc_b, const double *GRD_xc_b_pl,
                                  double *rho_b, double *rho_b_pl,
                                  double *rhovx_b, double *rhovx_b_pl,
                                  double *rhovy_b, double *rhovy_b_pl,
                                  double *rhovz_b, double *rhovz_b_pl,
                                  const double *rho, const double *rho_pl,
                                  const double *rhovx, const double *rhovx_pl,
                                  const double *rhovy, const double *rhovy_pl,
                                  const double *rhovz, const double *rhovz_pl,
                                  double dt);

/* horizontal_limiter_thuburn that also records its branch mask: the same
 * arguments and results as nicam_horizontal_limiter_thuburn, plus one uint8
 * per cell-level (gall,kall,lall): bit m-1 set where face m (1-3) of a cell
 * with i,j in [gmin-1,gmax] changed (0 elsewhere); mask_pl one uint8 per pole
 * point (gall_pl,kall,2), bit 0 set where the ring face changed. */
int nicam_horizontal_limiter_thuburn_mask(nicam_ctx ctx,
                                          double *q_a, double *q_a_pl,
                                          const double *q, const double *q_pl,
                                          const double *d, const double *d_pl,
                                          const double *ch, const double *ch_pl,
                                          const double *cmask, const double *cmask_pl,
                                          double *Qout_prev, double *Qout_prev_pl,
                                          const double *Qout_post, const double *Qout_post_pl,
                                          uint8_t *mask, uint8_t *mask_pl);

/* Adjoint of horizontal_limiter_thuburn with Qout_post supplied, at fixed d
 * and cmask (dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430):
 *   inout q_a_b        in: cotangent of the limited q_a; out: cotangent of
 *                      q_a's input (faces 4-6 the kernel overwrites get 0;
 *                      faces it does not touch pass through)
 *   out   q_b          cotangent of q
 *   in    Qout_prev_b  cotangent of Qout_prev (the ring of Qout_prev_pl is a
 *                      constant and is not read)
 *   out   Qout_post_b  cotangent of Qout_post
 *   in    q_a          the forward's output; q, d, ch, cmask, Qout_post its
 *                      inputs; mask its mask
 *   out   ch_b         cotangent of ch (NULL: not formed; with poles ch_b_pl
 *                      too): Qout of a cell depends on its six Courant
 *                      numbers through Cin, Cout and the inflow sums
 *                      (HLIM:216-284); at ch = 0 the right-hand derivative
 * A changed face value is attributed, in this order, to Qout_post min,
 * Qout_post max, the arc's Qin min, Qin max, by equality with q_a.  Every
 * cotangent is a sum in a fixed order (the order of a sequential sweep over
 * the cells), so the results do not vary between runs.
 * Half passes: with q_a_b NULL only the rows run (Qout_prev_b's terms; q_a_b,
 * Qout_post_b, q_a, Qout_post and mask may be NULL); with Qout_prev_b NULL only
 * the faces (q_a_b's; ch_b must then be NULL).  Each gives, bit for bit, the
 * full pass with the other cotangent zero; a caller that exchanges Qout
 * between the passes (the exchange mode's transpose) runs the faces, the
 * transpose of the exchange on Qout_post_b, then the rows. */
int nicam_horizontal_limiter_thuburn_adjoint(nicam_ctx ctx,
                                             double *q_a_b, double *q_a_b_pl,
                                             double *q_b, double *q_b_pl,
                                             const double *Qout_prev_b, const double *Qout_prev_b_pl,
                                             double *Qout_post_b, double *Qout_post_b_pl,
                                             const double *q_a, const double *q_a_pl,
                                             const double *q, const double *q_pl,
                                             const double *d, const double *d_pl,
                                             const double *ch, const double *ch_pl,
                                             const double *cmask, const double *cmask_pl,
                                             const double *Qout_post, const double *Qout_post_pl,
                                             const uint8_t *mask, const uint8_t *mask_pl,
                                             double *ch_b, double *ch_b_pl);

/* vertical_limiter_thuburn that also records its branch mask: the same
 * arguments and results as nicam_vertical_limiter_thuburn, plus one bit per
 * limited face (levels kmin+1..kmax), set where the limited value differs
 * from q_h's input.  Layout: uint32 word ((l*nword + p/32)*gall + g), bit
 * p%32, p = k-(kmin+1), nword = ceil((kall-3)/32); mask_pl the same over
 * gall_pl and the 2 pole regions. */
int nicam_vertical_limiter_thuburn_mask(nicam_ctx ctx,
                                        double *q_h, double *q_h_pl,
                                        const double *q, const double *q_pl,
                                        const double *d, const double *d_pl,
                                        const double *ck, const double *ck_pl,
                                        uint32_t *mask, uint32_t *mask_pl);

/* Adjoint of vertical_limiter_thuburn with respect to q_h's input, q and ck,
 * at fixed d and inflow flags (dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243):
 *   inout q_h_b  in: cotangent of the limited q_h; out: cotangent of q_h's
 *                input (faces outside kmin+1..kmax pass through)
 *   out   q_b    cotangent of q (every level written)
 *   in    q_h    the forward's output; q, d, ck its inputs; mask its mask
 *   out   ck_b   cotangent of ck(g,k,l,2) (NULL: not formed; with poles
 *                ck_b_pl too; every level written): the bounds of level k
 *                depend on ck(k,1) and ck(k,2) through Cin, Cout and CQin
 * A clipped face takes the lower bound when its value equals Qout_min of the
 * upwind level, else the upper bound. */
int nicam_vertical_limiter_thuburn_adjoint(nicam_ctx ctx,
                                           double *q_h_b, double *q_h_b_pl,
                                           double *q_b, double *q_b_pl,
                                           const double *q_h, const double *q_h_pl,
                                           const double *q, const double *q_pl,
                                           const double *d, const double *d_pl,
                                           const double *ck, const double *ck_pl,
                                           const uint32_t *mask, const uint32_t *mask_pl,
                                           double *ck_b, double *ck_b_pl);

/* ---- FP64 global tracer mass and the update that conserves it ----
 * Neither has a counterpart in the reference, whose kernels stop at the face
 * values; they complete the flux-form step the reference's kernels feed
 * (flx_h of dyn_horiz_adv_flux/src/mod_src_tracer.f90:234-309, the limited
 * q_a of dyn_horiz_adv_limiter/src/mod_src_tracer.f90:377-410). */

/* mass[k] (kall doubles, device or host per the pointer mode) = the sum over
 * the inner cells of every region of GMTR_p(AREA) * rhogq(k), plus both pole
 * centres' AREA * rhogq_pl(1,k).  Products and sums in FP64, in a fixed order
 * (the same result on every run).  The first call allocates a kall*lall
 * partial-sum buffer that the context keeps.  The _f32 form reads FP32
 * rhogq and widens it exactly. */
int nicam_tracer_mass(nicam_ctx ctx, const double *rhogq, const double *rhogq_pl, double *mass);
int nicam_tracer_mass_f32(nicam_ctx ctx, const float *rhogq, const float *rhogq_pl, double *mass);
/* nicam_tracer_mass with no allocation at all: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture from its first call, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums live in a static device array of
 * 65536 doubles per GPU (kall * lall must not exceed 65536, else
 * NICAM_ERR_ARG), whose address nicam_ctx_create looks up.  Device pointers
 * only; the result equals nicam_tracer_mass's bit for bit.  Calls that overlap
 * on one GPU share the array and race; calls on one stream are ordered. */
int nicam_tracer_mass_noalloc(nicam_ctx ctx, const double *rhogq, const double *rhogq_pl, double *mass);

/* rhogq -= RAREA * (sum of the six outward face transfers), in FP64, on the
 * inner cells and the pole centres (halos untouched).  Each arc's transfer
 * AREA(o)*flx_h(o,a)*q_a(o,a) is taken from its origin o (a = 1..3) and used
 * with opposite signs by its two cells; a face two regions share takes the
 * lower-indexed region's transfer, and the pole centres take the transfers of
 * the region faces that point at them, so the global sum of AREA*rhogq is
 * conserved to the rounding of the FP64 operations.  flx_h and q_a are
 * (gall,kall,lall,6); only faces 1-3 are read, and the pole arrays are not
 * read.  Needs nicam_comm_setup (all regions on this GPU).  The _f32 form
 * takes FP32 flx_h and q_a (the mixed step) and accumulates in FP64. */
int nicam_tracer_update_conservative(nicam_ctx ctx, double *rhogq, double *rhogq_pl,
                                     const double *flx_h, const double *q_a);
int nicam_tracer_update_conservative_f32(nicam_ctx ctx, double *rhogq, double *rhogq_pl,
                                         const float *flx_h, const float *q_a);

/* nicam_tracer_update_vertical with FP32 flx_v and q_h and an FP64 rhogq:
 * each half level's transfer flx_v*q_h is formed once in FP64 and used by the
 * levels on both sides, so column mass is conserved to FP64 rounding. */
int nicam_tracer_update_vertical_f32(nicam_ctx ctx, double *rhogq, double *rhogq_pl,
                                     const float *flx_v, const float *flx_v_pl,
                                     const float *q_h, const float *q_h_pl);

/* ---- many tracers, several grids: the batched tracer step ----
 * Batches the reference's per-call loop (dyn_horiz_adv_limiter/src/main.f90:208-214
 * calls horizontal_limiter_thuburn for one tracer at a time): the limiters
 * take ntr tracers in NICAM's tracer-last layout and read the fields the
 * tracers share (d, ch, cmask; d, ck) once per cell-level. */

/* horizontal_limiter_thuburn of ntr tracers: q_a(gall,kall,lall,6,ntr),
 * q(gall,kall,lall,ntr), Qout_prev and Qout_post (gall,kall,lall,2,ntr), the
 * pole arrays with the same trailing ntr dimension; d, ch, cmask as in the
 * single-tracer call.  Each tracer's result is the single-tracer call's, bit
 * for bit.  Qout_post = NULL exchanges all tracers' Qout on the device.
 * Device pointers only. */
int nicam_horizontal_limiter_thuburn_batch(nicam_ctx ctx, int ntr,
                                           double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                                           const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                                           const double *cmask, const double *cmask_pl,
                                           double *Qout_prev, double *Qout_prev_pl,
                                           const double *Qout_post, const double *Qout_post_pl);
int nicam_horizontal_limiter_thuburn_batch_f32(nicam_ctx ctx, int ntr,
                                               float *q_a, float *q_a_pl, const float *q, const float *q_pl,
                                               const float *d, const float *d_pl, const float *ch, const float *ch_pl,
                                               const float *cmask, const float *cmask_pl,
                                               float *Qout_prev, float *Qout_prev_pl,
                                               const float *Qout_post, const float *Qout_post_pl);
/* vertical_limiter_thuburn of ntr tracers: q_h and q (gall,kall,lall,ntr),
 * pole arrays likewise; d and ck shared.  Device pointers only. */
int nicam_vertical_limiter_thuburn_batch(nicam_ctx ctx, int ntr, double *q_h, double *q_h_pl,
                                         const double *q, const double *q_pl, const double *d, const double *d_pl,
                                         const double *ck, const double *ck_pl);
int nicam_vertical_limiter_thuburn_batch_f32(nicam_ctx ctx, int ntr, float *q_h, float *q_h_pl,
                                             const float *q, const float *q_pl, const float *d, const float *d_pl,
                                             const float *ck, const float *ck_pl);

/* The arrays of one grid's mixed-precision tracer step (device pointers):
 * the FP32 flux, the FP32 limiters of ntr tracers (layouts as in the _batch
 * calls; qv/dv are the vertical limiter's q and d) and the mixed rho*w solve
 * (FP32 state and sources, FP64 rhogw). */
typedef struct nicam_tracer_step_f32 {
    float *flx_h, *flx_h_pl, *GRD_xc, *GRD_xc_pl;
    const float *rho, *rho_pl, *rhovx, *rhovx_pl, *rhovy, *rhovy_pl, *rhovz, *rhovz_pl;
    float *q_a, *q_a_pl;
    const float *q, *q_pl, *d, *d_pl, *ch, *ch_pl, *cmask, *cmask_pl;
    float *Qout_prev, *Qout_prev_pl;
    const float *Qout_post, *Qout_post_pl;
    float *q_h, *q_h_pl;
    const float *qv, *qv_pl, *dv, *dv_pl, *ck, *ck_pl;
    double *rhogw, *rhogw_pl;
    const float *rhogw0, *rhogw0_pl, *preg0, *preg0_pl, *rhog0, *rhog0_pl;
    const float *Srho, *Srho_pl, *Sw, *Sw_pl, *Spre, *Spre_pl;
} nicam_tracer_step_f32;

/* A group of 1..8 grids (contexts on the current device) advanced together:
 * each kernel of the step runs once over every grid and tracer.  Launches
 * go on the group's stream (set one; the null stream cannot be captured). */
typedef struct nicam_group_s *nicam_group;
int nicam_group_create(nicam_group *grp, int n, const nicam_ctx *ctxs);
int nicam_group_destroy(nicam_group grp);
int nicam_group_set_stream(nicam_group grp, void *hip_stream);
/* One step on every grid: horizontal_flux_f32, the horizontal and vertical
 * limiters of ntr tracers, vi_rhow_solver_mixed -- four launches in all.
 * use_graph != 0: the first call captures the four launches into a HIP graph
 * that later calls with the same arguments replay (different arguments
 * recapture).  Each grid's outputs equal its single-grid calls bit for bit. */
int nicam_group_tracer_step_f32(nicam_group grp, int ntr, const nicam_tracer_step_f32 *grids,
                                float dt_large, double dt_small, int use_graph);
/* Picks each grid's launch shapes (levels per work-group of the flux, levels
 * per work-item of the vertical limiter, the rho*w elimination variant) by
 * timing the candidates on the given arrays (their outputs are overwritten),
 * and sets them with nicam_ctx_set_tuning.  With cache_path, the choices are
 * read from that text file when it holds this device and shape, else timed
 * and appended to it.  n_timed (may be NULL) returns the candidates timed. */
int nicam_group_autotune(nicam_group grp, int ntr, const nicam_tracer_step_f32 *grids, float dt_large,
                         double dt_small, const char *cache_path, int *n_timed);

#ifdef __cplusplus
}
#endif
#endif /* NICAM_TT_H */
