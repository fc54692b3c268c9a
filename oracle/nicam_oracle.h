/*
 * CPU oracle for the NICAM tracer-transport + vertical-implicit hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran kernels, used by the test suite, the package's smoke check and the
 * CPU-baseline leg of the benchmark as the *checker*.  Nothing in the product
 * library links, calls or falls back to it.
 *
 * Array layout follows the reference exactly (Fortran column-major):
 *   regular field  x(gall, kall, lall[, n])  -> x[((n*lall + l)*kall + k)*gall + g]
 *   pole field     x(gall_pl, kall, lall_pl[, n])
 *   metric GMTR_t  (gall, 1, lall, 2, 5)     -> [((w*2 + t)*lall + l)*gall + g]
 *   metric GMTR_a  (gall, 1, lall, 3, 12)    -> [((c*3 + a)*lall + l)*gall + g]
 *   metric GMTR_p  (gall, 1, lall, 8)        -> [(c*lall + l)*gall + g]
 *   GRD_xr         (gall, 1, lall, 3, 3)     -> [((d*3 + a)*lall + l)*gall + g]
 *   pole metrics   GMTR_t_pl(gall_pl,1,2,5), GMTR_a_pl(gall_pl,1,2,18),
 *                  GMTR_p_pl(gall_pl,1,2,8), GRD_xr_pl(gall_pl,1,2,3)
 * Indices in the restated loops are kept 1-based, as in the reference, and
 * converted to 0-based offsets only inside the accessor macros.
 *
 * The tangent-linear and adjoint functions (orc_*_tlm, orc_*_adjoint) have no
 * counterpart in the reference.  They are pinned by what the mathematics
 * fixes, in tests/test_adjoint.py:
 *   - <TLM v, u> = <v, ADJ u> (and, for the linear solve, against the forward
 *     solver itself);
 *   - central differences of the forward kernels, which are themselves pinned
 *     to the compiled reference.
 */
#ifndef NICAM_ORACLE_H
#include <stdint.h>

#define NICAM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int gall_1d;   /* ADM_gall_1d                                   */
    int kall;      /* ADM_kall                                      */
    int lall;      /* ADM_lall                                      */
    int gall_pl;   /* ADM_gall_pl (6 on an icosahedral grid)         */
    int have_pl;   /* ADM_have_pl                                   */
    double eps;    /* CONST_EPS                                     */
    double big;    /* CONST_HUGE                                    */
} orc_dims;

/* VI/src/mod_misc.f90:447-461 (CONST_setup): EPS = epsilon, HUGE = huge. */
void orc_const_setup(double *eps, double *huge_val);

/* VI/src/mod_misc.f90:464-524 (GRD_setup): vertical factors from gz/gzh. */
void orc_grd_setup(int kall, const double *gz, const double *gzh,
                   double *dgz, double *dgzh, double *rdgz, double *rdgzh,
                   double *afact, double *bfact, double *cfact, double *dfact);

/* FLX/src/mod_src_tracer.f90:42-379 */
void orc_horizontal_flux(const orc_dims *D, const int *have_sgp,
                         double *flx_h, double *flx_h_pl,
                         double *GRD_xc, double *GRD_xc_pl,
                         const double *rho, const double *rho_pl,
                         const double *rhovx, const double *rhovx_pl,
                         const double *rhovy, const double *rhovy_pl,
                         const double *rhovz, const double *rhovz_pl,
                         double dt,
                         const double *GRD_xr, const double *GRD_xr_pl,
                         const double *GMTR_p, const double *GMTR_p_pl,
                         const double *GMTR_t, const double *GMTR_t_pl,
                         const double *GMTR_a, const double *GMTR_a_pl);

/* HLIM/src/mod_src_tracer.f90:42-430 */
void orc_horizontal_limiter_thuburn(const orc_dims *D, const int *have_sgp,
                                    double *q_a, double *q_a_pl,
                                    const double *q, const double *q_pl,
                                    const double *d, const double *d_pl,
                                    const double *ch, const double *ch_pl,
                                    const double *cmask, const double *cmask_pl,
                                    double *Qout_prev, double *Qout_prev_pl,
                                    const double *Qout_post, const double *Qout_post_pl);

/* VLIM/src/mod_src_tracer.f90:41-243 */
void orc_vertical_limiter_thuburn(const orc_dims *D,
                                  double *q_h, double *q_h_pl,
                                  const double *q, const double *q_pl,
                                  const double *d, const double *d_pl,
                                  const double *ck, const double *ck_pl);

/* VI/src/mod_vi.f90:93-306; the coefficients Mc/Ml/Mu that the reference
 * keeps as module state after vi_rhow_solver_putcoef (:58-89) are passed in. */
void orc_vi_rhow_solver(const orc_dims *D,
                        double *rhogw, double *rhogw_pl,
                        const double *rhogw0, const double *rhogw0_pl,
                        const double *preg0, const double *preg0_pl,
                        const double *rhog0, const double *rhog0_pl,
                        const double *Srho, const double *Srho_pl,
                        const double *Sw, const double *Sw_pl,
                        const double *Spre, const double *Spre_pl,
                        double dt,
                        const double *Mc, const double *Mc_pl,
                        const double *Ml, const double *Ml_pl,
                        const double *Mu, const double *Mu_pl,
                        const double *RGAM, const double *RGAM_pl,
                        const double *RGAMH, const double *RGAMH_pl,
                        const double *RGSGAM2, const double *RGSGAM2_pl,
                        const double *RGSGAM2H, const double *RGSGAM2H_pl,
                        const double *GSGAM2H, const double *GSGAM2H_pl,
                        const double *rdgzh, const double *afact, const double *bfact,
                        double grav, double rdry, double cvdry, int non_hydro_alpha);

/* dyn_diffusion/src/mod_oprt.f90:41-296 (OPRT_diffusion).
 *   coef_intp(gall,3,3,2,lall)  coef_intp_pl(gall_pl,3,3,2)
 *   coef_diff(gall,6,3,lall)    coef_diff_pl(5,3,2) */
void orc_oprt_diffusion(const orc_dims *D, const int *have_sgp,
                        double *dscl, double *dscl_pl,
                        const double *scl, const double *scl_pl,
                        const double *kh, const double *kh_pl,
                        const double *coef_intp, const double *coef_intp_pl,
                        const double *coef_diff, const double *coef_diff_pl);

/* dyn_divdamp/src/mod_oprt3d.f90:41-402 (OPRT3D_divdamp).  The module
 * arrays it reads are passed in: VMTR_RGSQRTH, VMTR_RGAM, VMTR_RGAMH
 * (gall,kall,lall) and _pl, VMTR_C2WfactGz(gall,kall,6,lall) and
 * _pl(gall_pl,kall,6,2), GRD_rdgz(kall). */
void orc_oprt3d_divdamp(const orc_dims *D, const int *have_sgp,
                        double *ddivdx, double *ddivdx_pl,
                        double *ddivdy, double *ddivdy_pl,
                        double *ddivdz, double *ddivdz_pl,
                        const double *rhogvx, const double *rhogvx_pl,
                        const double *rhogvy, const double *rhogvy_pl,
                        const double *rhogvz, const double *rhogvz_pl,
                        const double *rhogw, const double *rhogw_pl,
                        const double *coef_intp, const double *coef_intp_pl,
                        const double *coef_diff, const double *coef_diff_pl,
                        const double *RGSQRTH, const double *RGSQRTH_pl,
                        const double *RGAM, const double *RGAM_pl,
                        const double *RGAMH, const double *RGAMH_pl,
                        const double *C2WfactGz, const double *C2WfactGz_pl,
                        const double *rdgz);

/* Number of OpenMP threads the oracle runs with (1 when built without it). */
/* Step glue (not in the reference; the NICAM forms of the Miura upwind
 * scheme, docs/NICAM/src/20_intro_nicam.tex:368-438):
 *   orc_oprt_gradient: grad(gall,kall,lall,3) = the 7-point stencil of
 *     coef_grad(gall,0:6,3,lall) (ij, ip1j, ip1jp1, ijp1, im1j, im1jm1, ijm1)
 *     on g in [gmin, gmax] of the flat index, zero elsewhere; the pole centre
 *     sums coef_grad_pl(0:gall_pl-1,3,2) over the centre and its ring.
 *   orc_tracer_courant: ch = flx_h/rhog, cmask = 0.5 - sign(0.5, ch - EPS)
 *     (1 where the face is an inflow); the pole ring against the centre.
 *   orc_horizontal_remap: for every arc AI/AIJ/AJ with origin in
 *     [gmin-1, gmax]^2, the face value is the upwind node's q plus its
 *     gradient dotted with (GRD_xc - GRD_x(node)), the upwind side chosen by
 *     the mask of the face's owner: face a of the origin by cmask(origin,a),
 *     face a+3 of the neighbour by cmask(neighbour,a+3); faces no arc
 *     reaches are kept.
 *   orc_tracer_update: rhogq -= sum over the six faces of flx_h*q_a, on the
 *     inner cells and the pole centres. */
void orc_oprt_gradient(const orc_dims *D, double *grad, double *grad_pl, const double *scl, const double *scl_pl,
                       const double *coef_grad, const double *coef_grad_pl);
void orc_tracer_courant(const orc_dims *D, double *ch, double *ch_pl, double *cmask, double *cmask_pl,
                        const double *flx_h, const double *flx_h_pl, const double *rhog, const double *rhog_pl);
void orc_horizontal_remap(const orc_dims *D, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                          const double *grad, const double *grad_pl, const double *cmask, const double *cmask_pl,
                          const double *GRD_xc, const double *GRD_xc_pl, const double *GRD_x,
                          const double *GRD_x_pl);
void orc_tracer_update(const orc_dims *D, double *rhogq, double *rhogq_pl, const double *flx_h,
                       const double *flx_h_pl, const double *q_a, const double *q_a_pl);
/* Adjoints of the four above (the transposes of their derivatives; cmask
 * fixed), for the differentiable step; outputs overwritten. */
void orc_oprt_gradient_adjoint(const orc_dims *D, double *scl_b, double *scl_b_pl, const double *grad_b,
                               const double *grad_b_pl, const double *coef_grad, const double *coef_grad_pl);
void orc_tracer_courant_adjoint(const orc_dims *D, double *flx_b, double *flx_b_pl, double *rhog_b, double *rhog_b_pl,
                                const double *ch_b, const double *ch_b_pl, const double *flx_h, const double *flx_h_pl,
                                const double *rhog, const double *rhog_pl);
void orc_horizontal_remap_adjoint(const orc_dims *D, double *q_b, double *q_b_pl, double *grad_b, double *grad_b_pl,
                                  double *xc_b, double *xc_b_pl, const double *qa_b, const double *qa_b_pl,
                                  const double *grad, const double *grad_pl, const double *cmask,
                                  const double *cmask_pl, const double *GRD_xc, const double *GRD_xc_pl,
                                  const double *GRD_x, const double *GRD_x_pl);
void orc_tracer_update_adjoint(const orc_dims *D, double *flx_b, double *flx_b_pl, double *qa_b, double *qa_b_pl,
                               const double *rhogq_b, const double *rhogq_b_pl, const double *flx_h,
                               const double *flx_h_pl, const double *q_a, const double *q_a_pl);

int orc_num_threads(void);

/* Vertical step glue (not in the reference), kmin = 2, kmax = kall-1:
 *   orc_vertical_mass_flux: flx_v = rhogw_vm*dt on kmin+1..kmax (rhogw_vm as
 *     dyn_divdamp/src/mod_oprt3d.f90:148-155 forms it), 0 elsewhere;
 *   orc_vertical_courant: ck(k,1) = -flx_v(k)/rhog(k)*rdgz(k),
 *     ck(k,2) = flx_v(k+1)/rhog(k)*rdgz(k) on the interior faces of
 *     kmin..kmax, 0 elsewhere, +0 in ck(kmin,1) and -0 in ck(kmax+1,1);
 *   orc_vertical_face: q_h(k) = afact(k) q(k) + bfact(k) q(k-1) on kmin+1..kmax;
 *   orc_tracer_update_vertical: rhogq(k) -= (flx_v(k+1) q_h(k+1) - flx_v(k) q_h(k)) rdgz(k)
 *     on kmin..kmax, the faces kmin and kmax+1 carrying nothing. */
void orc_vertical_mass_flux(const orc_dims *D, double *flx_v, double *flx_v_pl, const double *vx, const double *vx_pl,
                            const double *vy, const double *vy_pl, const double *vz, const double *vz_pl,
                            const double *w, const double *w_pl, const double *C2WfactGz,
                            const double *C2WfactGz_pl, const double *RGAMH, const double *RGAMH_pl,
                            const double *RGSQRTH, const double *RGSQRTH_pl, double dt);
void orc_vertical_courant(const orc_dims *D, const double *rdgz, double *ck, double *ck_pl, const double *flx_v,
                          const double *flx_v_pl, const double *rhog, const double *rhog_pl);
void orc_vertical_face(const orc_dims *D, const double *afact, const double *bfact, double *q_h, double *q_h_pl,
                       const double *q, const double *q_pl);
void orc_tracer_update_vertical(const orc_dims *D, const double *rdgz, double *rhogq, double *rhogq_pl,
                                const double *flx_v, const double *flx_v_pl, const double *q_h, const double *q_h_pl);
/* Adjoints of the three above (outputs overwritten), for the differentiable step. */
void orc_vertical_courant_adjoint(const orc_dims *D, const double *rdgz, double *flx_b, double *flx_b_pl,
                                  double *rhog_b, double *rhog_b_pl, const double *ck_b, const double *ck_b_pl,
                                  const double *flx_v, const double *flx_v_pl, const double *rhog,
                                  const double *rhog_pl);
void orc_vertical_face_adjoint(const orc_dims *D, const double *afact, const double *bfact, double *q_b,
                               double *q_b_pl, const double *q_h_b, const double *q_h_b_pl);
void orc_tracer_update_vertical_adjoint(const orc_dims *D, const double *rdgz, double *flx_b, double *flx_b_pl,
                                        double *q_h_b, double *q_h_b_pl, const double *rq_b, const double *rq_b_pl,
                                        const double *flx_v, const double *flx_v_pl, const double *q_h,
                                        const double *q_h_pl);

/* Mc/Ml/Mu of the rho*w system from h (eth) and g~ at the half levels, the
 * VMTR metrics and the GRD factors (no reference code). */
void orc_vi_rhow_update_matrix(const orc_dims *D, double *Mc, double *Mc_pl, double *Ml, double *Ml_pl, double *Mu,
                               double *Mu_pl, const double *eth, const double *eth_pl, const double *g_tilde,
                               const double *g_tilde_pl, const double *RGSGAM2, const double *RGSGAM2_pl,
                               const double *RGAMH, const double *RGAMH_pl, const double *GSGAM2H,
                               const double *GSGAM2H_pl, const double *rdgz, const double *rdgzh, const double *cfact,
                               const double *dfact, double dt, double grav, double rdry, double cvdry, int alpha_i);

/* Adjoint (transpose) of orc_vi_rhow_solver's linear map, coefficients fixed:
 * wb in = output cotangent, out = input cotangent; the six others written. */
void orc_vi_rhow_solver_adjoint(const orc_dims *D, double *wb, double *wb_pl, double *w0b, double *w0b_pl,
                                double *p0b, double *p0b_pl, double *r0b, double *r0b_pl, double *Srb,
                                double *Srb_pl, double *Swb, double *Swb_pl, double *Spb, double *Spb_pl, double dt,
                                const double *Mc, const double *Mc_pl, const double *Ml, const double *Ml_pl,
                                const double *Mu, const double *Mu_pl, const double *RGAM, const double *RGAM_pl,
                                const double *RGAMH, const double *RGAMH_pl, const double *RGSGAM2,
                                const double *RGSGAM2_pl, const double *RGSGAM2H, const double *RGSGAM2H_pl,
                                const double *GSGAM2H, const double *GSGAM2H_pl, const double *rdgzh,
                                const double *afact, const double *bfact, double grav, double rdry, double cvdry,
                                int non_hydro_alpha);

/* Tangent-linear and adjoint of orc_vertical_limiter_thuburn at fixed d (and
 * the signs of ck).  mask: one bit per limited face (set where the forward clipped it),
 * word ((l*nword + p/32)*gall + g), bit p%32, p = k-(kmin+1), nword =
 * ceil((kall-3)/32); q_h is the forward's output.  TLM: dqh from dqh_in, dq
 * and dck (NULL = zero).  Adjoint: qhb in = output cotangent, out =
 * cotangent of q_h's input; qb and ck_b (if not NULL, both planes) written
 * on every level. */
void orc_vertical_limiter_thuburn_tlm(const orc_dims *D, double *dqh, double *dqh_pl, const double *dqh_in,
                                      const double *dqh_in_pl, const double *dq, const double *dq_pl,
                                      const double *q_h, const double *q_h_pl, const double *q, const double *q_pl,
                                      const double *d, const double *d_pl, const double *ck, const double *ck_pl,
                                      const uint32_t *mask, const uint32_t *mask_pl, const double *dck,
                                      const double *dck_pl);
void orc_vertical_limiter_thuburn_adjoint(const orc_dims *D, double *qhb, double *qhb_pl, double *qb, double *qb_pl,
                                          const double *q_h, const double *q_h_pl, const double *q,
                                          const double *q_pl, const double *d, const double *d_pl, const double *ck,
                                          const double *ck_pl, const uint32_t *mask, const uint32_t *mask_pl,
                                          double *ck_b, double *ck_b_pl);

/* Tangent-linear and adjoint of orc_horizontal_limiter_thuburn with
 * Qout_post supplied, at fixed d and cmask.  mask: uint8 per cell-level,
 * bit m-1 set where face m (1-3) of a cell in [gmin-1,gmax]^2 changed; pole:
 * uint8 per point, bit 0 for ring faces.  q_a is the forward's output.
 * TLM: dq_a in = dq_a_in, out = dq_a; dQprev written from dq, dQpost and
 * dch (the Courant numbers' direction; NULL = zero).
 * Adjoint: q_a_b in/out; q_b, Qpost_b and ch_b (if not NULL; every plane and
 * level written) from q_a_b and Qprev_b.  The derivative of Qout in ch at
 * ch = 0 is the right-hand one. */
void orc_horizontal_limiter_thuburn_tlm(const orc_dims *D, const int *have_sgp, double *dq_a, double *dq_a_pl,
                                        const double *dq, const double *dq_pl, double *dQprev, double *dQprev_pl,
                                        const double *dQpost, const double *dQpost_pl, const double *q_a,
                                        const double *q_a_pl, const double *q, const double *q_pl, const double *d,
                                        const double *d_pl, const double *ch, const double *ch_pl,
                                        const double *cmask, const double *cmask_pl, const double *Qpost,
                                        const double *Qpost_pl, const uint8_t *mask, const uint8_t *mask_pl,
                                        const double *dch, const double *dch_pl);
void orc_horizontal_limiter_thuburn_adjoint(const orc_dims *D, const int *have_sgp, double *q_a_b,
                                            double *q_a_b_pl, double *q_b, double *q_b_pl, const double *Qprev_b,
                                            const double *Qprev_b_pl, double *Qpost_b, double *Qpost_b_pl,
                                            const double *q_a, const double *q_a_pl, const double *q,
                                            const double *q_pl, const double *d, const double *d_pl,
                                            const double *ch, const double *ch_pl, const double *cmask,
                                            const double *cmask_pl, const double *Qpost, const double *Qpost_pl,
                                            const uint8_t *mask, const uint8_t *mask_pl, double *ch_b,
                                            double *ch_b_pl);

/* Tangent-linear and adjoint of orc_horizontal_flux in (rho, rhovx, rhovy,
 * rhovz), metrics and dt fixed.  TLM: dflx (6 faces), dxc (3 arcs x 3)
 * written.  Adjoint: rho_b .. rhovz_b written from flx_h_b, GRD_xc_b. */
void orc_horizontal_flux_tlm(const orc_dims *D, const int *have_sgp, double *dflx, double *dflx_pl, double *dxc,
                             double *dxc_pl, const double *rho, const double *rho_pl, const double *rhovx,
                             const double *rhovx_pl, const double *rhovy, const double *rhovy_pl,
                             const double *rhovz, const double *rhovz_pl, const double *drho, const double *drho_pl,
                             const double *drhovx, const double *drhovx_pl, const double *drhovy,
                             const double *drhovy_pl, const double *drhovz, const double *drhovz_pl, double dt,
                             const double *GMTR_t, const double *GMTR_t_pl, const double *GMTR_a,
                             const double *GMTR_a_pl, const double *GMTR_p, const double *GMTR_p_pl);
void orc_horizontal_flux_adjoint(const orc_dims *D, const int *have_sgp, const double *flx_h_b,
                                 const double *flx_h_b_pl, const double *GRD_xc_b, const double *GRD_xc_b_pl,
                                 double *rho_b, double *rho_b_pl, double *rhovx_b, double *rhovx_b_pl,
                                 double *rhovy_b, double *rhovy_b_pl, double *rhovz_b, double *rhovz_b_pl,
                                 const double *rho, const double *rho_pl, const double *rhovx, const double *rhovx_pl,
                                 const double *rhovy, const double *rhovy_pl, const double *rhovz,
                                 const double *rhovz_pl, double dt, const double *GMTR_t, const double *GMTR_t_pl,
                                 const double *GMTR_a, const double *GMTR_a_pl, const double *GMTR_p,
                                 const double *GMTR_p_pl);

#ifdef __cplusplus
}
#endif
#endif
