/*
 * CPU oracle: plain-C restatement of the NICAM hot-path kernels.
 *
 * TEST INFRASTRUCTURE ONLY (checker for tests, smoke and the CPU baseline);
 * the product library never links or calls this file.
 *
 * Every loop follows the reference's loop bounds, write set and floating-point
 * evaluation order (left-to-right, no re-association).  Build with
 * -ffp-contract=off so that no multiply-add is fused: the reference builds
 * for x86-64 without -march, where no FMA instruction exists.
 *
 * Paths below are relative to the reference tree:
 *   FLX  = kernels/NICAM/dyn_horiz_adv_flux/src
 *   HLIM = kernels/NICAM/dyn_horiz_adv_limiter/src
 *   VLIM = kernels/NICAM/dyn_vert_adv_limiter/src
 *   VI   = kernels/NICAM/dyn_vi_rhow_solver/src
 */
#include "nicam_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LALL_PL 2
#define GSLF_PL 1
#define GMIN_PL 2
#define K0 1

/* Fortran MIN/MAX of two reals, first argument kept on ties. */
static inline double mn(double a, double b) { return (b < a) ? b : a; }
static inline double mx(double a, double b) { return (b > a) ? b : a; }
/* Fortran SIGN(a, b) = |a| carrying the sign bit of b (so SIGN(0.5,-0.0) = -0.5). */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* VI/src/mod_misc.f90:447-461 */
void orc_const_setup(double *eps, double *huge_val)
{
    *eps = DBL_EPSILON;
    *huge_val = DBL_MAX;
}

/* VI/src/mod_misc.f90:464-524.  Arrays are 1-based in the comments, 0-based here. */
void orc_grd_setup(int kall, const double *gz, const double *gzh,
                   double *dgz, double *dgzh, double *rdgz, double *rdgzh,
                   double *afact, double *bfact, double *cfact, double *dfact)
{
    const int kmin = 2, kmax = kall - 1;
#define V1(a, k) a[(k) - 1]
    for (int k = kmin - 1; k <= kmax; ++k)
        V1(dgz, k) = V1(gzh, k + 1) - V1(gzh, k);
    V1(dgz, kmax + 1) = V1(dgz, kmax);

    for (int k = kmin; k <= kmax + 1; ++k)
        V1(dgzh, k) = V1(gz, k) - V1(gz, k - 1);
    V1(dgzh, kmin - 1) = V1(dgzh, kmin);

    for (int k = 1; k <= kall; ++k) {
        V1(rdgz, k) = 1.0 / V1(dgz, k);
        V1(rdgzh, k) = 1.0 / V1(dgzh, k);
    }

    for (int k = kmin; k <= kmax + 1; ++k)
        V1(afact, k) = (V1(gzh, k) - V1(gz, k - 1)) / (V1(gz, k) - V1(gz, k - 1));
    V1(afact, kmin - 1) = 1.0;
    for (int k = 1; k <= kall; ++k) V1(bfact, k) = 1.0 - V1(afact, k);

    for (int k = kmin; k <= kmax; ++k)
        V1(cfact, k) = (V1(gz, k) - V1(gzh, k)) / (V1(gzh, k + 1) - V1(gzh, k));
    V1(cfact, kmin - 1) = 1.0;
    V1(cfact, kmax + 1) = 0.0;
    for (int k = 1; k <= kall; ++k) V1(dfact, k) = 1.0 - V1(cfact, k);
#undef V1
}

/* ---------------------------------------------------------------------------
 * Accessors (1-based indices, Fortran order)
 * ------------------------------------------------------------------------- */
#define F3(x, g, k, l) (x)[((size_t)((l) - 1) * kall + ((k) - 1)) * gall + ((g) - 1)]
#define F4(x, g, k, l, n) \
    (x)[(((size_t)((n) - 1) * lall + ((l) - 1)) * kall + ((k) - 1)) * gall + ((g) - 1)]
#define F5(x, g, k, l, a, n)                                                              \
    (x)[((((size_t)((n) - 1) * 3 + ((a) - 1)) * lall + ((l) - 1)) * kall + ((k) - 1)) * \
            gall + ((g) - 1)]
#define P3(x, v, k, l) (x)[((size_t)((l) - 1) * kall + ((k) - 1)) * gall_pl + ((v) - 1)]
#define P4(x, v, k, l, n) \
    (x)[(((size_t)((n) - 1) * LALL_PL + ((l) - 1)) * kall + ((k) - 1)) * gall_pl + ((v) - 1)]
/* metrics: K0 dimension has extent 1 */
#define MT(g, l, t, w) GMTR_t[(((size_t)((w) - 1) * 2 + ((t) - 1)) * lall + ((l) - 1)) * gall + ((g) - 1)]
#define MA(g, l, a, c) GMTR_a[(((size_t)((c) - 1) * 3 + ((a) - 1)) * lall + ((l) - 1)) * gall + ((g) - 1)]
#define MP(g, l, c) GMTR_p[((size_t)((c) - 1) * lall + ((l) - 1)) * gall + ((g) - 1)]
#define XR(g, l, a, d) GRD_xr[(((size_t)((d) - 1) * 3 + ((a) - 1)) * lall + ((l) - 1)) * gall + ((g) - 1)]
#define MT_PL(v, l, w) GMTR_t_pl[((size_t)((w) - 1) * LALL_PL + ((l) - 1)) * gall_pl + ((v) - 1)]
#define MA_PL(v, l, c) GMTR_a_pl[((size_t)((c) - 1) * LALL_PL + ((l) - 1)) * gall_pl + ((v) - 1)]
#define MP_PL(v, l, c) GMTR_p_pl[((size_t)((c) - 1) * LALL_PL + ((l) - 1)) * gall_pl + ((v) - 1)]
#define XR_PL(v, l, d) GRD_xr_pl[((size_t)((d) - 1) * LALL_PL + ((l) - 1)) * gall_pl + ((v) - 1)]

enum { TI = 1, TJ = 2 };
enum { AI = 1, AIJ = 2, AJ = 3 };
enum { W1 = 3, W2 = 4, W3 = 5 };
enum { HNX = 1, HNY = 2, HNZ = 3 };
enum { P_RAREA = 2 };
enum { XDIR = 1, YDIR = 2, ZDIR = 3 };
enum { I_MIN = 1, I_MAX = 2 };

/* ===========================================================================
 * horizontal_flux: FLX/mod_src_tracer.f90:42-379
 * ======================================================================== */
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
                         const double *GMTR_a, const double *GMTR_a_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall;
    const int gall_pl = D->gall_pl, gmax_pl = gall_pl;
    const int gmin = 2, gmax = iall - 1;
    const double EPS = D->eps;

    for (int l = 1; l <= lall; ++l) {
        /* FLX:130-137 -- the reference parallelises inside the k loop; here the
         * k iterations are independent, so each thread owns whole k slabs. */
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= kall; ++k) {
            double *rhot_TI = (double *)malloc(sizeof(double) * 8 * gall);
            double *rhot_TJ = rhot_TI + gall, *rhovxt_TI = rhot_TI + 2 * gall,
                   *rhovxt_TJ = rhot_TI + 3 * gall, *rhovyt_TI = rhot_TI + 4 * gall,
                   *rhovyt_TJ = rhot_TI + 5 * gall, *rhovzt_TI = rhot_TI + 6 * gall,
                   *rhovzt_TJ = rhot_TI + 7 * gall;
#define T1(a, g) (a)[(g) - 1]
            /* FLX:139-161  (i,j),(i+1,j) */
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ip1j = ij + 1;
                    T1(rhot_TI, ij) = F3(rho, ij, k, l) * MT(ij, l, TI, W1) + F3(rho, ip1j, k, l) * MT(ij, l, TI, W2);
                    T1(rhovxt_TI, ij) = F3(rhovx, ij, k, l) * MT(ij, l, TI, W1) + F3(rhovx, ip1j, k, l) * MT(ij, l, TI, W2);
                    T1(rhovyt_TI, ij) = F3(rhovy, ij, k, l) * MT(ij, l, TI, W1) + F3(rhovy, ip1j, k, l) * MT(ij, l, TI, W2);
                    T1(rhovzt_TI, ij) = F3(rhovz, ij, k, l) * MT(ij, l, TI, W1) + F3(rhovz, ip1j, k, l) * MT(ij, l, TI, W2);
                    T1(rhot_TJ, ij) = F3(rho, ij, k, l) * MT(ij, l, TJ, W1);
                    T1(rhovxt_TJ, ij) = F3(rhovx, ij, k, l) * MT(ij, l, TJ, W1);
                    T1(rhovyt_TJ, ij) = F3(rhovy, ij, k, l) * MT(ij, l, TJ, W1);
                    T1(rhovzt_TJ, ij) = F3(rhovz, ij, k, l) * MT(ij, l, TJ, W1);
                }
            /* FLX:163-186  (i,j+1),(i+1,j+1) */
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ijp1 = ij + iall, ip1jp1 = ij + iall + 1;
                    T1(rhot_TI, ij) = T1(rhot_TI, ij) + F3(rho, ip1jp1, k, l) * MT(ij, l, TI, W3);
                    T1(rhovxt_TI, ij) = T1(rhovxt_TI, ij) + F3(rhovx, ip1jp1, k, l) * MT(ij, l, TI, W3);
                    T1(rhovyt_TI, ij) = T1(rhovyt_TI, ij) + F3(rhovy, ip1jp1, k, l) * MT(ij, l, TI, W3);
                    T1(rhovzt_TI, ij) = T1(rhovzt_TI, ij) + F3(rhovz, ip1jp1, k, l) * MT(ij, l, TI, W3);
                    T1(rhot_TJ, ij) = T1(rhot_TJ, ij) + F3(rho, ip1jp1, k, l) * MT(ij, l, TJ, W2) + F3(rho, ijp1, k, l) * MT(ij, l, TJ, W3);
                    T1(rhovxt_TJ, ij) = T1(rhovxt_TJ, ij) + F3(rhovx, ip1jp1, k, l) * MT(ij, l, TJ, W2) + F3(rhovx, ijp1, k, l) * MT(ij, l, TJ, W3);
                    T1(rhovyt_TJ, ij) = T1(rhovyt_TJ, ij) + F3(rhovy, ip1jp1, k, l) * MT(ij, l, TJ, W2) + F3(rhovy, ijp1, k, l) * MT(ij, l, TJ, W3);
                    T1(rhovzt_TJ, ij) = T1(rhovzt_TJ, ij) + F3(rhovz, ip1jp1, k, l) * MT(ij, l, TJ, W2) + F3(rhovz, ijp1, k, l) * MT(ij, l, TJ, W3);
                }
            /* FLX:188-201  singular point */
            if (have_sgp[l - 1]) {
                const int ij = (gmin - 1 - 1) * iall + (gmin - 1), ip1j = ij + 1;
                T1(rhot_TI, ij) = T1(rhot_TJ, ip1j);
                T1(rhovxt_TI, ij) = T1(rhovxt_TJ, ip1j);
                T1(rhovyt_TI, ij) = T1(rhovyt_TJ, ip1j);
                T1(rhovzt_TI, ij) = T1(rhovzt_TJ, ip1j);
            }
            /* FLX:205-232  zero the halo */
            for (int j = 1; j <= iall; ++j)
                for (int i = 1; i <= iall; ++i)
                    if (i < gmin || i > gmax || j < gmin || j > gmax) {
                        const int ij = (j - 1) * iall + i;
                        for (int m = 1; m <= 6; ++m) F4(flx_h, ij, k, l, m) = 0.0;
                        for (int a = AI; a <= AJ; ++a)
                            for (int dd = XDIR; dd <= ZDIR; ++dd) F5(GRD_xc, ij, k, l, a, dd) = 0.0;
                    }
            /* FLX:234-258  arc AI */
            for (int j = gmin; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ip1j = ij + 1, ijm1 = ij - iall;
                    const double rrhoa2 = 1.0 / mx(T1(rhot_TJ, ijm1) + T1(rhot_TI, ij), EPS);
                    const double rhovxt2 = T1(rhovxt_TJ, ijm1) + T1(rhovxt_TI, ij);
                    const double rhovyt2 = T1(rhovyt_TJ, ijm1) + T1(rhovyt_TI, ij);
                    const double rhovzt2 = T1(rhovzt_TJ, ijm1) + T1(rhovzt_TI, ij);
                    const double flux = 0.5 * (rhovxt2 * MA(ij, l, AI, HNX) + rhovyt2 * MA(ij, l, AI, HNY) + rhovzt2 * MA(ij, l, AI, HNZ));
                    F4(flx_h, ij, k, l, 1) = flux * MP(ij, l, P_RAREA) * dt;
                    F4(flx_h, ip1j, k, l, 4) = -flux * MP(ip1j, l, P_RAREA) * dt;
                    F5(GRD_xc, ij, k, l, AI, XDIR) = XR(ij, l, AI, XDIR) - rhovxt2 * rrhoa2 * dt * 0.5;
                    F5(GRD_xc, ij, k, l, AI, YDIR) = XR(ij, l, AI, YDIR) - rhovyt2 * rrhoa2 * dt * 0.5;
                    F5(GRD_xc, ij, k, l, AI, ZDIR) = XR(ij, l, AI, ZDIR) - rhovzt2 * rrhoa2 * dt * 0.5;
                }
            /* FLX:260-283  arc AIJ */
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ip1jp1 = ij + iall + 1;
                    const double rrhoa2 = 1.0 / mx(T1(rhot_TI, ij) + T1(rhot_TJ, ij), EPS);
                    const double rhovxt2 = T1(rhovxt_TI, ij) + T1(rhovxt_TJ, ij);
                    const double rhovyt2 = T1(rhovyt_TI, ij) + T1(rhovyt_TJ, ij);
                    const double rhovzt2 = T1(rhovzt_TI, ij) + T1(rhovzt_TJ, ij);
                    const double flux = 0.5 * (rhovxt2 * MA(ij, l, AIJ, HNX) + rhovyt2 * MA(ij, l, AIJ, HNY) + rhovzt2 * MA(ij, l, AIJ, HNZ));
                    F4(flx_h, ij, k, l, 2) = flux * MP(ij, l, P_RAREA) * dt;
                    F4(flx_h, ip1jp1, k, l, 5) = -flux * MP(ip1jp1, l, P_RAREA) * dt;
                    F5(GRD_xc, ij, k, l, AIJ, XDIR) = XR(ij, l, AIJ, XDIR) - rhovxt2 * rrhoa2 * dt * 0.5;
                    F5(GRD_xc, ij, k, l, AIJ, YDIR) = XR(ij, l, AIJ, YDIR) - rhovyt2 * rrhoa2 * dt * 0.5;
                    F5(GRD_xc, ij, k, l, AIJ, ZDIR) = XR(ij, l, AIJ, ZDIR) - rhovzt2 * rrhoa2 * dt * 0.5;
                }
            /* FLX:285-309  arc AJ */
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ijp1 = ij + iall, im1j = ij - 1;
                    const double rrhoa2 = 1.0 / mx(T1(rhot_TJ, ij) + T1(rhot_TI, im1j), EPS);
                    const double rhovxt2 = T1(rhovxt_TJ, ij) + T1(rhovxt_TI, im1j);
                    const double rhovyt2 = T1(rhovyt_TJ, ij) + T1(rhovyt_TI, im1j);
                    const double rhovzt2 = T1(rhovzt_TJ, ij) + T1(rhovzt_TI, im1j);
                    const double flux = 0.5 * (rhovxt2 * MA(ij, l, AJ, HNX) + rhovyt2 * MA(ij, l, AJ, HNY) + rhovzt2 * MA(ij, l, AJ, HNZ));
                    F4(flx_h, ij, k, l, 3) = flux * MP(ij, l, P_RAREA) * dt;
                    F4(flx_h, ijp1, k, l, 6) = -flux * MP(ijp1, l, P_RAREA) * dt;
                    F5(GRD_xc, ij, k, l, AJ, XDIR) = XR(ij, l, AJ, XDIR) - rhovxt2 * rrhoa2 * dt * 0.5;
                    F5(GRD_xc, ij, k, l, AJ, YDIR) = XR(ij, l, AJ, YDIR) - rhovyt2 * rrhoa2 * dt * 0.5;
                    F5(GRD_xc, ij, k, l, AJ, ZDIR) = XR(ij, l, AJ, ZDIR) - rhovzt2 * rrhoa2 * dt * 0.5;
                }
            /* FLX:311-320 */
            if (have_sgp[l - 1]) {
                const int ij = (gmin - 1) * iall + gmin;
                F4(flx_h, ij, k, l, 6) = 0.0;
            }
#undef T1
            free(rhot_TI);
        }
    }

    /* FLX:326-374  pole regions */
    if (D->have_pl) {
        const int n = GSLF_PL;
        double rhot_pl[16], rhovxt_pl[16], rhovyt_pl[16], rhovzt_pl[16];
        for (int l = 1; l <= LALL_PL; ++l)
            for (int k = 1; k <= kall; ++k) {
                for (int v = GMIN_PL; v <= gmax_pl; ++v) {
                    const int ij = v;
                    int ijp1 = v + 1;
                    if (ijp1 == gmax_pl + 1) ijp1 = GMIN_PL;
                    rhot_pl[v] = P3(rho_pl, n, k, l) * MT_PL(ij, l, W1) + P3(rho_pl, ij, k, l) * MT_PL(ij, l, W2) + P3(rho_pl, ijp1, k, l) * MT_PL(ij, l, W3);
                    rhovxt_pl[v] = P3(rhovx_pl, n, k, l) * MT_PL(ij, l, W1) + P3(rhovx_pl, ij, k, l) * MT_PL(ij, l, W2) + P3(rhovx_pl, ijp1, k, l) * MT_PL(ij, l, W3);
                    rhovyt_pl[v] = P3(rhovy_pl, n, k, l) * MT_PL(ij, l, W1) + P3(rhovy_pl, ij, k, l) * MT_PL(ij, l, W2) + P3(rhovy_pl, ijp1, k, l) * MT_PL(ij, l, W3);
                    rhovzt_pl[v] = P3(rhovz_pl, n, k, l) * MT_PL(ij, l, W1) + P3(rhovz_pl, ij, k, l) * MT_PL(ij, l, W2) + P3(rhovz_pl, ijp1, k, l) * MT_PL(ij, l, W3);
                }
                for (int v = GMIN_PL; v <= gmax_pl; ++v) {
                    const int ij = v;
                    int ijm1 = v - 1;
                    if (ijm1 == GMIN_PL - 1) ijm1 = gmax_pl;
                    const double rrhoa2 = 1.0 / mx(rhot_pl[ijm1] + rhot_pl[ij], EPS);
                    const double rhovxt2 = rhovxt_pl[ijm1] + rhovxt_pl[ij];
                    const double rhovyt2 = rhovyt_pl[ijm1] + rhovyt_pl[ij];
                    const double rhovzt2 = rhovzt_pl[ijm1] + rhovzt_pl[ij];
                    const double flux = 0.5 * (rhovxt2 * MA_PL(ij, l, HNX) + rhovyt2 * MA_PL(ij, l, HNY) + rhovzt2 * MA_PL(ij, l, HNZ));
                    P3(flx_h_pl, v, k, l) = flux * MP_PL(n, l, P_RAREA) * dt;
                    P4(GRD_xc_pl, v, k, l, XDIR) = XR_PL(v, l, XDIR) - rhovxt2 * rrhoa2 * dt * 0.5;
                    P4(GRD_xc_pl, v, k, l, YDIR) = XR_PL(v, l, YDIR) - rhovyt2 * rrhoa2 * dt * 0.5;
                    P4(GRD_xc_pl, v, k, l, ZDIR) = XR_PL(v, l, ZDIR) - rhovzt2 * rrhoa2 * dt * 0.5;
                }
            }
    }
}

/* ===========================================================================
 * horizontal_limiter_thuburn: HLIM/mod_src_tracer.f90:42-430
 * The automatic arrays Qin/Qout of the reference (:98-101) are allocated per
 * (l, k) slab here; only one slab is live at a time in each thread.
 * ======================================================================== */
void orc_horizontal_limiter_thuburn(const orc_dims *D, const int *have_sgp,
                                    double *q_a, double *q_a_pl,
                                    const double *q, const double *q_pl,
                                    const double *d, const double *d_pl,
                                    const double *ch, const double *ch_pl,
                                    const double *cmask, const double *cmask_pl,
                                    double *Qout_prev, double *Qout_prev_pl,
                                    const double *Qout_post, const double *Qout_post_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall;
    const int gall_pl = D->gall_pl, gmax_pl = gall_pl;
    const int gmin = 2, gmax = iall - 1;
    const double EPS = D->eps, BIG = D->big;

    /* Qin_pl(gall_pl,kall,2,2,2) and Qout_pl(gall_pl,kall,2,2); :99,101 */
    const size_t npl = (size_t)gall_pl * kall * LALL_PL;
    double *Qin_pl = (double *)calloc(npl * 4, sizeof(double));
    double *Qout_pl = (double *)calloc(npl * 2, sizeof(double));
#define QINPL(v, k, l, mm, f) Qin_pl[((((f) - 1) * 2 + ((mm) - 1)) * npl) + (((size_t)((l) - 1) * kall + ((k) - 1)) * gall_pl + ((v) - 1))]
#define QOUTPL(v, k, l, mm) Qout_pl[(((mm) - 1) * npl) + (((size_t)((l) - 1) * kall + ((k) - 1)) * gall_pl + ((v) - 1))]

    for (int l = 1; l <= lall; ++l) {
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= kall; ++k) {
            /* per-slab Qin(gall,2,6) and Qout(gall,2) */
            double *Qin = (double *)malloc(sizeof(double) * gall * 12);
            double *Qout = (double *)malloc(sizeof(double) * gall * 2);
#define QIN(g, mm, f) Qin[(((f) - 1) * 2 + ((mm) - 1)) * (size_t)gall + ((g) - 1)]
#define QOUT(g, mm) Qout[((mm) - 1) * (size_t)gall + ((g) - 1)]
            /* HLIM:140-190  inflow bounds */
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ip1j = ij + 1, ip1jp1 = ij + iall + 1, ijp1 = ij + iall;
                    int im1j = ij - 1, ijm1 = ij - iall;
                    im1j = im1j > 1 ? im1j : 1; /* :152-153 */
                    ijm1 = ijm1 > 1 ? ijm1 : 1;
                    const double qij = F3(q, ij, k, l);
                    const double q_min_AI = mn(mn(mn(qij, F3(q, ijm1, k, l)), F3(q, ip1j, k, l)), F3(q, ip1jp1, k, l));
                    const double q_max_AI = mx(mx(mx(qij, F3(q, ijm1, k, l)), F3(q, ip1j, k, l)), F3(q, ip1jp1, k, l));
                    const double q_min_AIJ = mn(mn(mn(qij, F3(q, ip1j, k, l)), F3(q, ip1jp1, k, l)), F3(q, ijp1, k, l));
                    const double q_max_AIJ = mx(mx(mx(qij, F3(q, ip1j, k, l)), F3(q, ip1jp1, k, l)), F3(q, ijp1, k, l));
                    const double q_min_AJ = mn(mn(mn(qij, F3(q, ip1jp1, k, l)), F3(q, ijp1, k, l)), F3(q, im1j, k, l));
                    const double q_max_AJ = mx(mx(mx(qij, F3(q, ip1jp1, k, l)), F3(q, ijp1, k, l)), F3(q, im1j, k, l));
                    const double c1 = F4(cmask, ij, k, l, 1), c2 = F4(cmask, ij, k, l, 2), c3 = F4(cmask, ij, k, l, 3);

                    QIN(ij, I_MIN, 1) = (c1)*q_min_AI + (1.0 - c1) * BIG;
                    QIN(ip1j, I_MIN, 4) = (c1)*BIG + (1.0 - c1) * q_min_AI;
                    QIN(ij, I_MAX, 1) = (c1)*q_max_AI + (1.0 - c1) * (-BIG);
                    QIN(ip1j, I_MAX, 4) = (c1) * (-BIG) + (1.0 - c1) * q_max_AI;

                    QIN(ij, I_MIN, 2) = (c2)*q_min_AIJ + (1.0 - c2) * BIG;
                    QIN(ip1jp1, I_MIN, 5) = (c2)*BIG + (1.0 - c2) * q_min_AIJ;
                    QIN(ij, I_MAX, 2) = (c2)*q_max_AIJ + (1.0 - c2) * (-BIG);
                    QIN(ip1jp1, I_MAX, 5) = (c2) * (-BIG) + (1.0 - c2) * q_max_AIJ;

                    QIN(ij, I_MIN, 3) = (c3)*q_min_AJ + (1.0 - c3) * BIG;
                    QIN(ijp1, I_MIN, 6) = (c3)*BIG + (1.0 - c3) * q_min_AJ;
                    QIN(ij, I_MAX, 3) = (c3)*q_max_AJ + (1.0 - c3) * (-BIG);
                    QIN(ijp1, I_MAX, 6) = (c3) * (-BIG) + (1.0 - c3) * q_max_AJ;
                }
            /* HLIM:192-214  singular point */
            if (have_sgp[l - 1]) {
                const int ij = (gmin - 1 - 1) * iall + (gmin - 1);
                const int ijp1 = ij + iall, ip1jp1 = ij + iall + 1, ip2jp1 = ij + iall + 2;
                const double q_min_AIJ = mn(mn(mn(F3(q, ij, k, l), F3(q, ip1jp1, k, l)), F3(q, ip2jp1, k, l)), F3(q, ijp1, k, l));
                const double q_max_AIJ = mx(mx(mx(F3(q, ij, k, l), F3(q, ip1jp1, k, l)), F3(q, ip2jp1, k, l)), F3(q, ijp1, k, l));
                const double c2 = F4(cmask, ij, k, l, 2);
                QIN(ij, I_MIN, 2) = (c2)*q_min_AIJ + (1.0 - c2) * BIG;
                QIN(ip1jp1, I_MIN, 5) = (c2)*BIG + (1.0 - c2) * q_min_AIJ;
                QIN(ij, I_MAX, 2) = (c2)*q_max_AIJ + (1.0 - c2) * (-BIG);
                QIN(ip1jp1, I_MAX, 5) = (c2) * (-BIG) + (1.0 - c2) * q_max_AIJ;
            }
            /* HLIM:216-284  allowable range */
            for (int j = gmin; j <= gmax; ++j)
                for (int i = gmin; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i;
                    const double qij = F3(q, ij, k, l);
                    double qnext_min = qij, qnext_max = qij;
                    for (int f = 1; f <= 6; ++f) qnext_min = mn(qnext_min, QIN(ij, I_MIN, f));
                    for (int f = 1; f <= 6; ++f) qnext_max = mx(qnext_max, QIN(ij, I_MAX, f));
                    double chv[7], chm[7];
                    for (int f = 1; f <= 6; ++f) {
                        chv[f] = F4(ch, ij, k, l, f);
                        chm[f] = mn(chv[f], 0.0);
                    }
                    const double Cin_sum = chm[1] + chm[2] + chm[3] + chm[4] + chm[5] + chm[6];
                    const double Cout_sum = chv[1] - chm[1] + chv[2] - chm[2] + chv[3] - chm[3] + chv[4] - chm[4] + chv[5] - chm[5] + chv[6] - chm[6];
                    const double CQin_min_sum = chm[1] * QIN(ij, I_MIN, 1) + chm[2] * QIN(ij, I_MIN, 2) + chm[3] * QIN(ij, I_MIN, 3) + chm[4] * QIN(ij, I_MIN, 4) + chm[5] * QIN(ij, I_MIN, 5) + chm[6] * QIN(ij, I_MIN, 6);
                    const double CQin_max_sum = chm[1] * QIN(ij, I_MAX, 1) + chm[2] * QIN(ij, I_MAX, 2) + chm[3] * QIN(ij, I_MAX, 3) + chm[4] * QIN(ij, I_MAX, 4) + chm[5] * QIN(ij, I_MAX, 5) + chm[6] * QIN(ij, I_MAX, 6);
                    const double zerosw = 0.5 - fsign(0.5, fabs(Cout_sum) - EPS);
                    const double dij = F3(d, ij, k, l);
                    QOUT(ij, I_MIN) = (qij - CQin_max_sum - qnext_max * (1.0 - Cin_sum - Cout_sum + dij)) / (Cout_sum + zerosw) * (1.0 - zerosw) + qij * zerosw;
                    QOUT(ij, I_MAX) = (qij - CQin_min_sum - qnext_min * (1.0 - Cin_sum - Cout_sum + dij)) / (Cout_sum + zerosw) * (1.0 - zerosw) + qij * zerosw;
                }
            /* HLIM:286-301  halo: Qout = q */
            for (int j = 1; j <= iall; ++j)
                for (int i = 1; i <= iall; ++i)
                    if (i < gmin || i > gmax || j < gmin || j > gmax) {
                        const int ij = (j - 1) * iall + i;
                        QOUT(ij, I_MIN) = F3(q, ij, k, l);
                        QOUT(ij, I_MAX) = F3(q, ij, k, l);
                    }
            /* HLIM:369 Qout_prev = Qout;  :372 Qout = Qout_post */
            for (int g = 1; g <= gall; ++g) {
                F4(Qout_prev, g, k, l, I_MIN) = QOUT(g, I_MIN);
                F4(Qout_prev, g, k, l, I_MAX) = QOUT(g, I_MAX);
                QOUT(g, I_MIN) = F4(Qout_post, g, k, l, I_MIN);
                QOUT(g, I_MAX) = F4(Qout_post, g, k, l, I_MAX);
            }
            /* HLIM:377-410  apply inflow/outflow limiter */
            for (int j = gmin - 1; j <= gmax; ++j)
                for (int i = gmin - 1; i <= gmax; ++i) {
                    const int ij = (j - 1) * iall + i, ip1j = ij + 1, ip1jp1 = ij + iall + 1, ijp1 = ij + iall;
                    const int nb[4] = {0, ip1j, ip1jp1, ijp1};
                    for (int m = 1; m <= 3; ++m) {
                        const double c = F4(cmask, ij, k, l, m);
                        double qa = F4(q_a, ij, k, l, m);
                        qa = (c)*mn(mx(qa, QIN(ij, I_MIN, m)), QIN(ij, I_MAX, m)) + (1.0 - c) * mn(mx(qa, QIN(nb[m], I_MIN, m + 3)), QIN(nb[m], I_MAX, m + 3));
                        qa = (c)*mx(mn(qa, QOUT(nb[m], I_MAX)), QOUT(nb[m], I_MIN)) + (1.0 - c) * mx(mn(qa, QOUT(ij, I_MAX)), QOUT(ij, I_MIN));
                        F4(q_a, ij, k, l, m) = qa;
                        F4(q_a, nb[m], k, l, m + 3) = qa;
                    }
                }
#undef QIN
#undef QOUT
            free(Qin);
            free(Qout);
        }
    }

    if (D->have_pl) {
        const int n = GSLF_PL;
        /* HLIM:307-363 */
        for (int l = 1; l <= LALL_PL; ++l)
            for (int k = 1; k <= kall; ++k) {
                for (int v = GMIN_PL; v <= gmax_pl; ++v) {
                    const int ij = v;
                    int ijp1 = v + 1, ijm1 = v - 1;
                    if (ijp1 == gmax_pl + 1) ijp1 = GMIN_PL;
                    if (ijm1 == GMIN_PL - 1) ijm1 = gmax_pl;
                    const double q_min_pl = mn(mn(mn(P3(q_pl, n, k, l), P3(q_pl, ij, k, l)), P3(q_pl, ijm1, k, l)), P3(q_pl, ijp1, k, l));
                    const double q_max_pl = mx(mx(mx(P3(q_pl, n, k, l), P3(q_pl, ij, k, l)), P3(q_pl, ijm1, k, l)), P3(q_pl, ijp1, k, l));
                    const double c = P3(cmask_pl, ij, k, l);
                    QINPL(ij, k, l, I_MIN, 1) = (c)*q_min_pl + (1.0 - c) * BIG;
                    QINPL(ij, k, l, I_MIN, 2) = (c)*BIG + (1.0 - c) * q_min_pl;
                    QINPL(ij, k, l, I_MAX, 1) = (c)*q_max_pl + (1.0 - c) * (-BIG);
                    QINPL(ij, k, l, I_MAX, 2) = (c) * (-BIG) + (1.0 - c) * q_max_pl;
                }
                double qnext_min_pl = P3(q_pl, n, k, l), qnext_max_pl = P3(q_pl, n, k, l);
                for (int v = GMIN_PL; v <= gmax_pl; ++v) {
                    qnext_min_pl = mn(qnext_min_pl, QINPL(v, k, l, I_MIN, 1));
                    qnext_max_pl = mx(qnext_max_pl, QINPL(v, k, l, I_MAX, 1));
                }
                double Cin_sum_pl = 0.0, Cout_sum_pl = 0.0, CQin_max_sum_pl = 0.0, CQin_min_sum_pl = 0.0;
                for (int v = GMIN_PL; v <= gmax_pl; ++v) {
                    const double ch_masked = P3(cmask_pl, v, k, l) * P3(ch_pl, v, k, l);
                    Cin_sum_pl = Cin_sum_pl + ch_masked;
                    Cout_sum_pl = Cout_sum_pl - ch_masked + P3(ch_pl, v, k, l);
                    CQin_min_sum_pl = CQin_min_sum_pl + ch_masked * QINPL(v, k, l, I_MIN, 1);
                    CQin_max_sum_pl = CQin_max_sum_pl + ch_masked * QINPL(v, k, l, I_MAX, 1);
                }
                const double zerosw = 0.5 - fsign(0.5, fabs(Cout_sum_pl) - EPS);
                const double qn = P3(q_pl, n, k, l), dn = P3(d_pl, n, k, l);
                QOUTPL(n, k, l, I_MIN) = (qn - CQin_max_sum_pl - qnext_max_pl * (1.0 - Cin_sum_pl - Cout_sum_pl + dn)) / (Cout_sum_pl + zerosw) * (1.0 - zerosw) + qn * zerosw;
                QOUTPL(n, k, l, I_MAX) = (qn - CQin_min_sum_pl - qnext_min_pl * (1.0 - Cin_sum_pl - Cout_sum_pl + dn)) / (Cout_sum_pl + zerosw) * (1.0 - zerosw) + qn * zerosw;
            }
        /* HLIM:367-373  ring zeroed, copied out, replaced by the post-exchange values */
        for (int l = 1; l <= LALL_PL; ++l)
            for (int k = 1; k <= kall; ++k)
                for (int v = 1; v <= gall_pl; ++v)
                    for (int mm = I_MIN; mm <= I_MAX; ++mm) {
                        if (v >= GMIN_PL) QOUTPL(v, k, l, mm) = 0.0;
                        P4(Qout_prev_pl, v, k, l, mm) = QOUTPL(v, k, l, mm);
                        QOUTPL(v, k, l, mm) = P4(Qout_post_pl, v, k, l, mm);
                    }
        /* HLIM:412-425 */
        for (int l = 1; l <= LALL_PL; ++l)
            for (int k = 1; k <= kall; ++k)
                for (int v = GMIN_PL; v <= gmax_pl; ++v) {
                    const double c = P3(cmask_pl, v, k, l);
                    double qa = P3(q_a_pl, v, k, l);
                    qa = (c)*mn(mx(qa, QINPL(v, k, l, I_MIN, 1)), QINPL(v, k, l, I_MAX, 1)) + (1.0 - c) * mn(mx(qa, QINPL(v, k, l, I_MIN, 2)), QINPL(v, k, l, I_MAX, 2));
                    qa = (c)*mx(mn(qa, QOUTPL(v, k, l, I_MAX)), QOUTPL(v, k, l, I_MIN)) + (1.0 - c) * mx(mn(qa, QOUTPL(n, k, l, I_MAX)), QOUTPL(n, k, l, I_MIN));
                    P3(q_a_pl, v, k, l) = qa;
                }
    } else {
        /* HLIM:367-370 run without a pole region too: the ring of Qout_prev_pl
         * is zero and its centre point is left undefined by the reference. */
        for (int l = 1; l <= LALL_PL; ++l)
            for (int k = 1; k <= kall; ++k)
                for (int v = GMIN_PL; v <= gall_pl; ++v)
                    for (int mm = I_MIN; mm <= I_MAX; ++mm) P4(Qout_prev_pl, v, k, l, mm) = 0.0;
    }
#undef QINPL
#undef QOUTPL
    free(Qin_pl);
    free(Qout_pl);
}

/* ===========================================================================
 * vertical_limiter_thuburn: VLIM/mod_src_tracer.f90:41-243
 * ======================================================================== */
void orc_vertical_limiter_thuburn(const orc_dims *D,
                                  double *q_h, double *q_h_pl,
                                  const double *q, const double *q_pl,
                                  const double *d, const double *d_pl,
                                  const double *ck, const double *ck_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall;
    const int gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    const double EPS = D->eps, BIG = D->big;

    for (int l = 1; l <= lall; ++l) {
        /* The reference parallelises over g inside the k loop (:145-185); the
         * columns are independent, so threads own whole columns here. */
#pragma omp parallel for schedule(static)
        for (int g = 1; g <= gall; ++g) {
            double Qout_min_km1 = 0.0, Qout_max_km1 = 0.0;
            for (int k = kmin; k <= kmax; ++k) {
                const double inflagL = 0.5 - fsign(0.5, F4(ck, g, k, l, 1));
                const double inflagU = 0.5 + fsign(0.5, F4(ck, g, k + 1, l, 1));
                const double qk = F3(q, g, k, l);
                const double Qin_minL = mn(qk, F3(q, g, k - 1, l)) + (1.0 - inflagL) * BIG;
                const double Qin_minU = mn(qk, F3(q, g, k + 1, l)) + (1.0 - inflagU) * BIG;
                const double Qin_maxL = mx(qk, F3(q, g, k - 1, l)) - (1.0 - inflagL) * BIG;
                const double Qin_maxU = mx(qk, F3(q, g, k + 1, l)) - (1.0 - inflagU) * BIG;
                const double qnext_min = mn(mn(Qin_minL, Qin_minU), qk);
                const double qnext_max = mx(mx(Qin_maxL, Qin_maxU), qk);
                const double ck1 = F4(ck, g, k, l, 1), ck2 = F4(ck, g, k, l, 2);
                const double Cin = (inflagL)*ck1 + (inflagU)*ck2;
                const double Cout = (1.0 - inflagL) * ck1 + (1.0 - inflagU) * ck2;
                const double CQin_min = (inflagL)*ck1 * Qin_minL + (inflagU)*ck2 * Qin_minU;
                const double CQin_max = (inflagL)*ck1 * Qin_maxL + (inflagU)*ck2 * Qin_maxU;
                const double zerosw = 0.5 - fsign(0.5, fabs(Cout) - EPS);
                const double dk = F3(d, g, k, l);
                const double Qout_min_k = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) * (1.0 - zerosw) + qk * zerosw;
                const double Qout_max_k = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) * (1.0 - zerosw) + qk * zerosw;
                if (k > kmin) { /* :179-180; the kmin level is peeled (:106-143) */
                    const double qh = F3(q_h, g, k, l);
                    F3(q_h, g, k, l) = (inflagL)*mx(mn(qh, Qout_max_km1), Qout_min_km1) + (1.0 - inflagL) * mx(mn(qh, Qout_max_k), Qout_min_k);
                }
                Qout_min_km1 = Qout_min_k;
                Qout_max_km1 = Qout_max_k;
            }
        }
    }

    if (D->have_pl) {
        double *Qmin = (double *)malloc(sizeof(double) * gall_pl * kall);
        double *Qmax = (double *)malloc(sizeof(double) * gall_pl * kall);
#define QP(a, g, k) a[((k) - 1) * gall_pl + ((g) - 1)]
        for (int l = 1; l <= LALL_PL; ++l) {
            /* VLIM:194-226 (note the different product grouping of CQin here) */
            for (int k = kmin; k <= kmax; ++k)
                for (int g = 1; g <= gall_pl; ++g) {
                    const double inflagL = 0.5 - fsign(0.5, P4(ck_pl, g, k, l, 1));
                    const double inflagU = 0.5 + fsign(0.5, P4(ck_pl, g, k + 1, l, 1));
                    const double qk = P3(q_pl, g, k, l);
                    const double Qin_minL = mn(qk, P3(q_pl, g, k - 1, l)) + (1.0 - inflagL) * BIG;
                    const double Qin_minU = mn(qk, P3(q_pl, g, k + 1, l)) + (1.0 - inflagU) * BIG;
                    const double Qin_maxL = mx(qk, P3(q_pl, g, k - 1, l)) - (1.0 - inflagL) * BIG;
                    const double Qin_maxU = mx(qk, P3(q_pl, g, k + 1, l)) - (1.0 - inflagU) * BIG;
                    const double qnext_min = mn(mn(Qin_minL, Qin_minU), qk);
                    const double qnext_max = mx(mx(Qin_maxL, Qin_maxU), qk);
                    const double ck1 = P4(ck_pl, g, k, l, 1), ck2 = P4(ck_pl, g, k, l, 2);
                    const double Cin = (inflagL) * (ck1) + (inflagU) * (ck2);
                    const double Cout = (1.0 - inflagL) * (ck1) + (1.0 - inflagU) * (ck2);
                    const double CQin_max = (inflagL) * (ck1 * Qin_maxL) + (inflagU) * (ck2 * Qin_maxU);
                    const double CQin_min = (inflagL) * (ck1 * Qin_minL) + (inflagU) * (ck2 * Qin_minU);
                    const double zerosw = 0.5 - fsign(0.5, fabs(Cout) - EPS);
                    const double dk = P3(d_pl, g, k, l);
                    QP(Qmin, g, k) = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) * (1.0 - zerosw) + qk * zerosw;
                    QP(Qmax, g, k) = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) * (1.0 - zerosw) + qk * zerosw;
                }
            /* VLIM:228-235 */
            for (int k = kmin + 1; k <= kmax; ++k)
                for (int g = 1; g <= gall_pl; ++g) {
                    const double inflagL = 0.5 - fsign(0.5, P4(ck_pl, g, k, l, 1));
                    const double qh = P3(q_h_pl, g, k, l);
                    P3(q_h_pl, g, k, l) = (inflagL)*mx(mn(qh, QP(Qmax, g, k - 1)), QP(Qmin, g, k - 1)) + (1.0 - inflagL) * mx(mn(qh, QP(Qmax, g, k)), QP(Qmin, g, k));
                }
        }
#undef QP
        free(Qmin);
        free(Qmax);
    }
}

/* ===========================================================================
 * vi_rhow_solver: VI/mod_vi.f90:93-306
 * ======================================================================== */
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
                        double grav, double rdry, double cvdry, int non_hydro_alpha)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall;
    const int gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    const double CVovRt2 = cvdry / rdry / (dt * dt); /* :178 */
    const double alpha = (double)non_hydro_alpha;     /* :179 */

    /* One column (g, l) at a time; the reference's k-outer/g-inner sweeps
     * (:194-246) touch each column independently, so the order of columns
     * does not change any value. */
#define VI_COLUMN(G, L, NG, W, W0, PR, RH, SR, SWW, SP, MCC, MLL, MUU, RG, RGH, RGS, RGSH, GSH, A3)        \
    do {                                                                                                   \
        double Sall[kall + 1], gam[kall + 1];                                                              \
        for (int k = kmin + 1; k <= kmax; ++k) {                                                           \
            Sall[k] = ((A3(W0, G, k, L) * alpha + dt * A3(SWW, G, k, L)) * (A3(RGH, G, k, L) * A3(RGH, G, k, L)) \
                       - ((A3(PR, G, k, L) + dt * A3(SP, G, k, L)) * A3(RGS, G, k, L)                      \
                          - (A3(PR, G, k - 1, L) + dt * A3(SP, G, k - 1, L)) * A3(RGS, G, k - 1, L))       \
                             * dt * rdgzh[k - 1]                                                           \
                       - ((A3(RH, G, k, L) + dt * A3(SR, G, k, L)) * (A3(RG, G, k, L) * A3(RG, G, k, L)) * afact[k - 1] \
                          + (A3(RH, G, k - 1, L) + dt * A3(SR, G, k - 1, L)) * (A3(RG, G, k - 1, L) * A3(RG, G, k - 1, L)) * bfact[k - 1]) \
                             * dt * grav)                                                                  \
                      * CVovRt2;                                                                           \
        }                                                                                                  \
        A3(W, G, kmin, L) = A3(W, G, kmin, L) * A3(RGSH, G, kmin, L);                                     \
        A3(W, G, kmax + 1, L) = A3(W, G, kmax + 1, L) * A3(RGSH, G, kmax + 1, L);                         \
        Sall[kmin + 1] = Sall[kmin + 1] - A3(MLL, G, kmin + 1, L) * A3(W, G, kmin, L);                     \
        Sall[kmax] = Sall[kmax] - A3(MUU, G, kmax, L) * A3(W, G, kmax + 1, L);                             \
        double beta = A3(MCC, G, kmin + 1, L);                                                             \
        A3(W, G, kmin + 1, L) = Sall[kmin + 1] / beta;                                                     \
        for (int k = kmin + 2; k <= kmax; ++k) {                                                           \
            gam[k] = A3(MUU, G, k - 1, L) / beta;                                                          \
            beta = A3(MCC, G, k, L) - A3(MLL, G, k, L) * gam[k];                                           \
            A3(W, G, k, L) = (Sall[k] - A3(MLL, G, k, L) * A3(W, G, k - 1, L)) / beta;                     \
        }                                                                                                  \
        for (int k = kmax - 1; k >= kmin + 1; --k) {                                                       \
            A3(W, G, k, L) = A3(W, G, k, L) - gam[k + 1] * A3(W, G, k + 1, L);                             \
            A3(W, G, k + 1, L) = A3(W, G, k + 1, L) * A3(GSH, G, k + 1, L);                                \
        }                                                                                                  \
        A3(W, G, kmin, L) = A3(W, G, kmin, L) * A3(GSH, G, kmin, L);                                       \
        A3(W, G, kmin + 1, L) = A3(W, G, kmin + 1, L) * A3(GSH, G, kmin + 1, L);                           \
        A3(W, G, kmax + 1, L) = A3(W, G, kmax + 1, L) * A3(GSH, G, kmax + 1, L);                           \
    } while (0)

    for (int l = 1; l <= lall; ++l) {
#pragma omp parallel for schedule(static)
        for (int g = 1; g <= gall; ++g)
            VI_COLUMN(g, l, gall, rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, Mc, Ml, Mu,
                      RGAM, RGAMH, RGSGAM2, RGSGAM2H, GSGAM2H, F3);
    }
    if (D->have_pl) { /* :250-301 */
        for (int l = 1; l <= LALL_PL; ++l)
            for (int g = 1; g <= gall_pl; ++g)
                VI_COLUMN(g, l, gall_pl, rhogw_pl, rhogw0_pl, preg0_pl, rhog0_pl, Srho_pl, Sw_pl, Spre_pl,
                          Mc_pl, Ml_pl, Mu_pl, RGAM_pl, RGAMH_pl, RGSGAM2_pl, RGSGAM2H_pl, GSGAM2H_pl, P3);
    }
#undef VI_COLUMN
}

/* ------------------------------------------------------------------------- */
/* dyn_diffusion/src/mod_oprt.f90:41-296                                     */
/* ------------------------------------------------------------------------- */
void orc_oprt_diffusion(const orc_dims *D, const int *have_sgp,
                        double *dscl, double *dscl_pl,
                        const double *scl, const double *scl_pl,
                        const double *kh, const double *kh_pl,
                        const double *coef_intp, const double *coef_intp_pl,
                        const double *coef_diff, const double *coef_diff_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall;
    const int gall_pl = D->gall_pl;
    /* 1-based flattened bounds (:85-93) */
    const int gmin = (2 - 1) * iall + 2, gmax = (iall - 1 - 1) * iall + iall - 1;
    const int gminm1 = (2 - 1 - 1) * iall + 2 - 1;
#define CI(g, c, d, t, l) coef_intp[((((size_t)((l) - 1) * 2 + (t)) * 3 + ((d) - 1)) * 3 + ((c) - 1)) * gall + ((g) - 1)]
#define CD(g, m, d, l) coef_diff[(((size_t)((l) - 1) * 3 + ((d) - 1)) * 6 + ((m) - 1)) * gall + ((g) - 1)]
#define CI_PL(v, c, d, l) coef_intp_pl[((((size_t)((l) - 1)) * 3 + ((d) - 1)) * 3 + ((c) - 1)) * gall_pl + ((v) - 1)]
#define CD_PL(m, d, l) coef_diff_pl[(((size_t)((l) - 1)) * 3 + ((d) - 1)) * 5 + ((m) - 1)]

    for (int l = 1; l <= lall; ++l) {
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= kall; ++k) {
            double *vt = (double *)malloc(sizeof(double) * (size_t)gall * 6); /* vt(g, d, t) */
#define VT(g, d, t) vt[((size_t)(t) * 3 + ((d) - 1)) * gall + ((g) - 1)]
            for (int d = 1; d <= 3; ++d) {
                for (int g = gminm1; g <= gmax; ++g) { /* :97-110 */
                    const int ij = g, ip1j = g + 1, ip1jp1 = g + iall + 1;
                    const double c1 = CI(g, 1, d, 0, l), c2 = CI(g, 2, d, 0, l), c3 = CI(g, 3, d, 0, l);
                    VT(g, d, 0) = ((+2.0 * c1 - 1.0 * c2 - 1.0 * c3) * F3(scl, ij, k, l)
                                 + (-1.0 * c1 + 2.0 * c2 - 1.0 * c3) * F3(scl, ip1j, k, l)
                                 + (-1.0 * c1 - 1.0 * c2 + 2.0 * c3) * F3(scl, ip1jp1, k, l)) / 3.0;
                }
                for (int g = gminm1; g <= gmax; ++g) { /* :112-125 */
                    const int ij = g, ip1jp1 = g + iall + 1, ijp1 = g + iall;
                    const double c1 = CI(g, 1, d, 1, l), c2 = CI(g, 2, d, 1, l), c3 = CI(g, 3, d, 1, l);
                    VT(g, d, 1) = ((+2.0 * c1 - 1.0 * c2 - 1.0 * c3) * F3(scl, ij, k, l)
                                 + (-1.0 * c1 + 2.0 * c2 - 1.0 * c3) * F3(scl, ip1jp1, k, l)
                                 + (-1.0 * c1 - 1.0 * c2 + 2.0 * c3) * F3(scl, ijp1, k, l)) / 3.0;
                }
            }
            if (have_sgp[l - 1]) /* :128-132 */
                for (int d = 1; d <= 3; ++d) VT(gminm1, d, 0) = VT(gminm1 + 1, d, 1);
            for (int g = 1; g <= gmin - 1; ++g) F3(dscl, g, k, l) = 0.0;
            /* :138-216: the three passes; the Y and Z passes read the X
             * component at the first point of each pair, as the reference does */
            for (int d = 1; d <= 3; ++d)
                for (int g = gmin; g <= gmax; ++g) {
                    const int ij = g, ip1j = g + 1, ip1jp1 = g + iall + 1, ijp1 = g + iall;
                    const int im1j = g - 1, im1jm1 = g - iall - 1, ijm1 = g - iall;
                    const double kf1 = 0.5 * (F3(kh, ij, k, l) + F3(kh, ip1jp1, k, l));
                    const double kf2 = 0.5 * (F3(kh, ij, k, l) + F3(kh, ijp1, k, l));
                    const double kf3 = 0.5 * (F3(kh, im1j, k, l) + F3(kh, ij, k, l));
                    const double kf4 = 0.5 * (F3(kh, im1jm1, k, l) + F3(kh, ij, k, l));
                    const double kf5 = 0.5 * (F3(kh, ijm1, k, l) + F3(kh, ij, k, l));
                    const double kf6 = 0.5 * (F3(kh, ij, k, l) + F3(kh, ip1j, k, l));
                    const double s = kf1 * CD(g, 1, d, l) * (VT(ij, 1, 0) + VT(ij, d, 1))
                                   + kf2 * CD(g, 2, d, l) * (VT(ij, 1, 1) + VT(im1j, d, 0))
                                   + kf3 * CD(g, 3, d, l) * (VT(im1j, 1, 0) + VT(im1jm1, d, 1))
                                   + kf4 * CD(g, 4, d, l) * (VT(im1jm1, 1, 1) + VT(im1jm1, d, 0))
                                   + kf5 * CD(g, 5, d, l) * (VT(im1jm1, 1, 0) + VT(ijm1, d, 1))
                                   + kf6 * CD(g, 6, d, l) * (VT(ijm1, 1, 1) + VT(ij, d, 0));
                    if (d == 1) F3(dscl, g, k, l) = s;
                    else F3(dscl, g, k, l) = F3(dscl, g, k, l) + s;
                }
            for (int g = gmax + 1; g <= gall; ++g) F3(dscl, g, k, l) = 0.0;
#undef VT
            free(vt);
        }
    }

    if (D->have_pl) { /* :220-291 */
        const int n = GSLF_PL, gmin_pl = GMIN_PL, gmax_pl = gall_pl;
        double vt_pl[16][3];
        for (int l = 1; l <= LALL_PL; ++l)
            for (int k = 1; k <= kall; ++k) {
                for (int d = 1; d <= 3; ++d)
                    for (int v = gmin_pl; v <= gmax_pl; ++v) {
                        const int ij = v;
                        int ijp1 = v + 1;
                        if (ijp1 == gmax_pl + 1) ijp1 = gmin_pl;
                        const double c1 = CI_PL(v, 1, d, l), c2 = CI_PL(v, 2, d, l), c3 = CI_PL(v, 3, d, l);
                        vt_pl[ij][d - 1] = ((+2.0 * c1 - 1.0 * c2 - 1.0 * c3) * P3(scl_pl, n, k, l)
                                          + (-1.0 * c1 + 2.0 * c2 - 1.0 * c3) * P3(scl_pl, ij, k, l)
                                          + (-1.0 * c1 - 1.0 * c2 + 2.0 * c3) * P3(scl_pl, ijp1, k, l)) / 3.0;
                    }
                for (int v = 1; v <= gall_pl; ++v) P3(dscl_pl, v, k, l) = 0.0;
                for (int v = gmin_pl; v <= gmax_pl; ++v) {
                    const int ij = v;
                    int ijm1 = v - 1;
                    if (ijm1 == gmin_pl - 1) ijm1 = gmax_pl;
                    P3(dscl_pl, n, k, l) = P3(dscl_pl, n, k, l)
                        + (CD_PL(v - 1, 1, l) * (vt_pl[ijm1][0] + vt_pl[ij][0])
                         + CD_PL(v - 1, 2, l) * (vt_pl[ijm1][1] + vt_pl[ij][1])
                         + CD_PL(v - 1, 3, l) * (vt_pl[ijm1][2] + vt_pl[ij][2]))
                        * 0.5 * (P3(kh_pl, n, k, l) + P3(kh_pl, ij, k, l));
                }
            }
    } else {
        for (size_t i = 0; i < (size_t)gall_pl * kall * LALL_PL; ++i) dscl_pl[i] = 0.0;
    }
#undef CI
#undef CD
#undef CI_PL
#undef CD_PL
}

/* ------------------------------------------------------------------------- */
/* dyn_divdamp/src/mod_oprt3d.f90:41-402                                     */
/* ------------------------------------------------------------------------- */
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
                        const double *rdgz)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall;
    const int gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    const int gmin = (2 - 1) * iall + 2, gmax = (iall - 1 - 1) * iall + iall - 1;
    const int gminm1 = (2 - 1 - 1) * iall + 2 - 1;
#define CI(g, c, d, t, l) coef_intp[((((size_t)((l) - 1) * 2 + (t)) * 3 + ((d) - 1)) * 3 + ((c) - 1)) * gall + ((g) - 1)]
#define CD(g, m, d, l) coef_diff[(((size_t)((l) - 1) * 3 + ((d) - 1)) * 6 + ((m) - 1)) * gall + ((g) - 1)]
#define CI_PL(v, c, d, l) coef_intp_pl[((((size_t)((l) - 1)) * 3 + ((d) - 1)) * 3 + ((c) - 1)) * gall_pl + ((v) - 1)]
#define CD_PL(m, d, l) coef_diff_pl[(((size_t)((l) - 1)) * 3 + ((d) - 1)) * 5 + ((m) - 1)]
#define C2W(g, k, m, l) C2WfactGz[(((size_t)((l) - 1) * 6 + ((m) - 1)) * kall + ((k) - 1)) * gall + ((g) - 1)]
#define C2W_PL(g, k, m, l) C2WfactGz_pl[(((size_t)((l) - 1) * 6 + ((m) - 1)) * kall + ((k) - 1)) * gall_pl + ((g) - 1)]

    double *wvm = (double *)malloc(sizeof(double) * (size_t)gall * kall); /* rhogw_vm of one region */
#define WVM(g, k) wvm[((size_t)((k) - 1)) * gall + ((g) - 1)]
    for (int l = 1; l <= lall; ++l) {
        for (int k = kmin + 1; k <= kmax; ++k) /* :143-157 */
            for (int g = 1; g <= gall; ++g)
                WVM(g, k) = (C2W(g, k, 1, l) * F3(rhogvx, g, k, l) + C2W(g, k, 2, l) * F3(rhogvx, g, k - 1, l)
                           + C2W(g, k, 3, l) * F3(rhogvy, g, k, l) + C2W(g, k, 4, l) * F3(rhogvy, g, k - 1, l)
                           + C2W(g, k, 5, l) * F3(rhogvz, g, k, l) + C2W(g, k, 6, l) * F3(rhogvz, g, k - 1, l))
                          * F3(RGAMH, g, k, l)
                          + F3(rhogw, g, k, l) * F3(RGSQRTH, g, k, l);
        for (int g = 1; g <= gall; ++g) {
            WVM(g, kmin) = 0.0;
            WVM(g, kmax + 1) = 0.0;
        }
#pragma omp parallel for schedule(static)
        for (int k = kmin; k <= kmax; ++k) {
            double *buf = (double *)malloc(sizeof(double) * (size_t)gall * 5);
            double *vxm = buf, *vym = buf + gall, *vzm = buf + 2 * (size_t)gall, *st = buf + 3 * (size_t)gall;
#define ST(g, t) st[(size_t)(t) * gall + ((g) - 1)]
            for (int g = 1; g <= gall; ++g) { /* :176-181 */
                vxm[g - 1] = F3(rhogvx, g, k, l) * F3(RGAM, g, k, l);
                vym[g - 1] = F3(rhogvy, g, k, l) * F3(RGAM, g, k, l);
                vzm[g - 1] = F3(rhogvz, g, k, l) * F3(RGAM, g, k, l);
            }
            for (int g = gminm1; g <= gmax; ++g) { /* :183-203 */
                const int ij = g, ip1j = g + 1, ip1jp1 = g + iall + 1;
                const double sw = ((WVM(ij, k + 1) + WVM(ip1j, k + 1) + WVM(ip1jp1, k + 1))
                                 - (WVM(ij, k) + WVM(ip1j, k) + WVM(ip1jp1, k))) / 3.0 * rdgz[k - 1];
                ST(g, 0) = CI(g, 1, 1, 0, l) * vxm[ij - 1] + CI(g, 2, 1, 0, l) * vxm[ip1j - 1]
                         + CI(g, 3, 1, 0, l) * vxm[ip1jp1 - 1] + CI(g, 1, 2, 0, l) * vym[ij - 1]
                         + CI(g, 2, 2, 0, l) * vym[ip1j - 1] + CI(g, 3, 2, 0, l) * vym[ip1jp1 - 1]
                         + CI(g, 1, 3, 0, l) * vzm[ij - 1] + CI(g, 2, 3, 0, l) * vzm[ip1j - 1]
                         + CI(g, 3, 3, 0, l) * vzm[ip1jp1 - 1] + sw;
            }
            for (int g = gminm1; g <= gmax; ++g) { /* :205-225 */
                const int ij = g, ip1jp1 = g + iall + 1, ijp1 = g + iall;
                const double sw = ((WVM(ij, k + 1) + WVM(ip1jp1, k + 1) + WVM(ijp1, k + 1))
                                 - (WVM(ij, k) + WVM(ip1jp1, k) + WVM(ijp1, k))) / 3.0 * rdgz[k - 1];
                ST(g, 1) = CI(g, 1, 1, 1, l) * vxm[ij - 1] + CI(g, 2, 1, 1, l) * vxm[ip1jp1 - 1]
                         + CI(g, 3, 1, 1, l) * vxm[ijp1 - 1] + CI(g, 1, 2, 1, l) * vym[ij - 1]
                         + CI(g, 2, 2, 1, l) * vym[ip1jp1 - 1] + CI(g, 3, 2, 1, l) * vym[ijp1 - 1]
                         + CI(g, 1, 3, 1, l) * vzm[ij - 1] + CI(g, 2, 3, 1, l) * vzm[ip1jp1 - 1]
                         + CI(g, 3, 3, 1, l) * vzm[ijp1 - 1] + sw;
            }
            if (have_sgp[l - 1]) ST(gminm1, 0) = ST(gminm1 + 1, 1); /* :227-229 */
            double *out[3] = {ddivdx, ddivdy, ddivdz};
            for (int d = 0; d < 3; ++d) {
                for (int g = 1; g <= gmin - 1; ++g) F3(out[d], g, k, l) = 0.0;
                for (int g = gmin; g <= gmax; ++g) { /* :237-282 */
                    const int ij = g, im1j = g - 1, im1jm1 = g - iall - 1, ijm1 = g - iall;
                    F3(out[d], g, k, l) = CD(g, 1, d + 1, l) * (ST(ij, 0) + ST(ij, 1))
                                        + CD(g, 2, d + 1, l) * (ST(ij, 1) + ST(im1j, 0))
                                        + CD(g, 3, d + 1, l) * (ST(im1j, 0) + ST(im1jm1, 1))
                                        + CD(g, 4, d + 1, l) * (ST(im1jm1, 1) + ST(im1jm1, 0))
                                        + CD(g, 5, d + 1, l) * (ST(im1jm1, 0) + ST(ijm1, 1))
                                        + CD(g, 6, d + 1, l) * (ST(ijm1, 1) + ST(ij, 0));
                }
                for (int g = gmax + 1; g <= gall; ++g) F3(out[d], g, k, l) = 0.0;
            }
#undef ST
            free(buf);
        }
        for (int g = 1; g <= gall; ++g) { /* :303-310 */
            F3(ddivdx, g, kmin - 1, l) = 0.0;
            F3(ddivdy, g, kmin - 1, l) = 0.0;
            F3(ddivdz, g, kmin - 1, l) = 0.0;
            F3(ddivdx, g, kmax + 1, l) = 0.0;
            F3(ddivdy, g, kmax + 1, l) = 0.0;
            F3(ddivdz, g, kmax + 1, l) = 0.0;
        }
    }
#undef WVM
    free(wvm);

    if (D->have_pl) { /* :313-392 */
        const int n = GSLF_PL, gmin_pl = GMIN_PL, gmax_pl = gall_pl;
        double *wpl = (double *)malloc(sizeof(double) * (size_t)gall_pl * kall);
#define WPL(g, k) wpl[((size_t)((k) - 1)) * gall_pl + ((g) - 1)]
        for (int l = 1; l <= LALL_PL; ++l) {
            for (int k = kmin + 1; k <= kmax; ++k)
                for (int g = 1; g <= gall_pl; ++g)
                    WPL(g, k) = (C2W_PL(g, k, 1, l) * P3(rhogvx_pl, g, k, l) + C2W_PL(g, k, 2, l) * P3(rhogvx_pl, g, k - 1, l)
                               + C2W_PL(g, k, 3, l) * P3(rhogvy_pl, g, k, l) + C2W_PL(g, k, 4, l) * P3(rhogvy_pl, g, k - 1, l)
                               + C2W_PL(g, k, 5, l) * P3(rhogvz_pl, g, k, l) + C2W_PL(g, k, 6, l) * P3(rhogvz_pl, g, k - 1, l))
                              * P3(RGAMH_pl, g, k, l)
                              + P3(rhogw_pl, g, k, l) * P3(RGSQRTH_pl, g, k, l);
            for (int g = 1; g <= gall_pl; ++g) {
                WPL(g, kmin) = 0.0;
                WPL(g, kmax + 1) = 0.0;
            }
            for (int k = kmin; k <= kmax; ++k) {
                double vx[16], vy[16], vz[16], sp[16];
                for (int v = 1; v <= gall_pl; ++v) {
                    vx[v] = P3(rhogvx_pl, v, k, l) * P3(RGAM_pl, v, k, l);
                    vy[v] = P3(rhogvy_pl, v, k, l) * P3(RGAM_pl, v, k, l);
                    vz[v] = P3(rhogvz_pl, v, k, l) * P3(RGAM_pl, v, k, l);
                }
                for (int v = gmin_pl; v <= gmax_pl; ++v) {
                    const int ij = v;
                    int ijp1 = v + 1;
                    if (ijp1 == gmax_pl + 1) ijp1 = gmin_pl;
                    const double sw = ((WPL(n, k + 1) + WPL(ij, k + 1) + WPL(ijp1, k + 1))
                                     - (WPL(n, k) + WPL(ij, k) + WPL(ijp1, k))) / 3.0 * rdgz[k - 1];
                    sp[ij] = CI_PL(v, 1, 1, l) * vx[n] + CI_PL(v, 2, 1, l) * vx[ij] + CI_PL(v, 3, 1, l) * vx[ijp1]
                           + CI_PL(v, 1, 2, l) * vy[n] + CI_PL(v, 2, 2, l) * vy[ij] + CI_PL(v, 3, 2, l) * vy[ijp1]
                           + CI_PL(v, 1, 3, l) * vz[n] + CI_PL(v, 2, 3, l) * vz[ij] + CI_PL(v, 3, 3, l) * vz[ijp1]
                           + sw;
                }
                for (int v = 1; v <= gall_pl; ++v) {
                    P3(ddivdx_pl, v, k, l) = 0.0;
                    P3(ddivdy_pl, v, k, l) = 0.0;
                    P3(ddivdz_pl, v, k, l) = 0.0;
                }
                for (int v = gmin_pl; v <= gmax_pl; ++v) {
                    const int ij = v;
                    int ijm1 = v - 1;
                    if (ijm1 == gmin_pl - 1) ijm1 = gmax_pl;
                    P3(ddivdx_pl, n, k, l) = P3(ddivdx_pl, n, k, l) + CD_PL(v - 1, 1, l) * (sp[ijm1] + sp[ij]);
                    P3(ddivdy_pl, n, k, l) = P3(ddivdy_pl, n, k, l) + CD_PL(v - 1, 2, l) * (sp[ijm1] + sp[ij]);
                    P3(ddivdz_pl, n, k, l) = P3(ddivdz_pl, n, k, l) + CD_PL(v - 1, 3, l) * (sp[ijm1] + sp[ij]);
                }
            }
            for (int v = 1; v <= gall_pl; ++v) {
                P3(ddivdx_pl, v, kmin - 1, l) = 0.0;
                P3(ddivdx_pl, v, kmax + 1, l) = 0.0;
                P3(ddivdy_pl, v, kmin - 1, l) = 0.0;
                P3(ddivdy_pl, v, kmax + 1, l) = 0.0;
                P3(ddivdz_pl, v, kmin - 1, l) = 0.0;
                P3(ddivdz_pl, v, kmax + 1, l) = 0.0;
            }
        }
#undef WPL
        free(wpl);
    } else {
        for (size_t i = 0; i < (size_t)gall_pl * kall * LALL_PL; ++i) {
            ddivdx_pl[i] = 0.0;
            ddivdy_pl[i] = 0.0;
            ddivdz_pl[i] = 0.0;
        }
    }
#undef CI
#undef CD
#undef CI_PL
#undef CD_PL
#undef C2W
#undef C2W_PL
}

/* ---------------------------------------------------------------------------
 * Step glue.  The reference has none of this code (its kernels take ch,
 * cmask, q_a and ck as inputs, which its drivers describe as computed "in
 * advance by other subroutine"); these are the NICAM forms of the Miura upwind scheme
 * (docs/NICAM/src/20_intro_nicam.tex:368-438) that the library builds, and
 * this restatement defines them for the parity tests.
 * ------------------------------------------------------------------------- */
#define CG(g, m, d, l) coef_grad[((((size_t)((l) - 1) * 3 + ((d) - 1)) * 7 + (m)) * gall + ((g) - 1))]
#define CG_PL(m, d, l) coef_grad_pl[(((size_t)((l) - 1) * 3 + ((d) - 1)) * gall_pl + (m))]
#define GX(g, l, d) GRD_x[((size_t)((d) - 1) * lall + ((l) - 1)) * gall + ((g) - 1)]
#define GX_PL(v, l, d) GRD_x_pl[((size_t)((d) - 1) * LALL_PL + ((l) - 1)) * gall_pl + ((v) - 1)]

void orc_oprt_gradient(const orc_dims *D, double *grad, double *grad_pl, const double *scl, const double *scl_pl,
                       const double *coef_grad, const double *coef_grad_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int gmin = (2 - 1) * iall + 2, gmax = (iall - 1 - 1) * iall + iall - 1;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int d = 1; d <= 3; ++d) {
                for (int g = 1; g <= gall; ++g) F4(grad, g, k, l, d) = 0.0;
                for (int g = gmin; g <= gmax; ++g) {
                    const int ij = g, ip1j = g + 1, ip1jp1 = g + iall + 1, ijp1 = g + iall;
                    const int im1j = g - 1, im1jm1 = g - iall - 1, ijm1 = g - iall;
                    F4(grad, g, k, l, d) = CG(g, 0, d, l) * F3(scl, ij, k, l) + CG(g, 1, d, l) * F3(scl, ip1j, k, l) +
                                           CG(g, 2, d, l) * F3(scl, ip1jp1, k, l) + CG(g, 3, d, l) * F3(scl, ijp1, k, l) +
                                           CG(g, 4, d, l) * F3(scl, im1j, k, l) + CG(g, 5, d, l) * F3(scl, im1jm1, k, l) +
                                           CG(g, 6, d, l) * F3(scl, ijm1, k, l);
                }
            }
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int d = 1; d <= 3; ++d) {
                for (int v = 1; v <= gall_pl; ++v) P4(grad_pl, v, k, l, d) = 0.0;
                if (!D->have_pl) continue;
                double s = 0.0;
                for (int v = 1; v <= gall_pl; ++v) s = s + CG_PL(v - 1, d, l) * P3(scl_pl, v, k, l);
                P4(grad_pl, GSLF_PL, k, l, d) = s;
            }
}

void orc_tracer_courant(const orc_dims *D, double *ch, double *ch_pl, double *cmask, double *cmask_pl,
                        const double *flx_h, const double *flx_h_pl, const double *rhog, const double *rhog_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const double eps = D->eps;
    for (int m = 1; m <= 6; ++m)
        for (int l = 1; l <= lall; ++l)
            for (int k = 1; k <= kall; ++k)
                for (int g = 1; g <= gall; ++g) {
                    const double c = F4(flx_h, g, k, l, m) / F3(rhog, g, k, l);
                    F4(ch, g, k, l, m) = c;
                    F4(cmask, g, k, l, m) = 0.5 - copysign(0.5, c - eps);
                }
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k) {
            P3(ch_pl, GSLF_PL, k, l) = 0.0;
            P3(cmask_pl, GSLF_PL, k, l) = 0.0;
            for (int v = GMIN_PL; v <= gall_pl; ++v) {
                const double c = D->have_pl ? P3(flx_h_pl, v, k, l) / P3(rhog_pl, GSLF_PL, k, l) : 0.0;
                P3(ch_pl, v, k, l) = c;
                P3(cmask_pl, v, k, l) = D->have_pl ? 0.5 - copysign(0.5, c - eps) : 0.0;
            }
        }
}

void orc_horizontal_remap(const orc_dims *D, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                          const double *grad, const double *grad_pl, const double *cmask, const double *cmask_pl,
                          const double *GRD_xc, const double *GRD_xc_pl, const double *GRD_x,
                          const double *GRD_x_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int j = 1; j <= iall - 1; ++j)
                for (int i = 1; i <= iall - 1; ++i) {
                    const int ij = (j - 1) * iall + i;
                    const int nbs[3] = {ij + 1, ij + iall + 1, ij + iall};  /* AI, AIJ, AJ */
                    for (int a = 1; a <= 3; ++a) {
                        const int nb = nbs[a - 1];
                        const double x1 = F5(GRD_xc, ij, k, l, a, 1), x2 = F5(GRD_xc, ij, k, l, a, 2),
                                     x3 = F5(GRD_xc, ij, k, l, a, 3);
                        const double qs = F3(q, ij, k, l) + (F4(grad, ij, k, l, 1) * (x1 - GX(ij, l, 1)) +
                                                             F4(grad, ij, k, l, 2) * (x2 - GX(ij, l, 2)) +
                                                             F4(grad, ij, k, l, 3) * (x3 - GX(ij, l, 3)));
                        const double qn = F3(q, nb, k, l) + (F4(grad, nb, k, l, 1) * (x1 - GX(nb, l, 1)) +
                                                             F4(grad, nb, k, l, 2) * (x2 - GX(nb, l, 2)) +
                                                             F4(grad, nb, k, l, 3) * (x3 - GX(nb, l, 3)));
                        /* each face with its owner's mask: the origin's face a, the neighbour's face a+3 */
                        const double cs = F4(cmask, ij, k, l, a), cn = F4(cmask, nb, k, l, a + 3);
                        F4(q_a, ij, k, l, a) = cs * qn + (1.0 - cs) * qs;
                        F4(q_a, nb, k, l, a + 3) = cn * qs + (1.0 - cn) * qn;
                    }
                }
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int v = GMIN_PL; v <= gall_pl; ++v) {
                const int n = GSLF_PL;
                const double x1 = P4(GRD_xc_pl, v, k, l, 1), x2 = P4(GRD_xc_pl, v, k, l, 2),
                             x3 = P4(GRD_xc_pl, v, k, l, 3);
                const double qs = P3(q_pl, n, k, l) + (P4(grad_pl, n, k, l, 1) * (x1 - GX_PL(n, l, 1)) +
                                                       P4(grad_pl, n, k, l, 2) * (x2 - GX_PL(n, l, 2)) +
                                                       P4(grad_pl, n, k, l, 3) * (x3 - GX_PL(n, l, 3)));
                const double qn = P3(q_pl, v, k, l) + (P4(grad_pl, v, k, l, 1) * (x1 - GX_PL(v, l, 1)) +
                                                       P4(grad_pl, v, k, l, 2) * (x2 - GX_PL(v, l, 2)) +
                                                       P4(grad_pl, v, k, l, 3) * (x3 - GX_PL(v, l, 3)));
                const double c = P3(cmask_pl, v, k, l);
                P3(q_a_pl, v, k, l) = c * qn + (1.0 - c) * qs;
            }
}

void orc_tracer_update(const orc_dims *D, double *rhogq, double *rhogq_pl, const double *flx_h,
                       const double *flx_h_pl, const double *q_a, const double *q_a_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int j = 2; j <= iall - 1; ++j)
                for (int i = 2; i <= iall - 1; ++i) {
                    const int g = (j - 1) * iall + i;
                    F3(rhogq, g, k, l) =
                        F3(rhogq, g, k, l) -
                        (F4(flx_h, g, k, l, 1) * F4(q_a, g, k, l, 1) + F4(flx_h, g, k, l, 2) * F4(q_a, g, k, l, 2) +
                         F4(flx_h, g, k, l, 3) * F4(q_a, g, k, l, 3) + F4(flx_h, g, k, l, 4) * F4(q_a, g, k, l, 4) +
                         F4(flx_h, g, k, l, 5) * F4(q_a, g, k, l, 5) + F4(flx_h, g, k, l, 6) * F4(q_a, g, k, l, 6));
                }
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k) {
            double s = P3(flx_h_pl, GMIN_PL, k, l) * P3(q_a_pl, GMIN_PL, k, l);
            for (int v = GMIN_PL + 1; v <= gall_pl; ++v) s = s + P3(flx_h_pl, v, k, l) * P3(q_a_pl, v, k, l);
            P3(rhogq_pl, GSLF_PL, k, l) = P3(rhogq_pl, GSLF_PL, k, l) - s;
        }
}
/* Adjoints of the horizontal glue, for the differentiable step.  Each
 * overwrites its outputs with the transpose of its forward's derivative
 * applied to the given cotangents, summed in the order written here (the
 * order the GPU kernels reproduce).  Parity unpinned against the reference
 * (it has no glue); pinned by dot-product tests against the forwards above. */

/* scl_b(g) = sum over d = 1..3, m = 0..6 of coef_grad(g', m, d) grad_b(g', d),
 * g' = g - offset(m) in [gmin, gmax]; pole: scl_b_pl(v) = sum over d of
 * coef_grad_pl(v-1, d) grad_b_pl(centre, d). */
void orc_oprt_gradient_adjoint(const orc_dims *D, double *scl_b, double *scl_b_pl, const double *grad_b,
                               const double *grad_b_pl, const double *coef_grad, const double *coef_grad_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int gmin = (2 - 1) * iall + 2, gmax = (iall - 1 - 1) * iall + iall - 1;
    const int off[7] = {0, 1, iall + 1, iall, -1, -iall - 1, -iall};
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int g = 1; g <= gall; ++g) {
                double s = 0.0;
                for (int d = 1; d <= 3; ++d)
                    for (int m = 0; m < 7; ++m) {
                        const int gp = g - off[m];
                        if (gp >= gmin && gp <= gmax) s = s + CG(gp, m, d, l) * F4(grad_b, gp, k, l, d);
                    }
                F3(scl_b, g, k, l) = s;
            }
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int v = 1; v <= gall_pl; ++v) {
                double s = 0.0;
                if (D->have_pl)
                    for (int d = 1; d <= 3; ++d) s = s + CG_PL(v - 1, d, l) * P4(grad_b_pl, GSLF_PL, k, l, d);
                P3(scl_b_pl, v, k, l) = s;
            }
}

/* flx_b = ch_b / rhog; rhog_b = -(sum over faces of ch_b * (flx/rhog)) / rhog;
 * the pole ring against the centre's rhog_pl (cmask is piecewise constant). */
void orc_tracer_courant_adjoint(const orc_dims *D, double *flx_b, double *flx_b_pl, double *rhog_b, double *rhog_b_pl,
                                const double *ch_b, const double *ch_b_pl, const double *flx_h, const double *flx_h_pl,
                                const double *rhog, const double *rhog_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int g = 1; g <= gall; ++g) {
                const double r = F3(rhog, g, k, l);
                double t = 0.0;
                for (int m = 1; m <= 6; ++m) {
                    const double b = F4(ch_b, g, k, l, m);
                    F4(flx_b, g, k, l, m) = b / r;
                    t = t + b * (F4(flx_h, g, k, l, m) / r);
                }
                F3(rhog_b, g, k, l) = -t / r;
            }
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k) {
            for (int v = 1; v <= gall_pl; ++v) {
                P3(flx_b_pl, v, k, l) = 0.0;
                P3(rhog_b_pl, v, k, l) = 0.0;
            }
            if (!D->have_pl) continue;
            const double r = P3(rhog_pl, GSLF_PL, k, l);
            double t = 0.0;
            for (int v = GMIN_PL; v <= gall_pl; ++v) {
                const double b = P3(ch_b_pl, v, k, l);
                P3(flx_b_pl, v, k, l) = b / r;
                t = t + b * (P3(flx_h_pl, v, k, l) / r);
            }
            P3(rhog_b_pl, GSLF_PL, k, l) = -t / r;
        }
}

/* The transpose of orc_horizontal_remap in q, grad and GRD_xc (cmask fixed):
 * for every arc, in the forward's loop order, the cotangents of its two faces
 * fa (the origin's face a) and fb (the neighbour's face a+3) give those of its
 * two nodes, qs_b = (1-cs) fa + cn fb and qn_b = cs fa + (1-cn) fb, which are
 * added to q_b and grad_b of the origin and the neighbour; GRD_xc_b of the arc
 * is qs_b grad(origin) + qn_b grad(neighbour).  Other entries 0. */
void orc_horizontal_remap_adjoint(const orc_dims *D, double *q_b, double *q_b_pl, double *grad_b, double *grad_b_pl,
                                  double *xc_b, double *xc_b_pl, const double *qa_b, const double *qa_b_pl,
                                  const double *grad, const double *grad_pl, const double *cmask,
                                  const double *cmask_pl, const double *GRD_xc, const double *GRD_xc_pl,
                                  const double *GRD_x, const double *GRD_x_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const size_t n3 = (size_t)gall * kall * lall, n3p = (size_t)gall_pl * kall * LALL_PL;
    for (size_t x = 0; x < n3; ++x) q_b[x] = 0.0;
    for (size_t x = 0; x < 3 * n3; ++x) grad_b[x] = 0.0;
    for (size_t x = 0; x < 9 * n3; ++x) xc_b[x] = 0.0;
    for (size_t x = 0; x < n3p; ++x) q_b_pl[x] = 0.0;
    for (size_t x = 0; x < 3 * n3p; ++x) grad_b_pl[x] = 0.0;
    for (size_t x = 0; x < 3 * n3p; ++x) xc_b_pl[x] = 0.0;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int j = 1; j <= iall - 1; ++j)
                for (int i = 1; i <= iall - 1; ++i) {
                    const int ij = (j - 1) * iall + i;
                    const int nbs[3] = {ij + 1, ij + iall + 1, ij + iall};
                    for (int a = 1; a <= 3; ++a) {
                        const int nb = nbs[a - 1];
                        const double fa = F4(qa_b, ij, k, l, a), fb = F4(qa_b, nb, k, l, a + 3);
                        const double cs = F4(cmask, ij, k, l, a), cn = F4(cmask, nb, k, l, a + 3);
                        const double qs_b = (1.0 - cs) * fa + cn * fb;
                        const double qn_b = cs * fa + (1.0 - cn) * fb;
                        F3(q_b, ij, k, l) = F3(q_b, ij, k, l) + qs_b;
                        F3(q_b, nb, k, l) = F3(q_b, nb, k, l) + qn_b;
                        for (int d = 1; d <= 3; ++d) {
                            const double x = F5(GRD_xc, ij, k, l, a, d);
                            F4(grad_b, ij, k, l, d) = F4(grad_b, ij, k, l, d) + qs_b * (x - GX(ij, l, d));
                            F4(grad_b, nb, k, l, d) = F4(grad_b, nb, k, l, d) + qn_b * (x - GX(nb, l, d));
                            F5(xc_b, ij, k, l, a, d) = qs_b * F4(grad, ij, k, l, d) + qn_b * F4(grad, nb, k, l, d);
                        }
                    }
                }
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int v = GMIN_PL; v <= gall_pl; ++v) {
                const int n = GSLF_PL;
                const double f = P3(qa_b_pl, v, k, l), c = P3(cmask_pl, v, k, l);
                const double qs_b = (1.0 - c) * f, qn_b = c * f;
                P3(q_b_pl, n, k, l) = P3(q_b_pl, n, k, l) + qs_b;
                P3(q_b_pl, v, k, l) = P3(q_b_pl, v, k, l) + qn_b;
                for (int d = 1; d <= 3; ++d) {
                    const double x = P4(GRD_xc_pl, v, k, l, d);
                    P4(grad_b_pl, n, k, l, d) = P4(grad_b_pl, n, k, l, d) + qs_b * (x - GX_PL(n, l, d));
                    P4(grad_b_pl, v, k, l, d) = P4(grad_b_pl, v, k, l, d) + qn_b * (x - GX_PL(v, l, d));
                    P4(xc_b_pl, v, k, l, d) = qs_b * P4(grad_pl, n, k, l, d) + qn_b * P4(grad_pl, v, k, l, d);
                }
            }
}

/* The transpose of orc_tracer_update in flx_h and q_a (rhogq's cotangent passes
 * through unchanged): on the inner cells flx_b = -(rhogq_b q_a), q_a_b =
 * -(rhogq_b flx_h), 0 elsewhere; the pole ring against the centre's rhogq_b. */
void orc_tracer_update_adjoint(const orc_dims *D, double *flx_b, double *flx_b_pl, double *qa_b, double *qa_b_pl,
                               const double *rhogq_b, const double *rhogq_b_pl, const double *flx_h,
                               const double *flx_h_pl, const double *q_a, const double *q_a_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    for (int m = 1; m <= 6; ++m)
        for (int l = 1; l <= lall; ++l)
            for (int k = 1; k <= kall; ++k)
                for (int j = 1; j <= iall; ++j)
                    for (int i = 1; i <= iall; ++i) {
                        const int g = (j - 1) * iall + i;
                        const int inner = i >= 2 && i <= iall - 1 && j >= 2 && j <= iall - 1;
                        const double b = F3(rhogq_b, g, k, l);
                        F4(flx_b, g, k, l, m) = inner ? -(b * F4(q_a, g, k, l, m)) : 0.0;
                        F4(qa_b, g, k, l, m) = inner ? -(b * F4(flx_h, g, k, l, m)) : 0.0;
                    }
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k) {
            const double b = D->have_pl ? P3(rhogq_b_pl, GSLF_PL, k, l) : 0.0;
            for (int v = 1; v <= gall_pl; ++v) {
                const int ring = D->have_pl && v >= GMIN_PL;
                P3(flx_b_pl, v, k, l) = ring ? -(b * P3(q_a_pl, v, k, l)) : 0.0;
                P3(qa_b_pl, v, k, l) = ring ? -(b * P3(flx_h_pl, v, k, l)) : 0.0;
            }
        }
}
#undef CG
#undef CG_PL
#undef GX
#undef GX_PL

/* ---------------------------------------------------------------------------
 * Vertical step glue (no reference code; the reference's vertical limiter
 * takes ck and q_h as inputs).  kmin = 2, kmax = kall-1.
 * ------------------------------------------------------------------------- */
#define C2W(g, k, c, l) C2WfactGz[((((size_t)(l) - 1) * 6 + ((c) - 1)) * kall + ((k) - 1)) * gall + ((g) - 1)]
#define C2W_PL(v, k, c, l) \
    C2WfactGz_pl[((((size_t)(l) - 1) * 6 + ((c) - 1)) * kall + ((k) - 1)) * gall_pl + ((v) - 1)]

/* rhogw_vm as dyn_divdamp/src/mod_oprt3d.f90:148-155 forms it (and :325-332
 * for the poles), times dt, on kmin+1..kmax; 0 on the other levels. */
void orc_vertical_mass_flux(const orc_dims *D, double *flx_v, double *flx_v_pl, const double *vx, const double *vx_pl,
                            const double *vy, const double *vy_pl, const double *vz, const double *vz_pl,
                            const double *w, const double *w_pl, const double *C2WfactGz,
                            const double *C2WfactGz_pl, const double *RGAMH, const double *RGAMH_pl,
                            const double *RGSQRTH, const double *RGSQRTH_pl, double dt)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int g = 1; g <= gall; ++g) {
                double f = 0.0;
                if (k >= kmin + 1 && k <= kmax) {
                    const double vm = (C2W(g, k, 1, l) * F3(vx, g, k, l) + C2W(g, k, 2, l) * F3(vx, g, k - 1, l) +
                                       C2W(g, k, 3, l) * F3(vy, g, k, l) + C2W(g, k, 4, l) * F3(vy, g, k - 1, l) +
                                       C2W(g, k, 5, l) * F3(vz, g, k, l) + C2W(g, k, 6, l) * F3(vz, g, k - 1, l)) *
                                          F3(RGAMH, g, k, l) +
                                      F3(w, g, k, l) * F3(RGSQRTH, g, k, l);
                    f = vm * dt;
                }
                F3(flx_v, g, k, l) = f;
            }
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int v = 1; v <= gall_pl; ++v) {
                double f = 0.0;
                if (k >= kmin + 1 && k <= kmax) {
                    const double vm =
                        (C2W_PL(v, k, 1, l) * P3(vx_pl, v, k, l) + C2W_PL(v, k, 2, l) * P3(vx_pl, v, k - 1, l) +
                         C2W_PL(v, k, 3, l) * P3(vy_pl, v, k, l) + C2W_PL(v, k, 4, l) * P3(vy_pl, v, k - 1, l) +
                         C2W_PL(v, k, 5, l) * P3(vz_pl, v, k, l) + C2W_PL(v, k, 6, l) * P3(vz_pl, v, k - 1, l)) *
                            P3(RGAMH_pl, v, k, l) +
                        P3(w_pl, v, k, l) * P3(RGSQRTH_pl, v, k, l);
                    f = vm * dt;
                }
                P3(flx_v_pl, v, k, l) = f;
            }
}
#undef C2W
#undef C2W_PL

/* ck(k,1) = -flx_v(k)/rhog(k)*rdgz(k), ck(k,2) = flx_v(k+1)/rhog(k)*rdgz(k)
 * on the interior faces of kmin..kmax, 0 elsewhere; the closed faces carry
 * +0 in ck(kmin,1) and -0 in ck(kmax+1,1), which the vertical limiter's flags
 * (0.5 - sign(0.5, ck(k,1)) below, 0.5 + sign(0.5, ck(k+1,1)) above) read as
 * outflow. */
void orc_vertical_courant(const orc_dims *D, const double *rdgz, double *ck, double *ck_pl, const double *flx_v,
                          const double *flx_v_pl, const double *rhog, const double *rhog_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    for (int l = 1; l <= lall; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int g = 1; g <= gall; ++g) {
                double lo = 0.0, up = 0.0;
                if (k >= kmin && k <= kmax) {
                    if (k > kmin) lo = -F3(flx_v, g, k, l) / F3(rhog, g, k, l) * rdgz[k - 1];
                    if (k < kmax) up = F3(flx_v, g, k + 1, l) / F3(rhog, g, k, l) * rdgz[k - 1];
                } else if (k == kmax + 1) {
                    lo = -0.0;
                }
                F4(ck, g, k, l, 1) = lo;
                F4(ck, g, k, l, 2) = up;
            }
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = 1; k <= kall; ++k)
            for (int v = 1; v <= gall_pl; ++v) {
                double lo = 0.0, up = 0.0;
                if (k >= kmin && k <= kmax) {
                    if (k > kmin) lo = -P3(flx_v_pl, v, k, l) / P3(rhog_pl, v, k, l) * rdgz[k - 1];
                    if (k < kmax) up = P3(flx_v_pl, v, k + 1, l) / P3(rhog_pl, v, k, l) * rdgz[k - 1];
                } else if (k == kmax + 1) {
                    lo = -0.0;
                }
                P4(ck_pl, v, k, l, 1) = lo;
                P4(ck_pl, v, k, l, 2) = up;
            }
}

/* q_h(k) = afact(k) q(k) + bfact(k) q(k-1) on kmin+1..kmax. */
void orc_vertical_face(const orc_dims *D, const double *afact, const double *bfact, double *q_h, double *q_h_pl,
                       const double *q, const double *q_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    for (int l = 1; l <= lall; ++l)
        for (int k = kmin + 1; k <= kmax; ++k)
            for (int g = 1; g <= gall; ++g)
                F3(q_h, g, k, l) = afact[k - 1] * F3(q, g, k, l) + bfact[k - 1] * F3(q, g, k - 1, l);
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = kmin + 1; k <= kmax; ++k)
            for (int v = 1; v <= gall_pl; ++v)
                P3(q_h_pl, v, k, l) = afact[k - 1] * P3(q_pl, v, k, l) + bfact[k - 1] * P3(q_pl, v, k - 1, l);
}

/* rhogq(k) -= (flx_v(k+1) q_h(k+1) - flx_v(k) q_h(k)) rdgz(k) on kmin..kmax,
 * no flux through the faces kmin and kmax+1. */
void orc_tracer_update_vertical(const orc_dims *D, const double *rdgz, double *rhogq, double *rhogq_pl,
                                const double *flx_v, const double *flx_v_pl, const double *q_h, const double *q_h_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    for (int l = 1; l <= lall; ++l)
        for (int k = kmin; k <= kmax; ++k)
            for (int g = 1; g <= gall; ++g) {
                const double up = (k + 1 <= kmax) ? F3(flx_v, g, k + 1, l) * F3(q_h, g, k + 1, l) : 0.0;
                const double lo = (k >= kmin + 1) ? F3(flx_v, g, k, l) * F3(q_h, g, k, l) : 0.0;
                F3(rhogq, g, k, l) = F3(rhogq, g, k, l) - (up - lo) * rdgz[k - 1];
            }
    if (!D->have_pl) return;
    for (int l = 1; l <= LALL_PL; ++l)
        for (int k = kmin; k <= kmax; ++k)
            for (int v = 1; v <= gall_pl; ++v) {
                const double up = (k + 1 <= kmax) ? P3(flx_v_pl, v, k + 1, l) * P3(q_h_pl, v, k + 1, l) : 0.0;
                const double lo = (k >= kmin + 1) ? P3(flx_v_pl, v, k, l) * P3(q_h_pl, v, k, l) : 0.0;
                P3(rhogq_pl, v, k, l) = P3(rhogq_pl, v, k, l) - (up - lo) * rdgz[k - 1];
            }
}

/* Adjoints of the three above, for the differentiable step (outputs
 * overwritten; parity unpinned against the reference, which has no glue;
 * pinned by dot-product tests against the forwards).  Sums in the order
 * written. */

/* flx_b(k) = lower face k's term (-ck_b(k,1)/rhog(k)*rdgz(k)) + upper face
 * k-1's (ck_b(k-1,2)/rhog(k-1)*rdgz(k-1)); rhog_b(k) = -(ck_b(k,1) ck(k,1) +
 * ck_b(k,2) ck(k,2))/rhog(k), over the interior faces the forward forms. */
void orc_vertical_courant_adjoint(const orc_dims *D, const double *rdgz, double *flx_b, double *flx_b_pl,
                                  double *rhog_b, double *rhog_b_pl, const double *ck_b, const double *ck_b_pl,
                                  const double *flx_v, const double *flx_v_pl, const double *rhog,
                                  const double *rhog_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
#define VCA(X3, X4, fb, rb, cb, fv, rh, g, l)                                                               \
    for (int k = 1; k <= kall; ++k) {                                                                        \
        double f = 0.0, t = 0.0;                                                                             \
        if (k > kmin && k <= kmax) f = f + -X4(cb, g, k, l, 1) / X3(rh, g, k, l) * rdgz[k - 1];              \
        if (k - 1 >= kmin && k - 1 < kmax) f = f + X4(cb, g, k - 1, l, 2) / X3(rh, g, k - 1, l) * rdgz[k - 2]; \
        if (k > kmin && k <= kmax) t = t + X4(cb, g, k, l, 1) * (-X3(fv, g, k, l) / X3(rh, g, k, l) * rdgz[k - 1]); \
        if (k >= kmin && k < kmax)                                                                           \
            t = t + X4(cb, g, k, l, 2) * (X3(fv, g, k + 1, l) / X3(rh, g, k, l) * rdgz[k - 1]);              \
        X3(fb, g, k, l) = f;                                                                                 \
        X3(rb, g, k, l) = -t / X3(rh, g, k, l);                                                              \
    }
    for (int l = 1; l <= lall; ++l)
        for (int g = 1; g <= gall; ++g) VCA(F3, F4, flx_b, rhog_b, ck_b, flx_v, rhog, g, l)
    for (int l = 1; l <= LALL_PL; ++l)
        for (int v = 1; v <= gall_pl; ++v) {
            if (D->have_pl) {
                VCA(P3, P4, flx_b_pl, rhog_b_pl, ck_b_pl, flx_v_pl, rhog_pl, v, l)
            } else {
                for (int k = 1; k <= kall; ++k) P3(flx_b_pl, v, k, l) = P3(rhog_b_pl, v, k, l) = 0.0;
            }
        }
#undef VCA
}

/* q_b(k) = afact(k) q_h_b(k) [k in kmin+1..kmax] + bfact(k+1) q_h_b(k+1)
 * [k+1 in kmin+1..kmax]. */
void orc_vertical_face_adjoint(const orc_dims *D, const double *afact, const double *bfact, double *q_b,
                               double *q_b_pl, const double *q_h_b, const double *q_h_b_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
#define VFA(X3, qb, hb, g, l)                                                                  \
    for (int k = 1; k <= kall; ++k) {                                                           \
        double s = 0.0;                                                                         \
        if (k >= kmin + 1 && k <= kmax) s = s + afact[k - 1] * X3(hb, g, k, l);                 \
        if (k + 1 >= kmin + 1 && k + 1 <= kmax) s = s + bfact[k] * X3(hb, g, k + 1, l);         \
        X3(qb, g, k, l) = s;                                                                    \
    }
    for (int l = 1; l <= lall; ++l)
        for (int g = 1; g <= gall; ++g) VFA(F3, q_b, q_h_b, g, l)
    for (int l = 1; l <= LALL_PL; ++l)
        for (int v = 1; v <= gall_pl; ++v) {
            if (D->have_pl) {
                VFA(P3, q_b_pl, q_h_b_pl, v, l)
            } else {
                for (int k = 1; k <= kall; ++k) P3(q_b_pl, v, k, l) = 0.0;
            }
        }
#undef VFA
}

/* For the interior faces k (kmin+1..kmax), Fb = rhogq_b(k) rdgz(k) -
 * rhogq_b(k-1) rdgz(k-1); flx_b(k) = Fb q_h(k), q_h_b(k) = Fb flx_v(k); other
 * levels 0 (rhogq's cotangent passes through unchanged). */
void orc_tracer_update_vertical_adjoint(const orc_dims *D, const double *rdgz, double *flx_b, double *flx_b_pl,
                                        double *q_h_b, double *q_h_b_pl, const double *rq_b, const double *rq_b_pl,
                                        const double *flx_v, const double *flx_v_pl, const double *q_h,
                                        const double *q_h_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
#define UVA(X3, fb, hb, rb, fv, qh, g, l)                                                      \
    for (int k = 1; k <= kall; ++k) {                                                           \
        double a = 0.0, b = 0.0;                                                                \
        if (k >= kmin + 1 && k <= kmax) {                                                       \
            const double Fb = X3(rb, g, k, l) * rdgz[k - 1] - X3(rb, g, k - 1, l) * rdgz[k - 2]; \
            a = Fb * X3(qh, g, k, l);                                                           \
            b = Fb * X3(fv, g, k, l);                                                           \
        }                                                                                       \
        X3(fb, g, k, l) = a;                                                                    \
        X3(hb, g, k, l) = b;                                                                    \
    }
    for (int l = 1; l <= lall; ++l)
        for (int g = 1; g <= gall; ++g) UVA(F3, flx_b, q_h_b, rq_b, flx_v, q_h, g, l)
    for (int l = 1; l <= LALL_PL; ++l)
        for (int v = 1; v <= gall_pl; ++v) {
            if (D->have_pl) {
                UVA(P3, flx_b_pl, q_h_b_pl, rq_b_pl, flx_v_pl, q_h_pl, v, l)
            } else {
                for (int k = 1; k <= kall; ++k) P3(flx_b_pl, v, k, l) = P3(q_h_b_pl, v, k, l) = 0.0;
            }
        }
#undef UVA
}

/* ---------------------------------------------------------------------------
 * Mc/Ml/Mu of the vertically implicit rho*w system (no reference code: the
 * reference reads them from its input file).  Collected from the discretised
 * Helmholtz equation, docs/NICAM/src/20_intro_nicam.tex:510-545, at the half
 * level k (1-based kmin+1..kmax) between full levels k-1 and k; the docs'
 * expressions for M^C and M^U (:585-596) agree with this collection, the
 * docs' M^L (:597-599) writes A^o_{k+1}/dxi_{k+1} where collecting the terms
 * of W_{k-1/2} gives A^o_k/dxi_k.
 * ------------------------------------------------------------------------- */
void orc_vi_rhow_update_matrix(const orc_dims *D, double *Mc, double *Mc_pl, double *Ml, double *Ml_pl, double *Mu,
                               double *Mu_pl, const double *eth, const double *eth_pl, const double *g_tilde,
                               const double *g_tilde_pl, const double *RGSGAM2, const double *RGSGAM2_pl,
                               const double *RGAMH, const double *RGAMH_pl, const double *GSGAM2H,
                               const double *GSGAM2H_pl, const double *rdgz, const double *rdgzh, const double *cfact,
                               const double *dfact, double dt, double grav, double rdry, double cvdry, int alpha_i)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, gall_pl = D->gall_pl;
    const int kmin = 2, kmax = kall - 1;
    const double cvovr = cvdry / rdry, cvovrt2 = cvdry / rdry / (dt * dt), alpha = (double)alpha_i;
    for (int part = 0; part < 2; ++part) {
        if (part == 1 && !D->have_pl) break;
        const int ng = part ? gall_pl : gall, nl = part ? LALL_PL : lall;
        double *mc = part ? Mc_pl : Mc, *ml = part ? Ml_pl : Ml, *mu = part ? Mu_pl : Mu;
        const double *h = part ? eth_pl : eth, *gt = part ? g_tilde_pl : g_tilde;
        const double *ao_ = part ? RGSGAM2_pl : RGSGAM2, *rh_ = part ? RGAMH_pl : RGAMH;
        const double *gs_ = part ? GSGAM2H_pl : GSGAM2H;
#define IX(g, k, l) (((size_t)((l) - 1) * kall + ((k) - 1)) * ng + ((g) - 1))
        for (int l = 1; l <= nl; ++l)
            for (int k = 1; k <= kall; ++k)
                for (int g = 1; g <= ng; ++g) {
                    double c = 0.0, lo = 0.0, up = 0.0;
                    if (k >= kmin + 1 && k <= kmax) {
                        const double rhm = rh_[IX(g, k - 1, l)], rh = rh_[IX(g, k, l)], rhp = rh_[IX(g, k + 1, l)];
                        const double g2m = 1.0 / (rhm * rhm), g2 = 1.0 / (rh * rh), g2p = 1.0 / (rhp * rhp);
                        const double aim = g2m * h[IX(g, k - 1, l)], ai = g2 * h[IX(g, k, l)],
                                     aip = g2p * h[IX(g, k + 1, l)];
                        const double co = cvovr * grav * rh * rh;
                        const double dd = cvovrt2 * gs_[IX(g, k, l)] * rh * rh;
                        const double ao = ao_[IX(g, k, l)], aom = ao_[IX(g, k - 1, l)];
                        const double rzh = rdgzh[k - 1];
                        c = alpha * dd + (ao * ai * rdgz[k - 1] + aom * ai * rdgz[k - 2]) * rzh -
                            (dfact[k - 1] - cfact[k - 2]) * (gt[IX(g, k, l)] + co * g2) * rzh;
                        up = -ao * aip * rdgz[k - 1] * rzh - cfact[k - 1] * (gt[IX(g, k + 1, l)] + co * g2p) * rzh;
                        lo = -aom * aim * rdgz[k - 2] * rzh + dfact[k - 2] * (gt[IX(g, k - 1, l)] + co * g2m) * rzh;
                    }
                    mc[IX(g, k, l)] = c;
                    ml[IX(g, k, l)] = lo;
                    mu[IX(g, k, l)] = up;
                }
#undef IX
    }
}

/* ------------------------------------------------------------------------- */
/* Adjoint of vi_rhow_solver (dyn_vi_rhow_solver/src/mod_vi.f90:93-306).     */
/* ------------------------------------------------------------------------- */
/* With the coefficients, metrics and dt fixed, the solver is a linear map of
 * (rhogw at kmin and kmax+1, rhogw0, preg0, rhog0, Srho, Sw, Spre) to rhogw on
 * kmin..kmax+1; rhogw(1) passes through and its interior input values are
 * overwritten (:209-246).  This is the transpose of that map, column by
 * column: the transposed tridiagonal system (diagonal Mc, sub-diagonal
 * Mu(k-1), super-diagonal Ml(k+1)) solved by the same elimination, then the
 * transpose of the source term Sall (:194-206) and of the boundary terms
 * (:209-212).  wb holds the output cotangent on entry and the input cotangent
 * on return; the six others are written on every level (zero where the
 * forward map does not read them).  The level loop below is the order of
 * operations the GPU kernel follows. */
static void vi_adjoint_column(int kall, size_t base, size_t s, double *wb, double *w0b, double *p0b, double *r0b,
                              double *Srb, double *Swb, double *Spb, const double *Mc, const double *Ml,
                              const double *Mu, const double *RGAM, const double *RGAMH, const double *RGSGAM2,
                              const double *RGSGAM2H, const double *GSGAM2H, const double *rdgzh,
                              const double *afact, const double *bfact, double dt, double CVovRt2, double alpha,
                              double grav)
{
    const int kmin = 2, kmax = kall - 1;
#define AT(k) (base + (size_t)((k) - 1) * s)
    double y[kall + 2], gam[kall + 2];
    /* T^T y = GSGAM2H * wb on kmin+1..kmax */
    double beta = Mc[AT(kmin + 1)];
    y[kmin + 1] = (wb[AT(kmin + 1)] * GSGAM2H[AT(kmin + 1)]) / beta;
    for (int k = kmin + 2; k <= kmax; ++k) {
        gam[k] = Ml[AT(k)] / beta;
        beta = Mc[AT(k)] - Mu[AT(k - 1)] * gam[k];
        y[k] = (wb[AT(k)] * GSGAM2H[AT(k)] - Mu[AT(k - 1)] * y[k - 1]) / beta;
    }
    for (int k = kmax - 1; k >= kmin + 1; --k) y[k] = y[k] - gam[k + 1] * y[k + 1];
    /* boundary values: w(kmin), w(kmax+1) scaled by RGSGAM2H, entered in Sall
     * (:209-212) and scaled back by GSGAM2H (:242-246) */
    const double bb = wb[AT(kmin)] * GSGAM2H[AT(kmin)] - Ml[AT(kmin + 1)] * y[kmin + 1];
    const double bt = wb[AT(kmax + 1)] * GSGAM2H[AT(kmax + 1)] - Mu[AT(kmax)] * y[kmax];
    for (int k = kall; k >= 1; --k) {
        const int in_k = (k >= kmin + 1 && k <= kmax), in_n = (k + 1 >= kmin + 1 && k + 1 <= kmax);
        const double sb = in_k ? y[k] * CVovRt2 : 0.0;
        const double sbn = in_n ? y[k + 1] * CVovRt2 : 0.0;
        double pb = 0.0, rb = 0.0;
        if (in_k) {
            pb = -(sb * dt * rdgzh[k - 1]);
            rb = -(sb * dt * grav * afact[k - 1]);
        }
        if (in_n) {
            pb = pb + sbn * dt * rdgzh[k];
            rb = rb - sbn * dt * grav * bfact[k];
        }
        const size_t o = AT(k);
        p0b[o] = pb * RGSGAM2[o];
        Spb[o] = dt * p0b[o];
        r0b[o] = rb * (RGAM[o] * RGAM[o]);
        Srb[o] = dt * r0b[o];
        const double sw = in_k ? sb * (RGAMH[o] * RGAMH[o]) : 0.0;
        w0b[o] = sw * alpha;
        Swb[o] = dt * sw;
        if (k == kmin) wb[o] = bb * RGSGAM2H[o];
        else if (k == kmax + 1) wb[o] = bt * RGSGAM2H[o];
        else if (k > kmin) wb[o] = 0.0;
    }
#undef AT
}

void orc_vi_rhow_solver_adjoint(const orc_dims *D, double *wb, double *wb_pl, double *w0b, double *w0b_pl,
                                double *p0b, double *p0b_pl, double *r0b, double *r0b_pl, double *Srb,
                                double *Srb_pl, double *Swb, double *Swb_pl, double *Spb, double *Spb_pl, double dt,
                                const double *Mc, const double *Mc_pl, const double *Ml, const double *Ml_pl,
                                const double *Mu, const double *Mu_pl, const double *RGAM, const double *RGAM_pl,
                                const double *RGAMH, const double *RGAMH_pl, const double *RGSGAM2,
                                const double *RGSGAM2_pl, const double *RGSGAM2H, const double *RGSGAM2H_pl,
                                const double *GSGAM2H, const double *GSGAM2H_pl, const double *rdgzh,
                                const double *afact, const double *bfact, double grav, double rdry, double cvdry,
                                int non_hydro_alpha)
{
    const int gall = D->gall_1d * D->gall_1d, kall = D->kall;
    const double CVovRt2 = cvdry / rdry / (dt * dt), alpha = (double)non_hydro_alpha;
    for (int l = 0; l < D->lall; ++l)
#pragma omp parallel for schedule(static)
        for (int g = 0; g < gall; ++g)
            vi_adjoint_column(kall, (size_t)l * kall * gall + g, (size_t)gall, wb, w0b, p0b, r0b, Srb, Swb, Spb, Mc,
                              Ml, Mu, RGAM, RGAMH, RGSGAM2, RGSGAM2H, GSGAM2H, rdgzh, afact, bfact, dt, CVovRt2,
                              alpha, grav);
    if (D->have_pl)
        for (int l = 0; l < LALL_PL; ++l)
            for (int g = 0; g < D->gall_pl; ++g)
                vi_adjoint_column(kall, (size_t)l * kall * D->gall_pl + g, (size_t)D->gall_pl, wb_pl, w0b_pl, p0b_pl,
                                  r0b_pl, Srb_pl, Swb_pl, Spb_pl, Mc_pl, Ml_pl, Mu_pl, RGAM_pl, RGAMH_pl,
                                  RGSGAM2_pl, RGSGAM2H_pl, GSGAM2H_pl, rdgzh, afact, bfact, dt, CVovRt2, alpha, grav);
}

/* ------------------------------------------------------------------------- */
/* Tangent-linear and adjoint of vertical_limiter_thuburn                    */
/* (dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243), at fixed ck and d.  */
/* ------------------------------------------------------------------------- */
/* The limited face value is one of three things: the unlimited q_h, or the
 * lower or upper bound Qout of the upwind level j (j = k-1 for inflow from
 * below, else k).  The forward pass records one bit per face, set where the
 * value was clipped (output != input); a clipped face sits on the lower bound
 * if its value equals Qout_min(j), else on the upper bound.  Qout_min(j) and
 * Qout_max(j) are piecewise linear in q(j-1), q(j), q(j+1): the min/max
 * selections below follow the forward's own comparisons, mn(a,b) = b<a?b:a
 * and mx(a,b) = b>a?b:a.  vlim_level gives the two bounds of level j and
 * their rows of derivatives (index 0,1,2 = q(j-1), q(j), q(j+1)). */
typedef struct {
    double qmin, qmax;  /* Qout_min, Qout_max */
    double rmin[3], rmax[3];
    double cmin[2], cmax[2];  /* derivatives in ck(j,1), ck(j,2) */
} vlim_lvl;

static void vlim_level(const double *q, const double *d, const double *ck, size_t o, size_t s, size_t ck2,
                       double EPS, double BIG, int pole, vlim_lvl *r)
{
    const double ck1 = ck[o], ck1p = ck[o + s], ck2k = ck[ck2 + o];
    const double inflagL = 0.5 - fsign(0.5, ck1), inflagU = 0.5 + fsign(0.5, ck1p);
    const double qk = q[o], qkm1 = q[o - s], qkp1 = q[o + s];
    const double Qin_minL = mn(qk, qkm1) + (1.0 - inflagL) * BIG;
    const double Qin_minU = mn(qk, qkp1) + (1.0 - inflagU) * BIG;
    const double Qin_maxL = mx(qk, qkm1) - (1.0 - inflagL) * BIG;
    const double Qin_maxU = mx(qk, qkp1) - (1.0 - inflagU) * BIG;
    const double tmin = mn(Qin_minL, Qin_minU), tmax = mx(Qin_maxL, Qin_maxU);
    const double qnext_min = mn(tmin, qk), qnext_max = mx(tmax, qk);
    const double Cin = inflagL * ck1 + inflagU * ck2k;
    const double Cout = (1.0 - inflagL) * ck1 + (1.0 - inflagU) * ck2k;
    double CQin_min, CQin_max;
    if (pole) {
        CQin_max = inflagL * (ck1 * Qin_maxL) + inflagU * (ck2k * Qin_maxU);
        CQin_min = inflagL * (ck1 * Qin_minL) + inflagU * (ck2k * Qin_minU);
    } else {
        CQin_min = inflagL * ck1 * Qin_minL + inflagU * ck2k * Qin_minU;
        CQin_max = inflagL * ck1 * Qin_maxL + inflagU * ck2k * Qin_maxU;
    }
    const double zerosw = 0.5 - fsign(0.5, fabs(Cout) - EPS);
    const double dk = d[o];
    r->qmin = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) * (1.0 - zerosw) + qk * zerosw;
    r->qmax = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) * (1.0 - zerosw) + qk * zerosw;
    /* one-hot rows of the selections */
    const int maxL = (qkm1 > qk) ? 0 : 1, maxU = (qkp1 > qk) ? 2 : 1;  /* Qin_max* -> q index */
    const int minL = (qkm1 < qk) ? 0 : 1, minU = (qkp1 < qk) ? 2 : 1;  /* Qin_min* -> q index */
    const int nmax = (qk > tmax) ? 1 : ((Qin_maxU > Qin_maxL) ? maxU : maxL);
    const int nmin = (qk < tmin) ? 1 : ((Qin_minU < Qin_minL) ? minU : minL);
    const double fac = (1.0 - zerosw) / (Cout + zerosw);
    const double e = (Cin + Cout - dk) - 1.0;
    const double cL = inflagL * ck1, cU = inflagU * ck2k;
    for (int m = 0; m < 3; ++m) {
        const double dnx = (m == nmax) ? 1.0 : 0.0, dnn = (m == nmin) ? 1.0 : 0.0;
        const double cqx = ((m == maxL) ? cL : 0.0) + ((m == maxU) ? cU : 0.0);
        const double cqn = ((m == minL) ? cL : 0.0) + ((m == minU) ? cU : 0.0);
        r->rmin[m] = fac * (e * dnx - cqx);
        r->rmax[m] = fac * (e * dnn - cqn);
    }
    r->rmin[1] = (r->rmin[1] + fac) + zerosw;
    r->rmax[1] = (r->rmax[1] + fac) + zerosw;
    /* ck(j,1) enters Cin where inflagL = 1 (with CQin) and Cout elsewhere; ck(j,2) likewise with inflagU:
     * (qnext - Qin)/Cout or (qnext - Qout)/Cout; zero where zerosw = 1 */
    const int z = zerosw != 0.0;
    r->cmin[0] = z ? 0.0 : (inflagL > 0.0 ? qnext_max - Qin_maxL : qnext_max - r->qmin) / Cout;
    r->cmin[1] = z ? 0.0 : (inflagU > 0.0 ? qnext_max - Qin_maxU : qnext_max - r->qmin) / Cout;
    r->cmax[0] = z ? 0.0 : (inflagL > 0.0 ? qnext_min - Qin_minL : qnext_min - r->qmax) / Cout;
    r->cmax[1] = z ? 0.0 : (inflagU > 0.0 ? qnext_min - Qin_minU : qnext_min - r->qmax) / Cout;
}

/* Row of the limited face k: coefficient of q_h_in (*cqh), of q(k-2..k+1)
 * (row[0..3]) and of ck(j,1), ck(j,2) (crow) of the bound's level j (*lev:
 * 0 = k-1, 1 = k). */
static void vlim_face_row(const vlim_lvl *lo, const vlim_lvl *hi, double inflagL, int clipped, double out,
                          double *cqh, double row[4], double crow[2], int *lev)
{
    row[0] = row[1] = row[2] = row[3] = 0.0;
    crow[0] = crow[1] = 0.0;
    *lev = 0;
    *cqh = clipped ? 0.0 : 1.0;
    if (!clipped) return;
    const vlim_lvl *L = (inflagL > 0.0) ? lo : hi; /* level k-1 or k */
    const int off = (inflagL > 0.0) ? 0 : 1;        /* row index of q(j-1) */
    const int use_min = out == L->qmin;
    const double *r = use_min ? L->rmin : L->rmax;
    for (int m = 0; m < 3; ++m) row[off + m] = r[m];
    crow[0] = use_min ? L->cmin[0] : L->cmax[0];
    crow[1] = use_min ? L->cmin[1] : L->cmax[1];
    *lev = off;
}

/* One column: TLM (dqh_out from dqh_in, dq) or adjoint (qhb, qb) */
static void vlim_lin_column(int kall, size_t base, size_t s, size_t ck2, const uint32_t *mask, size_t mbase,
                            size_t mstride, const double *q_h, const double *q, const double *d, const double *ck,
                            double EPS, double BIG, int pole, int adjoint, double *a_qh, const double *a_dqh_in,
                            double *a_q, double *a_ck)
{
    const int kmin = 2, kmax = kall - 1;
#define AT(k) (base + (size_t)((k) - 1) * s)
    vlim_lvl lo, hi;
    if (adjoint)
        for (int k = 1; k <= kall; ++k) {
            a_q[AT(k)] = 0.0;
            if (a_ck) a_ck[AT(k)] = a_ck[ck2 + AT(k)] = 0.0;
        }
    vlim_level(q, d, ck, AT(kmin), s, ck2, EPS, BIG, pole, &lo);
    for (int k = kmin + 1; k <= kmax; ++k) {
        vlim_level(q, d, ck, AT(k), s, ck2, EPS, BIG, pole, &hi);
        const int p = k - (kmin + 1);
        const int clipped = (mask[mbase + (size_t)(p / 32) * mstride] >> (p % 32)) & 1u;
        const double inflagL = 0.5 - fsign(0.5, ck[AT(k)]);
        double cqh, row[4], crow[2];
        int lev;
        vlim_face_row(&lo, &hi, inflagL, clipped, q_h[AT(k)], &cqh, row, crow, &lev);
        const int j = k - 1 + lev;
        if (adjoint) {
            const double ob = a_qh[AT(k)];
            a_qh[AT(k)] = cqh * ob;
            for (int m = 0; m < 4; ++m) {
                const int kk = k - 2 + m;
                if (kk >= 1 && kk <= kall) a_q[AT(kk)] = a_q[AT(kk)] + row[m] * ob;
            }
            if (a_ck && clipped) {
                a_ck[AT(j)] = a_ck[AT(j)] + crow[0] * ob;
                a_ck[ck2 + AT(j)] = a_ck[ck2 + AT(j)] + crow[1] * ob;
            }
        } else {
            double v = cqh * a_dqh_in[AT(k)];
            for (int m = 0; m < 4; ++m) {
                const int kk = k - 2 + m;
                if (kk >= 1 && kk <= kall) v = v + row[m] * a_q[AT(kk)];
            }
            if (a_ck && clipped) v = (v + crow[0] * a_ck[AT(j)]) + crow[1] * a_ck[ck2 + AT(j)];
            a_qh[AT(k)] = v;
        }
        lo = hi;
    }
#undef AT
}

static void vlim_lin(const orc_dims *D, int adjoint, double *a_qh, double *a_qh_pl, const double *a_dqh_in,
                     const double *a_dqh_in_pl, double *a_q, double *a_q_pl, const double *q_h, const double *q_h_pl,
                     const double *q, const double *q_pl, const double *d, const double *d_pl, const double *ck,
                     const double *ck_pl, const uint32_t *mask, const uint32_t *mask_pl, double *a_ck,
                     double *a_ck_pl)
{
    const int gall = D->gall_1d * D->gall_1d, kall = D->kall, nword = (kall - 3 + 31) / 32;
    if (kall - 3 <= 0) return;
    for (int l = 0; l < D->lall; ++l)
#pragma omp parallel for schedule(static)
        for (int g = 0; g < gall; ++g)
            vlim_lin_column(kall, (size_t)l * kall * gall + g, gall, (size_t)gall * kall * D->lall, mask,
                            (size_t)l * nword * gall + g, gall, q_h, q, d, ck, D->eps, D->big, 0, adjoint, a_qh,
                            a_dqh_in, a_q, a_ck);
    if (D->have_pl)
        for (int l = 0; l < LALL_PL; ++l)
            for (int g = 0; g < D->gall_pl; ++g)
                vlim_lin_column(kall, (size_t)l * kall * D->gall_pl + g, D->gall_pl,
                                (size_t)D->gall_pl * kall * LALL_PL, mask_pl, (size_t)l * nword * D->gall_pl + g,
                                D->gall_pl, q_h_pl, q_pl, d_pl, ck_pl, D->eps, D->big, 1, adjoint, a_qh_pl,
                                a_dqh_in_pl, a_q_pl, a_ck_pl);
}

void orc_vertical_limiter_thuburn_tlm(const orc_dims *D, double *dqh, double *dqh_pl, const double *dqh_in,
                                      const double *dqh_in_pl, const double *dq, const double *dq_pl,
                                      const double *q_h, const double *q_h_pl, const double *q, const double *q_pl,
                                      const double *d, const double *d_pl, const double *ck, const double *ck_pl,
                                      const uint32_t *mask, const uint32_t *mask_pl, const double *dck,
                                      const double *dck_pl)
{
    vlim_lin(D, 0, dqh, dqh_pl, dqh_in, dqh_in_pl, (double *)dq, (double *)dq_pl, q_h, q_h_pl, q, q_pl, d, d_pl,
             ck, ck_pl, mask, mask_pl, (double *)dck, (double *)dck_pl);
}

void orc_vertical_limiter_thuburn_adjoint(const orc_dims *D, double *qhb, double *qhb_pl, double *qb, double *qb_pl,
                                          const double *q_h, const double *q_h_pl, const double *q,
                                          const double *q_pl, const double *d, const double *d_pl, const double *ck,
                                          const double *ck_pl, const uint32_t *mask, const uint32_t *mask_pl,
                                          double *ck_b, double *ck_b_pl)
{
    vlim_lin(D, 1, qhb, qhb_pl, NULL, NULL, qb, qb_pl, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl,
             ck_b, ck_b_pl);
}

/* ------------------------------------------------------------------------- */
/* Tangent-linear and adjoint of horizontal_limiter_thuburn                  */
/* (dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430) with Qout_post       */
/* supplied, at fixed d, ch and cmask.                                        */
/* ------------------------------------------------------------------------- */
/* Inputs that matter: q_a (faces 1-3 of cells [gmin-1,gmax]^2), q and
 * Qout_post; outputs: q_a and Qout_prev.  Qin of an arc is the min (max) of
 * its four tracer values, a selection; Qout of an inner cell is affine in q
 * given the selections of qnext and of its six arcs; a halo cell's Qout is q.
 * A face value is its input, a Qout_post bound of the downwind cell, or a Qin
 * bound of its arc.  The forward records one bit per face (bit m-1 of the
 * cell's mask byte: the value changed); a changed value is attributed, in
 * this order, to Qout_post min, Qout_post max, Qin min, Qin max, by equality
 * with the stored output.  Pole faces: bit 0 of the ring point's byte. */
typedef struct {
    int cell[4];
} hl_arc;

/* the 4 cells of arc m (1=AI, 2=AIJ, 3=AJ) of cell ij (HLIM:145-160, :197-203) */
static hl_arc hl_arc_cells(int ij, int m, int iall, int sgp_cell)
{
    hl_arc a;
    if (m == 1) {
        const int ijm1 = (ij - iall) > 1 ? ij - iall : 1;
        a.cell[0] = ij; a.cell[1] = ijm1; a.cell[2] = ij + 1; a.cell[3] = ij + iall + 1;
    } else if (m == 2) {
        if (sgp_cell) { a.cell[0] = ij; a.cell[1] = ij + iall + 1; a.cell[2] = ij + iall + 2; a.cell[3] = ij + iall; }
        else { a.cell[0] = ij; a.cell[1] = ij + 1; a.cell[2] = ij + iall + 1; a.cell[3] = ij + iall; }
    } else {
        const int im1j = (ij - 1) > 1 ? ij - 1 : 1;
        a.cell[0] = ij; a.cell[1] = ij + iall + 1; a.cell[2] = ij + iall; a.cell[3] = im1j;
    }
    return a;
}

/* cells holding the min and the max of the arc, following mn/mx's chains */
static void hl_arc_sel(const double *qs, const hl_arc *a, int *cmin, int *cmax, double *vmin, double *vmax)
{
    double lo = qs[a->cell[0] - 1], hi = lo;
    int il = a->cell[0], ih = a->cell[0];
    for (int t = 1; t < 4; ++t) {
        const double v = qs[a->cell[t] - 1];
        if (v < lo) { lo = v; il = a->cell[t]; }
        if (v > hi) { hi = v; ih = a->cell[t]; }
    }
    *cmin = il; *cmax = ih; *vmin = lo; *vmax = hi;
}

static int hl_off(int m, int iall) { return m == 1 ? 1 : (m == 2 ? iall + 1 : iall); }

/* mode 0 = TLM, 1 = adjoint; one (k, l) slab of a regular region.
 * TLM:     dqa (in dq_a_in, out dq_a), dq, dQpost -> dQprev
 * adjoint: qab (in/out), Qprevb -> qb (+=), Qpostb (+=) */
static void hl_slab(int iall, int sgp, const double *qs, const double *ds, const double *chs, const double *cms,
                    size_t cplane, const double *qas_out, const double *Qpost_s, size_t qplane,
                    const uint8_t *mk, double EPS, double BIG, int adjoint, double *xa, size_t aplane,
                    double *xq, double *xQprev, double *xQpost, double *xch)
{
    const int gall = iall * iall, gmin = 2, gmax = iall - 1;
    const int sgp_cell = sgp ? (gmin - 1 - 1) * iall + (gmin - 1) : -1;
#define CM(g, f) cms[(size_t)((f) - 1) * cplane + (g) - 1]
#define CH(g, f) chs[(size_t)((f) - 1) * cplane + (g) - 1]
    /* The adjoint's order: each row's terms are first summed per cell of the
     * owner's hexagon {g-iall-1, g-iall, g-1, g, g+1, g+iall, g+iall+1} (every
     * cell a row names lies there), in row order (bound min, then max; each
     * bound's terms in order), into seven partial sums; each cell then adds
     * the partial sums of the rows of its hexagon in cell order. */
    const int hexo[7] = {-iall - 1, -iall, -1, 0, 1, iall, iall + 1};
#define HEXADD(part)                                                              \
    do {                                                                          \
        for (int w = 0; w < 7; ++w) {                                             \
            const int tg = g + hexo[w];                                           \
            if (tg >= 1 && tg <= gall) xq[tg - 1] = xq[tg - 1] + (part)[w];       \
        }                                                                         \
    } while (0)
    /* Qout of each cell: coefficient rows (up to 14 terms) */
    for (int g = 1; g <= gall; ++g) {
        const int i = (g - 1) % iall + 1, j = (g - 1) / iall + 1;
        if (i < gmin || i > gmax || j < gmin || j > gmax) { /* halo: Qout = q (HLIM:286-301) */
            if (adjoint) {
                double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                for (int mm = 0; mm < 2; ++mm) part[3] = part[3] + xQprev[(size_t)mm * qplane + g - 1];
                HEXADD(part);
            } else {
                for (int mm = 0; mm < 2; ++mm) xQprev[(size_t)mm * qplane + g - 1] = xq[g - 1];
            }
            if (adjoint && xch)
                for (int f = 1; f <= 6; ++f) xch[(size_t)(f - 1) * cplane + g - 1] = 0.0;
            continue;
        }
        /* the six arcs of g: f=1..3 its own, f=4..6 the neighbours' (m = f-3) */
        int cmn[7], cmx[7];
        double vmn[7], vmx[7], cmin_c[7], cmax_c[7];
        for (int f = 1; f <= 6; ++f) {
            const int owner = f <= 3 ? g : g - hl_off(f - 3, iall);
            const int m = f <= 3 ? f : f - 3;
            hl_arc a = hl_arc_cells(owner, m, iall, owner == sgp_cell && m == 2);
            hl_arc_sel(qs, &a, &cmn[f], &cmx[f], &vmn[f], &vmx[f]);
            const double c = CM(owner, m);
            cmin_c[f] = f <= 3 ? c : 1.0 - c; /* d QIN / d (arc extreme) */
            cmax_c[f] = cmin_c[f];
            /* the forward's values, with the sentinels */
            if (f <= 3) { vmn[f] = c * vmn[f] + (1.0 - c) * BIG; vmx[f] = c * vmx[f] + (1.0 - c) * (-BIG); }
            else { vmn[f] = c * BIG + (1.0 - c) * vmn[f]; vmx[f] = c * (-BIG) + (1.0 - c) * vmx[f]; }
        }
        const double qij = qs[g - 1];
        int smin = 0, smax = 0;
        double nmin = qij, nmax = qij;
        for (int f = 1; f <= 6; ++f) {
            if (vmn[f] < nmin) { nmin = vmn[f]; smin = f; }
            if (vmx[f] > nmax) { nmax = vmx[f]; smax = f; }
        }
        double chv[7], chm[7];
        for (int f = 1; f <= 6; ++f) { chv[f] = CH(g, f); chm[f] = mn(chv[f], 0.0); }
        const double Cin = chm[1] + chm[2] + chm[3] + chm[4] + chm[5] + chm[6];
        const double Cout = chv[1] - chm[1] + chv[2] - chm[2] + chv[3] - chm[3] + chv[4] - chm[4] + chv[5] - chm[5] + chv[6] - chm[6];
        const double zerosw = 0.5 - fsign(0.5, fabs(Cout) - EPS);
        const double E = 1.0 - Cin - Cout + ds[g - 1];
        const double fac = (1.0 - zerosw) / (Cout + zerosw);
        /* derivatives of Qout min (mm=0) and max (mm=1) in ch(f): min(ch, 0) enters Cin and CQin, ch - min(ch, 0)
         * Cout, so an inflow face (ch < 0) gives (qnext - Qin)/Cout and any other (qnext - Qout)/Cout (the
         * right-hand derivative at ch = 0); zero where zerosw = 1, where Qout = q */
        double dch[2][7];
        {
            const double CQmin = chm[1] * vmn[1] + chm[2] * vmn[2] + chm[3] * vmn[3] + chm[4] * vmn[4] + chm[5] * vmn[5] + chm[6] * vmn[6];
            const double CQmax = chm[1] * vmx[1] + chm[2] * vmx[2] + chm[3] * vmx[3] + chm[4] * vmx[4] + chm[5] * vmx[5] + chm[6] * vmx[6];
            const double Qmin = (qij - CQmax - nmax * E) / (Cout + zerosw) * (1.0 - zerosw) + qij * zerosw;
            const double Qmax = (qij - CQmin - nmin * E) / (Cout + zerosw) * (1.0 - zerosw) + qij * zerosw;
            for (int f = 1; f <= 6; ++f) {
                dch[0][f] = zerosw != 0.0 ? 0.0 : (chv[f] < 0.0 ? nmax - vmx[f] : nmax - Qmin) / Cout;
                dch[1][f] = zerosw != 0.0 ? 0.0 : (chv[f] < 0.0 ? nmin - vmn[f] : nmin - Qmax) / Cout;
            }
        }
        /* a bound whose cotangent (direction) is zero contributes nothing: where an inflow face meets the
         * sentinel of an arc whose upwind side is the cell itself, the bound is huge and its derivative may
         * overflow, and 0 * inf must not turn into NaN */
        if (adjoint && xch) {
            const double b0 = xQprev[g - 1], b1 = xQprev[qplane + g - 1];
            for (int f = 1; f <= 6; ++f)
                xch[(size_t)(f - 1) * cplane + g - 1] = (b0 != 0.0 ? b0 * dch[0][f] : 0.0) + (b1 != 0.0 ? b1 * dch[1][f] : 0.0);
        }
        /* Qout min (mm=0) uses the max side, Qout max (mm=1) the min side */
        double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int mm = 0; mm < 2; ++mm) {
            const int sel = mm == 0 ? smax : smin;
            const int *cx = mm == 0 ? cmx : cmn;
            const double *cc = mm == 0 ? cmax_c : cmin_c;
            int cells[14];
            double coef[14];
            int n = 0;
            cells[n] = g; coef[n++] = fac + zerosw;
            cells[n] = sel == 0 ? g : cx[sel]; coef[n++] = -(fac * E) * (sel == 0 ? 1.0 : cc[sel]);
            for (int f = 1; f <= 6; ++f) { cells[n] = cx[f]; coef[n++] = -(fac * chm[f]) * cc[f]; }
            const size_t o = (size_t)mm * qplane + g - 1;
            if (adjoint) {
                const double b = xQprev[o];
                for (int t = 0; t < n; ++t) {
                    int w = 0;
                    while (w < 7 && g + hexo[w] != cells[t]) ++w;
                    if (w == 7) abort(); /* a row names only cells of its owner's hexagon */
                    part[w] = part[w] + coef[t] * b;
                }
            } else {
                double v = 0.0;
                for (int t = 0; t < n; ++t) v = v + coef[t] * xq[cells[t] - 1];
                if (xch)
                    for (int f = 1; f <= 6; ++f) {
                        const double dc = xch[(size_t)(f - 1) * cplane + g - 1];
                        if (dc != 0.0) v = v + dch[mm][f] * dc;
                    }
                xQprev[o] = v;
            }
        }
        if (adjoint) HEXADD(part);
    }
#undef HEXADD
    /* faces 1-3 of [gmin-1, gmax]^2 (HLIM:377-410) */
    for (int j = gmin - 1; j <= gmax; ++j)
        for (int i = gmin - 1; i <= gmax; ++i) {
            const int ij = (j - 1) * iall + i;
            for (int m = 1; m <= 3; ++m) {
                const int nb = ij + hl_off(m, iall);
                const size_t fo = (size_t)(m - 1) * aplane + ij - 1, fn = (size_t)(m + 2) * aplane + nb - 1;
                const int clipped = (mk[ij - 1] >> (m - 1)) & 1;
                const double c = CM(ij, m);
                const int T = c > 0.0 ? nb : ij;
                /* attribution: 0 input, 1 Qpost min, 2 Qpost max, 3 Qin min cell, 4 Qin max cell */
                int kind = 0, cell = 0;
                if (clipped) {
                    const double out = qas_out[fo];
                    hl_arc a = hl_arc_cells(ij, m, iall, ij == sgp_cell && m == 2);
                    int amn, amx;
                    double vmn, vmx;
                    hl_arc_sel(qs, &a, &amn, &amx, &vmn, &vmx);
                    if (out == Qpost_s[T - 1]) kind = 1;
                    else if (out == Qpost_s[qplane + T - 1]) kind = 2;
                    else if (out == vmn) { kind = 3; cell = amn; }
                    else { kind = 4; cell = amx; }
                }
                if (adjoint) {
                    const double ob = xa[fo] + xa[fn];
                    xa[fn] = 0.0;
                    xa[fo] = kind == 0 ? ob : 0.0;
                    if (kind == 1) xQpost[T - 1] = xQpost[T - 1] + ob;
                    else if (kind == 2) xQpost[qplane + T - 1] = xQpost[qplane + T - 1] + ob;
                    else if (kind >= 3) xq[cell - 1] = xq[cell - 1] + ob;
                } else {
                    double v;
                    if (kind == 0) v = xa[fo];
                    else if (kind == 1) v = xQpost[T - 1];
                    else if (kind == 2) v = xQpost[qplane + T - 1];
                    else v = xq[cell - 1];
                    xa[fo] = v;
                    xa[fn] = v;
                }
            }
        }
#undef CM
#undef CH
}

/* one (k, l) slab of a pole region (HLIM:307-363, 412-425); arrays at the slab */
static void hl_pole_slab(int npl, const double *qs, double dn, const double *chs, const double *cms,
                         const double *qas_out, const double *Qpost_s, size_t pplane, const uint8_t *mk, double EPS,
                         double BIG, int adjoint, double *xa, double *xq, double *xQprev, double *xQpost, double *xch)
{
    /* ring point v (1-based 2..npl): arc {centre, v, v-1, v+1} */
    int amn[16], amx[16];
    double vmn[16], vmx[16];
    for (int v = 2; v <= npl; ++v) {
        int vp1 = v + 1, vm1 = v - 1;
        if (vp1 == npl + 1) vp1 = 2;
        if (vm1 == 1) vm1 = npl;
        hl_arc a = {{1, v, vm1, vp1}};
        hl_arc_sel(qs, &a, &amn[v], &amx[v], &vmn[v], &vmx[v]);
    }
    const double qn = qs[0];
    int smin = 0, smax = 0;
    double nmin = qn, nmax = qn;
    for (int v = 2; v <= npl; ++v) {
        const double c = cms[v - 1];
        const double mn1 = c * vmn[v] + (1.0 - c) * BIG, mx1 = c * vmx[v] + (1.0 - c) * (-BIG);
        if (mn1 < nmin) { nmin = mn1; smin = v; }
        if (mx1 > nmax) { nmax = mx1; smax = v; }
    }
    double Cin = 0.0, Cout = 0.0, CQmin = 0.0, CQmax = 0.0;
    for (int v = 2; v <= npl; ++v) {
        const double chm = cms[v - 1] * chs[v - 1];
        const double c = cms[v - 1];
        Cin = Cin + chm;
        Cout = Cout - chm + chs[v - 1];
        CQmin = CQmin + chm * (c * vmn[v] + (1.0 - c) * BIG);
        CQmax = CQmax + chm * (c * vmx[v] + (1.0 - c) * (-BIG));
    }
    const double zerosw = 0.5 - fsign(0.5, fabs(Cout) - EPS);
    const double E = 1.0 - Cin - Cout + dn;
    const double fac = (1.0 - zerosw) / (Cout + zerosw);
    /* derivatives in ch(v): cmask*ch enters Cin and CQin, ch - cmask*ch Cout:
     * ((qnext - c Qin) - (1 - c) Qout) / Cout, zero where zerosw = 1 */
    double dch[2][16];
    {
        const double Qmin = (qn - CQmax - nmax * E) / (Cout + zerosw) * (1.0 - zerosw) + qn * zerosw;
        const double Qmax = (qn - CQmin - nmin * E) / (Cout + zerosw) * (1.0 - zerosw) + qn * zerosw;
        for (int v = 2; v <= npl; ++v) {
            const double c = cms[v - 1];
            const double imx = c * vmx[v] + (1.0 - c) * (-BIG), imn = c * vmn[v] + (1.0 - c) * BIG;
            dch[0][v - 1] = zerosw != 0.0 ? 0.0 : ((nmax - c * imx) - (1.0 - c) * Qmin) / Cout;
            dch[1][v - 1] = zerosw != 0.0 ? 0.0 : ((nmin - c * imn) - (1.0 - c) * Qmax) / Cout;
        }
    }
    if (adjoint && xch) {
        xch[0] = 0.0;
        const double b0 = xQprev[0], b1 = xQprev[pplane];
        for (int v = 2; v <= npl; ++v)
            xch[v - 1] = (b0 != 0.0 ? b0 * dch[0][v - 1] : 0.0) + (b1 != 0.0 ? b1 * dch[1][v - 1] : 0.0);
    }
    for (int mm = 0; mm < 2; ++mm) {
        const int sel = mm == 0 ? smax : smin;
        const int *cx = mm == 0 ? amx : amn;
        int cells[20];
        double coef[20];
        int n = 0;
        cells[n] = 1; coef[n++] = fac + zerosw;
        cells[n] = sel == 0 ? 1 : cx[sel]; coef[n++] = -(fac * E) * (sel == 0 ? 1.0 : cms[sel - 1]);
        for (int v = 2; v <= npl; ++v) { cells[n] = cx[v]; coef[n++] = -(fac * (cms[v - 1] * chs[v - 1])) * cms[v - 1]; }
        const size_t o = (size_t)mm * pplane;
        if (adjoint) {
            const double b = xQprev[o];
            for (int t = 0; t < n; ++t) xq[cells[t] - 1] = xq[cells[t] - 1] + coef[t] * b;
        } else {
            double v = 0.0;
            for (int t = 0; t < n; ++t) v = v + coef[t] * xq[cells[t] - 1];
            if (xch)
                for (int r = 2; r <= npl; ++r)
                    if (xch[r - 1] != 0.0) v = v + dch[mm][r - 1] * xch[r - 1];
            xQprev[o] = v;
            for (int r = 2; r <= npl; ++r) xQprev[o + r - 1] = 0.0; /* the ring is zeroed (HLIM:367-370) */
        }
    }
    for (int v = 2; v <= npl; ++v) {
        const int clipped = mk[v - 1] & 1;
        const double c = cms[v - 1];
        const int T = c > 0.0 ? v : 1;
        int kind = 0, cell = 0;
        if (clipped) {
            const double out = qas_out[v - 1];
            if (out == Qpost_s[T - 1]) kind = 1;
            else if (out == Qpost_s[pplane + T - 1]) kind = 2;
            else if (out == vmn[v]) { kind = 3; cell = amn[v]; }
            else { kind = 4; cell = amx[v]; }
        }
        if (adjoint) {
            const double ob = xa[v - 1];
            xa[v - 1] = kind == 0 ? ob : 0.0;
            if (kind == 1) xQpost[T - 1] = xQpost[T - 1] + ob;
            else if (kind == 2) xQpost[pplane + T - 1] = xQpost[pplane + T - 1] + ob;
            else if (kind >= 3) xq[cell - 1] = xq[cell - 1] + ob;
        } else {
            xa[v - 1] = kind == 0 ? xa[v - 1] : (kind == 1 ? xQpost[T - 1] : (kind == 2 ? xQpost[pplane + T - 1] : xq[cell - 1]));
        }
    }
}

/* TLM: dq_a (in: dq_a_in, out: dq_a), dQprev written from dq and dQpost.
 * Adjoint: q_a_b in/out, qb and Qpostb written (zeroed first), from
 * Qprevb.  q_a is the forward's output, mask its mask (uint8 per cell-level,
 * uint8 per pole point). */
static void hlim_lin(const orc_dims *D, const int *have_sgp, int adjoint, double *xa, double *xa_pl, double *xq,
                     double *xq_pl, double *xQprev, double *xQprev_pl, double *xQpost, double *xQpost_pl,
                     const double *q_a, const double *q_a_pl, const double *q, const double *q_pl, const double *d,
                     const double *d_pl, const double *ch, const double *ch_pl, const double *cmask,
                     const double *cmask_pl, const double *Qpost, const double *Qpost_pl, const uint8_t *mask,
                     const uint8_t *mask_pl, double *xch, double *xch_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, npl = D->gall_pl;
    const size_t nreg = (size_t)gall * kall * lall, npole = (size_t)npl * kall * LALL_PL;
    if (adjoint) {
        memset(xq, 0, nreg * sizeof(double));
        memset(xQpost, 0, 2 * nreg * sizeof(double));
        if (D->have_pl) {
            memset(xq_pl, 0, npole * sizeof(double));
            memset(xQpost_pl, 0, 2 * npole * sizeof(double));
        }
    }
    for (int l = 0; l < lall; ++l)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < kall; ++k) {
            const size_t s = ((size_t)l * kall + k) * gall;
            hl_slab(iall, have_sgp[l], q + s, d + s, ch + s, cmask + s, nreg, q_a + s, Qpost + s, nreg, mask + s,
                    D->eps, D->big, adjoint, xa + s, nreg, xq + s, xQprev + s, xQpost + s, xch ? xch + s : NULL);
        }
    if (D->have_pl)
        for (int l = 0; l < LALL_PL; ++l)
            for (int k = 0; k < kall; ++k) {
                const size_t s = ((size_t)l * kall + k) * npl;
                hl_pole_slab(npl, q_pl + s, d_pl[s], ch_pl + s, cmask_pl + s, q_a_pl + s, Qpost_pl + s, npole,
                             mask_pl + s, D->eps, D->big, adjoint, xa_pl + s, xq_pl + s, xQprev_pl + s,
                             xQpost_pl + s, xch_pl ? xch_pl + s : NULL);
            }
}

void orc_horizontal_limiter_thuburn_tlm(const orc_dims *D, const int *have_sgp, double *dq_a, double *dq_a_pl,
                                        const double *dq, const double *dq_pl, double *dQprev, double *dQprev_pl,
                                        const double *dQpost, const double *dQpost_pl, const double *q_a,
                                        const double *q_a_pl, const double *q, const double *q_pl, const double *d,
                                        const double *d_pl, const double *ch, const double *ch_pl,
                                        const double *cmask, const double *cmask_pl, const double *Qpost,
                                        const double *Qpost_pl, const uint8_t *mask, const uint8_t *mask_pl,
                                        const double *dch, const double *dch_pl)
{
    hlim_lin(D, have_sgp, 0, dq_a, dq_a_pl, (double *)dq, (double *)dq_pl, dQprev, dQprev_pl, (double *)dQpost,
             (double *)dQpost_pl, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qpost, Qpost_pl, mask,
             mask_pl, (double *)dch, (double *)dch_pl);
}

void orc_horizontal_limiter_thuburn_adjoint(const orc_dims *D, const int *have_sgp, double *q_a_b,
                                            double *q_a_b_pl, double *q_b, double *q_b_pl, const double *Qprev_b,
                                            const double *Qprev_b_pl, double *Qpost_b, double *Qpost_b_pl,
                                            const double *q_a, const double *q_a_pl, const double *q,
                                            const double *q_pl, const double *d, const double *d_pl,
                                            const double *ch, const double *ch_pl, const double *cmask,
                                            const double *cmask_pl, const double *Qpost, const double *Qpost_pl,
                                            const uint8_t *mask, const uint8_t *mask_pl, double *ch_b,
                                            double *ch_b_pl)
{
    hlim_lin(D, have_sgp, 1, q_a_b, q_a_b_pl, q_b, q_b_pl, (double *)Qprev_b, (double *)Qprev_b_pl, Qpost_b,
             Qpost_b_pl, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qpost, Qpost_pl, mask, mask_pl,
             ch_b, ch_b_pl);
}

/* ------------------------------------------------------------------------- */
/* Tangent-linear and adjoint of horizontal_flux                              */
/* (dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379), metrics and dt fixed.  */
/* ------------------------------------------------------------------------- */
/* flx_h is linear in rho*v; GRD_xc = GRD_xr - (rho*v)_arc / max(rho_arc, EPS)
 * * dt/2, whose derivative in rho_arc is (rho*v)_arc dt/2 / rho_arc^2 where
 * rho_arc > EPS... precisely where max() takes rho_arc.  The outputs that
 * carry a derivative are those the arc loops write last: faces 1-3 and
 * GRD_xc of the arc origins, faces 4-6 of their neighbours, except face 6 of
 * the cell the singular-point fix zeroes (FLX:311-320).  The triangle values
 * are the forward's (same sums, same order). */
#define FL_TW(g, t, w) GMTR_t[(((size_t)((w) - 1) * 2 + ((t) - 1)) * lall + l) * gall + (g)]
#define FL_A(g, a, c) GMTR_a[(((size_t)((c) - 1) * 3 + ((a) - 1)) * lall + l) * gall + (g)]
#define FL_RA(g) GMTR_p[((size_t)(P_RAREA - 1) * lall + l) * gall + (g)]

static void flux_lin_slab(int iall, int sgp, int l, int lall, const double *X[4], double dt, double EPS,
                          const double *GMTR_t, const double *GMTR_a, const double *GMTR_p, int adjoint,
                          const double *dX[4], double *dflx, double *dxc, size_t plane, const double *flxb,
                          const double *xcb, double *Xb[4])
{
    const int gall = iall * iall;
    double *T = (double *)calloc((size_t)16 * gall, sizeof(double)); /* TI, TJ (forward) and their tangents */
#define TIv(f, g) T[(size_t)(2 * (f)) * gall + (g)]
#define TJv(f, g) T[(size_t)(2 * (f) + 1) * gall + (g)]
#define TId(f, g) T[(size_t)(8 + 2 * (f)) * gall + (g)]
#define TJd(f, g) T[(size_t)(8 + 2 * (f) + 1) * gall + (g)]
    for (int f = 0; f < 4; ++f)
        for (int j = 0; j <= iall - 2; ++j)
            for (int i = 0; i <= iall - 2; ++i) {
                const int g = j * iall + i;
                const double *x = X[f];
                TIv(f, g) = x[g] * FL_TW(g, TI, W1) + x[g + 1] * FL_TW(g, TI, W2) + x[g + iall + 1] * FL_TW(g, TI, W3);
                TJv(f, g) = x[g] * FL_TW(g, TJ, W1) + x[g + iall + 1] * FL_TW(g, TJ, W2) + x[g + iall] * FL_TW(g, TJ, W3);
                if (!adjoint) {
                    const double *y = dX[f];
                    TId(f, g) = y[g] * FL_TW(g, TI, W1) + y[g + 1] * FL_TW(g, TI, W2) + y[g + iall + 1] * FL_TW(g, TI, W3);
                    TJd(f, g) = y[g] * FL_TW(g, TJ, W1) + y[g + iall + 1] * FL_TW(g, TJ, W2) + y[g + iall] * FL_TW(g, TJ, W3);
                }
            }
    if (sgp)
        for (int f = 0; f < 4; ++f) {
            TIv(f, 0) = TJv(f, 1);
            if (!adjoint) TId(f, 0) = TJd(f, 1);
        }
    const int sgp_face = sgp ? iall + 1 : -1; /* face 6 of this cell stays zero */
    if (!adjoint)
        for (size_t t = 0; t < 6 * plane; t += plane)
            for (int g = 0; g < gall; ++g) dflx[t + g] = 0.0;
    if (!adjoint)
        for (size_t t = 0; t < 9 * plane; t += plane)
            for (int g = 0; g < gall; ++g) dxc[t + g] = 0.0;
    double *Tb = adjoint ? (double *)calloc((size_t)8 * gall, sizeof(double)) : NULL;
#define TIb(f, g) Tb[(size_t)(2 * (f)) * gall + (g)]
#define TJb(f, g) Tb[(size_t)(2 * (f) + 1) * gall + (g)]
    for (int a = 1; a <= 3; ++a)
        for (int j = (a == 1 ? 1 : 0); j <= iall - 2; ++j)
            for (int i = (a == 3 ? 1 : 0); i <= iall - 2; ++i) {
                const int g = j * iall + i;
                const int nb = g + (a == 1 ? 1 : (a == 2 ? iall + 1 : iall));
                /* the two triangles of the arc: (kind, cell), kind 0 = TI, 1 = TJ */
                const int t1k = a == 1 ? 1 : (a == 2 ? 0 : 1), t1c = a == 1 ? g - iall : g;
                const int t2k = a == 1 ? 0 : (a == 2 ? 1 : 0), t2c = a == 3 ? g - 1 : g;
                double s2[4];
                for (int f = 0; f < 4; ++f)
                    s2[f] = (t1k ? TJv(f, t1c) : TIv(f, t1c)) + (t2k ? TJv(f, t2c) : TIv(f, t2c));
                const double rr = 1.0 / mx(s2[0], EPS);
                const int live_rho = !(EPS > s2[0]);
                const double hn[3] = {FL_A(g, a, HNX), FL_A(g, a, HNY), FL_A(g, a, HNZ)};
                const double ra = FL_RA(g) * dt, rn = FL_RA(nb) * dt;
                if (!adjoint) {
                    double d2[4];
                    for (int f = 0; f < 4; ++f)
                        d2[f] = (t1k ? TJd(f, t1c) : TId(f, t1c)) + (t2k ? TJd(f, t2c) : TId(f, t2c));
                    const double dflux = 0.5 * (d2[1] * hn[0] + d2[2] * hn[1] + d2[3] * hn[2]);
                    dflx[(size_t)(a - 1) * plane + g] = dflux * ra;
                    dflx[(size_t)(a + 2) * plane + nb] = (a == 3 && nb == sgp_face) ? 0.0 : -dflux * rn;
                    const double drr = live_rho ? -(d2[0] * rr * rr) : 0.0;
                    for (int dd = 0; dd < 3; ++dd)
                        dxc[((size_t)dd * 3 + (a - 1)) * plane + g] = -((d2[1 + dd] * rr + s2[1 + dd] * drr) * dt * 0.5);
                } else {
                    const double fb_nb = (a == 3 && nb == sgp_face) ? 0.0 : flxb[(size_t)(a + 2) * plane + nb];
                    const double fluxb = flxb[(size_t)(a - 1) * plane + g] * ra - fb_nb * rn;
                    double b[4];
                    double r2b = 0.0;
                    for (int dd = 0; dd < 3; ++dd) {
                        const double xb = xcb[((size_t)dd * 3 + (a - 1)) * plane + g];
                        b[1 + dd] = 0.5 * hn[dd] * fluxb - xb * rr * dt * 0.5;
                        r2b = r2b + xb * s2[1 + dd] * dt * 0.5 * rr * rr;
                    }
                    b[0] = live_rho ? r2b : 0.0;
                    for (int f = 0; f < 4; ++f) {
                        if (t1k) TJb(f, t1c) += b[f]; else TIb(f, t1c) += b[f];
                        if (t2k) TJb(f, t2c) += b[f]; else TIb(f, t2c) += b[f];
                    }
                }
            }
    if (adjoint) {
        if (sgp)
            for (int f = 0; f < 4; ++f) {
                TJb(f, 1) += TIb(f, 0);
                TIb(f, 0) = 0.0;
            }
        for (int f = 0; f < 4; ++f)
            for (int j = 0; j <= iall - 2; ++j)
                for (int i = 0; i <= iall - 2; ++i) {
                    const int g = j * iall + i;
                    double *y = Xb[f];
                    const double bi = TIb(f, g), bj = TJb(f, g);
                    y[g] += bi * FL_TW(g, TI, W1) + bj * FL_TW(g, TJ, W1);
                    y[g + 1] += bi * FL_TW(g, TI, W2);
                    y[g + iall + 1] += bi * FL_TW(g, TI, W3) + bj * FL_TW(g, TJ, W2);
                    y[g + iall] += bj * FL_TW(g, TJ, W3);
                }
        free(Tb);
    }
#undef TIv
#undef TJv
#undef TId
#undef TJd
#undef TIb
#undef TJb
    free(T);
}

static void flux_lin_pole(int npl, int l, const double *X[4], double dt, double EPS, const double *GMTR_t_pl,
                          const double *GMTR_a_pl, const double *GMTR_p_pl, int adjoint, const double *dX[4],
                          double *dflx, double *dxc, size_t pplane, const double *flxb, const double *xcb,
                          double *Xb[4])
{
    const int gall_pl = npl;
#define PT(v, w) GMTR_t_pl[((size_t)((w) - 1) * LALL_PL + l) * gall_pl + (v)]
#define PA(v, c) GMTR_a_pl[((size_t)((c) - 1) * LALL_PL + l) * gall_pl + (v)]
    const double ran = GMTR_p_pl[((size_t)(P_RAREA - 1) * LALL_PL + l) * gall_pl + 0] * dt;
    double Tv[4][16], Td[4][16], Tb[4][16];
    for (int v = 1; v < npl; ++v) {
        const int vp1 = (v + 1 == npl) ? 1 : v + 1;
        for (int f = 0; f < 4; ++f) {
            Tv[f][v] = X[f][0] * PT(v, W1) + X[f][v] * PT(v, W2) + X[f][vp1] * PT(v, W3);
            if (!adjoint) Td[f][v] = dX[f][0] * PT(v, W1) + dX[f][v] * PT(v, W2) + dX[f][vp1] * PT(v, W3);
            Tb[f][v] = 0.0;
        }
    }
    if (!adjoint) {
        for (int v = 0; v < npl; ++v) dflx[v] = 0.0;
        for (int dd = 0; dd < 3; ++dd)
            for (int v = 0; v < npl; ++v) dxc[(size_t)dd * pplane + v] = 0.0;
    }
    for (int v = 1; v < npl; ++v) {
        const int vm1 = (v == 1) ? npl - 1 : v - 1;
        double s2[4];
        for (int f = 0; f < 4; ++f) s2[f] = Tv[f][vm1] + Tv[f][v];
        const double rr = 1.0 / mx(s2[0], EPS);
        const int live_rho = !(EPS > s2[0]);
        const double hn[3] = {PA(v, HNX), PA(v, HNY), PA(v, HNZ)};
        if (!adjoint) {
            double d2[4];
            for (int f = 0; f < 4; ++f) d2[f] = Td[f][vm1] + Td[f][v];
            dflx[v] = 0.5 * (d2[1] * hn[0] + d2[2] * hn[1] + d2[3] * hn[2]) * ran;
            const double drr = live_rho ? -(d2[0] * rr * rr) : 0.0;
            for (int dd = 0; dd < 3; ++dd) dxc[(size_t)dd * pplane + v] = -((d2[1 + dd] * rr + s2[1 + dd] * drr) * dt * 0.5);
        } else {
            const double fluxb = flxb[v] * ran;
            double b[4], r2b = 0.0;
            for (int dd = 0; dd < 3; ++dd) {
                const double xb = xcb[(size_t)dd * pplane + v];
                b[1 + dd] = 0.5 * hn[dd] * fluxb - xb * rr * dt * 0.5;
                r2b = r2b + xb * s2[1 + dd] * dt * 0.5 * rr * rr;
            }
            b[0] = live_rho ? r2b : 0.0;
            for (int f = 0; f < 4; ++f) {
                Tb[f][vm1] += b[f];
                Tb[f][v] += b[f];
            }
        }
    }
    if (adjoint)
        for (int v = 1; v < npl; ++v) {
            const int vp1 = (v + 1 == npl) ? 1 : v + 1;
            for (int f = 0; f < 4; ++f) {
                Xb[f][0] += Tb[f][v] * PT(v, W1);
                Xb[f][v] += Tb[f][v] * PT(v, W2);
                Xb[f][vp1] += Tb[f][v] * PT(v, W3);
            }
        }
#undef PT
#undef PA
}

/* mode 0: TLM (dflx, dxc from drho, drhovx/y/z); mode 1: adjoint (rho_b...
 * written from flx_h_b, GRD_xc_b). */
static void flux_lin(const orc_dims *D, const int *have_sgp, int adjoint, double dt, const double *const X_reg[4],
                     const double *const X_pl[4], const double *const dX_reg[4], const double *const dX_pl[4],
                     double *dflx, double *dflx_pl, double *dxc, double *dxc_pl, const double *flxb,
                     const double *flxb_pl, const double *xcb, const double *xcb_pl, double *const Xb_reg[4],
                     double *const Xb_pl[4], const double *GMTR_t, const double *GMTR_t_pl, const double *GMTR_a,
                     const double *GMTR_a_pl, const double *GMTR_p, const double *GMTR_p_pl)
{
    const int iall = D->gall_1d, gall = iall * iall, kall = D->kall, lall = D->lall, npl = D->gall_pl;
    const size_t plane = (size_t)gall * kall * lall, pplane = (size_t)npl * kall * LALL_PL;
    if (adjoint) {
        for (int f = 0; f < 4; ++f) memset(Xb_reg[f], 0, plane * sizeof(double));
        if (D->have_pl)
            for (int f = 0; f < 4; ++f) memset(Xb_pl[f], 0, pplane * sizeof(double));
    }
    for (int l = 0; l < lall; ++l)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < kall; ++k) {
            const size_t s = ((size_t)l * kall + k) * gall;
            const double *X[4], *dX[4] = {0, 0, 0, 0};
            double *Xb[4] = {0, 0, 0, 0};
            for (int f = 0; f < 4; ++f) {
                X[f] = X_reg[f] + s;
                if (!adjoint) dX[f] = dX_reg[f] + s;
                else Xb[f] = Xb_reg[f] + s;
            }
            flux_lin_slab(iall, have_sgp[l], l, lall, X, dt, D->eps, GMTR_t, GMTR_a, GMTR_p, adjoint, dX,
                          adjoint ? NULL : dflx + s, adjoint ? NULL : dxc + s, plane, adjoint ? flxb + s : NULL,
                          adjoint ? xcb + s : NULL, Xb);
        }
    if (D->have_pl)
        for (int l = 0; l < LALL_PL; ++l)
            for (int k = 0; k < kall; ++k) {
                const size_t s = ((size_t)l * kall + k) * npl;
                const double *X[4], *dX[4] = {0, 0, 0, 0};
                double *Xb[4] = {0, 0, 0, 0};
                for (int f = 0; f < 4; ++f) {
                    X[f] = X_pl[f] + s;
                    if (!adjoint) dX[f] = dX_pl[f] + s;
                    else Xb[f] = Xb_pl[f] + s;
                }
                flux_lin_pole(npl, l, X, dt, D->eps, GMTR_t_pl, GMTR_a_pl, GMTR_p_pl, adjoint, dX,
                              adjoint ? NULL : dflx_pl + s, adjoint ? NULL : dxc_pl + s, pplane,
                              adjoint ? flxb_pl + s : NULL, adjoint ? xcb_pl + s : NULL, Xb);
            }
}

void orc_horizontal_flux_tlm(const orc_dims *D, const int *have_sgp, double *dflx, double *dflx_pl, double *dxc,
                             double *dxc_pl, const double *rho, const double *rho_pl, const double *rhovx,
                             const double *rhovx_pl, const double *rhovy, const double *rhovy_pl,
                             const double *rhovz, const double *rhovz_pl, const double *drho, const double *drho_pl,
                             const double *drhovx, const double *drhovx_pl, const double *drhovy,
                             const double *drhovy_pl, const double *drhovz, const double *drhovz_pl, double dt,
                             const double *GMTR_t, const double *GMTR_t_pl, const double *GMTR_a,
                             const double *GMTR_a_pl, const double *GMTR_p, const double *GMTR_p_pl)
{
    const double *Xr[4] = {rho, rhovx, rhovy, rhovz}, *Xp[4] = {rho_pl, rhovx_pl, rhovy_pl, rhovz_pl};
    const double *dXr[4] = {drho, drhovx, drhovy, drhovz}, *dXp[4] = {drho_pl, drhovx_pl, drhovy_pl, drhovz_pl};
    double *none[4] = {0, 0, 0, 0};
    flux_lin(D, have_sgp, 0, dt, Xr, Xp, dXr, dXp, dflx, dflx_pl, dxc, dxc_pl, NULL, NULL, NULL, NULL, none, none,
             GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, GMTR_p, GMTR_p_pl);
}

void orc_horizontal_flux_adjoint(const orc_dims *D, const int *have_sgp, const double *flx_h_b,
                                 const double *flx_h_b_pl, const double *GRD_xc_b, const double *GRD_xc_b_pl,
                                 double *rho_b, double *rho_b_pl, double *rhovx_b, double *rhovx_b_pl,
                                 double *rhovy_b, double *rhovy_b_pl, double *rhovz_b, double *rhovz_b_pl,
                                 const double *rho, const double *rho_pl, const double *rhovx, const double *rhovx_pl,
                                 const double *rhovy, const double *rhovy_pl, const double *rhovz,
                                 const double *rhovz_pl, double dt, const double *GMTR_t, const double *GMTR_t_pl,
                                 const double *GMTR_a, const double *GMTR_a_pl, const double *GMTR_p,
                                 const double *GMTR_p_pl)
{
    const double *Xr[4] = {rho, rhovx, rhovy, rhovz}, *Xp[4] = {rho_pl, rhovx_pl, rhovy_pl, rhovz_pl};
    const double *none[4] = {0, 0, 0, 0};
    double *Br[4] = {rho_b, rhovx_b, rhovy_b, rhovz_b}, *Bp[4] = {rho_b_pl, rhovx_b_pl, rhovy_b_pl, rhovz_b_pl};
    flux_lin(D, have_sgp, 1, dt, Xr, Xp, none, none, NULL, NULL, NULL, NULL, flx_h_b, flx_h_b_pl, GRD_xc_b,
             GRD_xc_b_pl, Br, Bp, GMTR_t, GMTR_t_pl, GMTR_a, GMTR_a_pl, GMTR_p, GMTR_p_pl);
}
#undef FL_TW
#undef FL_A
#undef FL_RA
