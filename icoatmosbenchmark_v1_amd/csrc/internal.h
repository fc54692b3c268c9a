// Internal declarations shared by the nicam_tt translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <string>
#include <vector>

#include "nicam_tt.h"

namespace nicam {

// Fixed indices of the reference's data contract
// (dyn_vi_rhow_solver/src/problem_size.inc:14-23, 38-41, 80-132).
constexpr int kLallPl = 2;     // ADM_lall_pl
constexpr int kGmtrPNmax = 8;  // GMTR_p_nmax
constexpr int kGmtrTNmax = 5;  // GMTR_t_nmax
constexpr int kGmtrANmax = 12; // GMTR_a_nmax
constexpr int kGmtrANmaxPl = 18;
constexpr int kMaxGroup = 8;  // grids in one grouped launch

// Columns of a grid in work-groups of BLOCK work-items: the regular regions'
// columns first, then, from a work-group of their own, the pole columns, so
// that the field set (A.reg or A.pl) a work-group works on is uniform across
// it and its pointers stay in scalar registers.  Args has reg, pl, gall, lall,
// kall, gall_pl and have_pl; body(F, base, s) gets the field set, the
// column's first element and the level stride.
template <int BLOCK>
inline long column_blocks(long gall, long lall, long gall_pl, bool have_pl)
{
    return (gall * lall + BLOCK - 1) / BLOCK + (have_pl ? (gall_pl * kLallPl + BLOCK - 1) / BLOCK : 0);
}
// blk: the work-group's index in its grid; single-grid launches pass
// xcd_contiguous(blockIdx.x, gridDim.x), so that neighbouring column blocks,
// which share the cache lines at each level row's ends, run on one XCD.
template <int BLOCK, class Args, class Body>
__device__ __forceinline__ void column_run(const Args &A, long blk, Body &&body)
{
    const long nreg = (long)A.gall * A.lall;
    const long nbr = (nreg + BLOCK - 1) / BLOCK;
    if (blk < nbr) {
        const long t = blk * BLOCK + threadIdx.x;
        if (t >= nreg) return;
        const int g = (int)(t % A.gall), l = (int)(t / A.gall);
        body(A.reg, (size_t)l * A.kall * A.gall + g, (size_t)A.gall);
    } else if (A.have_pl) {
        const long p = (blk - nbr) * BLOCK + threadIdx.x;
        if (p >= (long)A.gall_pl * kLallPl) return;
        const int g = (int)(p % A.gall_pl), l = (int)(p / A.gall_pl);
        body(A.pl, (size_t)l * A.kall * A.gall_pl + g, (size_t)A.gall_pl);
    }
}

// Sizes of one grid, as the kernels see them.  All indices inside the
// kernels are 0-based: g = j*gall_1d + i, halo at i or j = 0 and gall_1d-1.
struct Dims {
    int gall_1d;
    int gall;  // gall_1d^2
    int kall;
    int lall;
    int gall_pl;
    int have_pl;
    size_t nreg() const { return (size_t)gall * kall * lall; }  // one regular field
    size_t npl() const { return (size_t)gall_pl * kall * kLallPl; }
};

}  // namespace nicam

// Launch-shape choices (0 = the library's default); set with nicam_ctx_set_tuning.
struct nicam_tune {
    int hflux_kper = 0;   // levels per work-group in horizontal_flux
    int vlim_kchunk = 0;  // levels per work-item in vertical_limiter_thuburn
    int vi_variant = 0;   // elimination-factor storage in vi_rhow_solver (1..4)
    int diff_kper = 0;    // levels per work-group in OPRT_diffusion
    int ddmp_kper = 0;    // levels per work-group in OPRT3D_divdamp
    int hlim_variant = 0; // horizontal limiter: 1 = streamed loads after the tile barrier, 2 = before it
};

namespace nicam {
// horizontal metrics in one precision (device pointers)
template <class T>
struct HMetrics {
    T *xr = nullptr, *gp = nullptr, *gt = nullptr, *ga = nullptr;
    T *xr_pl = nullptr, *gp_pl = nullptr, *gt_pl = nullptr, *ga_pl = nullptr;
};
// CONST_EPS and CONST_HUGE of the reference built with -DSINGLE
constexpr float kEpsF32 = 1.1920928955078125e-07f;  // epsilon(0._SP) = 2^-23
constexpr float kHugeF32 = 3.4028234663852886e+38f; // huge(0._SP)
}  // namespace nicam

struct nicam_ctx_s {
    nicam::Dims d{};
    nicam_tune tune{};
    std::vector<int> have_sgp_h;
    int *have_sgp = nullptr;  // device [lall]

    double eps = 0, big = 0;
    // CONST_EPS / CONST_HUGE of the _f32 entry points (nicam_set_const_f32)
    float eps32 = nicam::kEpsF32, big32 = nicam::kHugeF32;
    double grav = 9.80616, rdry = 287.0, cvdry = 717.5;
    int alpha = 1;

    hipStream_t stream = nullptr;
    int ptr_mode = NICAM_PTR_DEVICE;
    int device = 0;

    // horizontal metrics (device copies of the mod_misc module arrays)
    double *GRD_xr = nullptr, *GRD_xr_pl = nullptr;
    double *GMTR_p = nullptr, *GMTR_p_pl = nullptr;
    double *GMTR_t = nullptr, *GMTR_t_pl = nullptr;
    double *GMTR_a = nullptr, *GMTR_a_pl = nullptr;
    bool have_hmetrics = false;
    nicam::HMetrics<float> hm32;  // the same metrics rounded to FP32, for the _f32 entry points

    // vertical metrics
    double *RGAM = nullptr, *RGAM_pl = nullptr;
    double *RGAMH = nullptr, *RGAMH_pl = nullptr;
    double *RGSGAM2 = nullptr, *RGSGAM2_pl = nullptr;
    double *RGSGAM2H = nullptr, *RGSGAM2H_pl = nullptr;
    double *GSGAM2H = nullptr, *GSGAM2H_pl = nullptr;
    bool have_vmetrics = false;
    // RGAM, RGAMH, RGSGAM2 rounded to FP32, for the mixed solver's source term
    float *RGAM32 = nullptr, *RGAM32_pl = nullptr, *RGAMH32 = nullptr, *RGAMH32_pl = nullptr;
    float *RGSGAM2_32 = nullptr, *RGSGAM2_32_pl = nullptr;
    // vertical metrics of OPRT3D_divdamp (RGAM/RGAMH above are shared)
    double *RGSQRTH = nullptr, *RGSQRTH_pl = nullptr;
    double *C2WfactGz = nullptr, *C2WfactGz_pl = nullptr;  // (gall,kall,6,lall), (gall_pl,kall,6,2)
    bool have_vmetrics3d = false;
    double *rdgz = nullptr;  // GRD_rdgz (kall)
    bool have_rdgz = false;

    // vertical grid factors (kall each)
    double *rdgzh = nullptr, *afact = nullptr, *bfact = nullptr;
    bool have_vgrid = false;
    double *cfact = nullptr, *dfact = nullptr;  // cell wall -> cell centre (nicam_grd_setup only)
    bool have_cdfact = false;

    // tridiagonal coefficients (module copies of mod_vi.f90:47-52)
    double *Mc = nullptr, *Mc_pl = nullptr;
    double *Ml = nullptr, *Ml_pl = nullptr;
    double *Mu = nullptr, *Mu_pl = nullptr;
    bool have_coef = false;

    // on-device halo fill (nicam_comm_setup): encoded destination/source
    // points, the two COMM_var pole-centre entries first, then halo_n entries
    int *halo_dst = nullptr, *halo_src = nullptr;
    int halo_n = 0;
    int rlevel = -1;
    // partners of the faces regions share (nicam_comm_setup; the conservative update)
    int *arc_slot = nullptr, *arc_tab = nullptr, *arc_pole = nullptr;
    // the halo fill's transpose (nicam_comm_setup): the sources in ascending
    // order, each with its destinations (CSR, ascending codes)
    int *hadj_src = nullptr, *hadj_off = nullptr, *hadj_dst = nullptr;
    int hadj_nsrc = 0;

    // partial sums of nicam_tracer_mass ([kall][lall]), allocated on its first call
    double *mass_part = nullptr;

    // the Thomas factorisation of Mc/Ml/Mu (beta(k), gamma(k)), formed on the
    // device whenever the coefficients are set
    double *vi_beta = nullptr, *vi_beta_pl = nullptr, *vi_gam = nullptr, *vi_gam_pl = nullptr;
    double *vi_tbeta = nullptr, *vi_tbeta_pl = nullptr, *vi_tgam = nullptr, *vi_tgam_pl = nullptr;  // transposed

    // staging buffers for NICAM_PTR_HOST, one per argument slot
    std::vector<void *> stage;
    std::vector<size_t> stage_bytes;

    size_t device_bytes = 0;
};

namespace nicam {

// error reporting (thread-local last message)
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define NICAM_HIP(call)                                          \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::nicam::hip_fail(e_, #call); \
    } while (0)

// Hardware dispatch sends work-group b to XCD b % 8, and each XCD has its own
// L2.  For a launch of n work-groups whose neighbours in logical order share
// data (halo rows, the level above and below), this maps the hardware index
// to a logical one so that each XCD runs one contiguous logical range; the
// last n % 8 work-groups keep their index.
__device__ __forceinline__ long xcd_contiguous(long b, long n)
{
    constexpr long nxcd = 8;
    const long per = n / nxcd;
    return (per > 0 && b < per * nxcd) ? (b % nxcd) * per + b / nxcd : b;
}

// Stores of outputs that no work-item of the same launch reads back carry the
// nontemporal hint, so output streams of hundreds of MB do not evict the halo
// rows and metrics that neighbouring work-groups re-read from L2.  Measured at
// gl7 it shortens the limiters and the VI solve by 6-10 %; the flux kernel,
// whose 15 output streams are most of its traffic, is faster with ordinary
// stores and does not use it.
template <class T>
__device__ __forceinline__ void stream_store(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// x / 3.0, correctly rounded, without a division: q0 = x*RN(1/3), then one
// FMA correction with the exact residual x - 3*q0 (the Markstein-style
// correction for a divisor known in advance).  The exact quotient of a double
// by 3 has denominator 3, so it is never within 2^-54 ulp of a rounding
// midpoint, while the corrected value differs from it by at most that: the
// result is the IEEE quotient.  Zero keeps its sign and infinities pass
// through, as in the division.
__device__ __forceinline__ double div3(double x)
{
    const double c = 1.0 / 3.0;
    const double q0 = x * c;
    const double r = __builtin_fma(-3.0, q0, x);
    const double q1 = __builtin_copysign(__builtin_fma(r, c, q0), x);
    return __builtin_isfinite(x) ? q1 : x;
}

// launchers (asynchronous on ctx->stream, device pointers)
int launch_gradient(nicam_ctx c, double *grad, double *grad_pl, const double *scl, const double *scl_pl,
                    const double *coef_grad, const double *coef_grad_pl);
int launch_courant(nicam_ctx c, double *ch, double *ch_pl, double *cmask, double *cmask_pl, const double *flx_h,
                   const double *flx_h_pl, const double *rhog, const double *rhog_pl);
int launch_remap(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl, const double *grad,
                 const double *grad_pl, const double *cmask, const double *cmask_pl, const double *xc,
                 const double *xc_pl, const double *x, const double *x_pl);
int launch_update(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_h, const double *flx_h_pl,
                  const double *q_a, const double *q_a_pl);
int launch_vertical_mass_flux(nicam_ctx c, double *flx_v, double *flx_v_pl, const double *vx, const double *vx_pl,
                              const double *vy, const double *vy_pl, const double *vz, const double *vz_pl,
                              const double *w, const double *w_pl, double dt);
int launch_vertical_courant(nicam_ctx c, double *ck, double *ck_pl, const double *flx_v, const double *flx_v_pl,
                            const double *rhog, const double *rhog_pl);
int launch_vertical_face(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl);
int launch_vertical_courant_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *rhog_b, double *rhog_b_pl,
                                    const double *ck_b, const double *ck_b_pl, const double *flx_v,
                                    const double *flx_v_pl, const double *rhog, const double *rhog_pl);
int launch_vertical_face_adjoint(nicam_ctx c, double *q_b, double *q_b_pl, const double *q_h_b, const double *q_h_b_pl);
int launch_vertical_update_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *q_h_b, double *q_h_b_pl,
                                   const double *rq_b, const double *rq_b_pl, const double *flx_v,
                                   const double *flx_v_pl, const double *q_h, const double *q_h_pl);
int launch_vi_matrix(nicam_ctx c, const double *eth, const double *eth_pl, const double *g_tilde,
                     const double *g_tilde_pl, double dt);
int launch_vertical_update(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_v, const double *flx_v_pl,
                           const double *q_h, const double *q_h_pl);
int launch_to_float(nicam_ctx c, const double *src, float *dst, size_t n);
int launch_hflux_f32(nicam_ctx c, float *flx_h, float *flx_h_pl, float *xc, float *xc_pl, const float *rho,
                     const float *rho_pl, const float *rvx, const float *rvx_pl, const float *rvy,
                     const float *rvy_pl, const float *rvz, const float *rvz_pl, float dt);
int launch_hflux(nicam_ctx c, double *flx_h, double *flx_h_pl, double *xc, double *xc_pl,
                 const double *rho, const double *rho_pl, const double *rvx, const double *rvx_pl,
                 const double *rvy, const double *rvy_pl, const double *rvz, const double *rvz_pl,
                 double dt);
int launch_hlim(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                const double *Qpost, const double *Qpost_pl);
int launch_vlim(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ck, const double *ck_pl);
// halo map (comm.cpp) and on-device halo fill (comm.hip)
// A halo point: kind 0 = regular (l, g), kind 1 = pole (l_pl, v); 0-based.
struct HaloPoint {
    int kind, l, g;
    bool operator<(const HaloPoint &o) const
    {
        return kind != o.kind ? kind < o.kind : l != o.l ? l < o.l : g < o.g;
    }
};
// dst/src: the exchange (later singular copies already composed); pole_src:
// the point each pole centre (north, south) takes before COMM_var's exchange.
int build_halo_map(int rlevel, int gall_1d, int gall_pl, std::vector<HaloPoint> &dst, std::vector<HaloPoint> &src,
                   std::vector<int> &have_sgp, HaloPoint pole_src[2]);
int comm_setup(nicam_ctx c, int rlevel);
// per edge cell 6 face codes (see comm.cpp), slot[l*gall+g] = edge-cell index or -1, pole faces
int build_arc_partners(int rlevel, int gall_1d, int gall_pl, const std::vector<HaloPoint> &dst,
                       const std::vector<HaloPoint> &src, const std::vector<int> &have_sgp, std::vector<int> &slot,
                       std::vector<int> &tab, std::vector<int> &pole_tab);
// corners W, N, E, S of the 10 diamonds -> icosahedron vertex class; returns the class count
int diamond_corner_classes(int cls[10][4]);
// var(gall,kmax,lall,nvar), var_pl(gall_pl,kmax,2,nvar); pole_fill: COMM_var
int launch_halo_fill(nicam_ctx c, double *var, double *var_pl, int kmax, int nvar, bool pole_fill);
// adjoints of the glue and of the halo fill (glue_adjoint.hip)
int halo_adjoint_setup(nicam_ctx c, const std::vector<int> &dst, const std::vector<int> &src);
int launch_halo_adjoint(nicam_ctx c, double *var, double *var_pl, int kmax, int nvar);
int launch_gradient_adjoint(nicam_ctx c, double *scl_b, double *scl_b_pl, const double *grad_b,
                            const double *grad_b_pl, const double *coef_grad, const double *coef_grad_pl);
int launch_courant_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *rhog_b, double *rhog_b_pl,
                           const double *ch_b, const double *ch_b_pl, const double *flx_h, const double *flx_h_pl,
                           const double *rhog, const double *rhog_pl);
int launch_remap_adjoint(nicam_ctx c, double *q_b, double *q_b_pl, double *grad_b, double *grad_b_pl, double *xc_b,
                         double *xc_b_pl, const double *q_a_b, const double *q_a_b_pl, const double *grad,
                         const double *grad_pl, const double *cmask, const double *cmask_pl, const double *xc,
                         const double *xc_pl, const double *x, const double *x_pl);
int launch_update_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *q_a_b, double *q_a_b_pl,
                          const double *rhogq_b, const double *rhogq_b_pl, const double *flx_h,
                          const double *flx_h_pl, const double *q_a, const double *q_a_pl);
int launch_halo_fill_f32(nicam_ctx c, float *var, float *var_pl, int kmax, int nvar, bool pole_fill);
int launch_hlim_f32(nicam_ctx c, float *q_a, float *q_a_pl, const float *q, const float *q_pl, const float *d,
                    const float *d_pl, const float *ch, const float *ch_pl, const float *cmask,
                    const float *cmask_pl, float *Qprev, float *Qprev_pl, const float *Qpost,
                    const float *Qpost_pl);
int launch_vlim_f32(nicam_ctx c, float *q_h, float *q_h_pl, const float *q, const float *q_pl, const float *d,
                    const float *d_pl, const float *ck, const float *ck_pl);

int launch_diffusion(nicam_ctx c, double *dscl, double *dscl_pl, const double *scl, const double *scl_pl,
                     const double *kh, const double *kh_pl, const double *coef_intp, const double *coef_intp_pl,
                     const double *coef_diff, const double *coef_diff_pl);
int launch_divdamp(nicam_ctx c, double *const out[3], double *const out_pl[3], const double *const v[3],
                   const double *const v_pl[3], const double *w, const double *w_pl, const double *coef_intp,
                   const double *coef_intp_pl, const double *coef_diff, const double *coef_diff_pl,
                   const double *rgsqrth, const double *rgsqrth_pl, const double *rgam, const double *rgam_pl,
                   const double *rgamh, const double *rgamh_pl, const double *c2w, const double *c2w_pl,
                   const double *rdgz_dev);
int launch_hflux_adjoint(nicam_ctx c, const double *flxb, const double *flxb_pl, const double *xcb,
                         const double *xcb_pl, double *const Xb[4], double *const Xb_pl[4], const double *const X[4],
                         const double *const X_pl[4], double dt);
int launch_hlim_mask(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                     const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                     const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                     const double *Qpost, const double *Qpost_pl, uint8_t *mask, uint8_t *mask_pl);
int launch_hlim_adjoint(nicam_ctx c, double *qab, double *qab_pl, double *qb, double *qb_pl, const double *Qprevb,
                        const double *Qprevb_pl, double *Qpostb, double *Qpostb_pl, const double *q_a,
                        const double *q_a_pl, const double *q, const double *q_pl, const double *d,
                        const double *d_pl, const double *ch, const double *ch_pl, const double *cmask,
                        const double *cmask_pl, const double *Qpost, const double *Qpost_pl, const uint8_t *mask,
                        const uint8_t *mask_pl, double *chb, double *chb_pl);
int launch_vlim_mask(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                     const double *d, const double *d_pl, const double *ck, const double *ck_pl, uint32_t *mask,
                     uint32_t *mask_pl);
int launch_vlim_adjoint(nicam_ctx c, double *qhb, double *qhb_pl, double *qb, double *qb_pl, const double *q_h,
                        const double *q_h_pl, const double *q, const double *q_pl, const double *d,
                        const double *d_pl, const double *ck, const double *ck_pl, const uint32_t *mask,
                        const uint32_t *mask_pl, double *ckb, double *ckb_pl);
int launch_vi_adjoint(nicam_ctx c, double *wb, double *wb_pl, double *w0b, double *w0b_pl, double *p0b,
                      double *p0b_pl, double *r0b, double *r0b_pl, double *Srb, double *Srb_pl, double *Swb,
                      double *Swb_pl, double *Spb, double *Spb_pl, double dt);
int launch_vi_mixed(nicam_ctx c, double *w, double *w_pl, const float *w0, const float *w0_pl, const float *p0,
                    const float *p0_pl, const float *r0, const float *r0_pl, const float *Sr, const float *Sr_pl,
                    const float *Sw, const float *Sw_pl, const float *Sp, const float *Sp_pl, double dt);
int launch_vi_factor(nicam_ctx c);
// many tracers, several grids
int launch_hlim_batch(nicam_ctx c, int ntr, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                      const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                      const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                      const double *Qpost, const double *Qpost_pl);
int launch_hlim_batch_f32(nicam_ctx c, int ntr, float *q_a, float *q_a_pl, const float *q, const float *q_pl,
                          const float *d, const float *d_pl, const float *ch, const float *ch_pl,
                          const float *cmask, const float *cmask_pl, float *Qprev, float *Qprev_pl,
                          const float *Qpost, const float *Qpost_pl);
int launch_vlim_batch(nicam_ctx c, int ntr, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                      const double *d, const double *d_pl, const double *ck, const double *ck_pl);
int launch_vlim_batch_f32(nicam_ctx c, int ntr, float *q_h, float *q_h_pl, const float *q, const float *q_pl,
                          const float *d, const float *d_pl, const float *ck, const float *ck_pl);
int launch_hflux_group_f32(const nicam_ctx *cs, int n, const nicam_tracer_step_f32 *g, float dt, hipStream_t s);
int launch_hlim_group_f32(const nicam_ctx *cs, int n, int ntr, const nicam_tracer_step_f32 *g, hipStream_t s);
int launch_vlim_group_f32(const nicam_ctx *cs, int n, int ntr, const nicam_tracer_step_f32 *g, hipStream_t s);
int launch_vi_group_mixed(const nicam_ctx *cs, int n, const nicam_tracer_step_f32 *g, double dt, hipStream_t s);
int mass_scratch_init(int device);  // the per-device address of nicam_tracer_mass_noalloc's scratch
int launch_mass(nicam_ctx c, const double *rq, const double *rq_pl, double *mass, double *part);
int launch_mass_f32(nicam_ctx c, const float *rq, const float *rq_pl, double *mass, double *part);
int launch_update_cons(nicam_ctx c, double *rq, double *rq_pl, const double *flx, const double *qa);
int launch_update_cons_f32(nicam_ctx c, double *rq, double *rq_pl, const float *flx, const float *qa);
int launch_vertical_update_f32(nicam_ctx c, double *rq, double *rq_pl, const float *flx_v, const float *flx_v_pl,
                               const float *q_h, const float *q_h_pl);
int launch_vi(nicam_ctx c, double *w, double *w_pl, const double *w0, const double *w0_pl,
              const double *p0, const double *p0_pl, const double *r0, const double *r0_pl,
              const double *Sr, const double *Sr_pl, const double *Sw, const double *Sw_pl,
              const double *Sp, const double *Sp_pl, double dt);

}  // namespace nicam
