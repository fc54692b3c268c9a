// vertical_limiter_thuburn on gfx950.
//
// Reference: kernels/NICAM/dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243.
//
// The limited half-level value q_h(k) needs the allowable ranges Qout of
// levels k-1 and k; the reference carries Qout(k-1) from one level to the
// next (:140-141, :182-183), but every (g, k) is independent.  One work-item
// handles a run of KC consecutive levels of one column: it recomputes Qout of
// the level just below its run (the reference's peel at kmin plays that role
// for the first run), then marches up carrying Qout, q(k-1), q(k) and
// ck(k,1) in registers.  A wavefront covers 64 consecutive g at one level, so
// every load and store is contiguous.  The pole columns are appended as the
// last work-items.
#include "internal.h"

namespace nicam {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double sign_half(double x) { return copysign(0.5, x); }

struct VlimArgs {
    double *q_h;
    const double *q, *d, *ck;
    double *q_h_pl;
    const double *q_pl, *d_pl, *ck_pl;
    int gall, kall, lall, gall_pl, have_pl;
    int kc, nrun;  // levels per work-item, runs per column
    double eps, big;
};

// Levels [kb, ke) of one column are limited (kb >= kmin+1, 0-based).
// base = offset of level 0, s = stride between levels, ck2 = offset of the
// upper-face plane of ck.  POLE selects the pole loop's product grouping of
// CQin (VLIM:212-215 versus :126-129).
template <bool POLE>
__device__ __forceinline__ void vlim_run(double *__restrict__ q_h, const double *__restrict__ q,
                                         const double *__restrict__ d, const double *__restrict__ ck, size_t base,
                                         size_t s, size_t ck2, int kb, int ke, double EPS, double BIG)
{
    int k = kb - 1;  // first level whose Qout is needed
    double qkm1 = q[base + (size_t)(k - 1) * s];
    double qk = q[base + (size_t)k * s];
    double ck1k = ck[base + (size_t)k * s];
    double Qmin_km1 = 0.0, Qmax_km1 = 0.0;
    for (; k < ke; ++k) {
        const size_t o = base + (size_t)k * s;
        const double qkp1 = q[o + s];
        const double ck1kp1 = ck[o + s];
        const double ck2k = ck[ck2 + o];
        const double dk = d[o];
        const double inflagL = 0.5 - sign_half(ck1k);
        const double inflagU = 0.5 + sign_half(ck1kp1);
        const double Qin_minL = dmin(qk, qkm1) + (1.0 - inflagL) * BIG;
        const double Qin_minU = dmin(qk, qkp1) + (1.0 - inflagU) * BIG;
        const double Qin_maxL = dmax(qk, qkm1) - (1.0 - inflagL) * BIG;
        const double Qin_maxU = dmax(qk, qkp1) - (1.0 - inflagU) * BIG;
        const double qnext_min = dmin(dmin(Qin_minL, Qin_minU), qk);
        const double qnext_max = dmax(dmax(Qin_maxL, Qin_maxU), qk);
        const double Cin = (inflagL)*ck1k + (inflagU)*ck2k;
        const double Cout = (1.0 - inflagL) * ck1k + (1.0 - inflagU) * ck2k;
        double CQin_min, CQin_max;
        if constexpr (POLE) {
            CQin_max = (inflagL) * (ck1k * Qin_maxL) + (inflagU) * (ck2k * Qin_maxU);
            CQin_min = (inflagL) * (ck1k * Qin_minL) + (inflagU) * (ck2k * Qin_minU);
        } else {
            CQin_min = (inflagL)*ck1k * Qin_minL + (inflagU)*ck2k * Qin_minU;
            CQin_max = (inflagL)*ck1k * Qin_maxL + (inflagU)*ck2k * Qin_maxU;
        }
        const double zerosw = 0.5 - sign_half(fabs(Cout) - EPS);
        const double Qmin_k = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) *
                                  (1.0 - zerosw) + qk * zerosw;
        const double Qmax_k = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) *
                                  (1.0 - zerosw) + qk * zerosw;
        if (k >= kb) {  // VLIM:179-180
            const double qh = q_h[o];
            stream_store(&q_h[o], (inflagL)*dmax(dmin(qh, Qmax_km1), Qmin_km1) + (1.0 - inflagL) * dmax(dmin(qh, Qmax_k), Qmin_k));
        }
        Qmin_km1 = Qmin_k;
        Qmax_km1 = Qmax_k;
        qkm1 = qk;
        qk = qkp1;
        ck1k = ck1kp1;
    }
}

// A work-group is BLOCK consecutive columns at one run of levels.  The runs
// of one column block are consecutive work-groups on one XCD: each run also
// reads the two levels below it, which the run below has just brought into L2.
__global__ void __launch_bounds__(BLOCK) vlim_kernel(VlimArgs A)
{
    const long ncol_reg = (long)A.gall * A.lall;
    const long ncol = ncol_reg + (A.have_pl ? (long)A.gall_pl * kLallPl : 0);
    const long ncb = (ncol + BLOCK - 1) / BLOCK;  // column blocks
    const long L = xcd_contiguous(blockIdx.x, ncb * A.nrun);
    const int run = (int)(L % A.nrun);
    const long col = (L / A.nrun) * BLOCK + threadIdx.x;
    if (col >= ncol) return;
    const int kmin = 1, kmax = A.kall - 2;  // 0-based
    const int kb = kmin + 1 + run * A.kc;
    const int ke = min(kmax + 1, kb + A.kc);
    if (kb >= ke) return;
    if (col < ncol_reg) {
        const int g = (int)(col % A.gall), l = (int)(col / A.gall);
        const size_t base = (size_t)l * A.kall * A.gall + g;
        const size_t ck2 = (size_t)A.gall * A.kall * A.lall;
        vlim_run<false>(A.q_h, A.q, A.d, A.ck, base, A.gall, ck2, kb, ke, A.eps, A.big);
    } else {
        const int p = (int)(col - ncol_reg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        const size_t base = (size_t)l * A.kall * A.gall_pl + g;
        const size_t ck2 = (size_t)A.gall_pl * A.kall * kLallPl;
        vlim_run<true>(A.q_h_pl, A.q_pl, A.d_pl, A.ck_pl, base, A.gall_pl, ck2, kb, ke, A.eps, A.big);
    }
}

}  // namespace

int launch_vlim(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ck, const double *ck_pl)
{
    const Dims &D = c->d;
    VlimArgs A;
    A.q_h = q_h;
    A.q = q;
    A.d = d;
    A.ck = ck;
    A.q_h_pl = q_h_pl;
    A.q_pl = q_pl;
    A.d_pl = d_pl;
    A.ck_pl = ck_pl;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.eps = c->eps;
    A.big = c->big;
    const int nlev = D.kall - 3;  // limited levels kmin+1..kmax
    if (nlev <= 0) return NICAM_OK;
    // short runs give more independent work-items (measured best at gl7: 4)
    int kc = c->tune.vlim_kchunk > 0 ? c->tune.vlim_kchunk : 4;
    kc = std::min(kc, nlev);
    A.kc = kc;
    A.nrun = (nlev + kc - 1) / kc;
    const long ncol = (long)D.gall * D.lall + (D.have_pl ? (long)D.gall_pl * kLallPl : 0);
    const long nblk = (ncol + BLOCK - 1) / BLOCK * A.nrun;
    if (nblk > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(vlim_kernel, dim3((unsigned)nblk), dim3(BLOCK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
