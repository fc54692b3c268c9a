// Adjoint (transposed) kernels of the tracer-transport and rho*w path.
//
// Each kernel computes the transpose of the forward kernel's derivative at
// fixed coefficients: the cotangent of every input that the forward kernel
// reads, from the cotangent of its outputs.  Operation orders are fixed so
// that the C restatement in the oracle gives the same bits.

#include "internal.h"

namespace nicam {
namespace {

// ---- vi_rhow_solver (dyn_vi_rhow_solver/src/mod_vi.f90:93-306) -------------
//
// With Mc/Ml/Mu, the metrics and dt fixed the solver is linear in
// (rhogw at kmin and kmax+1, rhogw0, preg0, rhog0, Srho, Sw, Spre).  Its
// transpose per column: the transposed tridiagonal system (diagonal Mc(k),
// sub-diagonal Mu(k-1), super-diagonal Ml(k+1)) by the same elimination,
// then the transpose of the source Sall (:194-206) and of the boundary terms
// (:209-212, :242-246).  One work-item per column, as in the forward solver.
// The forward sweep keeps its partial solution in w0b and its elimination
// factors in Srb; the backward sweep reads each before it writes that level's
// final cotangents, so no work array is needed.  The factors of the
// transposed system, g(k) = Ml(k)/beta(k-1) and beta(k) = Mc(k) - Mu(k-1) g(k),
// depend on the coefficients only: the factored variant (FACT) reads them
// from the factorisation nicam_vi_rhow_solver_putcoef stores, formed by the
// same operations, so its dependency chain per level holds one division.
struct VaField {
    double *wb, *w0b, *p0b, *r0b, *Srb, *Swb, *Spb;
    const double *Mc, *Ml, *Mu, *RGAM, *RGAMH, *RGSGAM2, *RGSGAM2H, *GSGAM2H;
    const double *tbeta, *tgam;  // the transposed system's stored factors
};

struct VaArgs {
    VaField reg, pl;
    const double *rdgzh, *afact, *bfact;
    long ncol;
    int gall, kall, lall, gall_pl, have_pl;
    double dt, CVovRt2, alpha, grav;
};

constexpr int BLOCK_VA = 256;

template <bool FACT>
__global__ void __launch_bounds__(BLOCK_VA) vi_adjoint_kernel(VaArgs A)
{
    column_run<BLOCK_VA>(A, xcd_contiguous(blockIdx.x, gridDim.x), [&](const VaField &F, size_t base, size_t s) {
        const int kall = A.kall, kmin = 1, kmax = kall - 2;  // 0-based
        const double dt = A.dt, CV = A.CVovRt2, grav = A.grav;
        auto at = [&](int k) { return base + (size_t)k * s; };

        // transposed elimination on kmin+1..kmax
        double beta = FACT ? F.tbeta[at(kmin + 1)] : F.Mc[at(kmin + 1)];
        double yp = (F.wb[at(kmin + 1)] * F.GSGAM2H[at(kmin + 1)]) / beta;
        F.w0b[at(kmin + 1)] = yp;
        for (int k = kmin + 2; k <= kmax; ++k) {
            const double mu_km1 = F.Mu[at(k - 1)];
            if constexpr (FACT) {
                beta = F.tbeta[at(k)];
            } else {
                const double g = F.Ml[at(k)] / beta;
                beta = F.Mc[at(k)] - mu_km1 * g;
                F.Srb[at(k)] = g;
            }
            yp = (F.wb[at(k)] * F.GSGAM2H[at(k)] - mu_km1 * yp) / beta;
            F.w0b[at(k)] = yp;
        }
        const double y_kmax = yp;

        // back substitution fused with the cotangents of each level, top down
        double yn = 0.0, g_k1 = 0.0;
        for (int k = kall - 1; k >= 0; --k) {
            const bool in_k = k >= kmin + 1 && k <= kmax, in_n = k + 1 >= kmin + 1 && k + 1 <= kmax;
            const size_t o = at(k);
            double yk = 0.0, gk = 0.0;
            if (in_k) {
                const double ypk = F.w0b[o];
                yk = (k == kmax) ? ypk : ypk - g_k1 * yn;
                if (k >= kmin + 2) gk = FACT ? F.tgam[o] : F.Srb[o];
            }
            const double sb = in_k ? yk * CV : 0.0;
            const double sbn = in_n ? yn * CV : 0.0;
            double pb = 0.0, rb = 0.0;
            if (in_k) {
                pb = -(sb * dt * A.rdgzh[k]);
                rb = -(sb * dt * grav * A.afact[k]);
            }
            if (in_n) {
                pb = pb + sbn * dt * A.rdgzh[k + 1];
                rb = rb - sbn * dt * grav * A.bfact[k + 1];
            }
            const double p0b = pb * F.RGSGAM2[o];
            const double rg = F.RGAM[o];
            const double r0b = rb * (rg * rg);
            const double rgh = F.RGAMH[o];
            const double sw = in_k ? sb * (rgh * rgh) : 0.0;
            F.p0b[o] = p0b;
            F.Spb[o] = dt * p0b;
            F.r0b[o] = r0b;
            F.Srb[o] = dt * r0b;
            F.w0b[o] = sw * A.alpha;
            F.Swb[o] = dt * sw;
            if (k == kmin) {
                const double bb = F.wb[o] * F.GSGAM2H[o] - F.Ml[at(kmin + 1)] * yn;
                F.wb[o] = bb * F.RGSGAM2H[o];
            } else if (k == kmax + 1) {
                const double bt = F.wb[o] * F.GSGAM2H[o] - F.Mu[at(kmax)] * y_kmax;
                F.wb[o] = bt * F.RGSGAM2H[o];
            } else if (in_k) {
                F.wb[o] = 0.0;
            }
            yn = yk;
            g_k1 = gk;
        }
    });
}

// The factored column in chunks of VC levels: each chunk's inputs are loaded
// before the previous chunk's outputs are stored, so a column waits for
// memory once per chunk rather than once per level (a launch holds about
// three waves per SIMD, too few to hide a load per level).  The operations
// and their order are those of vi_adjoint_kernel<true>.
constexpr int VC = 8;

__global__ void __launch_bounds__(BLOCK_VA) vi_adjoint_kernel_chunked(VaArgs A)
{
    column_run<BLOCK_VA>(A, xcd_contiguous(blockIdx.x, gridDim.x), [&](const VaField &F, size_t base, size_t s) {
        const int kall = A.kall, kmin = 1, kmax = kall - 2;  // 0-based
        const double dt = A.dt, CV = A.CVovRt2, grav = A.grav;
        auto at = [&](int k) { return base + (size_t)k * s; };

        // transposed elimination on kmin+1..kmax
        double beta = F.tbeta[at(kmin + 1)];
        double yp = (F.wb[at(kmin + 1)] * F.GSGAM2H[at(kmin + 1)]) / beta;
        F.w0b[at(kmin + 1)] = yp;
        {
            double mu[VC], tb[VC], wv[VC], gv[VC];
            auto load = [&](int k0, double *m_, double *t_, double *w_, double *g_) {
    #pragma unroll
                for (int c = 0; c < VC; ++c) {
                    const int k = k0 + c;
                    if (k <= kmax) {
                        m_[c] = F.Mu[at(k - 1)];
                        t_[c] = F.tbeta[at(k)];
                        w_[c] = F.wb[at(k)];
                        g_[c] = F.GSGAM2H[at(k)];
                    }
                }
            };
            load(kmin + 2, mu, tb, wv, gv);
            for (int k0 = kmin + 2; k0 <= kmax; k0 += VC) {
                double nmu[VC], ntb[VC], nwv[VC], ngv[VC];
                if (k0 + VC <= kmax) load(k0 + VC, nmu, ntb, nwv, ngv);
    #pragma unroll
                for (int c = 0; c < VC; ++c) {
                    const int k = k0 + c;
                    if (k <= kmax) {
                        yp = (wv[c] * gv[c] - mu[c] * yp) / tb[c];
                        F.w0b[at(k)] = yp;
                    }
                }
    #pragma unroll
                for (int c = 0; c < VC; ++c) {
                    mu[c] = nmu[c];
                    tb[c] = ntb[c];
                    wv[c] = nwv[c];
                    gv[c] = ngv[c];
                }
            }
        }
        const double y_kmax = yp;

        // back substitution fused with the cotangents of each level, top down, VC levels a chunk
        double yn = 0.0, g_k1 = 0.0;
        double w0[VC], tg[VC], rs[VC], rgv[VC], rhv[VC], rz[VC], rzn[VC], af[VC], bfn[VC];
        auto load = [&](int k0, double *w0_, double *tg_, double *rs_, double *rg_, double *rh_, double *rz_,
                        double *rzn_, double *af_, double *bfn_) {
    #pragma unroll
            for (int c = 0; c < VC; ++c) {
                const int k = k0 - c;
                if (k >= 0) {
                    const size_t o = at(k);
                    const bool in_k = k >= kmin + 1 && k <= kmax, in_n = k + 1 >= kmin + 1 && k + 1 <= kmax;
                    w0_[c] = in_k ? F.w0b[o] : 0.0;
                    tg_[c] = (in_k && k >= kmin + 2) ? F.tgam[o] : 0.0;
                    rs_[c] = F.RGSGAM2[o];
                    rg_[c] = F.RGAM[o];
                    rh_[c] = F.RGAMH[o];
                    rz_[c] = in_k ? A.rdgzh[k] : 0.0;
                    af_[c] = in_k ? A.afact[k] : 0.0;
                    rzn_[c] = in_n ? A.rdgzh[k + 1] : 0.0;
                    bfn_[c] = in_n ? A.bfact[k + 1] : 0.0;
                }
            }
        };
        load(kall - 1, w0, tg, rs, rgv, rhv, rz, rzn, af, bfn);
        for (int k0 = kall - 1; k0 >= 0; k0 -= VC) {
            double nw0[VC], ntg[VC], nrs[VC], nrg[VC], nrh[VC], nrz[VC], nrzn[VC], naf[VC], nbfn[VC];
            if (k0 - VC >= 0) load(k0 - VC, nw0, ntg, nrs, nrg, nrh, nrz, nrzn, naf, nbfn);
    #pragma unroll
            for (int c = 0; c < VC; ++c) {
                const int k = k0 - c;
                if (k < 0) break;
                const bool in_k = k >= kmin + 1 && k <= kmax, in_n = k + 1 >= kmin + 1 && k + 1 <= kmax;
                const size_t o = at(k);
                double yk = 0.0, gk = 0.0;
                if (in_k) {
                    yk = (k == kmax) ? w0[c] : w0[c] - g_k1 * yn;
                    if (k >= kmin + 2) gk = tg[c];
                }
                const double sb = in_k ? yk * CV : 0.0;
                const double sbn = in_n ? yn * CV : 0.0;
                double pb = 0.0, rb = 0.0;
                if (in_k) {
                    pb = -(sb * dt * rz[c]);
                    rb = -(sb * dt * grav * af[c]);
                }
                if (in_n) {
                    pb = pb + sbn * dt * rzn[c];
                    rb = rb - sbn * dt * grav * bfn[c];
                }
                const double p0b = pb * rs[c];
                const double rg = rgv[c];
                const double r0b = rb * (rg * rg);
                const double rgh = rhv[c];
                const double sw = in_k ? sb * (rgh * rgh) : 0.0;
                F.p0b[o] = p0b;
                F.Spb[o] = dt * p0b;
                F.r0b[o] = r0b;
                F.Srb[o] = dt * r0b;
                F.w0b[o] = sw * A.alpha;
                F.Swb[o] = dt * sw;
                if (k == kmin) {
                    const double bb = F.wb[o] * F.GSGAM2H[o] - F.Ml[at(kmin + 1)] * yn;
                    F.wb[o] = bb * F.RGSGAM2H[o];
                } else if (k == kmax + 1) {
                    const double bt = F.wb[o] * F.GSGAM2H[o] - F.Mu[at(kmax)] * y_kmax;
                    F.wb[o] = bt * F.RGSGAM2H[o];
                } else if (in_k) {
                    F.wb[o] = 0.0;
                }
                yn = yk;
                g_k1 = gk;
            }
    #pragma unroll
            for (int c = 0; c < VC; ++c) {
                w0[c] = nw0[c]; tg[c] = ntg[c]; rs[c] = nrs[c]; rgv[c] = nrg[c]; rhv[c] = nrh[c];
                rz[c] = nrz[c]; rzn[c] = nrzn[c]; af[c] = naf[c]; bfn[c] = nbfn[c];
            }
        }
    });
}

// The same column with kall fixed at compile time (used at 42): the partial
// solution and the elimination factors stay in registers (the loops are
// unrolled), so the output arrays are written once and never re-read.  The
// operations and their order are those of vi_adjoint_kernel.
constexpr int BLOCK_VAR = 64;

template <int KC>
__global__ void __launch_bounds__(BLOCK_VAR) vi_adjoint_kernel_reg(VaArgs A)
{
    column_run<BLOCK_VAR>(A, xcd_contiguous(blockIdx.x, gridDim.x), [&](const VaField &F, size_t base, size_t s) {
        constexpr int kmin = 1, kmax = KC - 2;
        const double dt = A.dt, CV = A.CVovRt2, grav = A.grav;
        auto at = [&](int k) { return base + (size_t)k * s; };
        double yp[KC], gm[KC];
        double beta = F.Mc[at(kmin + 1)];
        double ylast = (F.wb[at(kmin + 1)] * F.GSGAM2H[at(kmin + 1)]) / beta;
        yp[kmin + 1] = ylast;
    #pragma unroll
        for (int k = kmin + 2; k <= kmax; ++k) {
            const double mu_km1 = F.Mu[at(k - 1)];
            const double g = F.Ml[at(k)] / beta;
            beta = F.Mc[at(k)] - mu_km1 * g;
            ylast = (F.wb[at(k)] * F.GSGAM2H[at(k)] - mu_km1 * ylast) / beta;
            yp[k] = ylast;
            gm[k] = g;
            if constexpr (KC > 64) __builtin_amdgcn_sched_barrier(0);
        }
        const double y_kmax = ylast;
        double yn = 0.0;
    #pragma unroll
        for (int k = KC - 1; k >= 0; --k) {
            const bool in_k = k >= kmin + 1 && k <= kmax, in_n = k + 1 >= kmin + 1 && k + 1 <= kmax;
            const size_t o = at(k);
            double yk = 0.0;
            if (in_k) yk = (k == kmax) ? yp[k] : yp[k] - gm[k + 1] * yn;
            const double sb = in_k ? yk * CV : 0.0;
            const double sbn = in_n ? yn * CV : 0.0;
            double pb = 0.0, rb = 0.0;
            if (in_k) {
                pb = -(sb * dt * A.rdgzh[k]);
                rb = -(sb * dt * grav * A.afact[k]);
            }
            if (in_n) {
                pb = pb + sbn * dt * A.rdgzh[k + 1];
                rb = rb - sbn * dt * grav * A.bfact[k + 1];
            }
            const double p0b = pb * F.RGSGAM2[o];
            const double rg = F.RGAM[o];
            const double r0b = rb * (rg * rg);
            const double rgh = F.RGAMH[o];
            const double sw = in_k ? sb * (rgh * rgh) : 0.0;
            F.p0b[o] = p0b;
            F.Spb[o] = dt * p0b;
            F.r0b[o] = r0b;
            F.Srb[o] = dt * r0b;
            F.w0b[o] = sw * A.alpha;
            F.Swb[o] = dt * sw;
            if (k == kmin) {
                const double bb = F.wb[o] * F.GSGAM2H[o] - F.Ml[at(kmin + 1)] * yn;
                F.wb[o] = bb * F.RGSGAM2H[o];
            } else if (k == kmax + 1) {
                const double bt = F.wb[o] * F.GSGAM2H[o] - F.Mu[at(kmax)] * y_kmax;
                F.wb[o] = bt * F.RGSGAM2H[o];
            } else if (in_k) {
                F.wb[o] = 0.0;
            }
            if (in_k) yn = yk;
            if constexpr (KC > 64) __builtin_amdgcn_sched_barrier(0);
        }
    });
}

// ---- vertical_limiter_thuburn (dyn_vert_adv_limiter/src/mod_src_tracer.f90:41-243)
//
// At fixed ck and d the limited face value is q_h's input (bit clear in the
// forward's mask), or Qout_min or Qout_max of the upwind level j (bit set;
// the lower bound when the stored output equals Qout_min(j)).  Both bounds
// are piecewise linear in q(j-1), q(j), q(j+1); their rows follow the
// forward's own comparisons (dmin(a,b) = b<a?b:a, dmax(a,b) = b>a?b:a).
// One work-item per column marches up the faces and scatters each face's
// cotangent into a window of four q levels held in registers; a level is
// stored once no later face reaches it.
struct VlimLvl {
    double qmin, qmax;
    double n0, n1, n2;  // d Qout_min / d q(j-1), q(j), q(j+1)
    double x0, x1, x2;  // d Qout_max / d q(j-1), q(j), q(j+1)
    double cn1, cn2;    // d Qout_min / d ck(j,1), ck(j,2)
    double cx1, cx2;    // d Qout_max / d ck(j,1), ck(j,2)
};

__device__ __forceinline__ double sgn_half(double x) { return copysign(0.5, x); }
__device__ __forceinline__ double dmn(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmx(double a, double b) { return (b > a) ? b : a; }

template <bool POLE>
__device__ __forceinline__ VlimLvl vlim_level(double qkm1, double qk, double qkp1, double ck1, double ck1p,
                                              double ck2k, double dk, double EPS, double BIG)
{
    VlimLvl r;
    const double inflagL = 0.5 - sgn_half(ck1), inflagU = 0.5 + sgn_half(ck1p);
    const double Qin_minL = dmn(qk, qkm1) + (1.0 - inflagL) * BIG;
    const double Qin_minU = dmn(qk, qkp1) + (1.0 - inflagU) * BIG;
    const double Qin_maxL = dmx(qk, qkm1) - (1.0 - inflagL) * BIG;
    const double Qin_maxU = dmx(qk, qkp1) - (1.0 - inflagU) * BIG;
    const double tmin = dmn(Qin_minL, Qin_minU), tmax = dmx(Qin_maxL, Qin_maxU);
    const double qnext_min = dmn(tmin, qk), qnext_max = dmx(tmax, qk);
    const double Cin = inflagL * ck1 + inflagU * ck2k;
    const double Cout = (1.0 - inflagL) * ck1 + (1.0 - inflagU) * ck2k;
    double CQin_min, CQin_max;
    if constexpr (POLE) {
        CQin_max = inflagL * (ck1 * Qin_maxL) + inflagU * (ck2k * Qin_maxU);
        CQin_min = inflagL * (ck1 * Qin_minL) + inflagU * (ck2k * Qin_minU);
    } else {
        CQin_min = inflagL * ck1 * Qin_minL + inflagU * ck2k * Qin_minU;
        CQin_max = inflagL * ck1 * Qin_maxL + inflagU * ck2k * Qin_maxU;
    }
    const double zerosw = 0.5 - sgn_half(fabs(Cout) - EPS);
    r.qmin = ((qk - qnext_max) + qnext_max * (Cin + Cout - dk) - CQin_max) / (Cout + zerosw) * (1.0 - zerosw) +
             qk * zerosw;
    r.qmax = ((qk - qnext_min) + qnext_min * (Cin + Cout - dk) - CQin_min) / (Cout + zerosw) * (1.0 - zerosw) +
             qk * zerosw;
    const int maxL = (qkm1 > qk) ? 0 : 1, maxU = (qkp1 > qk) ? 2 : 1;
    const int minL = (qkm1 < qk) ? 0 : 1, minU = (qkp1 < qk) ? 2 : 1;
    const int nmax = (qk > tmax) ? 1 : ((Qin_maxU > Qin_maxL) ? maxU : maxL);
    const int nmin = (qk < tmin) ? 1 : ((Qin_minU < Qin_minL) ? minU : minL);
    const double fac = (1.0 - zerosw) / (Cout + zerosw);
    const double e = (Cin + Cout - dk) - 1.0;
    const double cL = inflagL * ck1, cU = inflagU * ck2k;
    auto rows = [&](int m, double &rn, double &rx) {
        const double dnx = (m == nmax) ? 1.0 : 0.0, dnn = (m == nmin) ? 1.0 : 0.0;
        const double cqx = ((m == maxL) ? cL : 0.0) + ((m == maxU) ? cU : 0.0);
        const double cqn = ((m == minL) ? cL : 0.0) + ((m == minU) ? cU : 0.0);
        rn = fac * (e * dnx - cqx);
        rx = fac * (e * dnn - cqn);
    };
    rows(0, r.n0, r.x0);
    rows(1, r.n1, r.x1);
    rows(2, r.n2, r.x2);
    r.n1 = (r.n1 + fac) + zerosw;
    r.x1 = (r.x1 + fac) + zerosw;
    // ck(j,1) enters Cin where inflagL = 1 (with CQin) and Cout elsewhere; ck(j,2) likewise with inflagU
    const bool z = zerosw != 0.0;
    r.cn1 = z ? 0.0 : (inflagL > 0.0 ? qnext_max - Qin_maxL : qnext_max - r.qmin) / Cout;
    r.cn2 = z ? 0.0 : (inflagU > 0.0 ? qnext_max - Qin_maxU : qnext_max - r.qmin) / Cout;
    r.cx1 = z ? 0.0 : (inflagL > 0.0 ? qnext_min - Qin_minL : qnext_min - r.qmax) / Cout;
    r.cx2 = z ? 0.0 : (inflagU > 0.0 ? qnext_min - Qin_minU : qnext_min - r.qmax) / Cout;
    return r;
}

struct VlaField {
    double *qhb, *qb, *ckb;  // ckb: cotangent of ck (NULL: not formed)
    const double *q_h, *q, *d, *ck;
    const uint32_t *mask;
};

struct VlaArgs {
    VlaField reg, pl;
    int gall, kall, lall, gall_pl, have_pl, nword;
    double eps, big;
};

// A work-group covers VCOL columns; each column is split into VRUN runs of
// faces.  The output cotangent of every face of the group's columns is first
// staged in LDS, so that q_h_b can be rewritten in place while other runs
// still read their neighbours' faces.  The work-item of run r writes q_h_b
// for its faces and q_b for its levels; a level's cotangent gathers faces
// m-1 .. m+2 in ascending order, the order of a sequential sweep, so the sums
// are the same bits whichever run computes them.
constexpr int VCOL = 32, VRUN = 8, BLOCK_VLA = VCOL * VRUN;

template <bool POLE>
__device__ __forceinline__ void vlim_adjoint_run(const VlaField &F, size_t base, size_t s, size_t ck2,
                                                 const uint32_t *mw, size_t mstride, int kall, double EPS,
                                                 double BIG, const double *s_ob, int r)
{
    const int kmin = 1, kmax = kall - 2;  // 0-based
    const int nf = kmax - kmin, kc = (nf + VRUN - 1) / VRUN;
    const int pb = min(r * kc, nf), pe = min(pb + kc, nf);
    if (pb >= pe) return;
    // levels written by this run; run 0 also takes the levels below its faces, the run with the last face
    // those above
    const int mb = (r == 0) ? 0 : kmin + 1 + pb;
    const int me = (pe == nf) ? kall : kmin + 1 + pe;
    const int kf = max(kmin + 1, mb - 1), kl = min(kmax, me + 1);  // faces that reach [mb, me)
    auto at = [&](int k) { return base + (size_t)k * s; };
    double qm = F.q[at(kf - 2)], q0 = F.q[at(kf - 1)], qp = F.q[at(kf)];
    double c1 = F.ck[at(kf - 1)], c1p = F.ck[at(kf)];
    VlimLvl lo = vlim_level<POLE>(qm, q0, qp, c1, c1p, F.ck[ck2 + at(kf - 1)], F.d[at(kf - 1)], EPS, BIG);
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;  // levels k-2 .. k+1
    double cp1 = 0.0, cp2 = 0.0, cc1 = 0.0, cc2 = 0.0;  // ck cotangents of levels k-1 and k
    const bool want_ck = F.ckb != nullptr;
    int wi = -1;
    uint32_t word = 0;
    for (int k = kf; k <= kl; ++k) {
        const int p = k - (kmin + 1);
        if ((p >> 5) != wi) {
            wi = p >> 5;
            word = mw[(size_t)wi * mstride];
        }
        qm = q0;
        q0 = qp;
        qp = F.q[at(k + 1)];
        c1 = c1p;
        c1p = F.ck[at(k + 1)];
        const VlimLvl hi = vlim_level<POLE>(qm, q0, qp, c1, c1p, F.ck[ck2 + at(k)], F.d[at(k)], EPS, BIG);
        const bool clipped = (word >> (p & 31)) & 1u;
        const double inflagL = 0.5 - sgn_half(c1);
        const double ob = s_ob[(size_t)k * VCOL];
        // row of the face: the selected bound's derivatives, placed at q(k-2..k) or q(k-1..k+1); written
        // with selects only (a pointer to either level's arrays would put both in scratch memory)
        double row0 = 0.0, row1 = 0.0, row2 = 0.0, row3 = 0.0;
        double cqh = 1.0;
        if (clipped) {
            cqh = 0.0;
            const double out = F.q_h[at(k)];
            const bool from_lo = inflagL > 0.0;
            const bool use_min = out == (from_lo ? lo.qmin : hi.qmin);
            const double e0 = from_lo ? (use_min ? lo.n0 : lo.x0) : (use_min ? hi.n0 : hi.x0);
            const double e1 = from_lo ? (use_min ? lo.n1 : lo.x1) : (use_min ? hi.n1 : hi.x1);
            const double e2 = from_lo ? (use_min ? lo.n2 : lo.x2) : (use_min ? hi.n2 : hi.x2);
            row0 = from_lo ? e0 : 0.0;
            row1 = from_lo ? e1 : e0;
            row2 = from_lo ? e2 : e1;
            row3 = from_lo ? 0.0 : e2;
            if (want_ck) {
                const double g1 = from_lo ? (use_min ? lo.cn1 : lo.cx1) : (use_min ? hi.cn1 : hi.cx1);
                const double g2 = from_lo ? (use_min ? lo.cn2 : lo.cx2) : (use_min ? hi.cn2 : hi.cx2);
                if (from_lo) {
                    cp1 = cp1 + g1 * ob;
                    cp2 = cp2 + g2 * ob;
                } else {
                    cc1 = cc1 + g1 * ob;
                    cc2 = cc2 + g2 * ob;
                }
            }
        }
        if (p >= pb && p < pe) F.qhb[at(k)] = cqh * ob;
        w0 = w0 + row0 * ob;
        w1 = w1 + row1 * ob;
        w2 = w2 + row2 * ob;
        w3 = w3 + row3 * ob;
        if (k - 2 >= mb && k - 2 < me) F.qb[at(k - 2)] = w0;  // no later face reaches level k-2
        w0 = w1;
        w1 = w2;
        w2 = w3;
        w3 = 0.0;
        if (want_ck) {  // faces k-1 and k reach level k-1's ck; none later
            if (k - 1 >= mb && k - 1 < me) {
                F.ckb[at(k - 1)] = cp1;
                F.ckb[ck2 + at(k - 1)] = cp2;
            }
            cp1 = cc1;
            cp2 = cc2;
            cc1 = cc2 = 0.0;
        }
        lo = hi;
    }
    if (kl - 1 >= mb && kl - 1 < me) F.qb[at(kl - 1)] = w0;
    if (kl >= mb && kl < me) F.qb[at(kl)] = w1;
    if (kl + 1 >= mb && kl + 1 < me) F.qb[at(kl + 1)] = w2;
    if (want_ck) {
        if (kl >= mb && kl < me) {
            F.ckb[at(kl)] = cp1;
            F.ckb[ck2 + at(kl)] = cp2;
        }
        // levels that no face reaches: below the first face's lower level and above the last face
        for (int j = mb; j < me; ++j)
            if (j < kmin || j > kmax) {
                F.ckb[at(j)] = 0.0;
                F.ckb[ck2 + at(j)] = 0.0;
            }
    }
}

// four waves per SIMD (36 bytes of spills): 0.315 ms at gl7 against 0.356 at three
__global__ void __launch_bounds__(BLOCK_VLA) __attribute__((amdgpu_waves_per_eu(4, 8))) vlim_adjoint_kernel(VlaArgs A)
{
    extern __shared__ double s_ob[];  // [kall][VCOL]
    const int c = threadIdx.x % VCOL, r = threadIdx.x / VCOL;
    // work-groups [0, nbr) take the regular columns, the rest the pole columns: the field set is uniform
    // across a work-group, its pointers in scalar registers
    const long nreg = (long)A.gall * A.lall;
    const long nbr = (nreg + VCOL - 1) / VCOL;
    auto run = [&](const VlaField &F, long t, int gall, int lall, auto pole) {
        const bool have = t < (long)gall * lall;
        const int g = (int)(t % gall), l = (int)(t / gall);
        const size_t base = (size_t)l * A.kall * gall + g, s = gall;
        if (have)
            for (int k = r; k < A.kall; k += VRUN) s_ob[(size_t)k * VCOL + c] = F.qhb[base + (size_t)k * s];
        __syncthreads();
        if (!have) return;
        const size_t ck2 = (size_t)gall * A.kall * lall;
        const uint32_t *mw = F.mask + (size_t)l * A.nword * gall + g;
        vlim_adjoint_run<decltype(pole)::value>(F, base, s, ck2, mw, (size_t)gall, A.kall, A.eps, A.big, s_ob + c, r);
    };
    if (blockIdx.x < nbr)
        run(A.reg, (long)blockIdx.x * VCOL + c, A.gall, A.lall, std::false_type{});
    else if (A.have_pl)
        run(A.pl, (long)(blockIdx.x - nbr) * VCOL + c, A.gall_pl, kLallPl, std::true_type{});
}

// ---- horizontal_limiter_thuburn (dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430)
//
// With Qout_post supplied and d, ch, cmask fixed the limiter is piecewise
// linear in (q_a faces 1-3, q, Qout_post).  Qout of an inner cell is affine in
// q given the selections of qnext and of its six arcs (14 terms), a halo
// cell's Qout is q, and a face value is its input, a Qout_post bound of the
// downwind cell or a Qin bound of its arc (the forward's mask says whether it
// changed; a changed value is attributed, in this order, to Qout_post min,
// Qout_post max, Qin min, Qin max, by equality with the stored output).
//
// One work-group per (k, l) slab, sweeping it in steps of HSW cells, with
// no atomics: every cotangent is a gather in a fixed order, the oracle's.
// Every cell a Qout row names lies in the hexagon of the row's owner
// {g-iall-1, g-iall, g-1, g, g+1, g+iall, g+iall+1} (the inner rows' arcs never
// reach further or hit the index clamps of HLIM:152-153), so each row sums its
// terms (bound mm = 0 then 1, each bound's terms in order) per hexagon cell
// into seven partial sums, and a cell's q_b is the sum of the partial sums of
// the rows of its hexagon, in cell order of the rows' owners, followed by the
// face cotangents attributed to it, in cell order of the faces' origins (arcs
// m = 0, 1, 2); Qout_post_b likewise.  An LDS ring of R = HSW + 2*iall + 3
// cells holds q and, per cell, its row's seven partial sums and the
// cotangents of its three own faces with their attribution (92 bytes).  Each
// step loads q 2*iall+3 cells ahead, forms rows and faces iall+1 cells ahead
// and gathers HSW cells; a face's q_a_b is read and rewritten by its origin's
// work-item only.
constexpr int HSW = 256;
constexpr int HMAXI = 160;
constexpr int HPOLE = 256;

struct HaArgs {
    double *qab, *qb, *Qpostb;
    double *chb;  // cotangent of ch (NULL: not formed)
    const double *Qprevb;
    const double *q_a, *q, *d, *ch, *cmask, *Qpost;
    const uint8_t *mask;
    const int *have_sgp;
    int iall, gall, kall, lall, ring;
    double eps, big;
    // poles
    double *qab_pl, *qb_pl, *Qpostb_pl, *chb_pl;
    const double *Qprevb_pl;
    const double *q_a_pl, *q_pl, *d_pl, *ch_pl, *cmask_pl, *Qpost_pl;
    const uint8_t *mask_pl;
    int gall_pl, have_pl;
};

// cells of arc m (0 = AI, 1 = AIJ, 2 = AJ) with origin p, in the forward's
// min/max chain order (HLIM:145-160, the index clamps of :152-153 and the
// singular-point stencil of :197-203)
__device__ __forceinline__ void ha_arc(int m, int p, int iall, bool sgp, int c[4])
{
    c[0] = p;
    if (m == 0) {
        c[1] = max(p - iall, 0); c[2] = p + 1; c[3] = p + iall + 1;
    } else if (m == 1) {
        if (sgp && p == 0) { c[1] = p + iall + 1; c[2] = p + iall + 2; c[3] = p + iall; }
        else { c[1] = p + 1; c[2] = p + iall + 1; c[3] = p + iall; }
    } else {
        c[1] = p + iall + 1; c[2] = p + iall; c[3] = max(p - 1, 0);
    }
}

// index 0..6 of cell g + off in the hexagon {-iall-1, -iall, -1, 0, 1, iall, iall+1}
__device__ __forceinline__ uint32_t hex_index(int off, int iall)
{
    return off < -1 ? (off == -iall - 1 ? 0u : 1u) : (off > 1 ? (off == iall ? 5u : 6u) : (uint32_t)(off + 3));
}

// Whether arc f (0..5, as ordered in a row) of a row's owner contains the
// owner's hexagon cell w: f0 {0,-iall,1,iall+1}, f1 {0,1,iall+1,iall},
// f2 {0,iall+1,iall,-1}, f3 {-1,-iall-1,0,iall}, f4 {-iall-1,-iall,0,-1},
// f5 {-iall,1,0,-iall-1} (not the singular-point row, whose f4 is
// {-iall-1,0,1,-1}).
__host__ __device__ constexpr bool arc_has(int f, int w)
{
    return w == 3 || (f == 0 && (w == 1 || w == 4 || w == 6)) || (f == 1 && (w == 4 || w == 5 || w == 6)) ||
           (f == 2 && (w == 2 || w == 5 || w == 6)) || (f == 3 && (w == 0 || w == 2 || w == 5)) ||
           (f == 4 && (w == 0 || w == 1 || w == 2)) || (f == 5 && (w == 0 || w == 1 || w == 4));
}

// The slot (0..3, in ha_arc's order) of arc m's cell at offset d from its
// origin, -1 if none; d by index: 0 iall+2, 1 iall+1, 2 iall, 3 1, 4 0, 5 -1,
// 6 -iall (the index clamps reach cell 0 only, which takes the general path).
__host__ __device__ constexpr int arc_slot(int m, int d)
{
    return d == 4 ? 0
                  : (m == 0 ? (d == 6 ? 1 : d == 3 ? 2 : d == 1 ? 3 : -1)
                            : (m == 1 ? (d == 3 ? 1 : d == 1 ? 2 : d == 2 ? 3 : -1)
                                      : (d == 1 ? 1 : d == 2 ? 2 : d == 5 ? 3 : -1)));
}
// ... of the singular-point arc (m = 1, origin 0): {0, iall+1, iall+2, iall}
__host__ __device__ constexpr int arc_slot_sgp(int d) { return d == 4 ? 0 : d == 1 ? 1 : d == 0 ? 2 : d == 2 ? 3 : -1; }

__global__ void __launch_bounds__(HPOLE) hlim_adjoint_pole(HaArgs A);

// four waves per SIMD (the LDS ring allows four work-groups per CU): 1.57 ms at gl7, against 1.87 ms at
// three and 2.25 ms at two, despite the registers the cap spills
// CHB: also the cotangent of ch.  PASSES: 1 the rows only (the cotangents of
// Qout_prev; no faces), 2 the faces only (of q_a; Qout_prev_b taken as zero),
// 3 both.  A pass left out adds nothing, so the two halves give the full
// pass's results, each bit for bit (its sums start at +0.0 and the other
// half's terms are +0.0 or -0.0).
template <bool CHB, int PASSES>
__global__ void __launch_bounds__(HSW) __attribute__((amdgpu_waves_per_eu(CHB ? 3 : 4, 8))) hlim_adjoint_kernel(HaArgs A)
{
    extern __shared__ double lds[];
    constexpr bool ROWS = (PASSES & 1) != 0, FACES = (PASSES & 2) != 0;
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall, R = A.ring;
    // q, with the first and last QP slots repeated past either end, so that a read within QP of a
    // cell needs no wrap
    const int QP = iall + 2;
    double *r_q = lds + QP;
    double *r_p = lds + R + 2 * QP;  // [7][R] the row's partial sums per hexagon cell
    double *r_ob = r_p + 7 * R;      // [3][R] face cotangents of the cell's own arcs
    uint32_t *r_fc = reinterpret_cast<uint32_t *>(r_ob + 3 * R);  // faces: attribution
    const int k = blockIdx.x % kall, l = blockIdx.x / kall;
    if (l >= lall) return;
    const size_t fb = ((size_t)l * kall + k) * gall;
    const size_t plane = (size_t)lall * kall * gall;
    const int tid = threadIdx.x;
    const bool sgp = A.have_sgp[l] != 0;
    const double BIG = A.big;

    // the hexagon index (3 bits per slot, in ha_arc's slot order) of the four cells of arc f of an
    // inner cell's row: f0..2 its own arcs, f3 the arc AI of (i-1,j), f4 AIJ of (i-1,j-1), f5 AJ of
    // (i,j-1); f4 of the singular-point row is {-iall-1, 0, 1, -1}
    constexpr uint32_t HEXTAB[6] = {3u | 1u << 3 | 4u << 6 | 6u << 9, 3u | 4u << 3 | 6u << 6 | 5u << 9,
                                    3u | 6u << 3 | 5u << 6 | 2u << 9, 2u | 0u << 3 | 3u << 6 | 5u << 9,
                                    0u | 1u << 3 | 3u << 6 | 2u << 9, 1u | 4u << 3 | 3u << 6 | 0u << 9};
    constexpr uint32_t HEXTAB4_SGP = 0u | 3u << 3 | 4u << 6 | 2u << 9;

    // one cell's row (if inner) or Qprev_b pass-through (halo), and its faces
    auto form = [&](int e) {
        const int se = e % R;
        auto Q = [&](int x) { return r_q[se + (x - e)]; };  // x within QP of e
        const int i = e % iall, j = e / iall;
        const bool inner = i >= 1 && i <= iall - 2 && j >= 1 && j <= iall - 2;
        const bool own = i <= iall - 2 && j <= iall - 2;  // origin of three arcs
        // ---- every global load of the cell before its first store: a load issued after a store
        // to memory it may alias is ordered behind it
        const double b0 = ROWS ? A.Qprevb[fb + e] : 0.0, b1 = ROWS ? A.Qprevb[plane + fb + e] : 0.0;
        uint32_t bits = 0;
        double cmo[3] = {0.0, 0.0, 0.0}, ob[3] = {0.0, 0.0, 0.0};  // own arcs' cmask, face cotangents
        if (own) {
            if constexpr (FACES) bits = A.mask[fb + e];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int nb = e + (m == 0 ? 1 : (m == 1 ? iall + 1 : iall));
                cmo[m] = A.cmask[(size_t)m * plane + fb + e];
                if constexpr (FACES)
                    ob[m] = A.qab[(size_t)m * plane + fb + e] + A.qab[(size_t)(m + 3) * plane + fb + nb];
            }
        }
        double cmw[3] = {0.0, 0.0, 0.0}, chv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, dc = 0.0;
        if (ROWS && inner) {
#pragma unroll
            for (int m = 0; m < 3; ++m)  // the arcs of (i-1,j), (i-1,j-1), (i,j-1)
                cmw[m] = A.cmask[(size_t)m * plane + fb + e - (m == 0 ? 1 : (m == 1 ? iall + 1 : iall))];
#pragma unroll
            for (int f = 0; f < 6; ++f) chv[f] = A.ch[(size_t)f * plane + fb + e];
            dc = A.d[fb + e];
        }
        double fout[3] = {0.0, 0.0, 0.0}, fq0[3] = {0.0, 0.0, 0.0}, fq1[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if ((bits >> m) & 1u) {  // a changed face: the forward's output and the bounds it may equal
                const int nb = e + (m == 0 ? 1 : (m == 1 ? iall + 1 : iall));
                const int T = cmo[m] > 0.0 ? nb : e;
                fout[m] = A.q_a[(size_t)m * plane + fb + e];
                fq0[m] = A.Qpost[fb + T];
                fq1[m] = A.Qpost[plane + fb + T];
            }
        }

        // min, max of the four cells of arc f (in the forward's chain order) and the slot of each,
        // the first in that order on a tie
        auto arc = [&](int f, double &lo, double &hi, uint32_t &sl, uint32_t &sh) {
            const int m = f < 3 ? f : f - 3;
            const int owner = f < 3 ? e : e - (m == 0 ? 1 : (m == 1 ? iall + 1 : iall));
            int c4[4];
            ha_arc(m, owner, iall, sgp, c4);
            lo = Q(c4[0]);
            hi = lo;
            sl = 0;
            sh = 0;
#pragma unroll
            for (int t = 1; t < 4; ++t) {
                const double v = Q(c4[t]);
                if (v < lo) { lo = v; sl = t; }
                if (v > hi) { hi = v; sh = t; }
            }
        };
        double olo[3], ohi[3];  // the own arcs'
        uint32_t osl[3], osh[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            olo[m] = 0.0; ohi[m] = 0.0; osl[m] = 0; osh[m] = 0;
            if (own) arc(m, olo[m], ohi[m], osl[m], osh[m]);
        }

        // ---- faces 1-3 of the own arcs (HLIM:377-410): attribution
        uint32_t fc = 0;
        if (FACES && own) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int nb = e + (m == 0 ? 1 : (m == 1 ? iall + 1 : iall));
                const int T = cmo[m] > 0.0 ? nb : e;
                uint32_t kind = 0, slot = 0;
                if ((bits >> m) & 1u) {
                    if (fout[m] == fq0[m]) kind = 1;
                    else if (fout[m] == fq1[m]) kind = 2;
                    else if (fout[m] == olo[m]) { kind = 3; slot = osl[m]; }
                    else { kind = 4; slot = osh[m]; }
                }
                A.qab[(size_t)(m + 3) * plane + fb + nb] = 0.0;
                A.qab[(size_t)m * plane + fb + e] = kind == 0 ? ob[m] : 0.0;
                r_ob[m * R + se] = ob[m];
                fc |= (kind | (slot << 3) | ((uint32_t)(T == nb) << 5)) << (6 * m);
            }
        }
        r_fc[se] = fc;

        // ---- the row of Qout (HLIM:216-301), summed per hexagon cell into the cell's seven slots of
        // r_p; the slots belong to this work-item alone, and its LDS adds are done in program order
        double *rp = r_p + se;
        if constexpr (!ROWS) {
            return;  // no partial sums: the gather reads none
        }
        if (!inner) {
#pragma unroll
            for (int w = 0; w < 7; ++w) rp[w * R] = w == 3 ? (0.0 + b0) + b1 : 0.0;  // halo: Qout = q
            if constexpr (CHB) {
#pragma unroll
                for (int f = 0; f < 6; ++f) A.chb[(size_t)f * plane + fb + e] = 0.0;
            }
            return;
        }
#pragma unroll
        for (int w = 0; w < 7; ++w) rp[w * R] = 0.0;
        const bool sgp_row = sgp && e == iall + 1;  // its f4 is the singular-point arc
        // the six arcs' bounds (Qin); the running extremes qnext, in face order from q(e)
        const double qij = Q(e);
        auto ccf = [&](int f) { return f < 3 ? cmo[f] : 1.0 - cmw[f - 3]; };
        auto bounds = [&](int f, double &mn, double &mx, uint32_t &wl, uint32_t &wh) {
            double lo, hi;
            uint32_t sl, sh;
            if (f < 3) {
                lo = olo[f]; hi = ohi[f]; sl = osl[f]; sh = osh[f];
            } else {
                arc(f, lo, hi, sl, sh);
            }
            const uint32_t tab = (f == 4 && sgp_row) ? HEXTAB4_SGP : HEXTAB[f];
            wl = (tab >> (3 * sl)) & 7u;
            wh = (tab >> (3 * sh)) & 7u;
            const double c = f < 3 ? cmo[f] : cmw[f - 3];
            if (f < 3) {
                mn = c * lo + (1.0 - c) * BIG;
                mx = c * hi + (1.0 - c) * (-BIG);
            } else {
                mn = c * BIG + (1.0 - c) * lo;
                mx = c * (-BIG) + (1.0 - c) * hi;
            }
        };
        double chm[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) chm[f] = dmn(chv[f], 0.0);
        double nmin = qij, nmax = qij, CQmin = 0.0, CQmax = 0.0;
        int smin = -1, smax = -1;
        uint32_t wmn = 0, wmx = 0;  // 3 bits per face: the hexagon cell of each arc's min, max
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            double mn, mx;
            uint32_t wl, wh;
            bounds(f, mn, mx, wl, wh);
            if (mn < nmin) { nmin = mn; smin = f; }
            if (mx > nmax) { nmax = mx; smax = f; }
            wmn |= wl << (3 * f);
            wmx |= wh << (3 * f);
            if constexpr (CHB) {
                CQmin = f == 0 ? chm[0] * mn : CQmin + chm[f] * mn;
                CQmax = f == 0 ? chm[0] * mx : CQmax + chm[f] * mx;
            }
        }
        double Cin = chm[0], Cout = chv[0] - chm[0];
#pragma unroll
        for (int f = 1; f < 6; ++f) {
            Cin = Cin + chm[f];
            Cout = Cout + chv[f] - chm[f];
        }
        const double zerosw = 0.5 - sgn_half(fabs(Cout) - A.eps);
        const double E = 1.0 - Cin - Cout + dc;
        const double fac = (1.0 - zerosw) / (Cout + zerosw);
        if constexpr (CHB) {
            // Qout in ch: an inflow face (ch < 0) gives (qnext - Qin)/Cout, any other (qnext - Qout)/Cout
            const double Qmin = (qij - CQmax - nmax * E) / (Cout + zerosw) * (1.0 - zerosw) + qij * zerosw;
            const double Qmax = (qij - CQmin - nmin * E) / (Cout + zerosw) * (1.0 - zerosw) + qij * zerosw;
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                double mn, mx;
                uint32_t wl, wh;
                bounds(f, mn, mx, wl, wh);
                const double d0 = zerosw != 0.0 ? 0.0 : (chv[f] < 0.0 ? nmax - mx : nmax - Qmin) / Cout;
                const double d1 = zerosw != 0.0 ? 0.0 : (chv[f] < 0.0 ? nmin - mn : nmin - Qmax) / Cout;
                // a zero cotangent contributes nothing (0 * an overflowed bound must not give NaN)
                A.chb[(size_t)f * plane + fb + e] = (b0 != 0.0 ? b0 * d0 : 0.0) + (b1 != 0.0 ? b1 * d1 : 0.0);
            }
        }
        // the row's terms, each added to the slot of the hexagon cell it names
        const double c0 = fac + zerosw;
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
            const int sel = mm == 0 ? smax : smin;
            const double b = mm == 0 ? b0 : b1;
            const uint32_t wb = mm == 0 ? wmx : wmn;
            const uint32_t selw = sel < 0 ? 3u : (wb >> (3 * sel)) & 7u;
            double selc = 1.0;
#pragma unroll
            for (int f = 0; f < 6; ++f) selc = sel == f ? ccf(f) : selc;
            atomicAdd(&rp[3 * R], c0 * b);
            atomicAdd(&rp[selw * R], (-(fac * E) * selc) * b);
#pragma unroll
            for (int f = 0; f < 6; ++f) atomicAdd(&rp[((wb >> (3 * f)) & 7u) * R], (-(fac * chm[f]) * ccf(f)) * b);
        }
    };

    // the cotangents of cell c: rows, then faces, in the sweep's order
    auto gather = [&](int c) {
        const int sc = c % R;
        auto sl = [&](int x) {
            int s = sc + (x - c);
            s += s < 0 ? R : 0;
            s -= s >= R ? R : 0;
            return s;
        };
        const int hx[7] = {-iall - 1, -iall, -1, 0, 1, iall, iall + 1};
        // Sums start at +0.0 and so never hold -0.0: adding +0.0 in place of a term that does not
        // apply leaves them unchanged, which lets every term below be a select rather than a branch.
        double acc = 0.0;
        if constexpr (ROWS) {
#pragma unroll
            for (int h = 0; h < 7; ++h) {  // the rows of c's hexagon, in cell order: their partial sums for c
                const int g = c + hx[h];
                const double v = r_p[(6 - h) * R + sl(g)];
                acc = acc + ((g >= 0 && g < gall) ? v : 0.0);
            }
        }
        double pn = 0.0, px = 0.0;
        // faces of origin p; d: index of c - p (arc_slot's), known after unrolling
        auto faces_at = [&](int p, int d) {
            const int s = sl(p);
            const uint32_t fc = (p >= 0 && p < gall) ? r_fc[s] : 0u;  // no origin, no faces (code 0)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int ws0 = arc_slot(m, d);
                const int ws = (m == 1 && arc_slot_sgp(d) != ws0) ? ((sgp && p == 0) ? arc_slot_sgp(d) : ws0) : ws0;
                const bool t_own = d == 4, t_nb = d == (m == 0 ? 3 : (m == 1 ? 1 : 2));
                if ((m != 1 || arc_slot_sgp(d) < 0) && ws0 < 0 && !t_own && !t_nb) continue;
                const uint32_t code = (fc >> (6 * m)) & 63u, kind = code & 7u;
                const double ob = r_ob[m * R + s];
                const bool hit = kind >= 3 && ws >= 0 && ((code >> 3) & 3u) == (uint32_t)ws;
                acc = acc + (hit ? ob : 0.0);
                if (t_own || t_nb) {
                    const bool at = (code >> 5) ? t_nb : t_own;
                    pn = pn + ((kind == 1 && at) ? ob : 0.0);
                    px = px + ((kind == 2 && at) ? ob : 0.0);
                }
            }
        };
        auto faces = [&](int p) {
            if (p < 0 || p >= gall) return;
            const int s = sl(p);
            const uint32_t fc = r_fc[s];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const uint32_t code = (fc >> (6 * m)) & 63u, kind = code & 7u;
                if (kind == 0) continue;
                const double ob = r_ob[m * R + s];
                if (kind >= 3) {
                    int c4[4];
                    ha_arc(m, p, iall, sgp, c4);
                    const uint32_t sl4 = (code >> 3) & 3u;
                    const int cell = sl4 == 0 ? c4[0] : (sl4 == 1 ? c4[1] : (sl4 == 2 ? c4[2] : c4[3]));
                    if (cell == c) acc = acc + ob;
                } else {
                    const int T = (code >> 5) ? p + (m == 0 ? 1 : (m == 1 ? iall + 1 : iall)) : p;
                    if (T == c) {
                        pn = kind == 1 ? pn + ob : pn;
                        px = kind == 1 ? px : px + ob;
                    }
                }
            }
        };
        if (!FACES) {
            // no faces: q_b is the rows' alone, Qout_post_b (if asked for) zero
        } else if (c == 0) {
            // the clamps send arc cells of row 0's origins to cell 0
            for (int p = 0; p <= iall; ++p) faces(p);
        } else {
            const int po[7] = {-iall - 2, -iall - 1, -iall, -1, 0, 1, iall};
#pragma unroll
            for (int h = 0; h < 7; ++h) faces_at(c + po[h], h);
        }
        A.qb[fb + c] = acc;
        if (A.Qpostb) {
            A.Qpostb[fb + c] = pn;
            A.Qpostb[plane + fb + c] = px;
        }
    };

    // q of cell e into its slot, and into the slot's copy past the ring's other end
    auto qput = [&](int e, double v) {
        const int se = e % R;
        r_q[se] = v;
        if (se < QP) r_q[se + R] = v;
        if (se >= R - QP) r_q[se - R] = v;
    };
    // prologue: q of cells [0, 2*iall+2], rows and faces of cells [0, iall]
    for (int e = tid; e <= 2 * iall + 2 && e < gall; e += HSW) qput(e, A.q[fb + e]);
    __syncthreads();
    for (int e = tid; e <= iall && e < gall; e += HSW) form(e);
    __syncthreads();
    for (int b0 = 0; b0 < gall; b0 += HSW) {
        const int eq = b0 + 2 * iall + 3 + tid;
        if (eq < gall) qput(eq, A.q[fb + eq]);
        __syncthreads();
        const int ef = b0 + iall + 1 + tid;
        if (ef < gall) form(ef);
        __syncthreads();
        const int c = b0 + tid;
        if (c < gall) gather(c);
    }
}

// Poles: one work-item per (k, l) slab, sequential as the oracle (HLIM:307-363,
// 412-425); it owns the slab's pole arrays, so no atomics.
__global__ void __launch_bounds__(HPOLE) hlim_adjoint_pole(HaArgs A)
{
    const int t = blockIdx.x * HPOLE + threadIdx.x;
    const int kall = A.kall, npl = A.gall_pl;
    if (!A.have_pl || t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const size_t pplane = (size_t)kLallPl * kall * npl;
    const double *qs = A.q_pl + fb, *chs = A.ch_pl + fb, *cms = A.cmask_pl + fb;
    // the rows run when Qout_prev_b is given, the faces when q_a_b is
    double *xq = A.qb_pl + fb, *xa = A.qab_pl ? A.qab_pl + fb : nullptr, *xp = A.Qpostb_pl ? A.Qpostb_pl + fb : nullptr;
    for (int v = 0; v < npl; ++v) {
        xq[v] = 0.0;
        if (xp) {
            xp[v] = 0.0;
            xp[pplane + v] = 0.0;
        }
    }
    const double BIG = A.big;
    auto arc = [&](int v, int &il, int &ih, double &lo_v, double &hi_v) {
        int vp1 = v + 1, vm1 = v - 1;
        if (vp1 == npl) vp1 = 1;
        if (vm1 == 0) vm1 = npl - 1;
        const int c4[4] = {0, v, vm1, vp1};
        lo_v = qs[0];
        hi_v = lo_v;
        il = 0;
        ih = 0;
        for (int u = 1; u < 4; ++u) {
            const double x = qs[c4[u]];
            if (x < lo_v) { lo_v = x; il = c4[u]; }
            if (x > hi_v) { hi_v = x; ih = c4[u]; }
        }
    };
    const double qn = qs[0];
    int smin = -1, smax = -1;
    double nmin = qn, nmax = qn;
    for (int v = 1; v < npl; ++v) {
        int il, ih;
        double lv, hv;
        arc(v, il, ih, lv, hv);
        const double c = cms[v];
        const double mn1 = c * lv + (1.0 - c) * BIG, mx1 = c * hv + (1.0 - c) * (-BIG);
        if (mn1 < nmin) { nmin = mn1; smin = v; }
        if (mx1 > nmax) { nmax = mx1; smax = v; }
    }
    double Cin = 0.0, Cout = 0.0;
    for (int v = 1; v < npl; ++v) {
        const double chm = cms[v] * chs[v];
        Cin = Cin + chm;
        Cout = Cout - chm + chs[v];
    }
    const double zerosw = 0.5 - sgn_half(fabs(Cout) - A.eps);
    const double E = 1.0 - Cin - Cout + A.d_pl[fb];
    const double fac = (1.0 - zerosw) / (Cout + zerosw);
    if (A.chb_pl) {
        // Qout in ch(v): cmask*ch enters Cin and CQin, ch - cmask*ch Cout
        double CQmin = 0.0, CQmax = 0.0;
        for (int v = 1; v < npl; ++v) {
            int il, ih;
            double lv, hv;
            arc(v, il, ih, lv, hv);
            const double chm = cms[v] * chs[v], c = cms[v];
            CQmin = CQmin + chm * (c * lv + (1.0 - c) * BIG);
            CQmax = CQmax + chm * (c * hv + (1.0 - c) * (-BIG));
        }
        const double Qmin = (qn - CQmax - nmax * E) / (Cout + zerosw) * (1.0 - zerosw) + qn * zerosw;
        const double Qmax = (qn - CQmin - nmin * E) / (Cout + zerosw) * (1.0 - zerosw) + qn * zerosw;
        const double b0 = A.Qprevb_pl[fb], b1 = A.Qprevb_pl[pplane + fb];
        double *xc = A.chb_pl + fb;
        xc[0] = 0.0;
        for (int v = 1; v < npl; ++v) {
            int il, ih;
            double lv, hv;
            arc(v, il, ih, lv, hv);
            const double c = cms[v];
            const double imx = c * hv + (1.0 - c) * (-BIG), imn = c * lv + (1.0 - c) * BIG;
            const double d0 = zerosw != 0.0 ? 0.0 : ((nmax - c * imx) - (1.0 - c) * Qmin) / Cout;
            const double d1 = zerosw != 0.0 ? 0.0 : ((nmin - c * imn) - (1.0 - c) * Qmax) / Cout;
            xc[v] = (b0 != 0.0 ? b0 * d0 : 0.0) + (b1 != 0.0 ? b1 * d1 : 0.0);
        }
    }
    for (int mm = 0; mm < 2 && A.Qprevb_pl; ++mm) {
        const double bq = A.Qprevb_pl[(size_t)mm * pplane + fb];
        const int sel = mm == 0 ? smax : smin;
        xq[0] = xq[0] + (fac + zerosw) * bq;
        int selc = 0;
        if (sel >= 0) {
            int il, ih;
            double lv, hv;
            arc(sel, il, ih, lv, hv);
            selc = mm == 0 ? ih : il;
        }
        xq[selc] = xq[selc] + (-(fac * E) * (sel < 0 ? 1.0 : cms[sel])) * bq;
        for (int v = 1; v < npl; ++v) {
            int il, ih;
            double lv, hv;
            arc(v, il, ih, lv, hv);
            const int cx = mm == 0 ? ih : il;
            xq[cx] = xq[cx] + (-(fac * (cms[v] * chs[v])) * cms[v]) * bq;
        }
    }
    for (int v = 1; v < npl && xa; ++v) {
        const double c = cms[v];
        const int T = c > 0.0 ? v : 0;
        int kind = 0, cell = 0;
        if (A.mask_pl[fb + v] & 1u) {
            const double out = A.q_a_pl[fb + v];
            int il, ih;
            double lv, hv;
            arc(v, il, ih, lv, hv);
            if (out == A.Qpost_pl[fb + T]) kind = 1;
            else if (out == A.Qpost_pl[pplane + fb + T]) kind = 2;
            else if (out == lv) { kind = 3; cell = il; }
            else { kind = 4; cell = ih; }
        }
        const double ob = xa[v];
        xa[v] = kind == 0 ? ob : 0.0;
        if (kind == 1) xp[T] = xp[T] + ob;
        else if (kind == 2) xp[pplane + T] = xp[pplane + T] + ob;
        else if (kind >= 3) xq[cell] = xq[cell] + ob;
    }
}

// ---- horizontal_flux (dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379) -------
//
// Linear in rho*v; GRD_xc carries 1/max(rho_arc, EPS).  The arc cotangents
// are gathered, not scattered: one work-group per (k, l) slab sweeps it in
// steps of FSW cells.  Each step computes the cotangents of the next FSW
// arc origins (each origin's three arcs, its four triangles formed once)
// into an LDS ring of R = FSW + 2*iall + 1 origins, then each work-item sums
// its cell's six triangles in the order of a sequential sweep over triangle
// origins, so the result is the oracle's, bit for bit, with no atomics and
// no arc computed twice.  The tracer fields are read from HBM directly
// (neighbouring arcs hit the caches), which keeps LDS to the arc cotangents.
// (Earlier forms: chunks of 1024 cells with their halo of arcs recomputed,
// one work-group per CU for its 96 KiB of LDS, 1.25 ms at gl7; with the
// tracer fields staged too, 256-cell chunks, 3.05 ms.)
constexpr int FSW = 256;

struct FaArgs {
    const double *flxb, *xcb;   // (gall,kall,lall,6), (gall,kall,lall,3,3)
    double *Xb[4];              // rho_b, rhovx_b, rhovy_b, rhovz_b
    const double *X[4];
    const double *gt, *ga, *gp;
    const int *have_sgp;
    int iall, gall, kall, lall, ring;
    double dt, eps;
    // poles
    const double *flxb_pl, *xcb_pl;
    double *Xb_pl[4];
    const double *X_pl[4];
    const double *gt_pl, *ga_pl, *gp_pl;
    int gall_pl, have_pl;
};

// The cotangents of one origin's arcs that HBM supplies: flx_h_b of its own
// face and of the neighbour's mirrored face, and GRD_xc_b (loaded a step
// ahead of their use).
struct ArcIn {
    double fa[3], fn[3], xb[3][3];
};

// R: the ring size, fixed at compile time so that a ring slot is one wrapped
// index per neighbour and every (arc, field) plane is an immediate offset from
// it.  R >= FSW + 2 * iall + 1.
template <int R>
__global__ void __launch_bounds__(FSW) __attribute__((amdgpu_waves_per_eu(2, 8))) hflux_adjoint_kernel(FaArgs A)
{
    extern __shared__ double lds[];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int k = blockIdx.x % kall, l = blockIdx.x / kall;
    if (l >= lall) return;
    double *s_a = lds;           // [3 arcs][4 fields][R] arc cotangents of origins [b0 - iall - 1, b0 + FSW + iall - 1]
    double *s_x = lds + 12 * R;  // [4 fields][R] rho, rho*v of cells [b0, b0 + FSW + 2 iall + 1)
    const size_t fb = ((size_t)l * kall + k) * gall;
    const size_t plane = (size_t)lall * kall * gall;
    const size_t mstr = (size_t)lall * gall, mb = (size_t)l * gall;
    const bool sgp = A.have_sgp[l] != 0;
    const int tid = threadIdx.x;
    auto W = [&](int c, int t, int w) { return A.gt[(size_t)(w * 2 + t) * mstr + mb + c]; };  // w: 2,3,4 = W1..W3
    const double dt = A.dt, EPS = A.eps;
    auto up = [](int q) { return q >= R ? q - R : q; };
    auto dn = [](int q) { return q < 0 ? q + R : q; };
    auto AB = [&](int a, int f, int q) { return s_a[((a - 1) * 4 + f) * R + q]; };
    auto valid = [&](int a, int i, int j) { return j >= (a == 1 ? 1 : 0) && i >= (a == 3 ? 1 : 0); };
    // is (i, j) an arc origin (i, j in [0, iall-2])
    auto is_org = [&](int i, int j) { return i <= iall - 2 && j <= iall - 2; };
    // loads of region-relative cells clamped into the region: lanes past its end or off its
    // origins read a cell of the region and discard what they read
    auto cl = [&](int c) { return c < 0 ? 0 : (c > gall - 1 ? gall - 1 : c); };
    auto arcin = [&](int a0, ArcIn &in) {
#pragma unroll
        for (int a = 1; a <= 3; ++a) {
            const int g = cl(a0), nb = cl(a0 + (a == 1 ? 1 : (a == 2 ? iall + 1 : iall)));
            in.fa[a - 1] = A.flxb[(size_t)(a - 1) * plane + fb + g];
            in.fn[a - 1] = A.flxb[(size_t)(a + 2) * plane + fb + nb];
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) in.xb[a - 1][dd] = A.xcb[(size_t)(dd * 3 + (a - 1)) * plane + fb + g];
        }
    };
    // the three arcs of origin a0 (slot s, position (i, j); zero where a0 is no origin).  GEN: a0 may
    // be 0 or 1 (the singular-point substitution TI(0) = TJ(1)) or lie in row 0 (no TJ(a0-iall));
    // otherwise a0 >= iall.
    auto arcs = [&](auto gen_tag, int a0, int s, int i, int j, const ArcIn &in) {
        constexpr bool GEN = decltype(gen_tag)::value;
        const bool org = is_org(i, j);
        double tI[4], tJ[4], tJs[4], tIw[4];
        if (!GEN || (org && !(sgp && a0 <= 1))) {
            // the four triangles the origin's arcs use: TI(a0), TJ(a0), TJ(a0-iall) (AI), TI(a0-1) (AJ).
            // They touch six cells (a0, its E, N, NE, S, W neighbours): each field is read there once,
            // and the sums are formed in the forward's order (FLX:139-201).
            const bool hs = !GEN || j >= 1, hw = i >= 1;
            const int c0 = cl(a0), cs = cl(a0 - iall), cw = cl(a0 - 1);
            const double wi1 = W(c0, 0, 2), wi2 = W(c0, 0, 3), wi3 = W(c0, 0, 4);
            const double wj1 = W(c0, 1, 2), wj2 = W(c0, 1, 3), wj3 = W(c0, 1, 4);
            const double ws1 = W(cs, 1, 2), ws2 = W(cs, 1, 3), ws3 = W(cs, 1, 4);
            const double ww1 = W(cw, 0, 2), ww2 = W(cw, 0, 3), ww3 = W(cw, 0, 4);
            const int qe = up(s + 1), qn = up(s + iall), qne = up(s + iall + 1), qs = dn(s - iall), qw = dn(s - 1);
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const double *x = s_x + f * R;
                const double x0 = x[s], xe = x[qe], xn = x[qn], xne = x[qne], xs = x[qs], xw = x[qw];
                tI[f] = x0 * wi1 + xe * wi2 + xne * wi3;
                tJ[f] = x0 * wj1 + xne * wj2 + xn * wj3;
                tJs[f] = hs ? xs * ws1 + xe * ws2 + x0 * ws3 : 0.0;
                tIw[f] = hw ? xw * ww1 + x0 * ww2 + xn * ww3 : 0.0;
            }
        } else {
            auto X = [&](int f, int p) { return s_x[f * R + dn(up(s + (p - a0)))]; };
            auto tri = [&](int f, int t, int c) {
                if (t == 0 && sgp && c == 0) t = 1, c = 1;
                if (t == 0) return X(f, c) * W(c, 0, 2) + X(f, c + 1) * W(c, 0, 3) + X(f, c + iall + 1) * W(c, 0, 4);
                return X(f, c) * W(c, 1, 2) + X(f, c + iall + 1) * W(c, 1, 3) + X(f, c + iall) * W(c, 1, 4);
            };
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                tI[f] = org ? tri(f, 0, a0) : 0.0;
                tJ[f] = org ? tri(f, 1, a0) : 0.0;
                tJs[f] = (org && j >= 1) ? tri(f, 1, a0 - iall) : 0.0;
                tIw[f] = (org && i >= 1) ? tri(f, 0, a0 - 1) : 0.0;
            }
        }
        const double ra = A.gp[1 * mstr + mb + cl(a0)] * dt;
#pragma unroll
        for (int a = 1; a <= 3; ++a) {
            // every lane forms the arc; those off the origins store zeros
            const bool ok = org && valid(a, i, j);
            const int g = cl(a0);
            const int nb = a0 + (a == 1 ? 1 : (a == 2 ? iall + 1 : iall));
            double s2[4];
#pragma unroll
            for (int f = 0; f < 4; ++f)  // FLX:234-309: the arc's two triangles, in the forward's order
                s2[f] = a == 1 ? tJs[f] + tI[f] : (a == 2 ? tI[f] + tJ[f] : tJ[f] + tIw[f]);
            const double rr = 1.0 / dmx(s2[0], EPS);
            const bool live_rho = !(EPS > s2[0]);
            const double rn = A.gp[1 * mstr + mb + cl(nb)] * dt;
            const bool sgp_face = sgp && a == 3 && nb == iall + 1;
            const double fb_nb = sgp_face ? 0.0 : in.fn[a - 1];
            const double fluxb = in.fa[a - 1] * ra - fb_nb * rn;
            double bv[4], r2b = 0.0;
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) {
                const double hn = A.ga[(size_t)(dd * 3 + (a - 1)) * mstr + mb + g];
                const double xb = in.xb[a - 1][dd];
                bv[1 + dd] = 0.5 * hn * fluxb - xb * rr * dt * 0.5;
                r2b = r2b + xb * s2[1 + dd] * dt * 0.5 * rr * rr;
            }
            bv[0] = live_rho ? r2b : 0.0;
#pragma unroll
            for (int f = 0; f < 4; ++f) s_a[((a - 1) * 4 + f) * R + s] = ok ? bv[f] : 0.0;
        }
    };
    // triangle origins of the cell at (i0, j0): its six triangles, and whether each is one
    auto origins = [&](int i0, int j0, bool &o_sw, bool &o_s, bool &o_w, bool &o_c) {
        o_sw = i0 >= 1 && j0 >= 1 && i0 - 1 <= iall - 2 && j0 - 1 <= iall - 2;  // X0-iall-1
        o_s = j0 >= 1 && i0 <= iall - 2 && j0 - 1 <= iall - 2;                  // X0-iall
        o_w = i0 >= 1 && i0 - 1 <= iall - 2 && j0 <= iall - 2;                  // X0-1
        o_c = i0 <= iall - 2 && j0 <= iall - 2;                                 // X0
    };
    // the weights of the cell in its six triangles (loaded a step ahead of their use)
    // (a weight is read where its triangle is no origin too, and then not used)
    auto cellw = [&](int X0, double w[6]) {
        const int c_sw = cl(X0 - iall - 1), c_s = cl(X0 - iall), c_w = cl(X0 - 1), c_c = cl(X0);
        w[0] = W(c_sw, 0, 4);
        w[1] = W(c_sw, 1, 3);
        w[2] = W(c_s, 1, 4);
        w[3] = W(c_w, 0, 3);
        w[4] = W(c_c, 0, 2);
        w[5] = W(c_c, 1, 2);
    };
    // the cell's four cotangents from its six triangles.  Triangle cotangents: TI(c) from AI(c),
    // AIJ(c), AJ(c+1); TJ(c) from AI(c+iall), AIJ(c), AJ(c); the six triangles use the twelve arcs
    // of origins c_sw, c_s, c_w, X0, X0+1, X0+iall.
    auto cell = [&](int X0, int sc, int i0, int j0, const double w[6]) {
        bool o_sw, o_s, o_w, o_c;
        origins(i0, j0, o_sw, o_s, o_w, o_c);
        const double w_sw_i = w[0], w_sw_j = w[1], w_s_j = w[2], w_w_i = w[3], w_c_i = w[4], w_c_j = w[5];
        if (sgp && (X0 <= 1 || X0 == iall + 1 || X0 == iall + 2)) {
            // the cells whose triangles include TI(0) or TJ(1): the singular-point substitution moves
            // TI(0)'s cotangent onto TJ(1)
            const int c_sw = X0 - iall - 1, c_s = X0 - iall, c_w = X0 - 1;
            auto ABg = [&](int a, int f, int c) { return AB(a, f, dn(up(sc + (c - X0)))); };
            auto tib0 = [&](int f, int c) { return ((0.0 + ABg(1, f, c)) + ABg(2, f, c)) + ABg(3, f, c + 1); };
            auto tjb0 = [&](int f, int c) { return ((0.0 + ABg(1, f, c + iall)) + ABg(2, f, c)) + ABg(3, f, c); };
            auto tib = [&](int f, int c) { return (c == 0) ? 0.0 : tib0(f, c); };
            auto tjb = [&](int f, int c) { return (c == 1) ? tjb0(f, 1) + tib0(f, 0) : tjb0(f, c); };
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                double y = 0.0;
                if (o_sw) y = y + (tib(f, c_sw) * w_sw_i + tjb(f, c_sw) * w_sw_j);
                if (o_s) y = y + tjb(f, c_s) * w_s_j;
                if (o_w) y = y + tib(f, c_w) * w_w_i;
                if (o_c) y = y + (tib(f, X0) * w_c_i + tjb(f, X0) * w_c_j);
                A.Xb[f][fb + X0] = y;
            }
            return;
        }
        const int q_c = sc, q_e = up(sc + 1), q_n = up(sc + iall);
        const int q_w = dn(sc - 1), q_s = dn(sc - iall), q_sw = dn(sc - iall - 1);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const double a1sw = AB(1, f, q_sw), a2sw = AB(2, f, q_sw), a3sw = AB(3, f, q_sw);
            const double a2s = AB(2, f, q_s), a3s = AB(3, f, q_s);
            const double a1w = AB(1, f, q_w), a2w = AB(2, f, q_w);
            const double a1c = AB(1, f, q_c), a2c = AB(2, f, q_c), a3c = AB(3, f, q_c);
            const double a3e = AB(3, f, q_e), a1n = AB(1, f, q_n);
            // a triangle that is no origin leaves the sum as it is (a select, not an added zero, which
            // would turn a -0.0 into +0.0); the terms are formed by every lane, from whatever the ring
            // and the clamped weights hold there
            double y = 0.0;
            const double t_sw = (((0.0 + a1sw) + a2sw) + a3s) * w_sw_i + (((0.0 + a1w) + a2sw) + a3sw) * w_sw_j;
            y = o_sw ? y + t_sw : y;
            const double t_s = (((0.0 + a1c) + a2s) + a3s) * w_s_j;
            y = o_s ? y + t_s : y;
            const double t_w = (((0.0 + a1w) + a2w) + a3c) * w_w_i;
            y = o_w ? y + t_w : y;
            const double t_c = (((0.0 + a1c) + a2c) + a3e) * w_c_i + (((0.0 + a1n) + a2c) + a3c) * w_c_j;
            y = o_c ? y + t_c : y;
            A.Xb[f][fb + X0] = y;
        }
    };
    // rho, rho*v of cell p into its slot of the ring (zero beyond the region)
    auto xload = [&](int p, double v[4]) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const double t = A.X[f][fb + cl(p)];
            v[f] = p < gall ? t : 0.0;
        }
    };
    auto xstore = [&](int sp, const double v[4]) {
#pragma unroll
        for (int f = 0; f < 4; ++f) s_x[f * R + sp] = v[f];
    };
    for (int p = tid; p < R; p += FSW) {
        double v[4];
        xload(p, v);
        xstore(p, v);
    }
    __syncthreads();
    for (int e = tid; e < iall; e += FSW) {  // origins [0, iall): row j = 0
        ArcIn in;
        arcin(e, in);
        arcs(std::true_type{}, e, e, e, 0, in);
    }
    // origins [b0 - iall - 1, b0 + FSW + iall - 1] are in the arc ring while cells [b0, b0 + FSW) gather;
    // each step replaces the cells [b0, b0 + FSW) of the cell ring, read by its arcs only, with the next
    // step's, loaded while the arcs are formed.  Positions advance by FSW cells a step: di columns and
    // dj rows.
    const int di = FSW % iall, dj = FSW / iall;
    auto adv = [&](int &i, int &j) {
        i += di;
        j += dj;
        if (i >= iall) i -= iall, j += 1;
    };
    int ia = tid % iall, ja = tid / iall + 1;  // origin iall + tid
    int ic = tid % iall, jc = tid / iall;      // cell tid
    int sa = up(iall + tid), sc = tid;         // their slots
    ArcIn cur;  // the arc cotangents of this step's origins, loaded during the previous step
    arcin(iall + tid, cur);
    for (int b0 = 0; b0 < gall; b0 += FSW) {
        const int a0 = b0 + iall + tid;
        double xn[4], w[6];
        xload(b0 + R + tid, xn);
        cellw(b0 + tid, w);
        arcs(std::false_type{}, a0, sa, ia, ja, cur);
        adv(ia, ja);
        arcin(a0 + FSW, cur);
        __syncthreads();
        if (b0 + tid < gall) cell(b0 + tid, sc, ic, jc, w);
        xstore(sc, xn);  // cell b0 + R + tid has cell b0 + tid's slot
        __syncthreads();
        adv(ic, jc);
        sa = up(sa + FSW);
        sc = up(sc + FSW);
    }
}

// Poles: one work-item per (k, l), in the oracle's order (FLX:326-374).
__global__ void hflux_adjoint_pole(FaArgs A)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int kall = A.kall, npl = A.gall_pl;
    if (!A.have_pl || t >= kall * kLallPl || npl > 16) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const size_t pplane = (size_t)kLallPl * kall * npl, mstr = (size_t)kLallPl * npl, mb = (size_t)l * npl;
    const double dt = A.dt, EPS = A.eps;
    auto PT = [&](int v, int w) { return A.gt_pl[(size_t)w * mstr + mb + v]; };  // w: 2,3,4 = W1..W3
    const double ran = A.gp_pl[1 * mstr + mb + 0] * dt;
    double Tv[4][16], Tb[4][16];
    for (int v = 1; v < npl; ++v) {
        const int vp1 = (v + 1 == npl) ? 1 : v + 1;
        for (int f = 0; f < 4; ++f) {
            const double *x = A.X_pl[f] + fb;
            Tv[f][v] = x[0] * PT(v, 2) + x[v] * PT(v, 3) + x[vp1] * PT(v, 4);
            Tb[f][v] = 0.0;
        }
    }
    for (int v = 1; v < npl; ++v) {
        const int vm1 = (v == 1) ? npl - 1 : v - 1;
        double s2[4];
        for (int f = 0; f < 4; ++f) s2[f] = Tv[f][vm1] + Tv[f][v];
        const double rr = 1.0 / dmx(s2[0], EPS);
        const bool live_rho = !(EPS > s2[0]);
        const double fluxb = A.flxb_pl[fb + v] * ran;
        double bv[4], r2b = 0.0;
        for (int dd = 0; dd < 3; ++dd) {
            const double hn = A.ga_pl[(size_t)dd * mstr + mb + v];
            const double xb = A.xcb_pl[(size_t)dd * pplane + fb + v];
            bv[1 + dd] = 0.5 * hn * fluxb - xb * rr * dt * 0.5;
            r2b = r2b + xb * s2[1 + dd] * dt * 0.5 * rr * rr;
        }
        bv[0] = live_rho ? r2b : 0.0;
        for (int f = 0; f < 4; ++f) {
            Tb[f][vm1] += bv[f];
            Tb[f][v] += bv[f];
        }
    }
    for (int f = 0; f < 4; ++f) {
        double *y = A.Xb_pl[f] + fb;
        for (int v = 0; v < npl; ++v) y[v] = 0.0;
        for (int v = 1; v < npl; ++v) {
            const int vp1 = (v + 1 == npl) ? 1 : v + 1;
            y[0] += Tb[f][v] * PT(v, 2);
            y[v] += Tb[f][v] * PT(v, 3);
            y[vp1] += Tb[f][v] * PT(v, 4);
        }
    }
}

}  // namespace

int launch_hflux_adjoint(nicam_ctx c, const double *flxb, const double *flxb_pl, const double *xcb,
                         const double *xcb_pl, double *const Xb[4], double *const Xb_pl[4], const double *const X[4],
                         const double *const X_pl[4], double dt)
{
    const Dims &D = c->d;
    if (D.gall_1d > 160) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (160)");
    if (D.have_pl && D.gall_pl > 16) return fail(NICAM_ERR_ARG, "gall_pl larger than 16");
    FaArgs A;
    A.flxb = flxb; A.xcb = xcb; A.flxb_pl = flxb_pl; A.xcb_pl = xcb_pl;
    for (int f = 0; f < 4; ++f) {
        A.Xb[f] = Xb[f]; A.X[f] = X[f]; A.Xb_pl[f] = Xb_pl[f]; A.X_pl[f] = X_pl[f];
    }
    A.gt = c->GMTR_t; A.ga = c->GMTR_a; A.gp = c->GMTR_p;
    A.gt_pl = c->GMTR_t_pl; A.ga_pl = c->GMTR_a_pl; A.gp_pl = c->GMTR_p_pl;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d; A.gall = D.gall; A.kall = D.kall; A.lall = D.lall;
    A.dt = dt; A.eps = c->eps;
    A.gall_pl = D.gall_pl; A.have_pl = D.have_pl;
    const long nblk = (long)D.kall * D.lall;
    if (nblk > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    // the ring covers the origins [b0 - iall - 1, b0 + FSW + iall - 1] a step reads: the smallest
    // compiled size of at least FSW + 2 iall + 1
    auto launch = [&](auto kern, int R) -> int {
        A.ring = R;
        const size_t lds = sizeof(double) * 16 * (size_t)R;
        if (lds > 64 * 1024)
            NICAM_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(FSW), lds, c->stream, A);
        return NICAM_OK;
    };
    const int need = FSW + 2 * D.gall_1d + 1;
    int rc;
    if (need <= FSW + 2 * 34 + 1) rc = launch(hflux_adjoint_kernel<FSW + 2 * 34 + 1>, FSW + 2 * 34 + 1);
    else if (need <= FSW + 2 * 66 + 1) rc = launch(hflux_adjoint_kernel<FSW + 2 * 66 + 1>, FSW + 2 * 66 + 1);
    else if (need <= FSW + 2 * 130 + 1) rc = launch(hflux_adjoint_kernel<FSW + 2 * 130 + 1>, FSW + 2 * 130 + 1);
    else rc = launch(hflux_adjoint_kernel<FSW + 2 * 160 + 1>, FSW + 2 * 160 + 1);
    if (rc != NICAM_OK) return rc;
    if (D.have_pl)
        hipLaunchKernelGGL(hflux_adjoint_pole, dim3((unsigned)((D.kall * kLallPl + 63) / 64)), dim3(64), 0, c->stream,
                           A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_hlim_adjoint(nicam_ctx c, double *qab, double *qab_pl, double *qb, double *qb_pl, const double *Qprevb,
                        const double *Qprevb_pl, double *Qpostb, double *Qpostb_pl, const double *q_a,
                        const double *q_a_pl, const double *q, const double *q_pl, const double *d,
                        const double *d_pl, const double *ch, const double *ch_pl, const double *cmask,
                        const double *cmask_pl, const double *Qpost, const double *Qpost_pl, const uint8_t *mask,
                        const uint8_t *mask_pl, double *chb, double *chb_pl)
{
    const Dims &D = c->d;
    if (D.gall_1d > HMAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (160)");
    HaArgs A;
    A.qab = qab; A.qb = qb; A.Qpostb = Qpostb; A.Qprevb = Qprevb; A.chb = chb; A.chb_pl = chb_pl;
    A.q_a = q_a; A.q = q; A.d = d; A.ch = ch; A.cmask = cmask; A.Qpost = Qpost; A.mask = mask;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d; A.gall = D.gall; A.kall = D.kall; A.lall = D.lall;
    A.eps = c->eps; A.big = c->big;
    A.qab_pl = qab_pl; A.qb_pl = qb_pl; A.Qpostb_pl = Qpostb_pl; A.Qprevb_pl = Qprevb_pl;
    A.q_a_pl = q_a_pl; A.q_pl = q_pl; A.d_pl = d_pl; A.ch_pl = ch_pl; A.cmask_pl = cmask_pl;
    A.Qpost_pl = Qpost_pl; A.mask_pl = mask_pl;
    A.gall_pl = D.gall_pl; A.have_pl = D.have_pl;
    // the ring covers the cells [b0 - iall - 2, b0 + HSW + iall] a step reads
    A.ring = HSW + 2 * D.gall_1d + 3;
    const long nslab = (long)D.kall * D.lall;
    if (nslab > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    const size_t lds = (size_t)(A.ring + 2 * (D.gall_1d + 2)) * sizeof(double) +
                       (size_t)A.ring * (10 * sizeof(double) + sizeof(uint32_t));
    // the rows run when Qout_prev_b is given (ch_b comes with them), the faces when q_a_b is
    const int passes = (Qprevb ? 1 : 0) | (qab ? 2 : 0);
    const bool wch = chb != nullptr;
    const void *kern = nullptr;
    if (passes == 3) kern = wch ? (const void *)hlim_adjoint_kernel<true, 3> : (const void *)hlim_adjoint_kernel<false, 3>;
    else if (passes == 1) kern = wch ? (const void *)hlim_adjoint_kernel<true, 1> : (const void *)hlim_adjoint_kernel<false, 1>;
    else kern = (const void *)hlim_adjoint_kernel<false, 2>;
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (passes == 3 && wch)
        hipLaunchKernelGGL((hlim_adjoint_kernel<true, 3>), dim3((unsigned)nslab), dim3(HSW), lds, c->stream, A);
    else if (passes == 3)
        hipLaunchKernelGGL((hlim_adjoint_kernel<false, 3>), dim3((unsigned)nslab), dim3(HSW), lds, c->stream, A);
    else if (passes == 1 && wch)
        hipLaunchKernelGGL((hlim_adjoint_kernel<true, 1>), dim3((unsigned)nslab), dim3(HSW), lds, c->stream, A);
    else if (passes == 1)
        hipLaunchKernelGGL((hlim_adjoint_kernel<false, 1>), dim3((unsigned)nslab), dim3(HSW), lds, c->stream, A);
    else
        hipLaunchKernelGGL((hlim_adjoint_kernel<false, 2>), dim3((unsigned)nslab), dim3(HSW), lds, c->stream, A);
    if (D.have_pl)
        hipLaunchKernelGGL(hlim_adjoint_pole, dim3((unsigned)((D.kall * kLallPl + HPOLE - 1) / HPOLE)), dim3(HPOLE),
                           0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_vlim_adjoint(nicam_ctx c, double *qhb, double *qhb_pl, double *qb, double *qb_pl, const double *q_h,
                        const double *q_h_pl, const double *q, const double *q_pl, const double *d,
                        const double *d_pl, const double *ck, const double *ck_pl, const uint32_t *mask,
                        const uint32_t *mask_pl, double *ckb, double *ckb_pl)
{
    const Dims &D = c->d;
    if (D.kall - 3 <= 0) return NICAM_OK;
    VlaArgs A;
    A.reg = VlaField{qhb, qb, ckb, q_h, q, d, ck, mask};
    A.pl = VlaField{qhb_pl, qb_pl, ckb_pl, q_h_pl, q_pl, d_pl, ck_pl, mask_pl};
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.nword = (D.kall - 3 + 31) / 32;
    A.eps = c->eps;
    A.big = c->big;
    const size_t lds = sizeof(double) * VCOL * D.kall;
    if (lds > 160 * 1024) return fail(NICAM_ERR_ARG, "kall too large for the staged cotangents");
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute((const void *)vlim_adjoint_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    const long nblk = ((long)D.gall * D.lall + VCOL - 1) / VCOL + (D.have_pl ? ((long)D.gall_pl * kLallPl + VCOL - 1) / VCOL : 0);
    hipLaunchKernelGGL(vlim_adjoint_kernel, dim3((unsigned)nblk), dim3(BLOCK_VLA), lds,
                       c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_vi_adjoint(nicam_ctx c, double *wb, double *wb_pl, double *w0b, double *w0b_pl, double *p0b,
                      double *p0b_pl, double *r0b, double *r0b_pl, double *Srb, double *Srb_pl, double *Swb,
                      double *Swb_pl, double *Spb, double *Spb_pl, double dt)
{
    const Dims &D = c->d;
    VaArgs A;
    A.reg = VaField{wb, w0b, p0b, r0b, Srb, Swb, Spb, c->Mc, c->Ml, c->Mu, c->RGAM, c->RGAMH, c->RGSGAM2,
                    c->RGSGAM2H, c->GSGAM2H, c->vi_tbeta, c->vi_tgam};
    A.pl = VaField{wb_pl, w0b_pl, p0b_pl, r0b_pl, Srb_pl, Swb_pl, Spb_pl, c->Mc_pl, c->Ml_pl, c->Mu_pl,
                   c->RGAM_pl, c->RGAMH_pl, c->RGSGAM2_pl, c->RGSGAM2H_pl, c->GSGAM2H_pl, c->vi_tbeta_pl,
                   c->vi_tgam_pl};
    A.rdgzh = c->rdgzh;
    A.afact = c->afact;
    A.bfact = c->bfact;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.dt = dt;
    A.CVovRt2 = c->cvdry / c->rdry / (dt * dt);
    A.alpha = (double)c->alpha;
    A.grav = c->grav;
    A.ncol = (long)D.gall * D.lall + (D.have_pl ? (long)D.gall_pl * kLallPl : 0);
    // kall = 42: the column in registers; at kall = 98 the two arrays spill (1.2 KB of scratch per lane), so
    // the partial solution stays in the outputs and the factors come from the stored factorisation
    // (NICAM_TUNE_VI_VARIANT 1 or 3: formed in the solve, as before)
    const int variant = c->tune.vi_variant;
    if (D.kall == 42 && variant != 2 && variant != 4) {
        hipLaunchKernelGGL(vi_adjoint_kernel_reg<42>, dim3((unsigned)column_blocks<BLOCK_VAR>(D.gall, D.lall, D.gall_pl, D.have_pl)),
                           dim3(BLOCK_VAR), 0, c->stream, A);
    } else if (variant == 1 || variant == 3) {
        hipLaunchKernelGGL(vi_adjoint_kernel<false>, dim3((unsigned)column_blocks<BLOCK_VA>(D.gall, D.lall, D.gall_pl, D.have_pl)),
                           dim3(BLOCK_VA), 0, c->stream, A);
    } else {
        hipLaunchKernelGGL(vi_adjoint_kernel_chunked, dim3((unsigned)column_blocks<BLOCK_VA>(D.gall, D.lall, D.gall_pl, D.have_pl)),
                           dim3(BLOCK_VA), 0, c->stream, A);
    }
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
