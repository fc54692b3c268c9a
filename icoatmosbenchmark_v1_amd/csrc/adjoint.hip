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
// final cotangents, so no work array is needed.
struct VaField {
    double *wb, *w0b, *p0b, *r0b, *Srb, *Swb, *Spb;
    const double *Mc, *Ml, *Mu, *RGAM, *RGAMH, *RGSGAM2, *RGSGAM2H, *GSGAM2H;
};

struct VaArgs {
    VaField reg, pl;
    const double *rdgzh, *afact, *bfact;
    long ncol;
    int gall, kall, lall, gall_pl, have_pl;
    double dt, CVovRt2, alpha, grav;
};

constexpr int BLOCK_VA = 256;

__global__ void __launch_bounds__(BLOCK_VA) vi_adjoint_kernel(VaArgs A)
{
    const long t = (long)blockIdx.x * BLOCK_VA + threadIdx.x;
    const long nreg = (long)A.gall * A.lall;
    const VaField *F;
    size_t base, s;
    if (t < nreg) {
        F = &A.reg;
        base = (size_t)(t / A.gall) * A.kall * A.gall + (size_t)(t % A.gall);
        s = A.gall;
    } else if (A.have_pl && t < nreg + (long)A.gall_pl * kLallPl) {
        const long p = t - nreg;
        F = &A.pl;
        base = (size_t)(p / A.gall_pl) * A.kall * A.gall_pl + (size_t)(p % A.gall_pl);
        s = A.gall_pl;
    } else {
        return;
    }
    const int kall = A.kall, kmin = 1, kmax = kall - 2;  // 0-based
    const double dt = A.dt, CV = A.CVovRt2, grav = A.grav;
    auto at = [&](int k) { return base + (size_t)k * s; };

    // transposed elimination on kmin+1..kmax
    double beta = F->Mc[at(kmin + 1)];
    double yp = (F->wb[at(kmin + 1)] * F->GSGAM2H[at(kmin + 1)]) / beta;
    F->w0b[at(kmin + 1)] = yp;
    for (int k = kmin + 2; k <= kmax; ++k) {
        const double mu_km1 = F->Mu[at(k - 1)];
        const double g = F->Ml[at(k)] / beta;
        beta = F->Mc[at(k)] - mu_km1 * g;
        yp = (F->wb[at(k)] * F->GSGAM2H[at(k)] - mu_km1 * yp) / beta;
        F->Srb[at(k)] = g;
        F->w0b[at(k)] = yp;
    }
    const double y_kmax = yp;

    // back substitution fused with the cotangents of each level, top down
    double yn = 0.0, g_k1 = 0.0;
    for (int k = kall - 1; k >= 0; --k) {
        const bool in_k = k >= kmin + 1 && k <= kmax, in_n = k + 1 >= kmin + 1 && k + 1 <= kmax;
        const size_t o = at(k);
        double yk = 0.0, gk = 0.0;
        if (in_k) {
            const double ypk = F->w0b[o];
            yk = (k == kmax) ? ypk : ypk - g_k1 * yn;
            if (k >= kmin + 2) gk = F->Srb[o];
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
        const double p0b = pb * F->RGSGAM2[o];
        const double rg = F->RGAM[o];
        const double r0b = rb * (rg * rg);
        const double rgh = F->RGAMH[o];
        const double sw = in_k ? sb * (rgh * rgh) : 0.0;
        F->p0b[o] = p0b;
        F->Spb[o] = dt * p0b;
        F->r0b[o] = r0b;
        F->Srb[o] = dt * r0b;
        F->w0b[o] = sw * A.alpha;
        F->Swb[o] = dt * sw;
        if (k == kmin) {
            const double bb = F->wb[o] * F->GSGAM2H[o] - F->Ml[at(kmin + 1)] * yn;
            F->wb[o] = bb * F->RGSGAM2H[o];
        } else if (k == kmax + 1) {
            const double bt = F->wb[o] * F->GSGAM2H[o] - F->Mu[at(kmax)] * y_kmax;
            F->wb[o] = bt * F->RGSGAM2H[o];
        } else if (in_k) {
            F->wb[o] = 0.0;
        }
        yn = yk;
        g_k1 = gk;
    }
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
    double qmin, qmax, rmin[3], rmax[3];
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
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double dnx = (m == nmax) ? 1.0 : 0.0, dnn = (m == nmin) ? 1.0 : 0.0;
        const double cqx = ((m == maxL) ? cL : 0.0) + ((m == maxU) ? cU : 0.0);
        const double cqn = ((m == minL) ? cL : 0.0) + ((m == minU) ? cU : 0.0);
        r.rmin[m] = fac * (e * dnx - cqx);
        r.rmax[m] = fac * (e * dnn - cqn);
    }
    r.rmin[1] = (r.rmin[1] + fac) + zerosw;
    r.rmax[1] = (r.rmax[1] + fac) + zerosw;
    return r;
}

struct VlaField {
    double *qhb, *qb;
    const double *q_h, *q, *d, *ck;
    const uint32_t *mask;
};

struct VlaArgs {
    VlaField reg, pl;
    int gall, kall, lall, gall_pl, have_pl, nword;
    double eps, big;
};

template <bool POLE>
__device__ __forceinline__ void vlim_adjoint_column(const VlaField &F, size_t base, size_t s, size_t ck2,
                                                    const uint32_t *mw, size_t mstride, int kall, double EPS,
                                                    double BIG)
{
    const int kmin = 1, kmax = kall - 2;  // 0-based
    auto at = [&](int k) { return base + (size_t)k * s; };
    // level kmin (the level below the first limited face)
    double qm = F.q[at(kmin - 1)], q0 = F.q[at(kmin)], qp = F.q[at(kmin + 1)];
    double c1 = F.ck[at(kmin)], c1p = F.ck[at(kmin + 1)];
    VlimLvl lo = vlim_level<POLE>(qm, q0, qp, c1, c1p, F.ck[ck2 + at(kmin)], F.d[at(kmin)], EPS, BIG);
    // q cotangents of levels k-2 .. k+1 while face k is processed
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
    uint32_t word = 0;
    for (int k = kmin + 1; k <= kmax; ++k) {
        const int p = k - (kmin + 1);
        if ((p & 31) == 0) word = mw[(size_t)(p >> 5) * mstride];
        qm = q0;
        q0 = qp;
        qp = F.q[at(k + 1)];
        c1 = c1p;
        c1p = F.ck[at(k + 1)];
        const VlimLvl hi = vlim_level<POLE>(qm, q0, qp, c1, c1p, F.ck[ck2 + at(k)], F.d[at(k)], EPS, BIG);
        const bool clipped = (word >> (p & 31)) & 1u;
        const double inflagL = 0.5 - sgn_half(c1);
        const size_t o = at(k);
        const double ob = F.qhb[o];
        double row[4] = {0.0, 0.0, 0.0, 0.0};
        double cqh = 1.0;
        if (clipped) {
            cqh = 0.0;
            const double out = F.q_h[o];
            const bool from_lo = inflagL > 0.0;
            const VlimLvl &L = from_lo ? lo : hi;
            const double *r = (out == L.qmin) ? L.rmin : L.rmax;
            if (from_lo) {
                row[0] = r[0];
                row[1] = r[1];
                row[2] = r[2];
            } else {
                row[1] = r[0];
                row[2] = r[1];
                row[3] = r[2];
            }
        }
        F.qhb[o] = cqh * ob;
        w0 = w0 + row[0] * ob;
        w1 = w1 + row[1] * ob;
        w2 = w2 + row[2] * ob;
        w3 = w3 + row[3] * ob;
        F.qb[at(k - 2)] = w0;  // no later face reaches level k-2
        w0 = w1;
        w1 = w2;
        w2 = w3;
        w3 = 0.0;
        lo = hi;
    }
    F.qb[at(kmax - 1)] = w0;
    F.qb[at(kmax)] = w1;
    F.qb[at(kmax + 1)] = w2;
}

constexpr int BLOCK_VLA = 256;

__global__ void __launch_bounds__(BLOCK_VLA) vlim_adjoint_kernel(VlaArgs A)
{
    const long t = (long)blockIdx.x * BLOCK_VLA + threadIdx.x;
    const long nreg = (long)A.gall * A.lall;
    if (t < nreg) {
        const int g = (int)(t % A.gall), l = (int)(t / A.gall);
        vlim_adjoint_column<false>(A.reg, (size_t)l * A.kall * A.gall + g, A.gall,
                                   (size_t)A.gall * A.kall * A.lall, A.reg.mask + (size_t)l * A.nword * A.gall + g,
                                   A.gall, A.kall, A.eps, A.big);
    } else if (A.have_pl && t < nreg + (long)A.gall_pl * kLallPl) {
        const int p = (int)(t - nreg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        vlim_adjoint_column<true>(A.pl, (size_t)l * A.kall * A.gall_pl + g, A.gall_pl,
                                  (size_t)A.gall_pl * A.kall * kLallPl,
                                  A.pl.mask + (size_t)l * A.nword * A.gall_pl + g, A.gall_pl, A.kall, A.eps, A.big);
    }
}

}  // namespace

int launch_vlim_adjoint(nicam_ctx c, double *qhb, double *qhb_pl, double *qb, double *qb_pl, const double *q_h,
                        const double *q_h_pl, const double *q, const double *q_pl, const double *d,
                        const double *d_pl, const double *ck, const double *ck_pl, const uint32_t *mask,
                        const uint32_t *mask_pl)
{
    const Dims &D = c->d;
    if (D.kall - 3 <= 0) return NICAM_OK;
    VlaArgs A;
    A.reg = VlaField{qhb, qb, q_h, q, d, ck, mask};
    A.pl = VlaField{qhb_pl, qb_pl, q_h_pl, q_pl, d_pl, ck_pl, mask_pl};
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    A.nword = (D.kall - 3 + 31) / 32;
    A.eps = c->eps;
    A.big = c->big;
    const long ncol = (long)D.gall * D.lall + (D.have_pl ? (long)D.gall_pl * kLallPl : 0);
    hipLaunchKernelGGL(vlim_adjoint_kernel, dim3((unsigned)((ncol + BLOCK_VLA - 1) / BLOCK_VLA)), dim3(BLOCK_VLA), 0,
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
                    c->RGSGAM2H, c->GSGAM2H};
    A.pl = VaField{wb_pl, w0b_pl, p0b_pl, r0b_pl, Srb_pl, Swb_pl, Spb_pl, c->Mc_pl, c->Ml_pl, c->Mu_pl,
                   c->RGAM_pl, c->RGAMH_pl, c->RGSGAM2_pl, c->RGSGAM2H_pl, c->GSGAM2H_pl};
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
    hipLaunchKernelGGL(vi_adjoint_kernel, dim3((unsigned)((A.ncol + BLOCK_VA - 1) / BLOCK_VA)), dim3(BLOCK_VA), 0,
                       c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
