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
    double beta = FACT ? F->tbeta[at(kmin + 1)] : F->Mc[at(kmin + 1)];
    double yp = (F->wb[at(kmin + 1)] * F->GSGAM2H[at(kmin + 1)]) / beta;
    F->w0b[at(kmin + 1)] = yp;
    for (int k = kmin + 2; k <= kmax; ++k) {
        const double mu_km1 = F->Mu[at(k - 1)];
        if constexpr (FACT) {
            beta = F->tbeta[at(k)];
        } else {
            const double g = F->Ml[at(k)] / beta;
            beta = F->Mc[at(k)] - mu_km1 * g;
            F->Srb[at(k)] = g;
        }
        yp = (F->wb[at(k)] * F->GSGAM2H[at(k)] - mu_km1 * yp) / beta;
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
            if (k >= kmin + 2) gk = FACT ? F->tgam[o] : F->Srb[o];
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

// The same column with kall fixed at compile time (used at 42): the partial
// solution and the elimination factors stay in registers (the loops are
// unrolled), so the output arrays are written once and never re-read.  The
// operations and their order are those of vi_adjoint_kernel.
constexpr int BLOCK_VAR = 64;

template <int KC>
__global__ void __launch_bounds__(BLOCK_VAR) vi_adjoint_kernel_reg(VaArgs A)
{
    const long t = (long)blockIdx.x * BLOCK_VAR + threadIdx.x;
    const long nreg = (long)A.gall * A.lall;
    VaField F;
    size_t base, s;
    if (t < nreg) {
        F = A.reg;
        base = (size_t)(t / A.gall) * KC * A.gall + (size_t)(t % A.gall);
        s = A.gall;
    } else if (A.have_pl && t < nreg + (long)A.gall_pl * kLallPl) {
        const long p = t - nreg;
        F = A.pl;
        base = (size_t)(p / A.gall_pl) * KC * A.gall_pl + (size_t)(p % A.gall_pl);
        s = A.gall_pl;
    } else {
        return;
    }
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
        lo = hi;
    }
    if (kl - 1 >= mb && kl - 1 < me) F.qb[at(kl - 1)] = w0;
    if (kl >= mb && kl < me) F.qb[at(kl)] = w1;
    if (kl + 1 >= mb && kl + 1 < me) F.qb[at(kl + 1)] = w2;
}

__global__ void __launch_bounds__(BLOCK_VLA) vlim_adjoint_kernel(VlaArgs A)
{
    extern __shared__ double s_ob[];  // [kall][VCOL]
    const int c = threadIdx.x % VCOL, r = threadIdx.x / VCOL;
    const long t = (long)blockIdx.x * VCOL + c;
    const long nreg = (long)A.gall * A.lall;
    const long ncol = nreg + (A.have_pl ? (long)A.gall_pl * kLallPl : 0);
    VlaField F{};  // by value: a pointer into the kernel arguments would copy them to scratch memory
    bool have = false;
    size_t base = 0, s = 0, ck2 = 0, mstride = 0;
    const uint32_t *mw = nullptr;
    if (t < nreg) {
        const int g = (int)(t % A.gall), l = (int)(t / A.gall);
        F = A.reg;
        have = true;
        base = (size_t)l * A.kall * A.gall + g;
        s = A.gall;
        ck2 = (size_t)A.gall * A.kall * A.lall;
        mw = A.reg.mask + (size_t)l * A.nword * A.gall + g;
        mstride = A.gall;
    } else if (t < ncol) {
        const int p = (int)(t - nreg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        F = A.pl;
        have = true;
        base = (size_t)l * A.kall * A.gall_pl + g;
        s = A.gall_pl;
        ck2 = (size_t)A.gall_pl * A.kall * kLallPl;
        mw = A.pl.mask + (size_t)l * A.nword * A.gall_pl + g;
        mstride = A.gall_pl;
    }
    if (have)
        for (int k = r; k < A.kall; k += VRUN) s_ob[(size_t)k * VCOL + c] = F.qhb[base + (size_t)k * s];
    __syncthreads();
    if (!have) return;
    if (t < nreg)
        vlim_adjoint_run<false>(F, base, s, ck2, mw, mstride, A.kall, A.eps, A.big, s_ob + c, r);
    else
        vlim_adjoint_run<true>(F, base, s, ck2, mw, mstride, A.kall, A.eps, A.big, s_ob + c, r);
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
// One work-group per CHUNK cells of one (k, l) slab, as in the forward: q of
// the window [g0-iall-1, g0+CHUNK+iall+2) staged in LDS, and the cotangents of
// q and Qout_post over the same window accumulated in LDS (every scatter
// target lies in it), then added to global memory once per window cell.
// Neighbouring windows overlap, so those adds are atomic and q_b, Qout_post_b
// are cleared first; sums of a cell's few terms may therefore round in any
// order.  Faces of q_a are owned by one work-item each.
constexpr int HCHUNK = 256;
constexpr int HMAXI = 160;

struct HaArgs {
    double *qab, *qb, *Qpostb;
    const double *Qprevb;
    const double *q_a, *q, *d, *ch, *cmask, *Qpost;
    const uint8_t *mask;
    const int *have_sgp;
    int iall, gall, kall, lall, nmain;
    double eps, big;
    // poles
    double *qab_pl, *qb_pl, *Qpostb_pl;
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

__global__ void __launch_bounds__(HCHUNK) hlim_adjoint_pole(HaArgs A);

__global__ void __launch_bounds__(HCHUNK) hlim_adjoint_kernel(HaArgs A)
{
    constexpr int SPAN = HCHUNK + 2 * HMAXI + 3;
    __shared__ double s_q[SPAN], s_b[SPAN], s_pn[SPAN], s_px[SPAN];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int nchunk = (gall + HCHUNK - 1) / HCHUNK;
    const int b = blockIdx.x;
    const int chunk = b % nchunk, kl = b / nchunk;
    const int k = kl % kall, l = kl / kall;
    const int g0 = chunk * HCHUNK, lo = g0 - iall - 1, n_in = HCHUNK + 2 * iall + 3;
    const size_t fb = ((size_t)l * kall + k) * gall;
    const size_t plane = (size_t)lall * kall * gall;
    const int tid = threadIdx.x;
    for (int s = tid; s < n_in; s += HCHUNK) {
        const int gg = lo + s;
        s_q[s] = (gg >= 0 && gg < gall) ? A.q[fb + gg] : 0.0;
        s_b[s] = 0.0;
        s_pn[s] = 0.0;
        s_px[s] = 0.0;
    }
    __syncthreads();
    auto Q = [&](int p) { return s_q[p - lo]; };
    auto addq = [&](int p, double v) { atomicAdd(&s_b[p - lo], v); };
    const int g = g0 + tid;
    if (g < gall) {
        const int i = g % iall, j = g / iall;
        const bool sgp = A.have_sgp[l] != 0;
        const double BIG = A.big;
        if (i < 1 || i > iall - 2 || j < 1 || j > iall - 2) {  // halo: Qout = q (HLIM:286-301)
            addq(g, A.Qprevb[fb + g]);
            addq(g, A.Qprevb[plane + fb + g]);
        } else {
            int cmn[6], cmx[6];
            double vmn[6], vmx[6], cc[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                const int m = f < 3 ? f : f - 3;
                const int owner = f < 3 ? g : g - (m == 0 ? 1 : (m == 1 ? iall + 1 : iall));
                int c4[4];
                ha_arc(m, owner, iall, sgp, c4);
                double lo_v = Q(c4[0]), hi_v = lo_v;
                int il = c4[0], ih = c4[0];
#pragma unroll
                for (int t = 1; t < 4; ++t) {
                    const double v = Q(c4[t]);
                    if (v < lo_v) { lo_v = v; il = c4[t]; }
                    if (v > hi_v) { hi_v = v; ih = c4[t]; }
                }
                const double c = A.cmask[(size_t)m * plane + fb + owner];
                cmn[f] = il;
                cmx[f] = ih;
                cc[f] = f < 3 ? c : 1.0 - c;
                if (f < 3) {
                    vmn[f] = c * lo_v + (1.0 - c) * BIG;
                    vmx[f] = c * hi_v + (1.0 - c) * (-BIG);
                } else {
                    vmn[f] = c * BIG + (1.0 - c) * lo_v;
                    vmx[f] = c * (-BIG) + (1.0 - c) * hi_v;
                }
            }
            const double qij = Q(g);
            int smin = -1, smax = -1;
            double nmin = qij, nmax = qij;
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                if (vmn[f] < nmin) { nmin = vmn[f]; smin = f; }
                if (vmx[f] > nmax) { nmax = vmx[f]; smax = f; }
            }
            double chv[6], chm[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                chv[f] = A.ch[(size_t)f * plane + fb + g];
                chm[f] = dmn(chv[f], 0.0);
            }
            const double Cin = chm[0] + chm[1] + chm[2] + chm[3] + chm[4] + chm[5];
            const double Cout = chv[0] - chm[0] + chv[1] - chm[1] + chv[2] - chm[2] + chv[3] - chm[3] + chv[4] -
                                chm[4] + chv[5] - chm[5];
            const double zerosw = 0.5 - sgn_half(fabs(Cout) - A.eps);
            const double E = 1.0 - Cin - Cout + A.d[fb + g];
            const double fac = (1.0 - zerosw) / (Cout + zerosw);
#pragma unroll
            for (int mm = 0; mm < 2; ++mm) {
                const double bq = A.Qprevb[(size_t)mm * plane + fb + g];
                const int sel = mm == 0 ? smax : smin;
                const int *cx = mm == 0 ? cmx : cmn;
                addq(g, (fac + zerosw) * bq);
                addq(sel < 0 ? g : cx[sel], (-(fac * E) * (sel < 0 ? 1.0 : cc[sel])) * bq);
#pragma unroll
                for (int f = 0; f < 6; ++f) addq(cx[f], (-(fac * chm[f]) * cc[f]) * bq);
            }
        }
        // faces 1-3 of the own arcs (HLIM:377-410)
        if (i <= iall - 2 && j <= iall - 2) {
            const uint32_t bits = A.mask[fb + g];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int nb = g + (m == 0 ? 1 : (m == 1 ? iall + 1 : iall));
                const size_t fo = (size_t)m * plane + fb + g, fn = (size_t)(m + 3) * plane + fb + nb;
                const double c = A.cmask[fo];
                const int T = c > 0.0 ? nb : g;
                const double ob = A.qab[fo] + A.qab[fn];
                int kind = 0, cell = 0;
                if ((bits >> m) & 1u) {
                    const double out = A.q_a[fo];
                    int c4[4];
                    ha_arc(m, g, iall, sgp, c4);
                    double lo_v = Q(c4[0]), hi_v = lo_v;
                    int il = c4[0], ih = c4[0];
#pragma unroll
                    for (int t = 1; t < 4; ++t) {
                        const double v = Q(c4[t]);
                        if (v < lo_v) { lo_v = v; il = c4[t]; }
                        if (v > hi_v) { hi_v = v; ih = c4[t]; }
                    }
                    if (out == A.Qpost[fb + T]) kind = 1;
                    else if (out == A.Qpost[plane + fb + T]) kind = 2;
                    else if (out == lo_v) { kind = 3; cell = il; }
                    else { kind = 4; cell = ih; }
                }
                A.qab[fn] = 0.0;
                A.qab[fo] = kind == 0 ? ob : 0.0;
                if (kind == 1) atomicAdd(&s_pn[T - lo], ob);
                else if (kind == 2) atomicAdd(&s_px[T - lo], ob);
                else if (kind >= 3) addq(cell, ob);
            }
        }
    }
    __syncthreads();
    for (int s = tid; s < n_in; s += HCHUNK) {
        const int gg = lo + s;
        if (gg < 0 || gg >= gall) continue;
        if (s_b[s] != 0.0) atomicAdd(&A.qb[fb + gg], s_b[s]);
        if (s_pn[s] != 0.0) atomicAdd(&A.Qpostb[fb + gg], s_pn[s]);
        if (s_px[s] != 0.0) atomicAdd(&A.Qpostb[plane + fb + gg], s_px[s]);
    }
}

// Poles: one work-item per (k, l) slab, sequential as the oracle (HLIM:307-363,
// 412-425); it owns the slab's pole arrays, so no atomics.
__global__ void __launch_bounds__(HCHUNK) hlim_adjoint_pole(HaArgs A)
{
    const int t = blockIdx.x * HCHUNK + threadIdx.x;
    const int kall = A.kall, npl = A.gall_pl;
    if (!A.have_pl || t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const size_t pplane = (size_t)kLallPl * kall * npl;
    const double *qs = A.q_pl + fb, *chs = A.ch_pl + fb, *cms = A.cmask_pl + fb;
    double *xq = A.qb_pl + fb, *xa = A.qab_pl + fb, *xp = A.Qpostb_pl + fb;
    for (int v = 0; v < npl; ++v) {
        xq[v] = 0.0;
        xp[v] = 0.0;
        xp[pplane + v] = 0.0;
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
    for (int mm = 0; mm < 2; ++mm) {
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
    for (int v = 1; v < npl; ++v) {
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

__global__ void fill_zero_kernel(double *p, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

// ---- horizontal_flux (dyn_horiz_adv_flux/src/mod_src_tracer.f90:42-379) -------
//
// Linear in rho*v; GRD_xc carries 1/max(rho_arc, EPS).  The arc cotangents
// are gathered, not scattered: a work-group of cells of one slab
// computes the cotangent of every arc whose triangles touch its cells
// (origins [g0-iall-1, g0+cells+iall)) into LDS, then each work-item sums
// its cell's six triangles in the order of a sequential sweep over triangle
// origins, so the result is the oracle's, bit for bit, with no atomics.  The
// tracer fields are read from HBM directly (neighbouring arcs hit the
// caches), which keeps LDS to the arc cotangents.  (Staging them too, with
// 256-cell chunks, ran at 3.05 ms at gl7; 512-cell chunks with two passes
// over the origins at 1.91 ms.)
// A work-group of FBLOCK work-items covers FBLOCK - (2*iall+1) cells, so the
// arc phase (FBLOCK origins) takes one pass.
constexpr int FBLOCK = 1024;

struct FaArgs {
    const double *flxb, *xcb;   // (gall,kall,lall,6), (gall,kall,lall,3,3)
    double *Xb[4];              // rho_b, rhovx_b, rhovy_b, rhovz_b
    const double *X[4];
    const double *gt, *ga, *gp;
    const int *have_sgp;
    int iall, gall, kall, lall, cells;
    double dt, eps;
    // poles
    const double *flxb_pl, *xcb_pl;
    double *Xb_pl[4];
    const double *X_pl[4];
    const double *gt_pl, *ga_pl, *gp_pl;
    int gall_pl, have_pl;
};

__global__ void __launch_bounds__(FBLOCK) hflux_adjoint_kernel(FaArgs A)
{
    extern __shared__ double lds[];
    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall, cells = A.cells;
    const int nchunk = (gall + cells - 1) / cells;
    const int b = blockIdx.x;
    const int chunk = b % nchunk, kl = b / nchunk;
    const int k = kl % kall, l = kl / kall;
    const int g0 = chunk * cells;
    const int alo = g0 - iall - 1, na = cells + 2 * iall + 1;
    double *s_a = lds;  // [3 arcs][4 fields][na]
    const size_t fb = ((size_t)l * kall + k) * gall;
    const size_t plane = (size_t)lall * kall * gall;
    const size_t mstr = (size_t)lall * gall, mb = (size_t)l * gall;
    const bool sgp = A.have_sgp[l] != 0;
    const int tid = threadIdx.x;
    auto W = [&](int c, int t, int w) { return A.gt[(size_t)(w * 2 + t) * mstr + mb + c]; };  // w: 2,3,4 = W1..W3
    auto X = [&](int f, int p) { return A.X[f][fb + p]; };
    auto tri = [&](int f, int t, int c) {  // forward triangle sums (FLX:139-201)
        if (t == 0 && sgp && c == 0) t = 1, c = 1;
        if (t == 0) return X(f, c) * W(c, 0, 2) + X(f, c + 1) * W(c, 0, 3) + X(f, c + iall + 1) * W(c, 0, 4);
        return X(f, c) * W(c, 1, 2) + X(f, c + iall + 1) * W(c, 1, 3) + X(f, c + iall) * W(c, 1, 4);
    };
    const double dt = A.dt, EPS = A.eps;
    for (int s = tid; s < na; s += FBLOCK) {
        const int a0 = alo + s;
        const int i = a0 >= 0 ? a0 % iall : -1, j = a0 >= 0 ? a0 / iall : -1;
        const bool org = a0 >= 0 && a0 < gall && i <= iall - 2 && j <= iall - 2;
        // the four triangles the origin's arcs use: TI(a0), TJ(a0), TJ(a0-iall) (AI), TI(a0-1) (AJ).
        // They touch six cells (a0, its E, N, NE, S, W neighbours): each field is read there once, and
        // the sums are formed in the forward's order.  The singular-point substitution (TI(0) = TJ(1))
        // reaches only origins 0 and 1 and takes the general path.
        double tI[4], tJ[4], tJs[4], tIw[4];
        if (org && !(sgp && a0 <= 1)) {
            const bool hs = j >= 1, hw = i >= 1;
            const double wi1 = W(a0, 0, 2), wi2 = W(a0, 0, 3), wi3 = W(a0, 0, 4);
            const double wj1 = W(a0, 1, 2), wj2 = W(a0, 1, 3), wj3 = W(a0, 1, 4);
            const int cs = a0 - iall, cw = a0 - 1;
            const double ws1 = hs ? W(cs, 1, 2) : 0.0, ws2 = hs ? W(cs, 1, 3) : 0.0, ws3 = hs ? W(cs, 1, 4) : 0.0;
            const double ww1 = hw ? W(cw, 0, 2) : 0.0, ww2 = hw ? W(cw, 0, 3) : 0.0, ww3 = hw ? W(cw, 0, 4) : 0.0;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const double x0 = X(f, a0), xe = X(f, a0 + 1), xn = X(f, a0 + iall), xne = X(f, a0 + iall + 1);
                const double xs = hs ? X(f, cs) : 0.0, xw = hw ? X(f, cw) : 0.0;
                tI[f] = x0 * wi1 + xe * wi2 + xne * wi3;
                tJ[f] = x0 * wj1 + xne * wj2 + xn * wj3;
                tJs[f] = hs ? xs * ws1 + xe * ws2 + x0 * ws3 : 0.0;
                tIw[f] = hw ? xw * ww1 + x0 * ww2 + xn * ww3 : 0.0;
            }
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                tI[f] = org ? tri(f, 0, a0) : 0.0;
                tJ[f] = org ? tri(f, 1, a0) : 0.0;
                tJs[f] = (org && j >= 1) ? tri(f, 1, a0 - iall) : 0.0;
                tIw[f] = (org && i >= 1) ? tri(f, 0, a0 - 1) : 0.0;
            }
        }
#pragma unroll
        for (int a = 1; a <= 3; ++a) {
            const bool valid = org && j >= (a == 1 ? 1 : 0) && i >= (a == 3 ? 1 : 0);
            double bv[4] = {0.0, 0.0, 0.0, 0.0};
            if (valid) {
                const int g = a0;
                const int nb = g + (a == 1 ? 1 : (a == 2 ? iall + 1 : iall));
                double s2[4];
#pragma unroll
                for (int f = 0; f < 4; ++f)  // FLX:234-309: the arc's two triangles, in the forward's order
                    s2[f] = a == 1 ? tJs[f] + tI[f] : (a == 2 ? tI[f] + tJ[f] : tJ[f] + tIw[f]);
                const double rr = 1.0 / dmx(s2[0], EPS);
                const bool live_rho = !(EPS > s2[0]);
                const double ra = A.gp[1 * mstr + mb + g] * dt, rn = A.gp[1 * mstr + mb + nb] * dt;
                const bool sgp_face = sgp && a == 3 && nb == iall + 1;
                const double fb_nb = sgp_face ? 0.0 : A.flxb[(size_t)(a + 2) * plane + fb + nb];
                const double fluxb = A.flxb[(size_t)(a - 1) * plane + fb + g] * ra - fb_nb * rn;
                double r2b = 0.0;
#pragma unroll
                for (int dd = 0; dd < 3; ++dd) {
                    const double hn = A.ga[(size_t)(dd * 3 + (a - 1)) * mstr + mb + g];
                    const double xb = A.xcb[(size_t)(dd * 3 + (a - 1)) * plane + fb + g];
                    bv[1 + dd] = 0.5 * hn * fluxb - xb * rr * dt * 0.5;
                    r2b = r2b + xb * s2[1 + dd] * dt * 0.5 * rr * rr;
                }
                bv[0] = live_rho ? r2b : 0.0;
            }
#pragma unroll
            for (int f = 0; f < 4; ++f) s_a[((a - 1) * 4 + f) * na + s] = bv[f];
        }
    }
    __syncthreads();
    const int X0 = g0 + tid;
    if (tid >= cells || X0 >= gall) return;
    auto AB = [&](int a, int f, int c) { return s_a[((a - 1) * 4 + f) * na + c - alo]; };
    // triangle origins of the cell's six triangles, and whether each is one (i, j in [0, iall-2])
    const int i0 = X0 % iall, j0 = X0 / iall;
    const bool o_sw = i0 >= 1 && j0 >= 1 && i0 - 1 <= iall - 2 && j0 - 1 <= iall - 2;  // X0-iall-1
    const bool o_s = j0 >= 1 && i0 <= iall - 2 && j0 - 1 <= iall - 2;                  // X0-iall
    const bool o_w = i0 >= 1 && i0 - 1 <= iall - 2 && j0 <= iall - 2;                  // X0-1
    const bool o_c = i0 <= iall - 2 && j0 <= iall - 2;                                 // X0
    const int c_sw = X0 - iall - 1, c_s = X0 - iall, c_w = X0 - 1;
    const double w_sw_i = o_sw ? W(c_sw, 0, 4) : 0.0, w_sw_j = o_sw ? W(c_sw, 1, 3) : 0.0;
    const double w_s_j = o_s ? W(c_s, 1, 4) : 0.0, w_w_i = o_w ? W(c_w, 0, 3) : 0.0;
    const double w_c_i = o_c ? W(X0, 0, 2) : 0.0, w_c_j = o_c ? W(X0, 1, 2) : 0.0;
    // triangle cotangents: TI(c) from AI(c), AIJ(c), AJ(c+1); TJ(c) from AI(c+iall), AIJ(c), AJ(c)
    auto tib0 = [&](int f, int c) { return ((0.0 + AB(1, f, c)) + AB(2, f, c)) + AB(3, f, c + 1); };
    auto tjb0 = [&](int f, int c) { return ((0.0 + AB(1, f, c + iall)) + AB(2, f, c)) + AB(3, f, c); };
    auto tib = [&](int f, int c) { return (sgp && c == 0) ? 0.0 : tib0(f, c); };
    auto tjb = [&](int f, int c) { return (sgp && c == 1) ? tjb0(f, 1) + tib0(f, 0) : tjb0(f, c); };
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double y = 0.0;
        if (o_sw) y = y + (tib(f, c_sw) * w_sw_i + tjb(f, c_sw) * w_sw_j);
        if (o_s) y = y + tjb(f, c_s) * w_s_j;
        if (o_w) y = y + tib(f, c_w) * w_w_i;
        if (o_c) y = y + (tib(f, X0) * w_c_i + tjb(f, X0) * w_c_j);
        A.Xb[f][fb + X0] = y;
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
    const int iall = D.gall_1d;
    A.cells = FBLOCK - (2 * iall + 1);
    const long nblk = (long)((D.gall + A.cells - 1) / A.cells) * D.kall * D.lall;
    if (nblk > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    const size_t lds = sizeof(double) * 12 * (size_t)FBLOCK;
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute((const void *)hflux_adjoint_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(hflux_adjoint_kernel, dim3((unsigned)nblk), dim3(FBLOCK), lds, c->stream, A);
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
                        const uint8_t *mask_pl)
{
    const Dims &D = c->d;
    if (D.gall_1d > HMAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (160)");
    HaArgs A;
    A.qab = qab; A.qb = qb; A.Qpostb = Qpostb; A.Qprevb = Qprevb;
    A.q_a = q_a; A.q = q; A.d = d; A.ch = ch; A.cmask = cmask; A.Qpost = Qpost; A.mask = mask;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d; A.gall = D.gall; A.kall = D.kall; A.lall = D.lall;
    A.eps = c->eps; A.big = c->big;
    A.qab_pl = qab_pl; A.qb_pl = qb_pl; A.Qpostb_pl = Qpostb_pl; A.Qprevb_pl = Qprevb_pl;
    A.q_a_pl = q_a_pl; A.q_pl = q_pl; A.d_pl = d_pl; A.ch_pl = ch_pl; A.cmask_pl = cmask_pl;
    A.Qpost_pl = Qpost_pl; A.mask_pl = mask_pl;
    A.gall_pl = D.gall_pl; A.have_pl = D.have_pl;
    const long nmain = (long)((D.gall + HCHUNK - 1) / HCHUNK) * D.kall * D.lall;
    if (nmain > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    const size_t nreg = (size_t)D.gall * D.kall * D.lall;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(2048), dim3(256), 0, c->stream, qb, nreg);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(2048), dim3(256), 0, c->stream, Qpostb, 2 * nreg);
    hipLaunchKernelGGL(hlim_adjoint_kernel, dim3((unsigned)nmain), dim3(HCHUNK), 0, c->stream, A);
    if (D.have_pl)
        hipLaunchKernelGGL(hlim_adjoint_pole, dim3((unsigned)((D.kall * kLallPl + HCHUNK - 1) / HCHUNK)),
                           dim3(HCHUNK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

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
    const size_t lds = sizeof(double) * VCOL * D.kall;
    if (lds > 160 * 1024) return fail(NICAM_ERR_ARG, "kall too large for the staged cotangents");
    if (lds > 64 * 1024)
        NICAM_HIP(hipFuncSetAttribute((const void *)vlim_adjoint_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(vlim_adjoint_kernel, dim3((unsigned)((ncol + VCOL - 1) / VCOL)), dim3(BLOCK_VLA), lds,
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
        hipLaunchKernelGGL(vi_adjoint_kernel_reg<42>, dim3((unsigned)((A.ncol + BLOCK_VAR - 1) / BLOCK_VAR)),
                           dim3(BLOCK_VAR), 0, c->stream, A);
    } else if (variant == 1 || variant == 3) {
        hipLaunchKernelGGL(vi_adjoint_kernel<false>, dim3((unsigned)((A.ncol + BLOCK_VA - 1) / BLOCK_VA)),
                           dim3(BLOCK_VA), 0, c->stream, A);
    } else {
        hipLaunchKernelGGL(vi_adjoint_kernel<true>, dim3((unsigned)((A.ncol + BLOCK_VA - 1) / BLOCK_VA)),
                           dim3(BLOCK_VA), 0, c->stream, A);
    }
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
