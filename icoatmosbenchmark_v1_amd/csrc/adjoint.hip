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

}  // namespace

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
