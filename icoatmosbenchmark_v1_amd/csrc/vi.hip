// vi_rhow_solver on gfx950.
//
// Reference: kernels/NICAM/dyn_vi_rhow_solver/src/mod_vi.f90:93-306.
//
// The reference builds the Helmholtz source Sall for a whole region, applies
// the boundary conditions, then runs a Thomas elimination per column, all with
// g innermost.  Here one work-item owns one column (g, l): the source term is
// formed level by level inside the forward sweep (the level k-1 products are
// carried in registers, so no Sall array exists), the elimination factors
// gamma(k) are kept in LDS (one slot per lane and level), and the forward
// solution is written to rhogw and read back by the back substitution.  A
// wavefront covers 64 consecutive g at one level, so all global traffic is
// contiguous.  The pole columns are appended as the last work-items.
#include "internal.h"

namespace nicam {
namespace {

constexpr int BLOCK = 64;

struct ViField {
    double *w;
    const double *w0, *p0, *r0, *Sr, *Sw, *Sp;
    const double *Mc, *Ml, *Mu;
    const double *RGAM, *RGAMH, *RGSGAM2, *RGSGAM2H, *GSGAM2H;
};

struct ViArgs {
    ViField reg, pl;
    const double *rdgzh, *afact, *bfact;
    int gall, kall, lall, gall_pl, have_pl;
    double dt, CVovRt2, alpha, grav;
};

__device__ void vi_column(const ViField &F, size_t base, size_t s, int kall, const double *rdgzh,
                          const double *afact, const double *bfact, double dt, double CVovRt2,
                          double alpha, double grav, double *gam /* LDS, stride BLOCK */)
{
    const int kmin = 1, kmax = kall - 2;  // 0-based
    auto at = [&](int k) { return base + (size_t)k * s; };

    // mod_vi.f90:209-210: boundary values of rhogw scaled in place
    const double w_kmin = F.w[at(kmin)] * F.RGSGAM2H[at(kmin)];
    const double w_kmax1 = F.w[at(kmax + 1)] * F.RGSGAM2H[at(kmax + 1)];

    // level kmin products for the k = kmin+1 source (mod_vi.f90:196-203)
    double P_km1, R_km1;
    {
        const size_t o = at(kmin);
        const double rg = F.RGAM[o];
        P_km1 = (F.p0[o] + dt * F.Sp[o]) * F.RGSGAM2[o];
        R_km1 = (F.r0[o] + dt * F.Sr[o]) * (rg * rg);
    }
    double beta = 0.0, w_km1 = 0.0, Mu_km1 = 0.0;
    for (int k = kmin + 1; k <= kmax; ++k) {
        const size_t o = at(k);
        const double rgh = F.RGAMH[o], rg = F.RGAM[o];
        const double P_k = (F.p0[o] + dt * F.Sp[o]) * F.RGSGAM2[o];
        const double R_k = (F.r0[o] + dt * F.Sr[o]) * (rg * rg);
        double S = ((F.w0[o] * alpha + dt * F.Sw[o]) * (rgh * rgh) - (P_k - P_km1) * dt * rdgzh[k] -
                    (R_k * afact[k] + R_km1 * bfact[k]) * dt * grav) *
                   CVovRt2;
        const double Mc = F.Mc[o], Ml = F.Ml[o], Mu = F.Mu[o];
        if (k == kmin + 1) S = S - Ml * w_kmin;  // mod_vi.f90:211
        if (k == kmax) S = S - Mu * w_kmax1;     // mod_vi.f90:212
        double w;
        if (k == kmin + 1) {  // mod_vi.f90:218-222
            beta = Mc;
            w = S / beta;
        } else {  // mod_vi.f90:225-231
            const double gk = Mu_km1 / beta;
            gam[k * BLOCK] = gk;
            beta = Mc - Ml * gk;
            w = (S - Ml * w_km1) / beta;
        }
        F.w[o] = w;  // forward value, replaced by the back substitution
        w_km1 = w;
        Mu_km1 = Mu;
        P_km1 = P_k;
        R_km1 = R_k;
    }
    // mod_vi.f90:234-239
    double wn = w_km1;  // w(kmax), not yet scaled
    for (int k = kmax - 1; k >= kmin + 1; --k) {
        const double wk = F.w[at(k)] - gam[(k + 1) * BLOCK] * wn;
        F.w[at(k + 1)] = wn * F.GSGAM2H[at(k + 1)];
        wn = wk;
    }
    // mod_vi.f90:242-246
    F.w[at(kmin)] = w_kmin * F.GSGAM2H[at(kmin)];
    F.w[at(kmin + 1)] = wn * F.GSGAM2H[at(kmin + 1)];
    F.w[at(kmax + 1)] = w_kmax1 * F.GSGAM2H[at(kmax + 1)];
}

__global__ void __launch_bounds__(BLOCK) vi_kernel(ViArgs A)
{
    extern __shared__ double s_gam[];
    double *gam = s_gam + threadIdx.x;
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long nreg = (long)A.gall * A.lall;
    if (t < nreg) {
        const int g = (int)(t % A.gall), l = (int)(t / A.gall);
        vi_column(A.reg, (size_t)l * A.kall * A.gall + g, A.gall, A.kall, A.rdgzh, A.afact, A.bfact, A.dt,
                  A.CVovRt2, A.alpha, A.grav, gam);
    } else if (A.have_pl && t < nreg + (long)A.gall_pl * kLallPl) {
        const int p = (int)(t - nreg);
        const int g = p % A.gall_pl, l = p / A.gall_pl;
        vi_column(A.pl, (size_t)l * A.kall * A.gall_pl + g, A.gall_pl, A.kall, A.rdgzh, A.afact, A.bfact,
                  A.dt, A.CVovRt2, A.alpha, A.grav, gam);
    }
}

}  // namespace

int launch_vi(nicam_ctx c, double *w, double *w_pl, const double *w0, const double *w0_pl,
              const double *p0, const double *p0_pl, const double *r0, const double *r0_pl,
              const double *Sr, const double *Sr_pl, const double *Sw, const double *Sw_pl,
              const double *Sp, const double *Sp_pl, double dt)
{
    const Dims &D = c->d;
    const size_t lds = sizeof(double) * BLOCK * D.kall;
    if (lds > 160 * 1024) return fail(NICAM_ERR_ARG, "kall too large for the LDS elimination factors");
    ViArgs A;
    A.reg = ViField{w, w0, p0, r0, Sr, Sw, Sp, c->Mc, c->Ml, c->Mu, c->RGAM, c->RGAMH, c->RGSGAM2,
                    c->RGSGAM2H, c->GSGAM2H};
    A.pl = ViField{w_pl, w0_pl, p0_pl, r0_pl, Sr_pl, Sw_pl, Sp_pl, c->Mc_pl, c->Ml_pl, c->Mu_pl,
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
    A.CVovRt2 = c->cvdry / c->rdry / (dt * dt);  // mod_vi.f90:178
    A.alpha = (double)c->alpha;                   // mod_vi.f90:179
    A.grav = c->grav;
    const long n = (long)D.gall * D.lall + (D.have_pl ? (long)D.gall_pl * kLallPl : 0);
    const long nblk = (n + BLOCK - 1) / BLOCK;
    if (lds > 64 * 1024) NICAM_HIP(hipFuncSetAttribute((const void *)vi_kernel,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(vi_kernel, dim3((unsigned)nblk), dim3(BLOCK), lds, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
