// vi_rhow_solver on gfx950.
//
// Reference: kernels/NICAM/dyn_vi_rhow_solver/src/mod_vi.f90:93-306.
//
// The reference builds the Helmholtz source Sall for a whole region, applies
// the boundary conditions, then runs a Thomas elimination per column, all with
// g innermost.  Here one work-item owns one column (g, l): the source term is
// formed level by level inside the forward sweep (the level k-1 products are
// carried in registers, so no Sall array exists), and a wavefront covers 64
// consecutive g at one level, so all global traffic is contiguous.  The pole
// columns are appended as the last work-items.
//
// The elimination factors gamma(k) = Mu(k-1)/beta(k-1) and
// beta(k) = Mc(k) - Ml(k)*gamma(k) depend on the coefficients only, not on the
// source term.  nicam_vi_rhow_solver_putcoef (and the on-device matrix build)
// therefore run the factorisation once, per column with the reference's
// operations in its order (vi_factor_kernel), and keep beta and gamma beside
// Mc/Ml/Mu.  A factored solve reads Ml, beta and gamma where the reference
// reads Mc, Ml and Mu (the same traffic), and its dependency chain per level
// holds one division instead of two.  Storage variants of the solve:
//   1  unfactored, gamma in LDS (one slot per lane and level), forward
//      solution in rhogw
//   2  factored, forward solution in rhogw (any kall)
//   3  unfactored, gamma and the forward solution in registers (fully
//      unrolled levels; built for kall = 42 and 98)
//   4  factored, forward solution in registers (kall = 42 and 98)
// All of them produce the same bits: the factors are formed by the same
// operations in the same order whether they are stored or not.
#include <type_traits>

#include "internal.h"

namespace nicam {
namespace {

// Src is the type of the prognostic and source inputs and of the metrics that
// only enter the source term (double, or float for the mixed solver: loads
// are widened exactly, every operation stays FP64).
template <class Src>
struct ViFieldT {
    double *w;
    const Src *w0, *p0, *r0, *Sr, *Sw, *Sp;
    const double *Mc, *Ml, *Mu;
    const Src *RGAM, *RGAMH, *RGSGAM2;
    const double *RGSGAM2H, *GSGAM2H;
    const double *beta, *gam;  // the stored factorisation (factored variants)
};

template <class Src>
struct ViArgsT {
    ViFieldT<Src> reg, pl;
    const double *rdgzh, *afact, *bfact;
    long ncol;  // columns including the pole columns
    int gall, kall, lall, gall_pl, have_pl;
    double dt, CVovRt2, alpha, grav;
};

// gamma(k) / forward-solution storage policies
struct LdsGamma {
    double *p;  // this lane's slot 0; stride = block size
    int stride;
    __device__ void set(int k, double v) { p[k * stride] = v; }
    __device__ double get(int k) const { return p[k * stride]; }
};
struct NoGamma {  // factored variants: gamma is read from the stored factorisation
    __device__ void set(int, double) {}
};
template <int N>
struct RegArray {
    double v[N];
    __device__ void set(int k, double x) { v[k] = x; }
    __device__ double get(int k) const { return v[k]; }
};

// One column.  KC > 0 fixes kall at compile time (loops fully unrolled, W
// may then hold the forward solution in registers); KC == 0 reads kall at run
// time and keeps the forward solution in rhogw.  FACT reads beta and gamma
// from the stored factorisation instead of forming them.
template <int KC, bool WREG, bool FACT, class G, class WS, class Src>
__device__ __forceinline__ void vi_column(const ViFieldT<Src> &F, size_t base, size_t s, int kall_rt,
                                          const double *__restrict__ rdgzh, const double *__restrict__ afact,
                                          const double *__restrict__ bfact, double dt, double CVovRt2,
                                          double alpha, double grav, G &gam, WS &wf)
{
    const int kall = KC > 0 ? KC : kall_rt;
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
    // the inputs of one level; the next level's are loaded before this
    // level's stores are issued, so the loads overlap the elimination
    struct Lvl {
        double rgh, rg, p0, Sp, rgs2, r0, Sr, w0, Sw, Mc, Ml, Mu;  // Mc holds beta when FACT
    };
    auto load = [&](int k) {
        const size_t o = at(k);
        Lvl v;
        v.rgh = F.RGAMH[o];
        v.rg = F.RGAM[o];
        v.p0 = F.p0[o];
        v.Sp = F.Sp[o];
        v.rgs2 = F.RGSGAM2[o];
        v.r0 = F.r0[o];
        v.Sr = F.Sr[o];
        v.w0 = F.w0[o];
        v.Sw = F.Sw[o];
        if constexpr (FACT) {
            v.Mc = F.beta[o];
            v.Ml = F.Ml[o];
            v.Mu = 0.0;
        } else {
            v.Mc = F.Mc[o];
            v.Ml = F.Ml[o];
            v.Mu = F.Mu[o];
        }
        return v;
    };
    double beta = 0.0, w_km1 = 0.0, Mu_km1 = 0.0;
    constexpr int UNROLL = KC > 0 ? KC : 1;  // full unroll only at compile-time kall
    Lvl cur = load(kmin + 1);
#pragma unroll UNROLL
    for (int k = kmin + 1; k <= kmax; ++k) {
        const size_t o = at(k);
        Lvl nxt = cur;
        if (k < kmax) nxt = load(k + 1);
        const double P_k = (cur.p0 + dt * cur.Sp) * cur.rgs2;
        const double R_k = (cur.r0 + dt * cur.Sr) * (cur.rg * cur.rg);
        double S = ((cur.w0 * alpha + dt * cur.Sw) * (cur.rgh * cur.rgh) - (P_k - P_km1) * dt * rdgzh[k] -
                    (R_k * afact[k] + R_km1 * bfact[k]) * dt * grav) *
                   CVovRt2;
        if (k == kmin + 1) S = S - cur.Ml * w_kmin;  // mod_vi.f90:211
        if (k == kmax) S = S - (FACT ? F.Mu[o] : cur.Mu) * w_kmax1;  // mod_vi.f90:212
        double w;
        if (k == kmin + 1) {  // mod_vi.f90:218-222
            beta = cur.Mc;
            w = S / beta;
        } else if constexpr (FACT) {  // mod_vi.f90:225-231 with beta(k) stored
            w = (S - cur.Ml * w_km1) / cur.Mc;
        } else {  // mod_vi.f90:225-231
            const double gk = Mu_km1 / beta;
            gam.set(k, gk);
            beta = cur.Mc - cur.Ml * gk;
            w = (S - cur.Ml * w_km1) / beta;
        }
        if constexpr (WREG) wf.set(k, w);
        else F.w[o] = w;  // forward value, replaced by the back substitution
        w_km1 = w;
        Mu_km1 = cur.Mu;
        P_km1 = P_k;
        R_km1 = R_k;
        cur = nxt;
        // with deep columns in registers, keep the scheduler from hoisting
        // later levels' loads (it would run out of registers)
        if constexpr (KC > 64) __builtin_amdgcn_sched_barrier(0);
    }
    // mod_vi.f90:234-239
    double wn = w_km1;  // w(kmax), not yet scaled
    auto load_b = [&](int k, double &wf_k, double &g_k1, double &gs_k1) {
        if constexpr (WREG) wf_k = wf.get(k);
        else wf_k = F.w[at(k)];
        if constexpr (FACT) g_k1 = F.gam[at(k + 1)];
        else g_k1 = gam.get(k + 1);
        gs_k1 = F.GSGAM2H[at(k + 1)];
    };
    double wf_c = 0.0, g_c = 0.0, gs_c = 0.0;
    if (kmax - 1 >= kmin + 1) load_b(kmax - 1, wf_c, g_c, gs_c);
#pragma unroll UNROLL
    for (int k = kmax - 1; k >= kmin + 1; --k) {
        double wf_n = 0.0, g_n = 0.0, gs_n = 0.0;
        if (k - 1 >= kmin + 1) load_b(k - 1, wf_n, g_n, gs_n);
        const double wk = wf_c - g_c * wn;
        stream_store(&F.w[at(k + 1)], wn * gs_c);
        wn = wk;
        wf_c = wf_n;
        g_c = g_n;
        gs_c = gs_n;
        if constexpr (KC > 64) __builtin_amdgcn_sched_barrier(0);
    }
    // mod_vi.f90:242-246
    F.w[at(kmin)] = w_kmin * F.GSGAM2H[at(kmin)];
    F.w[at(kmin + 1)] = wn * F.GSGAM2H[at(kmin + 1)];
    F.w[at(kmax + 1)] = w_kmax1 * F.GSGAM2H[at(kmax + 1)];
}

constexpr int BLOCK_LDS = 64;
constexpr int BLOCK_FAC = 256;
constexpr int BLOCK_REG = 64;

template <class Src>
__global__ void __launch_bounds__(BLOCK_LDS) vi_kernel_lds(ViArgsT<Src> A)
{
    extern __shared__ double s_gam[];
    column_run<BLOCK_LDS>(A, xcd_contiguous(blockIdx.x, gridDim.x), [&](const ViFieldT<Src> &F, size_t base, size_t s) {
        LdsGamma gam{s_gam + threadIdx.x, BLOCK_LDS};
        int unused = 0;
        vi_column<0, false, false>(F, base, s, A.kall, A.rdgzh, A.afact, A.bfact, A.dt, A.CVovRt2, A.alpha, A.grav,
                                   gam, unused);
    });
}

// factored, any kall: the forward solution goes through rhogw
template <class Src>
__global__ void __launch_bounds__(BLOCK_FAC) vi_kernel_fact(ViArgsT<Src> A)
{
    column_run<BLOCK_FAC>(A, xcd_contiguous(blockIdx.x, gridDim.x), [&](const ViFieldT<Src> &F, size_t base, size_t s) {
        NoGamma gam;
        int unused = 0;
        vi_column<0, false, true>(F, base, s, A.kall, A.rdgzh, A.afact, A.bfact, A.dt, A.CVovRt2, A.alpha, A.grav,
                                  gam, unused);
    });
}

// several grids in one launch (factored, any kall): work-groups
// [off[n], off[n+1]) belong to grid n
template <class Src>
struct ViGroup {
    ViArgsT<Src> a[kMaxGroup];
    int off[kMaxGroup + 1];
    int n;
};

template <class Src>
__global__ void __launch_bounds__(BLOCK_FAC) vi_group_kernel(ViGroup<Src> G)
{
    const int b = blockIdx.x;
    int n = 0;
    while (n + 1 < G.n && b >= G.off[n + 1]) ++n;
    const ViArgsT<Src> &A = G.a[n];
    column_run<BLOCK_FAC>(A, b - G.off[n], [&](const ViFieldT<Src> &F, size_t base, size_t s) {
        NoGamma gam;
        int unused = 0;
        vi_column<0, false, true>(F, base, s, A.kall, A.rdgzh, A.afact, A.bfact, A.dt, A.CVovRt2, A.alpha, A.grav,
                                  gam, unused);
    });
}

template <int KC, bool FACT, class Src>
__global__ void __launch_bounds__(BLOCK_REG) vi_kernel_reg(ViArgsT<Src> A)
{
    auto column = [&](const ViFieldT<Src> &F, size_t base, size_t s) {
        RegArray<KC> wf;
        if constexpr (FACT) {
            NoGamma gam;
            vi_column<KC, true, true>(F, base, s, KC, A.rdgzh, A.afact, A.bfact, A.dt, A.CVovRt2, A.alpha, A.grav,
                                      gam, wf);
        } else {
            RegArray<KC> gam;
            vi_column<KC, true, false>(F, base, s, KC, A.rdgzh, A.afact, A.bfact, A.dt, A.CVovRt2, A.alpha, A.grav,
                                       gam, wf);
        }
    };
    if constexpr (std::is_same<Src, double>::value) {
        // FP64 sources: the field set chosen per work-item, its pointers in vector registers, schedules
        // better here (gl7, variant 4: 368 us against 401 us with the work-group's)
        // (one copy of the column, on a per-work-item pointer)
        const long t = xcd_contiguous(blockIdx.x, gridDim.x) * BLOCK_REG + threadIdx.x;
        const long nreg = (long)A.gall * A.lall;
        const ViFieldT<Src> *F;
        size_t base, s;
        if (t < nreg) {
            F = &A.reg;
            base = (size_t)(t / A.gall) * A.kall * A.gall + (size_t)(t % A.gall);
            s = A.gall;
        } else {
            // the grid is column_blocks' (the pole columns from a work-group of their own)
            const long p = t - ((nreg + BLOCK_REG - 1) / BLOCK_REG) * BLOCK_REG;
            if (!A.have_pl || p < 0 || p >= (long)A.gall_pl * kLallPl) return;
            F = &A.pl;
            base = (size_t)(p / A.gall_pl) * A.kall * A.gall_pl + (size_t)(p % A.gall_pl);
            s = A.gall_pl;
        }
        column(*F, base, s);
    } else {
        column_run<BLOCK_REG>(A, xcd_contiguous(blockIdx.x, gridDim.x), column);  // FP32 sources: 309 us against 550 us per work-item
    }
}

// The Thomas factorisation of one column (mod_vi.f90:218-231 without the
// right-hand side): beta(kmin+1) = Mc(kmin+1); gamma(k) = Mu(k-1)/beta(k-1),
// beta(k) = Mc(k) - Ml(k)*gamma(k) for k = kmin+2..kmax.  Other levels are
// left as they are; no solve reads them.
struct FactorPart {
    const double *Mc, *Ml, *Mu;
    double *beta, *gam;    // the system's factors
    double *tbeta, *tgam;  // the transposed system's (the adjoint solve)
    int gall, ncols;       // level stride and columns of this part
};
struct FactorArgs {
    FactorPart reg, pl;
    int kall;
};

__global__ void __launch_bounds__(BLOCK_FAC) vi_factor_kernel(FactorArgs A)
{
    long t = (long)blockIdx.x * BLOCK_FAC + threadIdx.x;
    const FactorPart *P = &A.reg;
    if (t >= A.reg.ncols) {
        t -= A.reg.ncols;
        P = &A.pl;
        if (t >= A.pl.ncols) return;
    }
    const int kall = A.kall, kmin = 1, kmax = kall - 2;
    const size_t base = (size_t)(t / P->gall) * kall * P->gall + (size_t)(t % P->gall), s = P->gall;
    double beta = P->Mc[base + (kmin + 1) * s];
    P->beta[base + (kmin + 1) * s] = beta;
    for (int k = kmin + 2; k <= kmax; ++k) {
        const size_t o = base + k * s;
        const double g = P->Mu[o - s] / beta;
        beta = P->Mc[o] - P->Ml[o] * g;
        P->gam[o] = g;
        P->beta[o] = beta;
    }
    // the transposed system (diagonal Mc(k), sub-diagonal Mu(k-1), super-diagonal
    // Ml(k+1)): g(k) = Ml(k)/beta(k-1), beta(k) = Mc(k) - Mu(k-1) g(k)
    beta = P->Mc[base + (kmin + 1) * s];
    P->tbeta[base + (kmin + 1) * s] = beta;
    for (int k = kmin + 2; k <= kmax; ++k) {
        const size_t o = base + k * s;
        const double g = P->Ml[o] / beta;
        beta = P->Mc[o] - P->Mu[o - s] * g;
        P->tgam[o] = g;
        P->tbeta[o] = beta;
    }
}

template <class Src>
void vi_make_args(nicam_ctx c, ViArgsT<Src> &A, double *w, double *w_pl, const Src *w0, const Src *w0_pl,
                  const Src *p0, const Src *p0_pl, const Src *r0, const Src *r0_pl, const Src *Sr, const Src *Sr_pl,
                  const Src *Sw, const Src *Sw_pl, const Src *Sp, const Src *Sp_pl, double dt, const Src *RGAM,
                  const Src *RGAM_pl, const Src *RGAMH, const Src *RGAMH_pl, const Src *RGSGAM2,
                  const Src *RGSGAM2_pl)
{
    const Dims &D = c->d;
    A.reg = ViFieldT<Src>{w, w0, p0, r0, Sr, Sw, Sp, c->Mc, c->Ml, c->Mu, RGAM, RGAMH, RGSGAM2, c->RGSGAM2H,
                          c->GSGAM2H, c->vi_beta, c->vi_gam};
    A.pl = ViFieldT<Src>{w_pl, w0_pl, p0_pl, r0_pl, Sr_pl, Sw_pl, Sp_pl, c->Mc_pl, c->Ml_pl, c->Mu_pl, RGAM_pl,
                         RGAMH_pl, RGSGAM2_pl, c->RGSGAM2H_pl, c->GSGAM2H_pl, c->vi_beta_pl, c->vi_gam_pl};
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
    A.ncol = (long)D.gall * D.lall + (D.have_pl ? (long)D.gall_pl * kLallPl : 0);
}

template <class Src>
int launch_vi_t(nicam_ctx c, double *w, double *w_pl, const Src *w0, const Src *w0_pl, const Src *p0, const Src *p0_pl,
                const Src *r0, const Src *r0_pl, const Src *Sr, const Src *Sr_pl, const Src *Sw, const Src *Sw_pl, const Src *Sp,
                const Src *Sp_pl, double dt, const Src *RGAM, const Src *RGAM_pl, const Src *RGAMH, const Src *RGAMH_pl,
                const Src *RGSGAM2, const Src *RGSGAM2_pl)
{
    const Dims &D = c->d;
    ViArgsT<Src> A;
    vi_make_args(c, A, w, w_pl, w0, w0_pl, p0, p0_pl, r0, r0_pl, Sr, Sr_pl, Sw, Sw_pl, Sp, Sp_pl, dt, RGAM, RGAM_pl,
                 RGAMH, RGAMH_pl, RGSGAM2, RGSGAM2_pl);

    int variant = c->tune.vi_variant;
    const bool reg_ok = (D.kall == 42 || D.kall == 98);
    if ((variant == 3 || variant == 4) && !reg_ok) variant = 0;
    // measured at gl7 (40 x 66x66 x 98): FP64 inputs: 4 (factored, registers) 368 us, 3 379, 2 469, 1 623;
    // FP32 inputs: 4 309 us, 2 (factored, through rhogw) 324, 3 759, 1 633
    if (variant == 0) variant = reg_ok ? 4 : 2;

    if (variant == 1) {
        const size_t lds = sizeof(double) * BLOCK_LDS * D.kall;
        if (lds > 160 * 1024) return fail(NICAM_ERR_ARG, "kall too large for the LDS elimination factors");
        if (lds > 64 * 1024)
            NICAM_HIP(hipFuncSetAttribute((const void *)vi_kernel_lds<Src>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds));
        hipLaunchKernelGGL(vi_kernel_lds<Src>, dim3((unsigned)column_blocks<BLOCK_LDS>(D.gall, D.lall, D.gall_pl, D.have_pl)),
                           dim3(BLOCK_LDS), lds, c->stream, A);
    } else if (variant == 2) {
        hipLaunchKernelGGL(vi_kernel_fact<Src>, dim3((unsigned)column_blocks<BLOCK_FAC>(D.gall, D.lall, D.gall_pl, D.have_pl)),
                           dim3(BLOCK_FAC), 0, c->stream, A);
    } else {
        const dim3 grid((unsigned)column_blocks<BLOCK_REG>(D.gall, D.lall, D.gall_pl, D.have_pl));
        // variant 3 holds gamma and the forward solution of a whole column in
        // registers (252 VGPRs at kall = 42; 256 VGPRs + 256 AGPRs, one
        // wavefront per SIMD, at kall = 98); variant 4 only the forward solution
        if (variant == 3) {
            if (D.kall == 42) hipLaunchKernelGGL((vi_kernel_reg<42, false, Src>), grid, dim3(BLOCK_REG), 0, c->stream, A);
            else hipLaunchKernelGGL((vi_kernel_reg<98, false, Src>), grid, dim3(BLOCK_REG), 0, c->stream, A);
        } else {
            if (D.kall == 42) hipLaunchKernelGGL((vi_kernel_reg<42, true, Src>), grid, dim3(BLOCK_REG), 0, c->stream, A);
            else hipLaunchKernelGGL((vi_kernel_reg<98, true, Src>), grid, dim3(BLOCK_REG), 0, c->stream, A);
        }
    }
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace

int launch_vi_factor(nicam_ctx c)
{
    const Dims &D = c->d;
    FactorArgs A;
    A.reg = FactorPart{c->Mc, c->Ml, c->Mu, c->vi_beta, c->vi_gam, c->vi_tbeta, c->vi_tgam, D.gall, D.gall * D.lall};
    A.pl = FactorPart{c->Mc_pl, c->Ml_pl, c->Mu_pl, c->vi_beta_pl, c->vi_gam_pl, c->vi_tbeta_pl, c->vi_tgam_pl,
                      D.gall_pl, D.have_pl ? D.gall_pl * kLallPl : 0};
    A.kall = D.kall;
    const long n = (long)A.reg.ncols + A.pl.ncols;
    hipLaunchKernelGGL(vi_factor_kernel, dim3((unsigned)((n + BLOCK_FAC - 1) / BLOCK_FAC)), dim3(BLOCK_FAC), 0,
                       c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_vi(nicam_ctx c, double *w, double *w_pl, const double *w0, const double *w0_pl,
              const double *p0, const double *p0_pl, const double *r0, const double *r0_pl,
              const double *Sr, const double *Sr_pl, const double *Sw, const double *Sw_pl,
              const double *Sp, const double *Sp_pl, double dt)
{
    return launch_vi_t<double>(c, w, w_pl, w0, w0_pl, p0, p0_pl, r0, r0_pl, Sr, Sr_pl, Sw, Sw_pl, Sp, Sp_pl, dt,
                               c->RGAM, c->RGAM_pl, c->RGAMH, c->RGAMH_pl, c->RGSGAM2, c->RGSGAM2_pl);
}

int launch_vi_mixed(nicam_ctx c, double *w, double *w_pl, const float *w0, const float *w0_pl, const float *p0,
                    const float *p0_pl, const float *r0, const float *r0_pl, const float *Sr, const float *Sr_pl,
                    const float *Sw, const float *Sw_pl, const float *Sp, const float *Sp_pl, double dt)
{
    return launch_vi_t<float>(c, w, w_pl, w0, w0_pl, p0, p0_pl, r0, r0_pl, Sr, Sr_pl, Sw, Sw_pl, Sp, Sp_pl, dt,
                              c->RGAM32, c->RGAM32_pl, c->RGAMH32, c->RGAMH32_pl, c->RGSGAM2_32, c->RGSGAM2_32_pl);
}

int launch_vi_group_mixed(const nicam_ctx *cs, int n, const nicam_tracer_step_f32 *g, double dt, hipStream_t stream)
{
    if (n < 1 || n > kMaxGroup) return fail(NICAM_ERR_ARG, "a group holds 1..8 grids");
    ViGroup<float> G;
    G.n = n;
    G.off[0] = 0;
    for (int i = 0; i < n; ++i) {
        const nicam_ctx c = cs[i];
        const nicam_tracer_step_f32 &p = g[i];
        vi_make_args<float>(c, G.a[i], p.rhogw, p.rhogw_pl, p.rhogw0, p.rhogw0_pl, p.preg0, p.preg0_pl, p.rhog0,
                            p.rhog0_pl, p.Srho, p.Srho_pl, p.Sw, p.Sw_pl, p.Spre, p.Spre_pl, dt, c->RGAM32,
                            c->RGAM32_pl, c->RGAMH32, c->RGAMH32_pl, c->RGSGAM2_32, c->RGSGAM2_32_pl);
        const long nb = column_blocks<BLOCK_FAC>(G.a[i].gall, G.a[i].lall, G.a[i].gall_pl, G.a[i].have_pl);
        if ((long)G.off[i] + nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "group too large for one launch");
        G.off[i + 1] = G.off[i] + (int)nb;
    }
    hipLaunchKernelGGL(vi_group_kernel<float>, dim3((unsigned)G.off[n]), dim3(BLOCK_FAC), 0, stream, G);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
