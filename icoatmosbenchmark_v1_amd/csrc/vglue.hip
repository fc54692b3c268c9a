// The vertical half of the step glue, on gfx950.
//
// The reference's vertical limiter takes the Courant numbers ck and the face
// values q_h as inputs and has no code for them.  These kernels build them
// and the tracer update they feed, in the NICAM forms of the Miura scheme in
// the vertical (flux form on the Lorenz grid, docs/NICAM/src/20_intro_nicam.tex:440-470):
//   mass flux  flx_v(k) = rhogw_vm(k) * dt on the interior half levels
//              kmin+1..kmax and 0 on the two boundary faces, rhogw_vm being
//              the contravariant vertical mass flux the reference forms in
//              OPRT3D_divdamp (dyn_divdamp/src/mod_oprt3d.f90:148-155);
//   courant    ck(k,1) = -flx_v(k) / rhog(k) * rdgz(k) (the lower face),
//              ck(k,2) =  flx_v(k+1) / rhog(k) * rdgz(k) (the upper face), on
//              kmin..kmax, 0 elsewhere: negative on an inflow face, as the
//              limiter's flags read them (dyn_vert_adv_limiter/src/mod_src_tracer.f90:113-114);
//              the closed faces kmin and kmax+1 carry signed zeros that
//              those flags read as outflow;
//   face       q_h(k) = afact(k) q(k) + bfact(k) q(k-1) on kmin+1..kmax, the
//              half-level interpolation of the vertical grid;
//   update     rhogq(k) -= (flx_v(k+1) q_h(k+1) - flx_v(k) q_h(k)) rdgz(k) on
//              kmin..kmax, the boundary faces carrying nothing.
// Arrays are (gall,kall,lall) with the poles (gall_pl,kall,2); one work-item
// per cell-level, its column neighbours one level plane away.
#include "internal.h"

namespace nicam {
namespace {

constexpr int BLK = 256;

// One of the two parts of a launch: the regular regions or the poles.
struct VPart {
    double *out;                   // flx_v, ck, q_h or rhogq
    double *out2;                  // the adjoints' second output
    const double *in[4];           // per operation, see the kernels
    const double *c2w, *rgamh, *rgsqrth;  // vertical metrics (mass flux only)
    int gall, lall;
    long n;  // cell-levels of this part
};

struct VCommon {
    int kall, kmin, kmax;  // 0-based kmin = 1, kmax = kall-2
    const double *rdgz, *afact, *bfact;
    double dt;
};

enum { V_FLUX = 0, V_COURANT = 1, V_FACE = 2, V_UPDATE = 3, V_COURANT_ADJ = 4, V_FACE_ADJ = 5, V_UPDATE_ADJ = 6 };

// Work-groups [0, nR) cover the regular regions, the rest the poles.  The
// levels above and below a cell are gall cells away, a few work-groups on:
// the regular part keeps consecutive work-groups on one XCD.
template <int OP>
__global__ void __launch_bounds__(BLK) vglue_kernel(VPart R, VPart P, VCommon C)
{
    const long nR = (R.n + BLK - 1) / BLK;
    const bool pole = (long)blockIdx.x >= nR;
    const long t = pole ? ((long)blockIdx.x - nR) * BLK + threadIdx.x
                        : xcd_contiguous(blockIdx.x, nR) * BLK + threadIdx.x;
    const VPart &A = pole ? P : R;
    if (t >= A.n) return;
    const int gall = A.gall, kall = C.kall;
    const int g = (int)(t % gall);
    const long kl = t / gall;
    const int k = (int)(kl % kall), l = (int)(kl / kall);
    const long dn = gall;  // one level up or down
    if (OP == V_FLUX) {
        // in: rhogvx, rhogvy, rhogvz, rhogw
        double f = 0.0;
        if (k >= C.kmin + 1 && k <= C.kmax) {
            const size_t cstr = (size_t)kall * gall, cb = (size_t)l * 6 * cstr + (size_t)k * gall + g;
            const double vm = (A.c2w[0 * cstr + cb] * A.in[0][t] + A.c2w[1 * cstr + cb] * A.in[0][t - dn] +
                               A.c2w[2 * cstr + cb] * A.in[1][t] + A.c2w[3 * cstr + cb] * A.in[1][t - dn] +
                               A.c2w[4 * cstr + cb] * A.in[2][t] + A.c2w[5 * cstr + cb] * A.in[2][t - dn]) *
                                  A.rgamh[t] +
                              A.in[3][t] * A.rgsqrth[t];
            f = vm * C.dt;
        }
        stream_store(&A.out[t], f);
    } else if (OP == V_COURANT) {
        // in: flx_v, rhog; out: ck(...,2), the two planes A.n apart.  The
        // closed faces get zeros whose signs the limiter reads as outflow:
        // +0 in ck(kmin,1), -0 in ck(kmax+1,1) (its flags are
        // 0.5 -/+ sign(0.5, ck(.,1))), so no ghost level enters its bounds.
        double lo = 0.0, up = 0.0;
        if (k >= C.kmin && k <= C.kmax) {
            const double r = A.in[1][t];
            if (k > C.kmin) lo = -A.in[0][t] / r * C.rdgz[k];
            if (k < C.kmax) up = A.in[0][t + dn] / r * C.rdgz[k];
        } else if (k == C.kmax + 1) {
            lo = -0.0;
        }
        stream_store(&A.out[t], lo);
        stream_store(&A.out[A.n + t], up);
    } else if (OP == V_FACE) {
        // in: q; q_h kept outside kmin+1..kmax
        if (k >= C.kmin + 1 && k <= C.kmax)
            stream_store(&A.out[t], C.afact[k] * A.in[0][t] + C.bfact[k] * A.in[0][t - dn]);
    } else if (OP == V_UPDATE) {
        // in: flx_v, q_h; rhogq updated on kmin..kmax
        if (k >= C.kmin && k <= C.kmax) {
            const double up = (k + 1 <= C.kmax) ? A.in[0][t + dn] * A.in[1][t + dn] : 0.0;
            const double lo = (k >= C.kmin + 1) ? A.in[0][t] * A.in[1][t] : 0.0;
            A.out[t] = A.out[t] - (up - lo) * C.rdgz[k];
        }
    } else if (OP == V_COURANT_ADJ) {
        // in: ck_b (two planes A.n apart), flx_v, rhog; out: flx_v_b, rhog_b.  flx_v(k) enters the
        // lower face of k and the upper face of k-1, rhog(k) both faces of k (-ck/rhog each)
        const double *cb = A.in[0], *fv = A.in[1], *rh = A.in[2];
        const double r = rh[t];
        double f = 0.0, s = 0.0;
        if (k > C.kmin && k <= C.kmax) f = f + -cb[t] / r * C.rdgz[k];
        if (k - 1 >= C.kmin && k - 1 < C.kmax) f = f + cb[A.n + t - dn] / rh[t - dn] * C.rdgz[k - 1];
        if (k > C.kmin && k <= C.kmax) s = s + cb[t] * (-fv[t] / r * C.rdgz[k]);
        if (k >= C.kmin && k < C.kmax) s = s + cb[A.n + t] * (fv[t + dn] / r * C.rdgz[k]);
        A.out[t] = f;
        A.out2[t] = -s / r;
    } else if (OP == V_FACE_ADJ) {
        // in: q_h_b; out: q_b (q(k) enters q_h(k) and q_h(k+1))
        const double *hb = A.in[0];
        double s = 0.0;
        if (k >= C.kmin + 1 && k <= C.kmax) s = s + C.afact[k] * hb[t];
        if (k + 1 >= C.kmin + 1 && k + 1 <= C.kmax) s = s + C.bfact[k + 1] * hb[t + dn];
        A.out[t] = s;
    } else {
        // V_UPDATE_ADJ, in: rhogq_b, flx_v, q_h; out: flx_v_b, q_h_b.  The flux through the interior face
        // k leaves level k-1 and enters level k.
        double a = 0.0, b = 0.0;
        if (k >= C.kmin + 1 && k <= C.kmax) {
            const double Fb = A.in[0][t] * C.rdgz[k] - A.in[0][t - dn] * C.rdgz[k - 1];
            a = Fb * A.in[2][t];
            b = Fb * A.in[1][t];
        }
        A.out[t] = a;
        A.out2[t] = b;
    }
}

VPart part(int gall, int lall, int kall, bool on, double *out, std::initializer_list<const double *> in)
{
    VPart p{};
    p.out = out;
    int i = 0;
    for (const double *x : in) p.in[i++] = x;
    p.gall = gall;
    p.lall = lall;
    p.n = on ? (long)gall * kall * lall : 0;
    return p;
}

template <int OP>
int launch_v(nicam_ctx c, const VPart &R, const VPart &P, double dt)
{
    const Dims &D = c->d;
    VCommon C;
    C.kall = D.kall;
    C.kmin = 1;
    C.kmax = D.kall - 2;
    C.rdgz = c->rdgz;
    C.afact = c->afact;
    C.bfact = c->bfact;
    C.dt = dt;
    const long nb = (R.n + BLK - 1) / BLK + (P.n + BLK - 1) / BLK;
    if (nb == 0) return NICAM_OK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(vglue_kernel<OP>, dim3((unsigned)nb), dim3(BLK), 0, c->stream, R, P, C);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

// Coefficients of the tridiagonal system for rho*w (the vertically implicit
// Helmholtz equation, docs/NICAM/src/20_intro_nicam.tex:480-600), collected
// from its discretised form at the half level k between full levels k-1 and
// k (0-based storage, the row Ml(k) W(k-1) + Mc(k) W(k) + Mu(k) W(k+1) of
// vi_rhow_solver, dyn_vi_rhow_solver/src/mod_vi.f90:218-232):
//   A^o = RGSGAM2 (full), A^i = gam2 h (half), B = g_tilde (half),
//   C^o = (Cv/R) g / gam2 (half k), C^i = gam2 (half), D = Cv G^1/2 / (R dt^2) (half k)
// with gam2 = 1/RGAMH^2 and G^1/2 = GSGAM2H / gam2;
//   Mc = alpha D + (A^o(k) A^i(k) rdgz(k) + A^o(k-1) A^i(k) rdgz(k-1)) rdgzh(k)
//        - (dfact(k) - cfact(k-1)) (B(k) + C^o C^i(k)) rdgzh(k)
//   Mu = -A^o(k) A^i(k+1) rdgz(k) rdgzh(k) - cfact(k) (B(k+1) + C^o C^i(k+1)) rdgzh(k)
//   Ml = -A^o(k-1) A^i(k-1) rdgz(k-1) rdgzh(k) + dfact(k-1) (B(k-1) + C^o C^i(k-1)) rdgzh(k)
// on kmin+1..kmax, 0 on the other levels.
struct MatPart {
    double *Mc, *Ml, *Mu;
    const double *eth, *gt, *rgsgam2, *rgamh, *gsgam2h;
    int gall;
    long n;
};

struct MatCommon {
    int kall, kmin, kmax;
    const double *rdgz, *rdgzh, *cfact, *dfact;
    double cvovr, cvovrt2, grav, alpha;
};

__global__ void __launch_bounds__(BLK) vi_matrix_kernel(MatPart R, MatPart P, MatCommon C)
{
    const long nR = (R.n + BLK - 1) / BLK;
    const bool pole = (long)blockIdx.x >= nR;
    const long t = pole ? ((long)blockIdx.x - nR) * BLK + threadIdx.x
                        : xcd_contiguous(blockIdx.x, nR) * BLK + threadIdx.x;
    const MatPart &A = pole ? P : R;
    if (t >= A.n) return;
    const int k = (int)((t / A.gall) % C.kall);
    const long dn = A.gall;
    double mc = 0.0, ml = 0.0, mu = 0.0;
    if (k >= C.kmin + 1 && k <= C.kmax) {
        const double rh = A.rgamh[t];
        const double g2m = 1.0 / (A.rgamh[t - dn] * A.rgamh[t - dn]), g2 = 1.0 / (rh * rh),
                     g2p = 1.0 / (A.rgamh[t + dn] * A.rgamh[t + dn]);
        const double aim = g2m * A.eth[t - dn], ai = g2 * A.eth[t], aip = g2p * A.eth[t + dn];
        const double co = C.cvovr * C.grav * rh * rh;
        const double d = C.cvovrt2 * A.gsgam2h[t] * rh * rh;
        const double ao = A.rgsgam2[t], aom = A.rgsgam2[t - dn];
        const double rzh = C.rdgzh[k];
        mc = C.alpha * d + (ao * ai * C.rdgz[k] + aom * ai * C.rdgz[k - 1]) * rzh -
             (C.dfact[k] - C.cfact[k - 1]) * (A.gt[t] + co * g2) * rzh;
        mu = -ao * aip * C.rdgz[k] * rzh - C.cfact[k] * (A.gt[t + dn] + co * g2p) * rzh;
        ml = -aom * aim * C.rdgz[k - 1] * rzh + C.dfact[k - 1] * (A.gt[t - dn] + co * g2m) * rzh;
    }
    A.Mc[t] = mc;
    A.Ml[t] = ml;
    A.Mu[t] = mu;
}

}  // namespace

int launch_vi_matrix(nicam_ctx c, const double *eth, const double *eth_pl, const double *g_tilde,
                     const double *g_tilde_pl, double dt)
{
    const Dims &D = c->d;
    MatPart R{c->Mc, c->Ml, c->Mu, eth, g_tilde, c->RGSGAM2, c->RGAMH, c->GSGAM2H, D.gall,
              (long)D.gall * D.kall * D.lall};
    MatPart P{c->Mc_pl, c->Ml_pl, c->Mu_pl, eth_pl, g_tilde_pl, c->RGSGAM2_pl, c->RGAMH_pl, c->GSGAM2H_pl,
              D.gall_pl, D.have_pl ? (long)D.gall_pl * D.kall * kLallPl : 0};
    MatCommon C;
    C.kall = D.kall;
    C.kmin = 1;
    C.kmax = D.kall - 2;
    C.rdgz = c->rdgz;
    C.rdgzh = c->rdgzh;
    C.cfact = c->cfact;
    C.dfact = c->dfact;
    C.cvovr = c->cvdry / c->rdry;
    C.cvovrt2 = c->cvdry / c->rdry / (dt * dt);
    C.grav = c->grav;
    C.alpha = (double)c->alpha;
    const long nb = (R.n + BLK - 1) / BLK + (P.n + BLK - 1) / BLK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(vi_matrix_kernel, dim3((unsigned)nb), dim3(BLK), 0, c->stream, R, P, C);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_vertical_mass_flux(nicam_ctx c, double *flx_v, double *flx_v_pl, const double *vx, const double *vx_pl,
                              const double *vy, const double *vy_pl, const double *vz, const double *vz_pl,
                              const double *w, const double *w_pl, double dt)
{
    const Dims &D = c->d;
    VPart R = part(D.gall, D.lall, D.kall, true, flx_v, {vx, vy, vz, w});
    R.c2w = c->C2WfactGz;
    R.rgamh = c->RGAMH;
    R.rgsqrth = c->RGSQRTH;
    VPart P = part(D.gall_pl, kLallPl, D.kall, D.have_pl && flx_v_pl, flx_v_pl, {vx_pl, vy_pl, vz_pl, w_pl});
    P.c2w = c->C2WfactGz_pl;
    P.rgamh = c->RGAMH_pl;
    P.rgsqrth = c->RGSQRTH_pl;
    return launch_v<V_FLUX>(c, R, P, dt);
}

int launch_vertical_courant(nicam_ctx c, double *ck, double *ck_pl, const double *flx_v, const double *flx_v_pl,
                            const double *rhog, const double *rhog_pl)
{
    const Dims &D = c->d;
    const VPart R = part(D.gall, D.lall, D.kall, true, ck, {flx_v, rhog});
    const VPart P = part(D.gall_pl, kLallPl, D.kall, D.have_pl && ck_pl, ck_pl, {flx_v_pl, rhog_pl});
    return launch_v<V_COURANT>(c, R, P, 0.0);
}

int launch_vertical_face(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl)
{
    const Dims &D = c->d;
    const VPart R = part(D.gall, D.lall, D.kall, true, q_h, {q});
    const VPart P = part(D.gall_pl, kLallPl, D.kall, D.have_pl && q_h_pl, q_h_pl, {q_pl});
    return launch_v<V_FACE>(c, R, P, 0.0);
}

int launch_vertical_update(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_v, const double *flx_v_pl,
                           const double *q_h, const double *q_h_pl)
{
    const Dims &D = c->d;
    const VPart R = part(D.gall, D.lall, D.kall, true, rhogq, {flx_v, q_h});
    const VPart P = part(D.gall_pl, kLallPl, D.kall, D.have_pl && rhogq_pl, rhogq_pl, {flx_v_pl, q_h_pl});
    return launch_v<V_UPDATE>(c, R, P, 0.0);
}

int launch_vertical_courant_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *rhog_b, double *rhog_b_pl,
                                    const double *ck_b, const double *ck_b_pl, const double *flx_v,
                                    const double *flx_v_pl, const double *rhog, const double *rhog_pl)
{
    const Dims &D = c->d;
    VPart R = part(D.gall, D.lall, D.kall, true, flx_b, {ck_b, flx_v, rhog});
    VPart P = part(D.gall_pl, kLallPl, D.kall, flx_b_pl != nullptr, flx_b_pl, {ck_b_pl, flx_v_pl, rhog_pl});
    R.out2 = rhog_b;
    P.out2 = rhog_b_pl;
    return launch_v<V_COURANT_ADJ>(c, R, P, 0.0);
}

int launch_vertical_face_adjoint(nicam_ctx c, double *q_b, double *q_b_pl, const double *q_h_b, const double *q_h_b_pl)
{
    const Dims &D = c->d;
    VPart R = part(D.gall, D.lall, D.kall, true, q_b, {q_h_b});
    VPart P = part(D.gall_pl, kLallPl, D.kall, q_b_pl != nullptr, q_b_pl, {q_h_b_pl});
    return launch_v<V_FACE_ADJ>(c, R, P, 0.0);
}

int launch_vertical_update_adjoint(nicam_ctx c, double *flx_b, double *flx_b_pl, double *q_h_b, double *q_h_b_pl,
                                   const double *rq_b, const double *rq_b_pl, const double *flx_v,
                                   const double *flx_v_pl, const double *q_h, const double *q_h_pl)
{
    const Dims &D = c->d;
    VPart R = part(D.gall, D.lall, D.kall, true, flx_b, {rq_b, flx_v, q_h});
    VPart P = part(D.gall_pl, kLallPl, D.kall, flx_b_pl != nullptr, flx_b_pl, {rq_b_pl, flx_v_pl, q_h_pl});
    R.out2 = q_h_b;
    P.out2 = q_h_b_pl;
    return launch_v<V_UPDATE_ADJ>(c, R, P, 0.0);
}

}  // namespace nicam
