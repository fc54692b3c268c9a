// The flux-form tracer update that conserves mass exactly, on gfx950.
//
// rhogq(c) -= RAREA(c) * sum over the six faces m of F(c, m), with F the
// outward tracer-mass transfer through the face, in FP64.  Every arc's
// transfer is one number that both of its cells use with opposite signs:
//   * an arc between two inner cells of a region, from origin o in direction
//     a = 1..3 (the arcs horizontal_flux forms,
//     FLX/src/mod_src_tracer.f90:234-309), transfers
//     X = AREA(o) * flx_h(o, a) * q_a(o, a) out of o (flx_h carries RAREA(o),
//     so the product is the arc's mass flux times the face value); face a+3
//     of the other cell uses -X, read from the origin, never its own
//     flx_h(., a+3) (which carries that cell's RAREA and, in FP32, rounds
//     differently);
//   * a face toward a halo point is taken from the inner cell's own side,
//     AREA(c) * flx_h(c, m) * q_a(c, m): halo cells may carry no metrics
//     (RAREA = 0 there gives flx_h = 0 on their own faces);
//   * a face two regions share uses the transfer of the region with the lower
//     index (regions before the poles): the partner table nicam_comm_setup
//     builds from the halo map;
//   * face 6 of a singular-point cell carries nothing, as the flux sets it
//     (FLX/src/mod_src_tracer.f90:311-320).
// The pole centre takes minus the transfers of the five region faces that
// point at it.  The vertical update of the mixed step (FP32 flx_v and q_h,
// FP64 rhogq) is here too: each half level's transfer flx_v*q_h is formed
// once in FP64 and used by the levels above and below, as the FP64 form in
// vglue.hip does.  The sum of AREA * rhogq over the sphere then changes by the
// rounding of the FP64 operations alone.  flx_h and q_a may be FP32 (the
// mixed step: FP32 fluxes and limiter, FP64 accumulation) or FP64.
#include "internal.h"

namespace nicam {
namespace {

constexpr int CBLK = 256;

template <class T>
struct ConsArgs {
    double *rq, *rq_pl;
    const T *flx, *qa;
    const double *gp, *gp_pl;  // GMTR_p (gall,1,lall,8), (gall_pl,1,2,8): AREA, RAREA, ...
    const int *slot, *tab, *pole_tab;
    long nreg;  // one plane of flx_h / q_a
    int iall, gall, kall, lall, gall_pl, have_pl;
};

// AREA(g) * flx_h(g, m) * q_a(g, m) of region l, level k, face m = 1..6
template <class T>
__device__ __forceinline__ double own_x(const ConsArgs<T> &A, int l, int g, int m, int k)
{
    const size_t f = (size_t)(m - 1) * A.nreg + ((size_t)l * A.kall + k) * A.gall + g;
    return A.gp[(size_t)l * A.gall + g] * (double)A.flx[f] * (double)A.qa[f];
}

// outward transfer of inner cell g of region l through its face m = 1..6, by
// this region's arcs
template <class T>
__device__ __forceinline__ double face_x(const ConsArgs<T> &A, int l, int g, int m, int k)
{
    if (m <= 3) return own_x(A, l, g, m, k);
    const int i = g % A.iall, j = g / A.iall;
    const int oi = (m == 6) ? i : i - 1, oj = (m == 4) ? j : j - 1;  // the arc's origin: im1j, im1jm1, ijm1
    if (oi < 1 || oj < 1) return own_x(A, l, g, m, k);              // a halo origin: this cell's side
    return -own_x(A, l, oj * A.iall + oi, m - 3, k);
}

template <class T>
__device__ __forceinline__ double partner_x(const ConsArgs<T> &A, int code, int k)
{
    const int m = code % 8, lg = code / 8;
    return -face_x(A, lg / A.gall, lg % A.gall, m, k);
}

template <class T>
__global__ void __launch_bounds__(CBLK) cons_update_kernel(ConsArgs<T> A)
{
    const long t = (long)blockIdx.x * CBLK + threadIdx.x;
    if (t < A.nreg) {
        const int g = (int)(t % A.gall), k = (int)((t / A.gall) % A.kall), l = (int)(t / ((long)A.gall * A.kall));
        const int i = g % A.iall, j = g / A.iall;
        if (i < 1 || j < 1 || i > A.iall - 2 || j > A.iall - 2) return;  // halo: left to the exchange
        const int s = A.slot[(size_t)l * A.gall + g];
        double sum = 0.0;
#pragma unroll
        for (int m = 1; m <= 6; ++m) {
            const int code = s >= 0 ? A.tab[(size_t)s * 6 + m - 1] : -1;
            double f;
            if (code == -1) f = face_x(A, l, g, m, k);
            else if (code == -2) f = 0.0;
            else f = partner_x(A, code, k);
            sum = sum + f;
        }
        const double rarea = A.gp[((size_t)A.lall + l) * A.gall + g];
        A.rq[t] = A.rq[t] - rarea * sum;
        return;
    }
    const long p = t - A.nreg;  // one work-item per (level, pole centre)
    if (!A.have_pl || p >= (long)A.kall * kLallPl) return;
    const int k = (int)(p % A.kall), pl = (int)(p / A.kall), npl = A.gall_pl - 1;
    double sum = 0.0;
    for (int v = 0; v < npl; ++v) sum = sum + partner_x(A, A.pole_tab[pl * npl + v], k);
    const double rarea = A.gp_pl[((size_t)kLallPl + pl) * A.gall_pl];
    const size_t o = ((size_t)pl * A.kall + k) * A.gall_pl;
    A.rq_pl[o] = A.rq_pl[o] - rarea * sum;
}

template <class T>
int launch_update_cons_t(nicam_ctx c, double *rq, double *rq_pl, const T *flx, const T *qa)
{
    const Dims &D = c->d;
    if (!c->arc_slot) return fail(NICAM_ERR_STATE, "the conservative update needs the halo map (nicam_comm_setup)");
    ConsArgs<T> A;
    A.rq = rq;
    A.rq_pl = rq_pl;
    A.flx = flx;
    A.qa = qa;
    A.gp = c->GMTR_p;
    A.gp_pl = c->GMTR_p_pl;
    A.slot = c->arc_slot;
    A.tab = c->arc_tab;
    A.pole_tab = c->arc_pole;
    A.nreg = (long)D.gall * D.kall * D.lall;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl && rq_pl;
    const long n = A.nreg + (A.have_pl ? (long)D.kall * kLallPl : 0), nb = (n + CBLK - 1) / CBLK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(cons_update_kernel<T>, dim3((unsigned)nb), dim3(CBLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

// rhogq(k) -= (flx_v(k+1) q_h(k+1) - flx_v(k) q_h(k)) rdgz(k) on kmin..kmax,
// the products widened to FP64; one work-item per cell-level
struct VUpdArgs {
    double *rq, *rq_pl;
    const float *flx, *flx_pl, *qh, *qh_pl;
    const double *rdgz;
    long nR, nP;
    int gall, gall_pl, kall;
};

__global__ void __launch_bounds__(CBLK) vupdate_mixed_kernel(VUpdArgs A)
{
    long t = (long)blockIdx.x * CBLK + threadIdx.x;
    const bool pole = t >= A.nR;
    if (pole) t -= A.nR;
    if (t >= (pole ? A.nP : A.nR)) return;
    const int gall = pole ? A.gall_pl : A.gall;
    const int k = (int)((t / gall) % A.kall), kmin = 1, kmax = A.kall - 2;
    if (k < kmin || k > kmax) return;
    const float *flx = pole ? A.flx_pl : A.flx, *qh = pole ? A.qh_pl : A.qh;
    double *rq = pole ? A.rq_pl : A.rq;
    const double up = (k + 1 <= kmax) ? (double)flx[t + gall] * (double)qh[t + gall] : 0.0;
    const double lo = (k >= kmin + 1) ? (double)flx[t] * (double)qh[t] : 0.0;
    rq[t] = rq[t] - (up - lo) * A.rdgz[k];
}

}  // namespace

int launch_vertical_update_f32(nicam_ctx c, double *rq, double *rq_pl, const float *flx_v, const float *flx_v_pl,
                               const float *q_h, const float *q_h_pl)
{
    const Dims &D = c->d;
    VUpdArgs A;
    A.rq = rq;
    A.rq_pl = rq_pl;
    A.flx = flx_v;
    A.flx_pl = flx_v_pl;
    A.qh = q_h;
    A.qh_pl = q_h_pl;
    A.rdgz = c->rdgz;
    A.nR = (long)D.gall * D.kall * D.lall;
    A.nP = (D.have_pl && rq_pl) ? (long)D.gall_pl * D.kall * kLallPl : 0;
    A.gall = D.gall;
    A.gall_pl = D.gall_pl;
    A.kall = D.kall;
    const long nb = (A.nR + A.nP + CBLK - 1) / CBLK;
    if (nb > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    hipLaunchKernelGGL(vupdate_mixed_kernel, dim3((unsigned)nb), dim3(CBLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

int launch_update_cons(nicam_ctx c, double *rq, double *rq_pl, const double *flx, const double *qa)
{
    return launch_update_cons_t<double>(c, rq, rq_pl, flx, qa);
}

int launch_update_cons_f32(nicam_ctx c, double *rq, double *rq_pl, const float *flx, const float *qa)
{
    return launch_update_cons_t<float>(c, rq, rq_pl, flx, qa);
}

}  // namespace nicam
