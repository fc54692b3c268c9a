// horizontal_limiter_thuburn on gfx950.
//
// Reference: kernels/NICAM/dyn_horiz_adv_limiter/src/mod_src_tracer.f90:42-430.
//
// The reference stores Qin(gall,kall,lall,2,6) (the inflow bounds of the six
// faces of every cell) and Qout(gall,kall,lall,2) in automatic arrays.  Qin of
// a face is a function of the 4 tracer values around its arc and of that
// arc's cmask, so it is recomputed here where it is used instead of being
// stored: one work-item per cell computes the min/max of its 3 own arcs and of
// the 3 arcs it shares with its lower-left neighbours from a tracer tile held
// in LDS, derives Qout of the cell (phase 1, written to Qout_prev) and clamps
// the face values of its own 3 arcs into the inflow bounds and the
// post-exchange range Qout_post (phase 2), writing both the arc's own face and
// the neighbour's mirrored face.  With Qout_post given, the two phases are
// independent, so one launch does both.
#include "internal.h"

namespace nicam {
namespace {

#ifndef NICAM_HLIM_CHUNK
#define NICAM_HLIM_CHUNK 512  // 256: 678 us, 512: 659 us, 1024: 684 us at gl7
#endif
#ifndef NICAM_HLIM_NTLOAD
#define NICAM_HLIM_NTLOAD 0
#endif
#ifndef NICAM_HLIM_STORE
#define NICAM_HLIM_STORE 0  // 0: nontemporal stores (668 us at gl7), 1: plain mirrored-face stores (697), 2: plain throughout (710)
#endif
constexpr int CHUNK = NICAM_HLIM_CHUNK;  // cells (work-items) per work-group
constexpr int MAXI = 160;

// A load of a streamed input read once (optionally with the nontemporal hint).
__device__ __forceinline__ double ld_once(const double *p)
{
#if NICAM_HLIM_NTLOAD
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (b > a) ? b : a; }
// Fortran SIGN(0.5, x): 0.5 with the sign bit of x
__device__ __forceinline__ double sign_half(double x) { return copysign(0.5, x); }

struct HlimArgs {
    double *q_a;          // (gall,kall,lall,6) inout
    const double *q, *d;  // (gall,kall,lall)
    const double *ch, *cmask;  // (gall,kall,lall,6)
    double *Qprev;        // (gall,kall,lall,2)
    const double *Qpost;  // (gall,kall,lall,2)
    const int *have_sgp;
    int iall, gall, kall, lall;
    int phases;  // 1: allowable range only, 2: clamp only, 3: both
    int nmain;   // work-groups of the regular regions; the rest do the poles
    double eps, big;
    // pole regions (HLIM:307-363, 367-373, 412-425)
    double *q_a_pl, *Qprev_pl;
    const double *q_pl, *d_pl, *ch_pl, *cmask_pl, *Qpost_pl;
    int gall_pl, have_pl;
};

__device__ void hlim_pole(const HlimArgs &A, int t);

__global__ void __launch_bounds__(CHUNK) hlim_kernel(HlimArgs A)
{
    // q for cells [g0 - iall - 1, g0 + CHUNK + iall + 2)
    constexpr int SPAN = CHUNK + 2 * MAXI + 3;
    __shared__ double s_q[SPAN];

    const int iall = A.iall, gall = A.gall, kall = A.kall, lall = A.lall;
    const int nchunk = (gall + CHUNK - 1) / CHUNK;
    int b = blockIdx.x;
    if (b >= A.nmain) {  // pole work-groups
        hlim_pole(A, (b - A.nmain) * CHUNK + threadIdx.x);
        return;
    }
    {
        const int nxcd = 8, per = A.nmain / nxcd;
        if (per > 0 && b < per * nxcd) b = (b % nxcd) * per + b / nxcd;
    }
    const int chunk = b % nchunk;
    const int kl = b / nchunk;
    const int k = kl % kall, l = kl / kall;
    if (l >= lall) return;
    const int g0 = chunk * CHUNK;
    const int lo = g0 - iall - 1;
    const int n_in = CHUNK + 2 * iall + 3;
    const size_t fb = ((size_t)l * kall + k) * gall;
    const int tid = threadIdx.x;
    for (int s = tid; s < n_in; s += CHUNK) {
        const int g = lo + s;
        s_q[s] = (g >= 0 && g < gall) ? A.q[fb + g] : 0.0;
    }
    __syncthreads();

    const int g = g0 + tid;
    if (g >= gall) return;
    const int i = g % iall, j = g / iall;
    const bool sgp = A.have_sgp[l] != 0;
    const size_t plane = (size_t)lall * kall * gall;
    const double BIG = A.big;

    // tracer at absolute cell index p (must lie in the staged window)
    auto Q = [&](int p) { return s_q[p - lo]; };

    // min/max of the 4 tracer values around arc m (0=AI,1=AIJ,2=AJ) with
    // origin p; HLIM:145-160 with the index clamps of :152-153 and the
    // singular-point stencil of :197-203.
    auto arc_minmax = [&](int m, int p, double &qmn, double &qmx) {
        double a, b2, c2, d2;
        if (m == 0) {
            const int ijm1 = max(p - iall, 0);
            a = Q(p); b2 = Q(ijm1); c2 = Q(p + 1); d2 = Q(p + iall + 1);
        } else if (m == 1) {
            if (sgp && p == 0) {
                a = Q(p); b2 = Q(p + iall + 1); c2 = Q(p + iall + 2); d2 = Q(p + iall);
            } else {
                a = Q(p); b2 = Q(p + 1); c2 = Q(p + iall + 1); d2 = Q(p + iall);
            }
        } else {
            const int im1j = max(p - 1, 0);
            a = Q(p); b2 = Q(p + iall + 1); c2 = Q(p + iall); d2 = Q(im1j);
        }
        qmn = dmin(dmin(dmin(a, b2), c2), d2);
        qmx = dmax(dmax(dmax(a, b2), c2), d2);
    };

    // own arcs, origin g (valid for i,j <= iall-2)
    const bool own = (i <= iall - 2 && j <= iall - 2);
    double omn[3], omx[3], ocm[3];
    if (own) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            arc_minmax(m, g, omn[m], omx[m]);
            ocm[m] = A.cmask[(size_t)m * plane + fb + g];  // re-read by the neighbours
        }
    }

    // ---- phase 1: allowable range of the cell (HLIM:216-301) ----
    double qout_min, qout_max;
    const double qc = Q(g);
    if ((A.phases & 1) && i >= 1 && i <= iall - 2 && j >= 1 && j <= iall - 2) {
        double qinmn[6], qinmx[6];
#pragma unroll
        for (int m = 0; m < 3; ++m) {  // faces 1..3: own side of own arcs (:162-187)
            const double c = ocm[m];
            qinmn[m] = (c)*omn[m] + (1.0 - c) * BIG;
            qinmx[m] = (c)*omx[m] + (1.0 - c) * (-BIG);
        }
        // faces 4..6: neighbour side of the arcs at (i-1,j), (i-1,j-1), (i,j-1)
        const int org[3] = {g - 1, g - 1 - iall, g - iall};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double mn, mx;
            arc_minmax(m, org[m], mn, mx);
            const double c = A.cmask[(size_t)m * plane + fb + org[m]];
            qinmn[m + 3] = (c)*BIG + (1.0 - c) * mn;
            qinmx[m + 3] = (c) * (-BIG) + (1.0 - c) * mx;
        }
        double qnext_min = qc, qnext_max = qc;
#pragma unroll
        for (int f = 0; f < 6; ++f) qnext_min = dmin(qnext_min, qinmn[f]);
#pragma unroll
        for (int f = 0; f < 6; ++f) qnext_max = dmax(qnext_max, qinmx[f]);
        double chv[6], chm[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            chv[f] = ld_once(&A.ch[(size_t)f * plane + fb + g]);
            chm[f] = dmin(chv[f], 0.0);
        }
        const double Cin_sum = chm[0] + chm[1] + chm[2] + chm[3] + chm[4] + chm[5];
        const double Cout_sum = chv[0] - chm[0] + chv[1] - chm[1] + chv[2] - chm[2] + chv[3] - chm[3] +
                                chv[4] - chm[4] + chv[5] - chm[5];
        const double CQin_min_sum = chm[0] * qinmn[0] + chm[1] * qinmn[1] + chm[2] * qinmn[2] +
                                    chm[3] * qinmn[3] + chm[4] * qinmn[4] + chm[5] * qinmn[5];
        const double CQin_max_sum = chm[0] * qinmx[0] + chm[1] * qinmx[1] + chm[2] * qinmx[2] +
                                    chm[3] * qinmx[3] + chm[4] * qinmx[4] + chm[5] * qinmx[5];
        const double zerosw = 0.5 - sign_half(fabs(Cout_sum) - A.eps);
        const double dc = ld_once(&A.d[fb + g]);
        qout_min = (qc - CQin_max_sum - qnext_max * (1.0 - Cin_sum - Cout_sum + dc)) / (Cout_sum + zerosw) *
                       (1.0 - zerosw) + qc * zerosw;
        qout_max = (qc - CQin_min_sum - qnext_min * (1.0 - Cin_sum - Cout_sum + dc)) / (Cout_sum + zerosw) *
                       (1.0 - zerosw) + qc * zerosw;
    } else {  // halo (:286-301)
        qout_min = qc;
        qout_max = qc;
    }
    if (A.phases & 1) {
#if NICAM_HLIM_STORE == 2
        A.Qprev[fb + g] = qout_min;
        A.Qprev[plane + fb + g] = qout_max;
#else
        stream_store(&A.Qprev[fb + g], qout_min);
        stream_store(&A.Qprev[plane + fb + g], qout_max);
#endif
    }

    // ---- phase 2: clamp the face values of the own arcs (HLIM:377-410) ----
    if (own && (A.phases & 2)) {
        const int nb[3] = {g + 1, g + iall + 1, g + iall};
        const double qo_min_c = A.Qpost[fb + g], qo_max_c = A.Qpost[plane + fb + g];  // re-read by neighbours
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const double c = ocm[m];
            const double in_mn_s = (c)*omn[m] + (1.0 - c) * BIG;
            const double in_mx_s = (c)*omx[m] + (1.0 - c) * (-BIG);
            const double in_mn_n = (c)*BIG + (1.0 - c) * omn[m];
            const double in_mx_n = (c) * (-BIG) + (1.0 - c) * omx[m];
            const double qo_min_n = A.Qpost[fb + nb[m]], qo_max_n = A.Qpost[plane + fb + nb[m]];
            double qa = ld_once(&A.q_a[(size_t)m * plane + fb + g]);
            qa = (c)*dmin(dmax(qa, in_mn_s), in_mx_s) + (1.0 - c) * dmin(dmax(qa, in_mn_n), in_mx_n);
            qa = (c)*dmax(dmin(qa, qo_max_n), qo_min_n) + (1.0 - c) * dmax(dmin(qa, qo_max_c), qo_min_c);
#if NICAM_HLIM_STORE == 2
            A.q_a[(size_t)m * plane + fb + g] = qa;
#else
            stream_store(&A.q_a[(size_t)m * plane + fb + g], qa);
#endif
#if NICAM_HLIM_STORE >= 1
            A.q_a[(size_t)(m + 3) * plane + fb + nb[m]] = qa;  // one element off the wave's alignment
#else
            stream_store(&A.q_a[(size_t)(m + 3) * plane + fb + nb[m]], qa);
#endif
        }
    }
}

// Pole regions: one work-item per (k, l) (HLIM:307-363, 367-373, 412-425).
// Qin of ring point v depends only on q around it and its cmask, so it is
// recomputed in each of the three passes over the ring instead of stored.
__device__ void hlim_pole(const HlimArgs &A, int t)
{
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const size_t plane = (size_t)kLallPl * kall * npl;
    if (!A.have_pl) {  // only the ring zeroing of :367-370 happens
        if (!(A.phases & 1)) return;
        for (int v = 1; v < npl; ++v) {
            A.Qprev_pl[fb + v] = 0.0;
            A.Qprev_pl[plane + fb + v] = 0.0;
        }
        return;
    }
    const double BIG = A.big;
    // [min,max][side 1,2] of ring point v
    auto qin = [&](int v, double &mn1, double &mn2, double &mx1, double &mx2) {
        int vp1 = v + 1, vm1 = v - 1;
        if (vp1 == npl) vp1 = 1;
        if (vm1 == 0) vm1 = npl - 1;
        const double a = A.q_pl[fb + 0], b = A.q_pl[fb + v], c2 = A.q_pl[fb + vm1], d2 = A.q_pl[fb + vp1];
        const double mn = dmin(dmin(dmin(a, b), c2), d2);
        const double mx = dmax(dmax(dmax(a, b), c2), d2);
        const double c = A.cmask_pl[fb + v];
        mn1 = (c)*mn + (1.0 - c) * BIG;
        mn2 = (c)*BIG + (1.0 - c) * mn;
        mx1 = (c)*mx + (1.0 - c) * (-BIG);
        mx2 = (c) * (-BIG) + (1.0 - c) * mx;
    };
    const double qn = A.q_pl[fb + 0];
    double qnext_min = qn, qnext_max = qn;
    for (int v = 1; v < npl; ++v) {
        double mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        qnext_min = dmin(qnext_min, mn1);
        qnext_max = dmax(qnext_max, mx1);
    }
    double Cin = 0.0, Cout = 0.0, CQmax = 0.0, CQmin = 0.0;
    for (int v = 1; v < npl; ++v) {
        double mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        const double chm = A.cmask_pl[fb + v] * A.ch_pl[fb + v];
        Cin = Cin + chm;
        Cout = Cout - chm + A.ch_pl[fb + v];
        CQmin = CQmin + chm * mn1;
        CQmax = CQmax + chm * mx1;
    }
    const double zerosw = 0.5 - sign_half(fabs(Cout) - A.eps);
    const double dn = A.d_pl[fb + 0];
    const double qo_min = (qn - CQmax - qnext_max * (1.0 - Cin - Cout + dn)) / (Cout + zerosw) * (1.0 - zerosw) + qn * zerosw;
    const double qo_max = (qn - CQmin - qnext_min * (1.0 - Cin - Cout + dn)) / (Cout + zerosw) * (1.0 - zerosw) + qn * zerosw;
    if (A.phases & 1) {
        A.Qprev_pl[fb + 0] = qo_min;
        A.Qprev_pl[plane + fb + 0] = qo_max;
        for (int v = 1; v < npl; ++v) {
            A.Qprev_pl[fb + v] = 0.0;
            A.Qprev_pl[plane + fb + v] = 0.0;
        }
    }
    if (!(A.phases & 2)) return;
    const double pn_min = A.Qpost_pl[fb + 0], pn_max = A.Qpost_pl[plane + fb + 0];
    for (int v = 1; v < npl; ++v) {
        double mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        const double c = A.cmask_pl[fb + v];
        double qa = A.q_a_pl[fb + v];
        qa = (c)*dmin(dmax(qa, mn1), mx1) + (1.0 - c) * dmin(dmax(qa, mn2), mx2);
        qa = (c)*dmax(dmin(qa, A.Qpost_pl[plane + fb + v]), A.Qpost_pl[fb + v]) + (1.0 - c) * dmax(dmin(qa, pn_max), pn_min);
        A.q_a_pl[fb + v] = qa;
    }
}

}  // namespace

int launch_hlim(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                const double *Qpost, const double *Qpost_pl)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (160)");
    HlimArgs A;
    A.q_a = q_a;
    A.q = q;
    A.d = d;
    A.ch = ch;
    A.cmask = cmask;
    A.Qprev = Qprev;
    A.Qpost = Qpost;
    A.have_sgp = c->have_sgp;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.eps = c->eps;
    A.big = c->big;
    const long nmain = (long)((D.gall + CHUNK - 1) / CHUNK) * D.kall * D.lall;
    const long npole = Qprev_pl ? (D.kall * kLallPl + CHUNK - 1) / CHUNK : 0;
    if (nmain + npole > 0x7fffffffL) return fail(NICAM_ERR_ARG, "grid too large for one launch");
    A.nmain = (int)nmain;
    A.q_a_pl = q_a_pl;
    A.q_pl = q_pl;
    A.d_pl = d_pl;
    A.ch_pl = ch_pl;
    A.cmask_pl = cmask_pl;
    A.Qprev_pl = Qprev_pl;
    A.Qpost_pl = Qpost_pl;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl;
    auto run = [&](int phases) -> int {
        A.phases = phases;
        hipLaunchKernelGGL(hlim_kernel, dim3((unsigned)(nmain + npole)), dim3(CHUNK), 0, c->stream, A);
        NICAM_HIP(hipGetLastError());
        return NICAM_OK;
    };
    if (Qpost) return run(3);  // the reference's kernel: Qout_post supplied

    // No Qout_post: do the halo exchange the reference cut out (HLIM:365-375,
    // the commented COMM_data_transfer of Qout) on the device, in place on
    // Qout_prev as NICAM itself exchanges Qout: sources are inner points and
    // pole centres, destinations halo points and pole rings, so the gather
    // reads nothing it writes.  The clamp then reads the exchanged range.
    if (!c->halo_n) return fail(NICAM_ERR_STATE, "Qout_post is NULL but no halo map is set (nicam_comm_setup)");
    if (!Qprev_pl) return fail(NICAM_ERR_ARG, "Qout_prev_pl is needed to exchange Qout with the poles");
    int rc = run(1);
    if (rc != NICAM_OK) return rc;
    rc = launch_halo_fill(c, Qprev, Qprev_pl, D.kall, 2, false);
    if (rc != NICAM_OK) return rc;
    A.Qpost = Qprev;
    A.Qpost_pl = Qprev_pl;
    return run(2);
}

}  // namespace nicam
