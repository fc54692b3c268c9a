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
template <class T>
__device__ __forceinline__ T ld_once(const T *p)
{
#if NICAM_HLIM_NTLOAD
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

template <class T>
__device__ __forceinline__ T dmin(T a, T b) { return (b < a) ? b : a; }
template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return (b > a) ? b : a; }
// Fortran SIGN(0.5, x): 0.5 with the sign bit of x
__device__ __forceinline__ double sign_half(double x) { return copysign(0.5, x); }
__device__ __forceinline__ float sign_half(float x) { return copysignf(0.5f, x); }
__device__ __forceinline__ double tabs(double x) { return fabs(x); }
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }

template <class T>
struct HlimArgs {
    T *q_a;          // (gall,kall,lall,6) inout
    const T *q, *d;  // (gall,kall,lall)
    const T *ch, *cmask;  // (gall,kall,lall,6)
    T *Qprev;        // (gall,kall,lall,2)
    const T *Qpost;  // (gall,kall,lall,2)
    const int *have_sgp;
    int iall, gall, kall, lall;
    int phases;  // 1: allowable range only, 2: clamp only, 3: both
    int nmain;   // work-groups of the regular regions; the rest do the poles
    T eps, big;
    // pole regions (HLIM:307-363, 367-373, 412-425)
    T *q_a_pl, *Qprev_pl;
    const T *q_pl, *d_pl, *ch_pl, *cmask_pl, *Qpost_pl;
    int gall_pl, have_pl;
    // branch mask (masked variant): uint8 per cell-level, bit m set where face
    // m+1 of an own arc changed; uint8 per pole point, bit 0 for ring faces
    uint8_t *mask, *mask_pl;
};

template <class T>
__device__ void hlim_pole(const HlimArgs<T> &A, int t);

template <class T>
__global__ void __launch_bounds__(CHUNK) hlim_kernel(HlimArgs<T> A)
{
    // q for cells [g0 - iall - 1, g0 + CHUNK + iall + 2)
    constexpr int SPAN = CHUNK + 2 * MAXI + 3;
    __shared__ T s_q[SPAN];

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
        s_q[s] = (g >= 0 && g < gall) ? A.q[fb + g] : T(0);
    }
    __syncthreads();

    const int g = g0 + tid;
    if (g >= gall) return;
    const int i = g % iall, j = g / iall;
    const bool sgp = A.have_sgp[l] != 0;
    const size_t plane = (size_t)lall * kall * gall;
    const T BIG = A.big;

    // tracer at absolute cell index p (must lie in the staged window)
    auto Q = [&](int p) { return s_q[p - lo]; };

    // min/max of the 4 tracer values around arc m (0=AI,1=AIJ,2=AJ) with
    // origin p; HLIM:145-160 with the index clamps of :152-153 and the
    // singular-point stencil of :197-203.
    auto arc_minmax = [&](int m, int p, T &qmn, T &qmx) {
        T a, b2, c2, d2;
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
    T omn[3], omx[3], ocm[3];
    if (own) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            arc_minmax(m, g, omn[m], omx[m]);
            ocm[m] = A.cmask[(size_t)m * plane + fb + g];  // re-read by the neighbours
        }
    }

    // ---- phase 1: allowable range of the cell (HLIM:216-301) ----
    T qout_min, qout_max;
    const T qc = Q(g);
    if ((A.phases & 1) && i >= 1 && i <= iall - 2 && j >= 1 && j <= iall - 2) {
        T qinmn[6], qinmx[6];
#pragma unroll
        for (int m = 0; m < 3; ++m) {  // faces 1..3: own side of own arcs (:162-187)
            const T c = ocm[m];
            qinmn[m] = (c)*omn[m] + (T(1) - c) * BIG;
            qinmx[m] = (c)*omx[m] + (T(1) - c) * (-BIG);
        }
        // faces 4..6: neighbour side of the arcs at (i-1,j), (i-1,j-1), (i,j-1)
        const int org[3] = {g - 1, g - 1 - iall, g - iall};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            T mn, mx;
            arc_minmax(m, org[m], mn, mx);
            const T c = A.cmask[(size_t)m * plane + fb + org[m]];
            qinmn[m + 3] = (c)*BIG + (T(1) - c) * mn;
            qinmx[m + 3] = (c) * (-BIG) + (T(1) - c) * mx;
        }
        T qnext_min = qc, qnext_max = qc;
#pragma unroll
        for (int f = 0; f < 6; ++f) qnext_min = dmin(qnext_min, qinmn[f]);
#pragma unroll
        for (int f = 0; f < 6; ++f) qnext_max = dmax(qnext_max, qinmx[f]);
        T chv[6], chm[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            chv[f] = ld_once(&A.ch[(size_t)f * plane + fb + g]);
            chm[f] = dmin(chv[f], T(0));
        }
        const T Cin_sum = chm[0] + chm[1] + chm[2] + chm[3] + chm[4] + chm[5];
        const T Cout_sum = chv[0] - chm[0] + chv[1] - chm[1] + chv[2] - chm[2] + chv[3] - chm[3] +
                                chv[4] - chm[4] + chv[5] - chm[5];
        const T CQin_min_sum = chm[0] * qinmn[0] + chm[1] * qinmn[1] + chm[2] * qinmn[2] +
                                    chm[3] * qinmn[3] + chm[4] * qinmn[4] + chm[5] * qinmn[5];
        const T CQin_max_sum = chm[0] * qinmx[0] + chm[1] * qinmx[1] + chm[2] * qinmx[2] +
                                    chm[3] * qinmx[3] + chm[4] * qinmx[4] + chm[5] * qinmx[5];
        const T zerosw = T(0.5) - sign_half(tabs(Cout_sum) - A.eps);
        const T dc = ld_once(&A.d[fb + g]);
        qout_min = (qc - CQin_max_sum - qnext_max * (T(1) - Cin_sum - Cout_sum + dc)) / (Cout_sum + zerosw) *
                       (T(1) - zerosw) + qc * zerosw;
        qout_max = (qc - CQin_min_sum - qnext_min * (T(1) - Cin_sum - Cout_sum + dc)) / (Cout_sum + zerosw) *
                       (T(1) - zerosw) + qc * zerosw;
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
    if (A.mask && (A.phases & 2) && !own) A.mask[fb + g] = 0;
    if (own && (A.phases & 2)) {
        uint32_t bits = 0;
        const int nb[3] = {g + 1, g + iall + 1, g + iall};
        const T qo_min_c = A.Qpost[fb + g], qo_max_c = A.Qpost[plane + fb + g];  // re-read by neighbours
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const T c = ocm[m];
            const T in_mn_s = (c)*omn[m] + (T(1) - c) * BIG;
            const T in_mx_s = (c)*omx[m] + (T(1) - c) * (-BIG);
            const T in_mn_n = (c)*BIG + (T(1) - c) * omn[m];
            const T in_mx_n = (c) * (-BIG) + (T(1) - c) * omx[m];
            const T qo_min_n = A.Qpost[fb + nb[m]], qo_max_n = A.Qpost[plane + fb + nb[m]];
            T qa = ld_once(&A.q_a[(size_t)m * plane + fb + g]);
            const T qa_in = qa;
            qa = (c)*dmin(dmax(qa, in_mn_s), in_mx_s) + (T(1) - c) * dmin(dmax(qa, in_mn_n), in_mx_n);
            qa = (c)*dmax(dmin(qa, qo_max_n), qo_min_n) + (T(1) - c) * dmax(dmin(qa, qo_max_c), qo_min_c);
            bits |= (qa != qa_in ? 1u : 0u) << m;
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
        if (A.mask) A.mask[fb + g] = (uint8_t)bits;
    }
}

// Pole regions: one work-item per (k, l) (HLIM:307-363, 367-373, 412-425).
// Qin of ring point v depends only on q around it and its cmask, so it is
// recomputed in each of the three passes over the ring instead of stored.
template <class T>
__device__ void hlim_pole(const HlimArgs<T> &A, int t)
{
    const int kall = A.kall, npl = A.gall_pl;
    if (t >= kall * kLallPl) return;
    const int k = t % kall, l = t / kall;
    const size_t fb = ((size_t)l * kall + k) * npl;
    const size_t plane = (size_t)kLallPl * kall * npl;
    if (!A.have_pl) {  // only the ring zeroing of :367-370 happens
        if (!(A.phases & 1)) return;
        for (int v = 1; v < npl; ++v) {
            A.Qprev_pl[fb + v] = T(0);
            A.Qprev_pl[plane + fb + v] = T(0);
        }
        return;
    }
    const T BIG = A.big;
    // [min,max][side 1,2] of ring point v
    auto qin = [&](int v, T &mn1, T &mn2, T &mx1, T &mx2) {
        int vp1 = v + 1, vm1 = v - 1;
        if (vp1 == npl) vp1 = 1;
        if (vm1 == 0) vm1 = npl - 1;
        const T a = A.q_pl[fb + 0], b = A.q_pl[fb + v], c2 = A.q_pl[fb + vm1], d2 = A.q_pl[fb + vp1];
        const T mn = dmin(dmin(dmin(a, b), c2), d2);
        const T mx = dmax(dmax(dmax(a, b), c2), d2);
        const T c = A.cmask_pl[fb + v];
        mn1 = (c)*mn + (T(1) - c) * BIG;
        mn2 = (c)*BIG + (T(1) - c) * mn;
        mx1 = (c)*mx + (T(1) - c) * (-BIG);
        mx2 = (c) * (-BIG) + (T(1) - c) * mx;
    };
    const T qn = A.q_pl[fb + 0];
    T qnext_min = qn, qnext_max = qn;
    for (int v = 1; v < npl; ++v) {
        T mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        qnext_min = dmin(qnext_min, mn1);
        qnext_max = dmax(qnext_max, mx1);
    }
    T Cin = T(0), Cout = T(0), CQmax = T(0), CQmin = T(0);
    for (int v = 1; v < npl; ++v) {
        T mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        const T chm = A.cmask_pl[fb + v] * A.ch_pl[fb + v];
        Cin = Cin + chm;
        Cout = Cout - chm + A.ch_pl[fb + v];
        CQmin = CQmin + chm * mn1;
        CQmax = CQmax + chm * mx1;
    }
    const T zerosw = T(0.5) - sign_half(tabs(Cout) - A.eps);
    const T dn = A.d_pl[fb + 0];
    const T qo_min = (qn - CQmax - qnext_max * (T(1) - Cin - Cout + dn)) / (Cout + zerosw) * (T(1) - zerosw) + qn * zerosw;
    const T qo_max = (qn - CQmin - qnext_min * (T(1) - Cin - Cout + dn)) / (Cout + zerosw) * (T(1) - zerosw) + qn * zerosw;
    if (A.phases & 1) {
        A.Qprev_pl[fb + 0] = qo_min;
        A.Qprev_pl[plane + fb + 0] = qo_max;
        for (int v = 1; v < npl; ++v) {
            A.Qprev_pl[fb + v] = T(0);
            A.Qprev_pl[plane + fb + v] = T(0);
        }
    }
    if (!(A.phases & 2)) return;
    const T pn_min = A.Qpost_pl[fb + 0], pn_max = A.Qpost_pl[plane + fb + 0];
    if (A.mask_pl) A.mask_pl[fb + 0] = 0;
    for (int v = 1; v < npl; ++v) {
        T mn1, mn2, mx1, mx2;
        qin(v, mn1, mn2, mx1, mx2);
        const T c = A.cmask_pl[fb + v];
        T qa = A.q_a_pl[fb + v];
        const T qa_in = qa;
        qa = (c)*dmin(dmax(qa, mn1), mx1) + (T(1) - c) * dmin(dmax(qa, mn2), mx2);
        qa = (c)*dmax(dmin(qa, A.Qpost_pl[plane + fb + v]), A.Qpost_pl[fb + v]) + (T(1) - c) * dmax(dmin(qa, pn_max), pn_min);
        A.q_a_pl[fb + v] = qa;
        if (A.mask_pl) A.mask_pl[fb + v] = (qa != qa_in) ? 1 : 0;
    }
}

}  // namespace

namespace {

int halo_fill(nicam_ctx c, double *v, double *v_pl, int kmax, int nvar)
{
    return launch_halo_fill(c, v, v_pl, kmax, nvar, false);
}
int halo_fill(nicam_ctx c, float *v, float *v_pl, int kmax, int nvar)
{
    return launch_halo_fill_f32(c, v, v_pl, kmax, nvar, false);
}

template <class T>
int launch_hlim_t(nicam_ctx c, T *q_a, T *q_a_pl, const T *q, const T *q_pl, const T *d, const T *d_pl,
                  const T *ch, const T *ch_pl, const T *cmask, const T *cmask_pl, T *Qprev, T *Qprev_pl,
                  const T *Qpost, const T *Qpost_pl, T eps, T big, uint8_t *mask = nullptr,
                  uint8_t *mask_pl = nullptr)
{
    const Dims &D = c->d;
    if (D.gall_1d > MAXI) return fail(NICAM_ERR_ARG, "gall_1d larger than the LDS tiles (160)");
    HlimArgs<T> A;
    A.mask = mask;
    A.mask_pl = mask_pl;
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
    A.eps = eps;
    A.big = big;
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
        hipLaunchKernelGGL(hlim_kernel<T>, dim3((unsigned)(nmain + npole)), dim3(CHUNK), 0, c->stream, A);
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
    rc = halo_fill(c, Qprev, Qprev_pl, D.kall, 2);
    if (rc != NICAM_OK) return rc;
    A.Qpost = Qprev;
    A.Qpost_pl = Qprev_pl;
    return run(2);
}

}  // namespace

int launch_hlim(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                const double *Qpost, const double *Qpost_pl)
{
    return launch_hlim_t<double>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                 Qpost_pl, c->eps, c->big);
}

int launch_hlim_mask(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                     const double *d, const double *d_pl, const double *ch, const double *ch_pl,
                     const double *cmask, const double *cmask_pl, double *Qprev, double *Qprev_pl,
                     const double *Qpost, const double *Qpost_pl, uint8_t *mask, uint8_t *mask_pl)
{
    return launch_hlim_t<double>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                 Qpost_pl, c->eps, c->big, mask, mask_pl);
}

int launch_hlim_f32(nicam_ctx c, float *q_a, float *q_a_pl, const float *q, const float *q_pl, const float *d,
                    const float *d_pl, const float *ch, const float *ch_pl, const float *cmask,
                    const float *cmask_pl, float *Qprev, float *Qprev_pl, const float *Qpost,
                    const float *Qpost_pl)
{
    // CONST_EPS and CONST_HUGE of the reference built with -DSINGLE
    return launch_hlim_t<float>(c, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl, Qprev, Qprev_pl, Qpost,
                                Qpost_pl, kEpsF32, kHugeF32);
}

}  // namespace nicam
