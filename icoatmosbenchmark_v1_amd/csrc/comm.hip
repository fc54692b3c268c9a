// On-device halo fill: COMM_data_transfer and COMM_var (kernels/NICAM/
// communication/src/mod_comm.f90:1506-1891, :1895-2101) for a grid whose
// regions and poles all live on one GPU.  The host builds the map once
// (comm.cpp); each call is one gather over (map entry, level, variable).
// Sources are inner points or pole centres and destinations are halo points or
// pole rings, so no destination is also a source within one gather and the
// launch has no ordering hazard.  COMM_var first copies one halo point into
// each pole centre, a source of the main gather, so it is a launch of its own.
#include <vector>

#include "internal.h"

namespace nicam {

namespace {

template <class T>
struct GatherArgs {
    T *var;     // (gall,kmax,lall,nvar)
    T *var_pl;  // (gall_pl,kmax,2,nvar)
    const int *dst, *src;  // encoded points: l*gall + g, or lall*gall + l_pl*gall_pl + v
    int n, gall, kmax, lall, gall_pl, nvar;
};

template <class T>
__device__ __forceinline__ T *locate(const GatherArgs<T> &A, int code, int k, int v)
{
    const int nreg = A.gall * A.lall;
    if (code < nreg) {
        const int l = code / A.gall, g = code - l * A.gall;
        return A.var + (((size_t)v * A.lall + l) * A.kmax + k) * A.gall + g;
    }
    const int p = code - nreg;
    const int l = p / A.gall_pl, g = p - l * A.gall_pl;
    return A.var_pl + (((size_t)v * kLallPl + l) * A.kmax + k) * A.gall_pl + g;
}

// One work-item per (entry, level); entries vary fastest so that the
// consecutive halo points of one edge share cache lines of their rows.
template <class T>
__global__ void __launch_bounds__(256) halo_gather_kernel(GatherArgs<T> A)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)A.n * A.kmax) return;
    const int e = (int)(t % A.n), k = (int)(t / A.n);
    const int d = A.dst[e], s = A.src[e];
    for (int v = 0; v < A.nvar; ++v) *locate(A, d, k, v) = *locate(A, s, k, v);
}

template <class T>
int gather(nicam_ctx c, GatherArgs<T> A)
{
    const long n = (long)A.n * A.kmax;
    if (n == 0) return NICAM_OK;
    if ((n + 255) / 256 > 0x7fffffffL) return fail(NICAM_ERR_ARG, "halo fill too large for one launch");
    hipLaunchKernelGGL(halo_gather_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace

int comm_setup(nicam_ctx c, int rlevel)
{
    const Dims &D = c->d;
    if (rlevel < 0 || rlevel > 8) return fail(NICAM_ERR_ARG, "rlevel must be 0..8");
    const long rall = 10L << (2 * rlevel);
    if (D.lall != rall)
        return fail(NICAM_ERR_ARG, "the on-device halo fill needs every region on this GPU: lall = 10*4^rlevel");
    if (!D.have_pl) return fail(NICAM_ERR_ARG, "the on-device halo fill needs the pole regions (have_pl)");
    std::vector<HaloPoint> dst, src;
    std::vector<int> sgp;
    HaloPoint pole_src[2];
    int rc = build_halo_map(rlevel, D.gall_1d, D.gall_pl, dst, src, sgp, pole_src);
    if (rc != NICAM_OK) return rc;
    for (int l = 0; l < D.lall; ++l)
        if (sgp[l] != c->have_sgp_h[l])
            return fail(NICAM_ERR_ARG, "have_sgp given at context creation disagrees with the icosahedral topology");
    const int nreg = D.gall * D.lall;
    auto code = [&](const HaloPoint &p) { return p.kind == 0 ? p.l * D.gall + p.g : nreg + p.l * D.gall_pl + p.g; };
    std::vector<int> hd, hs;
    for (int p = 0; p < 2; ++p) {  // pole centres first (COMM_var)
        hd.push_back(nreg + p * D.gall_pl);
        hs.push_back(code(pole_src[p]));
    }
    for (size_t e = 0; e < dst.size(); ++e) {
        hd.push_back(code(dst[e]));
        hs.push_back(code(src[e]));
    }
    if (c->halo_dst) (void)hipFree(c->halo_dst);
    if (c->halo_src) (void)hipFree(c->halo_src);
    c->halo_dst = c->halo_src = nullptr;
    c->halo_n = 0;
    NICAM_HIP(hipMalloc(&c->halo_dst, sizeof(int) * hd.size()));
    NICAM_HIP(hipMalloc(&c->halo_src, sizeof(int) * hs.size()));
    NICAM_HIP(hipMemcpy(c->halo_dst, hd.data(), sizeof(int) * hd.size(), hipMemcpyHostToDevice));
    NICAM_HIP(hipMemcpy(c->halo_src, hs.data(), sizeof(int) * hs.size(), hipMemcpyHostToDevice));
    c->halo_n = (int)dst.size();
    c->rlevel = rlevel;
    rc = halo_adjoint_setup(c, std::vector<int>(hd.begin() + 2, hd.end()), std::vector<int>(hs.begin() + 2, hs.end()));
    if (rc != NICAM_OK) return rc;

    std::vector<int> slot, tab, pole_tab;
    rc = build_arc_partners(rlevel, D.gall_1d, D.gall_pl, dst, src, sgp, slot, tab, pole_tab);
    if (rc != NICAM_OK) return rc;
    int **arr[] = {&c->arc_slot, &c->arc_tab, &c->arc_pole};
    const std::vector<int> *vec[] = {&slot, &tab, &pole_tab};
    for (int a = 0; a < 3; ++a) {
        if (*arr[a]) (void)hipFree(*arr[a]);
        *arr[a] = nullptr;
        NICAM_HIP(hipMalloc(arr[a], sizeof(int) * std::max<size_t>(1, vec[a]->size())));
        NICAM_HIP(hipMemcpy(*arr[a], vec[a]->data(), sizeof(int) * vec[a]->size(), hipMemcpyHostToDevice));
    }
    return NICAM_OK;
}

template <class T>
int halo_fill_t(nicam_ctx c, T *var, T *var_pl, int kmax, int nvar, bool pole_fill)
{
    if (!c->halo_n) return fail(NICAM_ERR_STATE, "halo map not set (nicam_comm_setup)");
    if (kmax < 1 || nvar < 1) return fail(NICAM_ERR_ARG, "kmax and vmax must be positive");
    const Dims &D = c->d;
    GatherArgs<T> A;
    A.var = var;
    A.var_pl = var_pl;
    A.gall = D.gall;
    A.kmax = kmax;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.nvar = nvar;
    if (pole_fill) {
        A.dst = c->halo_dst;
        A.src = c->halo_src;
        A.n = 2;
        int rc = gather(c, A);
        if (rc != NICAM_OK) return rc;
    }
    A.dst = c->halo_dst + 2;
    A.src = c->halo_src + 2;
    A.n = c->halo_n;
    return gather(c, A);
}

int launch_halo_fill(nicam_ctx c, double *var, double *var_pl, int kmax, int nvar, bool pole_fill)
{
    return halo_fill_t<double>(c, var, var_pl, kmax, nvar, pole_fill);
}

int launch_halo_fill_f32(nicam_ctx c, float *var, float *var_pl, int kmax, int nvar, bool pole_fill)
{
    return halo_fill_t<float>(c, var, var_pl, kmax, nvar, pole_fill);
}

}  // namespace nicam
