// Global tracer-mass sums in FP64 on gfx950.
//
// M(k) = sum over the inner cells of every region of GMTR_p(AREA)*rhogq(k),
// plus both pole centres' area times their rhogq(k).  The summands are formed
// in FP64 (an FP32 rhogq is widened exactly), and the order of the additions
// is fixed: every (level, region) pair is one work-group that sums its cells
// in a fixed strided order and then by a fixed tree in LDS; one more
// work-group adds each level's region partial sums in region order and the
// two pole terms.  The result is therefore the same on every run and every
// launch shape of the device, which a conservation check needs.
//
// Two entry points share the kernels.  nicam_tracer_mass keeps its partial
// sums in a buffer the context allocates on the first call.
// nicam_tracer_mass_noalloc is for callers that cannot allocate device memory
// (code under HIP graph capture from its first call, serving loops, harnesses
// that forbid hipMalloc): its partial sums live in `scratch`, a file-scope
// device array of kMassScratch doubles, one copy per GPU, whose device address
// nicam_ctx_create looks up once per device (mass_scratch_init).
#include <mutex>

#include "internal.h"

namespace nicam {
namespace {

constexpr int MBLK = 256;
constexpr size_t kMassScratch = 1u << 16;  // partial sums: kall * lall <= 65536 (512 KiB per GPU)

__device__ __align__(256) double scratch[kMassScratch];

template <class T>
struct MassArgs {
    const T *rq, *rq_pl;
    const double *area, *area_pl;  // GMTR_p(:,1,:,AREA) of the regions and of the poles
    double *part;                  // [kall][lall]
    double *mass;                  // [kall]
    int iall, gall, kall, lall, gall_pl, have_pl;
};

// one work-group per (level, region): its inner cells in a fixed order
template <class T>
__global__ void __launch_bounds__(MBLK) mass_part_kernel(MassArgs<T> A)
{
    __shared__ double s[MBLK];
    const int k = blockIdx.x % A.kall, l = blockIdx.x / A.kall;
    const int ni = A.iall - 2, n = ni * ni;
    const T *rq = A.rq + ((size_t)l * A.kall + k) * A.gall;
    const double *area = A.area + (size_t)l * A.gall;
    double acc = 0.0;
    for (int idx = threadIdx.x; idx < n; idx += MBLK) {
        const int g = (1 + idx / ni) * A.iall + 1 + idx % ni;
        acc = acc + area[g] * (double)rq[g];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = MBLK / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) A.part[(size_t)k * A.lall + l] = s[0];
}

// each level: the regions' partial sums in region order, then the poles
template <class T>
__global__ void __launch_bounds__(MBLK) mass_level_kernel(MassArgs<T> A)
{
    for (int k = threadIdx.x; k < A.kall; k += MBLK) {
        double m = 0.0;
        for (int l = 0; l < A.lall; ++l) m = m + A.part[(size_t)k * A.lall + l];
        if (A.have_pl)
            for (int p = 0; p < kLallPl; ++p)
                m = m + A.area_pl[(size_t)p * A.gall_pl] * (double)A.rq_pl[((size_t)p * A.kall + k) * A.gall_pl];
        A.mass[k] = m;
    }
}

template <class T>
int launch_mass_t(nicam_ctx c, const T *rq, const T *rq_pl, double *mass, double *part)
{
    const Dims &D = c->d;
    MassArgs<T> A;
    A.rq = rq;
    A.rq_pl = rq_pl;
    A.area = c->GMTR_p;  // AREA is the first of the GMTR_p planes
    A.area_pl = c->GMTR_p_pl;
    A.part = part;
    A.mass = mass;
    A.iall = D.gall_1d;
    A.gall = D.gall;
    A.kall = D.kall;
    A.lall = D.lall;
    A.gall_pl = D.gall_pl;
    A.have_pl = D.have_pl && rq_pl != nullptr;
    hipLaunchKernelGGL(mass_part_kernel<T>, dim3((unsigned)(D.kall * D.lall)), dim3(MBLK), 0, c->stream, A);
    hipLaunchKernelGGL(mass_level_kernel<T>, dim3(1), dim3(MBLK), 0, c->stream, A);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

// the device address of `scratch` on each GPU (the host-side name is not one)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Once per device, with that device current: the address of its `scratch`.
int mass_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(NICAM_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return NICAM_OK;
    int cur = 0;
    NICAM_HIP(hipGetDevice(&cur));
    if (cur != device) NICAM_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double *>(p);
    return NICAM_OK;
}

int launch_mass(nicam_ctx c, const double *rq, const double *rq_pl, double *mass, double *part)
{
    return launch_mass_t<double>(c, rq, rq_pl, mass, part);
}

int launch_mass_f32(nicam_ctx c, const float *rq, const float *rq_pl, double *mass, double *part)
{
    return launch_mass_t<float>(c, rq, rq_pl, mass, part);
}

}  // namespace nicam

using namespace nicam;

// nicam_tracer_mass without any allocation: device pointers only.  Calls that
// overlap on one GPU share its `scratch` and race; calls on one stream are
// ordered, and each call writes every partial sum it reads.
extern "C" int nicam_tracer_mass_noalloc(nicam_ctx c, const double *rhogq, const double *rhogq_pl, double *mass)
{
    if (!c) return fail(NICAM_ERR_ARG, "null context");
    if (!rhogq || !mass) return fail(NICAM_ERR_ARG, "null argument");
    if (c->d.have_pl && !rhogq_pl) return fail(NICAM_ERR_ARG, "null pole array");
    if ((size_t)c->d.kall * c->d.lall > kMassScratch)
        return fail(NICAM_ERR_ARG, "kall * lall exceeds the 65536 partial sums of the static scratch array");
    if (c->ptr_mode != NICAM_PTR_DEVICE)
        return fail(NICAM_ERR_STATE, "nicam_tracer_mass_noalloc takes device pointers only (NICAM_PTR_DEVICE)");
    if (!c->have_hmetrics) return fail(NICAM_ERR_STATE, "horizontal metrics not set (nicam_set_hmetrics)");
    double *p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxDevices) p = g_scratch_ptr[c->device];
    }
    if (!p) return fail(NICAM_ERR_STATE, "the scratch address of this device is not set (nicam_ctx_create)");
    return launch_mass(c, rhogq, c->d.have_pl ? rhogq_pl : nullptr, mass, p);
}
