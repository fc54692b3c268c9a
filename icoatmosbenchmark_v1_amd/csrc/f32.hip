// Rounding of double-precision arrays to FP32 on the device, for the _f32
// entry points: round to nearest even, as Fortran's real(x, kind=SP) rounds
// the reference's double-precision input data when it is built with -DSINGLE
// (e.g. kernels/NICAM/dyn_horiz_adv_limiter/src/main.f90:153-162).
#include "internal.h"

namespace nicam {
namespace {

__global__ void __launch_bounds__(256) to_float_kernel(const double *__restrict__ src, float *__restrict__ dst,
                                                       size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

}  // namespace

int launch_to_float(nicam_ctx c, const double *src, float *dst, size_t n)
{
    if (n == 0) return NICAM_OK;
    const size_t nb = (n + 255) / 256;
    if (nb > 0x7fffffffUL) return fail(NICAM_ERR_ARG, "array too large for one launch");
    hipLaunchKernelGGL(to_float_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, src, dst, n);
    NICAM_HIP(hipGetLastError());
    return NICAM_OK;
}

}  // namespace nicam
