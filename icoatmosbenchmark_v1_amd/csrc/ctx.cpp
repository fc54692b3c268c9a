// Context, set-up calls, pointer-mode staging and the exported C entry points.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include <type_traits>

#include "internal.h"

namespace nicam {

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory) return fail(NICAM_ERR_NOMEM, m);
    return fail(NICAM_ERR_HIP, m);
}

static int dev_alloc(nicam_ctx c, double **p, size_t n)
{
    if (*p) return NICAM_OK;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    *p = static_cast<double *>(q);
    c->device_bytes += n * sizeof(double);
    return NICAM_OK;
}

static int dev_alloc_f(nicam_ctx c, float **p, size_t n)
{
    if (*p) return NICAM_OK;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    *p = static_cast<float *>(q);
    c->device_bytes += n * sizeof(float);
    return NICAM_OK;
}

static void dev_free(double *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Copies n doubles from a host or device pointer (unified addressing decides).
static int copy_in(nicam_ctx c, double *dst, const double *src, size_t n)
{
    NICAM_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDefault, c->stream));
    return NICAM_OK;
}

// One argument of an entry point in NICAM_PTR_HOST mode.
struct HostArg {
    const void *host;  // may be null (absent pole array)
    size_t bytes;
    bool out;          // copied back after the call
};

// Maps host arguments onto per-slot device staging buffers; every argument is
// copied in (outputs too, since the kernels leave parts of them untouched).
static int stage_in(nicam_ctx c, const std::vector<HostArg> &args, std::vector<void *> &dev)
{
    if (c->stage.size() < args.size()) {
        c->stage.resize(args.size(), nullptr);
        c->stage_bytes.resize(args.size(), 0);
    }
    dev.assign(args.size(), nullptr);
    for (size_t i = 0; i < args.size(); ++i) {
        if (!args[i].host) continue;
        if (c->stage_bytes[i] < args[i].bytes) {
            if (c->stage[i]) {
                (void)hipFree(c->stage[i]);
                c->device_bytes -= c->stage_bytes[i];
            }
            c->stage[i] = nullptr;
            c->stage_bytes[i] = 0;
            NICAM_HIP(hipMalloc(&c->stage[i], args[i].bytes));
            c->stage_bytes[i] = args[i].bytes;
            c->device_bytes += args[i].bytes;
        }
        dev[i] = c->stage[i];
        NICAM_HIP(hipMemcpyAsync(dev[i], args[i].host, args[i].bytes, hipMemcpyHostToDevice, c->stream));
    }
    return NICAM_OK;
}

static int stage_out(nicam_ctx c, const std::vector<HostArg> &args, const std::vector<void *> &dev)
{
    for (size_t i = 0; i < args.size(); ++i)
        if (args[i].host && args[i].out)
            NICAM_HIP(hipMemcpyAsync(const_cast<void *>(args[i].host), dev[i], args[i].bytes,
                                     hipMemcpyDeviceToHost, c->stream));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    return NICAM_OK;
}

static int check_ctx(nicam_ctx c)
{
    if (!c) return fail(NICAM_ERR_ARG, "null context");
    NICAM_HIP(hipSetDevice(c->device));
    return NICAM_OK;
}

static int need(const void *p, const char *name)
{
    if (!p) return fail(NICAM_ERR_ARG, std::string("null pointer argument: ") + name);
    return NICAM_OK;
}

#define NICAM_TRY(x)                  \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != NICAM_OK) return rc_; \
    } while (0)

}  // namespace nicam

using namespace nicam;

extern "C" {

const char *nicam_version(void) { return "nicam_tt 0.1 gfx950"; }

const char *nicam_last_error(void) { return nicam::g_last_error.c_str(); }

int nicam_ctx_create(nicam_ctx *out, int gall_1d, int kall, int lall, int gall_pl, int have_pl,
                     const int *have_sgp)
{
    if (!out) return fail(NICAM_ERR_ARG, "null output handle");
    *out = nullptr;
    if (gall_1d < 4 || kall < 4 || lall < 1 || gall_pl < 3 || gall_pl > 15)
        return fail(NICAM_ERR_ARG, "sizes out of range: need gall_1d>=4, kall>=4, lall>=1, 3<=gall_pl<=15");
    if ((size_t)gall_1d * gall_1d * kall * lall * 12 > ((size_t)1 << 40))
        return fail(NICAM_ERR_ARG, "grid too large");
    if (!have_sgp) return fail(NICAM_ERR_ARG, "null have_sgp");
    auto *c = new (std::nothrow) nicam_ctx_s();
    if (!c) return fail(NICAM_ERR_NOMEM, "host allocation failed");
    c->d.gall_1d = gall_1d;
    c->d.gall = gall_1d * gall_1d;
    c->d.kall = kall;
    c->d.lall = lall;
    c->d.gall_pl = gall_pl;
    c->d.have_pl = have_pl ? 1 : 0;
    c->have_sgp_h.assign(have_sgp, have_sgp + lall);
    for (int &v : c->have_sgp_h) v = v ? 1 : 0;
    c->eps = DBL_EPSILON;
    c->big = DBL_MAX;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipMalloc(&c->have_sgp, sizeof(int) * lall);
    if (e == hipSuccess)
        e = hipMemcpy(c->have_sgp, c->have_sgp_h.data(), sizeof(int) * lall, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "context set-up");
        nicam_ctx_destroy(c);
        return rc;
    }
    const int rc_s = mass_scratch_init(c->device);
    if (rc_s != NICAM_OK) {
        nicam_ctx_destroy(c);
        return rc_s;
    }
    *out = c;
    return NICAM_OK;
}

int nicam_ctx_destroy(nicam_ctx c)
{
    if (!c) return NICAM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    double **all[] = {&c->GRD_xr, &c->GRD_xr_pl, &c->GMTR_p, &c->GMTR_p_pl, &c->GMTR_t,
                      &c->GMTR_t_pl, &c->GMTR_a, &c->GMTR_a_pl, &c->RGAM, &c->RGAM_pl,
                      &c->RGAMH, &c->RGAMH_pl, &c->RGSGAM2, &c->RGSGAM2_pl, &c->RGSGAM2H,
                      &c->RGSGAM2H_pl, &c->GSGAM2H, &c->GSGAM2H_pl, &c->rdgzh, &c->afact,
                      &c->bfact, &c->Mc, &c->Mc_pl, &c->Ml, &c->Ml_pl, &c->Mu, &c->Mu_pl, &c->vi_beta,
                      &c->vi_beta_pl, &c->vi_gam, &c->vi_gam_pl, &c->vi_tbeta, &c->vi_tbeta_pl, &c->vi_tgam,
                      &c->vi_tgam_pl,
                      &c->RGSQRTH, &c->RGSQRTH_pl, &c->C2WfactGz, &c->C2WfactGz_pl, &c->rdgz,
                      &c->cfact, &c->dfact};
    for (double **p : all) dev_free(*p);
    float *f32[] = {c->hm32.xr, c->hm32.gp, c->hm32.gt, c->hm32.ga,
                    c->hm32.xr_pl, c->hm32.gp_pl, c->hm32.gt_pl, c->hm32.ga_pl,
                    c->RGAM32, c->RGAM32_pl, c->RGAMH32, c->RGAMH32_pl, c->RGSGAM2_32, c->RGSGAM2_32_pl};
    for (float *p : f32)
        if (p) (void)hipFree(p);
    for (void *p : c->stage)
        if (p) (void)hipFree(p);
    if (c->have_sgp) (void)hipFree(c->have_sgp);
    if (c->halo_dst) (void)hipFree(c->halo_dst);
    if (c->halo_src) (void)hipFree(c->halo_src);
    for (int *p : {c->arc_slot, c->arc_tab, c->arc_pole, c->hadj_src, c->hadj_off, c->hadj_dst})
        if (p) (void)hipFree(p);
    if (c->mass_part) (void)hipFree(c->mass_part);
    delete c;
    return NICAM_OK;
}

int nicam_ctx_set_stream(nicam_ctx c, void *s)
{
    NICAM_TRY(check_ctx(c));
    c->stream = static_cast<hipStream_t>(s);
    return NICAM_OK;
}

int nicam_ctx_set_pointer_mode(nicam_ctx c, int mode)
{
    NICAM_TRY(check_ctx(c));
    if (mode != NICAM_PTR_DEVICE && mode != NICAM_PTR_HOST)
        return fail(NICAM_ERR_ARG, "pointer mode must be NICAM_PTR_DEVICE or NICAM_PTR_HOST");
    c->ptr_mode = mode;
    return NICAM_OK;
}

int nicam_ctx_set_tuning(nicam_ctx c, int param, int value)
{
    NICAM_TRY(check_ctx(c));
    if (value < 0) return fail(NICAM_ERR_ARG, "tuning values must be >= 0");
    switch (param) {
    case NICAM_TUNE_HFLUX_KPER: c->tune.hflux_kper = value; break;
    case NICAM_TUNE_VLIM_KCHUNK: c->tune.vlim_kchunk = value; break;
    case NICAM_TUNE_DIFF_KPER: c->tune.diff_kper = value; break;
    case NICAM_TUNE_DDMP_KPER: c->tune.ddmp_kper = value; break;
    case NICAM_TUNE_VI_VARIANT:
        if (value > 4) return fail(NICAM_ERR_ARG, "NICAM_TUNE_VI_VARIANT must be 0..4");
        c->tune.vi_variant = value;
        break;
    case NICAM_TUNE_HLIM_VARIANT:
        if (value > 2) return fail(NICAM_ERR_ARG, "NICAM_TUNE_HLIM_VARIANT must be 0..2");
        c->tune.hlim_variant = value;
        break;
    default: return fail(NICAM_ERR_ARG, "unknown tuning parameter");
    }
    return NICAM_OK;
}

size_t nicam_ctx_device_bytes(nicam_ctx c) { return c ? c->device_bytes : 0; }

int nicam_const_setup(nicam_ctx c)
{
    NICAM_TRY(check_ctx(c));
    c->eps = DBL_EPSILON;
    c->big = DBL_MAX;
    return NICAM_OK;
}

int nicam_set_const(nicam_ctx c, double eps, double huge_val)
{
    NICAM_TRY(check_ctx(c));
    if (!(eps >= 0.0) || !(huge_val > 0.0)) return fail(NICAM_ERR_ARG, "eps must be >= 0 and huge > 0");
    c->eps = eps;
    c->big = huge_val;
    return NICAM_OK;
}

int nicam_set_const_f32(nicam_ctx c, float eps, float huge_val)
{
    NICAM_TRY(check_ctx(c));
    if (!(eps >= 0.0f) || !(huge_val > 0.0f)) return fail(NICAM_ERR_ARG, "eps must be >= 0 and huge > 0");
    c->eps32 = eps;
    c->big32 = huge_val;
    return NICAM_OK;
}

int nicam_set_phys_const(nicam_ctx c, double grav, double rdry, double cvdry, int non_hydro_alpha)
{
    NICAM_TRY(check_ctx(c));
    if (!(rdry != 0.0)) return fail(NICAM_ERR_ARG, "Rdry must be nonzero");
    c->grav = grav;
    c->rdry = rdry;
    c->cvdry = cvdry;
    c->alpha = non_hydro_alpha;
    return NICAM_OK;
}

// ---- vertical grid ---------------------------------------------------------

static uint32_t be32(const unsigned char *p)
{
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

static double be64d(const unsigned char *p)
{
    uint64_t u = 0;
    for (int i = 0; i < 8; ++i) u = u << 8 | p[i];
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

// One Fortran sequential unformatted record with 4-byte big-endian markers.
static int read_record(FILE *f, std::vector<unsigned char> &buf, const char *fname)
{
    unsigned char m[4], m2[4];
    if (std::fread(m, 1, 4, f) != 4) return fail(NICAM_ERR_IO, std::string("short vgrid file: ") + fname);
    uint32_t n = be32(m);
    if (n > (1u << 26)) return fail(NICAM_ERR_IO, std::string("bad record marker in ") + fname);
    buf.resize(n);
    if (std::fread(buf.data(), 1, n, f) != n || std::fread(m2, 1, 4, f) != 4 || be32(m2) != n)
        return fail(NICAM_ERR_IO, std::string("truncated record in ") + fname);
    return NICAM_OK;
}

int nicam_grd_input_vgrid(const char *fname, int kall, double *gz, double *gzh)
{
    if (!fname || !gz || !gzh) return fail(NICAM_ERR_ARG, "null argument");
    if (kall < 4) return fail(NICAM_ERR_ARG, "kall must be >= 4");
    FILE *f = std::fopen(fname, "rb");
    if (!f) return fail(NICAM_ERR_IO, std::string("No vertical grid file : ") + fname);
    std::vector<unsigned char> r;
    int rc = read_record(f, r, fname);
    if (rc == NICAM_OK && r.size() != 4) rc = fail(NICAM_ERR_IO, "first vgrid record is not one int32");
    if (rc == NICAM_OK && (int)be32(r.data()) != kall - 2)
        rc = fail(NICAM_ERR_IO, "inconsistency in number of vertical layers");
    for (int rec = 0; rec < 2 && rc == NICAM_OK; ++rec) {
        rc = read_record(f, r, fname);
        if (rc == NICAM_OK && r.size() != (size_t)kall * 8) rc = fail(NICAM_ERR_IO, "vgrid record size != 8*kall");
        if (rc == NICAM_OK)
            for (int k = 0; k < kall; ++k) (rec == 0 ? gz : gzh)[k] = be64d(&r[(size_t)k * 8]);
    }
    std::fclose(f);
    return rc;
}

int nicam_grd_setup(nicam_ctx c, const double *gz, const double *gzh, double *rdgzh_out,
                    double *afact_out, double *bfact_out)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(need(gz, "gz"));
    NICAM_TRY(need(gzh, "gzh"));
    const int kall = c->d.kall, kmin = 1, kmax = kall - 2;  // 0-based kmin/kmax
    std::vector<double> dgzh(kall), rdgzh(kall), afact(kall), bfact(kall);
    // mod_misc.f90:486-490 and :493-496 (cell-wall intervals and their inverse)
    for (int k = kmin; k <= kmax + 1; ++k) dgzh[k] = gz[k] - gz[k - 1];
    dgzh[kmin - 1] = dgzh[kmin];
    for (int k = 0; k < kall; ++k) rdgzh[k] = 1.0 / dgzh[k];
    // mod_misc.f90:505-510 (cell-centre -> cell-wall interpolation factors)
    for (int k = kmin; k <= kmax + 1; ++k) afact[k] = (gzh[k] - gz[k - 1]) / (gz[k] - gz[k - 1]);
    afact[kmin - 1] = 1.0;
    for (int k = 0; k < kall; ++k) bfact[k] = 1.0 - afact[k];
    // mod_misc.f90:481-484 (cell-centre intervals) for GRD_rdgz
    std::vector<double> dgz(kall), rdgz(kall);
    for (int k = kmin - 1; k <= kmax; ++k) dgz[k] = gzh[k + 1] - gzh[k];
    dgz[kmax + 1] = dgz[kmax];
    for (int k = 0; k < kall; ++k) rdgz[k] = 1.0 / dgz[k];
    // mod_misc.f90:513-521 (cell-wall -> cell-centre factors, for the VI matrix)
    std::vector<double> cfact(kall), dfact(kall);
    for (int k = kmin; k <= kmax; ++k) cfact[k] = (gz[k] - gzh[k]) / (gzh[k + 1] - gzh[k]);
    cfact[kmin - 1] = 1.0;
    cfact[kmax + 1] = 0.0;
    for (int k = 0; k < kall; ++k) dfact[k] = 1.0 - cfact[k];
    for (int k = 0; k < kall; ++k) {
        if (!std::isfinite(rdgzh[k]) || !std::isfinite(afact[k]) || !std::isfinite(rdgz[k]) ||
            !std::isfinite(cfact[k]))
            return fail(NICAM_ERR_ARG, "vertical grid has coincident levels");
    }
    NICAM_TRY(dev_alloc(c, &c->cfact, kall));
    NICAM_TRY(dev_alloc(c, &c->dfact, kall));
    NICAM_TRY(copy_in(c, c->cfact, cfact.data(), kall));
    NICAM_TRY(copy_in(c, c->dfact, dfact.data(), kall));
    c->have_cdfact = true;
    NICAM_TRY(nicam_set_rdgz(c, rdgz.data()));
    if (rdgzh_out) std::memcpy(rdgzh_out, rdgzh.data(), sizeof(double) * kall);
    if (afact_out) std::memcpy(afact_out, afact.data(), sizeof(double) * kall);
    if (bfact_out) std::memcpy(bfact_out, bfact.data(), sizeof(double) * kall);
    return nicam_set_vgrid_factors(c, rdgzh.data(), afact.data(), bfact.data());
}

int nicam_set_rdgz(nicam_ctx c, const double *rdgz)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(need(rdgz, "rdgz"));
    NICAM_TRY(dev_alloc(c, &c->rdgz, c->d.kall));
    NICAM_TRY(copy_in(c, c->rdgz, rdgz, c->d.kall));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_rdgz = true;
    return NICAM_OK;
}

int nicam_set_vgrid_factors(nicam_ctx c, const double *rdgzh, const double *afact, const double *bfact)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(need(rdgzh, "rdgzh"));
    NICAM_TRY(need(afact, "afact"));
    NICAM_TRY(need(bfact, "bfact"));
    const size_t k = c->d.kall;
    NICAM_TRY(dev_alloc(c, &c->rdgzh, k));
    NICAM_TRY(dev_alloc(c, &c->afact, k));
    NICAM_TRY(dev_alloc(c, &c->bfact, k));
    NICAM_TRY(copy_in(c, c->rdgzh, rdgzh, k));
    NICAM_TRY(copy_in(c, c->afact, afact, k));
    NICAM_TRY(copy_in(c, c->bfact, bfact, k));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_vgrid = true;
    return NICAM_OK;
}

// ---- metrics ------------------------------------------------------------------

int nicam_set_hmetrics(nicam_ctx c, const double *xr, const double *xr_pl, const double *gp,
                       const double *gp_pl, const double *gt, const double *gt_pl,
                       const double *ga, const double *ga_pl)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(need(xr, "GRD_xr"));
    NICAM_TRY(need(gp, "GMTR_p"));
    NICAM_TRY(need(gt, "GMTR_t"));
    NICAM_TRY(need(ga, "GMTR_a"));
    const size_t g = (size_t)c->d.gall * c->d.lall, p = (size_t)c->d.gall_pl * kLallPl;
    NICAM_TRY(dev_alloc(c, &c->GRD_xr, g * 9));
    NICAM_TRY(dev_alloc(c, &c->GMTR_p, g * kGmtrPNmax));
    NICAM_TRY(dev_alloc(c, &c->GMTR_t, g * 2 * kGmtrTNmax));
    NICAM_TRY(dev_alloc(c, &c->GMTR_a, g * 3 * kGmtrANmax));
    NICAM_TRY(copy_in(c, c->GRD_xr, xr, g * 9));
    NICAM_TRY(copy_in(c, c->GMTR_p, gp, g * kGmtrPNmax));
    NICAM_TRY(copy_in(c, c->GMTR_t, gt, g * 2 * kGmtrTNmax));
    NICAM_TRY(copy_in(c, c->GMTR_a, ga, g * 3 * kGmtrANmax));
    if (c->d.have_pl) {
        NICAM_TRY(need(xr_pl, "GRD_xr_pl"));
        NICAM_TRY(need(gp_pl, "GMTR_p_pl"));
        NICAM_TRY(need(gt_pl, "GMTR_t_pl"));
        NICAM_TRY(need(ga_pl, "GMTR_a_pl"));
        NICAM_TRY(dev_alloc(c, &c->GRD_xr_pl, p * 3));
        NICAM_TRY(dev_alloc(c, &c->GMTR_p_pl, p * kGmtrPNmax));
        NICAM_TRY(dev_alloc(c, &c->GMTR_t_pl, p * kGmtrTNmax));
        NICAM_TRY(dev_alloc(c, &c->GMTR_a_pl, p * kGmtrANmaxPl));
        NICAM_TRY(copy_in(c, c->GRD_xr_pl, xr_pl, p * 3));
        NICAM_TRY(copy_in(c, c->GMTR_p_pl, gp_pl, p * kGmtrPNmax));
        NICAM_TRY(copy_in(c, c->GMTR_t_pl, gt_pl, p * kGmtrTNmax));
        NICAM_TRY(copy_in(c, c->GMTR_a_pl, ga_pl, p * kGmtrANmaxPl));
    }
    // FP32 copies for the _f32 entry points
    nicam::HMetrics<float> &h = c->hm32;
    NICAM_TRY(dev_alloc_f(c, &h.xr, g * 9));
    NICAM_TRY(dev_alloc_f(c, &h.gp, g * kGmtrPNmax));
    NICAM_TRY(dev_alloc_f(c, &h.gt, g * 2 * kGmtrTNmax));
    NICAM_TRY(dev_alloc_f(c, &h.ga, g * 3 * kGmtrANmax));
    NICAM_TRY(launch_to_float(c, c->GRD_xr, h.xr, g * 9));
    NICAM_TRY(launch_to_float(c, c->GMTR_p, h.gp, g * kGmtrPNmax));
    NICAM_TRY(launch_to_float(c, c->GMTR_t, h.gt, g * 2 * kGmtrTNmax));
    NICAM_TRY(launch_to_float(c, c->GMTR_a, h.ga, g * 3 * kGmtrANmax));
    if (c->d.have_pl) {
        NICAM_TRY(dev_alloc_f(c, &h.xr_pl, p * 3));
        NICAM_TRY(dev_alloc_f(c, &h.gp_pl, p * kGmtrPNmax));
        NICAM_TRY(dev_alloc_f(c, &h.gt_pl, p * kGmtrTNmax));
        NICAM_TRY(dev_alloc_f(c, &h.ga_pl, p * kGmtrANmaxPl));
        NICAM_TRY(launch_to_float(c, c->GRD_xr_pl, h.xr_pl, p * 3));
        NICAM_TRY(launch_to_float(c, c->GMTR_p_pl, h.gp_pl, p * kGmtrPNmax));
        NICAM_TRY(launch_to_float(c, c->GMTR_t_pl, h.gt_pl, p * kGmtrTNmax));
        NICAM_TRY(launch_to_float(c, c->GMTR_a_pl, h.ga_pl, p * kGmtrANmaxPl));
    }
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_hmetrics = true;
    return NICAM_OK;
}

static int set_pair(nicam_ctx c, double **dst, double **dst_pl, const double *src, const double *src_pl,
                    const char *name)
{
    NICAM_TRY(need(src, name));
    NICAM_TRY(dev_alloc(c, dst, c->d.nreg()));
    NICAM_TRY(copy_in(c, *dst, src, c->d.nreg()));
    if (c->d.have_pl) {
        NICAM_TRY(need(src_pl, name));
        NICAM_TRY(dev_alloc(c, dst_pl, c->d.npl()));
        NICAM_TRY(copy_in(c, *dst_pl, src_pl, c->d.npl()));
    }
    return NICAM_OK;
}

// FP32 copies of the metrics that enter only the solver's source term, kept
// in step with the FP64 ones (both setters write RGAM and RGAMH).
static int refresh_vmetrics32(nicam_ctx c)
{
    const size_t R = c->d.nreg(), P = c->d.npl();
    struct {
        double *src, *src_pl;
        float **dst, **dst_pl;
    } m[] = {{c->RGAM, c->RGAM_pl, &c->RGAM32, &c->RGAM32_pl},
             {c->RGAMH, c->RGAMH_pl, &c->RGAMH32, &c->RGAMH32_pl},
             {c->RGSGAM2, c->RGSGAM2_pl, &c->RGSGAM2_32, &c->RGSGAM2_32_pl}};
    for (auto &e : m) {
        if (!e.src) continue;
        NICAM_TRY(dev_alloc_f(c, e.dst, R));
        NICAM_TRY(launch_to_float(c, e.src, *e.dst, R));
        if (c->d.have_pl && e.src_pl) {
            NICAM_TRY(dev_alloc_f(c, e.dst_pl, P));
            NICAM_TRY(launch_to_float(c, e.src_pl, *e.dst_pl, P));
        }
    }
    return NICAM_OK;
}

int nicam_set_vmetrics(nicam_ctx c, const double *RGAM, const double *RGAM_pl, const double *RGAMH,
                       const double *RGAMH_pl, const double *RGSGAM2, const double *RGSGAM2_pl,
                       const double *RGSGAM2H, const double *RGSGAM2H_pl, const double *GSGAM2H,
                       const double *GSGAM2H_pl)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(set_pair(c, &c->RGAM, &c->RGAM_pl, RGAM, RGAM_pl, "VMTR_RGAM"));
    NICAM_TRY(set_pair(c, &c->RGAMH, &c->RGAMH_pl, RGAMH, RGAMH_pl, "VMTR_RGAMH"));
    NICAM_TRY(set_pair(c, &c->RGSGAM2, &c->RGSGAM2_pl, RGSGAM2, RGSGAM2_pl, "VMTR_RGSGAM2"));
    NICAM_TRY(set_pair(c, &c->RGSGAM2H, &c->RGSGAM2H_pl, RGSGAM2H, RGSGAM2H_pl, "VMTR_RGSGAM2H"));
    NICAM_TRY(set_pair(c, &c->GSGAM2H, &c->GSGAM2H_pl, GSGAM2H, GSGAM2H_pl, "VMTR_GSGAM2H"));
    NICAM_TRY(refresh_vmetrics32(c));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_vmetrics = true;
    return NICAM_OK;
}

int nicam_set_vmetrics3d(nicam_ctx c, const double *RGSQRTH, const double *RGSQRTH_pl, const double *RGAM,
                         const double *RGAM_pl, const double *RGAMH, const double *RGAMH_pl, const double *C2WfactGz,
                         const double *C2WfactGz_pl)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(set_pair(c, &c->RGSQRTH, &c->RGSQRTH_pl, RGSQRTH, RGSQRTH_pl, "VMTR_RGSQRTH"));
    NICAM_TRY(set_pair(c, &c->RGAM, &c->RGAM_pl, RGAM, RGAM_pl, "VMTR_RGAM"));
    NICAM_TRY(set_pair(c, &c->RGAMH, &c->RGAMH_pl, RGAMH, RGAMH_pl, "VMTR_RGAMH"));
    NICAM_TRY(need(C2WfactGz, "VMTR_C2WfactGz"));
    NICAM_TRY(dev_alloc(c, &c->C2WfactGz, c->d.nreg() * 6));
    NICAM_TRY(copy_in(c, c->C2WfactGz, C2WfactGz, c->d.nreg() * 6));
    if (c->d.have_pl) {
        NICAM_TRY(need(C2WfactGz_pl, "VMTR_C2WfactGz_pl"));
        NICAM_TRY(dev_alloc(c, &c->C2WfactGz_pl, c->d.npl() * 6));
        NICAM_TRY(copy_in(c, c->C2WfactGz_pl, C2WfactGz_pl, c->d.npl() * 6));
    }
    NICAM_TRY(refresh_vmetrics32(c));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_vmetrics3d = true;
    return NICAM_OK;
}

// Allocates the factor arrays (set-up only; the solves allocate nothing) and
// factors the coefficients just installed.
static int factor_coef(nicam_ctx c)
{
    double **reg[] = {&c->vi_beta, &c->vi_gam, &c->vi_tbeta, &c->vi_tgam};
    double **pl[] = {&c->vi_beta_pl, &c->vi_gam_pl, &c->vi_tbeta_pl, &c->vi_tgam_pl};
    for (double **p : reg) NICAM_TRY(dev_alloc(c, p, c->d.nreg()));
    if (c->d.have_pl)
        for (double **p : pl) NICAM_TRY(dev_alloc(c, p, c->d.npl()));
    return launch_vi_factor(c);
}

int nicam_vi_rhow_solver_putcoef(nicam_ctx c, const double *Mc, const double *Mc_pl, const double *Ml,
                                 const double *Ml_pl, const double *Mu, const double *Mu_pl)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(set_pair(c, &c->Mc, &c->Mc_pl, Mc, Mc_pl, "Mc"));
    NICAM_TRY(set_pair(c, &c->Ml, &c->Ml_pl, Ml, Ml_pl, "Ml"));
    NICAM_TRY(set_pair(c, &c->Mu, &c->Mu_pl, Mu, Mu_pl, "Mu"));
    NICAM_TRY(factor_coef(c));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_coef = true;
    return NICAM_OK;
}

int nicam_ctx_synchronize(nicam_ctx c)
{
    NICAM_TRY(check_ctx(c));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    return NICAM_OK;
}

int nicam_comm_setup(nicam_ctx c, int rlevel)
{
    NICAM_TRY(check_ctx(c));
    return comm_setup(c, rlevel);
}

// COMM_data_transfer / COMM_var on var(gall,kmax,lall,vmax), var_pl(gall_pl,kmax,2,vmax)
static int halo_entry(nicam_ctx c, double *var, double *var_pl, int kmax, int vmax, bool pole_fill)
{
    NICAM_TRY(check_ctx(c));
    if (kmax < 1 || vmax < 1) return fail(NICAM_ERR_ARG, "kmax and vmax must be positive");
    NICAM_TRY(need(var, "var"));
    NICAM_TRY(need(var_pl, "var_pl"));
    if (!c->halo_n) return fail(NICAM_ERR_STATE, "halo map not set (nicam_comm_setup)");
    if (c->ptr_mode == NICAM_PTR_DEVICE) return launch_halo_fill(c, var, var_pl, kmax, vmax, pole_fill);
    const size_t n = (size_t)kmax * vmax * 8;
    const size_t R = (size_t)c->d.gall * c->d.lall * n, P = (size_t)c->d.gall_pl * kLallPl * n;
    std::vector<HostArg> a = {{var, R, true}, {var_pl, P, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    NICAM_TRY(launch_halo_fill(c, static_cast<double *>(v[0]), static_cast<double *>(v[1]), kmax, vmax, pole_fill));
    return stage_out(c, a, v);
}

int nicam_comm_data_transfer(nicam_ctx c, double *var, double *var_pl, int kmax, int vmax)
{
    return halo_entry(c, var, var_pl, kmax, vmax, false);
}

int nicam_comm_data_transfer_adjoint(nicam_ctx c, double *var_b, double *var_b_pl, int kmax, int vmax)
{
    NICAM_TRY(check_ctx(c));
    if (kmax < 1 || vmax < 1) return fail(NICAM_ERR_ARG, "kmax and vmax must be positive");
    NICAM_TRY(need(var_b, "var_b"));
    NICAM_TRY(need(var_b_pl, "var_b_pl"));
    if (!c->hadj_nsrc) return fail(NICAM_ERR_STATE, "halo map not set (nicam_comm_setup)");
    if (c->ptr_mode == NICAM_PTR_DEVICE) return launch_halo_adjoint(c, var_b, var_b_pl, kmax, vmax);
    const size_t n = (size_t)kmax * vmax * 8;
    const size_t R = (size_t)c->d.gall * c->d.lall * n, P = (size_t)c->d.gall_pl * kLallPl * n;
    std::vector<HostArg> a = {{var_b, R, true}, {var_b_pl, P, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    NICAM_TRY(launch_halo_adjoint(c, static_cast<double *>(v[0]), static_cast<double *>(v[1]), kmax, vmax));
    return stage_out(c, a, v);
}

int nicam_comm_var(nicam_ctx c, double *var, double *var_pl, int kmax, int vmax)
{
    return halo_entry(c, var, var_pl, kmax, vmax, true);
}

// ---- hot path entry points ---------------------------------------------------

#define PL(p) (c->d.have_pl ? (p) : nullptr)

int nicam_horizontal_flux(nicam_ctx c, double *flx_h, double *flx_h_pl, double *xc, double *xc_pl,
                          const double *rho, const double *rho_pl, const double *rvx,
                          const double *rvx_pl, const double *rvy, const double *rvy_pl,
                          const double *rvz, const double *rvz_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_hmetrics) return fail(NICAM_ERR_STATE, "horizontal metrics not set (nicam_set_hmetrics)");
    const void *req[] = {flx_h, xc, rho, rvx, rvy, rvz};
    const char *nm[] = {"flx_h", "GRD_xc", "rho", "rhovx", "rhovy", "rhovz"};
    for (int i = 0; i < 6; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {flx_h_pl, xc_pl, rho_pl, rvx_pl, rvy_pl, rvz_pl};
        for (int i = 0; i < 6; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_hflux(c, flx_h, PL(flx_h_pl), xc, PL(xc_pl), rho, PL(rho_pl), rvx, PL(rvx_pl),
                            rvy, PL(rvy_pl), rvz, PL(rvz_pl), dt);
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{flx_h, R * 6, true},  {PL(flx_h_pl), P, true}, {xc, R * 9, true},
                              {PL(xc_pl), P * 3, true}, {rho, R, false},       {PL(rho_pl), P, false},
                              {rvx, R, false},         {PL(rvx_pl), P, false}, {rvy, R, false},
                              {PL(rvy_pl), P, false},  {rvz, R, false},        {PL(rvz_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_hflux(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11), dt));
    return stage_out(c, a, v);
}

int nicam_horizontal_limiter_thuburn(nicam_ctx c, double *q_a, double *q_a_pl, const double *q,
                                     const double *q_pl, const double *d, const double *d_pl,
                                     const double *ch, const double *ch_pl, const double *cmask,
                                     const double *cmask_pl, double *Qprev, double *Qprev_pl,
                                     const double *Qpost, const double *Qpost_pl)
{
    NICAM_TRY(check_ctx(c));
    // Qout_post NULL: the library exchanges Qout_prev itself (nicam_comm_setup
    // must have run); Qout_post_pl is then ignored.
    const void *req[] = {q_a, q, d, ch, cmask, Qprev};
    const char *nm[] = {"q_a", "q", "d", "ch", "cmask", "Qout_prev"};
    for (int i = 0; i < 6; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_a_pl, q_pl, d_pl, ch_pl, cmask_pl, Qprev_pl};
        for (int i = 0; i < 6; ++i) NICAM_TRY(need(rp[i], "pole array"));
        if (Qpost) NICAM_TRY(need(Qpost_pl, "Qout_post_pl"));
    }
    if (!Qpost) Qpost_pl = nullptr;
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_hlim(c, q_a, PL(q_a_pl), q, PL(q_pl), d, PL(d_pl), ch, PL(ch_pl), cmask,
                           PL(cmask_pl), Qprev, Qprev_pl, Qpost, PL(Qpost_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{q_a, R * 6, true},       {PL(q_a_pl), P, true},   {q, R, false},
                              {PL(q_pl), P, false},     {d, R, false},           {PL(d_pl), P, false},
                              {ch, R * 6, false},       {PL(ch_pl), P, false},   {cmask, R * 6, false},
                              {PL(cmask_pl), P, false}, {Qprev, R * 2, true},    {Qprev_pl, P * 2, true},
                              {Qpost, R * 2, false},    {PL(Qpost_pl), P * 2, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_hlim(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11),
                          D(12), D(13)));
    return stage_out(c, a, v);
}

int nicam_vertical_limiter_thuburn(nicam_ctx c, double *q_h, double *q_h_pl, const double *q,
                                   const double *q_pl, const double *d, const double *d_pl,
                                   const double *ck, const double *ck_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {q_h, q, d, ck};
    const char *nm[] = {"q_h", "q", "d", "ck"};
    for (int i = 0; i < 4; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_h_pl, q_pl, d_pl, ck_pl};
        for (int i = 0; i < 4; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vlim(c, q_h, PL(q_h_pl), q, PL(q_pl), d, PL(d_pl), ck, PL(ck_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{q_h, R, true},  {PL(q_h_pl), P, true}, {q, R, false},      {PL(q_pl), P, false},
                              {d, R, false},   {PL(d_pl), P, false},  {ck, R * 2, false}, {PL(ck_pl), P * 2, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vlim(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7)));
    return stage_out(c, a, v);
}

int nicam_oprt_diffusion(nicam_ctx c, double *dscl, double *dscl_pl, const double *scl, const double *scl_pl,
                         const double *kh, const double *kh_pl, const double *coef_intp, const double *coef_intp_pl,
                         const double *coef_diff, const double *coef_diff_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {dscl, scl, kh, coef_intp, coef_diff};
    const char *nm[] = {"dscl", "scl", "kh", "coef_intp", "coef_diff"};
    for (int i = 0; i < 5; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {dscl_pl, scl_pl, kh_pl, coef_intp_pl, coef_diff_pl};
        for (int i = 0; i < 5; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    // dscl_pl is zeroed without pole regions, as the reference does, if given
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_diffusion(c, dscl, dscl_pl, scl, PL(scl_pl), kh, PL(kh_pl), coef_intp, PL(coef_intp_pl),
                                coef_diff, PL(coef_diff_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    const size_t CI = (size_t)c->d.gall * c->d.lall * 18 * 8, CIP = (size_t)c->d.gall_pl * 18 * 8;
    const size_t CD = (size_t)c->d.gall * c->d.lall * 18 * 8, CDP = (size_t)5 * 3 * 2 * 8;
    std::vector<HostArg> a = {{dscl, R, true},         {dscl_pl, P, true},      {scl, R, false},
                              {PL(scl_pl), P, false},  {kh, R, false},          {PL(kh_pl), P, false},
                              {coef_intp, CI, false},  {PL(coef_intp_pl), CIP, false},
                              {coef_diff, CD, false},  {PL(coef_diff_pl), CDP, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_diffusion(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9)));
    return stage_out(c, a, v);
}

int nicam_oprt3d_divdamp(nicam_ctx c, double *ddivdx, double *ddivdx_pl, double *ddivdy, double *ddivdy_pl,
                         double *ddivdz, double *ddivdz_pl, const double *rhogvx, const double *rhogvx_pl,
                         const double *rhogvy, const double *rhogvy_pl, const double *rhogvz,
                         const double *rhogvz_pl, const double *rhogw, const double *rhogw_pl,
                         const double *coef_intp, const double *coef_intp_pl, const double *coef_diff,
                         const double *coef_diff_pl)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_vmetrics3d) return fail(NICAM_ERR_STATE, "vertical metrics not set (nicam_set_vmetrics3d)");
    if (!c->have_rdgz) return fail(NICAM_ERR_STATE, "GRD_rdgz not set (nicam_grd_setup or nicam_set_rdgz)");
    const void *req[] = {ddivdx, ddivdy, ddivdz, rhogvx, rhogvy, rhogvz, rhogw, coef_intp, coef_diff};
    const char *nm[] = {"ddivdx", "ddivdy", "ddivdz", "rhogvx", "rhogvy", "rhogvz", "rhogw", "coef_intp", "coef_diff"};
    for (int i = 0; i < 9; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {ddivdx_pl, ddivdy_pl, ddivdz_pl, rhogvx_pl, rhogvy_pl, rhogvz_pl, rhogw_pl,
                            coef_intp_pl, coef_diff_pl};
        for (int i = 0; i < 9; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if ((ddivdx_pl == nullptr) != (ddivdy_pl == nullptr) || (ddivdx_pl == nullptr) != (ddivdz_pl == nullptr))
        return fail(NICAM_ERR_ARG, "give all three or none of ddivdx_pl, ddivdy_pl, ddivdz_pl");
    auto run = [&](double *const *o, double *const *opl, const double *const *v, const double *const *vpl,
                   const double *w, const double *wpl, const double *ci, const double *cipl, const double *cd,
                   const double *cdpl) {
        return launch_divdamp(c, o, opl, v, vpl, w, wpl, ci, cipl, cd, cdpl, c->RGSQRTH, c->RGSQRTH_pl, c->RGAM,
                              c->RGAM_pl, c->RGAMH, c->RGAMH_pl, c->C2WfactGz, c->C2WfactGz_pl, c->rdgz);
    };
    if (c->ptr_mode == NICAM_PTR_DEVICE) {
        double *o[3] = {ddivdx, ddivdy, ddivdz}, *opl[3] = {ddivdx_pl, ddivdy_pl, ddivdz_pl};
        const double *v[3] = {rhogvx, rhogvy, rhogvz}, *vpl[3] = {PL(rhogvx_pl), PL(rhogvy_pl), PL(rhogvz_pl)};
        return run(o, opl, v, vpl, rhogw, PL(rhogw_pl), coef_intp, PL(coef_intp_pl), coef_diff, PL(coef_diff_pl));
    }
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    const size_t CI = (size_t)c->d.gall * c->d.lall * 18 * 8, CIP = (size_t)c->d.gall_pl * 18 * 8;
    const size_t CD = CI, CDP = (size_t)5 * 3 * 2 * 8;
    std::vector<HostArg> a = {{ddivdx, R, true},          {ddivdx_pl, P, true},           {ddivdy, R, true},
                              {ddivdy_pl, P, true},       {ddivdz, R, true},              {ddivdz_pl, P, true},
                              {rhogvx, R, false},         {PL(rhogvx_pl), P, false},      {rhogvy, R, false},
                              {PL(rhogvy_pl), P, false},  {rhogvz, R, false},             {PL(rhogvz_pl), P, false},
                              {rhogw, R, false},          {PL(rhogw_pl), P, false},       {coef_intp, CI, false},
                              {PL(coef_intp_pl), CIP, false}, {coef_diff, CD, false},     {PL(coef_diff_pl), CDP, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    double *o[3] = {D(0), D(2), D(4)}, *opl[3] = {D(1), D(3), D(5)};
    const double *vv[3] = {D(6), D(8), D(10)}, *vpl[3] = {D(7), D(9), D(11)};
    NICAM_TRY(run(o, opl, vv, vpl, D(12), D(13), D(14), D(15), D(16), D(17)));
    return stage_out(c, a, v);
}

int nicam_vi_rhow_solver(nicam_ctx c, double *w, double *w_pl, const double *w0, const double *w0_pl,
                         const double *p0, const double *p0_pl, const double *r0, const double *r0_pl,
                         const double *Sr, const double *Sr_pl, const double *Sw, const double *Sw_pl,
                         const double *Sp, const double *Sp_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_coef) return fail(NICAM_ERR_STATE, "tridiagonal coefficients not set (nicam_vi_rhow_solver_putcoef)");
    if (!c->have_vmetrics) return fail(NICAM_ERR_STATE, "vertical metrics not set (nicam_set_vmetrics)");
    if (!c->have_vgrid) return fail(NICAM_ERR_STATE, "vertical grid not set (nicam_grd_setup)");
    if (!(dt != 0.0)) return fail(NICAM_ERR_ARG, "dt must be nonzero");
    const void *req[] = {w, w0, p0, r0, Sr, Sw, Sp};
    const char *nm[] = {"rhogw", "rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre"};
    for (int i = 0; i < 7; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {w_pl, w0_pl, p0_pl, r0_pl, Sr_pl, Sw_pl, Sp_pl};
        for (int i = 0; i < 7; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vi(c, w, PL(w_pl), w0, PL(w0_pl), p0, PL(p0_pl), r0, PL(r0_pl), Sr, PL(Sr_pl), Sw,
                         PL(Sw_pl), Sp, PL(Sp_pl), dt);
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{w, R, true},  {PL(w_pl), P, true},   {w0, R, false}, {PL(w0_pl), P, false},
                              {p0, R, false}, {PL(p0_pl), P, false}, {r0, R, false}, {PL(r0_pl), P, false},
                              {Sr, R, false}, {PL(Sr_pl), P, false}, {Sw, R, false}, {PL(Sw_pl), P, false},
                              {Sp, R, false}, {PL(Sp_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vi(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11), D(12),
                        D(13), dt));
    return stage_out(c, a, v);
}

int nicam_vi_rhow_solver_mixed(nicam_ctx c, double *w, double *w_pl, const float *w0, const float *w0_pl,
                               const float *p0, const float *p0_pl, const float *r0, const float *r0_pl,
                               const float *Sr, const float *Sr_pl, const float *Sw, const float *Sw_pl,
                               const float *Sp, const float *Sp_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_coef) return fail(NICAM_ERR_STATE, "tridiagonal coefficients not set (nicam_vi_rhow_solver_putcoef)");
    if (!c->have_vmetrics) return fail(NICAM_ERR_STATE, "vertical metrics not set (nicam_set_vmetrics)");
    if (!c->have_vgrid) return fail(NICAM_ERR_STATE, "vertical grid not set (nicam_grd_setup)");
    if (!(dt != 0.0)) return fail(NICAM_ERR_ARG, "dt must be nonzero");
    const void *req[] = {w, w0, p0, r0, Sr, Sw, Sp};
    const char *nm[] = {"rhogw", "rhogw0", "preg0", "rhog0", "Srho", "Sw", "Spre"};
    for (int i = 0; i < 7; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {w_pl, w0_pl, p0_pl, r0_pl, Sr_pl, Sw_pl, Sp_pl};
        for (int i = 0; i < 7; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vi_mixed(c, w, PL(w_pl), w0, PL(w0_pl), p0, PL(p0_pl), r0, PL(r0_pl), Sr, PL(Sr_pl), Sw,
                               PL(Sw_pl), Sp, PL(Sp_pl), dt);
    const size_t R = c->d.nreg(), P = c->d.npl();
    std::vector<HostArg> a = {{w, R * 8, true},       {PL(w_pl), P * 8, true},   {w0, R * 4, false},
                              {PL(w0_pl), P * 4, false}, {p0, R * 4, false},      {PL(p0_pl), P * 4, false},
                              {r0, R * 4, false},     {PL(r0_pl), P * 4, false}, {Sr, R * 4, false},
                              {PL(Sr_pl), P * 4, false}, {Sw, R * 4, false},      {PL(Sw_pl), P * 4, false},
                              {Sp, R * 4, false},     {PL(Sp_pl), P * 4, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto F = [&](int i) { return static_cast<const float *>(v[i]); };
    NICAM_TRY(launch_vi_mixed(c, static_cast<double *>(v[0]), static_cast<double *>(v[1]), F(2), F(3), F(4), F(5),
                              F(6), F(7), F(8), F(9), F(10), F(11), F(12), F(13), dt));
    return stage_out(c, a, v);
}

// ---- adjoints ----
int nicam_vi_rhow_solver_adjoint(nicam_ctx c, double *wb, double *wb_pl, double *w0b, double *w0b_pl, double *p0b,
                                 double *p0b_pl, double *r0b, double *r0b_pl, double *Srb, double *Srb_pl,
                                 double *Swb, double *Swb_pl, double *Spb, double *Spb_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_coef) return fail(NICAM_ERR_STATE, "tridiagonal coefficients not set (nicam_vi_rhow_solver_putcoef)");
    if (!c->have_vmetrics) return fail(NICAM_ERR_STATE, "vertical metrics not set (nicam_set_vmetrics)");
    if (!c->have_vgrid) return fail(NICAM_ERR_STATE, "vertical grid not set (nicam_grd_setup)");
    if (!(dt != 0.0)) return fail(NICAM_ERR_ARG, "dt must be nonzero");
    double *req[] = {wb, w0b, p0b, r0b, Srb, Swb, Spb};
    const char *nm[] = {"rhogw_b", "rhogw0_b", "preg0_b", "rhog0_b", "Srho_b", "Sw_b", "Spre_b"};
    for (int i = 0; i < 7; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {wb_pl, w0b_pl, p0b_pl, r0b_pl, Srb_pl, Swb_pl, Spb_pl};
        for (int i = 0; i < 7; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vi_adjoint(c, wb, PL(wb_pl), w0b, PL(w0b_pl), p0b, PL(p0b_pl), r0b, PL(r0b_pl), Srb,
                                 PL(Srb_pl), Swb, PL(Swb_pl), Spb, PL(Spb_pl), dt);
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{wb, R, true},  {PL(wb_pl), P, true},  {w0b, R, true}, {PL(w0b_pl), P, true},
                              {p0b, R, true}, {PL(p0b_pl), P, true}, {r0b, R, true}, {PL(r0b_pl), P, true},
                              {Srb, R, true}, {PL(Srb_pl), P, true}, {Swb, R, true}, {PL(Swb_pl), P, true},
                              {Spb, R, true}, {PL(Spb_pl), P, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vi_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11), D(12),
                                D(13), dt));
    return stage_out(c, a, v);
}

int nicam_horizontal_flux_adjoint(nicam_ctx c, const double *flx_h_b, const double *flx_h_b_pl,
                                  const double *GRD_xc_b, const double *GRD_xc_b_pl, double *rho_b, double *rho_b_pl,
                                  double *rhovx_b, double *rhovx_b_pl, double *rhovy_b, double *rhovy_b_pl,
                                  double *rhovz_b, double *rhovz_b_pl, const double *rho, const double *rho_pl,
                                  const double *rhovx, const double *rhovx_pl, const double *rhovy,
                                  const double *rhovy_pl, const double *rhovz, const double *rhovz_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_hmetrics) return fail(NICAM_ERR_STATE, "horizontal metrics not set (nicam_set_hmetrics)");
    const void *req[] = {flx_h_b, GRD_xc_b, rho_b, rhovx_b, rhovy_b, rhovz_b, rho, rhovx, rhovy, rhovz};
    const char *nm[] = {"flx_h_b", "GRD_xc_b", "rho_b", "rhovx_b", "rhovy_b", "rhovz_b", "rho", "rhovx", "rhovy",
                        "rhovz"};
    for (int i = 0; i < 10; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {flx_h_b_pl, GRD_xc_b_pl, rho_b_pl, rhovx_b_pl, rhovy_b_pl, rhovz_b_pl, rho_pl, rhovx_pl,
                            rhovy_pl, rhovz_pl};
        for (int i = 0; i < 10; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE) {
        double *const Xb[4] = {rho_b, rhovx_b, rhovy_b, rhovz_b};
        double *const Xb_pl[4] = {PL(rho_b_pl), PL(rhovx_b_pl), PL(rhovy_b_pl), PL(rhovz_b_pl)};
        const double *const X[4] = {rho, rhovx, rhovy, rhovz};
        const double *const X_pl[4] = {PL(rho_pl), PL(rhovx_pl), PL(rhovy_pl), PL(rhovz_pl)};
        return launch_hflux_adjoint(c, flx_h_b, PL(flx_h_b_pl), GRD_xc_b, PL(GRD_xc_b_pl), Xb, Xb_pl, X, X_pl, dt);
    }
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{flx_h_b, 6 * R, false},  {PL(flx_h_b_pl), P, false},  {GRD_xc_b, 9 * R, false},
                              {PL(GRD_xc_b_pl), 3 * P, false}, {rho_b, R, true},      {PL(rho_b_pl), P, true},
                              {rhovx_b, R, true},       {PL(rhovx_b_pl), P, true},   {rhovy_b, R, true},
                              {PL(rhovy_b_pl), P, true}, {rhovz_b, R, true},         {PL(rhovz_b_pl), P, true},
                              {rho, R, false},          {PL(rho_pl), P, false},      {rhovx, R, false},
                              {PL(rhovx_pl), P, false}, {rhovy, R, false},           {PL(rhovy_pl), P, false},
                              {rhovz, R, false},        {PL(rhovz_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    double *const Xb[4] = {D(4), D(6), D(8), D(10)};
    double *const Xb_pl[4] = {D(5), D(7), D(9), D(11)};
    const double *const X[4] = {D(12), D(14), D(16), D(18)};
    const double *const X_pl[4] = {D(13), D(15), D(17), D(19)};
    NICAM_TRY(launch_hflux_adjoint(c, D(0), D(1), D(2), D(3), Xb, Xb_pl, X, X_pl, dt));
    return stage_out(c, a, v);
}

int nicam_horizontal_limiter_thuburn_mask(nicam_ctx c, double *q_a, double *q_a_pl, const double *q,
                                          const double *q_pl, const double *d, const double *d_pl, const double *ch,
                                          const double *ch_pl, const double *cmask, const double *cmask_pl,
                                          double *Qprev, double *Qprev_pl, const double *Qpost,
                                          const double *Qpost_pl, uint8_t *mask, uint8_t *mask_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {q_a, q, d, ch, cmask, Qprev, mask};
    const char *nm[] = {"q_a", "q", "d", "ch", "cmask", "Qout_prev", "mask"};
    for (int i = 0; i < 7; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_a_pl, q_pl, d_pl, ch_pl, cmask_pl, Qprev_pl, mask_pl};
        for (int i = 0; i < 7; ++i) NICAM_TRY(need(rp[i], "pole array"));
        if (Qpost) NICAM_TRY(need(Qpost_pl, "Qout_post_pl"));
    }
    if (!Qpost) Qpost_pl = nullptr;
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_hlim_mask(c, q_a, PL(q_a_pl), q, PL(q_pl), d, PL(d_pl), ch, PL(ch_pl), cmask, PL(cmask_pl),
                                Qprev, Qprev_pl, Qpost, PL(Qpost_pl), mask, PL(mask_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{q_a, R * 6, true},       {PL(q_a_pl), P, true},       {q, R, false},
                              {PL(q_pl), P, false},     {d, R, false},               {PL(d_pl), P, false},
                              {ch, R * 6, false},       {PL(ch_pl), P, false},       {cmask, R * 6, false},
                              {PL(cmask_pl), P, false}, {Qprev, R * 2, true},        {Qprev_pl, P * 2, true},
                              {Qpost, R * 2, false},    {PL(Qpost_pl), P * 2, false}, {mask, R / 8, true},
                              {PL(mask_pl), P / 8, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_hlim_mask(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11), D(12),
                               D(13), static_cast<uint8_t *>(v[14]), static_cast<uint8_t *>(v[15])));
    return stage_out(c, a, v);
}

int nicam_horizontal_limiter_thuburn_adjoint(nicam_ctx c, double *q_a_b, double *q_a_b_pl, double *q_b,
                                             double *q_b_pl, const double *Qprev_b, const double *Qprev_b_pl,
                                             double *Qpost_b, double *Qpost_b_pl, const double *q_a,
                                             const double *q_a_pl, const double *q, const double *q_pl,
                                             const double *d, const double *d_pl, const double *ch,
                                             const double *ch_pl, const double *cmask, const double *cmask_pl,
                                             const double *Qpost, const double *Qpost_pl, const uint8_t *mask,
                                             const uint8_t *mask_pl, double *ch_b, double *ch_b_pl)
{
    NICAM_TRY(check_ctx(c));
    // the faces run when q_a_b is given, the rows when Qout_prev_b is; ch_b comes from the rows
    if (!q_a_b && !Qprev_b) return fail(NICAM_ERR_ARG, "neither q_a_b nor Qout_prev_b given");
    if (ch_b && !Qprev_b) return fail(NICAM_ERR_ARG, "ch_b needs Qout_prev_b (it comes from the rows)");
    if (ch_b && c->d.have_pl && !ch_b_pl) return fail(NICAM_ERR_ARG, "ch_b given without ch_b_pl");
    const void *req[] = {q_b, q, d, ch, cmask};
    const char *nm[] = {"q_b", "q", "d", "ch", "cmask"};
    for (int i = 0; i < 5; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (q_a_b) {
        const void *fr[] = {Qpost_b, q_a, Qpost, mask};
        const char *fn[] = {"Qout_post_b", "q_a", "Qout_post", "mask"};
        for (int i = 0; i < 4; ++i) NICAM_TRY(need(fr[i], fn[i]));
    }
    if (c->d.have_pl) {
        const void *rp[] = {q_b_pl, q_pl, d_pl, ch_pl, cmask_pl};
        for (int i = 0; i < 5; ++i) NICAM_TRY(need(rp[i], "pole array"));
        if (q_a_b) {
            const void *fp[] = {q_a_b_pl, Qpost_b_pl, q_a_pl, Qpost_pl, mask_pl};
            for (int i = 0; i < 5; ++i) NICAM_TRY(need(fp[i], "pole array"));
        }
        if (Qprev_b) NICAM_TRY(need(Qprev_b_pl, "pole array"));
    }
    if (!q_a_b) q_a_b_pl = Qpost_b = Qpost_b_pl = nullptr;
    if (!Qprev_b) Qprev_b_pl = nullptr;
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_hlim_adjoint(c, q_a_b, PL(q_a_b_pl), q_b, PL(q_b_pl), Qprev_b, PL(Qprev_b_pl), Qpost_b,
                                   PL(Qpost_b_pl), q_a, PL(q_a_pl), q, PL(q_pl), d, PL(d_pl), ch, PL(ch_pl), cmask,
                                   PL(cmask_pl), Qpost, PL(Qpost_pl), mask, PL(mask_pl), ch_b,
                                   ch_b ? PL(ch_b_pl) : nullptr);
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{q_a_b, R * 6, true},      {PL(q_a_b_pl), P, true},       {q_b, R, true},
                              {PL(q_b_pl), P, true},     {Qprev_b, R * 2, false},       {PL(Qprev_b_pl), P * 2, false},
                              {Qpost_b, R * 2, true},    {PL(Qpost_b_pl), P * 2, true}, {q_a, R * 6, false},
                              {PL(q_a_pl), P, false},    {q, R, false},                 {PL(q_pl), P, false},
                              {d, R, false},             {PL(d_pl), P, false},          {ch, R * 6, false},
                              {PL(ch_pl), P, false},     {cmask, R * 6, false},         {PL(cmask_pl), P, false},
                              {Qpost, R * 2, false},     {PL(Qpost_pl), P * 2, false},  {mask, R / 8, false},
                              {PL(mask_pl), P / 8, false}, {ch_b, R * 6, true},
                              {ch_b ? PL(ch_b_pl) : nullptr, P, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_hlim_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11),
                                  D(12), D(13), D(14), D(15), D(16), D(17), D(18), D(19),
                                  static_cast<const uint8_t *>(v[20]), static_cast<const uint8_t *>(v[21]), D(22),
                                  D(23)));
    return stage_out(c, a, v);
}

static size_t vlim_mask_words(const Dims &d, bool pole)
{
    const size_t nword = d.kall > 3 ? (size_t)(d.kall - 3 + 31) / 32 : 0;
    return pole ? nword * d.gall_pl * kLallPl : nword * d.gall * d.lall;
}

int nicam_vertical_limiter_thuburn_mask(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl,
                                        const double *d, const double *d_pl, const double *ck, const double *ck_pl,
                                        uint32_t *mask, uint32_t *mask_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {q_h, q, d, ck, mask};
    const char *nm[] = {"q_h", "q", "d", "ck", "mask"};
    for (int i = 0; i < 5; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_h_pl, q_pl, d_pl, ck_pl, mask_pl};
        for (int i = 0; i < 5; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vlim_mask(c, q_h, PL(q_h_pl), q, PL(q_pl), d, PL(d_pl), ck, PL(ck_pl), mask, PL(mask_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    const size_t M = vlim_mask_words(c->d, false) * 4, MP = vlim_mask_words(c->d, true) * 4;
    std::vector<HostArg> a = {{q_h, R, true},       {PL(q_h_pl), P, true}, {q, R, false},
                              {PL(q_pl), P, false},  {d, R, false},
                              {PL(d_pl), P, false},  {ck, 2 * R, false},
                              {PL(ck_pl), 2 * P, false}, {mask, M, true},
                              {PL(mask_pl), MP, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vlim_mask(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), static_cast<uint32_t *>(v[8]),
                               static_cast<uint32_t *>(v[9])));
    return stage_out(c, a, v);
}

int nicam_vertical_limiter_thuburn_adjoint(nicam_ctx c, double *q_h_b, double *q_h_b_pl, double *q_b, double *q_b_pl,
                                           const double *q_h, const double *q_h_pl, const double *q,
                                           const double *q_pl, const double *d, const double *d_pl, const double *ck,
                                           const double *ck_pl, const uint32_t *mask, const uint32_t *mask_pl,
                                           double *ck_b, double *ck_b_pl)
{
    NICAM_TRY(check_ctx(c));
    if (ck_b && c->d.have_pl && !ck_b_pl) return fail(NICAM_ERR_ARG, "ck_b given without ck_b_pl");
    const void *req[] = {q_h_b, q_b, q_h, q, d, ck, mask};
    const char *nm[] = {"q_h_b", "q_b", "q_h", "q", "d", "ck", "mask"};
    for (int i = 0; i < 7; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_h_b_pl, q_b_pl, q_h_pl, q_pl, d_pl, ck_pl, mask_pl};
        for (int i = 0; i < 7; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vlim_adjoint(c, q_h_b, PL(q_h_b_pl), q_b, PL(q_b_pl), q_h, PL(q_h_pl), q, PL(q_pl), d,
                                   PL(d_pl), ck, PL(ck_pl), mask, PL(mask_pl), ck_b, ck_b ? PL(ck_b_pl) : nullptr);
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    const size_t M = vlim_mask_words(c->d, false) * 4, MP = vlim_mask_words(c->d, true) * 4;
    std::vector<HostArg> a = {{q_h_b, R, true},        {PL(q_h_b_pl), P, true},      {q_b, R, true},
                              {PL(q_b_pl), P, true},   {q_h, R, false},          {PL(q_h_pl), P, false},
                              {q, R, false},       {PL(q_pl), P, false},     {d, R, false},
                              {PL(d_pl), P, false}, {ck, 2 * R, false},      {PL(ck_pl), 2 * P, false},
                              {mask, M, false},    {PL(mask_pl), MP, false}, {ck_b, 2 * R, true},
                              {ck_b ? PL(ck_b_pl) : nullptr, 2 * P, true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vlim_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11),
                                  static_cast<const uint32_t *>(v[12]), static_cast<const uint32_t *>(v[13]), D(14),
                                  D(15)));
    return stage_out(c, a, v);
}

// ---- step glue (NICAM forms; not in the reference) ----
int nicam_oprt_gradient(nicam_ctx c, double *grad, double *grad_pl, const double *scl, const double *scl_pl,
                        const double *coef_grad, const double *coef_grad_pl)
{
    NICAM_TRY(check_ctx(c));
    NICAM_TRY(need(grad, "grad"));
    NICAM_TRY(need(scl, "scl"));
    NICAM_TRY(need(coef_grad, "coef_grad"));
    if (c->d.have_pl) {
        const void *rp[] = {grad_pl, scl_pl, coef_grad_pl};
        for (int i = 0; i < 3; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_gradient(c, grad, PL(grad_pl), scl, PL(scl_pl), coef_grad, PL(coef_grad_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    const size_t CG = (size_t)c->d.gall * c->d.lall * 21 * 8, CGP = (size_t)c->d.gall_pl * 3 * 2 * 8;
    std::vector<HostArg> a = {{grad, 3 * R, true},          {PL(grad_pl), 3 * P, true}, {scl, R, false},
                              {PL(scl_pl), P, false},       {coef_grad, CG, false},     {PL(coef_grad_pl), CGP, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_gradient(c, D(0), D(1), D(2), D(3), D(4), D(5)));
    return stage_out(c, a, v);
}

int nicam_tracer_courant(nicam_ctx c, double *ch, double *ch_pl, double *cmask, double *cmask_pl,
                         const double *flx_h, const double *flx_h_pl, const double *rhog, const double *rhog_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {ch, cmask, flx_h, rhog};
    const char *nm[] = {"ch", "cmask", "flx_h", "rhog"};
    for (int i = 0; i < 4; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {ch_pl, cmask_pl, flx_h_pl, rhog_pl};
        for (int i = 0; i < 4; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if ((ch_pl == nullptr) != (cmask_pl == nullptr)) return fail(NICAM_ERR_ARG, "give both or none of ch_pl, cmask_pl");
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_courant(c, ch, ch_pl, cmask, cmask_pl, flx_h, PL(flx_h_pl), rhog, PL(rhog_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{ch, 6 * R, true},    {ch_pl, P, true},         {cmask, 6 * R, true},
                              {cmask_pl, P, true},  {flx_h, 6 * R, false},    {PL(flx_h_pl), P, false},
                              {rhog, R, false},     {PL(rhog_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_courant(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7)));
    return stage_out(c, a, v);
}

int nicam_horizontal_remap(nicam_ctx c, double *q_a, double *q_a_pl, const double *q, const double *q_pl,
                           const double *grad, const double *grad_pl, const double *cmask, const double *cmask_pl,
                           const double *GRD_xc, const double *GRD_xc_pl, const double *GRD_x,
                           const double *GRD_x_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {q_a, q, grad, cmask, GRD_xc, GRD_x};
    const char *nm[] = {"q_a", "q", "grad", "cmask", "GRD_xc", "GRD_x"};
    for (int i = 0; i < 6; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_a_pl, q_pl, grad_pl, cmask_pl, GRD_xc_pl, GRD_x_pl};
        for (int i = 0; i < 6; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_remap(c, q_a, PL(q_a_pl), q, PL(q_pl), grad, PL(grad_pl), cmask, PL(cmask_pl), GRD_xc,
                            PL(GRD_xc_pl), GRD_x, PL(GRD_x_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    const size_t X = (size_t)c->d.gall * c->d.lall * 3 * 8, XP = (size_t)c->d.gall_pl * 2 * 3 * 8;
    std::vector<HostArg> a = {{q_a, 6 * R, true},          {PL(q_a_pl), P, true},      {q, R, false},
                              {PL(q_pl), P, false},        {grad, 3 * R, false},       {PL(grad_pl), 3 * P, false},
                              {cmask, 6 * R, false},       {PL(cmask_pl), P, false},   {GRD_xc, 9 * R, false},
                              {PL(GRD_xc_pl), 3 * P, false}, {GRD_x, X, false},       {PL(GRD_x_pl), XP, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_remap(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11)));
    return stage_out(c, a, v);
}

int nicam_tracer_update(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_h, const double *flx_h_pl,
                        const double *q_a, const double *q_a_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {rhogq, flx_h, q_a};
    const char *nm[] = {"rhogq", "flx_h", "q_a"};
    for (int i = 0; i < 3; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {rhogq_pl, flx_h_pl, q_a_pl};
        for (int i = 0; i < 3; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_update(c, rhogq, PL(rhogq_pl), flx_h, PL(flx_h_pl), q_a, PL(q_a_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{rhogq, R, true},        {PL(rhogq_pl), P, true}, {flx_h, 6 * R, false},
                              {PL(flx_h_pl), P, false}, {q_a, 6 * R, false},     {PL(q_a_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_update(c, D(0), D(1), D(2), D(3), D(4), D(5)));
    return stage_out(c, a, v);
}

// ---- adjoints of the horizontal glue ------------------------------------------
// Shared by the four below: check the arrays (pole arrays when the context has
// poles), then launch directly (device pointers) or stage host arrays.
static int glue_adjoint_call(nicam_ctx c, std::vector<HostArg> a, const char *const *names,
                             const std::function<int(const std::vector<void *> &)> &launch)
{
    NICAM_TRY(check_ctx(c));
    for (size_t i = 0; i < a.size(); i += 2) {
        NICAM_TRY(need(a[i].host, names[i / 2]));
        if (c->d.have_pl) NICAM_TRY(need(a[i + 1].host, "pole array"));
        else a[i + 1].host = nullptr;
    }
    std::vector<void *> v;
    if (c->ptr_mode == NICAM_PTR_DEVICE) {
        for (const HostArg &h : a) v.push_back(const_cast<void *>(h.host));
        return launch(v);
    }
    NICAM_TRY(stage_in(c, a, v));
    NICAM_TRY(launch(v));
    return stage_out(c, a, v);
}

int nicam_oprt_gradient_adjoint(nicam_ctx c, double *scl_b, double *scl_b_pl, const double *grad_b,
                                const double *grad_b_pl, const double *coef_grad, const double *coef_grad_pl)
{
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const size_t CG = c ? (size_t)c->d.gall * c->d.lall * 21 * 8 : 0, CGP = c ? (size_t)c->d.gall_pl * 3 * 2 * 8 : 0;
    const char *nm[] = {"scl_b", "grad_b", "coef_grad"};
    return glue_adjoint_call(c, {{scl_b, R, true}, {scl_b_pl, P, true}, {grad_b, 3 * R, false},
                                 {grad_b_pl, 3 * P, false}, {coef_grad, CG, false}, {coef_grad_pl, CGP, false}},
                             nm, [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_gradient_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5));
                             });
}

int nicam_tracer_courant_adjoint(nicam_ctx c, double *flx_h_b, double *flx_h_b_pl, double *rhog_b,
                                 double *rhog_b_pl, const double *ch_b, const double *ch_b_pl, const double *flx_h,
                                 const double *flx_h_pl, const double *rhog, const double *rhog_pl)
{
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const char *nm[] = {"flx_h_b", "rhog_b", "ch_b", "flx_h", "rhog"};
    return glue_adjoint_call(c, {{flx_h_b, 6 * R, true}, {flx_h_b_pl, P, true}, {rhog_b, R, true},
                                 {rhog_b_pl, P, true}, {ch_b, 6 * R, false}, {ch_b_pl, P, false},
                                 {flx_h, 6 * R, false}, {flx_h_pl, P, false}, {rhog, R, false}, {rhog_pl, P, false}},
                             nm, [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_courant_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7),
                                                               D(8), D(9));
                             });
}

int nicam_horizontal_remap_adjoint(nicam_ctx c, double *q_b, double *q_b_pl, double *grad_b, double *grad_b_pl,
                                   double *GRD_xc_b, double *GRD_xc_b_pl, const double *q_a_b, const double *q_a_b_pl,
                                   const double *grad, const double *grad_pl, const double *cmask,
                                   const double *cmask_pl, const double *GRD_xc, const double *GRD_xc_pl,
                                   const double *GRD_x, const double *GRD_x_pl)
{
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const size_t X = c ? (size_t)c->d.gall * c->d.lall * 3 * 8 : 0, XP = c ? (size_t)c->d.gall_pl * 2 * 3 * 8 : 0;
    const char *nm[] = {"q_b", "grad_b", "GRD_xc_b", "q_a_b", "grad", "cmask", "GRD_xc", "GRD_x"};
    return glue_adjoint_call(c, {{q_b, R, true}, {q_b_pl, P, true}, {grad_b, 3 * R, true}, {grad_b_pl, 3 * P, true},
                                 {GRD_xc_b, 9 * R, true}, {GRD_xc_b_pl, 3 * P, true}, {q_a_b, 6 * R, false},
                                 {q_a_b_pl, P, false}, {grad, 3 * R, false}, {grad_pl, 3 * P, false},
                                 {cmask, 6 * R, false}, {cmask_pl, P, false}, {GRD_xc, 9 * R, false},
                                 {GRD_xc_pl, 3 * P, false}, {GRD_x, X, false}, {GRD_x_pl, XP, false}},
                             nm, [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_remap_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8),
                                                             D(9), D(10), D(11), D(12), D(13), D(14), D(15));
                             });
}

int nicam_tracer_update_adjoint(nicam_ctx c, double *flx_h_b, double *flx_h_b_pl, double *q_a_b, double *q_a_b_pl,
                                const double *rhogq_b, const double *rhogq_b_pl, const double *flx_h,
                                const double *flx_h_pl, const double *q_a, const double *q_a_pl)
{
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const char *nm[] = {"flx_h_b", "q_a_b", "rhogq_b", "flx_h", "q_a"};
    return glue_adjoint_call(c, {{flx_h_b, 6 * R, true}, {flx_h_b_pl, P, true}, {q_a_b, 6 * R, true},
                                 {q_a_b_pl, P, true}, {rhogq_b, R, false}, {rhogq_b_pl, P, false},
                                 {flx_h, 6 * R, false}, {flx_h_pl, P, false}, {q_a, 6 * R, false}, {q_a_pl, P, false}},
                             nm, [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_update_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8),
                                                              D(9));
                             });
}

// ---- adjoints of the vertical glue ---------------------------------------------
int nicam_vertical_courant_adjoint(nicam_ctx c, double *flx_v_b, double *flx_v_b_pl, double *rhog_b,
                                   double *rhog_b_pl, const double *ck_b, const double *ck_b_pl, const double *flx_v,
                                   const double *flx_v_pl, const double *rhog, const double *rhog_pl)
{
    if (c && !c->have_rdgz) return fail(NICAM_ERR_STATE, "GRD_rdgz not set (nicam_grd_setup or nicam_set_rdgz)");
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const char *nm[] = {"flx_v_b", "rhog_b", "ck_b", "flx_v", "rhog"};
    return glue_adjoint_call(c, {{flx_v_b, R, true}, {flx_v_b_pl, P, true}, {rhog_b, R, true}, {rhog_b_pl, P, true},
                                 {ck_b, 2 * R, false}, {ck_b_pl, 2 * P, false}, {flx_v, R, false},
                                 {flx_v_pl, P, false}, {rhog, R, false}, {rhog_pl, P, false}},
                             nm, [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_vertical_courant_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6),
                                                                        D(7), D(8), D(9));
                             });
}

int nicam_vertical_face_adjoint(nicam_ctx c, double *q_b, double *q_b_pl, const double *q_h_b, const double *q_h_b_pl)
{
    if (c && !c->have_vgrid) return fail(NICAM_ERR_STATE, "vertical grid not set (nicam_grd_setup)");
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const char *nm[] = {"q_b", "q_h_b"};
    return glue_adjoint_call(c, {{q_b, R, true}, {q_b_pl, P, true}, {q_h_b, R, false}, {q_h_b_pl, P, false}}, nm,
                             [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_vertical_face_adjoint(c, D(0), D(1), D(2), D(3));
                             });
}

int nicam_tracer_update_vertical_adjoint(nicam_ctx c, double *flx_v_b, double *flx_v_b_pl, double *q_h_b,
                                         double *q_h_b_pl, const double *rhogq_b, const double *rhogq_b_pl,
                                         const double *flx_v, const double *flx_v_pl, const double *q_h,
                                         const double *q_h_pl)
{
    if (c && !c->have_rdgz) return fail(NICAM_ERR_STATE, "GRD_rdgz not set (nicam_grd_setup or nicam_set_rdgz)");
    const size_t R = c ? c->d.nreg() * 8 : 0, P = c ? c->d.npl() * 8 : 0;
    const char *nm[] = {"flx_v_b", "q_h_b", "rhogq_b", "flx_v", "q_h"};
    return glue_adjoint_call(c, {{flx_v_b, R, true}, {flx_v_b_pl, P, true}, {q_h_b, R, true}, {q_h_b_pl, P, true},
                                 {rhogq_b, R, false}, {rhogq_b_pl, P, false}, {flx_v, R, false},
                                 {flx_v_pl, P, false}, {q_h, R, false}, {q_h_pl, P, false}},
                             nm, [&](const std::vector<void *> &v) {
                                 auto D = [&](int i) { return static_cast<double *>(v[i]); };
                                 return launch_vertical_update_adjoint(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6),
                                                                       D(7), D(8), D(9));
                             });
}

int nicam_vertical_mass_flux(nicam_ctx c, double *flx_v, double *flx_v_pl, const double *rhogvx,
                             const double *rhogvx_pl, const double *rhogvy, const double *rhogvy_pl,
                             const double *rhogvz, const double *rhogvz_pl, const double *rhogw,
                             const double *rhogw_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_vmetrics3d) return fail(NICAM_ERR_STATE, "vertical metrics not set (nicam_set_vmetrics3d)");
    const void *req[] = {flx_v, rhogvx, rhogvy, rhogvz, rhogw};
    const char *nm[] = {"flx_v", "rhogvx", "rhogvy", "rhogvz", "rhogw"};
    for (int i = 0; i < 5; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {flx_v_pl, rhogvx_pl, rhogvy_pl, rhogvz_pl, rhogw_pl};
        for (int i = 0; i < 5; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vertical_mass_flux(c, flx_v, PL(flx_v_pl), rhogvx, PL(rhogvx_pl), rhogvy, PL(rhogvy_pl), rhogvz,
                                         PL(rhogvz_pl), rhogw, PL(rhogw_pl), dt);
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{flx_v, R, true},   {PL(flx_v_pl), P, true},   {rhogvx, R, false},
                              {PL(rhogvx_pl), P, false}, {rhogvy, R, false}, {PL(rhogvy_pl), P, false},
                              {rhogvz, R, false}, {PL(rhogvz_pl), P, false}, {rhogw, R, false},
                              {PL(rhogw_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vertical_mass_flux(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), dt));
    return stage_out(c, a, v);
}

int nicam_vertical_courant(nicam_ctx c, double *ck, double *ck_pl, const double *flx_v, const double *flx_v_pl,
                           const double *rhog, const double *rhog_pl)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_rdgz) return fail(NICAM_ERR_STATE, "GRD_rdgz not set (nicam_grd_setup or nicam_set_rdgz)");
    const void *req[] = {ck, flx_v, rhog};
    const char *nm[] = {"ck", "flx_v", "rhog"};
    for (int i = 0; i < 3; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {ck_pl, flx_v_pl, rhog_pl};
        for (int i = 0; i < 3; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vertical_courant(c, ck, PL(ck_pl), flx_v, PL(flx_v_pl), rhog, PL(rhog_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{ck, 2 * R, true},  {PL(ck_pl), 2 * P, true}, {flx_v, R, false},
                              {PL(flx_v_pl), P, false}, {rhog, R, false}, {PL(rhog_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vertical_courant(c, D(0), D(1), D(2), D(3), D(4), D(5)));
    return stage_out(c, a, v);
}

int nicam_vertical_face(nicam_ctx c, double *q_h, double *q_h_pl, const double *q, const double *q_pl)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_vgrid) return fail(NICAM_ERR_STATE, "vertical grid not set (nicam_grd_setup)");
    NICAM_TRY(need(q_h, "q_h"));
    NICAM_TRY(need(q, "q"));
    if (c->d.have_pl) {
        NICAM_TRY(need(q_h_pl, "pole array"));
        NICAM_TRY(need(q_pl, "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE) return launch_vertical_face(c, q_h, PL(q_h_pl), q, PL(q_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{q_h, R, true}, {PL(q_h_pl), P, true}, {q, R, false}, {PL(q_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vertical_face(c, D(0), D(1), D(2), D(3)));
    return stage_out(c, a, v);
}

int nicam_tracer_update_vertical(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_v,
                                 const double *flx_v_pl, const double *q_h, const double *q_h_pl)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_rdgz) return fail(NICAM_ERR_STATE, "GRD_rdgz not set (nicam_grd_setup or nicam_set_rdgz)");
    const void *req[] = {rhogq, flx_v, q_h};
    const char *nm[] = {"rhogq", "flx_v", "q_h"};
    for (int i = 0; i < 3; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {rhogq_pl, flx_v_pl, q_h_pl};
        for (int i = 0; i < 3; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vertical_update(c, rhogq, PL(rhogq_pl), flx_v, PL(flx_v_pl), q_h, PL(q_h_pl));
    const size_t R = c->d.nreg() * 8, P = c->d.npl() * 8;
    std::vector<HostArg> a = {{rhogq, R, true},       {PL(rhogq_pl), P, true}, {flx_v, R, false},
                              {PL(flx_v_pl), P, false}, {q_h, R, false},      {PL(q_h_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<double *>(v[i]); };
    NICAM_TRY(launch_vertical_update(c, D(0), D(1), D(2), D(3), D(4), D(5)));
    return stage_out(c, a, v);
}

int nicam_tracer_update_vertical_f32(nicam_ctx c, double *rhogq, double *rhogq_pl, const float *flx_v,
                                     const float *flx_v_pl, const float *q_h, const float *q_h_pl)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_rdgz) return fail(NICAM_ERR_STATE, "GRD_rdgz not set (nicam_grd_setup or nicam_set_rdgz)");
    const void *req[] = {rhogq, flx_v, q_h};
    const char *nm[] = {"rhogq", "flx_v", "q_h"};
    for (int i = 0; i < 3; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {rhogq_pl, flx_v_pl, q_h_pl};
        for (int i = 0; i < 3; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vertical_update_f32(c, rhogq, PL(rhogq_pl), flx_v, PL(flx_v_pl), q_h, PL(q_h_pl));
    const size_t R = c->d.nreg(), P = c->d.npl();
    std::vector<HostArg> a = {{rhogq, R * 8, true},       {PL(rhogq_pl), P * 8, true}, {flx_v, R * 4, false},
                              {PL(flx_v_pl), P * 4, false}, {q_h, R * 4, false},      {PL(q_h_pl), P * 4, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    NICAM_TRY(launch_vertical_update_f32(c, static_cast<double *>(v[0]), static_cast<double *>(v[1]),
                                         static_cast<const float *>(v[2]), static_cast<const float *>(v[3]),
                                         static_cast<const float *>(v[4]), static_cast<const float *>(v[5])));
    return stage_out(c, a, v);
}

int nicam_vi_rhow_update_matrix(nicam_ctx c, const double *eth, const double *eth_pl, const double *g_tilde,
                                const double *g_tilde_pl, double dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_vmetrics) return fail(NICAM_ERR_STATE, "vertical metrics not set (nicam_set_vmetrics)");
    if (!c->have_cdfact || !c->have_rdgz || !c->have_vgrid)
        return fail(NICAM_ERR_STATE, "vertical grid factors not set (nicam_grd_setup)");
    if (!(dt > 0.0)) return fail(NICAM_ERR_ARG, "dt must be positive");
    NICAM_TRY(need(eth, "eth"));
    NICAM_TRY(need(g_tilde, "g_tilde"));
    if (c->d.have_pl) {
        NICAM_TRY(need(eth_pl, "pole array"));
        NICAM_TRY(need(g_tilde_pl, "pole array"));
    }
    const size_t nr = c->d.nreg(), np = c->d.npl();
    NICAM_TRY(dev_alloc(c, &c->Mc, nr));
    NICAM_TRY(dev_alloc(c, &c->Ml, nr));
    NICAM_TRY(dev_alloc(c, &c->Mu, nr));
    if (c->d.have_pl) {
        NICAM_TRY(dev_alloc(c, &c->Mc_pl, np));
        NICAM_TRY(dev_alloc(c, &c->Ml_pl, np));
        NICAM_TRY(dev_alloc(c, &c->Mu_pl, np));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE) {
        NICAM_TRY(launch_vi_matrix(c, eth, PL(eth_pl), g_tilde, PL(g_tilde_pl), dt));
    } else {
        const size_t R = nr * 8, P = np * 8;
        std::vector<HostArg> a = {{eth, R, false}, {PL(eth_pl), P, false}, {g_tilde, R, false},
                                  {PL(g_tilde_pl), P, false}};
        std::vector<void *> v;
        NICAM_TRY(stage_in(c, a, v));
        auto D = [&](int i) { return static_cast<const double *>(v[i]); };
        NICAM_TRY(launch_vi_matrix(c, D(0), D(1), D(2), D(3), dt));
        NICAM_TRY(stage_out(c, a, v));
    }
    NICAM_TRY(factor_coef(c));
    NICAM_HIP(hipStreamSynchronize(c->stream));
    c->have_coef = true;
    return NICAM_OK;
}

int nicam_vi_rhow_getcoef(nicam_ctx c, double *Mc, double *Mc_pl, double *Ml, double *Ml_pl, double *Mu,
                          double *Mu_pl)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_coef) return fail(NICAM_ERR_STATE, "no coefficients (nicam_vi_rhow_solver_putcoef or _update_matrix)");
    const size_t nr = c->d.nreg() * sizeof(double), np = c->d.npl() * sizeof(double);
    const hipMemcpyKind kind = c->ptr_mode == NICAM_PTR_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    double *dst[] = {Mc, Ml, Mu};
    double *src[] = {c->Mc, c->Ml, c->Mu};
    double *dst_pl[] = {Mc_pl, Ml_pl, Mu_pl};
    double *src_pl[] = {c->Mc_pl, c->Ml_pl, c->Mu_pl};
    for (int i = 0; i < 3; ++i) {
        NICAM_TRY(need(dst[i], "coefficient array"));
        NICAM_HIP(hipMemcpyAsync(dst[i], src[i], nr, kind, c->stream));
        if (c->d.have_pl) {
            NICAM_TRY(need(dst_pl[i], "pole array"));
            NICAM_HIP(hipMemcpyAsync(dst_pl[i], src_pl[i], np, kind, c->stream));
        }
    }
    NICAM_HIP(hipStreamSynchronize(c->stream));
    return NICAM_OK;
}

int nicam_horizontal_flux_f32(nicam_ctx c, float *flx_h, float *flx_h_pl, float *xc, float *xc_pl,
                          const float *rho, const float *rho_pl, const float *rvx,
                          const float *rvx_pl, const float *rvy, const float *rvy_pl,
                          const float *rvz, const float *rvz_pl, float dt)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_hmetrics) return fail(NICAM_ERR_STATE, "horizontal metrics not set (nicam_set_hmetrics)");
    const void *req[] = {flx_h, xc, rho, rvx, rvy, rvz};
    const char *nm[] = {"flx_h", "GRD_xc", "rho", "rhovx", "rhovy", "rhovz"};
    for (int i = 0; i < 6; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {flx_h_pl, xc_pl, rho_pl, rvx_pl, rvy_pl, rvz_pl};
        for (int i = 0; i < 6; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_hflux_f32(c, flx_h, PL(flx_h_pl), xc, PL(xc_pl), rho, PL(rho_pl), rvx, PL(rvx_pl),
                            rvy, PL(rvy_pl), rvz, PL(rvz_pl), dt);
    const size_t R = c->d.nreg() * 4, P = c->d.npl() * 4;
    std::vector<HostArg> a = {{flx_h, R * 6, true},  {PL(flx_h_pl), P, true}, {xc, R * 9, true},
                              {PL(xc_pl), P * 3, true}, {rho, R, false},       {PL(rho_pl), P, false},
                              {rvx, R, false},         {PL(rvx_pl), P, false}, {rvy, R, false},
                              {PL(rvy_pl), P, false},  {rvz, R, false},        {PL(rvz_pl), P, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<float *>(v[i]); };
    NICAM_TRY(launch_hflux_f32(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11), dt));
    return stage_out(c, a, v);
}

int nicam_horizontal_limiter_thuburn_f32(nicam_ctx c, float *q_a, float *q_a_pl, const float *q,
                                     const float *q_pl, const float *d, const float *d_pl,
                                     const float *ch, const float *ch_pl, const float *cmask,
                                     const float *cmask_pl, float *Qprev, float *Qprev_pl,
                                     const float *Qpost, const float *Qpost_pl)
{
    NICAM_TRY(check_ctx(c));
    // Qout_post NULL: the library exchanges Qout_prev itself (nicam_comm_setup
    // must have run); Qout_post_pl is then ignored.
    const void *req[] = {q_a, q, d, ch, cmask, Qprev};
    const char *nm[] = {"q_a", "q", "d", "ch", "cmask", "Qout_prev"};
    for (int i = 0; i < 6; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_a_pl, q_pl, d_pl, ch_pl, cmask_pl, Qprev_pl};
        for (int i = 0; i < 6; ++i) NICAM_TRY(need(rp[i], "pole array"));
        if (Qpost) NICAM_TRY(need(Qpost_pl, "Qout_post_pl"));
    }
    if (!Qpost) Qpost_pl = nullptr;
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_hlim_f32(c, q_a, PL(q_a_pl), q, PL(q_pl), d, PL(d_pl), ch, PL(ch_pl), cmask,
                           PL(cmask_pl), Qprev, Qprev_pl, Qpost, PL(Qpost_pl));
    const size_t R = c->d.nreg() * 4, P = c->d.npl() * 4;
    std::vector<HostArg> a = {{q_a, R * 6, true},       {PL(q_a_pl), P, true},   {q, R, false},
                              {PL(q_pl), P, false},     {d, R, false},           {PL(d_pl), P, false},
                              {ch, R * 6, false},       {PL(ch_pl), P, false},   {cmask, R * 6, false},
                              {PL(cmask_pl), P, false}, {Qprev, R * 2, true},    {Qprev_pl, P * 2, true},
                              {Qpost, R * 2, false},    {PL(Qpost_pl), P * 2, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<float *>(v[i]); };
    NICAM_TRY(launch_hlim_f32(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9), D(10), D(11),
                          D(12), D(13)));
    return stage_out(c, a, v);
}

int nicam_vertical_limiter_thuburn_f32(nicam_ctx c, float *q_h, float *q_h_pl, const float *q,
                                   const float *q_pl, const float *d, const float *d_pl,
                                   const float *ck, const float *ck_pl)
{
    NICAM_TRY(check_ctx(c));
    const void *req[] = {q_h, q, d, ck};
    const char *nm[] = {"q_h", "q", "d", "ck"};
    for (int i = 0; i < 4; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) {
        const void *rp[] = {q_h_pl, q_pl, d_pl, ck_pl};
        for (int i = 0; i < 4; ++i) NICAM_TRY(need(rp[i], "pole array"));
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE)
        return launch_vlim_f32(c, q_h, PL(q_h_pl), q, PL(q_pl), d, PL(d_pl), ck, PL(ck_pl));
    const size_t R = c->d.nreg() * 4, P = c->d.npl() * 4;
    std::vector<HostArg> a = {{q_h, R, true},  {PL(q_h_pl), P, true}, {q, R, false},      {PL(q_pl), P, false},
                              {d, R, false},   {PL(d_pl), P, false},  {ck, R * 2, false}, {PL(ck_pl), P * 2, false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    auto D = [&](int i) { return static_cast<float *>(v[i]); };
    NICAM_TRY(launch_vlim_f32(c, D(0), D(1), D(2), D(3), D(4), D(5), D(6), D(7)));
    return stage_out(c, a, v);
}

#undef PL

}  // extern "C"

// ---- FP64 global tracer-mass sums and the conservative update ---------------

template <class T>
static int tracer_mass_t(nicam_ctx c, const T *rhogq, const T *rhogq_pl, double *mass)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_hmetrics) return fail(NICAM_ERR_STATE, "horizontal metrics not set (nicam_set_hmetrics)");
    NICAM_TRY(need(rhogq, "rhogq"));
    NICAM_TRY(need(mass, "mass"));
    if (c->d.have_pl) NICAM_TRY(need(rhogq_pl, "pole array"));
    const size_t np = (size_t)c->d.kall * c->d.lall;
    if (!c->mass_part) {  // the partial sums: allocated on the first call, kept by the context
        NICAM_HIP(hipMalloc(&c->mass_part, np * sizeof(double)));
        c->device_bytes += np * sizeof(double);
    }
    if (c->ptr_mode == NICAM_PTR_DEVICE) {
        if constexpr (std::is_same<T, float>::value) return launch_mass_f32(c, rhogq, (c->d.have_pl ? rhogq_pl : nullptr), mass, c->mass_part);
        else return launch_mass(c, rhogq, (c->d.have_pl ? rhogq_pl : nullptr), mass, c->mass_part);
    }
    const size_t R = c->d.nreg() * sizeof(T), P = c->d.npl() * sizeof(T);
    std::vector<HostArg> a = {{rhogq, R, false}, {(c->d.have_pl ? rhogq_pl : nullptr), P, false}, {mass, c->d.kall * sizeof(double), true}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    if constexpr (std::is_same<T, float>::value)
        NICAM_TRY(launch_mass_f32(c, static_cast<float *>(v[0]), static_cast<float *>(v[1]),
                                  static_cast<double *>(v[2]), c->mass_part));
    else
        NICAM_TRY(launch_mass(c, static_cast<double *>(v[0]), static_cast<double *>(v[1]),
                              static_cast<double *>(v[2]), c->mass_part));
    return stage_out(c, a, v);
}

template <class T>
static int update_cons_t(nicam_ctx c, double *rhogq, double *rhogq_pl, const T *flx_h, const T *q_a)
{
    NICAM_TRY(check_ctx(c));
    if (!c->have_hmetrics) return fail(NICAM_ERR_STATE, "horizontal metrics not set (nicam_set_hmetrics)");
    if (!c->arc_slot) return fail(NICAM_ERR_STATE, "the conservative update needs the halo map (nicam_comm_setup)");
    const void *req[] = {rhogq, flx_h, q_a};
    const char *nm[] = {"rhogq", "flx_h", "q_a"};
    for (int i = 0; i < 3; ++i) NICAM_TRY(need(req[i], nm[i]));
    if (c->d.have_pl) NICAM_TRY(need(rhogq_pl, "pole array"));
    auto launch = [&](double *r, double *rp, const T *f, const T *q) {
        if constexpr (std::is_same<T, float>::value) return launch_update_cons_f32(c, r, rp, f, q);
        else return launch_update_cons(c, r, rp, f, q);
    };
    if (c->ptr_mode == NICAM_PTR_DEVICE) return launch(rhogq, (c->d.have_pl ? rhogq_pl : nullptr), flx_h, q_a);
    const size_t R = c->d.nreg(), P = c->d.npl();
    std::vector<HostArg> a = {{rhogq, R * 8, true}, {(c->d.have_pl ? rhogq_pl : nullptr), P * 8, true}, {flx_h, 6 * R * sizeof(T), false},
                              {q_a, 6 * R * sizeof(T), false}};
    std::vector<void *> v;
    NICAM_TRY(stage_in(c, a, v));
    NICAM_TRY(launch(static_cast<double *>(v[0]), static_cast<double *>(v[1]), static_cast<const T *>(v[2]),
                     static_cast<const T *>(v[3])));
    return stage_out(c, a, v);
}

extern "C" {

int nicam_tracer_mass(nicam_ctx c, const double *rhogq, const double *rhogq_pl, double *mass)
{
    return tracer_mass_t<double>(c, rhogq, rhogq_pl, mass);
}

int nicam_tracer_mass_f32(nicam_ctx c, const float *rhogq, const float *rhogq_pl, double *mass)
{
    return tracer_mass_t<float>(c, rhogq, rhogq_pl, mass);
}

int nicam_tracer_update_conservative(nicam_ctx c, double *rhogq, double *rhogq_pl, const double *flx_h,
                                     const double *q_a)
{
    return update_cons_t<double>(c, rhogq, rhogq_pl, flx_h, q_a);
}

int nicam_tracer_update_conservative_f32(nicam_ctx c, double *rhogq, double *rhogq_pl, const float *flx_h,
                                         const float *q_a)
{
    return update_cons_t<float>(c, rhogq, rhogq_pl, flx_h, q_a);
}

}  // extern "C"
