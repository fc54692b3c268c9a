// Many tracers and several grids: the batched limiters, the grid group, its
// four-launch tracer step (optionally captured as one HIP graph) and the
// autotuner of its launch shapes.
//
// The reference advances one tracer of one grid per call
// (dyn_horiz_adv_limiter/src/main.f90:208-214).  A group holds up to eight
// grids (contexts of one device); each kernel of the step runs once over all
// of them, work-groups [off[n], off[n+1]) of the launch belonging to grid n,
// and the limiters take every tracer of a grid in one pass over the fields
// the tracers share.
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "internal.h"

using namespace nicam;

struct nicam_group_s {
    std::vector<nicam_ctx> ctxs;
    hipStream_t stream = nullptr;
    int device = 0;
    // the captured step and the arguments it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<nicam_tracer_step_f32> key;
    int key_ntr = 0;
    float key_dtl = 0.0f;
    double key_dts = 0.0;
};

namespace {

#define NICAM_TRY(x)                      \
    do {                                  \
        int rc_ = (x);                    \
        if (rc_ != NICAM_OK) return rc_;  \
    } while (0)

int check_group(nicam_group g)
{
    if (!g) return fail(NICAM_ERR_ARG, "null group");
    return NICAM_OK;
}

void drop_graph(nicam_group g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
    g->key.clear();
}

int check_step_args(nicam_group g, int ntr, const nicam_tracer_step_f32 *grids)
{
    if (ntr < 1) return fail(NICAM_ERR_ARG, "ntr must be >= 1");
    if (!grids) return fail(NICAM_ERR_ARG, "null grid arguments");
    for (size_t i = 0; i < g->ctxs.size(); ++i) {
        const nicam_ctx c = g->ctxs[i];
        const nicam_tracer_step_f32 &p = grids[i];
        if (!c->have_hmetrics || !c->have_vmetrics || !c->have_vgrid || !c->have_coef)
            return fail(NICAM_ERR_STATE, "a grid of the group lacks metrics, vertical grid or coefficients");
        if (c->ptr_mode != NICAM_PTR_DEVICE) return fail(NICAM_ERR_STATE, "group steps take device pointers only");
        const void *req[] = {p.flx_h, p.GRD_xc, p.rho, p.rhovx, p.rhovy, p.rhovz, p.q_a, p.q, p.d, p.ch, p.cmask,
                             p.Qout_prev, p.Qout_post, p.q_h, p.qv, p.dv, p.ck, p.rhogw, p.rhogw0, p.preg0, p.rhog0,
                             p.Srho, p.Sw, p.Spre};
        for (const void *x : req)
            if (!x) return fail(NICAM_ERR_ARG, "null array in the group step arguments");
        if (c->d.have_pl) {
            const void *rp[] = {p.flx_h_pl, p.GRD_xc_pl, p.rho_pl, p.rhovx_pl, p.rhovy_pl, p.rhovz_pl, p.q_a_pl,
                                p.q_pl, p.d_pl, p.ch_pl, p.cmask_pl, p.Qout_prev_pl, p.Qout_post_pl, p.q_h_pl,
                                p.qv_pl, p.dv_pl, p.ck_pl, p.rhogw_pl, p.rhogw0_pl, p.preg0_pl, p.rhog0_pl,
                                p.Srho_pl, p.Sw_pl, p.Spre_pl};
            for (const void *x : rp)
                if (!x) return fail(NICAM_ERR_ARG, "null pole array in the group step arguments");
        }
    }
    return NICAM_OK;
}

int run_step(nicam_group g, int ntr, const nicam_tracer_step_f32 *grids, float dtl, double dts)
{
    const int n = (int)g->ctxs.size();
    const nicam_ctx *cs = g->ctxs.data();
    NICAM_TRY(launch_hflux_group_f32(cs, n, grids, dtl, g->stream));
    NICAM_TRY(launch_hlim_group_f32(cs, n, ntr, grids, g->stream));
    NICAM_TRY(launch_vlim_group_f32(cs, n, ntr, grids, g->stream));
    return launch_vi_group_mixed(cs, n, grids, dts, g->stream);
}

float time_ms(hipStream_t s, int reps, const std::function<int()> &f, int &rc)
{
    hipEvent_t e0, e1;
    rc = NICAM_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(NICAM_ERR_HIP, "hipEventCreate");
        return 0.0f;
    }
    rc = f();  // warm-up
    float ms = 0.0f;
    if (rc == NICAM_OK) {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && rc == NICAM_OK; ++r) rc = f();
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

std::string tune_key(nicam_ctx c, int ntr)
{
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
    for (char &ch : arch)
        if (ch == ' ') ch = '_';
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s %d %d %d %d %d %d", arch.c_str(), c->d.gall_1d, c->d.kall, c->d.lall,
                  c->d.gall_pl, c->d.have_pl, ntr);
    return buf;
}

}  // namespace

extern "C" {

int nicam_horizontal_limiter_thuburn_batch(nicam_ctx c, int ntr, double *q_a, double *q_a_pl, const double *q,
                                           const double *q_pl, const double *d, const double *d_pl, const double *ch,
                                           const double *ch_pl, const double *cmask, const double *cmask_pl,
                                           double *Qprev, double *Qprev_pl, const double *Qpost,
                                           const double *Qpost_pl)
{
    if (!c) return fail(NICAM_ERR_ARG, "null context");
    if (c->ptr_mode != NICAM_PTR_DEVICE) return fail(NICAM_ERR_STATE, "batched calls take device pointers only");
    if (!q_a || !q || !d || !ch || !cmask || !Qprev) return fail(NICAM_ERR_ARG, "null argument");
    const bool pl = c->d.have_pl;
    return launch_hlim_batch(c, ntr, q_a, pl ? q_a_pl : nullptr, q, pl ? q_pl : nullptr, d, pl ? d_pl : nullptr, ch,
                             pl ? ch_pl : nullptr, cmask, pl ? cmask_pl : nullptr, Qprev, Qprev_pl, Qpost,
                             pl ? Qpost_pl : nullptr);
}

int nicam_horizontal_limiter_thuburn_batch_f32(nicam_ctx c, int ntr, float *q_a, float *q_a_pl, const float *q,
                                               const float *q_pl, const float *d, const float *d_pl, const float *ch,
                                               const float *ch_pl, const float *cmask, const float *cmask_pl,
                                               float *Qprev, float *Qprev_pl, const float *Qpost,
                                               const float *Qpost_pl)
{
    if (!c) return fail(NICAM_ERR_ARG, "null context");
    if (c->ptr_mode != NICAM_PTR_DEVICE) return fail(NICAM_ERR_STATE, "batched calls take device pointers only");
    if (!q_a || !q || !d || !ch || !cmask || !Qprev) return fail(NICAM_ERR_ARG, "null argument");
    const bool pl = c->d.have_pl;
    return launch_hlim_batch_f32(c, ntr, q_a, pl ? q_a_pl : nullptr, q, pl ? q_pl : nullptr, d, pl ? d_pl : nullptr,
                                 ch, pl ? ch_pl : nullptr, cmask, pl ? cmask_pl : nullptr, Qprev, Qprev_pl, Qpost,
                                 pl ? Qpost_pl : nullptr);
}

int nicam_vertical_limiter_thuburn_batch(nicam_ctx c, int ntr, double *q_h, double *q_h_pl, const double *q,
                                         const double *q_pl, const double *d, const double *d_pl, const double *ck,
                                         const double *ck_pl)
{
    if (!c) return fail(NICAM_ERR_ARG, "null context");
    if (c->ptr_mode != NICAM_PTR_DEVICE) return fail(NICAM_ERR_STATE, "batched calls take device pointers only");
    if (!q_h || !q || !d || !ck) return fail(NICAM_ERR_ARG, "null argument");
    if (c->d.have_pl && (!q_h_pl || !q_pl || !d_pl || !ck_pl)) return fail(NICAM_ERR_ARG, "null pole array");
    const bool pl = c->d.have_pl;
    return launch_vlim_batch(c, ntr, q_h, pl ? q_h_pl : nullptr, q, pl ? q_pl : nullptr, d, pl ? d_pl : nullptr, ck,
                             pl ? ck_pl : nullptr);
}

int nicam_vertical_limiter_thuburn_batch_f32(nicam_ctx c, int ntr, float *q_h, float *q_h_pl, const float *q,
                                             const float *q_pl, const float *d, const float *d_pl, const float *ck,
                                             const float *ck_pl)
{
    if (!c) return fail(NICAM_ERR_ARG, "null context");
    if (c->ptr_mode != NICAM_PTR_DEVICE) return fail(NICAM_ERR_STATE, "batched calls take device pointers only");
    if (!q_h || !q || !d || !ck) return fail(NICAM_ERR_ARG, "null argument");
    if (c->d.have_pl && (!q_h_pl || !q_pl || !d_pl || !ck_pl)) return fail(NICAM_ERR_ARG, "null pole array");
    const bool pl = c->d.have_pl;
    return launch_vlim_batch_f32(c, ntr, q_h, pl ? q_h_pl : nullptr, q, pl ? q_pl : nullptr, d, pl ? d_pl : nullptr,
                                 ck, pl ? ck_pl : nullptr);
}

int nicam_group_create(nicam_group *out, int n, const nicam_ctx *ctxs)
{
    if (!out) return fail(NICAM_ERR_ARG, "null output");
    *out = nullptr;
    if (n < 1 || n > kMaxGroup) return fail(NICAM_ERR_ARG, "a group holds 1..8 grids");
    if (!ctxs) return fail(NICAM_ERR_ARG, "null context list");
    auto *g = new (std::nothrow) nicam_group_s();
    if (!g) return fail(NICAM_ERR_NOMEM, "host allocation failed");
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) {
            delete g;
            return fail(NICAM_ERR_ARG, "null context in the group");
        }
        if (ctxs[i]->device != ctxs[0]->device) {
            delete g;
            return fail(NICAM_ERR_ARG, "the grids of a group must live on one device");
        }
        g->ctxs.push_back(ctxs[i]);
    }
    g->device = ctxs[0]->device;
    g->stream = ctxs[0]->stream;
    *out = g;
    return NICAM_OK;
}

int nicam_group_destroy(nicam_group g)
{
    if (!g) return NICAM_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    drop_graph(g);
    delete g;
    return NICAM_OK;
}

int nicam_group_set_stream(nicam_group g, void *s)
{
    NICAM_TRY(check_group(g));
    drop_graph(g);  // a graph replays on any stream, but keep capture and replay on the same one
    g->stream = static_cast<hipStream_t>(s);
    return NICAM_OK;
}

int nicam_group_tracer_step_f32(nicam_group g, int ntr, const nicam_tracer_step_f32 *grids, float dt_large,
                                double dt_small, int use_graph)
{
    NICAM_TRY(check_group(g));
    NICAM_TRY(check_step_args(g, ntr, grids));
    if (!use_graph) return run_step(g, ntr, grids, dt_large, dt_small);
    if (!g->stream) return fail(NICAM_ERR_STATE, "graph capture needs a stream (nicam_group_set_stream)");
    const size_t n = g->ctxs.size();
    const bool same = g->exec && g->key_ntr == ntr && g->key_dtl == dt_large && g->key_dts == dt_small &&
                      g->key.size() == n && std::memcmp(g->key.data(), grids, n * sizeof(*grids)) == 0;
    if (!same) {
        drop_graph(g);
        NICAM_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed));
        const int rc = run_step(g, ntr, grids, dt_large, dt_small);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(g->stream, &graph);
        if (rc != NICAM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        g->graph = graph;
        const hipError_t ei = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            drop_graph(g);
            return hip_fail(ei, "hipGraphInstantiate");
        }
        g->key.assign(grids, grids + n);
        g->key_ntr = ntr;
        g->key_dtl = dt_large;
        g->key_dts = dt_small;
    }
    NICAM_HIP(hipGraphLaunch(g->exec, g->stream));
    return NICAM_OK;
}

int nicam_group_autotune(nicam_group g, int ntr, const nicam_tracer_step_f32 *grids, float dt_large,
                         double dt_small, const char *cache_path, int *n_timed)
{
    NICAM_TRY(check_group(g));
    NICAM_TRY(check_step_args(g, ntr, grids));
    (void)dt_small;
    int timed = 0;
    std::vector<std::string> cache;
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "r")) {
            char line[512];
            while (std::fgets(line, sizeof line, f)) cache.emplace_back(line);
            std::fclose(f);
        }
    }
    drop_graph(g);  // launch shapes change: a captured step is stale
    for (size_t i = 0; i < g->ctxs.size(); ++i) {
        const nicam_ctx c = g->ctxs[i];
        const nicam_tracer_step_f32 &p = grids[i];
        const std::string key = "nicam_tt-tune v1 " + tune_key(c, ntr) + " ";
        bool found = false;
        for (const std::string &ln : cache) {
            int kper = 0, kch = 0;
            if (ln.compare(0, key.size(), key) == 0 &&
                std::sscanf(ln.c_str() + key.size(), "hflux_kper=%d vlim_kchunk=%d", &kper, &kch) == 2) {
                c->tune.hflux_kper = kper;
                c->tune.vlim_kchunk = kch;
                found = true;
            }
        }
        if (found) continue;
        const hipStream_t s = g->stream;
        const hipStream_t s0 = c->stream;
        c->stream = s;
        int rc = NICAM_OK;
        // levels per work-group of the flux (0 = the library's default count)
        int best_kper = 0;
        float best = 1e30f;
        for (int kper : {0, 1, 2, 4, 8, 16, 32}) {
            if (kper > c->d.kall) continue;
            c->tune.hflux_kper = kper;
            const float ms = time_ms(s, 3, [&]() {
                return launch_hflux_f32(c, p.flx_h, p.flx_h_pl, p.GRD_xc, p.GRD_xc_pl, p.rho, p.rho_pl, p.rhovx,
                                        p.rhovx_pl, p.rhovy, p.rhovy_pl, p.rhovz, p.rhovz_pl, dt_large);
            }, rc);
            if (rc != NICAM_OK) break;
            ++timed;
            if (ms < best) {
                best = ms;
                best_kper = kper;
            }
        }
        c->tune.hflux_kper = best_kper;
        // levels per work-item of the batched vertical limiter
        int best_kc = 0;
        best = 1e30f;
        for (int kc : {2, 4, 8, 16}) {
            if (rc != NICAM_OK) break;
            c->tune.vlim_kchunk = kc;
            const float ms = time_ms(s, 3, [&]() {
                return launch_vlim_batch_f32(c, ntr, p.q_h, p.q_h_pl, p.qv, p.qv_pl, p.dv, p.dv_pl, p.ck, p.ck_pl);
            }, rc);
            if (rc != NICAM_OK) break;
            ++timed;
            if (ms < best) {
                best = ms;
                best_kc = kc;
            }
        }
        c->tune.vlim_kchunk = best_kc;
        c->stream = s0;
        if (rc != NICAM_OK) return rc;
        if (cache_path) {
            FILE *f = std::fopen(cache_path, "a");
            if (!f) return fail(NICAM_ERR_IO, std::string("cannot append to ") + cache_path);
            std::fprintf(f, "%shflux_kper=%d vlim_kchunk=%d\n", key.c_str(), best_kper, best_kc);
            std::fclose(f);
        }
    }
    if (n_timed) *n_timed = timed;
    return NICAM_OK;
}

}  // extern "C"
