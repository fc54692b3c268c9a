#!/usr/bin/env python3
"""NICAM tracer-transport + vertical-implicit step on one MI355X.

A step is one pass of the hot path over the glevel-7 workload: 40 regions of
66x66 cells (1-cell halo), 98 levels (96 layers, dz stretched 100 m -> 2 km),
plus the two pole regions, FP64 throughout:

    horizontal_flux -> horizontal_limiter_thuburn -> vertical_limiter_thuburn
    -> vi_rhow_solver

The grid is the generated glevel-7 icosahedral grid with its GMTR metrics
(the reference's dyn_metrics formulas); the fields are seeded synthetic
fields, all resident in HBM when the timed region starts.  Prints one JSON line (metric, value, roofline, cpu_baseline...).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Beside the FP64 step it times the mixed-precision step (the FP32 flux and
limiters, bit-identical to the reference built with -DSINGLE, with the FP64
VI solve) on the same inputs, with each field's error against FP64
("mixed" in the JSON line).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "NICAM adv+limiter+VI step: cell-levels/s & %HBM peak; adjoint/fwd ratio"
KERNELS = ("horizontal_flux", "horizontal_limiter_thuburn", "vertical_limiter_thuburn", "vi_rhow_solver")
DT_LARGE, DT_SMALL = 360.0, 60.0  # SET_dt_large / SET_dt_small (problem_size.inc:136-137)
EPS64, FLT_MAX = 2.0 ** -52, 3.4028234663852886e38


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class _CallRecorder:
    """Stands in for the library context and records which entry points a
    step calls (the calls go through to the context)."""

    def __init__(self, ctx):
        self._ctx, self.names = ctx, []

    def __getattr__(self, name):
        f = getattr(self._ctx, name)
        if not callable(f):
            return f

        def call(*a, **k):
            self.names.append(name)
            return f(*a, **k)

        return call


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-regions", type=int, default=40, help="regions in the CPU baseline sample")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip timing the neighbouring operators, the halo fill and the limiter's exchange mode")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint kernels' timing")
    ap.add_argument("--no-batched", action="store_true", help="skip the 32-tracer step (eager and HIP graph)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision step (FP32 flux and limiters, FP64 VI)")
    return ap.parse_args()


def _cpu_pass(impl, sub, geo, flux, hlim, vlim, vi, fac):
    """One pass of the four kernels; seconds per kernel."""
    times = {}
    t = time.perf_counter()
    impl.horizontal_flux(sub, geo, flux, DT_LARGE)
    times["horizontal_flux"] = time.perf_counter() - t
    t = time.perf_counter()
    impl.horizontal_limiter_thuburn(sub, geo["have_sgp"], hlim)
    times["horizontal_limiter_thuburn"] = time.perf_counter() - t
    t = time.perf_counter()
    impl.vertical_limiter_thuburn(sub, vlim)
    times["vertical_limiter_thuburn"] = time.perf_counter() - t
    t = time.perf_counter()
    impl.vi_rhow_solver(sub, vi, DT_SMALL, fac["rdgzh"], fac["afact"], fac["bfact"])
    times["vi_rhow_solver"] = time.perf_counter() - t
    return times


_MP = {}  # the sample the forked region workers read (inherited, not pickled)


def _mp_worker(regions):
    """One process's share of the regions, through the single-thread
    reference build; returns its wall time."""
    from oracle import binding
    from parity import region_slice

    from icoatmosbenchmark_v1_amd import synth

    spec, case, fac = _MP["spec"], _MP["case"], _MP["fac"]
    sub = synth.GridSpec(spec.gall_1d, spec.kall, len(regions), have_pl=regions[0] == 0, seed=spec.seed)
    pick = lambda d: region_slice(d, list(regions), pole=True)  # noqa: E731  (poles computed by the first share only)
    parts = {k: pick(case[k]) for k in ("geo", "flux", "hlim", "vlim", "vi")}
    impl = binding.RefLibs()
    t = time.perf_counter()
    _cpu_pass(impl, sub, parts["geo"], parts["flux"], parts["hlim"], parts["vlim"], parts["vi"], fac)
    return time.perf_counter() - t


def cpu_baseline(spec, case, regions: int, reps: int = 3):
    """The reference's own Fortran kernels (oracle/_ref) on the host cores of
    the machine the GPU sits in, on all 40 regions: one thread (the reference's documented runs), its
    -fopenmp build on every host core, and the regions split over one process
    per core (as NICAM's MPI decomposition splits them), each running the
    single-thread build.  The reported value is the fastest; the C
    restatement stands in when oracle/_ref is absent."""
    import multiprocessing as mp

    import numpy as np  # noqa: F401

    sys.path.insert(0, str(REPO / "tests"))
    from oracle import binding
    from parity import region_slice

    from icoatmosbenchmark_v1_amd import synth

    sub = synth.GridSpec(spec.gall_1d, spec.kall, regions, seed=spec.seed)
    pick = lambda d: region_slice(d, list(range(regions)), pole=True)  # noqa: E731
    geo = pick(case["geo"])
    flux, hlim, vlim, vi = pick(case["flux"]), pick(case["hlim"]), pick(case["vlim"]), pick(case["vi"])
    if not binding.ORACLE_SO.exists():
        binding.build(with_ref=False)
    orc = binding.Oracle()
    fac = orc.grd_setup(case["vi"]["gz"], case["vi"]["gzh"])  # GRD_setup, outside the timed part
    cpu_model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu_model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    ncpu = max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu))))
    if not binding.RefLibs.available():
        runs = sorted((_cpu_pass(orc, sub, geo, flux, hlim, vlim, vi, fac) for _ in range(reps)),
                      key=lambda r: sum(r.values()))
        total = sum(runs[len(runs) // 2].values())
        return dict(value=sub.inner_cell_levels() / total, unit="cell-levels/s", cores=orc.num_threads(),
                    kind="port", sample=f"{regions} regions, C restatement with OpenMP, {total:.2f} s per pass; "
                                        f"{cpu_model}")
    legs = {}
    # one thread
    runs = sorted((_cpu_pass(binding.RefLibs(), sub, geo, flux, hlim, vlim, vi, fac) for _ in range(reps)),
                  key=lambda r: sum(r.values()))
    legs["one_thread"] = dict(seconds=round(sum(runs[len(runs) // 2].values()), 4), cores=1,
                              kernels={k: round(v, 4) for k, v in runs[len(runs) // 2].items()})
    # the reference's own OpenMP loops on every core
    if binding.RefLibs.available(omp=True):
        ro = binding.RefLibs(omp=True)
        ro.set_threads(ncpu)
        runs = sorted((_cpu_pass(ro, sub, geo, flux, hlim, vlim, vi, fac) for _ in range(reps)),
                      key=lambda r: sum(r.values()))
        legs["openmp"] = dict(seconds=round(sum(runs[len(runs) // 2].values()), 4), cores=ncpu,
                              kernels={k: round(v, 4) for k, v in runs[len(runs) // 2].items()})
    # the regions over one process per core (fork: the sample is inherited)
    if ncpu > 1:
        _MP.update(spec=sub, case=dict(geo=geo, flux=flux, hlim=hlim, vlim=vlim, vi=vi), fac=fac)
        shares = [list(r) for r in np.array_split(np.arange(regions), min(ncpu, regions))]
        ctx = mp.get_context("fork")
        with ctx.Pool(len(shares)) as pool:
            pool.map(_mp_worker, shares)  # warm-up: libraries loaded in every worker
            walls = []
            for _ in range(reps):
                t = time.perf_counter()
                pool.map(_mp_worker, shares)
                walls.append(time.perf_counter() - t)
        _MP.clear()
        legs["region_processes"] = dict(seconds=round(sorted(walls)[len(walls) // 2], 4), cores=len(shares))
    best = min(legs, key=lambda k: legs[k]["seconds"])
    work = sub.inner_cell_levels()
    desc = {"one_thread": "1 thread", "openmp": "the reference's -fopenmp build",
            "region_processes": "the regions split over one process per core, single-thread build"}
    return dict(value=work / legs[best]["seconds"], unit="cell-levels/s", cores=legs[best]["cores"],
                kind="reference",
                sample=(f"{regions} of 40 regions (66x66 cells, 98 levels) + poles, one pass of the four kernels, "
                        f"median of {reps}; reference Fortran (flang -O2) on the host: "
                        + "; ".join(f"{desc[k]} ({v['cores']} core(s)) {v['seconds']:.3f} s" for k, v in legs.items())
                        + f"; reported: {desc[best]}; host {cpu_model}, {ncpu} usable core(s)"),
                legs={k: dict(v, value=work / v["seconds"]) for k, v in legs.items()})


def time_extras(ctx, spec, case, coef, stream, steps):
    """Per-launch times of what sits beside the step on the same gl7 grid:
    the limiter with its own Qout exchange (Qout_post = NULL), the halo fill
    of one field, OPRT_diffusion and OPRT3D_divdamp.  Reported beside the
    headline, not part of it."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, perfmodel, synth

    dev = torch.device("cuda:0")

    def up(d, keys):
        return {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in keys}

    def z(shape):
        return torch.zeros(shape, dtype=torch.float64, device=dev)

    def timed(fn):
        for _ in range(2):
            fn()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record(stream)
        for _ in range(steps):
            fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        return e[0].elapsed_time(e[1]) / steps

    out = {}
    t0 = time.perf_counter()
    ctx.comm_setup(1)
    H = up(case["hlim"], drivers.HLIM_IN + ("q_a", "q_a_pl"))
    H["Qout_prev"], H["Qout_prev_pl"] = z((2,) + spec.shape3), z((2,) + spec.shape_pl)
    ms = timed(lambda: ctx.horizontal_limiter_thuburn(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"],
                                                       H["ch"], H["ch_pl"], H["cmask"], H["cmask_pl"],
                                                       H["Qout_prev"], H["Qout_prev_pl"], None, None))
    out["horizontal_limiter_thuburn_exchange"] = dict(ms=round(ms, 5))
    del H
    var, var_pl = z(spec.shape3), z(spec.shape_pl)
    ms = timed(lambda: ctx.comm_data_transfer(var, var_pl, 1))
    out["comm_data_transfer_1var"] = dict(ms=round(ms, 5))
    del var, var_pl
    ob = perfmodel.operator_bytes(spec)
    f = synth.diffusion_inputs(spec, case["geo"], coef)
    D = up(f, drivers.DIFF_IN)
    D["dscl"], D["dscl_pl"] = z(spec.shape3), z(spec.shape_pl)
    ms = timed(lambda: ctx.oprt_diffusion(D["dscl"], D["dscl_pl"], *[D[k] for k in drivers.DIFF_IN]))
    gbs = ob["oprt_diffusion"] / (ms * 1e-3) / 1e9
    out["oprt_diffusion"] = dict(ms=round(ms, 5), alg_bytes=ob["oprt_diffusion"], gbs=round(gbs, 1),
                                 frac=round(gbs / perfmodel.HBM_PEAK_GBS, 4))
    del D, f
    f = synth.divdamp_inputs(spec, case["geo"], coef)
    drivers.setup_divdamp(ctx, f)
    G = up(f, drivers.DDMP_IN)
    for k in drivers.DDMP_OUT:
        G[k] = z(spec.shape_pl if k.endswith("_pl") else spec.shape3)
    ms = timed(lambda: ctx.oprt3d_divdamp(*[G[k] for k in drivers.DDMP_OUT], *[G[k] for k in drivers.DDMP_IN]))
    gbs = ob["oprt3d_divdamp"] / (ms * 1e-3) / 1e9
    out["oprt3d_divdamp"] = dict(ms=round(ms, 5), alg_bytes=ob["oprt3d_divdamp"], gbs=round(gbs, 1),
                                 frac=round(gbs / perfmodel.HBM_PEAK_GBS, 4))
    del G, f
    geo = case["geo"]
    if "coef_grad" in geo:  # the step glue between the flux and the limiter, on the generated grid
        F = up(case["flux"], drivers.FLUX_IN)
        flx, flx_pl = z((6,) + spec.shape3), z(spec.shape_pl)
        xc, xc_pl = z((3, 3) + spec.shape3), z((3,) + spec.shape_pl)
        ctx.horizontal_flux(flx, flx_pl, xc, xc_pl, *[F[k] for k in drivers.FLUX_IN], 360.0)
        Q = up(case["hlim"], ("q", "q_pl"))
        M = up(geo, ("coef_grad", "coef_grad_pl", "GRD_x", "GRD_x_pl"))
        ch, ch_pl, cm, cm_pl = z((6,) + spec.shape3), z(spec.shape_pl), z((6,) + spec.shape3), z(spec.shape_pl)
        grad, grad_pl = z((3,) + spec.shape3), z((3,) + spec.shape_pl)
        qa, qa_pl = z((6,) + spec.shape3), z(spec.shape_pl)
        rq, rq_pl = F["rho"] * Q["q"], F["rho_pl"] * Q["q_pl"]
        calls = dict(
            tracer_courant=lambda: ctx.tracer_courant(ch, ch_pl, cm, cm_pl, flx, flx_pl, F["rho"], F["rho_pl"]),
            oprt_gradient=lambda: ctx.oprt_gradient(grad, grad_pl, Q["q"], Q["q_pl"], M["coef_grad"],
                                                    M["coef_grad_pl"]),
            horizontal_remap=lambda: ctx.horizontal_remap(qa, qa_pl, Q["q"], Q["q_pl"], grad, grad_pl, cm, cm_pl, xc,
                                                          xc_pl, M["GRD_x"], M["GRD_x_pl"]),
            tracer_update=lambda: ctx.tracer_update(rq, rq_pl, flx, flx_pl, qa, qa_pl))
        for name, fn in calls.items():
            ms = timed(fn)
            gbs = ob[name] / (ms * 1e-3) / 1e9
            out[name] = dict(ms=round(ms, 5), alg_bytes=ob[name], gbs=round(gbs, 1),
                             frac=round(gbs / perfmodel.HBM_PEAK_GBS, 4))
        del F, Q, M, flx, xc, ch, cm, grad, qa, rq
    torch.cuda.synchronize()
    log("extras: " + ", ".join(f"{k}={v['ms']:.4f} ms" for k, v in out.items())
        + f" ({time.perf_counter() - t0:.1f} s)")
    return out


def time_chained(ctx, spec, case, stream, steps):
    """One chained tracer step, FP64 and mixed (drivers.chain_step): the flux,
    the horizontal limiter, the conservative horizontal update of rhogq, the
    vertical limiter, the vertical update, the rho*w solve, and the FP64
    global mass sum.  Reports each step's time, the conservation of the
    global tracer mass, and each field's error of the mixed step against the
    FP64 step: max |mixed - f64| / max |f64| over the field, and the share of
    points where |mixed - f64| exceeds 1e-5 of that maximum."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers

    ctx.comm_setup(1)
    ctx.set_const_f32(EPS64, FLT_MAX)
    ctx.grd_setup(case["vi"]["gz"], case["vi"]["gzh"])  # GRD_rdgz of this grid (the divdamp extras set their own)
    gzh = case["vi"]["gzh"]
    S = {m: drivers.chain_state(ctx, spec, case, mixed=m) for m in (False, True)}
    m0 = drivers.total_mass(ctx, S[False]["rhogq"], S[False]["rhogq_pl"], gzh)
    out = {}
    for m in (False, True):
        drivers.chain_step(ctx, S[m], DT_LARGE, DT_SMALL)
        mass = drivers.total_mass(ctx, S[m]["rhogq"], S[m]["rhogq_pl"], gzh)
        out["mixed" if m else "f64"] = dict(mass_rel_change=abs(mass - m0) / abs(m0))
    torch.cuda.synchronize()
    rho = torch.from_numpy(np.ascontiguousarray(case["flux"]["rho"])).to(S[False]["rhogq"].device)
    inner = torch.zeros(spec.gall_1d, spec.gall_1d, dtype=torch.bool, device=rho.device)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    pairs = dict(flx_h=("flx_h", None), GRD_xc=("GRD_xc", None), q_a=("h_q_a", None), Qout_prev=("Qout_prev", None),
                 q_h=("v_q_h", None), q=("rhogq", rho), rhogw=("rhogw", None))
    errors, share = {}, {}
    for name, (key, div) in pairs.items():
        a, b = S[True][key].double(), S[False][key]
        if div is not None:  # the updated mixing ratio on the inner cells
            a, b = (a / div)[..., inner], (b / div)[..., inner]
        scale = float(b.abs().max())
        err = (a - b).abs() / scale
        errors[name] = float(err.max())
        share[name] = float((err > 1e-5).double().mean())
    mass_part = torch.zeros(spec.kall, dtype=torch.float64, device=rho.device)

    def step(m):
        drivers.chain_step(ctx, S[m], DT_LARGE, DT_SMALL)
        ctx.tracer_mass(S[m]["rhogq"], S[m]["rhogq_pl"], mass_part)

    for m in (False, True):
        for _ in range(2):
            step(m)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        e[0].record(stream)
        for _ in range(steps):
            step(m)
        e[1].record(stream)
        torch.cuda.synchronize()
        out["mixed" if m else "f64"]["ms_per_step"] = round(e[0].elapsed_time(e[1]) / steps, 5)
    out["speedup_mixed_vs_f64"] = round(out["f64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    out["rel_error_vs_f64"] = errors
    out["share_above_1em5"] = share
    out["steps"] = ("horizontal_flux, horizontal_limiter_thuburn, tracer_update_conservative, "
                    "vertical_limiter_thuburn, tracer_update_vertical, vi_rhow_solver, tracer_mass")
    del S
    torch.cuda.empty_cache()
    log(f"chained step: f64 {out['f64']['ms_per_step']} ms, mixed {out['mixed']['ms_per_step']} ms; "
        f"mass change f64 {out['f64']['mass_rel_change']:.2e}, mixed {out['mixed']['mass_rel_change']:.2e}; "
        f"q error {errors['q']:.2e}, rhogw error {errors['rhogw']:.2e}")
    return out


def time_mixed(ctx, spec, case, stream, steps):
    """The mixed-precision step: the FP32 flux and limiters (with FP64's
    CONST_EPS in the limiters' switches) with the mixed rho*w solve (FP32
    state and sources, FP64 elimination), on the same inputs rounded to FP32;
    per-kernel times and algorithmic bytes.  The errors against FP64 come from
    the chained step (time_chained)."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, perfmodel

    dev = torch.device("cuda:0")

    def up32(d, keys):
        return {k: torch.from_numpy(np.ascontiguousarray(d[k], dtype=np.float32)).to(dev) for k in keys}

    def z(shape):
        return torch.zeros(shape, dtype=torch.float32, device=dev)

    ctx.set_const_f32(EPS64, FLT_MAX)
    F = up32(case["flux"], drivers.FLUX_IN)
    F.update(flx_h=z((6,) + spec.shape3), flx_h_pl=z(spec.shape_pl), GRD_xc=z((3, 3) + spec.shape3),
             GRD_xc_pl=z((3,) + spec.shape_pl))
    H = up32(case["hlim"], drivers.HLIM_IN + ("q_a", "q_a_pl"))
    H.update(Qout_prev=z((2,) + spec.shape3), Qout_prev_pl=z((2,) + spec.shape_pl))
    V = up32(case["vlim"], drivers.VLIM_IN + ("q_h", "q_h_pl"))

    def k_flux():
        ctx.horizontal_flux_f32(F["flx_h"], F["flx_h_pl"], F["GRD_xc"], F["GRD_xc_pl"],
                                *[F[k] for k in drivers.FLUX_IN], DT_LARGE)

    def k_hlim():
        ctx.horizontal_limiter_thuburn_f32(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                           H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_prev"],
                                           H["Qout_prev_pl"], H["Qout_post"], H["Qout_post_pl"])

    def k_vlim():
        ctx.vertical_limiter_thuburn_f32(V["q_h"], V["q_h_pl"], V["q"], V["q_pl"], V["d"], V["d_pl"], V["ck"],
                                         V["ck_pl"])

    # the solve with FP32 state and sources, FP64 rho*w and elimination
    vkeys = list(drivers.VI_IN)
    W32 = {k + s: torch.from_numpy(np.ascontiguousarray(case["vi"][k + s], dtype=np.float32)).to(dev)
           for k in vkeys for s in ("", "_pl")}
    w64 = {s: torch.from_numpy(np.ascontiguousarray(case["vi"]["rhogw" + s])).to(dev) for s in ("", "_pl")}

    def k_vi_mixed():
        a = []
        for k in vkeys:
            a += [W32[k], W32[k + "_pl"]]
        ctx.vi_rhow_solver_mixed(w64[""], w64["_pl"], *a, DT_SMALL)

    calls = (k_flux, k_hlim, k_vlim, k_vi_mixed)
    for _ in range(2):
        for f in calls:
            f()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(calls) + 1)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for st in range(steps):
        ev[st][0].record(stream)
        for i, f in enumerate(calls):
            f()
            ev[st][i + 1].record(stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    kms = {name: float(np.mean([ev[st][i].elapsed_time(ev[st][i + 1]) for st in range(steps)]))
           for i, name in enumerate(KERNELS)}
    b32, b64 = perfmodel.kernel_bytes(spec, word=4), perfmodel.kernel_bytes(spec)
    kb = dict(b32, vi_rhow_solver=perfmodel.vi_mixed_bytes(spec))
    kernels = {}
    for name in KERNELS:
        gbs = kb[name] / (kms[name] * 1e-3) / 1e9
        kernels[name] = dict(ms=round(kms[name], 5), alg_bytes=kb[name], gbs=round(gbs, 1),
                             frac=round(gbs / perfmodel.HBM_PEAK_GBS, 4))
    step_gbs = sum(kb.values()) / (ms_step * 1e-3) / 1e9
    log(f"mixed step: {ms_step:.4f} ms; " + ", ".join(f"{k}={v:.4f}" for k, v in kms.items()))
    return dict(dtype="f32 flux, limiters and VI sources; f64 rho*w and tridiagonal elimination", ms_per_step=ms_step,
                value=spec.inner_cell_levels() / (ms_step * 1e-3), unit="cell-levels/s",
                hbm_frac_step=round(step_gbs / perfmodel.HBM_PEAK_GBS, 4), step_alg_bytes=sum(kb.values()),
                kernels=kernels, eps="FP64 CONST_EPS (2^-52) in the FP32 limiters")


def time_adjoint(ctx, spec, case, stream, steps, fwd_ms):
    """The masked forward limiters and the adjoint kernels on the same gl7
    grid: per-launch times, algorithmic bytes, and each adjoint's cost
    against its forward kernel's."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, perfmodel

    dev = torch.device("cuda:0")

    def up(d, keys):
        return {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in keys}

    def timed(fn):
        for _ in range(2):
            fn()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record(stream)
        for _ in range(steps):
            fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        return e[0].elapsed_time(e[1]) / steps

    R, P = ctx.nreg, ctx.npl
    out, masked = {}, {}
    H = up(case["hlim"], drivers.HLIM_IN + ("q_a", "q_a_pl"))
    qa0, qa0_pl = H["q_a"].clone(), H["q_a_pl"].clone()
    H["Qout_prev"] = torch.zeros((2,) + spec.shape3, dtype=torch.float64, device=dev)
    H["Qout_prev_pl"] = torch.zeros((2,) + spec.shape_pl, dtype=torch.float64, device=dev)
    hm = torch.zeros(R, dtype=torch.uint8, device=dev)
    hm_pl = torch.zeros(P, dtype=torch.uint8, device=dev)

    def hlim_mask():
        ctx.horizontal_limiter_thuburn_mask(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                            H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_prev"], H["Qout_prev_pl"],
                                            H["Qout_post"], H["Qout_post_pl"], hm, hm_pl)

    masked["horizontal_limiter_thuburn"] = timed(hlim_mask)
    H["q_a"].copy_(qa0)
    H["q_a_pl"].copy_(qa0_pl)
    hlim_mask()
    gen = torch.Generator(device=dev).manual_seed(5)
    rnd = lambda shp: torch.randn(shp, dtype=torch.float64, device=dev, generator=gen)  # noqa: E731
    qab, qab_pl = rnd((6,) + spec.shape3), rnd(spec.shape_pl)
    qpb, qpb_pl = rnd((2,) + spec.shape3), rnd((2,) + spec.shape_pl)
    qb, qb_pl = torch.empty(spec.shape3, dtype=torch.float64, device=dev), torch.empty(spec.shape_pl,
                                                                                       dtype=torch.float64, device=dev)
    pob, pob_pl = torch.empty_like(qpb), torch.empty_like(qpb_pl)
    out["horizontal_limiter_thuburn"] = timed(lambda: ctx.horizontal_limiter_thuburn_adjoint(
        qab, qab_pl, qb, qb_pl, qpb, qpb_pl, pob, pob_pl, H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"],
        H["d_pl"], H["ch"], H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_post"], H["Qout_post_pl"], hm, hm_pl))
    del H, qab, qab_pl, qpb, qpb_pl, qb, qb_pl, pob, pob_pl

    V = up(case["vlim"], drivers.VLIM_IN + ("q_h", "q_h_pl"))
    qh0, qh0_pl = V["q_h"].clone(), V["q_h_pl"].clone()
    M, MP = ctx.vlim_mask_words()
    vm = torch.zeros(M, dtype=torch.int32, device=dev)
    vm_pl = torch.zeros(max(MP, 1), dtype=torch.int32, device=dev)[:MP]

    def vlim_mask():
        ctx.vertical_limiter_thuburn_mask(V["q_h"], V["q_h_pl"], V["q"], V["q_pl"], V["d"], V["d_pl"], V["ck"],
                                          V["ck_pl"], vm, vm_pl)

    masked["vertical_limiter_thuburn"] = timed(vlim_mask)
    V["q_h"].copy_(qh0)
    V["q_h_pl"].copy_(qh0_pl)
    vlim_mask()
    qhb, qhb_pl = rnd(spec.shape3), rnd(spec.shape_pl)
    vqb, vqb_pl = torch.empty_like(qhb), torch.empty_like(qhb_pl)
    out["vertical_limiter_thuburn"] = timed(lambda: ctx.vertical_limiter_thuburn_adjoint(
        qhb, qhb_pl, vqb, vqb_pl, V["q_h"], V["q_h_pl"], V["q"], V["q_pl"], V["d"], V["d_pl"], V["ck"], V["ck_pl"],
        vm, vm_pl))
    del V, qhb, qhb_pl, vqb, vqb_pl

    F = up(case["flux"], drivers.FLUX_IN)
    fb = [rnd((6,) + spec.shape3), rnd(spec.shape_pl), rnd((3, 3) + spec.shape3), rnd((3,) + spec.shape_pl)]
    xb = [torch.empty(spec.shape3 if i % 2 == 0 else spec.shape_pl, dtype=torch.float64, device=dev)
          for i in range(8)]
    out["horizontal_flux"] = timed(lambda: ctx.horizontal_flux_adjoint(*fb, *xb, *[F[k] for k in drivers.FLUX_IN],
                                                                       DT_LARGE))
    del F, fb, xb

    wb = {s: rnd(spec.shape3 if s == "" else spec.shape_pl) for s in ("", "_pl")}
    bufs = [torch.empty(spec.shape3 if i % 2 == 0 else spec.shape_pl, dtype=torch.float64, device=dev)
            for i in range(12)]
    out["vi_rhow_solver"] = timed(lambda: ctx.vi_rhow_solver_adjoint(wb[""], wb["_pl"], *bufs, DT_SMALL))
    del wb, bufs
    torch.cuda.synchronize()

    ab = perfmodel.adjoint_bytes(spec)
    kernels = {}
    for name, ms in out.items():
        gbs = ab[name] / (ms * 1e-3) / 1e9
        kernels[name] = dict(ms=round(ms, 5), fwd_ms=round(fwd_ms[name], 5), ratio=round(ms / fwd_ms[name], 3),
                             alg_bytes=ab[name], gbs=round(gbs, 1), frac=round(gbs / perfmodel.HBM_PEAK_GBS, 4))
    adj = sum(out.values())
    fwd = sum(fwd_ms[k] for k in out)
    # the reverse pass of a checkpointed run recomputes the step's forward (the limiters masked) and then
    # runs the adjoints: its cost over the plain forward's is the ratio the gradient pays per step
    recompute = fwd + sum(masked.values()) - sum(fwd_ms[k] for k in masked)
    return dict(kernels=kernels,
                masked_forward_ms={k: round(v, 5) for k, v in masked.items()},
                covered=sorted(out), ratio_covered=round(adj / fwd, 3),
                recompute_ms=round(recompute, 5),
                ratio_with_recompute=round((recompute + adj) / fwd, 3))


BATCH_GRIDS = ((5, 98), (6, 78), (7, 58), (8, 42))  # (glevel, kall) of the four rlevel-1 grids


def time_batched(geo7, stream, steps, ntr: int = 32):
    """ntr tracers on four grids of uneven size (glevel 5..8, rlevel 1, kall
    98/78/58/42), FP32 flux and limiters with FP64's CONST_EPS, the mixed
    rho*w solve: the group step (four launches over all grids and tracers,
    launch shapes from the autotuner) timed eagerly and as its captured HIP
    graph, against the same work as per-call launches (one grid and one
    tracer per call, 4 x (2 + 2 ntr) launches).  Each tracer's fields are the
    base fields shifted along the cells."""
    import tempfile

    import torch

    from icoatmosbenchmark_v1_amd import drivers, perfmodel, synth
    from icoatmosbenchmark_v1_amd.group import Group, batched_state

    t0 = time.perf_counter()
    side = torch.cuda.Stream()  # graph capture needs a stream of its own (not the null stream)
    side.wait_stream(torch.cuda.current_stream())
    stream = side
    specs, ctxs, states = [], [], []
    for gl, kall in BATCH_GRIDS:
        geo = geo7 if gl == 7 else synth.icosahedral_geometry(gl, 1)[0]
        spec = synth.GridSpec(2 ** (gl - 1) + 2, kall, 40, seed=100 + gl)
        case = synth.make_case(spec, geo=geo)
        c = drivers.setup_context(spec, geo, case["vi"])
        c.set_stream(stream)
        c.set_const_f32(EPS64, FLT_MAX)
        specs.append(spec)
        ctxs.append(c)
        states.append(batched_state(spec, case, ntr))
        del case
    torch.cuda.synchronize()
    log(f"batched: {len(specs)} grids x {ntr} tracers set up in {time.perf_counter() - t0:.1f} s, "
        f"device memory {torch.cuda.memory_allocated() / 1e9:.1f} GB")
    grp = Group(ctxs)
    grp.set_stream(stream)
    with tempfile.TemporaryDirectory() as tmp:
        n_timed = grp.autotune(ntr, states, DT_LARGE, DT_SMALL, cache_path=f"{tmp}/tune.txt")
        tuned = open(f"{tmp}/tune.txt").read().strip().splitlines()

    def timed(fn):
        for _ in range(2):
            fn()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        e[0].record(stream)
        for _ in range(steps):
            fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        return e[0].elapsed_time(e[1]) / steps

    eager = timed(lambda: grp.tracer_step_f32(ntr, states, DT_LARGE, DT_SMALL, use_graph=False))
    ref = [S[k].clone() for S in states for k in ("q_a", "q_h")]
    graph = timed(lambda: grp.tracer_step_f32(ntr, states, DT_LARGE, DT_SMALL, use_graph=True))
    # the limiters are idempotent at fixed bounds: the replays reproduce the eager faces
    same = all(torch.equal(a, S[k]) for a, (S, k) in zip(ref, [(S, k) for S in states for k in ("q_a", "q_h")]))

    def per_call():
        for c, S in zip(ctxs, states):
            c.horizontal_flux_f32(S["flx_h"], S["flx_h_pl"], S["GRD_xc"], S["GRD_xc_pl"], S["rho"], S["rho_pl"],
                                  S["rhovx"], S["rhovx_pl"], S["rhovy"], S["rhovy_pl"], S["rhovz"], S["rhovz_pl"],
                                  DT_LARGE)
            for t in range(ntr):
                c.horizontal_limiter_thuburn_f32(S["q_a"][t], S["q_a_pl"][t], S["q"][t], S["q_pl"][t], S["d"],
                                                 S["d_pl"], S["ch"], S["ch_pl"], S["cmask"], S["cmask_pl"],
                                                 S["Qout_prev"][t], S["Qout_prev_pl"][t], S["Qout_post"][t],
                                                 S["Qout_post_pl"][t])
                c.vertical_limiter_thuburn_f32(S["q_h"][t], S["q_h_pl"][t], S["qv"][t], S["qv_pl"][t], S["dv"],
                                               S["dv_pl"], S["ck"], S["ck_pl"])
            a = []
            for k in drivers.VI_IN:
                a += [S[k], S[k + "_pl"]]
            c.vi_rhow_solver_mixed(S["rhogw"], S["rhogw_pl"], *a, DT_SMALL)

    calls = timed(per_call)
    alg = sum(sum(perfmodel.batched_bytes(sp, ntr).values()) for sp in specs)
    work = sum(ntr * sp.inner_cell_levels() for sp in specs)
    grp.close()
    for c in ctxs:
        c.close()
    del states
    torch.cuda.empty_cache()
    return dict(tracers=ntr, grids=[dict(glevel=gl, rlevel=1, gall_1d=sp.gall_1d, kall=sp.kall, lall=sp.lall)
                                    for (gl, _), sp in zip(BATCH_GRIDS, specs)],
                dtype="f32 flux and limiters (FP64 CONST_EPS), mixed rho*w solve",
                graph_ms=round(graph, 4), eager_ms=round(eager, 4), per_call_ms=round(calls, 4),
                launches_per_step=4, per_call_launches=len(specs) * (2 + 2 * ntr),
                value=work / (graph * 1e-3), unit="tracer cell-levels/s", alg_bytes=alg,
                hbm_frac=round(alg / (graph * 1e-3) / 1e9 / perfmodel.HBM_PEAK_GBS, 4),
                graph_matches_eager=bool(same), autotuner=dict(candidates_timed=n_timed, choices=tuned))


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the hot path is built and "
              f"measured on a single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, perfmodel, runtime, synth

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if not runtime.LIB_PATH.exists():
        from icoatmosbenchmark_v1_amd import _build

        _build.build_library()
    dev = torch.device("cuda:0")
    spec = synth.GridSpec(66, 98, 40, seed=args.seed)
    t0 = time.perf_counter()
    geo, coef = synth.icosahedral_geometry(7, 1)  # the grid, its GMTR metrics and operator coefficients
    case = synth.make_case(spec, geo=geo)
    log(f"inputs generated in {time.perf_counter() - t0:.1f} s")
    geo, vi = case["geo"], case["vi"]
    ctx = drivers.setup_context(spec, geo, vi)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream)

    def up(d, keys):
        return {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in keys}

    F = up(case["flux"], drivers.FLUX_IN)
    F["flx_h"] = torch.zeros((6,) + spec.shape3, dtype=torch.float64, device=dev)
    F["flx_h_pl"] = torch.zeros(spec.shape_pl, dtype=torch.float64, device=dev)
    F["GRD_xc"] = torch.zeros((3, 3) + spec.shape3, dtype=torch.float64, device=dev)
    F["GRD_xc_pl"] = torch.zeros((3,) + spec.shape_pl, dtype=torch.float64, device=dev)
    H = up(case["hlim"], drivers.HLIM_IN + ("q_a", "q_a_pl"))
    H["Qout_prev"] = torch.zeros((2,) + spec.shape3, dtype=torch.float64, device=dev)
    H["Qout_prev_pl"] = torch.zeros((2,) + spec.shape_pl, dtype=torch.float64, device=dev)
    V = up(case["vlim"], drivers.VLIM_IN + ("q_h", "q_h_pl"))
    vkeys = [k for k in drivers.VI_IN] + [k + "_pl" for k in drivers.VI_IN] + ["rhogw", "rhogw_pl"]
    W = up(vi, vkeys)
    torch.cuda.synchronize()
    log(f"device memory in use: {torch.cuda.memory_allocated() / 1e9:.2f} GB + context {ctx.device_bytes() / 1e9:.2f} GB")

    # each headline call takes the library handle, so one pass can be checked
    # through a recorder: the step must call exactly the four FP64 entry points
    def k_flux(c):
        c.horizontal_flux(F["flx_h"], F["flx_h_pl"], F["GRD_xc"], F["GRD_xc_pl"], *[F[k] for k in drivers.FLUX_IN],
                          DT_LARGE)

    def k_hlim(c):
        c.horizontal_limiter_thuburn(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                     H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_prev"], H["Qout_prev_pl"],
                                     H["Qout_post"], H["Qout_post_pl"])

    def k_vlim(c):
        c.vertical_limiter_thuburn(V["q_h"], V["q_h_pl"], V["q"], V["q_pl"], V["d"], V["d_pl"], V["ck"],
                                   V["ck_pl"])

    def k_vi(c):
        a = []
        for k in drivers.VI_IN:
            a += [W[k], W[k + "_pl"]]
        c.vi_rhow_solver(W["rhogw"], W["rhogw_pl"], *a, DT_SMALL)

    calls = (k_flux, k_hlim, k_vlim, k_vi)
    rec = _CallRecorder(ctx)
    for f in calls:
        f(rec)
    if rec.names != list(KERNELS):
        raise AssertionError(f"the FP64 headline step must call {list(KERNELS)}, it called {rec.names}")
    fields = [t for d in (F, H, V, W) for t in d.values() if t.is_floating_point()]
    if any(t.dtype != torch.float64 for t in fields):
        raise AssertionError("the FP64 headline step was handed a non-FP64 field")
    for _ in range(args.warmup):
        for f in calls:
            f(ctx)
    torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(calls) + 1)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ev[s][0].record(stream)
        for i, f in enumerate(calls):
            f(ctx)
            ev[s][i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    kms = {name: float(np.mean([ev[s][i].elapsed_time(ev[s][i + 1]) for s in range(args.steps)]))
           for i, name in enumerate(KERNELS)}
    log("per-kernel ms: " + ", ".join(f"{k}={v:.4f}" for k, v in kms.items()))

    inner = spec.inner_cell_levels()
    kb = perfmodel.kernel_bytes(spec)
    step_bytes = sum(kb.values())
    kernels = {}
    for name in KERNELS:
        gbs = kb[name] / (kms[name] * 1e-3) / 1e9
        kernels[name] = dict(ms=round(kms[name], 5), alg_bytes=kb[name], gbs=round(gbs, 1),
                             frac=round(gbs / perfmodel.HBM_PEAK_GBS, 4))
    dom = max(KERNELS, key=lambda k: kms[k])
    traffic = None
    tfile = REPO / "profiles" / "traffic_gl7_fp64.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(dom)
        except (ValueError, OSError):
            traffic = None
    roof = dict(bound="hbm", kernel=dom, achieved=kernels[dom]["gbs"], peak=perfmodel.HBM_PEAK_GBS, unit="GB/s",
                frac=kernels[dom]["frac"], traffic=traffic)
    step_gbs = step_bytes / (ms_step * 1e-3) / 1e9

    extras = None if args.no_extras else time_extras(ctx, spec, case, coef, stream, args.steps)

    mixed = chained = None
    if not args.no_mixed:
        mixed = time_mixed(ctx, spec, case, stream, args.steps)
        mixed["speedup_vs_f64"] = round(ms_step / mixed["ms_per_step"], 3)
        chained = time_chained(ctx, spec, case, stream, max(3, args.steps // 2))

    batched = None
    if not args.no_batched:
        batched = time_batched(geo, stream, max(3, args.steps // 4))
        log(f"batched {batched['tracers']} tracers on {len(batched['grids'])} grids: graph {batched['graph_ms']} ms, "
            f"eager {batched['eager_ms']} ms, per-call {batched['per_call_ms']} ms, "
            f"{batched['hbm_frac']:.3f} of 8 TB/s")

    adjoint = None
    if not args.no_adjoint:
        adjoint = time_adjoint(ctx, spec, case, stream, args.steps, kms)
        log(f"adjoint/forward over {adjoint['covered']}: {adjoint['ratio_covered']} (adjoints only), "
            f"{adjoint['ratio_with_recompute']} (with the recomputed masked forward)")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            log(f"CPU baseline on {args.cpu_regions} regions ...")
            cpu = cpu_baseline(spec, case, args.cpu_regions)
            log(f"CPU baseline: {cpu['value']:.4g} cell-levels/s ({cpu['kind']}, {cpu['cores']} core(s))")
        except Exception as e:  # reported, never fatal
            log(f"CPU baseline failed: {e!r}")
            cpu = dict(value=None, unit="cell-levels/s", cores=None, kind=None, sample=f"failed: {e!r}")

    out = dict(
        metric=METRIC,
        value=inner / (ms_step * 1e-3),
        unit="cell-levels/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=ms_step,
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f64",
        data="synthetic (seeded fields on the generated gl7 icosahedral grid and its metrics)",
        config=dict(workload="NICAM gl7 rl1: 40 regions x 66x66 cells x 98 levels + 2 poles, FP64, "
                             "horizontal_flux + horizontal_limiter_thuburn + vertical_limiter_thuburn + vi_rhow_solver",
                    gall_1d=66, kall=98, lall=40, inner_cell_levels=inner, array_cell_levels=spec.array_cell_levels(),
                    tracers=1),
        hbm_frac_step=round(step_gbs / perfmodel.HBM_PEAK_GBS, 4),
        step_alg_bytes=step_bytes,
        adjoint_fwd_ratio=None if adjoint is None else adjoint["ratio_with_recompute"],
        adjoint=adjoint,
        batched=batched,
        kernels=kernels,
        extras=extras,
        mixed=mixed,
        chained=chained,
        roofline=roof,
        cpu_baseline=cpu,
        host=platform.node(),
    )
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
