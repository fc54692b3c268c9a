"""Profiling driver: runs the benchmark step a few times for rocprofv3.

    rocprofv3 --kernel-trace --stats -d OUT -- python3 -m icoatmosbenchmark_v1_amd.profile_step
    rocprofv3 --pmc FETCH_SIZE      -d OUT -- python3 -m icoatmosbenchmark_v1_amd.profile_step
    rocprofv3 --pmc WRITE_SIZE      -d OUT -- python3 -m icoatmosbenchmark_v1_amd.profile_step

Before the step it launches one plain device-to-device copy of exactly
--copy-mib MiB (torch's copy kernel), whose FETCH_SIZE / WRITE_SIZE readings
calibrate the counters: bytes moved / counter reading gives the factor that
turns the kernels' counter readings into HBM bytes.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--gall-1d", type=int, default=66)
    ap.add_argument("--kall", type=int, default=98)
    ap.add_argument("--lall", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()

    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, synth

    dev = torch.device("cuda:0")
    n = args.copy_mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"[profile_step] copy of {args.copy_mib} MiB done", flush=True)
    del a, b

    spec = synth.GridSpec(args.gall_1d, args.kall, args.lall, seed=args.seed)
    lv = synth.icosahedral_levels(spec.gall_1d, spec.lall)
    geo = synth.icosahedral_geometry(*lv)[0] if lv else None
    case = synth.make_case(spec, geo=geo)
    ctx = drivers.setup_context(spec, case["geo"], case["vi"])
    ctx.set_stream(torch.cuda.current_stream(dev))

    def up(d, keys):
        return {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in keys}

    F = up(case["flux"], drivers.FLUX_IN)
    z = lambda shape: torch.zeros(shape, dtype=torch.float64, device=dev)  # noqa: E731
    F.update(flx_h=z((6,) + spec.shape3), flx_h_pl=z(spec.shape_pl), GRD_xc=z((3, 3) + spec.shape3),
             GRD_xc_pl=z((3,) + spec.shape_pl))
    H = up(case["hlim"], drivers.HLIM_IN + ("q_a", "q_a_pl"))
    H.update(Qout_prev=z((2,) + spec.shape3), Qout_prev_pl=z((2,) + spec.shape_pl))
    V = up(case["vlim"], drivers.VLIM_IN + ("q_h", "q_h_pl"))
    W = up(case["vi"], [k for k in drivers.VI_IN] + [k + "_pl" for k in drivers.VI_IN] + ["rhogw", "rhogw_pl"])
    torch.cuda.synchronize()
    for s in range(args.steps):
        ctx.horizontal_flux(F["flx_h"], F["flx_h_pl"], F["GRD_xc"], F["GRD_xc_pl"], *[F[k] for k in drivers.FLUX_IN],
                            360.0)
        ctx.horizontal_limiter_thuburn(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                       H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_prev"], H["Qout_prev_pl"],
                                       H["Qout_post"], H["Qout_post_pl"])
        ctx.vertical_limiter_thuburn(V["q_h"], V["q_h_pl"], V["q"], V["q_pl"], V["d"], V["d_pl"], V["ck"], V["ck_pl"])
        args_vi = []
        for k in drivers.VI_IN:
            args_vi += [W[k], W[k + "_pl"]]
        ctx.vi_rhow_solver(W["rhogw"], W["rhogw_pl"], *args_vi, 60.0)
        torch.cuda.synchronize()
        print(f"[profile_step] step {s + 1}/{args.steps} done", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
