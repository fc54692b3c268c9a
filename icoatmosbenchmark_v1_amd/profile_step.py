"""Profiling driver: runs the benchmark step a few times for rocprofv3.

    rocprofv3 --kernel-trace --stats -d OUT -- python3 -m icoatmosbenchmark_v1_amd.profile_step
    rocprofv3 --pmc FETCH_SIZE      -d OUT -- python3 -m icoatmosbenchmark_v1_amd.profile_step
    rocprofv3 --pmc WRITE_SIZE      -d OUT -- python3 -m icoatmosbenchmark_v1_amd.profile_step

--mode picks what each step launches:
  step      the four FP64 kernels (default)
  chained   the chained FP64 and mixed steps: the _f32 flux and limiters, the
            mixed rho*w solve, the conservative and vertical updates, the
            mass sums; the limiter with the library's own Qout exchange
            (FP64 and FP32) and one COMM_data_transfer
  vi        every rho*w solve variant (1-4) with FP64 and FP32 inputs, the
            factorisation (putcoef) and both adjoint variants
  batched   the grouped multi-tracer step on the gl5-gl8 grids (32 tracers)
  extras    OPRT_diffusion, OPRT3D_divdamp, the horizontal and vertical step
            glue and the Mc/Ml/Mu update
  adjoint   the masked limiters and the four adjoints
  glue_adjoint  the adjoints of the horizontal and vertical glue and of the
            halo fill (COMM_data_transfer), and the h-limiter adjoint's
            two half passes
  coverage  what the other modes leave out: the flux and h-limiter
            instances for wide regions (gall_1d 258, 322, 514; FP64 and
            FP32, one and two tracers), the rho*w adjoint with its column in
            registers (kall 42) and the multi-tracer vertical limiter

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
    ap.add_argument("--mode", default="step", choices=("step", "chained", "vi", "batched", "extras", "adjoint",
                                                         "glue_adjoint", "coverage"))
    ap.add_argument("--extras", action="store_true", help="same as --mode extras")
    ap.add_argument("--adjoint", action="store_true", help="same as --mode adjoint")
    args = ap.parse_args()
    if args.extras:
        args.mode = "extras"
    if args.adjoint:
        args.mode = "adjoint"

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
    if args.mode == "batched":
        return batched(args.steps)
    if args.mode == "coverage":
        return coverage(args.steps)

    spec = synth.GridSpec(args.gall_1d, args.kall, args.lall, seed=args.seed)
    lv = synth.icosahedral_levels(spec.gall_1d, spec.lall)
    geo, coef = synth.icosahedral_geometry(*lv) if lv else (None, None)
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
    if args.mode == "extras":
        return extras(ctx, spec, case, coef, args.steps, up, z)
    if args.mode == "adjoint":
        return adjoint(ctx, spec, args.steps, F, H, V, W, z)
    if args.mode == "glue_adjoint":
        if lv is None:
            print("[profile_step] --mode glue_adjoint needs an icosahedral grid size", file=sys.stderr)
            return 2
        ctx.comm_setup(lv[1])
        return glue_adjoint(ctx, spec, case, args.steps, F, H, up, z)
    if args.mode == "chained":
        return chained(ctx, spec, case, args.steps, z)
    if args.mode == "vi":
        return vi_variants(ctx, spec, args.steps, W)
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


def adjoint(ctx, spec, steps, F, H, V, W, z) -> int:
    """One reverse pass per step: the masked limiters, then the adjoints of
    the flux, both limiters and the solve, each on random cotangents."""
    import torch

    from icoatmosbenchmark_v1_amd import drivers

    dev = F["rho"].device
    R, P = ctx.nreg, ctx.npl
    hm, hm_pl = torch.zeros(R, dtype=torch.uint8, device=dev), torch.zeros(P, dtype=torch.uint8, device=dev)
    M, MP = ctx.vlim_mask_words()
    vm = torch.zeros(M, dtype=torch.int32, device=dev)
    vm_pl = torch.zeros(max(MP, 1), dtype=torch.int32, device=dev)[:MP]
    r = lambda shape: torch.randn(shape, dtype=torch.float64, device=dev)  # noqa: E731
    fb = [r((6,) + spec.shape3), r(spec.shape_pl), r((3, 3) + spec.shape3), r((3,) + spec.shape_pl)]
    xb = [z(spec.shape3 if i % 2 == 0 else spec.shape_pl) for i in range(8)]
    qab, qab_pl, qpb, qpb_pl = r((6,) + spec.shape3), r(spec.shape_pl), r((2,) + spec.shape3), r((2,) + spec.shape_pl)
    qb, qb_pl, pob, pob_pl = z(spec.shape3), z(spec.shape_pl), z((2,) + spec.shape3), z((2,) + spec.shape_pl)
    qhb, qhb_pl, vqb, vqb_pl = r(spec.shape3), r(spec.shape_pl), z(spec.shape3), z(spec.shape_pl)
    wb, wb_pl = r(spec.shape3), r(spec.shape_pl)
    bufs = [z(spec.shape3 if i % 2 == 0 else spec.shape_pl) for i in range(12)]
    for s in range(steps):
        ctx.horizontal_limiter_thuburn_mask(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                            H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_prev"], H["Qout_prev_pl"],
                                            H["Qout_post"], H["Qout_post_pl"], hm, hm_pl)
        ctx.vertical_limiter_thuburn_mask(V["q_h"], V["q_h_pl"], V["q"], V["q_pl"], V["d"], V["d_pl"], V["ck"],
                                          V["ck_pl"], vm, vm_pl)
        ctx.horizontal_flux_adjoint(*fb, *xb, *[F[k] for k in drivers.FLUX_IN], 360.0)
        ctx.horizontal_limiter_thuburn_adjoint(qab, qab_pl, qb, qb_pl, qpb, qpb_pl, pob, pob_pl, H["q_a"],
                                               H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                               H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_post"],
                                               H["Qout_post_pl"], hm, hm_pl)
        ctx.vertical_limiter_thuburn_adjoint(qhb, qhb_pl, vqb, vqb_pl, V["q_h"], V["q_h_pl"], V["q"], V["q_pl"],
                                             V["d"], V["d_pl"], V["ck"], V["ck_pl"], vm, vm_pl)
        ctx.vi_rhow_solver_adjoint(wb, wb_pl, *bufs, 60.0)
        torch.cuda.synchronize()
        print(f"[profile_step] adjoint {s + 1}/{steps} done", flush=True)
    ctx.close()
    return 0


def glue_adjoint(ctx, spec, case, steps, F, H, up, z) -> int:
    """The glue's adjoints on random cotangents at the forward's own inputs:
    gradient, Courant numbers, upwind remap, tracer update, the halo fill's
    transpose (one and three variables), the vertical Courant numbers, face
    values and update; then the h-limiter adjoint's rows and faces passes."""
    import torch

    from icoatmosbenchmark_v1_amd import drivers

    dev = F["rho"].device
    M = up(case["geo"], ("coef_grad", "coef_grad_pl", "GRD_x", "GRD_x_pl"))
    X, X_pl = M["GRD_x"][:, :, 0, :].contiguous(), M["GRD_x_pl"][:, :, 0, :].contiguous()
    r = lambda shape: torch.randn(shape, dtype=torch.float64, device=dev)  # noqa: E731
    e = torch.empty_like
    rho, rho_pl = F["rho"], F["rho_pl"]
    flx, flx_pl, xc, xc_pl = F["flx_h"], F["flx_h_pl"], F["GRD_xc"], F["GRD_xc_pl"]
    ctx.horizontal_flux(flx, flx_pl, xc, xc_pl, *[F[k] for k in drivers.FLUX_IN], 360.0)
    q, q_pl = H["q"], H["q_pl"]
    ch, ch_pl, cm, cm_pl = z((6,) + spec.shape3), z(spec.shape_pl), z((6,) + spec.shape3), z(spec.shape_pl)
    ctx.tracer_courant(ch, ch_pl, cm, cm_pl, flx, flx_pl, rho, rho_pl)
    grad, grad_pl = z((3,) + spec.shape3), z((3,) + spec.shape_pl)
    ctx.oprt_gradient(grad, grad_pl, q, q_pl, M["coef_grad"], M["coef_grad_pl"])
    qa, qa_pl = z((6,) + spec.shape3), z(spec.shape_pl)
    ctx.horizontal_remap(qa, qa_pl, q, q_pl, grad, grad_pl, cm, cm_pl, xc, xc_pl, X, X_pl)
    fv, fv_pl = 1.0e-3 * r(spec.shape3), 1.0e-3 * r(spec.shape_pl)
    # cotangents and outputs
    gb, gb_pl = r((3,) + spec.shape3), r((3,) + spec.shape_pl)
    sb, sb_pl = e(q), e(q_pl)
    chb, chb_pl = r((6,) + spec.shape3), r(spec.shape_pl)
    fb, fb_pl, rb, rb_pl = e(flx), e(flx_pl), e(rho), e(rho_pl)
    qab, qab_pl = r((6,) + spec.shape3), r(spec.shape_pl)
    gb2, gb2_pl, xb, xb_pl = e(grad), e(grad_pl), e(xc), e(xc_pl)
    rqb, rqb_pl = r(spec.shape3), r(spec.shape_pl)
    qab2, qab2_pl = e(qa), e(qa_pl)
    h1, h1_pl, h3, h3_pl = r(spec.shape3), r(spec.shape_pl), r((3,) + spec.shape3), r((3,) + spec.shape_pl)
    ckb, ckb_pl = r((2,) + spec.shape3), r((2,) + spec.shape_pl)
    vfb, vfb_pl, vrb, vrb_pl = e(fv), e(fv_pl), e(fv), e(fv_pl)
    qhb, qhb_pl, vqb, vqb_pl = r(spec.shape3), r(spec.shape_pl), e(fv), e(fv_pl)
    qh, qh_pl, vqhb, vqhb_pl = r(spec.shape3), r(spec.shape_pl), e(fv), e(fv_pl)
    R, P = ctx.nreg, ctx.npl
    hm, hm_pl = torch.zeros(R, dtype=torch.uint8, device=dev), torch.zeros(P, dtype=torch.uint8, device=dev)
    ctx.horizontal_limiter_thuburn_mask(H["q_a"], H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                        H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_prev"], H["Qout_prev_pl"],
                                        H["Qout_post"], H["Qout_post_pl"], hm, hm_pl)
    lqab, lqab_pl, lqpb, lqpb_pl = r((6,) + spec.shape3), r(spec.shape_pl), r((2,) + spec.shape3), r((2,) + spec.shape_pl)
    lqb, lqb_pl, lpob, lpob_pl = z(spec.shape3), z(spec.shape_pl), z((2,) + spec.shape3), z((2,) + spec.shape_pl)
    torch.cuda.synchronize()
    for s in range(steps):
        ctx.oprt_gradient_adjoint(sb, sb_pl, gb, gb_pl, M["coef_grad"], M["coef_grad_pl"])
        ctx.tracer_courant_adjoint(fb, fb_pl, rb, rb_pl, chb, chb_pl, flx, flx_pl, rho, rho_pl)
        ctx.horizontal_remap_adjoint(sb, sb_pl, gb2, gb2_pl, xb, xb_pl, qab, qab_pl, grad, grad_pl, cm, cm_pl, xc,
                                     xc_pl, X, X_pl)
        ctx.tracer_update_adjoint(fb, fb_pl, qab2, qab2_pl, rqb, rqb_pl, flx, flx_pl, qa, qa_pl)
        ctx.comm_data_transfer_adjoint(h1, h1_pl, vmax=1, kmax=spec.kall)
        ctx.comm_data_transfer_adjoint(h3, h3_pl, vmax=3, kmax=spec.kall)
        ctx.vertical_courant_adjoint(vfb, vfb_pl, vrb, vrb_pl, ckb, ckb_pl, fv, fv_pl, rho, rho_pl)
        ctx.vertical_face_adjoint(vqb, vqb_pl, qhb, qhb_pl)
        ctx.tracer_update_vertical_adjoint(vfb, vfb_pl, vqhb, vqhb_pl, rqb, rqb_pl, fv, fv_pl, qh, qh_pl)
        # the h-limiter adjoint as two half passes: faces (q_a_b, Qout_post_b) then rows (Qout_prev_b)
        ctx.horizontal_limiter_thuburn_adjoint(lqab, lqab_pl, lqb, lqb_pl, None, None, lpob, lpob_pl, H["q_a"],
                                               H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                               H["ch_pl"], H["cmask"], H["cmask_pl"], H["Qout_post"],
                                               H["Qout_post_pl"], hm, hm_pl)
        ctx.horizontal_limiter_thuburn_adjoint(None, None, lqb, lqb_pl, lqpb, lqpb_pl, None, None, H["q_a"],
                                               H["q_a_pl"], H["q"], H["q_pl"], H["d"], H["d_pl"], H["ch"],
                                               H["ch_pl"], H["cmask"], H["cmask_pl"], None, None, hm, hm_pl)
        torch.cuda.synchronize()
        print(f"[profile_step] glue adjoint {s + 1}/{steps} done", flush=True)
    ctx.close()
    return 0


def coverage(steps) -> int:
    """Per step: flux and h-limiter (FP64 and FP32, the limiter for one and
    for two tracers) on regions 258, 322 and 514 wide, the rho*w adjoint at
    kall 42 (the register column), the vertical limiter for four tracers."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, synth

    dev = torch.device("cuda:0")
    T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    grids = []
    for n, kall, lall in ((258, 16, 8), (322, 16, 4), (514, 16, 1)):
        spec = synth.GridSpec(n, kall, lall, seed=n)
        case = synth.make_case(spec)
        ctx = drivers.setup_context(spec, case["geo"])
        ctx.set_stream(torch.cuda.current_stream(dev))
        f = {k: T(case["flux"][k]) for k in drivers.FLUX_IN}
        h = {k: T(case["hlim"][k]) for k in drivers.HLIM_IN + ("q_a", "q_a_pl")}
        grids.append((spec, ctx, f, h))
    specv = synth.GridSpec(66, 42, 10, seed=42)
    casev = synth.make_case(specv)
    ctxv = drivers.setup_context(specv, casev["geo"], casev["vi"])
    ctxv.set_stream(torch.cuda.current_stream(dev))
    r = lambda shape: torch.randn(shape, dtype=torch.float64, device=dev)  # noqa: E731
    wb, wb_pl = r(specv.shape3), r(specv.shape_pl)
    bufs = [torch.zeros(specv.shape3 if i % 2 == 0 else specv.shape_pl, dtype=torch.float64, device=dev)
            for i in range(12)]
    ntr = 4
    v = casev["vlim"]
    tr = lambda a: T(np.stack([a] * ntr, axis=0))  # Fortran's tracer-last: the outermost C index  # noqa: E731
    vq_h, vq_h_pl, vq, vq_pl = tr(v["q_h"]), tr(v["q_h_pl"]), tr(v["q"]), tr(v["q_pl"])
    vd, vd_pl, vck, vck_pl = T(v["d"]), T(v["d_pl"]), T(v["ck"]), T(v["ck_pl"])
    torch.cuda.synchronize()
    for s in range(steps):
        for spec, ctx, f, h in grids:
            z = lambda shape, dt=torch.float64: torch.zeros(shape, dtype=dt, device=dev)  # noqa: E731
            for dt in (torch.float64, torch.float32):
                x = {k: t.to(dt) for k, t in f.items()}
                flx, flx_pl = z((6,) + spec.shape3, dt), z(spec.shape_pl, dt)
                xc, xc_pl = z((3, 3) + spec.shape3, dt), z((3,) + spec.shape_pl, dt)
                run = ctx.horizontal_flux if dt == torch.float64 else ctx.horizontal_flux_f32
                run(flx, flx_pl, xc, xc_pl, *[x[k] for k in drivers.FLUX_IN], 360.0)
                y = {k: t.to(dt) for k, t in h.items()}
                qp, qp_pl = z((2,) + spec.shape3, dt), z((2,) + spec.shape_pl, dt)
                run = ctx.horizontal_limiter_thuburn if dt == torch.float64 else ctx.horizontal_limiter_thuburn_f32
                run(y["q_a"], y["q_a_pl"], y["q"], y["q_pl"], y["d"], y["d_pl"], y["ch"], y["ch_pl"], y["cmask"],
                    y["cmask_pl"], qp, qp_pl, y["Qout_post"], y["Qout_post_pl"])
            # two tracers (Fortran's tracer-last layout: the outermost C index)
            two = lambda a: torch.stack([a, a], dim=0).contiguous()  # noqa: E731
            qp2, qp2_pl = z((2, 2) + spec.shape3), z((2, 2) + spec.shape_pl)
            ctx.horizontal_limiter_thuburn_batch(2, two(h["q_a"]), two(h["q_a_pl"]), two(h["q"]), two(h["q_pl"]),
                                                 h["d"], h["d_pl"], h["ch"], h["ch_pl"], h["cmask"], h["cmask_pl"],
                                                 qp2, qp2_pl, two(h["Qout_post"]), two(h["Qout_post_pl"]))
        ctxv.vi_rhow_solver_adjoint(wb, wb_pl, *bufs, 60.0)
        ctxv.vertical_limiter_thuburn_batch(ntr, vq_h, vq_h_pl, vq, vq_pl, vd, vd_pl, vck, vck_pl)
        torch.cuda.synchronize()
        print(f"[profile_step] coverage {s + 1}/{steps} done", flush=True)
    for _, ctx, _, _ in grids:
        ctx.close()
    ctxv.close()
    return 0


def chained(ctx, spec, case, steps, z) -> int:
    """The chained FP64 and mixed steps with their mass sums, then the limiter
    with the library's Qout exchange and one halo transfer, per step."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, synth

    lv = synth.icosahedral_levels(spec.gall_1d, spec.lall)
    if not lv:
        print("[profile_step] --mode chained needs an icosahedral grid size", file=sys.stderr)
        return 2
    ctx.comm_setup(lv[1])
    ctx.set_const_f32(float(np.finfo(np.float64).eps), float(np.finfo(np.float32).max))
    S64 = drivers.chain_state(ctx, spec, case, mixed=False)
    S32 = drivers.chain_state(ctx, spec, case, mixed=True)
    m = torch.zeros(spec.kall, dtype=torch.float64, device=S64["rho"].device)
    var, var_pl = z(spec.shape3), z(spec.shape_pl)
    for s in range(steps):
        for S in (S64, S32):
            drivers.chain_step(ctx, S, 360.0, 60.0)
            ctx.tracer_mass(S["rhogq"], S["rhogq_pl"], m)
        for S, single in ((S64, False), (S32, True)):
            sfx = "_f32" if single else ""
            getattr(ctx, "horizontal_limiter_thuburn" + sfx)(
                S["h_q_a"], S["h_q_a_pl"], S["h_q"], S["h_q_pl"], S["h_d"], S["h_d_pl"], S["h_ch"], S["h_ch_pl"],
                S["h_cmask"], S["h_cmask_pl"], S["Qout_prev"], S["Qout_prev_pl"], None, None)
        ctx.comm_data_transfer(var, var_pl, 1)
        ctx.tracer_mass(S32["h_q"], S32["h_q_pl"], m, single=True)
        torch.cuda.synchronize()
        print(f"[profile_step] chained {s + 1}/{steps} done", flush=True)
    ctx.close()
    return 0


def vi_variants(ctx, spec, steps, W) -> int:
    """Each rho*w solve variant, FP64 and FP32 inputs; putcoef's
    factorisation; the adjoint forming its factors and reading them."""
    import torch

    from icoatmosbenchmark_v1_amd import drivers
    from icoatmosbenchmark_v1_amd.runtime import TUNE_VI_VARIANT

    W32 = {k: (v.float() if k not in ("rhogw", "rhogw_pl") else v) for k, v in W.items()}
    a64, a32 = [], []
    for k in drivers.VI_IN:
        a64 += [W[k], W[k + "_pl"]]
        a32 += [W32[k], W32[k + "_pl"]]
    bufs = [torch.zeros_like(W["rhogw"] if i % 2 == 0 else W["rhogw_pl"]) for i in range(12)]
    wb, wb_pl = torch.randn_like(W["rhogw"]), torch.randn_like(W["rhogw_pl"])
    Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl = (torch.empty_like(W["rhogw"] if i % 2 == 0 else W["rhogw_pl"])
                                       for i in range(6))
    ctx.vi_rhow_getcoef(Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl)
    for s in range(steps):
        ctx.vi_rhow_solver_putcoef(Mc, Mc_pl, Ml, Ml_pl, Mu, Mu_pl)
        for v in (1, 2, 3, 4):
            ctx.set_tuning(TUNE_VI_VARIANT, v)
            ctx.vi_rhow_solver(W["rhogw"], W["rhogw_pl"], *a64, 60.0)
            ctx.vi_rhow_solver_mixed(W["rhogw"], W["rhogw_pl"], *a32, 60.0)
        for v in (1, 2):
            ctx.set_tuning(TUNE_VI_VARIANT, v)
            ctx.vi_rhow_solver_adjoint(wb, wb_pl, *bufs, 60.0)
        torch.cuda.synchronize()
        print(f"[profile_step] vi variants {s + 1}/{steps} done", flush=True)
    ctx.close()
    return 0


def batched(steps) -> int:
    """The grouped step of 32 tracers on the gl5-gl8 rl1 grids (98, 78, 58, 42
    levels), eager launches on a side stream."""
    import numpy as np
    import torch

    from icoatmosbenchmark_v1_amd import drivers, group, synth

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)
    grids = ((5, 98), (6, 78), (7, 58), (8, 42))
    ctxs, states = [], []
    for gl, kall in grids:
        geo, _ = synth.icosahedral_geometry(gl, 1)
        spec = synth.GridSpec(2 ** (gl - 1) + 2, kall, 40, seed=gl)
        case = synth.make_case(spec, geo=geo)
        ctx = drivers.setup_context(spec, geo, case["vi"])
        ctx.set_stream(stream)
        ctx.set_const_f32(float(np.finfo(np.float64).eps), float(np.finfo(np.float32).max))
        ctxs.append(ctx)
        states.append(group.batched_state(spec, case, 32, dev))
    G = group.Group(ctxs)
    G.set_stream(stream)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for s in range(steps):
            G.tracer_step_f32(32, states, 360.0, 60.0, use_graph=False)
            stream.synchronize()
            print(f"[profile_step] batched {s + 1}/{steps} done", flush=True)
    G.close()
    for c in ctxs:
        c.close()
    return 0


def extras(ctx, spec, case, coef, steps, up, z) -> int:
    """The vertical glue (mass flux, Courant numbers, face values, update),
    the Mc/Ml/Mu update, OPRT_diffusion, OPRT3D_divdamp and the horizontal
    glue (courant, gradient, remap, update), each launched once per step."""
    import torch

    from icoatmosbenchmark_v1_amd import drivers, synth
    from icoatmosbenchmark_v1_amd.drivers import DDMP_IN

    if coef is None or "coef_grad" not in case["geo"]:
        print("[profile_step] --extras needs an icosahedral grid size", file=sys.stderr)
        return 2
    geo = case["geo"]
    D = up(synth.diffusion_inputs(spec, geo, coef), drivers.DIFF_IN)
    D["dscl"], D["dscl_pl"] = z(spec.shape3), z(spec.shape_pl)
    f = synth.divdamp_inputs(spec, geo, coef)
    drivers.setup_divdamp(ctx, f)
    G = up(f, drivers.DDMP_IN)
    for k in drivers.DDMP_OUT:
        G[k] = z(spec.shape_pl if k.endswith("_pl") else spec.shape3)
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
    fv, fv_pl = z(spec.shape3), z(spec.shape_pl)
    ck, ck_pl = z((2,) + spec.shape3), z((2,) + spec.shape_pl)
    qh, qh_pl = z(spec.shape3), z(spec.shape_pl)
    eth, eth_pl = torch.full_like(fv, 3.0e5), torch.full_like(fv_pl, 3.0e5)
    gt, gt_pl = torch.full_like(fv, 9.80616), torch.full_like(fv_pl, 9.80616)
    torch.cuda.synchronize()
    for s in range(steps):
        ctx.vertical_mass_flux(fv, fv_pl, *[G[k] for k in DDMP_IN[:8]], 360.0)
        ctx.vertical_courant(ck, ck_pl, fv, fv_pl, F["rho"], F["rho_pl"])
        ctx.vertical_face(qh, qh_pl, Q["q"], Q["q_pl"])
        ctx.tracer_update_vertical(rq, rq_pl, fv, fv_pl, qh, qh_pl)
        ctx.vi_rhow_update_matrix(eth, eth_pl, gt, gt_pl, 60.0)
        ctx.oprt_diffusion(D["dscl"], D["dscl_pl"], *[D[k] for k in drivers.DIFF_IN])
        ctx.oprt3d_divdamp(*[G[k] for k in drivers.DDMP_OUT], *[G[k] for k in drivers.DDMP_IN])
        ctx.tracer_courant(ch, ch_pl, cm, cm_pl, flx, flx_pl, F["rho"], F["rho_pl"])
        ctx.oprt_gradient(grad, grad_pl, Q["q"], Q["q_pl"], M["coef_grad"], M["coef_grad_pl"])
        ctx.horizontal_remap(qa, qa_pl, Q["q"], Q["q_pl"], grad, grad_pl, cm, cm_pl, xc, xc_pl, M["GRD_x"],
                             M["GRD_x_pl"])
        ctx.tracer_update(rq, rq_pl, flx, flx_pl, qa, qa_pl)
        torch.cuda.synchronize()
        print(f"[profile_step] extras {s + 1}/{steps} done", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
