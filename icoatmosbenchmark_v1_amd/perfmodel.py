"""Algorithmic HBM traffic of the four kernels (bytes per launch).

"Algorithmic" counts every array element the kernel must read or write at
least once -- exactly the elements the reference loops touch -- and nothing
else: re-reads from cache, scratch traffic and unused planes of an argument
(cmask faces 4..6, q_a faces 4..6 on input) are excluded.  The counts follow
the loop bounds of the reference (paths relative to kernels/NICAM/):

horizontal_flux (dyn_horiz_adv_flux/src/mod_src_tracer.f90:137-320)
    reads  rho, rhovx, rhovy, rhovz at every cell of a slab      4 x gall
    writes flx_h (6 faces) and GRD_xc (3 arcs x 3) at every cell 15 x gall
    reads  25 K-invariant metrics per cell once per region
           (GMTR_t W1..W3 x TI/TJ, GMTR_a HN x 3 arcs, GMTR_p RAREA, GRD_xr)
horizontal_limiter_thuburn (dyn_horiz_adv_limiter/src/mod_src_tracer.f90:139-410)
    reads  q (all cells), Qout_post (2, all cells), d and ch (6) on the
           interior, cmask and q_a faces 1..3 on the arc origins
           i,j in [gmin-1, gmax]
    writes Qout_prev (2, all cells), q_a faces 1..3 on the arc origins and
           faces 4..6 on their neighbours
vertical_limiter_thuburn (dyn_vert_adv_limiter/src/mod_src_tracer.f90:99-186)
    per column: q (kall), d (kall-2), ck lower (kall-1), ck upper (kall-2),
    q_h read and written on kmin+1..kmax
vi_rhow_solver (dyn_vi_rhow_solver/src/mod_vi.f90:192-246)
    per column: rhogw0, Sw, RGAMH, Mc, Ml, Mu on kmin+1..kmax; preg0, Spre,
    RGSGAM2, rhog0, Srho, RGAM on kmin..kmax; RGSGAM2H and rhogw at kmin,
    kmax+1; GSGAM2H and the written rhogw on kmin..kmax+1.
Pole regions are included; they are below 0.01 % of the total.
"""
from __future__ import annotations

from typing import Dict

from .synth import GridSpec

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def kernel_bytes(spec: GridSpec, word: int = 8) -> Dict[str, int]:
    n, kall, lall, npl = spec.gall_1d, spec.kall, spec.lall, spec.gall_pl
    gall = n * n
    inner = (n - 2) ** 2
    arcs = (n - 1) ** 2  # arc origins i,j in [gmin-1, gmax]
    ncol = gall * lall
    pl = 2 if spec.have_pl else 0
    ring = npl - 1

    flux = (19 * gall * kall * lall + 25 * gall * lall) * word
    flux += pl * (4 * npl * kall + 4 * ring * kall) * word

    hlim_r = (gall * (1 + 2) + inner * (1 + 6) + arcs * (3 + 3)) * kall * lall
    hlim_w = (gall * 2 + arcs * 6) * kall * lall
    hlim = (hlim_r + hlim_w) * word
    hlim += pl * (4 * npl * kall + 2 * 2 * npl * kall + ring * kall * 2) * word

    vcol = kall + (kall - 2) + (kall - 1) + (kall - 2) + 2 * (kall - 3)
    vlim = vcol * (ncol + pl * npl) * word

    nk3, nk2, nk1 = kall - 3, kall - 2, kall - 1
    vicol = 6 * nk3 + 6 * nk2 + 2 * 2 + 2 * nk1
    vi = vicol * (ncol + pl * npl) * word
    return dict(horizontal_flux=flux, horizontal_limiter_thuburn=hlim, vertical_limiter_thuburn=vlim,
                vi_rhow_solver=vi)


def vi_mixed_bytes(spec: GridSpec) -> int:
    """Algorithmic bytes of nicam_vi_rhow_solver_mixed: the per-column streams
    of kernel_bytes' VI with the state, the sources and RGAM, RGAMH, RGSGAM2
    at 4 bytes; Mc/Ml/Mu, RGSGAM2H, GSGAM2H and rhogw at 8."""
    kall, ncol = spec.kall, spec.gall * spec.lall + (2 if spec.have_pl else 0) * spec.gall_pl
    nk3, nk2, nk1 = kall - 3, kall - 2, kall - 1
    # levels kmin+1..kmax: rhogw0, Sw, RGAMH (4 B) and Mc, Ml, Mu (8 B); kmin..kmax: preg0, Spre, rhog0, Srho,
    # RGSGAM2, RGAM (4 B); rhogw and RGSGAM2H at kmin and kmax+1; GSGAM2H and the rhogw result on kall-1 levels
    return (nk3 * (3 * 4 + 3 * 8) + nk2 * 6 * 4 + 2 * 2 * 8 + nk1 * 2 * 8) * ncol


def adjoint_bytes(spec: GridSpec, word: int = 8) -> Dict[str, int]:
    """Algorithmic bytes of the adjoint kernels: each array the kernel must
    read or write counted once over the levels and points it touches (the
    masks at their stored width; clearing q_b and Qout_post_b before the
    atomic sums and the solver's reuse of its outputs as work space are not
    counted)."""
    n, kall, lall, npl = spec.gall_1d, spec.kall, spec.lall, spec.gall_pl
    gall = n * n
    inner = (n - 2) ** 2
    arcs = (n - 1) ** 2
    ncol = gall * lall
    pl = 2 if spec.have_pl else 0
    ring = npl - 1
    # h-limiter: reads q, Qout_post (2), Qout_prev_b (2) on every cell; d and ch (6) on inner cells; q_a_b (6),
    # q_a (3) and cmask (3) on the arcs; writes q_b, Qout_post_b (2) and q_a_b (6)
    hr = (gall * (1 + 2 + 2) + inner * (1 + 6) + arcs * (6 + 3 + 3)) * word + gall
    hw = (gall * (1 + 2) + arcs * 6) * word
    hlim = (hr + hw) * kall * lall
    hlim += pl * kall * ((npl * (1 + 2 + 2 + 1 + 2) + ring * 3) * word + npl)
    # v-limiter per column: q, d, ck (2 planes), q_h, q_h_b, mask; writes q_h_b, q_b
    nf = kall - 3
    vcol = (kall + (kall - 2) + (kall - 1) + (kall - 2) + 2 * nf) * word + 4 * ((nf + 31) // 32)
    vcol += (nf + kall) * word
    vlim = vcol * (ncol + pl * npl)
    # VI per column: rhogw_b, GSGAM2H (kall-1), Mc, Ml, Mu (kall-3 each), RGSGAM2H (2), RGSGAM2, RGAM,
    # RGAMH (kall); writes the seven cotangents
    vicol = (2 * (kall - 1) + 3 * (kall - 3) + 2 + 3 * kall + 6 * kall + (kall - 1)) * word
    vi = vicol * (ncol + pl * npl)
    # flux: reads rho, rhovx/y/z on every cell, flx_h_b (6) and GRD_xc_b (9) on the arcs and faces; writes the
    # four cotangents; metrics GMTR_t (6), GMTR_a (9), GMTR_p (1) once per column
    flux = ((4 + 4) * gall + (6 + 9) * arcs) * kall * lall * word + 16 * gall * lall * word
    flux += pl * kall * (npl * (4 + 4) + ring * (1 + 3)) * word
    return dict(horizontal_flux=flux, horizontal_limiter_thuburn=hlim, vertical_limiter_thuburn=vlim,
                vi_rhow_solver=vi)


def kernel_flops(spec: GridSpec) -> Dict[str, int]:
    """Floating-point operations per launch (reported, not bounding): flux
    ~125, h-limiter ~170 (min/max included), v-limiter ~45, VI ~45 per
    array cell-level."""
    cl = spec.array_cell_levels()
    return dict(horizontal_flux=125 * cl, horizontal_limiter_thuburn=170 * cl, vertical_limiter_thuburn=45 * cl,
                vi_rhow_solver=45 * cl)


def operator_bytes(spec: GridSpec, word: int = 8) -> Dict[str, int]:
    """Algorithmic bytes of the neighbouring operators and the halo fill.

    OPRT_diffusion (dyn_diffusion/src/mod_oprt.f90:97-216): scl and kh read
    and dscl written on every cell of every level; coef_intp (18) and
    coef_diff (18) per cell, K-invariant, once per region.
    OPRT3D_divdamp (dyn_divdamp/src/mod_oprt3d.f90:143-310): per column the
    momenta and RGAM on kmin..kmax, rhogw, RGSQRTH, RGAMH and C2WfactGz (6)
    on kmin+1..kmax, the three outputs on all levels; the same 36
    coefficients per cell.
    halo fill (nicam_comm_data_transfer, one variable): every map entry
    reads one value and writes one per level.
    step glue, per cell-level of the array: courant reads flx_h (6) and rhog
    and writes ch (6) and cmask (6); gradient reads q and writes grad (3),
    plus 21 coef_grad per cell once; remap reads q, grad (3), cmask (6) and
    GRD_xc (9) and writes q_a (6), plus GRD_x (3) per cell once; update reads
    rhogq, flx_h (6) and q_a (6) and writes rhogq."""
    n, kall, lall, npl = spec.gall_1d, spec.kall, spec.lall, spec.gall_pl
    gall = n * n
    pl = 2 if spec.have_pl else 0
    diff = (3 * gall * kall + 36 * gall) * lall * word + pl * (3 * npl * kall + 9 * npl + 15) * word
    dcol = 4 * (kall - 2) + 9 * (kall - 3) + 3 * kall
    ddmp = (dcol * gall + 36 * gall) * lall * word + pl * (dcol * npl + 9 * npl + 15) * word
    cl = gall * kall * lall + pl * npl * kall
    glue = dict(tracer_courant=19 * cl * word, oprt_gradient=(4 * cl + 21 * gall * lall) * word,
                horizontal_remap=(25 * cl + 3 * gall * lall) * word, tracer_update=14 * cl * word)
    return dict(oprt_diffusion=diff, oprt3d_divdamp=ddmp, **glue)


def batched_bytes(spec: GridSpec, ntr: int, word: int = 4) -> Dict[str, int]:
    """Algorithmic bytes of the batched step of one grid (ntr tracers): the
    flux and the mixed rho*w solve once; the limiters' shared fields (d, ch
    on the interior, cmask on the arc origins; d and the two ck planes) once,
    and their per-tracer streams (kernel_bytes' other terms) ntr times."""
    n, kall, lall, npl = spec.gall_1d, spec.kall, spec.lall, spec.gall_pl
    gall, inner, arcs = n * n, (n - 2) ** 2, (n - 1) ** 2
    ncol = gall * lall
    pl = 2 if spec.have_pl else 0
    ring = npl - 1
    kb = kernel_bytes(spec, word)
    h_shared = (inner * (1 + 6) + arcs * 3) * kall * lall * word + pl * (2 * npl * kall + ring * kall) * word
    h_tracer = kb["horizontal_limiter_thuburn"] - h_shared
    v_shared = ((kall - 2) + (kall - 1) + (kall - 2)) * (ncol + pl * npl) * word
    v_tracer = kb["vertical_limiter_thuburn"] - v_shared
    return dict(horizontal_flux=kb["horizontal_flux"], vi_rhow_solver=vi_mixed_bytes(spec),
                horizontal_limiter_thuburn=h_shared + ntr * h_tracer,
                vertical_limiter_thuburn=v_shared + ntr * v_tracer)


def chain_bytes(spec: GridSpec, mixed: bool) -> Dict[str, int]:
    """Algorithmic bytes of the chained step's kernels beyond the four of
    kernel_bytes: the conservative horizontal update (reads rhogq, flx_h and
    q_a faces 1..3, writes rhogq, plus AREA and RAREA per column), the
    vertical update (reads rhogq, flx_v, q_h, writes rhogq) and the FP64 mass
    sum (reads rhogq and AREA).  rhogq is FP64 in both precisions."""
    w = 4 if mixed else 8
    cl = spec.array_cell_levels() + (2 * spec.gall_pl * spec.kall if spec.have_pl else 0)
    ncol = spec.gall * spec.lall
    return dict(tracer_update_conservative=cl * (16 + 6 * w) + 2 * ncol * 8,
                tracer_update_vertical=cl * (16 + 2 * w), tracer_mass=cl * 8 + ncol * 8)
