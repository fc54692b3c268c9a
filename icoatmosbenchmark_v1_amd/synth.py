"""Seeded synthetic inputs for the four hot-path kernels.

Arrays use the reference's Fortran layout with the dimensions reversed for
C/numpy: Fortran ``x(gall,kall,lall,n)`` is ``x[n, l, k, g]`` here.  All
values come from ``numpy.random.default_rng(seed)`` and closed-form fields, so
the same (sizes, seed) give the same bits on any machine with this numpy.

The reference ships no inputs (its snapshot files are downloaded), and
nothing in it computes the limiter inputs (Courant numbers, upwind masks,
face values) or the tridiagonal coefficients, so the fields below are this
package's own construction; they are shaped like the reference's data
(kernels/NICAM/*/reference/*.log ranges): densities around 1, tracers in
[0, 1] with exact zeros, Courant numbers below 0.5 in magnitude, diagonally
dominant tridiagonal systems.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

RADIUS = 6.37122e6  # GRD_rscale (problem_size.inc:69)


@dataclass(frozen=True)
class GridSpec:
    gall_1d: int
    kall: int
    lall: int
    gall_pl: int = 6
    have_pl: bool = True
    seed: int = 0

    @property
    def gall(self) -> int:
        return self.gall_1d * self.gall_1d

    @property
    def shape3(self):
        return (self.lall, self.kall, self.gall)

    @property
    def shape_pl(self):
        return (2, self.kall, self.gall_pl)

    def inner_cell_levels(self) -> int:
        """lall*(gall_1d-2)^2*(kall-2) + 2*(kall-2): the interior cells of
        all regions plus the two pole points, over the physical levels."""
        return self.lall * (self.gall_1d - 2) ** 2 * (self.kall - 2) + (2 * (self.kall - 2) if self.have_pl else 0)

    def array_cell_levels(self) -> int:
        return self.gall * self.kall * self.lall


def stretched_vgrid(kall: int, dz_bottom: float = 100.0, dz_top: float = 2000.0):
    """gz, gzh (kall) for layers whose thickness grows geometrically from
    dz_bottom to dz_top; ghost levels follow the layout of the shipped
    vgrid40_600m_24km.dat (gzh(kmin)=0 at the surface, one ghost level
    below and above)."""
    nl = kall - 2
    if nl == 1:
        dz = np.array([dz_bottom])
    else:
        dz = dz_bottom * (dz_top / dz_bottom) ** (np.arange(nl) / (nl - 1))
    gzh = np.empty(kall)
    gz = np.empty(kall)
    gzh[1] = 0.0
    for k in range(1, kall - 1):
        gzh[k + 1] = gzh[k] + dz[k - 1]
    for k in range(1, kall - 1):
        gz[k] = 0.5 * (gzh[k] + gzh[k + 1])
    gzh[0] = -dz[0]
    gz[0] = -0.5 * dz[0]
    gz[kall - 1] = gzh[kall - 1] + 0.5 * dz[-1]
    return gz, gzh


def uniform_vgrid(kall: int, dz: float = 600.0):
    """The layout of the reference's vgrid40_600m_24km.dat for kall levels."""
    k = np.arange(kall, dtype=np.float64)
    return (k - 0.5) * dz, (k - 1.0) * dz


# ----------------------------------------------------------------------------
# geometry
# ----------------------------------------------------------------------------
def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def make_geometry(spec: GridSpec, have_sgp=None) -> Dict[str, np.ndarray]:
    """Per-region patches of a sphere with plausible NICAM metrics.

    Region l is a gall_1d x gall_1d patch of points (halo included) around a
    seeded centre; GRD_xr holds the midpoints of the three arcs of every
    point, GMTR_a the arc normals scaled by the arc length, GMTR_t positive
    triangle weights summing to one, GMTR_p the cell areas and their
    inverses.  The two pole regions are 5-point rings around the poles.
    """
    rng = np.random.default_rng(spec.seed * 7919 + 11)
    n, lall, gall = spec.gall_1d, spec.lall, spec.gall
    dx = 2.0 * np.pi * RADIUS / (5.0 * max(n - 2, 1) * 2)  # spacing of a gl-(log2(n-2)+1) grid
    centres = _unit(rng.normal(size=(lall, 3)))
    tmp = _unit(rng.normal(size=(lall, 3)))
    e1 = _unit(np.cross(centres, tmp))
    e2 = np.cross(centres, e1)
    ii, jj = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="xy")
    ii = ii.ravel() - (n - 1) / 2.0
    jj = jj.ravel() - (n - 1) / 2.0

    def pos(di, dj):
        u = (ii + di + 0.5 * (jj + dj))[None, :, None] * dx
        v = (jj + dj)[None, :, None] * (dx * np.sqrt(3.0) / 2.0)
        p = centres[:, None, :] * RADIUS + u * e1[:, None, :] + v * e2[:, None, :]
        return _unit(p) * RADIUS  # (lall, gall, 3)

    p00 = pos(0, 0)
    ends = {0: pos(1, 0), 1: pos(1, 1), 2: pos(0, 1)}  # AI, AIJ, AJ far ends
    GRD_xr = np.empty((3, 3, lall, 1, gall))  # [d, a, l, 1, g]
    GMTR_a = np.zeros((12, 3, lall, 1, gall))  # [c, a, l, 1, g]
    for a in range(3):
        mid = 0.5 * (p00 + ends[a])
        GRD_xr[:, a, :, 0, :] = np.moveaxis(mid, -1, 0)
        edge = ends[a] - p00
        hn = np.cross(_unit(mid), edge) / np.sqrt(3.0)  # normal to the arc, |hn| ~ dual edge length
        hn *= 1.0 + 0.02 * rng.uniform(-1, 1, size=hn.shape[:2])[..., None]
        GMTR_a[0:3, a, :, 0, :] = np.moveaxis(hn, -1, 0)
        GMTR_a[3:6, a, :, 0, :] = np.moveaxis(edge, -1, 0) / np.sqrt(3.0)
    area = (np.sqrt(3.0) / 2.0) * dx * dx * (1.0 + 0.05 * rng.uniform(-1, 1, size=(lall, gall)))
    GMTR_p = np.zeros((8, lall, 1, gall))
    GMTR_p[0, :, 0, :] = area
    GMTR_p[1, :, 0, :] = 1.0 / area
    GMTR_t = np.zeros((5, 2, lall, 1, gall))  # [w, t, l, 1, g]
    for t in range(2):
        w = 1.0 / 3.0 + 0.05 * rng.uniform(-1, 1, size=(3, lall, gall))
        w /= w.sum(axis=0, keepdims=True)
        GMTR_t[2:5, t, :, 0, :] = w
        GMTR_t[0, t, :, 0, :] = area / 3.0
        GMTR_t[1, t, :, 0, :] = 3.0 / area

    npl = spec.gall_pl
    GRD_xr_pl = np.zeros((3, 2, 1, npl))
    GMTR_p_pl = np.zeros((8, 2, 1, npl))
    GMTR_t_pl = np.zeros((5, 2, 1, npl))
    GMTR_a_pl = np.zeros((18, 2, 1, npl))
    for l, zsign in enumerate((1.0, -1.0)):
        ang = 2.0 * np.pi * np.arange(npl - 1) / (npl - 1)
        th = dx / RADIUS
        ring = np.stack([np.sin(th) * np.cos(ang), np.sin(th) * np.sin(ang), zsign * np.cos(th) * np.ones_like(ang)], -1)
        pole = np.array([0.0, 0.0, zsign])
        for v in range(1, npl):
            nxt = ring[v % (npl - 1)]
            mid = _unit(pole + ring[v - 1] + nxt) * RADIUS
            GRD_xr_pl[:, l, 0, v] = mid
            hn = np.cross(_unit(mid), (nxt - ring[v - 1]) * RADIUS) / np.sqrt(3.0)
            GMTR_a_pl[0:3, l, 0, v] = hn
            w = 1.0 / 3.0 + 0.05 * rng.uniform(-1, 1, size=3)
            GMTR_t_pl[2:5, l, 0, v] = w / w.sum()
        a_pl = (npl - 1) * 0.5 * dx * dx * (1.0 + 0.05 * rng.uniform(-1, 1))
        GMTR_p_pl[0, l, 0, :] = a_pl
        GMTR_p_pl[1, l, 0, :] = 1.0 / a_pl
    if have_sgp is None:
        have_sgp = [1 if (l % 4 == 0) else 0 for l in range(lall)]
    return dict(GRD_xr=GRD_xr, GRD_xr_pl=GRD_xr_pl, GMTR_p=GMTR_p, GMTR_p_pl=GMTR_p_pl, GMTR_t=GMTR_t,
                GMTR_t_pl=GMTR_t_pl, GMTR_a=GMTR_a, GMTR_a_pl=GMTR_a_pl,
                have_sgp=np.asarray(have_sgp, dtype=np.int32), points=p00)


def icosahedral_levels(gall_1d: int, lall: int):
    """(glevel, rlevel) when gall_1d = 2^(glevel-rlevel)+2 and lall = 10*4^rlevel
    (all regions of an icosahedral grid on one process), else None."""
    r = 0
    while 10 * 4 ** r < lall:
        r += 1
    n = gall_1d - 2
    if 10 * 4 ** r != lall or n < 2 or n & (n - 1):
        return None
    return n.bit_length() - 1 + r, r


def icosahedral_geometry(glevel: int, rlevel: int, radius: float = RADIUS):
    """The metrics of the real icosahedral grid, from the library's host
    routines: positions (nicam_grid_icosahedral), triangle vertices and arc
    midpoints (nicam_grid_vertices), GMTR_p/t/a (nicam_gmtr_setup, the
    reference's dyn_metrics formulas) and the OPRT_diffusion coefficients
    (nicam_oprt_diffusion_setup).  Returns (geo, coef): geo in the layout of
    make_geometry plus the positions GRD_x/GRD_x_pl and the gradient
    coefficients coef_grad/coef_grad_pl (nicam_oprt_gradient_setup) that the
    step glue reads, coef the coefficient dict of runtime.oprt_diffusion_setup.
    For gall_1d = 2^(glevel-rlevel)+2 and lall = 10*4^rlevel."""
    from . import runtime

    n1 = 2 ** (glevel - rlevel) + 2
    x, x_pl = runtime.grid_icosahedral(glevel, rlevel, radius)
    sgp = runtime.halo_map(rlevel, n1)["have_sgp"]
    v = runtime.grid_vertices(x, x_pl, sgp, radius)
    m = runtime.gmtr_setup(x, x_pl, v["xt"], v["xt_pl"], v["s"], v["s_pl"], sgp, radius)
    coef = runtime.oprt_diffusion_setup(m, sgp)
    cg, cg_pl = runtime.oprt_coef_setup("grad", m, sgp)
    points = np.ascontiguousarray(np.moveaxis(x[:, :, 0, :], 0, -1))  # (lall, gall, 3)
    # the unused halo corners hold no point: give them the first inner point
    empty = np.linalg.norm(points, axis=-1) == 0.0
    first = points[:, n1 + 1, :]
    points = np.where(empty[..., None], first[:, None, :], points)
    geo = dict(GRD_xr=v["xr"], GRD_xr_pl=v["xr_pl"], GMTR_p=m["p"], GMTR_p_pl=m["p_pl"], GMTR_t=m["t"],
               GMTR_t_pl=m["t_pl"], GMTR_a=m["a"], GMTR_a_pl=m["a_pl"], have_sgp=np.asarray(sgp, dtype=np.int32),
               points=points, GRD_x=x, GRD_x_pl=x_pl, coef_grad=cg, coef_grad_pl=cg_pl)
    return geo, coef


# ----------------------------------------------------------------------------
# fields
# ----------------------------------------------------------------------------
def _tracer(points: np.ndarray, rng) -> np.ndarray:
    """3 cosine bells + 1 slotted cylinder in [0, 1] on unit-sphere points (..., 3)."""
    x = _unit(points)
    q = np.zeros(x.shape[:-1])
    rb = 0.35
    for _ in range(3):
        c = _unit(rng.normal(size=3))
        dist = np.arccos(np.clip(x @ c, -1.0, 1.0))
        q = np.maximum(q, np.where(dist < rb, 0.5 * (1.0 + np.cos(np.pi * dist / rb)), 0.0))
    c = _unit(rng.normal(size=3))
    dist = np.arccos(np.clip(x @ c, -1.0, 1.0))
    e = _unit(np.cross(c, [0.0, 0.0, 1.0]))
    slot = np.abs(x @ e) < 0.05
    cyl = (dist < rb) & ~(slot & (x @ c > np.cos(rb) + 0.02))
    q = np.where(cyl, 1.0, q)
    return q


def flux_inputs(spec: GridSpec, geo: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """rho and rho*v at cell centres (solid-body rotation plus a random vortex)."""
    rng = np.random.default_rng(spec.seed * 7919 + 23)
    gz, _ = stretched_vgrid(spec.kall)
    x = geo["points"]  # (lall, gall, 3)
    omega = _unit(rng.normal(size=3)) * (20.0 / RADIUS)
    v = np.cross(omega, x)  # m/s
    vort_c = _unit(rng.normal(size=3))
    v += 8.0 * np.cross(vort_c, _unit(x)) * np.exp(-((_unit(x) @ vort_c - 1.0) ** 2) * 4.0)[..., None]
    prof = np.exp(-np.maximum(gz, 0.0) / 8000.0)
    pert = 1.0 + 0.1 * np.sin(3.0 * _unit(x)[..., 0] + 2.0 * _unit(x)[..., 1]) * np.cos(2.0 * _unit(x)[..., 2])
    rho = prof[None, :, None] * pert[:, None, :]  # (lall, kall, gall)
    rho = rho * (1.0 + 0.01 * rng.uniform(-1, 1, size=rho.shape))
    shear = (1.0 + 0.3 * np.tanh(gz / 5000.0))[None, :, None]
    out = dict(rho=rho)
    for c, nm in enumerate(("rhovx", "rhovy", "rhovz")):
        out[nm] = rho * v[:, None, :, c] * shear
    npl = spec.gall_pl
    for nm in ("rho", "rhovx", "rhovy", "rhovz"):
        base = rho.mean() if nm == "rho" else out[nm].std()
        out[nm + "_pl"] = base * (1.0 + 0.1 * rng.uniform(-1, 1, size=(2, spec.kall, npl)))
    return out


def hlim_inputs(spec: GridSpec, geo: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """Tracer, face values, horizontal Courant numbers and upwind masks, the
    divergence term d and a post-exchange allowable range Qout_post."""
    rng = np.random.default_rng(spec.seed * 7919 + 37)
    n, lall, kall, gall = spec.gall_1d, spec.lall, spec.kall, spec.gall
    q2 = _tracer(geo["points"], rng)  # (lall, gall)
    tilt = 0.5 + 0.5 * np.cos(np.linspace(0.0, np.pi, kall))
    q = np.clip(q2[:, None, :] * tilt[None, :, None], 0.0, 1.0)
    q = np.where(q < 1e-3, 0.0, q)
    ch = np.empty((6, lall, kall, gall))
    ch[:3] = rng.uniform(-0.3, 0.3, size=(3, lall, kall, gall))
    ch[3:] = rng.uniform(-0.3, 0.3, size=(3, lall, kall, gall))
    # mirrored faces carry the opposite Courant number of the shared arc
    idx = np.arange(gall).reshape(n, n)  # [j, i]
    for m, (dj, di) in enumerate(((0, 1), (1, 1), (1, 0))):
        src = idx[: n - dj, : n - di].ravel()
        dst = idx[dj:, di:].ravel()
        ratio = 1.0 + 0.05 * rng.uniform(-1, 1, size=(lall, kall, src.size))
        ch[m + 3][:, :, dst] = -ch[m][:, :, src] * ratio
    cmask = np.where(np.signbit(ch), 1.0, 0.0)  # 1 where the face is an inflow (ch < 0)
    # face values: the cell value pushed towards the neighbour, with overshoots
    q_a = np.empty((6, lall, kall, gall))
    for m in range(6):
        q_a[m] = q + rng.uniform(-0.35, 0.35, size=q.shape) * (np.abs(q) + 0.05)
    d = 2e-5 * rng.uniform(-1, 1, size=(lall, kall, gall))
    spread = rng.uniform(0.0, 0.2, size=(2, lall, kall, gall))
    Qout_post = np.stack([q - spread[0], q + spread[1]])
    npl = spec.gall_pl
    q_pl = rng.uniform(0.0, 1.0, size=(2, kall, npl))
    ch_pl = rng.uniform(-0.3, 0.3, size=(2, kall, npl))
    out = dict(q=q, q_a=q_a, ch=ch, cmask=cmask, d=d, Qout_post=Qout_post,
               q_pl=q_pl, q_a_pl=q_pl + rng.uniform(-0.3, 0.3, size=q_pl.shape),
               ch_pl=ch_pl, cmask_pl=np.where(np.signbit(ch_pl), 1.0, 0.0),
               d_pl=2e-5 * rng.uniform(-1, 1, size=q_pl.shape))
    sp = rng.uniform(0.0, 0.2, size=(2, 2, kall, npl))
    out["Qout_post_pl"] = np.stack([q_pl - sp[0], q_pl + sp[1]])
    return out


def vlim_inputs(spec: GridSpec, geo: Dict[str, np.ndarray], signed_zero_d: bool = True) -> Dict[str, np.ndarray]:
    """Tracer, half-level values, vertical Courant numbers (lower/upper face)."""
    rng = np.random.default_rng(spec.seed * 7919 + 41)
    lall, kall, gall = spec.lall, spec.kall, spec.gall
    q2 = _tracer(geo["points"], rng)
    prof = np.sin(np.linspace(0.2, 2.8, kall)) ** 2
    q = q2[:, None, :] * prof[None, :, None]
    q_h = np.empty_like(q)
    q_h[:, 1:, :] = 0.5 * (q[:, 1:, :] + q[:, :-1, :])
    q_h[:, 0, :] = q[:, 0, :]
    q_h += rng.uniform(-0.2, 0.2, size=q.shape) * (np.abs(q_h) + 0.02)
    ck1 = rng.uniform(-0.4, 0.4, size=(lall, kall, gall))
    zero = rng.uniform(size=ck1.shape) < 0.02
    ck1 = np.where(zero, np.where(rng.uniform(size=ck1.shape) < 0.5, 0.0, -0.0), ck1)
    ck2 = np.empty_like(ck1)
    ck2[:, :-1, :] = -ck1[:, 1:, :] * (1.0 + 0.03 * rng.uniform(-1, 1, size=ck1[:, 1:, :].shape))
    ck2[:, -1, :] = rng.uniform(-0.4, 0.4, size=ck1[:, -1, :].shape)
    d = np.full((lall, kall, gall), -0.0) if signed_zero_d else 1e-5 * rng.uniform(-1, 1, size=(lall, kall, gall))
    npl = spec.gall_pl
    q_pl = rng.uniform(0.0, 1.0, size=(2, kall, npl))
    ck_pl = rng.uniform(-0.4, 0.4, size=(2, 2, kall, npl))
    return dict(q=q, q_h=q_h, ck=np.stack([ck1, ck2]), d=d, q_pl=q_pl,
                q_h_pl=q_pl + rng.uniform(-0.2, 0.2, size=q_pl.shape), ck_pl=ck_pl,
                d_pl=np.full(q_pl.shape, -0.0) if signed_zero_d else 1e-5 * rng.uniform(-1, 1, size=q_pl.shape))


def vi_inputs(spec: GridSpec, dt: float = 60.0, vgrid=None) -> Dict[str, np.ndarray]:
    """Prognostic state, sources, vertical metrics and a diagonally dominant
    tridiagonal system for vi_rhow_solver.

    Mc/Ml/Mu follow the structure of the implicit rho*w equation
    (docs/NICAM/src/20_intro_nicam.tex:582-656): off-diagonals of size
    (c_s*dt/dz)^2 with c_s in [300, 350] m/s, and a diagonal of
    1 + |Ml| + |Mu| plus a margin, so every system is strictly diagonally
    dominant."""
    rng = np.random.default_rng(spec.seed * 7919 + 53)
    lall, kall, gall = spec.lall, spec.kall, spec.gall
    gz, gzh = vgrid if vgrid is not None else stretched_vgrid(kall)
    dzh = np.diff(gz, prepend=gz[0] - (gz[1] - gz[0]))
    dzh[0] = dzh[1]

    def fields(shape):
        cs = rng.uniform(300.0, 350.0, size=shape)
        a = (cs * dt / dzh[None, :, None] if len(shape) == 3 else cs * dt / dzh[None, :, None]) ** 2 * 1e-2
        Ml = -a * (1.0 + 0.05 * rng.uniform(-1, 1, size=shape))
        Mu = -a * (1.0 + 0.05 * rng.uniform(-1, 1, size=shape))
        Mc = 1.0 + np.abs(Ml) + np.abs(Mu) + rng.uniform(0.05, 0.5, size=shape)
        z = np.maximum(gz, 0.0)[None, :, None]
        gam = 1.0 + z / RADIUS
        G = 1.0 + 0.01 * rng.uniform(-1, 1, size=shape)
        zh = np.maximum(gzh, 0.0)[None, :, None]
        gamh = 1.0 + zh / RADIUS
        Gh = 1.0 + 0.01 * rng.uniform(-1, 1, size=shape)
        rgsgam2h = 1.0 / (Gh * gamh * gamh)
        rho0 = np.exp(-z / 8000.0) * (1.0 + 0.02 * rng.uniform(-1, 1, size=shape))
        return dict(
            Mc=Mc, Ml=Ml, Mu=Mu,
            RGAM=np.broadcast_to(1.0 / gam, shape).copy(),
            RGAMH=np.broadcast_to(1.0 / gamh, shape).copy(),
            RGSGAM2=1.0 / (G * gam * gam),
            RGSGAM2H=rgsgam2h,
            GSGAM2H=1.0 / rgsgam2h,
            rhogw=0.5 * rng.uniform(-1, 1, size=shape) * rho0,
            rhogw0=0.5 * rng.uniform(-1, 1, size=shape) * rho0,
            preg0=100.0 * rng.uniform(-1, 1, size=shape) * rho0,
            rhog0=rho0,
            Srho=1e-5 * rng.uniform(-1, 1, size=shape),
            Sw=1e-3 * rng.uniform(-1, 1, size=shape),
            Spre=1e-1 * rng.uniform(-1, 1, size=shape),
        )

    reg = fields((lall, kall, gall))
    pl = fields((2, kall, spec.gall_pl))
    out = dict(reg)
    for k_, v_ in pl.items():
        out[k_ + "_pl"] = v_
    out["gz"], out["gzh"] = gz, gzh
    return out


def diffusion_inputs(spec: GridSpec, geo: Dict[str, np.ndarray], coef=None) -> Dict[str, np.ndarray]:
    """Scalar, diffusion coefficient and the stencil coefficients of
    OPRT_diffusion (dyn_diffusion/src/mod_oprt.f90:41-60):
    coef_intp(gall,3,3,2,lall) -> [lall][2][3][3][gall], coef_intp_pl(gall_pl,3,3,2),
    coef_diff(gall,6,3,lall) -> [lall][3][6][gall], coef_diff_pl(5,3,2).
    The coefficients are those of coef (icosahedral_geometry's) when given,
    else seeded values of the size 1/dx that OPRT_diffusion_setup gives them;
    parity does not depend on their geometry."""
    rng = np.random.default_rng(spec.seed * 7919 + 61)
    lall, kall, gall, npl = spec.lall, spec.kall, spec.gall, spec.gall_pl
    q2 = _tracer(geo["points"], rng)
    prof = 0.5 + 0.5 * np.cos(np.linspace(0.0, np.pi, kall))
    scl = q2[:, None, :] * prof[None, :, None] + 0.01 * rng.standard_normal((lall, kall, gall))
    kh = 1e5 * (1.0 + 0.5 * rng.uniform(0.0, 1.0, size=(lall, kall, gall)))
    inv_dx = 1.0 / 1e5
    out = dict(scl=scl, scl_pl=rng.uniform(0.0, 1.0, size=(2, kall, npl)),
               kh=kh, kh_pl=1e5 * (1.0 + 0.5 * rng.uniform(0.0, 1.0, size=(2, kall, npl))),
               coef_intp=inv_dx * rng.uniform(-1.0, 1.0, size=(lall, 2, 3, 3, gall)),
               coef_intp_pl=inv_dx * rng.uniform(-1.0, 1.0, size=(2, 3, 3, npl)),
               coef_diff=inv_dx * rng.uniform(-1.0, 1.0, size=(lall, 3, 6, gall)),
               coef_diff_pl=inv_dx * rng.uniform(-1.0, 1.0, size=(2, 3, 5)))
    if coef is not None:
        for k in ("coef_intp", "coef_intp_pl", "coef_diff", "coef_diff_pl"):
            if np.shape(coef[k]) != out[k].shape:
                raise ValueError(f"{k}: shape {np.shape(coef[k])}, the grid needs {out[k].shape}")
            out[k] = np.ascontiguousarray(coef[k], dtype=np.float64)
    return out


def divdamp_inputs(spec: GridSpec, geo: Dict[str, np.ndarray], coef=None) -> Dict[str, np.ndarray]:
    """Momentum, the stencil coefficients and the vertical metrics that
    OPRT3D_divdamp reads (dyn_divdamp/src/mod_oprt3d.f90:41-84):
    VMTR_C2WfactGz(gall,kall,6,lall) -> [lall][6][kall][gall] and
    _pl(gall_pl,kall,6,2), GRD_rdgz(kall), the rest as diffusion_inputs."""
    rng = np.random.default_rng(spec.seed * 7919 + 67)
    lall, kall, gall, npl = spec.lall, spec.kall, spec.gall, spec.gall_pl
    base = diffusion_inputs(spec, geo, coef)
    gz, _ = stretched_vgrid(kall)
    dz = np.diff(gz, append=gz[-1] + (gz[-1] - gz[-2]))
    x = geo["points"]
    v = np.cross(_unit(rng.normal(size=3)) * (20.0 / RADIUS), x)  # solid body, m/s
    prof = np.exp(-np.maximum(gz, 0.0) / 8000.0)

    def reg(c):
        return prof[None, :, None] * v[:, None, :, c] * (1.0 + 0.05 * rng.uniform(-1, 1, size=(lall, kall, gall)))

    def pol(scale=1.0):
        return scale * rng.uniform(-1.0, 1.0, size=(2, kall, npl))

    def metric(shape):
        return 1.0 + 0.01 * rng.uniform(-1.0, 1.0, size=shape)

    return dict(rhogvx=reg(0), rhogvx_pl=pol(10.0), rhogvy=reg(1), rhogvy_pl=pol(10.0),
                rhogvz=reg(2), rhogvz_pl=pol(10.0),
                rhogw=0.1 * rng.standard_normal((lall, kall, gall)), rhogw_pl=pol(0.1),
                coef_intp=base["coef_intp"], coef_intp_pl=base["coef_intp_pl"],
                coef_diff=base["coef_diff"], coef_diff_pl=base["coef_diff_pl"],
                RGSQRTH=metric((lall, kall, gall)), RGSQRTH_pl=metric((2, kall, npl)),
                RGAM=metric((lall, kall, gall)), RGAM_pl=metric((2, kall, npl)),
                RGAMH=metric((lall, kall, gall)), RGAMH_pl=metric((2, kall, npl)),
                C2WfactGz=0.5 * rng.uniform(-1.0, 1.0, size=(lall, 6, kall, gall)) / dz[None, None, :, None],
                C2WfactGz_pl=0.5 * rng.uniform(-1.0, 1.0, size=(2, 6, kall, npl)) / dz[None, None, :, None],
                rdgz=1.0 / dz)


def make_case(spec: GridSpec, have_sgp=None, vgrid=None, geo=None) -> Dict[str, Dict[str, np.ndarray]]:
    """All inputs of the four kernels for one grid (on the given geometry, else
    on make_geometry's seeded patches)."""
    if geo is None:
        geo = make_geometry(spec, have_sgp)
    return dict(geo=geo, flux=flux_inputs(spec, geo), hlim=hlim_inputs(spec, geo),
                vlim=vlim_inputs(spec, geo), vi=vi_inputs(spec, vgrid=vgrid))
