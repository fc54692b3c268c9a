"""CPU oracle of the FP64 tracer-mass sums and of the flux-form update that
conserves mass exactly -- TEST INFRASTRUCTURE ONLY.

Neither routine exists in the reference: its kernels stop at the face fluxes
(FLX/src/mod_src_tracer.f90:234-309, flx_h(o,a) = flux*RAREA(o)*dt) and the
limited face values (HLIM/src/mod_src_tracer.f90:377-410, mirrored onto the
neighbour's face).  The update completes the flux form they feed:

    rhogq(c) -= RAREA(c) * sum_{m=1..6} F(c, m)

where F(c, m) is the outward mass transfer through face m.  An arc between
two inner cells, from origin o in direction a = 1..3, transfers
X = AREA(o)*flx_h(o,a)*q_a(o,a) out of o and -X out of its other cell; a face
toward a halo point takes the inner cell's own AREA*flx_h*q_a (halo cells may
carry no metrics); a face two regions share takes the transfer of the
lower-indexed region; the pole centre takes minus the transfers of the region
faces pointing at it; face 6 of a singular-point cell carries nothing
(FLX/src/mod_src_tracer.f90:311-320).  The partner of a shared face is found
here from the halo map by a search of our own, written independently of the
library's (csrc/comm.cpp), so the two are checked against each other.

Arrays are C arrays with the Fortran dimensions reversed: rhogq[l][k][g],
flx_h[m][l][k][g], GMTR_p[n][l][1][g], GMTR_p_pl[n][p][1][v].
"""
from __future__ import annotations

import math

import numpy as np

DI = (0, 1, 1, 0, -1, -1, 0)  # face m = 1..6: ip1j ip1jp1 ijp1 im1j im1jm1 ijm1
DJ = (0, 0, 1, 1, 0, -1, -1)


def arc_partners(hm, gall_1d: int, gall_pl: int = 6):
    """Face partners from the halo map hm (runtime.halo_map's dict).

    Returns (faces, pole): faces[(l, g, m)] = (l2, g2, m2) for every face of a
    region cell that takes its partner's transfer; (l, g, 6) -> None for the
    zero face of a singular-point cell; pole[(p, v)] = (l2, g2, m2), v = 1..5.
    """
    src = {}
    for e in range(len(hm["dst_kind"])):
        src[(int(hm["dst_kind"][e]), int(hm["dst_l"][e]), int(hm["dst_g"][e]))] = (
            int(hm["src_kind"][e]), int(hm["src_l"][e]), int(hm["src_g"][e]))
    sgp = hm["have_sgp"]
    lall, n1 = len(sgp), gall_1d
    lo, hi = 1, n1 - 2

    def zero_face(l, g, m):
        return m == 6 and sgp[l] and g == lo * n1 + lo

    def toward(l, g, target):
        i, j = g % n1, g // n1
        hits = []
        for m in range(1, 7):
            ni, nj = i + DI[m], j + DJ[m]
            if lo <= ni <= hi and lo <= nj <= hi or zero_face(l, g, m):
                continue
            if src.get((0, l, nj * n1 + ni)) == target:
                hits.append(m)
        assert len(hits) == 1, (l, g, target, hits)
        return hits[0]

    faces = {}
    for l in range(lall):
        for j in range(lo, hi + 1):
            for i in range(lo, hi + 1):
                g = j * n1 + i
                for m in range(1, 7):
                    ni, nj = i + DI[m], j + DJ[m]
                    if lo <= ni <= hi and lo <= nj <= hi:
                        continue
                    if zero_face(l, g, m):
                        faces[(l, g, m)] = None
                        continue
                    kind, l2, g2 = src[(0, l, nj * n1 + ni)]
                    if kind == 0 and l2 < l:
                        faces[(l, g, m)] = (l2, g2, toward(l2, g2, (0, l, g)))
    pole = {}
    for p in range(2):
        for v in range(1, gall_pl):
            kind, l2, g2 = src[(1, p, v)]
            assert kind == 0
            pole[(p, v)] = (l2, g2, toward(l2, g2, (1, p, 0)))
    return faces, pole


def _own_x(gp, flx, qa, l, g, m):
    """[k] AREA(g)*flx_h(g,m)*q_a(g,m) in FP64, the products in that order."""
    return (gp[0, l, 0, g] * flx[m - 1, l, :, g].astype(np.float64)) * qa[m - 1, l, :, g].astype(np.float64)


def _face_x(gp, flx, qa, n1, l, g, m):
    """[k] outward transfer of inner cell g through face m by region l's arcs:
    faces 1-3 its own; faces 4-6 minus the origin's arc when the origin is
    inner, else its own side (halo cells may carry no metrics)."""
    if m <= 3:
        return _own_x(gp, flx, qa, l, g, m)
    i, j = g % n1, g // n1
    oi, oj = i + DI[m], j + DJ[m]
    if oi < 1 or oj < 1:
        return _own_x(gp, flx, qa, l, g, m)
    return -_own_x(gp, flx, qa, l, oj * n1 + oi, m - 3)


def update_conservative(rhogq, rhogq_pl, flx_h, q_a, gp, gp_pl, partners, gall_1d: int):
    """The conservative update on copies of rhogq[l][k][g], rhogq_pl[p][k][v]
    (FP64), with flx_h, q_a [6][l][k][g] (FP32 or FP64); the same operations
    in the same order as the library's kernel: each face's transfer, the six
    summed from face 1 to 6, then rhogq - RAREA*sum."""
    faces, pole = partners
    rq, rq_pl = rhogq.astype(np.float64).copy(), rhogq_pl.astype(np.float64).copy()
    n1 = gall_1d
    lall = rq.shape[0]
    for l in range(lall):
        for j in range(1, n1 - 1):
            for i in range(1, n1 - 1):
                g = j * n1 + i
                s = np.zeros(rq.shape[1])
                for m in range(1, 7):
                    key = (l, g, m)
                    if key in faces:
                        pr = faces[key]
                        f = np.zeros(rq.shape[1]) if pr is None else -_face_x(gp, flx_h, q_a, n1, *pr)
                    else:
                        f = _face_x(gp, flx_h, q_a, n1, l, g, m)
                    s = s + f
                rq[l, :, g] = rq[l, :, g] - gp[1, l, 0, g] * s
    npl = rq_pl.shape[2] - 1
    for p in range(2):
        s = np.zeros(rq.shape[1])
        for v in range(1, npl + 1):
            s = s + -_face_x(gp, flx_h, q_a, n1, *pole[(p, v)])
        rq_pl[p, :, 0] = rq_pl[p, :, 0] - gp_pl[1, p, 0, 0] * s
    return rq, rq_pl


def tracer_mass(rhogq, rhogq_pl, gp, gp_pl, gall_1d: int, have_pl: bool = True):
    """[k] sum of AREA*rhogq over the inner cells and both pole centres,
    summed exactly (math.fsum) and rounded once."""
    n1 = gall_1d
    lall, kall = rhogq.shape[0], rhogq.shape[1]
    inner = np.zeros((n1, n1), dtype=bool)
    inner[1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    out = np.empty(kall)
    for k in range(kall):
        terms = (gp[0, :, 0, :][:, inner] * rhogq[:, k, :][:, inner].astype(np.float64)).ravel().tolist()
        if have_pl:
            terms += [gp_pl[0, p, 0, 0] * float(rhogq_pl[p, k, 0]) for p in range(2)]
        out[k] = math.fsum(terms)
    return out
