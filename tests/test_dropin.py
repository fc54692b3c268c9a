"""The reference's own kernel drivers on our snapshot files, and the same
drivers relinked against the Fortran shim over libnicam_tt.

Each driver (kernels/NICAM/<kernel>/src/main.f90, unchanged) reads
snapshot.<kernel>.pe000000 -- inputs plus the outputs it checks its result
against -- runs the kernel once and prints the reference's DEBUG_valuecheck
lines (max, min, sum of every array), then the "point-by-point diff" between
its result and the snapshot's outputs.  oracle/Makefile (target drivers)
builds two executables per kernel into oracle/_ref: drv_<k>_ref.exe from
the reference sources alone, and drv_<k>_gpu.exe with the kernel module
replaced by icoatmosbenchmark_v1_amd/fortran/shim_<k>.f90.

We write the snapshot (icoatmosbenchmark_v1_amd.snapshot) from seeded inputs
at the drivers' compiled-in size (gall_1d 130, kall 42, one region with a
singular point, poles on) with the oracle's outputs as the expected values:
* CPU: the reference driver reads our file as we wrote it (its input checks
  match the arrays) and reproduces the oracle's outputs to the bit (diff 0);
* GPU: the shim-linked driver prints exactly the reference driver's lines.
"""
import re
import resource
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from icoatmosbenchmark_v1_amd import runtime, snapshot, synth

REPO = Path(__file__).resolve().parent.parent
REF = REPO / "oracle" / "_ref"
VGRID = REPO / "tests" / "golden" / "vgrid40_600m_24km.dat"
KERNELS = {"flux": "dyn_horiz_adv_flux", "hlim": "dyn_horiz_adv_limiter", "vlim": "dyn_vert_adv_limiter",
           "vi": "dyn_vi_rhow_solver", "diff": "dyn_diffusion", "ddmp": "dyn_divdamp"}
# the metrics driver's shim runs the library's host set-up routines: no GPU
HOST_KERNELS = {"met": "dyn_metrics"}
N1, KALL, LALL = 130, 42, 1  # problem_size.inc:28-45 of every driver
CHECK = re.compile(r"\+check\[(.{16})\] max=(.*),min=(.*),sum=(.*)$")


def _exe(short, kind):
    p = REF / f"drv_{short}_{kind}.exe"
    if not p.exists():
        pytest.skip(f"{p.name} not built (oracle/Makefile target drivers needs the reference tree)")
    return p


def _unlimited_stack():
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run_driver(exe: Path, workdir: Path, timeout=300):
    out = subprocess.run([str(exe)], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=timeout, preexec_fn=_unlimited_stack)
    assert out.returncode == 0, f"{exe.name} failed ({out.returncode}):\n{out.stdout[-3000:]}"
    sections, cur = {}, None
    for ln in out.stdout.splitlines():
        if "### Input ###" in ln:
            cur = "input"
        elif "### Output ###" in ln:
            cur = "output"
        elif "point-by-point diff" in ln:
            cur = "diff"
        m = CHECK.search(ln)
        if m and cur:
            sections.setdefault(cur, []).append((m.group(1).strip(), m.group(2).strip(), m.group(3).strip(),
                                                 m.group(4).strip()))
    assert set(sections) == {"input", "output", "diff"}, out.stdout[-3000:]
    return sections


def make_snapshot(short, oracle, workdir: Path):
    """Writes the snapshot (and the vertical grid file) into workdir; returns
    the fields as written."""
    spec = synth.GridSpec(N1, KALL, LALL, have_pl=True, seed=91)
    gz, gzh = runtime.grd_input_vgrid(str(VGRID), KALL)
    case = synth.make_case(spec, have_sgp=[1], vgrid=(gz, gzh))
    geo = case["geo"]
    if short == "flux":
        f = case["flux"]
        out = oracle.horizontal_flux(spec, geo, f, 360.0)  # SET_dt_large
        fields = {**out, **{k: f[k] for k in ("rho", "rho_pl", "rhovx", "rhovx_pl", "rhovy", "rhovy_pl",
                                              "rhovz", "rhovz_pl")},
                  **{k: geo[k] for k in ("GRD_xr", "GRD_xr_pl", "GMTR_p", "GMTR_p_pl", "GMTR_t", "GMTR_t_pl",
                                          "GMTR_a", "GMTR_a_pl")}}
    elif short == "hlim":
        f = case["hlim"]
        out = oracle.horizontal_limiter_thuburn(spec, [1], f)
        fields = {**f, **out, "q_a_prev": f["q_a"], "q_a_prev_pl": f["q_a_pl"]}
    elif short == "vlim":
        f = case["vlim"]
        out = oracle.vertical_limiter_thuburn(spec, f)
        fields = {**f, **out, "q_h_prev": f["q_h"], "q_h_prev_pl": f["q_h_pl"]}
    elif short == "diff":
        f = synth.diffusion_inputs(spec, geo)
        fields = {**f, **oracle.oprt_diffusion(spec, [1], f)}
    elif short == "ddmp":  # the driver's GRD_setup reads GRD_rdgz from the vertical grid file
        f = dict(synth.divdamp_inputs(spec, geo), rdgz=oracle.grd_setup(gz, gzh)["rdgz"])
        fields = {**f, **oracle.oprt3d_divdamp(spec, [1], f)}
        shutil.copy(VGRID, workdir / VGRID.name)
    else:
        f = case["vi"]
        g = oracle.grd_setup(gz, gzh)
        out = oracle.vi_rhow_solver(spec, f, 60.0, g["rdgzh"], g["afact"], g["bfact"])  # SET_dt_small
        fields = {**f, **out, "rhogw_prev": f["rhogw"], "rhogw_prev_pl": f["rhogw_pl"]}
        shutil.copy(VGRID, workdir / VGRID.name)
    snapshot.write_snapshot(workdir / snapshot.snapshot_name(KERNELS[short]), KERNELS[short], fields, N1, KALL, LALL)
    return fields


def make_metrics_snapshot(reflibs, workdir: Path):
    """One region with a singular point of the generated gl8/rl1 grid
    (gall_1d 130, the driver's compiled-in size) and the poles.  The driver
    restores the halo and pole-ring entries of GMTR_p from the snapshot after
    GMTR_p_setup (dyn_metrics/src/main.f90:336-350), as NICAM fills them by a
    halo exchange, so GMTR_p is written with its halo filled through the halo
    map; the expected metrics are the library's, the expected stencil
    coefficients those of the compiled reference on that GMTR_p."""
    sys_path = str(REPO / "tests")
    import sys

    if sys_path not in sys.path:
        sys.path.insert(0, sys_path)
    from halo_sim import apply_map

    gl, rl = 8, 1
    x, xpl = runtime.grid_icosahedral(gl, rl)
    hm = runtime.halo_map(rl, N1)
    sgp = hm["have_sgp"]
    v = runtime.grid_vertices(x, xpl, sgp)
    m = runtime.gmtr_setup(x, xpl, v["xt"], v["xt_pl"], v["s"], v["s_pl"], sgp)
    # GMTR_p(gall,1,lall,8): the 8 components are the variables of one exchange
    p_full, p_pl = apply_map(hm, m["p"][:, :, 0:1, :], m["p_pl"][:, :, 0:1, :])
    L = int(np.flatnonzero(sgp)[0])
    reg = lambda a: np.ascontiguousarray(a[..., L:L + 1, :, :])  # noqa: E731  (..., lall, 1, gall)
    one = [1]
    fields = {"GRD_x": reg(x), "GRD_x_pl": xpl, "GRD_xt": reg(v["xt"]), "GRD_xt_pl": v["xt_pl"],
              "GRD_s": reg(v["s"]), "GRD_s_pl": v["s_pl"], "GMTR_p": reg(p_full), "GMTR_p_pl": p_pl,
              "GMTR_t": reg(m["t"]), "GMTR_t_pl": m["t_pl"], "GMTR_a": reg(m["a"]), "GMTR_a_pl": m["a_pl"]}
    mr = {"p": fields["GMTR_p"], "p_pl": p_pl, "t": fields["GMTR_t"], "t_pl": m["t_pl"],
          "a": fields["GMTR_a"], "a_pl": m["a_pl"]}
    for which, name in (("div", "div"), ("rot", "rot"), ("grad", "grad"), ("lap", "lap")):
        c, c_pl = reflibs.oprt_coef_setup(which, mr, one)
        fields[f"OPRT_coef_{name}"], fields[f"OPRT_coef_{name}_pl"] = c, c_pl
    d = reflibs.oprt_diffusion_setup(mr, one)
    for k in ("coef_intp", "coef_intp_pl", "coef_diff", "coef_diff_pl"):
        fields["OPRT_" + k] = d[k]
    snapshot.write_snapshot(workdir / snapshot.snapshot_name("dyn_metrics"), "dyn_metrics", fields, N1, KALL, LALL)
    return fields


def test_snapshot_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    sizes = snapshot.field_sizes("dyn_vert_adv_limiter", 6, 5, 2)
    fields = {k: rng.standard_normal(n) for k, n in sizes}
    p = snapshot.write_snapshot(tmp_path / "s", "dyn_vert_adv_limiter", fields, 6, 5, 2)
    assert p.stat().st_size == 8 * sum(n for _, n in sizes)
    raw = p.read_bytes()
    assert raw[:8] == struct.pack(">d", fields["q_h_prev"][0])  # big-endian, no header
    back = snapshot.read_snapshot(p, "dyn_vert_adv_limiter", 6, 5, 2)
    for k in fields:
        assert np.array_equal(back[k], fields[k])
    with pytest.raises(ValueError):
        snapshot.write_snapshot(tmp_path / "t", "dyn_vert_adv_limiter", {**fields, "ck": fields["q"]}, 6, 5, 2)


def test_vgrid_writer_reproduces_the_shipped_file(tmp_path):
    gz, gzh = runtime.grd_input_vgrid(str(VGRID), KALL)
    p = snapshot.write_vgrid(tmp_path / "v.dat", gz, gzh)
    assert p.read_bytes() == VGRID.read_bytes()


def _close_sum(printed: str, a: np.ndarray) -> bool:
    """A printed sum against ours: the driver sums in storage order, numpy
    pairwise, so allow the recursive-summation error bound n*eps*sum|a|."""
    return abs(float(printed) - float(np.sum(a))) <= a.size * 2.3e-16 * float(np.sum(np.abs(a)))


@pytest.mark.ref
@pytest.mark.parametrize("short", list(KERNELS))
def test_reference_driver_on_our_snapshot(short, oracle, tmp_path):
    fields = make_snapshot(short, oracle, tmp_path)
    sec = run_driver(_exe(short, "ref"), tmp_path)
    # the driver's view of what it read: max and min exact, sum to its own order
    for name, mx, mn, sm in sec["input"]:
        key = name.removeprefix("check_")
        if key.startswith("VMTR_"):
            key = key[5:]
        if key not in fields:
            continue
        a = np.asarray(fields[key], dtype=np.float64)
        assert float(mx) == a.max() and float(mn) == a.min(), name
        assert _close_sum(sm, a), name
    # the reference reproduces the oracle's outputs exactly
    for name, mx, mn, sm in sec["diff"]:
        assert float(mx) == 0.0 and float(mn) == 0.0 and float(sm) == 0.0, f"{short}: {name} {mx} {mn} {sm}"


@pytest.mark.gpu
@pytest.mark.parametrize("short", list(KERNELS))
def test_shim_driver_matches_reference_driver(short, gpu, oracle, tmp_path):
    make_snapshot(short, oracle, tmp_path)
    ref = run_driver(_exe(short, "ref"), tmp_path)
    got = run_driver(_exe(short, "gpu"), tmp_path)
    assert got == ref
    for name, mx, mn, sm in got["diff"]:
        assert float(mx) == 0.0 and float(mn) == 0.0 and float(sm) == 0.0, f"{short}: {name}"


@pytest.mark.ref
def test_metrics_drivers(lib_built, reflibs, tmp_path):
    """The reference's metrics driver (dyn_metrics/src/main.f90, unchanged)
    on our snapshot reproduces every metric and stencil coefficient (diff 0);
    relinked against the shim over the library's host routines it prints
    exactly the same check lines."""
    fields = make_metrics_snapshot(reflibs, tmp_path)
    ref = run_driver(_exe("met", "ref"), tmp_path)
    for name, mx, mn, sm in ref["input"]:
        key = name.removeprefix("check_")
        if key in fields:
            a = np.asarray(fields[key], dtype=np.float64)
            assert float(mx) == a.max() and float(mn) == a.min(), name
    for name, mx, mn, sm in ref["diff"]:
        assert float(mx) == 0.0 and float(mn) == 0.0 and float(sm) == 0.0, name
    got = run_driver(_exe("met", "gpu"), tmp_path)
    assert got == ref


@pytest.fixture(scope="module")
def lib_built():
    if not runtime.LIB_PATH.exists():
        from icoatmosbenchmark_v1_amd import _build

        _build.build_library()
    return runtime.load_library()
