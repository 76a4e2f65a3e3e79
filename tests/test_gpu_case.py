"""Case files end to end on the GPU: the reference's 3D_PEC test case
(testData/maxwellInputs/3D_PEC) read by the library, run with the
reference's integrator (RK4) in FP64 and compared with the oracle on the same
mesh; point probes against the nodal interpolant of the returned state; the
C++ launcher's PointProbe files against the same numbers."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import dgtd_amd as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PEC = os.path.join(ROOT, "tests", "golden", "cases", "3D_PEC.json")
LAUNCH = os.path.join(ROOT, "dgtd_amd", "lib", "dgtd_launch")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_pec_case_fp64_rk4_matches_oracle():
    c = D.Case(PEC)
    T = 0.1  # 10 steps of the case's dt = 1e-2
    x, info = D.run_case(PEC, precision=64, integrator="rk4", final_time=T)
    assert info["steps"] == 10 and info["dt"] == 1e-2
    ref = O.Operator(c.vx, c.EToV, c.face_bc, c.order, alpha=c.info["alpha"])
    xr = c.initial_state()
    steps, t = ref.run(xr, info["dt"], T, "rk4")
    assert steps == info["steps"] and t == pytest.approx(info["time"], abs=1e-14)
    assert rel(x, xr) <= 1e-11


@pytest.mark.parametrize("precision", [64, 32])
def test_point_probes_match_interpolant(precision):
    c = D.Case(PEC)
    op = c.operator(precision=precision, integrator="lserk4")
    try:
        op.set_state(c.initial_state())
        op.advance(c.time_step(op), 7)
        elem = op.probe_set(c.probes)
        got = op.probe_sample()
        x = op.get_state().reshape(6, c.K, -1)
    finally:
        op.close()
    e2, w = D.mesh_locate(c.vx, c.EToV, c.order, c.probes)
    np.testing.assert_array_equal(elem, e2)
    want = np.einsum("pn,cpn->pc", w, x[:, elem, :])
    np.testing.assert_allclose(got, want, rtol=1e-12 if precision == 64 else 1e-6, atol=1e-14)


def test_launcher_writes_reference_probe_files(tmp_path):
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, PEC, "--final-time", "0.05", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["steps"] == 5 and summary["K"] == 1227 and summary["nonfinite"] == 0
    # the same run through the library from Python, probes sampled at the end
    c = D.Case(PEC)
    op = c.operator(precision=64, integrator="rk4")
    try:
        op.set_state(c.initial_state())
        op.run(c.time_step(op), 0.05)
        op.probe_set(c.probes)
        last = op.probe_sample()
    finally:
        op.close()
    for p in range(3):
        lines = (out / "PointProbes" / f"PointProbe{p}.dat").read_text().splitlines()
        assert lines[0] == f"PointProbe ID {p}"
        assert lines[3].startswith("Time (s) // Ex")
        rows = np.array([[float(v) for v in ln.split()] for ln in lines[4:]])
        assert rows.shape == (6, 7)  # t = 0 and after each of the 5 steps (interval 1)
        # times are written as t / c in seconds with 6 significant digits
        np.testing.assert_allclose(rows[:, 0] * 299792458.0, np.arange(6) * 0.01, rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(rows[-1, 1:], last[p], rtol=1e-4, atol=1e-9)


def test_launcher_field_probes_match_point_probe_columns(tmp_path):
    """FieldProbe files (ProbesManager.cpp:222-259, 344-371): one component of
    the point probe at the same position, same times, every second step."""
    case = json.load(open(PEC))
    pts = [p["position"] for p in case["probes"]["point"]]
    case["probes"]["field"] = [{"field_type": "electric", "polarization": "Z", "position": pts[0]},
                               {"field_type": "magnetic", "polarization": "Y", "position": pts[1], "steps": 2}]
    shutil.copy(os.path.join(os.path.dirname(PEC), "3D_PEC.msh"), tmp_path / "3D_PEC.msh")
    path = tmp_path / "3D_PEC.json"
    path.write_text(json.dumps(case))
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, str(path), "--final-time", "0.05", "--out", str(out)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr

    def rows(f, skip=4):
        return np.array([[float(v) for v in ln.split()] for ln in f.read_text().splitlines()[skip:]])

    for fid, (pid, comp, every) in enumerate([(0, 2, 1), (1, 4, 2)]):
        f = out / "FieldProbes" / f"FieldProbe{fid}.dat"
        head = f.read_text().splitlines()[:4]
        assert head[0] == f"FieldProbe ID {fid}" and head[3] == "Time (s) // " + ["Ez", "Hy"][fid]
        fr, pr = rows(f), rows(out / "PointProbes" / f"PointProbe{pid}.dat")
        sel = np.r_[np.arange(0, 6, every), 5] if every > 1 else np.arange(6)  # the final time is always sampled
        sel = np.unique(sel)
        assert fr.shape == (sel.size, 2)
        np.testing.assert_array_equal(fr[:, 0], pr[sel, 0])
        np.testing.assert_array_equal(fr[:, 1], pr[sel, 1 + comp])


def test_launcher_domain_snapshots(tmp_path):
    """3D_Resonant_Box_TM55_H1_P2 (refinement 1, domain_snapshot every 2828
    steps): 3 steps give snapshots at cycle 0 and at the final time."""
    tm55 = os.path.join(ROOT, "tests", "golden", "cases", "3D_Resonant_Box_TM55_H1_P2.json")
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, tm55, "--final-time", "3e-4", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["steps"] == 3 and summary["snapshots"] == 2
    snap = out / "DomainSnapshotProbes"
    assert (snap / "meshes" / "mesh_rank0").exists()
    c = D.Case(tm55)
    x, info = D.run_case(tm55, precision=64, integrator="rk4", final_time=3e-4)
    n = x.size // 6
    for cyc, want, t in (("cycle_000000", c.initial_state(), 0.0), ("cycle_000003", x, info["time"])):
        d = snap / "rank_0" / cyc
        assert float((d / "time.txt").read_text()) == pytest.approx(t, abs=1e-15)
        for k, name in enumerate(["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]):
            lines = (d / f"{name}.gf").read_text().split("\n")
            assert lines[1] == "FiniteElementCollection: L2_T1_3D_P2"
            vals = np.array([float(v) for v in lines[5:] if v.strip()])
            np.testing.assert_allclose(vals, want[k * n:(k + 1) * n], rtol=1e-14, atol=1e-15)


def test_launcher_tfsf_case(tmp_path):
    """3D_TFSF (plane wave through the TF/SF plane x = 1 of a PEC/PMC
    waveguide, refined once): the launcher's run equals the library run, and
    the probe at x = 0 on the scattered-field side stays at rest while the
    pulse (peak at x = t - 1) crosses the total-field side."""
    case = os.path.join(ROOT, "tests", "golden", "cases", "3D_TFSF.json")
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, case, "--final-time", "2.5", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    x, info = D.run_case(case, precision=64, integrator="rk4", final_time=2.5)
    assert summary["steps"] == info["steps"] == 2500
    assert summary["norm2"] == pytest.approx(np.linalg.norm(x), rel=1e-10)
    assert np.linalg.norm(x) > 1.0  # the wave is in the total-field region
    rows = np.array([[float(v) for v in ln.split()] for ln in
                     (out / "PointProbes" / "PointProbe0.dat").read_text().splitlines()[4:]])
    assert rows.shape[0] == 2501 and np.abs(rows[:, 1:]).max() < 1e-3


PEC2D = os.path.join(ROOT, "tests", "golden", "cases", "2D_PEC.json")


def _case_variant(tmp_path, src, **opts):
    case = json.load(open(src))
    case["solver_options"].update(opts)
    d = os.path.dirname(src)
    for f in os.listdir(d):
        if f.startswith(os.path.splitext(os.path.basename(src))[0]) and f.endswith(".msh"):
            shutil.copy(os.path.join(d, f), tmp_path / f)
    path = tmp_path / os.path.basename(src)
    path.write_text(json.dumps(case))
    return str(path)


def test_2d_pec_case_fp64_rk4_matches_oracle():
    from oracle import oracle2d as O2
    c = D.Case(PEC2D)
    x, info = D.run_case(PEC2D, precision=64, integrator="rk4", final_time=0.01)
    assert info["steps"] == 10
    ref = O2.Operator2D(c.vx, c.EToV, c.face_bc, c.order, alpha=c.info["alpha"])
    xr, steps, t = ref.run(c.initial_state(), info["dt"], 0.01, "rk4")
    assert steps == 10 and t == pytest.approx(info["time"], abs=1e-14)
    assert rel(x, xr) <= 1e-11


@pytest.mark.parametrize("alpha,dt,norm_tol", [(0.0, 1e-3, 1e-3), (1.0, 5e-3, 5e-3)])
def test_2d_pec_hesthaven_like_reference(tmp_path, alpha, dt, norm_tol):
    """ExtensiveCasesTest 2D_PEC_Centered_Hesthaven / 2D_PEC_Upwind_Hesthaven
    (test/cases/ExtensiveCasesTest.cpp:205-270, 339-404): a 1D Gaussian in Ez
    between PEC walls at x = 0, 1 (PMC at y = 0, 1) splits, reflects and
    returns after t = 2; the probes on the PEC walls see E ~ 0 and H_y = +-1 at
    t = 1/2, the centre probe E_z ~ 1 at t = 0 and t = 2 (tolerance 2e-2)."""
    path = _case_variant(tmp_path, PEC2D, upwind_alpha=alpha, order=3, evolution_operator="hesthaven",
                         time_step=dt)
    c = D.Case(path)
    op = c.operator(precision=64, integrator="rk4")
    try:
        x0 = c.initial_state()
        op.set_state(x0, 0.0)
        op.probe_set(c.probes)
        movies = [(0.0, op.probe_sample())]
        n = int(round(2.0 / dt))
        for _ in range(n):
            op.advance(dt, 1)
            movies.append((op.GetTime(), op.probe_sample()))
        x = op.get_state()
    finally:
        op.close()
    tol = 2e-2
    nd = x.size // 6
    assert np.linalg.norm(x[2 * nd:3 * nd] - x0[2 * nd:3 * nd]) <= tol  # eOld.DistanceTo(eNew)
    assert abs(np.linalg.norm(x) - np.linalg.norm(x0)) <= norm_tol
    seen = 0
    for t, f in movies:
        for p, sign in ((0, 1.0), (1, -1.0)):
            assert np.abs(f[p, :3]).max() <= tol
            if abs(t - 0.5) <= 1e-3:
                assert f[p, 4] == pytest.approx(sign * 1.0, abs=tol)
                seen += 1
        if abs(t) <= 1e-3 or abs(t - 2.0) <= 1e-3:
            assert f[2, 2] == pytest.approx(1.0, abs=tol) and abs(f[2, 4]) <= tol
    assert seen >= 2


def test_launcher_runs_2d_case(tmp_path):
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, PEC2D, "--final-time", "0.05", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["steps"] == 50 and summary["K"] == 246 and summary["nonfinite"] == 0
    x, info = D.run_case(PEC2D, precision=64, integrator="rk4", final_time=0.05)
    assert summary["norm2"] == pytest.approx(np.linalg.norm(x), rel=1e-10)  # printed with 11 digits
    rows = (out / "PointProbes" / "PointProbe2.dat").read_text().splitlines()
    assert rows[0] == "PointProbe ID 2" and len(rows) == 4 + 51


TFSF2D = os.path.join(ROOT, "tests", "golden", "cases", "2D_TFSF.json")
TFSF2D_INT = os.path.join(ROOT, "tests", "golden", "cases", "2D_TFSF_IntBoundary.json")
EX, EY, EZ, HX, HY, HZ = range(6)


def _probe_movies(path, dt=None):
    c = D.Case(path)
    op = c.operator(precision=64, integrator="rk4")
    try:
        dt = c.time_step(op) if dt is None else dt
        x0 = c.initial_state()
        op.set_state(x0, 0.0)
        op.probe_set(c.probes)
        movies = [(0.0, op.probe_sample())]
        T = c.info["final_time"]
        while op.GetTime() <= T - 1e-8 * dt:
            op.advance(min(dt, T - op.GetTime()), 1)
            movies.append((op.GetTime(), op.probe_sample()))
        x = op.get_state()
    finally:
        op.close()
    return x0, x, movies


def _check_movies(movies, zero_probes, expect, tol):
    hits = 0
    for t, f in movies:
        for p in zero_probes:
            assert np.abs(f[p, [EX, EZ, HX, HY]]).max() <= tol, (t, p, f[p])
        for p, te, want in expect:
            if abs(t - te) <= 1e-3:
                for comp, v in want.items():
                    assert f[p, comp] == pytest.approx(v, abs=tol), (t, p, comp, f[p])
                hits += 1
    assert hits >= len(expect)


@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_2d_tfsf_hesthaven_like_reference(tmp_path, alpha):
    """ExtensiveCasesTest 2D_TFSF_Centered_Hesthaven / 2D_TFSF_Upwind_Hesthaven
    (test/cases/ExtensiveCasesTest.cpp:533-611, 691-769): a Gaussian plane
    wave (E along y) enters through the TF/SF line and bounces in the PEC/PMC
    guide; tolerance 1e-2."""
    path = _case_variant(tmp_path, TFSF2D, upwind_alpha=alpha, evolution_operator="hesthaven")
    x0, x, movies = _probe_movies(path)
    assert np.linalg.norm(x0) == 0.0
    assert movies[-1][0] + 1e-2 > 9.5
    _check_movies(movies, (0, 1, 2),
                  [(0, 9.8, {HZ: 0.0, EY: 2.0}), (1, 2.8, {HZ: 1.0, EY: 1.0}), (1, 8.8, {HZ: -1.0, EY: 1.0}),
                   (2, 5.8, {HZ: 0.0, EY: 2.0})], 1e-2)


def test_2d_tfsf_interior_pec_like_reference(tmp_path):
    """ExtensiveCasesTest 2D_TFSF_InteriorPEC_Hesthaven (:85-200): the wave is
    injected, reflected by an interior PEC line and leaves through SMA; behind
    the interior PEC the field stays at rest; tolerance 2e-2."""
    path = _case_variant(tmp_path, TFSF2D_INT, evolution_operator="hesthaven")
    x0, x, movies = _probe_movies(path)
    tol = 2e-2
    assert np.linalg.norm(x0) <= tol
    assert abs(np.linalg.norm(x) - np.linalg.norm(x0)) <= tol
    _check_movies(movies, (1, 3),
                  [(1, 0.0, {EY: 0.0, HZ: 0.0}), (3, 0.0, {EY: 0.0, HZ: 0.0}),
                   (1, 2.1, {EY: 1.0, HZ: 1.0}), (3, 2.1, {EY: 0.0, HZ: 0.0}),
                   (2, 2.9, {EY: 0.0, HZ: 2.0}), (3, 2.9, {EY: 0.0, HZ: 0.0}),
                   (1, 3.7, {EY: -1.0, HZ: 1.0}), (3, 3.7, {EY: 0.0, HZ: 0.0})], tol)


QUADS2D = os.path.join(ROOT, "tests", "golden", "cases", "2D_TFSF_Quads.json")


@pytest.mark.parametrize("alpha,dt", [(0.0, 5e-3), (1.0, None)])
def test_2d_tfsf_quads_like_reference(tmp_path, alpha, dt):
    """ExtensiveCasesTest 2D_TFSF_Centered_Quads (alpha 0, time step 5e-3) and
    2D_TFSF_Upwind_Quads (the case as written: alpha 1, time step 0.01)
    (test/cases/ExtensiveCasesTest.cpp:1003-1157): quadrilaterals on the
    Global (assembled) operator; the Gaussian plane wave enters through the
    TF/SF line x = 1 and bounces in the PEC/PMC guide; tolerance 1e-2."""
    opts = dict(upwind_alpha=alpha)
    if dt is not None:
        opts["time_step"] = dt
    path = _case_variant(tmp_path, QUADS2D, **opts)
    c = D.Case(path)
    assert c.quads and c.info["evolution_operator"] == "global"
    x0, x, movies = _probe_movies(path)
    assert np.linalg.norm(x0) == 0.0
    assert movies[-1][0] + 1e-2 > 9.5
    _check_movies(movies, (0, 1, 2),
                  [(0, 9.8, {HZ: 0.0, EY: 2.0}), (1, 2.8, {HZ: 1.0, EY: 1.0}), (1, 8.8, {HZ: -1.0, EY: 1.0}),
                   (2, 5.8, {HZ: 0.0, EY: 2.0})], 1e-2)


def test_launcher_runs_quad_case(tmp_path):
    """The C++ launcher takes the quadrilateral case to the assembled
    operator and matches the library run."""
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, QUADS2D, "--final-time", "0.5", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["steps"] == 50 and summary["K"] == 60 and summary["nonfinite"] == 0
    x, info = D.run_case(QUADS2D, precision=64, integrator="rk4", final_time=0.5)
    assert info["steps"] == 50
    assert summary["norm2"] == pytest.approx(np.linalg.norm(x), rel=1e-10)
    rows = (out / "PointProbes" / "PointProbe0.dat").read_text().splitlines()
    assert rows[0] == "PointProbe ID 0"


def test_launcher_runs_hexahedral_case(tmp_path):
    """sma_upwind_hexa_1dot5D (Solver3DTest.cpp:155-215) as a case file: a
    Gmsh mesh of 10 x 2 x 2 hexahedra on [0,1] x [0,0.4] x [0,0.4] with
    MakeCartesian3D's side tags (PEC z, PMC y, SMA x), the Gaussian E_z pulse,
    order 3, time step 7e-3, final time 1, point probes at x = 0 and x = 1.
    The launcher runs it on the assembled operator; its PointProbe files show
    the two halves leaving through the SMA ends (E_z = 0.5, H_y = +-0.5 at the
    sample nearest t = 0.5; zero at t = 1; tolerance 1e-2), and its final norm
    matches the library run."""
    from test_case_files import write_gmsh_hexes

    write_gmsh_hexes(str(tmp_path / "h.msh"), 10, 2, 2, sx=1.0, sy=0.4, sz=0.4)
    case = {"solver_options": {"upwind_alpha": 1.0, "evolution_operator": "global", "time_step": 7e-3,
                               "final_time": 1.0, "order": 3},
            "model": {"filename": "h.msh", "materials": [{"tags": [1], "type": "vacuum"}],
                      "boundaries": [{"tags": [1, 6], "type": "PEC"}, {"tags": [2, 4], "type": "PMC"},
                                     {"tags": [3, 5], "type": "SMA"}]},
            "probes": {"exporter": {"steps": 1},
                       "point": [{"position": [0.0, 0.2, 0.2]}, {"position": [1.0, 0.2, 0.2]}]},
            "sources": [{"type": "initial", "field_type": "electric", "center": [0.5, 0.2, 0.2],
                         "polarization": [0.0, 0.0, 1.0], "dimension": 1,
                         "magnitude": {"type": "gaussian", "spread": 0.1}}]}
    path = tmp_path / "hexa.json"
    path.write_text(json.dumps(case))
    out = tmp_path / "exp"
    r = subprocess.run([LAUNCH, str(path), "--out", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["K"] == 40 and summary["steps"] == 143 and summary["nonfinite"] == 0
    assert summary["norm2"] <= 1e-2
    x, info = D.run_case(str(path), precision=64, integrator="rk4")
    assert summary["norm2"] == pytest.approx(np.linalg.norm(x), rel=1e-9, abs=1e-12)
    tol = 1e-2
    for p, sign in ((0, 1.0), (1, -1.0)):
        lines = (out / "PointProbes" / f"PointProbe{p}.dat").read_text().splitlines()
        rows = np.array([[float(v) for v in ln.split()] for ln in lines[4:]])
        t = rows[:, 0] * 299792458.0
        assert rows.shape == (144, 7) and t[-1] == pytest.approx(1.0, abs=1e-5)
        mid = rows[np.argmin(np.abs(t - 0.5))]
        assert mid[1 + EZ] == pytest.approx(0.5, abs=tol) and mid[1 + HY] == pytest.approx(0.5 * sign, abs=tol)
        assert abs(rows[-1, 1 + EZ]) <= tol and abs(rows[-1, 1 + HY]) <= tol
