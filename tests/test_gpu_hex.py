"""Hexahedra on the GPU: the assembled operator (DGTD_KERNEL_CSR), the
operator the reference runs on hexahedral meshes (its Global operator,
GlobalEvolution::Mult, GlobalEvolution.cpp:628-1088).

Bars: FP64 Mult equals the hexahedron oracle (oracle/oracle_hex.py, pinned
in tests/test_oracle_hex.py) to 1e-12 relative, TF/SF plane waves included;
FP64 RK4 / LSERK4 runs equal the oracle's to 1e-11 relative L2; FP32 agrees
with FP64 to 1e-5; the reference's Solver3DTest cases pec_global_1dot5D
(hexahedra and tetrahedra) and sma_upwind_hexa_1dot5D
(Solver3DTest.cpp:57-215) meet their own tolerances.  The matrix-free
hexahedron stage (kernel "valu", the default on hexahedra) meets the same
bars; the matrix-core kernels refuse hexahedra.
"""
import numpy as np
import pytest

from oracle import oracle_hex as OH
from test_oracle_hex import sheared_hexes

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402

EX, EY, EZ, HX, HY, HZ = range(6)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def hex_case(seed):
    vx, E, bc = sheared_hexes(3, 2, 2, bc=("pec", "pmc", "sma", "pec", "pmc", "sma"), sx=1.5)
    bc = bc.copy()
    bc[0, 2], bc[1, 4] = 4, 4
    bc[1, 3], bc[4, 1] = 7, 7
    bc[4, 5], bc[10, 0] = 5, 5
    bc[2, 3], bc[5, 1] = 6, 6
    rng = np.random.default_rng(seed)
    return vx, E, bc, rng.uniform(1, 4, E.shape[0]), rng.uniform(1, 2, E.shape[0])


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_hex_mult_matches_oracle(order):
    vx, E, bc, eps, mu = hex_case(order)
    for alpha in (1.0, 0.0):
        ref = OH.OperatorHex(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu)
        x = np.random.default_rng(order).standard_normal(ref.ndofs())
        yr = ref.mult(x)
        g64 = D.GlobalEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=64)
        assert g64.kernel_variant() == "csr" and g64.Np == (order + 1) ** 3
        assert np.abs(g64.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
        g32 = D.GlobalEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=32)
        assert rel(g32.mult(x), yr) <= 1e-5
        g64.close()
        g32.close()


def test_hex_kernel_choice_positions_and_probes():
    vx, E, bc = sheared_hexes(2, 2, 1)
    op = D.HesthavenEvolution(vx, E, bc, 3)  # kernel "auto" takes the matrix-free hexahedron stage
    assert op.kernel_variant() == "valu"
    ref = OH.OperatorHex(vx, E, bc, 3)
    pos = op.node_positions()
    for d in range(3):
        np.testing.assert_allclose(pos[d], ref.pos[..., d], atol=1e-14)
    assert op.max_fscale() == pytest.approx(ref.max_fscale(), rel=1e-14)
    x = np.random.default_rng(1).standard_normal(op.ndofs)
    op.set_state(x)
    pts = np.array([[0.4, 0.3, 0.2], [1.2, 0.7, 0.9]])
    elem = op.probe_set(pts)
    assert (elem >= 0).all()
    _, w = D.mesh_locate(vx, E, 3, pts)
    q = x.reshape(6, op.K, op.Np)
    want = np.einsum("pn,cpn->pc", w, q[:, elem])
    np.testing.assert_allclose(op.probe_sample(), want, atol=1e-13)
    with pytest.raises(D.DgtdError, match="matrix cores") as ei:
        D.HesthavenEvolution(vx, E, bc, 3, kernel="mfma")
    assert ei.value.code == -4
    with pytest.raises(D.DgtdError):
        op.set_candidate(0)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_hex_matrix_free_matches_oracle(order):
    """The matrix-free hexahedron stage (kernel "valu", hex_stage.hip)
    against the oracle: every face code, random eps, mu, conductivity, both
    fluxes; FP64 to 1e-12, FP32 to 1e-5; and its RK4 / LSERK4 steps equal
    the assembled operator's to 1e-12."""
    vx, E, bc, eps, mu = hex_case(order)
    sigma = np.where(np.arange(E.shape[0]) % 3 == 0, 0.7, 0.0)
    for alpha in (1.0, 0.0, 0.5):
        ref = OH.OperatorHex(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu)
        x = np.random.default_rng(order).standard_normal(ref.ndofs())
        yr = ref.mult(x)
        m64 = D.HesthavenEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=64, kernel="valu")
        assert m64.kernel_variant() == "valu"
        assert np.abs(m64.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
        m32 = D.HesthavenEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=32, kernel="valu")
        assert rel(m32.mult(x), yr) <= 1e-5
        m64.close()
        m32.close()
    for integrator in ("rk4", "lserk4"):
        a = D.HesthavenEvolution(vx, E, bc, order, eps=eps, mu=mu, sigma=sigma, precision=64, kernel="valu",
                                 integrator=integrator)
        g = D.GlobalEvolution(vx, E, bc, order, eps=eps, mu=mu, sigma=sigma, precision=64, integrator=integrator)
        x0 = np.random.default_rng(7).standard_normal(a.ndofs)
        assert rel(a.mult(x0), g.mult(x0)) <= 1e-12  # conductivity included
        dt = 0.3 / (a.max_fscale() * (order + 1) ** 2)
        a.set_state(x0)
        g.set_state(x0)
        assert a.run(dt, 5.5 * dt) == g.run(dt, 5.5 * dt) == 6
        assert rel(a.get_state(), g.get_state()) <= 1e-12
        a.close()
        g.close()


def _hex_waveguide(nx=16, length=4.0, x_tfsf=1.0):
    """[0, length] x [0, 0.5] x [0, 0.5] hexahedra: SMA ends (x), PEC z walls,
    PMC y walls, TF/SF plane x = x_tfsf."""
    vx, E, bc = D.cartesian_3d(nx, 1, 1, "hex", sx=length, sy=0.5, sz=0.5,
                               bc=("pec", "pmc", "sma", "pmc", "sma", "pec"))
    bc = bc.copy()
    for e in range(E.shape[0]):
        for f, cyc in enumerate(OH.HEX_FACES):
            if np.all(np.abs(vx[E[e, list(cyc)], 0] - x_tfsf) < 1e-12):
                bc[e, f] = 7
    assert (bc == 7).sum() == 2
    return vx, E, bc


@pytest.mark.parametrize("kernel", ["csr", "valu"])
@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
def test_hex_tfsf_run_matches_oracle(integrator, kernel):
    vx, E, bc = _hex_waveguide()
    wave = dict(polarization=(0, 0, 1), propagation=(1, 0, 0), spread=0.3, mean=-1.0)
    ref = OH.OperatorHex(vx, E, bc, 3)
    ref.set_planewaves([wave])
    gpu = D.HesthavenEvolution(vx, E, bc, 3, precision=64, integrator=integrator, kernel=kernel)
    assert gpu.kernel_variant() == kernel
    gpu.set_planewaves([D.planewave((0, 0, 1), (1, 0, 0), 0.3, -1.0)])
    x = np.random.default_rng(2).standard_normal(ref.ndofs())
    gpu.SetTime(0.9)
    yr = ref.mult(x, 0.9)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
    dt = 0.4 / (ref.max_fscale() * 9)
    xr, steps, _ = ref.run(np.zeros(ref.ndofs()), dt, 2.2, integrator)
    gpu.set_state(np.zeros(ref.ndofs()))
    assert gpu.run(dt, 2.2) == steps
    assert rel(gpu.get_state(), xr) <= 1e-11
    # the pulse peak (x = 1.2 at t = 2.2) is on the total-field side, which
    # carries the incident wave (7e-4 relative on the oracle); the scattered
    # side stays empty
    q = gpu.get_state().reshape(6, gpu.K, gpu.Np)
    Ei, Hi = ref.incident(ref.pos, gpu.GetTime())
    total = ref.pos[..., 0].mean(axis=1) > 1.0
    assert rel(q[EZ][total], Ei[2][total]) < 2e-3 and rel(q[HY][total], Hi[1][total]) < 2e-3
    assert np.abs(q[:, ~total]).max() < 1e-12


def _solver3d(shape, nx, ny, nz, s, bcs, order, alpha, dt, T, centre, probes=None):
    """Solver3DTest's set-up (Solver3DTest.cpp:19-42): MakeCartesian3D mesh,
    boundary attributes 1..6, Gaussian E_z of spread 0.1 varying along x
    (SourceFixtures.h:11-26), the Global operator, RK4, fixed time step."""
    vx, E, fbc = D.cartesian_3d(nx, ny, nz, shape, sx=s[0], sy=s[1], sz=s[2], bc=bcs)
    op = D.GlobalEvolution(vx, E, fbc, order, alpha=alpha, precision=64, integrator="rk4")
    x0 = D.project_initial_fields(vx, E, order, [D.gaussian("E", centre, 0.1, (0, 0, 1), 1)])
    sol = D.Solver(op, x0, final_time=T, time_step=dt, evolution_operator="global")
    movie = []
    if probes is not None:
        op.probe_set(probes)
        movie.append((0.0, op.probe_sample()))
    while sol.time <= T - 1e-8 * sol.dt:
        sol.step()
        if probes is not None:
            movie.append((sol.time, op.probe_sample()))
    return op, x0, sol.getFields(), movie


@pytest.mark.parametrize("shape", ["hex", "tet"])
@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_solver3d_pec_global_1dot5d(shape, alpha):
    """pec_global_1dot5D (Solver3DTest.cpp:57-102): 10 x 1 x 1 cells,
    PEC/PMC walls, order 3, time step 1e-2, final time 2: E_y and H_z end
    as they began (6e-2; the tetrahedra's centred flux reaches 0.054 on the
    host matrix).  Beyond the reference's check: after one round trip the
    E_z pulse is back where it started (1e-2 relative; 0.7% at most on the
    host matrix)."""
    op, x0, x, _ = _solver3d(shape, 10, 1, 1, (1.0, 1.0, 1.0), ("pec", "pmc", "pec", "pmc", "pec", "pec"), 3,
                             alpha, 1e-2, 2.0, (0.5, 0.5, 0.5))
    a, b = x0.reshape(6, op.K, op.Np), x.reshape(6, op.K, op.Np)
    assert np.linalg.norm(a[EY] - b[EY]) <= 6e-2
    assert np.linalg.norm(a[HZ] - b[HZ]) <= 6e-2
    assert np.linalg.norm(a[EZ] - b[EZ]) <= 1e-2 * np.linalg.norm(a[EZ])


def test_solver3d_sma_upwind_hexa_1dot5d():
    """sma_upwind_hexa_1dot5D (Solver3DTest.cpp:155-215): 10 x 2 x 2
    hexahedra on [0,1] x [0,0.4] x [0,0.4], SMA at x = 0 and x = 1, order 3,
    time step 7e-3, final time 1; point probes at x = 0 and x = 1: at t = 0.5
    E_z = 0.5 and H_y = +-0.5 (the halves of the pulse leaving), at t = 1
    both zero; the final state norm near zero (1e-2)."""
    op, x0, x, movie = _solver3d("hex", 10, 2, 2, (1.0, 0.4, 0.4), ("pec", "pmc", "sma", "pmc", "sma", "pec"), 3,
                                 1.0, 7e-3, 1.0, (0.5, 0.2, 0.2), probes=[[0.0, 0.2, 0.2], [1.0, 0.2, 0.2]])
    tol = 1e-2
    assert np.linalg.norm(x) <= tol
    assert movie[-1][0] == pytest.approx(1.0, abs=1e-12)  # the last step is clipped to the final time
    for t, f in movie:  # the reference's windows (no sample falls within 1e-3 of 0.5 at this step)
        if abs(t - 0.5) <= 1e-3:
            assert f[0, EZ] == pytest.approx(0.5, abs=tol) and f[0, HY] == pytest.approx(0.5, abs=tol)
            assert f[1, EZ] == pytest.approx(0.5, abs=tol) and f[1, HY] == pytest.approx(-0.5, abs=tol)
        if abs(t - 1.0) <= 1e-3:
            assert abs(f[0, EZ]) <= tol and abs(f[0, HY]) <= tol
            assert abs(f[1, EZ]) <= tol and abs(f[1, HY]) <= tol
    # the sample nearest t = 0.5 (t = 0.497: the Gaussian peak is 0.003 away, 2e-4 below 0.5)
    t, f = min(movie, key=lambda s: abs(s[0] - 0.5))
    assert abs(t - 0.5) < 7e-3 / 2
    assert f[0, EZ] == pytest.approx(0.5, abs=tol) and f[0, HY] == pytest.approx(0.5, abs=tol)
    assert f[1, EZ] == pytest.approx(0.5, abs=tol) and f[1, HY] == pytest.approx(-0.5, abs=tol)


def test_hexahedral_tfsf_case_matches_oracle(tmp_path):
    """The hexahedral TF/SF case file (interior quadrilateral surface x = 1 as
    the plane-wave tag, automatic delay and time step) run through the case
    reader's operator equals the oracle's RK4 run with the same wave, and
    the scattered-field side stays empty."""
    import json

    from test_case_files import write_gmsh_hexes

    write_gmsh_hexes(str(tmp_path / "g.msh"), 16, 1, 1, sx=4.0, sy=0.5, sz=0.5, x_planes=[(1.0, 7)])
    case = {"solver_options": {"time_step": 0.0, "final_time": 2.2, "order": 3},
            "model": {"filename": "g.msh", "materials": [{"tags": [1], "type": "vacuum"}],
                      "boundaries": [{"tags": [1, 6], "type": "PEC"}, {"tags": [2, 4], "type": "PMC"},
                                     {"tags": [3, 5], "type": "SMA"}]},
            "sources": [{"type": "planewave", "polarization": [0.0, 0.0, 1.0], "propagation": [1.0, 0.0, 0.0],
                         "tags": [7], "magnitude": {"type": "gaussian", "spread": 0.3}}]}
    (tmp_path / "t.json").write_text(json.dumps(case))
    c = D.Case(str(tmp_path / "t.json"))
    (w,) = c.planewaves
    op = c.operator(precision=64, integrator="rk4")
    try:
        assert op.kernel_variant() == "valu"  # the matrix-free hexahedron stage
        dt = c.time_step(op)
        op.set_state(np.zeros(c.ndofs))
        steps = op.run(dt, 2.2)
        x = op.get_state()
    finally:
        op.close()
    ref = OH.OperatorHex(c.vx, c.EToV, c.face_bc, 3)
    ref.set_planewaves([dict(polarization=(0, 0, 1), propagation=(1, 0, 0), spread=0.3, mean=w.mean)])
    xr, steps_ref, _ = ref.run(np.zeros(ref.ndofs()), dt, 2.2, "rk4")
    assert steps == steps_ref
    assert rel(x, xr) <= 1e-11
    q = x.reshape(6, c.K, c.Np)
    total = ref.pos[..., 0].mean(axis=1) > 1.0
    assert np.abs(q[:, total]).max() > 0.1 and np.abs(q[:, ~total]).max() < 1e-12
