"""Quadrilaterals on the GPU: the assembled operator (DGTD_KERNEL_CSR), the
operator the reference runs on quadrilateral meshes (its Global operator,
GlobalEvolution::Mult, GlobalEvolution.cpp:628-1088).

Bars: FP64 Mult equals the quadrilateral oracle (oracle/oracle2d.py,
Operator2D(quads=True), pinned in tests/test_oracle_quad.py) to 1e-12
relative, TF/SF plane waves included; FP64 RK4 / LSERK4 runs equal the
oracle's runs to 1e-11 relative L2; FP32 agrees with FP64 to 1e-5; the
reference's Solver2DTest quadrilateral cases (Solver2DTest.cpp:143-246) meet
their own tolerances.  The matrix-free hesthaven kernels refuse
quadrilaterals (the reference disables its hesthaven quadrilateral case,
ExtensiveCasesTest.cpp:1159).
"""
import numpy as np
import pytest

import meshgen as MG
from oracle import oracle2d as O2

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402

EX, EY, EZ, HX, HY, HZ = range(6)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def quad_case(order, seed):
    """Sheared parallelograms of unequal sizes with every face code: boundary
    PEC/PMC/SMA, interior PEC/PMC/SMA and a TF/SF edge."""
    vx, E, bc = MG.quad_mesh(4, 3, Lx=2.0, bc=["pec", "pmc", "sma", "pec"], stretch=0.3, shear=0.3, seed=seed)
    bc = bc.copy()
    bc[5, 1], bc[6, 3] = 4, 4
    bc[1, 2], bc[5, 0] = 5, 5
    bc[6, 2], bc[10, 0] = 6, 6
    bc[2, 1], bc[3, 3] = 7, 7
    rng = np.random.default_rng(seed)
    return vx, E, bc, rng.uniform(1, 4, E.shape[0]), rng.uniform(1, 2, E.shape[0])


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_quad_mult_matches_oracle(order):
    vx, E, bc, eps, mu = quad_case(order, order)
    for alpha in (1.0, 0.0):
        ref = O2.Operator2D(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, global_sma=True, quads=True)
        x = np.random.default_rng(order).standard_normal(ref.ndofs())
        yr = ref.mult(x)
        g64 = D.GlobalEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=64, quads=True)
        assert g64.kernel_variant() == "csr" and g64.Np == (order + 1) ** 2
        assert np.abs(g64.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
        g32 = D.GlobalEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=32, quads=True)
        assert rel(g32.mult(x), yr) <= 1e-5
        g64.close()
        g32.close()


def test_quad_kernel_choice_and_node_positions():
    vx, E, bc = MG.quad_mesh(3, 2, bc="pec", stretch=0.2, shear=0.2)
    op = D.HesthavenEvolution(vx, E, bc, 3, quads=True)  # kernel "auto" takes the assembled operator
    assert op.kernel_variant() == "csr"
    ref = O2.Operator2D(vx, E, bc, 3, quads=True)
    pos = op.node_positions()
    np.testing.assert_allclose(pos[0], ref.pos[..., 0], atol=1e-14)
    np.testing.assert_allclose(pos[1], ref.pos[..., 1], atol=1e-14)
    assert op.max_fscale() == pytest.approx(ref.max_fscale(), rel=1e-14)
    for kernel in ("valu", "mfma"):
        with pytest.raises(D.DgtdError, match="assembled operator") as ei:
            D.HesthavenEvolution(vx, E, bc, 3, kernel=kernel, quads=True)
        assert ei.value.code == -4


def _quad_waveguide(nx=12, ny=3, length=4.0, x_tfsf=1.0):
    """[0, length] x [0, 1]: SMA ends, PEC walls, TF/SF line x = x_tfsf."""
    vx, E, bc = MG.quad_mesh(nx, ny, Lx=length, bc=["sma", "sma", "pec", "pec"])
    bc = bc.copy()
    for e in range(E.shape[0]):
        for f in range(4):
            a, b = vx[E[e, f]], vx[E[e, (f + 1) % 4]]
            if abs(a[0] - x_tfsf) < 1e-12 and abs(b[0] - x_tfsf) < 1e-12:
                bc[e, f] = 7
    assert (bc == 7).sum() == 2 * ny
    return vx, E, bc


@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
def test_quad_tfsf_run_matches_oracle(integrator):
    vx, E, bc = _quad_waveguide()
    wave = dict(polarization=(0, 1, 0), propagation=(1, 0, 0), spread=0.3, mean=-1.0)
    ref = O2.Operator2D(vx, E, bc, 3, alpha=1.0, global_sma=True, quads=True)
    ref.set_planewaves([wave])
    gpu = D.GlobalEvolution(vx, E, bc, 3, precision=64, integrator=integrator, quads=True)
    gpu.set_planewaves([D.planewave((0, 1, 0), (1, 0, 0), 0.3, -1.0)])
    x = np.random.default_rng(2).standard_normal(ref.ndofs())
    gpu.SetTime(0.9)
    yr = ref.mult(x, 0.9)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
    dt = 0.4 / (ref.max_fscale() * 9)
    xr, steps, _ = ref.run(np.zeros(ref.ndofs()), dt, 1.2, integrator)
    gpu.set_state(np.zeros(ref.ndofs()))
    assert gpu.run(dt, 1.2) == steps
    assert rel(gpu.get_state(), xr) <= 1e-11
    # the total-field side carries the incident wave
    q = gpu.get_state().reshape(6, gpu.K, gpu.Np)
    Ei, _ = ref.incident(ref.pos, gpu.GetTime())
    total = ref.pos[..., 0].mean(axis=1) > 1.0
    assert rel(q[EY][total], Ei[1][total]) < 2e-2


def _solver2d(nx, ny, shape, bcs, order, alpha, time_step=0.0, final_time=2.0):
    """Solver2DTest's set-up (Solver2DTest.cpp:18-69): MakeCartesian2D mesh
    with boundary attributes bottom, right, top, left; Gaussian E_z of spread
    0.1 varying along x, centred at x = 0.5 (SourceFixtures.h:11-26); field
    probes E_z and H_y at (0, 0.5) and (1, 0.5) sampled every step; the
    Global operator; time_step 0 takes the automatic estimate
    (Solver.cpp:302-353)."""
    vx, E, bc = D.cartesian_2d(nx, ny, shape, bc=bcs)
    quads = shape == "quad"
    op = D.GlobalEvolution(vx, E, bc, order, alpha=alpha, precision=64, integrator="rk4", quads=quads)
    x0 = D.project_initial_fields(vx, E, order, [D.gaussian("E", (0.5, 0.5, 0.0), 0.1, (0, 0, 1), 1)],
                                  quads=quads)
    s = D.Solver(op, x0, final_time=final_time, time_step=time_step, evolution_operator="global")
    op.probe_set([[0.0, 0.5, 0.0], [1.0, 0.5, 0.0]])
    movie = [op.probe_sample()]
    while s.time <= final_time - 1e-8 * s.dt:
        s.step()
        movie.append(op.probe_sample())
    return op, x0, s.getFields(), np.array(movie), s.dt


def _expect_1dot5d(x0, x, movie, K, Np, tol=1e-2):
    """expectFieldsAreNearAfterEvolution_1dot5D (Solver2DTest.cpp:73-91)."""
    a, b = x0.reshape(6, K, Np), x.reshape(6, K, Np)
    assert np.abs(a[EY] - b[EY]).max() <= tol and np.abs(a[HZ] - b[HZ]).max() <= tol
    assert abs(movie[:, 0, EZ].max()) == pytest.approx(0.0, abs=tol)
    assert abs(movie[:, 1, EZ].max()) == pytest.approx(0.0, abs=tol)
    assert abs(movie[:, 0, HY].max()) == pytest.approx(1.0, abs=tol)
    assert abs(movie[:, 1, HY].max()) == pytest.approx(1.0, abs=tol)


PMC_PEC = ("pmc", "pec", "pmc", "pec")  # bottom, right, top, left


@pytest.mark.parametrize("shape", ["quad", "triangle"])
def test_solver2d_pec_centered_1dot5d(shape):
    """pec_centered_quads_1dot5D / pec_centered_tris_1dot5D
    (Solver2DTest.cpp:125-156): 14 x 1 cells, order 3, centred flux,
    automatic time step, final time 2."""
    op, x0, x, movie, dt = _solver2d(14, 1, shape, PMC_PEC, 3, 0.0)
    assert dt == pytest.approx(D.mesh_time_step(*D.cartesian_2d(14, 1, shape)[:2], 3, "global"), rel=1e-14)
    _expect_1dot5d(x0, x, movie, op.K, op.Np)


@pytest.mark.parametrize("shape,nx,ny,dt", [("quad", 5, 5, 5e-3), ("triangle", 5, 3, 0.0)])
def test_solver2d_pec_upwind_1dot5d(shape, nx, ny, dt):
    """pec_upwind_quads_1dot5D (time step 5e-3) / pec_upwind_tris_1dot5D
    (Solver2DTest.cpp:160-189): order 4, upwind flux."""
    op, x0, x, movie, _ = _solver2d(nx, ny, shape, PMC_PEC, 4, 1.0, time_step=dt)
    _expect_1dot5d(x0, x, movie, op.K, op.Np)


@pytest.mark.parametrize("shape,n,dt", [("quad", 6, 3e-3), ("triangle", 5, 0.0)])
def test_solver2d_sma_upwind_1dot5d(shape, n, dt):
    """sma_upwind_quads_1dot5D (6 x 6, time step 3e-3) / sma_upwind_tris_1dot5D
    (5 x 3) (Solver2DTest.cpp:191-246): SMA left and right, final time 1: the
    pulse leaves; the E_z degrees of freedom are tol close to zero (l2
    distance, GridFunction::DistanceTo), E_z and H_y at x = 0 never dip
    below -tol and H_y at x = 1 never rises above tol; tol is 1e-3 on
    quadrilaterals and 1e-2 on triangles, as in the reference."""
    ny = n if shape == "quad" else 3
    op, x0, x, movie, _ = _solver2d(n, ny, shape, ("pmc", "sma", "pmc", "sma"), 3, 1.0, time_step=dt,
                                    final_time=1.0)
    assert np.linalg.norm(x0.reshape(6, op.K, op.Np)[EZ]) > 1e-2
    tol = 1e-3 if shape == "quad" else 1e-2
    assert np.linalg.norm(x.reshape(6, op.K, op.Np)[EZ]) <= tol
    assert abs(movie[:, 0, EZ].min()) <= tol and abs(movie[:, 1, EZ].min()) <= tol
    assert abs(movie[:, 0, HY].min()) <= tol and abs(movie[:, 1, HY].max()) <= tol
