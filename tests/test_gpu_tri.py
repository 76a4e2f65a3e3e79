"""Parity of the HIP triangle (2D) operator, through the C ABI, with the 2D
oracle (oracle/oracle2d.py).

Both stage variants (per-node VALU and matrix-core MFMA, 8-triangle wave
tasks).  Bars as for tetrahedra: FP64 Mult and RK steps agree with the oracle to
1e-12 relative (summation order only), FP64 runs to 1e-11 relative L2, FP32
agrees with FP64 to 1e-5 relative L2.
"""
import numpy as np
import pytest

import meshgen as G
from oracle import oracle2d as O

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def mixed_case(n, seed=0, media=True, interior=None):
    """Jittered square with PEC / PMC / SMA / PEC sides, random media and,
    optionally, an interior condition on the line x = 1/2."""
    rng = np.random.default_rng(seed)
    vx, E, bc = G.square_mesh(n, bc=["pec", "pmc", "sma", "pec"], jitter=0.2 if interior is None else 0.0,
                              seed=seed)
    if interior is not None:
        for e in range(E.shape[0]):
            for f, (p, q) in enumerate(G.TRI_EDGES):
                if abs(vx[E[e, p], 0] - 0.5) < 1e-12 and abs(vx[E[e, q], 0] - 0.5) < 1e-12:
                    bc[e, f] = interior
    K = E.shape[0]
    eps = rng.uniform(1, 4, K) if media else None
    mu = rng.uniform(1, 2, K) if media else None
    return vx, E, bc, eps, mu


KERNELS = ["valu", "mfma"]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_tri_mult_fp64_matches_oracle(order, alpha, kernel):
    vx, E, bc, eps, mu = mixed_case(4, seed=order)
    ref = O.Operator2D(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu)
    gpu = D.HesthavenEvolution(vx, E, bc, order, alpha=alpha, eps=eps, mu=mu, precision=64, kernel=kernel)
    assert gpu.kernel_variant() == kernel
    assert gpu.Np == ref.Np and gpu.ndofs == ref.ndofs()
    x = np.random.default_rng(7).standard_normal(ref.ndofs())
    y = gpu.mult(x)
    yr = ref.mult(x)
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()
    assert gpu.max_fscale() == pytest.approx(ref.max_fscale(), rel=1e-14)
    xyz = gpu.node_positions()
    np.testing.assert_allclose(xyz[0], ref.pos[..., 0], atol=1e-14)
    np.testing.assert_allclose(xyz[1], ref.pos[..., 1], atol=1e-14)
    assert np.abs(xyz[2]).max() == 0.0


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("code", [4, 5, 6])
def test_tri_interior_conditions_match_oracle(code, kernel):
    vx, E, bc, eps, mu = mixed_case(4, seed=code, interior=code)
    ref = O.Operator2D(vx, E, bc, 3, alpha=1.0, eps=eps, mu=mu)
    gpu = D.HesthavenEvolution(vx, E, bc, 3, alpha=1.0, eps=eps, mu=mu, precision=64, kernel=kernel)
    x = np.random.default_rng(1).standard_normal(ref.ndofs())
    yr = ref.mult(x)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_tri_mult_fp32_close_to_fp64(order, kernel):
    vx, E, bc, eps, mu = mixed_case(6, seed=3)
    x = np.random.default_rng(2).standard_normal(6 * (order + 1) * (order + 2) // 2 * E.shape[0])
    y64 = D.HesthavenEvolution(vx, E, bc, order, eps=eps, mu=mu, precision=64, kernel="valu").mult(x)
    y32 = D.HesthavenEvolution(vx, E, bc, order, eps=eps, mu=mu, precision=32, kernel=kernel).mult(x)
    assert rel(y32, y64) <= 1e-5


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("integrator", ["lserk4", "rk4"])
def test_tri_run_matches_oracle(integrator, kernel):
    vx, E, bc = G.square_mesh(6, bc="pec")
    ref = O.Operator2D(vx, E, bc, 3)
    x0 = G.tm_mode(ref.pos, 0.0).reshape(-1)
    dt = 0.5 / (ref.max_fscale() * 9)
    T = 40.5 * dt  # 41 steps, the last truncated (Solver::run)
    xr, steps, t = ref.run(x0, dt, T, integrator)
    gpu = D.HesthavenEvolution(vx, E, bc, 3, precision=64, integrator=integrator, kernel=kernel)
    gpu.set_state(x0)
    n = gpu.run(dt, T)
    assert n == steps == 41
    x = gpu.get_state()
    assert rel(x, xr) <= 1e-11
    # energy from the device reduction equals the oracle's
    assert gpu.norms()["energy"] == pytest.approx(ref.energy(x), rel=1e-12)
    # and the run follows the analytic TM_11 mode
    assert rel(x, G.tm_mode(ref.pos, t).reshape(-1)) < 2e-3


def test_tri_fp32_run_close_to_fp64():
    vx, E, bc = G.square_mesh(8, bc=["pec", "pmc", "pec", "pmc"])
    ref = O.Operator2D(vx, E, bc, 4)
    r2 = (ref.pos[..., 0] - 0.4) ** 2 + (ref.pos[..., 1] - 0.6) ** 2
    x0 = np.zeros((6,) + r2.shape)
    x0[2] = np.exp(-r2 / (2 * 0.1 ** 2))
    x0 = x0.reshape(-1)
    dt = 0.5 / (ref.max_fscale() * 16)
    out = {}
    for prec in (64, 32):
        g = D.HesthavenEvolution(vx, E, bc, 4, precision=prec)
        g.set_state(x0)
        g.advance(dt, 200)
        out[prec] = g.get_state()
    assert rel(out[32], out[64]) <= 1e-5


def test_tri_probes_match_interpolant():
    vx, E, bc = G.square_mesh(5, bc="pec", jitter=0.2, seed=4)
    gpu = D.HesthavenEvolution(vx, E, bc, 3, precision=64)
    x = np.random.default_rng(5).standard_normal(gpu.ndofs)
    gpu.set_state(x)
    pts = np.array([[0.13, 0.77, 0.0], [0.5, 0.5, 0.0], [0.91, 0.02, 0.0], [1.5, 0.5, 0.0]])
    elem = gpu.probe_set(pts)
    got = gpu.probe_sample()
    e2, w = D.mesh_locate(vx, E, 3, pts)
    np.testing.assert_array_equal(elem, e2)
    assert elem[3] == -1 and (elem[:3] >= 0).all()
    q = x.reshape(6, gpu.K, gpu.Np)
    want = np.einsum("pn,cpn->pc", w[:3], q[:, elem[:3], :])
    np.testing.assert_allclose(got[:3], want, rtol=1e-12, atol=1e-14)


def test_tri_refusals():
    vx, E, bc = G.square_mesh(2, bc="pec")
    bad = bc.copy()
    bad[bad == 0] = 1  # PEC on interior edges
    with pytest.raises(D.DgtdError):
        D.HesthavenEvolution(vx, E, bad, 2)
    tf = bc.copy()
    tf[bc == 1] = 7  # TF/SF on boundary edges
    with pytest.raises(D.DgtdError):
        D.HesthavenEvolution(vx, E, tf, 2)


def _tfsf_case(n=8, order=3, alpha=1.0):
    vx, E, bc = G.square_mesh(n, L=1.0, bc=["sma", "sma", "pec", "pec"], jitter=0.0)
    vx = vx.copy()
    vx[:, 0] *= 4.0
    for e in range(E.shape[0]):
        for f, (p, q) in enumerate(G.TRI_EDGES):
            if abs(vx[E[e, p], 0] - 1.0) < 1e-12 and abs(vx[E[e, q], 0] - 1.0) < 1e-12:
                bc[e, f] = 7
    waves = [dict(polarization=(0, 0.6, 0.8), propagation=(1, 0.2, 0), spread=0.4, mean=-0.5),
             dict(polarization=(0, 1, 0), propagation=(1, 0, 0), spread=0.3, mean=0.2, kind="modulated",
                  frequency=0.8)]
    gw = [D.planewave(w["polarization"], w["propagation"], w["spread"], w["mean"], w.get("frequency"))
          for w in waves]
    return vx, E, bc, waves, gw


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [2, 4])
def test_tri_tfsf_mult_matches_oracle(order, kernel):
    vx, E, bc, waves, gw = _tfsf_case(order=order)
    ref = O.Operator2D(vx, E, bc, order, alpha=0.7)
    ref.set_planewaves(waves)
    gpu = D.HesthavenEvolution(vx, E, bc, order, alpha=0.7, precision=64, kernel=kernel)
    gpu.set_planewaves(gw)
    x = np.random.default_rng(3).standard_normal(ref.ndofs())
    for t in (0.0, 0.8, 1.7):
        gpu.SetTime(t)
        y = np.empty_like(x)
        gpu.Mult(x, y)
        yr = ref.mult(x, t)
        assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("integrator", ["lserk4", "rk4"])
def test_tri_tfsf_run_matches_oracle(integrator):
    vx, E, bc, waves, gw = _tfsf_case(order=3)
    ref = O.Operator2D(vx, E, bc, 3)
    ref.set_planewaves(waves)
    dt = 0.4 / (ref.max_fscale() * 9)
    xr, steps, t = ref.run(np.zeros(ref.ndofs()), dt, 1.0, integrator)
    for prec, tol in ((64, 1e-11), (32, 1e-5)):
        gpu = D.HesthavenEvolution(vx, E, bc, 3, precision=prec, integrator=integrator)
        gpu.set_planewaves(gw)
        gpu.set_state(np.zeros(ref.ndofs()))
        assert gpu.run(dt, 1.0) == steps
        assert rel(gpu.get_state(), xr) <= tol
