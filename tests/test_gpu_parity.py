"""Parity of the HIP operator (through the C ABI) with the CPU FP64 oracle.

Bars: FP64 Mult and RK steps agree with the oracle to 1e-12 relative (only the
summation order differs); the FP64 run of the 8^3 TE110 cavity (3,072 tets,
p=3, 500 LSERK4 steps) agrees to 1e-11 relative L2 at the final time; FP32
agrees with FP64 to 1e-5 relative L2.
"""
import numpy as np
import pytest

import meshgen as MG
from oracle import oracle as O

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def mixed_case(n, order, alpha=1.0, seed=0, media=True, bc=("pec", "pmc", "sma", "pec", "pmc", "sma")):
    rng = np.random.default_rng(seed)
    vx, E, fbc = MG.cube_mesh(n, 1.0, list(bc))
    K = E.shape[0]
    eps = rng.uniform(1, 4, K) if media else None
    mu = rng.uniform(1, 2, K) if media else None
    return vx, E, fbc, eps, mu


KERNELS = ["valu", "mfma"]


def make_gpu(*args, kernel="valu", **kw):
    return D.HesthavenEvolution(*args, kernel=kernel, **kw)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_mult_fp64_matches_oracle(order, alpha, kernel):
    vx, E, fbc, eps, mu = mixed_case(2, order, alpha, seed=order)
    ref = O.Operator(vx, E, fbc, order, alpha=alpha, eps=eps, mu=mu)
    gpu = make_gpu(vx, E, fbc, order, alpha=alpha, eps=eps, mu=mu, precision=64, kernel=kernel)
    x = np.random.default_rng(7).standard_normal(ref.ndofs)
    y = gpu.mult(x)
    yr = ref.mult(x)
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 3])
def test_mult_centred_pec_pmc(order, kernel):
    vx, E, fbc, eps, mu = mixed_case(2, order, bc=("pec", "pmc", "pec", "pmc", "none", "pec"))
    ref = O.Operator(vx, E, fbc, order, alpha=0.0)
    gpu = make_gpu(vx, E, fbc, order, alpha=0.0, precision=64, kernel=kernel)
    x = np.random.default_rng(3).standard_normal(ref.ndofs)
    yr = ref.mult(x)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_mult_fp32_close_to_oracle(order, kernel):
    vx, E, fbc, eps, mu = mixed_case(3, order, seed=11)
    ref = O.Operator(vx, E, fbc, order, eps=eps, mu=mu)
    gpu = make_gpu(vx, E, fbc, order, eps=eps, mu=mu, precision=32, kernel=kernel)
    pos = ref.positions()
    x = MG.gaussian(pos, (0.45, 0.5, 0.55), 0.2, comps=range(6))
    assert rel(gpu.mult(x), ref.mult(x)) <= 1e-5


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
@pytest.mark.parametrize("order", [2, 4])
def test_steps_fp64_match_oracle(integrator, order, kernel):
    vx, E, fbc, eps, mu = mixed_case(2, order, seed=5)
    ref = O.Operator(vx, E, fbc, order, eps=eps, mu=mu)
    gpu = make_gpu(vx, E, fbc, order, eps=eps, mu=mu, precision=64, integrator=integrator, kernel=kernel)
    x0 = np.random.default_rng(1).standard_normal(ref.ndofs)
    dt = 0.25 / (ref.max_fscale() * order ** 2)
    xr = x0.copy()
    res = np.zeros_like(xr)
    for s in range(3):
        if integrator == "rk4":
            ref.rk4_step(xr, s * dt, dt)
        else:
            ref.lserk4_step(xr, res, s * dt, dt)
    gpu.set_state(x0)
    gpu.advance(dt, 3)
    xg = gpu.get_state()
    assert rel(xg, xr) <= 1e-12
    assert abs(gpu.GetTime() - 3 * dt) < 1e-15


@pytest.mark.parametrize("kernel", KERNELS)
def test_te110_cavity_fp64_end_to_end(kernel):
    """8^3 PEC cavity, p=3, TE110, LSERK4, dt=2e-3, T=1 (500 steps)."""
    vx, E, fbc = MG.cube_mesh(8, 1.0, "pec")
    ref = O.Operator(vx, E, fbc, 3)
    gpu = make_gpu(vx, E, fbc, 3, precision=64, integrator="lserk4", kernel=kernel)
    pos = ref.positions()
    q0 = MG.te110(pos)
    xr = q0.copy()
    steps_r, tr = ref.run(xr, 2e-3, 1.0, "lserk4")
    gpu.set_state(q0)
    steps_g = gpu.run(2e-3, 1.0)
    xg = gpu.get_state()
    assert steps_g == steps_r == 500
    assert abs(gpu.GetTime() - tr) < 1e-12
    assert rel(xg, xr) <= 1e-11
    exact = MG.te110_exact(pos, tr)
    assert rel(xg, exact) < 5e-4  # discretisation error at h=1/8, p=3


def test_rk4_run_matches_oracle_with_truncated_last_step():
    vx, E, fbc = MG.cube_mesh(3, 1.0, "pec")
    ref = O.Operator(vx, E, fbc, 2)
    gpu = D.HesthavenEvolution(vx, E, fbc, 2, precision=64, integrator="rk4")
    q0 = MG.te110(ref.positions())
    xr = q0.copy()
    steps_r, tr = ref.run(xr, 7e-3, 0.1, "rk4")  # 0.1/7e-3 is not an integer
    gpu.set_state(q0)
    assert gpu.run(7e-3, 0.1) == steps_r
    assert abs(gpu.GetTime() - tr) < 1e-15
    assert rel(gpu.get_state(), xr) <= 1e-12


def test_norms_and_energy():
    vx, E, fbc, eps, mu = mixed_case(3, 3, seed=2)
    ref = O.Operator(vx, E, fbc, 3, eps=eps, mu=mu)
    for prec in (64, 32):
        gpu = D.HesthavenEvolution(vx, E, fbc, 3, eps=eps, mu=mu, precision=prec)
        x = MG.gaussian(ref.positions(), (0.5, 0.5, 0.5), 0.2, comps=range(6))
        gpu.set_state(x)
        nm = gpu.norms()
        tol = 1e-13 if prec == 64 else 1e-6
        assert abs(nm["norm2"] - np.linalg.norm(x)) <= tol * np.linalg.norm(x)
        assert abs(nm["energy"] - ref.energy(x)) <= tol * ref.energy(x)
        assert nm["nonfinite"] == 0
        x[17] = np.nan
        gpu.set_state(x)
        assert gpu.norms()["nonfinite"] == 1


@pytest.mark.parametrize("kernel", KERNELS)
def test_fp32_cavity_tracks_fp64(kernel):
    vx, E, fbc = MG.cube_mesh(6, 1.0, "pec")
    g64 = D.HesthavenEvolution(vx, E, fbc, 4, precision=64, kernel="valu")
    g32 = make_gpu(vx, E, fbc, 4, precision=32, kernel=kernel)
    pos = g64.node_positions()
    q0 = MG.gaussian(pos, (0.45, 0.5, 0.55), 0.1, comps=range(6))
    dt = D.estimate_time_step(g64)
    for g in (g64, g32):
        g.set_state(q0)
        g.run(dt, 0.3)
    assert rel(g32.get_state(), g64.get_state()) <= 1e-5


def test_partial_last_task_and_odd_sizes():
    """K not a multiple of 16 (or of the VALU block): 5 hexes per edge -> 750 tets."""
    vx, E, fbc, eps, mu = mixed_case(5, 3, seed=9)
    ref = O.Operator(vx, E, fbc, 3, eps=eps, mu=mu)
    x = np.random.default_rng(4).standard_normal(ref.ndofs)
    yr = ref.mult(x)
    for kernel in KERNELS:
        for prec, tol in ((64, 1e-12), (32, 1e-5)):
            g = D.HesthavenEvolution(vx, E, fbc, 3, eps=eps, mu=mu, precision=prec, kernel=kernel)
            assert rel(g.mult(x), yr) <= tol, (kernel, prec)


def test_autotune_picks_a_variant():
    vx, E, fbc = MG.cube_mesh(4, 1.0, "pec")
    g = D.HesthavenEvolution(vx, E, fbc, 4, precision=32, kernel="auto")
    assert g.kernel_variant() in ("valu", "mfma")
    t = g.tuning()
    assert t["variant"] == g.kernel_variant() and t["candidates_us"]
    g2 = D.HesthavenEvolution(vx, E, fbc, 4, precision=32, kernel="auto")
    assert g2.tuning()["from_cache"] and g2.kernel_variant() == g.kernel_variant()


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_every_candidate_same_bits(order, precision):
    """Every matrix-core candidate of the autotuner (configuration x task
    order) stores the same bits after LSERK4 steps (the same per-node
    arithmetic in the same order); the per-node candidate agrees to the
    precision's bar, and FP64 to the oracle at 1e-12.  750 tetrahedra: the
    last 8-element task is partial."""
    vx, E, fbc, eps, mu = mixed_case(5, order, seed=20 + order)
    ref = O.Operator(vx, E, fbc, order, eps=eps, mu=mu)
    g = D.HesthavenEvolution(vx, E, fbc, order, eps=eps, mu=mu, precision=precision, integrator="lserk4",
                             kernel="mfma")
    x0 = np.random.default_rng(order).standard_normal(ref.ndofs)
    dt = 0.5 / (ref.max_fscale() * order ** 2)
    out = {}
    for c, name in enumerate(D.CANDIDATES):
        try:
            g.set_candidate(c)
        except D.DgtdError:
            continue
        g.set_state(x0)
        g.advance(dt, 4)
        out[name] = g.get_state()
    mf = [n for n in out if n.startswith("mfma")]
    assert "valu" in out and "mfma_c0_morton" in out and "mfma_c0_element" in out
    for n in mf:
        np.testing.assert_array_equal(out[n], out["mfma_c0_morton"], err_msg=n)
    xr = x0.copy()
    res = np.zeros_like(xr)
    for s in range(4):
        ref.lserk4_step(xr, res, s * dt, dt)
    tol = 1e-12 if precision == 64 else 1e-5
    for n in out:
        assert rel(out[n], xr) <= tol, n


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_interior_conditions_match_oracle(order, kernel):
    """Interior PEC / PMC / SMA planes (x, y, z = 0.5) with media, FP64 Mult and
    three LSERK4 steps against the oracle; FP32 against FP64."""
    vx, E, fbc, eps, mu = mixed_case(2, order, seed=20 + order)
    fbc = fbc.copy()
    for axis, code in ((0, 4), (1, 5), (2, 6)):
        fbc[MG.plane_faces(vx, E, axis, 0.5)] = code
    ref = O.Operator(vx, E, fbc, order, eps=eps, mu=mu)
    gpu = make_gpu(vx, E, fbc, order, eps=eps, mu=mu, precision=64, kernel=kernel)
    x = np.random.default_rng(order).standard_normal(ref.ndofs)
    yr = ref.mult(x)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
    pos = ref.positions()
    q0 = MG.gaussian(pos, (0.4, 0.55, 0.5), 0.15, comps=range(6))
    dt = 0.5 / (ref.max_fscale() * order ** 2)
    xr = q0.copy()
    res = np.zeros_like(xr)
    for s in range(3):
        ref.lserk4_step(xr, res, s * dt, dt)
    gpu.set_state(q0)
    gpu.advance(dt, 3)
    x64 = gpu.get_state()
    assert rel(x64, xr) <= 1e-12
    g32 = make_gpu(vx, E, fbc, order, eps=eps, mu=mu, precision=32, kernel=kernel)
    g32.set_state(q0)
    g32.advance(dt, 3)
    assert rel(g32.get_state(), x64) <= 1e-5
