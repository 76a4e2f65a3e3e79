"""The assembled ("csr") operator on the GPU, through the C ABI: the
reference's Global path (GlobalEvolution::Mult, GlobalEvolution.cpp:628-1088)
as one fused CSR SpMV + stage-update kernel.

Bars: FP64 Mult equals the oracle's Mult (no conductivity; SMA faces as the
Global operator weights them, see tests/test_global_assembly.py) and the host CSR
product (with conductivity) to 1e-12 relative; FP64 RK4 / LSERK4 runs equal
the oracle's runs to 1e-11 relative L2 (TF/SF plane waves included); FP32
agrees with FP64 to 1e-5; a conductive cavity mode decays as the damped
oscillator a'' + (sigma/eps) a' + w^2 a = 0 predicts.
"""
import numpy as np
import pytest

import meshgen as MG
from oracle import oracle as O
from oracle import oracle2d as O2
from test_global_assembly import csr_matvec, global_expected
from test_gpu_tfsf import WAVE
from test_tfsf import MEAN, SPREAD, waveguide

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cube_case(n, seed):
    rng = np.random.default_rng(seed)
    vx, E, fbc = MG.cube_mesh(n, 1.0, ["pec", "pmc", "sma", "pec", "pmc", "sma"])
    K = E.shape[0]
    return vx, E, fbc, rng.uniform(1, 4, K), rng.uniform(1, 2, K)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_tet_mult_matches_oracle(order):
    vx, E, fbc, eps, mu = cube_case(2, order)
    gpu = D.GlobalEvolution(vx, E, fbc, order, alpha=0.5, eps=eps, mu=mu, precision=64)
    assert gpu.kernel_variant() == "csr"
    x = np.random.default_rng(order).standard_normal(gpu.ndofs)
    yr = global_expected(lambda b, a: O.Operator(vx, E, b, order, alpha=a, eps=eps, mu=mu), fbc, x, 0.5)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("alpha", [0.0, 0.5])
def test_centred_and_partial_upwind_sma_follow_the_global_operator(alpha):
    """SMA faces weighted as at alpha = 1 (centred SMA runs, as in the
    reference's Global operator; the hesthaven operator refuses it)."""
    vx, E, fbc, eps, mu = cube_case(2, 11)
    gpu = D.GlobalEvolution(vx, E, fbc, 3, alpha=alpha, eps=eps, mu=mu, precision=64)
    x = np.random.default_rng(5).standard_normal(gpu.ndofs)
    yr = global_expected(lambda b, a: O.Operator(vx, E, b, 3, alpha=a, eps=eps, mu=mu), fbc, x, alpha)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
    if alpha == 0.0:  # the matrix-free hesthaven operator refuses centred SMA
        with pytest.raises(D.DgtdError, match="SMA"):
            D.HesthavenEvolution(vx, E, fbc, 3, alpha=0.0, precision=64)


@pytest.mark.parametrize("order", [2, 4])
def test_tri_mult_matches_oracle(order):
    vx, E, bc = MG.square_mesh(5, bc=["pec", "pmc", "sma", "pec"], jitter=0.2, seed=order)
    ref = O2.Operator2D(vx, E, bc, order, alpha=1.0)
    gpu = D.GlobalEvolution(vx, E, bc, order, precision=64)
    x = np.random.default_rng(order).standard_normal(ref.ndofs())
    yr = ref.mult(x)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("dim", [3, 2])
def test_conductive_mult_matches_host_matrix(dim):
    if dim == 3:
        vx, E, fbc, eps, mu = cube_case(2, 7)
        order = 3
    else:
        vx, E, fbc = MG.square_mesh(4, bc="sma")
        order = 4
        eps = mu = None
    sigma = np.random.default_rng(3).uniform(0.0, 5.0, E.shape[0])
    A = D.global_assemble(vx, E, fbc, order, eps=eps, mu=mu, sigma=sigma)
    x = np.random.default_rng(4).standard_normal(A[0].size - 1)
    yr = csr_matvec(A, x)
    g64 = D.GlobalEvolution(vx, E, fbc, order, eps=eps, mu=mu, sigma=sigma, precision=64)
    assert np.abs(g64.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
    g32 = D.GlobalEvolution(vx, E, fbc, order, eps=eps, mu=mu, sigma=sigma, precision=32)
    assert rel(g32.mult(x), yr) <= 1e-5


@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
def test_runs_match_oracle(integrator):
    vx, E, fbc, eps, mu = cube_case(3, 1)
    ref = O.Operator(vx, E, fbc, 3, eps=eps, mu=mu)
    pos = ref.positions()
    x0 = MG.gaussian(pos, (0.4, 0.5, 0.6), 0.15)
    dt = 0.5 / (ref.max_fscale() * 9)
    xr = x0.copy()
    steps, _ = ref.run(xr, dt, 0.6, integrator)
    out = {}
    for prec in (64, 32):
        g = D.GlobalEvolution(vx, E, fbc, 3, eps=eps, mu=mu, precision=prec, integrator=integrator)
        g.set_state(x0)
        assert g.run(dt, 0.6) == steps
        out[prec] = g.get_state()
    assert rel(out[64], xr) <= 1e-11
    assert rel(out[32], out[64]) <= 1e-5


@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
def test_tfsf_run_matches_oracle(integrator):
    vx, E, fbc = waveguide(3)
    ref = O.Operator(vx, E, fbc, 2)
    ref.set_planewaves([WAVE])
    gpu = D.GlobalEvolution(vx, E, fbc, 2, precision=64, integrator=integrator)
    gpu.set_planewaves([D.planewave((0, 0, 1), (1, 0, 0), SPREAD, MEAN)])
    x = np.random.default_rng(2).standard_normal(ref.ndofs)
    gpu.SetTime(0.9)
    yr = ref.mult(x, 0.9)
    assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()
    dt = 0.5 / (ref.max_fscale() * 4)
    xr = np.zeros(ref.ndofs)
    steps, _ = ref.run(xr, dt, 1.0, integrator)
    gpu.set_state(np.zeros(ref.ndofs))
    assert gpu.run(dt, 1.0) == steps
    assert rel(gpu.get_state(), xr) <= 1e-11


def damped(t, w, sigma):
    """a(t) of a'' + sigma a' + w^2 a = 0, a(0) = 1, a'(0) = -sigma (eps = 1,
    H = 0 at t = 0)."""
    g = 0.5 * sigma
    wd = np.sqrt(w * w - g * g)
    return np.exp(-g * t) * (np.cos(wd * t) - (g / wd) * np.sin(wd * t))


@pytest.mark.parametrize("dim", [3, 2])
def test_conductive_cavity_mode_decays_as_damped_oscillator(dim):
    sigma = 0.8
    if dim == 3:
        vx, E, fbc = MG.cube_mesh(4, 1.0, "pec")
        pos = O.Operator(vx, E, fbc, 3).positions()
        x0 = MG.te110(pos)
        ez0 = x0.reshape(6, -1)[2].copy()
    else:
        vx, E, fbc = MG.square_mesh(6, bc="pec")
        pos = O2.Operator2D(vx, E, fbc, 3).pos
        x0 = MG.tm_mode(pos, 0.0).reshape(-1)
        ez0 = x0.reshape(6, -1)[2].copy()
    g = D.GlobalEvolution(vx, E, fbc, 3, sigma=np.full(E.shape[0], sigma), precision=64, integrator="rk4")
    g.set_state(x0)
    dt = 0.5 / (g.max_fscale() * 9)
    T = 1.0
    g.run(dt, T)
    x, t = g.get_state(), g.GetTime()
    ez = x.reshape(6, -1)[2]
    a = damped(t, np.pi * np.sqrt(2.0), sigma)
    assert rel(ez, a * ez0) < 5e-3
    # the undamped mode would be far off: conductivity is acting
    assert rel(ez, np.cos(np.pi * np.sqrt(2.0) * t) * ez0) > 0.1
