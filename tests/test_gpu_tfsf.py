"""TF/SF plane-wave injection on the GPU against the oracle: the PEC/PMC
waveguide with SMA ends and a TF/SF plane (tests/test_tfsf.py), FP64 Mult at
a given time and RK4 / LSERK4 steps to 1e-12 relative, FP32 against FP64 to
1e-5, and the end-to-end physics (incident wave on the total-field side, rest
on the scattered-field side)."""
import numpy as np
import pytest

import dgtd_amd as D
from oracle import oracle as O
from test_tfsf import MEAN, SPREAD, incident, waveguide

pytestmark = pytest.mark.gpu

WAVE = dict(pol=(0, 0, 1), dir=(1, 0, 0), spread=SPREAD, mean=MEAN)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make(order, precision=64, integrator="lserk4", kernel="valu", n=3):
    vx, E, fbc = waveguide(n)
    ref = O.Operator(vx, E, fbc, order)
    ref.set_planewaves([WAVE])
    gpu = D.HesthavenEvolution(vx, E, fbc, order, precision=precision, integrator=integrator, kernel=kernel)
    gpu.set_planewaves([D.planewave((0, 0, 1), (1, 0, 0), SPREAD, MEAN)])
    return ref, gpu


@pytest.mark.parametrize("kernel", ["valu", "mfma"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_mult_at_time_matches_oracle(order, kernel):
    ref, gpu = make(order, kernel=kernel)
    x = np.random.default_rng(order).standard_normal(ref.ndofs)
    for t in (0.0, 0.9, 1.4):
        gpu.SetTime(t)
        yr = ref.mult(x, t)
        assert np.abs(gpu.mult(x) - yr).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
def test_steps_match_oracle(integrator):
    ref, gpu = make(3, integrator=integrator, kernel="mfma")
    x0 = np.zeros(ref.ndofs)
    dt = 0.5 / (ref.max_fscale() * 9)
    xr = x0.copy()
    steps, t = ref.run(xr, dt, 1.2, integrator)
    gpu.set_state(x0)
    assert gpu.run(dt, 1.2) == steps
    x64 = gpu.get_state()
    assert rel(x64, xr) <= 1e-11
    _, g32 = make(3, precision=32, integrator=integrator, kernel="mfma")
    g32.set_state(x0)
    g32.run(dt, 1.2)
    assert rel(g32.get_state(), x64) <= 1e-5


def test_plane_wave_physics_fp32():
    vx, E, fbc = waveguide(6)
    gpu = D.HesthavenEvolution(vx, E, fbc, 3, precision=32)
    gpu.set_planewaves([D.planewave((0, 0, 1), (1, 0, 0), SPREAD, MEAN)])
    gpu.set_state(np.zeros(gpu.ndofs))
    gpu.run(0.5 / (gpu.max_fscale() * 9), 1.8)
    pos = gpu.node_positions().reshape(3, -1)
    f = gpu.get_state().reshape(6, -1)
    g = incident(pos[0], 1.8)
    tf, sf = pos[0] > 0.85, pos[0] < 0.5
    assert np.linalg.norm(f[2][tf] - g[tf]) / np.linalg.norm(g[tf]) < 0.03
    assert np.linalg.norm(f[4][tf] + g[tf]) / np.linalg.norm(g[tf]) < 0.03
    assert np.abs(f[:, sf]).max() < 0.02


def test_no_planewaves_is_plain_interior():
    vx, E, fbc = waveguide(3)
    plain = fbc.copy()
    plain[plain == 7] = 0
    a = D.HesthavenEvolution(vx, E, fbc, 2, kernel="valu")
    b = D.HesthavenEvolution(vx, E, plain, 2, kernel="valu")
    x = np.random.default_rng(2).standard_normal(a.ndofs)
    np.testing.assert_array_equal(a.mult(x), b.mult(x))
    a.set_planewaves([D.planewave((0, 0, 1), (1, 0, 0), 0.3)])
    a.set_planewaves([])
    np.testing.assert_array_equal(a.mult(x), b.mult(x))
