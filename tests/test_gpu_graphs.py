"""Time steps replayed from a captured HIP graph give the same state, bit for
bit, as launching every stage kernel directly: several replays in a row, with
the program's own work (state reads, reductions, a probe) and synchronisation
between them, odd step counts (LSERK buffer parity) and a change of dt (a new
capture)."""
import numpy as np
import pytest

import meshgen as G

pytestmark = pytest.mark.gpu

import dgtd_amd as D  # noqa: E402


def _pair(make):
    a, b = make(), make()
    a.set_graphs(True)
    b.set_graphs(False)
    return a, b


@pytest.mark.parametrize("precision,integrator,kernel", [(32, "lserk4", "mfma"), (64, "rk4", "valu"),
                                                         (64, "lserk4", "mfma")])
def test_graph_steps_equal_direct_launches_tets(precision, integrator, kernel):
    vx, E, fbc = D.cube_mesh(4, 1.0, "pec")
    x0 = np.random.default_rng(0).standard_normal(6 * D.num_nodes(3) * E.shape[0])
    a, b = _pair(lambda: D.HesthavenEvolution(vx, E, fbc, 3, precision=precision, integrator=integrator,
                                              kernel=kernel))
    for op in (a, b):
        op.set_state(x0)
    dt = 0.5 / (a.max_fscale() * 9)
    for rep, (n, h) in enumerate([(3, dt), (1, dt), (4, dt), (2, 0.7 * dt), (3, 0.7 * dt)]):
        a.advance(h, n)
        b.advance(h, n)
        xa, xb = a.get_state(), b.get_state()
        assert np.array_equal(xa, xb), rep
        assert a.norms()["energy"] == b.norms()["energy"]
        assert a.GetTime() == b.GetTime()


def test_graph_steps_equal_direct_launches_triangles():
    vx, E, bc = G.square_mesh(6, bc=["pec", "pmc", "sma", "pec"], jitter=0.2, seed=2)
    x0 = np.random.default_rng(1).standard_normal(6 * 15 * E.shape[0])
    a, b = _pair(lambda: D.HesthavenEvolution(vx, E, bc, 4, precision=64, integrator="lserk4"))
    for op in (a, b):
        op.set_state(x0)
        op.probe_set(np.array([[0.3, 0.4, 0.0]]))
    dt = 0.5 / (a.max_fscale() * 16)
    for rep in range(4):
        a.advance(dt, 2 * rep + 1)
        b.advance(dt, 2 * rep + 1)
        assert np.array_equal(a.get_state(), b.get_state()), rep
        assert np.array_equal(a.probe_sample(), b.probe_sample())
    # run() with the last step truncated matches too
    na, nb = a.run(dt, a.GetTime() + 5.5 * dt), b.run(dt, b.GetTime() + 5.5 * dt)
    assert na == nb == 6 and np.array_equal(a.get_state(), b.get_state())
