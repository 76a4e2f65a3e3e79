"""Multi-instance batch on the GPU: uneven sizes and orders in one persistent
launch per stage, hipGraph-replayed steps with host polls in between,
per-instance dt and final time, admission by footprint, continuous batching.

Bars: FP64 instances agree with the oracle run to 1e-11 relative L2 and take
the same number of steps; an FP32 instance's fields are bitwise identical
alone and in a mix (no cross-instance interference); FP32 vs FP64 <= 1e-5.
"""
import numpy as np
import pytest

import dgtd_amd as D
import meshgen as MG
from oracle import oracle as O
from oracle import oracle2d as O2

pytestmark = pytest.mark.gpu

CASES = [  # n, order, final_time, bc
    (2, 1, 0.30, "pec"),
    (3, 2, 0.25, ["pec", "pmc", "sma", "pec", "pmc", "sma"]),
    (2, 3, 0.20, "sma"),
    (3, 4, 0.15, "pec"),
    (2, 5, 0.10, "pmc"),
]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case_input(n, order, bc, seed):
    vx, E, fbc = MG.cube_mesh(n, 1.0, bc)
    ref = O.Operator(vx, E, fbc, order)
    rng = np.random.default_rng(seed)
    q0 = MG.gaussian(ref.positions(), rng.uniform(0.3, 0.7, 3), 0.15, comps=range(6))
    return vx, E, fbc, ref, q0


def test_mixed_orders_fp64_match_oracle():
    b = D.Batch(precision=64, max_instances=16)
    ids, refs = [], []
    for k, (n, order, T, bc) in enumerate(CASES):
        vx, E, fbc, ref, q0 = case_input(n, order, bc, k)
        i = b.add(vx, E, fbc, order, x0=q0, dt=0.0, final_time=T)
        dt = b.info(i)["dt"]
        assert dt == pytest.approx(1.0 / (ref.max_fscale() * order ** 2), rel=1e-14)
        x = q0.copy()
        steps, t = ref.run(x, dt, T, "lserk4")
        ids.append(i)
        refs.append((x, steps, t))
    total = 0
    while True:
        done, running = b.poll()
        if running == 0:
            break
        b.step(3)  # graph replays, eager polls between them
        total += 3
    assert total >= max(r[1] for r in refs)
    for i, (x, steps, t) in zip(ids, refs):
        info = b.info(i)
        assert info["state"] == "done"
        assert info["steps"] == steps
        assert info["t"] == pytest.approx(t, abs=1e-14)
        assert rel(b.get_state(i), x) <= 1e-11


def test_alone_equals_in_mix_fp32_and_tracks_fp64():
    n, order, T, bc = 3, 3, 0.2, "pec"
    vx, E, fbc, ref, q0 = case_input(n, order, bc, 42)
    alone = D.Batch(precision=32)
    a = alone.add(vx, E, fbc, order, x0=q0, final_time=T)
    alone.run(poll_every=4)
    xa = alone.get_state(a)

    mix = D.Batch(precision=32)
    others = []
    for k, (n2, o2, T2, bc2) in enumerate(CASES):
        vx2, E2, fbc2, _, q2 = case_input(n2, o2, bc2, 100 + k)
        others.append(mix.add(vx2, E2, fbc2, o2, x0=q2, final_time=T2))
    m = mix.add(vx, E, fbc, order, x0=q0, final_time=T)
    mix.run(poll_every=5)
    xm = mix.get_state(m)
    np.testing.assert_array_equal(xm, xa)  # same result alone and in the mix

    g64 = D.Batch(precision=64)
    r = g64.add(vx, E, fbc, order, x0=q0, final_time=T)
    g64.run()
    assert rel(xm, g64.get_state(r)) <= 1e-5


def test_continuous_admission_and_removal():
    b = D.Batch(precision=32, max_instances=4)
    vx, E, fbc, ref, q0 = case_input(2, 2, "pec", 7)
    first = b.add(vx, E, fbc, 2, x0=q0, final_time=0.3)
    b.step(5)
    late = b.add(vx, E, fbc, 2, x0=q0, final_time=0.3)  # joins while the first runs
    b.run(poll_every=2)
    assert b.info(first)["steps"] == b.info(late)["steps"]
    np.testing.assert_array_equal(b.get_state(first), b.get_state(late))
    b.remove(first)
    b.remove(late)
    assert b.stats()["bytes_in_use"] == 0


def test_add_refuses_quadrilaterals_and_hexahedra():
    """Tensor-product elements run the Global operator only; the batch
    refuses them at admission and stays empty."""
    b = D.Batch(precision=32, max_instances=4)
    for vx, E, fbc in (D.cartesian_2d(3, 2, "quad"), D.cartesian_3d(2, 2, 1, "hex")):
        with pytest.raises(D.DgtdError, match="Global operator only") as ei:
            b.add(vx, E, fbc, 2, final_time=0.1)
        assert ei.value.code == -4
    assert b.stats()["bytes_in_use"] == 0


def test_new_order_admitted_while_steps_are_in_flight():
    """A request of an order the batch has not run yet (its stage choice is
    tuned at admission) joins while queued steps of another instance are in
    flight: the running instance keeps its progress and ends bitwise as
    alone."""
    vx, E, fbc, ref, q0 = case_input(3, 2, "pec", 11)
    alone = D.Batch(precision=32, max_instances=1)
    a = alone.add(vx, E, fbc, 2, x0=q0, final_time=0.3)
    alone.run(poll_every=3)
    xa, sa = alone.get_state(a), alone.info(a)["steps"]
    alone.close()
    b = D.Batch(precision=32, max_instances=4)
    first = b.add(vx, E, fbc, 2, x0=q0, final_time=0.3)
    b.step(6)  # queued, not waited for
    vx3, E3, fbc3, _, q3 = case_input(2, 3, "sma", 12)
    b.add(vx3, E3, fbc3, 3, x0=q3, final_time=0.2)
    b.run(poll_every=3)
    assert b.info(first)["steps"] == sa
    np.testing.assert_array_equal(b.get_state(first), xa)


def test_admission_by_footprint():
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    fp = D.footprint(vx, E, fbc, 3, precision=32)
    b = D.Batch(precision=32, max_instances=8, hbm_budget=int(2.5 * fp))
    i1 = b.add(vx, E, fbc, 3, final_time=0.1)
    b.add(vx, E, fbc, 3, final_time=0.1)
    with pytest.raises(D.CapacityError):
        b.add(vx, E, fbc, 3, final_time=0.1)
    b.remove(i1)
    b.add(vx, E, fbc, 3, final_time=0.1)
    assert b.stats()["bytes_in_use"] == 2 * fp


def test_unstable_instances_fail_and_the_rest_finish_as_alone():
    """The per-step stability test (Solver::run, Solver.cpp:499-516): a NaN
    in one request's initial fields (order 1: the per-node stage) and an
    over-CFL dt in another (order 3, 10x the Hesthaven rule: the matrix-core
    stage) retire those instances as failed at the step whose stored state
    first holds a non-finite entry or one above 1e20; the other instances
    run to their final time bitwise as alone, and the slots free up."""
    good = [(3, 2, 0.25, "pec"), (2, 4, 0.15, "sma")]
    alone = []
    for k, (n, order, T, bc) in enumerate(good):
        vx, E, fbc, ref, q0 = case_input(n, order, bc, 200 + k)
        b1 = D.Batch(precision=32, max_instances=1)
        i = b1.add(vx, E, fbc, order, x0=q0, final_time=T)
        b1.run(poll_every=3)
        alone.append((b1.get_state(i), b1.info(i)["steps"]))
        b1.close()
    b = D.Batch(precision=32, max_instances=8)
    ids = []
    for k, (n, order, T, bc) in enumerate(good):
        vx, E, fbc, ref, q0 = case_input(n, order, bc, 200 + k)
        ids.append(b.add(vx, E, fbc, order, x0=q0, final_time=T))
    vx, E, fbc, ref, q0 = case_input(2, 1, "pec", 300)
    qn = q0.copy()
    qn[7] = np.nan
    nan_id = b.add(vx, E, fbc, 1, x0=qn, final_time=0.3)
    vx, E, fbc, ref, q0 = case_input(2, 3, "pec", 301)
    dt_rule = 1.0 / (ref.max_fscale() * 9)
    # the oracle's own run crosses 1e20 at step 6 with this dt (the state's largest entry)
    x, res = q0.copy(), np.zeros_like(q0)
    first = None
    for s in range(40):
        ref.lserk4_step(x, res, s * 10 * dt_rule, 10 * dt_rule)
        if not np.abs(x).max() <= 1e20:
            first = s + 1
            break
    assert first is not None
    cfl_id = b.add(vx, E, fbc, 3, x0=q0, dt=10 * dt_rule, final_time=400 * dt_rule)
    ended = {}
    while True:
        done, running = b.poll()
        for i in done:
            ended[i] = b.info(i)
        if running == 0:
            break
        b.step(2)
    assert ended[nan_id]["state"] == "failed" and ended[nan_id]["steps"] == 1
    assert ended[cfl_id]["state"] == "failed"
    assert abs(ended[cfl_id]["steps"] - first) <= 1  # FP32 vs FP64 growth
    for i, (xa, steps) in zip(ids, alone):
        assert ended[i]["state"] == "done" and ended[i]["steps"] == steps
        np.testing.assert_array_equal(b.get_state(i), xa)
    for i in (nan_id, cfl_id):
        b.remove(i)
    assert b.stats()["running"] == 0


def test_cancel_stops_one_instance():
    vx, E, fbc, ref, q0 = case_input(2, 2, "pec", 8)
    b = D.Batch(precision=32, max_instances=4)
    keep = b.add(vx, E, fbc, 2, x0=q0, final_time=0.3)
    stop = b.add(vx, E, fbc, 2, x0=q0, final_time=0.3)
    b.step(3)
    b.cancel(stop)
    done, running = b.poll()
    assert done == [stop] and running == 1
    info = b.info(stop)
    assert info["state"] == "cancelled" and info["steps"] == 3
    x3 = b.get_state(stop)
    b.run(poll_every=2)
    np.testing.assert_array_equal(b.get_state(stop), x3)  # did not move after the cancel
    assert b.info(keep)["state"] == "done" and b.info(keep)["steps"] > 3
    b.cancel(keep)  # already ended: left as it is
    assert b.info(keep)["state"] == "done"


@pytest.mark.parametrize("kernel", ["valu", "mfma"])
def test_operator_stability_flag(kernel):
    vx, E, fbc, ref, q0 = case_input(2, 3, "pec", 9)
    op = D.HesthavenEvolution(vx, E, fbc, 3, precision=32, integrator="lserk4", kernel=kernel)
    dt = 1.0 / (op.max_fscale() * 9)
    op.set_state(q0)
    op.advance(dt, 5)
    assert not op.unstable()
    bad = q0.copy()
    bad[3] = np.inf
    op.set_state(bad)
    op.advance(dt, 1)
    assert op.unstable()
    op.set_state(q0)  # cleared by a new state
    assert not op.unstable()
    op.close()


TRI_CASES = [  # n, order, final_time, bc (x=0, x=L, y=0, y=L)
    (3, 1, 0.30, "pec"),
    (4, 2, 0.25, ["pec", "pmc", "sma", "pec"]),
    (3, 3, 0.20, "sma"),
    (5, 4, 0.15, "pmc"),
    (2, 5, 0.10, "pec"),
]


def tri_case(n, order, bc, seed):
    vx, E, fbc = MG.square_mesh(n, bc=bc, jitter=0.15, seed=seed)
    rng = np.random.default_rng(seed)
    ref = O2.Operator2D(vx, E, fbc, order)
    c = rng.uniform(0.3, 0.7, 2)
    r2 = (ref.pos[..., 0] - c[0]) ** 2 + (ref.pos[..., 1] - c[1]) ** 2
    q0 = np.zeros((6,) + r2.shape)
    for k in range(6):
        q0[k] = rng.uniform(0.5, 1.0) * np.exp(-r2 / (2 * 0.15 ** 2))
    return vx, E, fbc, ref, q0.reshape(-1)


def test_triangles_and_tetrahedra_fp64_match_oracles():
    """Triangle meshes of orders 1-5 (the matrix-core stage, DIM = 2) in one
    batch with tetrahedral meshes: each instance takes the oracle's number of
    steps and ends within 1e-11 of its oracle run (HesthavenEvolution::Mult on
    triangles through the 2D oracle)."""
    b = D.Batch(precision=64, max_instances=16)
    runs = []
    for k, (n, order, T, bc) in enumerate(TRI_CASES):
        vx, E, fbc, ref, q0 = tri_case(n, order, bc, 50 + k)
        i = b.add(vx, E, fbc, order, x0=q0, final_time=T)
        dt = b.info(i)["dt"]
        assert dt == pytest.approx(1.0 / (ref.max_fscale() * order ** 2), rel=1e-14)
        x, steps, t = ref.run(q0, dt, T, "lserk4")
        runs.append((i, x, steps))
    for k, (n, order, T, bc) in enumerate(CASES[1:3]):
        vx, E, fbc, ref, q0 = case_input(n, order, bc, 60 + k)
        i = b.add(vx, E, fbc, order, x0=q0, final_time=T)
        x = q0.copy()
        steps, t = ref.run(x, b.info(i)["dt"], T, "lserk4")
        runs.append((i, x, steps))
    b.run(poll_every=4)
    for i, x, steps in runs:
        info = b.info(i)
        assert info["state"] == "done" and info["steps"] == steps
        assert rel(b.get_state(i), x) <= 1e-11, i


def test_triangle_alone_equals_in_mix_fp32():
    n, order, T, bc = 6, 4, 0.2, "pec"
    vx, E, fbc, ref, q0 = tri_case(n, order, bc, 70)
    alone = D.Batch(precision=32, max_instances=1)
    a = alone.add(vx, E, fbc, order, x0=q0, final_time=T)
    alone.run(poll_every=3)
    xa = alone.get_state(a)
    alone.close()
    mix = D.Batch(precision=32)
    for k, (n2, o2, T2, bc2) in enumerate(TRI_CASES):
        vx2, E2, fbc2, _, q2 = tri_case(n2, o2, bc2, 80 + k)
        mix.add(vx2, E2, fbc2, o2, x0=q2, final_time=T2)
    for k, (n2, o2, T2, bc2) in enumerate(CASES):
        vx2, E2, fbc2, _, q2 = case_input(n2, o2, bc2, 90 + k)
        mix.add(vx2, E2, fbc2, o2, x0=q2, final_time=T2)
    m = mix.add(vx, E, fbc, order, x0=q0, final_time=T)
    mix.run(poll_every=5)
    np.testing.assert_array_equal(mix.get_state(m), xa)
    g64 = D.Batch(precision=64)
    r = g64.add(vx, E, fbc, order, x0=q0, final_time=T)
    g64.run()
    assert rel(xa, g64.get_state(r)) <= 1e-5


def _tfsf_instances():
    """A tetrahedral PEC/PMC waveguide with SMA ends and a TF/SF plane, and a
    triangle strip with a TF/SF line (the single-operator TF/SF cases), each
    with its oracle run over its own dt and final time."""
    from test_gpu_tri import _tfsf_case
    from test_tfsf import MEAN, SPREAD, waveguide

    vx, E, fbc = waveguide(3)
    ref = O.Operator(vx, E, fbc, 3)
    ref.set_planewaves([dict(pol=(0, 0, 1), dir=(1, 0, 0), spread=SPREAD, mean=MEAN)])
    dt = 0.5 / (ref.max_fscale() * 9)
    x = np.zeros(ref.ndofs)
    steps, _ = ref.run(x, dt, 1.2, "lserk4")
    tet = (vx, E, fbc, 3, dt, 1.2, [D.planewave((0, 0, 1), (1, 0, 0), SPREAD, MEAN)], x, steps)
    vx2, E2, bc2, waves, gw = _tfsf_case(order=3)
    ref2 = O2.Operator2D(vx2, E2, bc2, 3)
    ref2.set_planewaves(waves)
    dt2 = 0.4 / (ref2.max_fscale() * 9)
    x2, steps2, _ = ref2.run(np.zeros(ref2.ndofs()), dt2, 1.0, "lserk4")
    tri = (vx2, E2, bc2, 3, dt2, 1.0, gw, x2, steps2)
    return [tet, tri]


def test_tfsf_instances_fp64_match_oracles():
    """Plane waves injected on TF/SF faces of batched instances (tetrahedra
    and triangles, each at its own stage times t + c_s dt), beside plain
    instances: every instance matches its oracle run to 1e-11."""
    b = D.Batch(precision=64, max_instances=8)
    runs = []
    for vx, E, fbc, order, dt, T, gw, x, steps in _tfsf_instances():
        i = b.add(vx, E, fbc, order, x0=np.zeros(x.size), dt=dt, final_time=T, planewaves=gw)
        runs.append((i, x, steps))
    vx, E, fbc, ref, q0 = case_input(3, 2, "sma", 5)
    i = b.add(vx, E, fbc, 2, x0=q0, final_time=0.2)
    x = q0.copy()
    steps, _ = ref.run(x, b.info(i)["dt"], 0.2, "lserk4")
    runs.append((i, x, steps))
    b.run(poll_every=8)
    for i, x, steps in runs:
        info = b.info(i)
        assert info["state"] == "done" and info["steps"] == steps
        assert rel(b.get_state(i), x) <= 1e-11, i


def test_tfsf_instance_alone_equals_in_mix_fp32():
    tet, tri = _tfsf_instances()
    out = []
    for mixed in (False, True):
        b = D.Batch(precision=32, max_instances=16)
        if mixed:
            for k, (n, order, T, bc) in enumerate(CASES):
                vx, E, fbc, _, q0 = case_input(n, order, bc, 110 + k)
                b.add(vx, E, fbc, order, x0=q0, final_time=T)
        ids = [b.add(vx, E, fbc, order, x0=np.zeros(x.size), dt=dt, final_time=T, planewaves=gw)
               for vx, E, fbc, order, dt, T, gw, x, steps in (tet, tri)]
        b.run(poll_every=6)
        out.append([b.get_state(i) for i in ids])
        b.close()
    for a, m, inst in zip(out[0], out[1], (tet, tri)):
        np.testing.assert_array_equal(a, m)
        assert rel(a, inst[7]) <= 1e-5
