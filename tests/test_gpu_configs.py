"""The batched-instances and serving configurations at their stated sizes.

Batch: the 64 seeded PEC cavities (n log-uniform over {2..24} hexahedra x 6,
order uniform over {1..5}, up to 63,888 tetrahedra at order 5) advance in one
batch, FP32 and FP64, to a short final time.  Bars: FP32 vs FP64 <= 1e-5
relative L2 for every instance; FP64 vs the oracle's LSERK4 run (Solver::run
loop, Solver.cpp:483-551) <= 1e-11 for the largest instance (n = 22, order
5), an n = 20 one and the smallest; two instances bitwise equal alone and in
the mix.

Serving: a seeded Poisson stream drawn from the stated ranges (K log-uniform
1k-400k tetrahedra, order in {2,3,4}, final time U[0.5,2], eps_r ~ U[1,4] per
element, PEC cavities and SMA boxes) with a request above 200k tetrahedra;
every result is bitwise equal to the same request solved alone."""
import numpy as np
import pytest

import dgtd_amd as D
from dgtd_amd import serve as S
from dgtd_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

T_BATCH = 0.01


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def batch64():
    insts = W.cavity_batch(2024, 64)
    assert max(n for n, *_ in insts) == 22 and {o for _, o, *_ in insts} == {1, 2, 3, 4, 5}
    out = {}
    for prec in (32, 64):
        b = D.Batch(precision=prec, max_instances=64)
        ids = [b.add(vx, E, fbc, o, x0=x0, final_time=T_BATCH) for (n, o, vx, E, fbc, x0) in insts]
        b.run(poll_every=8)
        out[prec] = [(b.get_state(i), b.info(i)) for i in ids]
        b.close()
    return insts, out


def test_batch64_fp32_tracks_fp64(batch64):
    insts, out = batch64
    for k in range(64):
        (x32, i32), (x64, i64) = out[32][k], out[64][k]
        assert i32["state"] == i64["state"] == "done" and i32["steps"] == i64["steps"]
        assert rel(x32, x64) <= 1e-5, (k, insts[k][:2])


def test_batch64_fp64_matches_oracle(batch64):
    insts, out = batch64
    sizes = [n for n, *_ in insts]
    pick = {int(np.argmax(sizes)), sizes.index(20), int(np.argmin(sizes))}
    assert len(pick) == 3
    for k in sorted(pick):
        n, order, vx, E, fbc, x0 = insts[k]
        x64, info = out[64][k]
        ref = O.Operator(vx, E, fbc, order)
        assert info["dt"] == pytest.approx(1.0 / (ref.max_fscale() * order ** 2), rel=1e-14)
        x = x0.copy()
        steps, t = ref.run(x, info["dt"], T_BATCH, "lserk4")
        assert steps == info["steps"]
        assert rel(x64, x) <= 1e-11, (n, order)


def test_batch64_alone_equals_in_mix(batch64):
    insts, out = batch64
    sizes = [n for n, *_ in insts]
    for k in (sizes.index(21), int(np.argmin(sizes))):
        n, order, vx, E, fbc, x0 = insts[k]
        b = D.Batch(precision=32, max_instances=1)
        i = b.add(vx, E, fbc, order, x0=x0, final_time=T_BATCH)
        b.run(poll_every=8)
        np.testing.assert_array_equal(b.get_state(i), out[32][k][0])
        b.close()


def test_serving_stream_at_stated_sizes_equals_alone():
    reqs = S.poisson_stream(rate=40.0, duration=0.25, seed=297)
    assert max(r.K for r in reqs) > 200_000
    assert {r.order for r in reqs} == {2, 3, 4} and {r.bc for r in reqs} == {"pec", "sma"}
    assert all(1000 <= r.K <= 400_000 and 0.5 <= r.final_time <= 2.0 for r in reqs)
    srv = S.Server(precision=32, hbm_budget=int(64e9), max_instances=64, poll_every=4, workers=8)
    srv.keep_fields = True
    recs = srv.serve(reqs)
    srv.close()
    assert S.summarize(recs, srv.wall)["completed"] == len(reqs)
    for r in reqs:
        rec = recs[r.rid]
        assert rec.status == "done"
        vx, E, fbc, eps, x0 = S.build_request(r)
        b = D.Batch(precision=32, max_instances=1)
        i = b.add(vx, E, fbc, r.order, x0=x0, final_time=r.final_time, eps=eps)
        b.run(poll_every=8)
        assert b.info(i)["steps"] == rec.steps
        np.testing.assert_array_equal(b.get_state(i), rec.extra["x"])
        b.close()
