"""Serving on the GPU: a Poisson stream of uneven solves admitted into the
resident batch while others run; every result equals the same request solved
alone (FP32, bitwise: the batch kernels never mix instances), and each
request took the step count of its own Solver::run loop."""
import math

import numpy as np
import pytest

import dgtd_amd as D
from dgtd_amd import serve as S

pytestmark = pytest.mark.gpu


def test_stream_results_equal_alone_results():
    reqs = S.poisson_stream(rate=40.0, duration=0.25, seed=3, k_range=(48, 3000))
    assert len(reqs) >= 5
    srv = S.Server(precision=32, hbm_budget=int(2e9), max_instances=8, poll_every=2, workers=4)
    srv.keep_fields = True
    recs = srv.serve(reqs)
    srv.close()
    summ = S.summarize(recs, srv.wall)
    assert summ["completed"] == len(reqs)
    for r in reqs:
        rec = recs[r.rid]
        assert rec.finished >= rec.admitted >= rec.arrival
        vx, E, fbc, eps, x0 = S.build_request(r)
        b = D.Batch(precision=32, max_instances=1)
        i = b.add(vx, E, fbc, r.order, x0=x0, final_time=r.final_time, eps=eps)
        b.run(poll_every=3)
        info = b.info(i)
        assert info["steps"] == rec.steps
        np.testing.assert_array_equal(b.get_state(i), rec.extra["x"])
        # steps of the reference loop: while t <= T - 1e-8 dt
        dt = info["dt"]
        assert rec.steps == math.ceil(r.final_time / dt - 1e-8)
        b.close()


def test_admission_waits_for_capacity():
    """With room for one request at a time the stream still completes, FIFO."""
    reqs = S.poisson_stream(rate=100.0, duration=0.05, seed=5, k_range=(48, 400), orders=(2,))
    big = max(D.footprint(*S.build_request(r)[:3], r.order) for r in reqs)
    srv = S.Server(precision=32, hbm_budget=big, max_instances=4, poll_every=2, workers=2)
    recs = srv.serve(reqs)
    srv.close()
    adm = [recs[r.rid].admitted for r in sorted(reqs, key=lambda r: r.arrival)]
    assert all(recs[r.rid].finished > 0 for r in reqs)
    assert adm == sorted(adm)
