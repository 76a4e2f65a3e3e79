"""Host logic of the request loop (dgtd_amd/serve.py) on a stand-in batch that
needs no GPU: admission by footprint, refusal of a request that can never
fit, and retirement of failed solves."""
import numpy as np
import pytest

import dgtd_amd as D
from dgtd_amd import serve as S
from dgtd_amd.batch import CapacityError, footprint


class FakeBatch:
    """Admits while the summed footprint fits; an instance ends one step
    after admission, "failed" when its x0 holds a NaN."""

    def __init__(self, budget):
        self.budget, self.used = budget, 0
        self.inst, self.next = {}, 0

    def add(self, vx, E, fbc, order, x0=None, dt=0.0, final_time=1.0, eps=None):
        fp = footprint(vx, E, fbc, order, 32)
        if self.used + fp > self.budget:
            raise CapacityError(-6, "HBM budget exceeded")
        self.used += fp
        i = self.next
        self.next += 1
        bad = x0 is not None and not np.all(np.isfinite(x0))
        self.inst[i] = {"fp": fp, "steps": 0, "state": "running", "bad": bad}
        return i

    def step(self, n):
        for r in self.inst.values():
            if r["state"] == "running":
                r["steps"] += n
                r["state"] = "failed" if r["bad"] else "done"

    def poll(self):
        ended = [i for i, r in self.inst.items() if r["state"] != "running" and not r.get("seen")]
        for i in ended:
            self.inst[i]["seen"] = True
        return ended, sum(r["state"] == "running" for r in self.inst.values())

    def info(self, i):
        r = self.inst[i]
        return {"t": 0.0, "steps": r["steps"], "state": r["state"], "dt": 0.0}

    def get_state(self, i):
        return np.zeros(1)

    def remove(self, i):
        self.used -= self.inst.pop(i)["fp"]

    def close(self):
        pass


def test_request_that_never_fits_is_rejected_and_the_rest_complete():
    small = S.Request(0, 0.0, 2, 2, 0.5)
    big = S.Request(1, 0.0, 12, 4, 0.5)   # 10,368 tets at order 4: far above the budget
    after = S.Request(2, 0.001, 2, 3, 0.5)
    vx, E, fbc, eps, x0 = S.build_request(small)
    budget = 4 * footprint(vx, E, fbc, 3, 32)
    srv = S.Server(hbm_budget=budget, batch=FakeBatch(budget), poll_every=1, workers=2)
    recs = srv.serve([small, big, after])  # returns: nothing waits forever behind `big`
    srv.close()
    assert recs[1].status == "rejected"
    assert recs[0].status == "done" and recs[2].status == "done"
    summ = S.summarize(recs, srv.wall)
    assert summ["completed"] == 2 and summ["rejected"] == 1 and summ["failed"] == 0


def test_failed_solve_is_retired_and_counted(monkeypatch):
    reqs = [S.Request(i, 0.0, 2, 2, 0.5, seed=i) for i in range(3)]
    real_build = S.build_request

    def build(r):
        vx, E, fbc, eps, x0 = real_build(r)
        if r.rid == 1:
            x0 = x0.copy()
            x0[5] = np.nan
        return vx, E, fbc, eps, x0

    monkeypatch.setattr(S, "build_request", build)
    srv = S.Server(hbm_budget=1 << 40, batch=FakeBatch(1 << 40), poll_every=1, workers=2)
    recs = srv.serve(reqs)
    srv.close()
    assert [recs[i].status for i in range(3)] == ["done", "failed", "done"]
    summ = S.summarize(recs, srv.wall)
    assert summ["completed"] == 2 and summ["failed"] == 1


def test_footprint_counts_the_element_shape():
    """Admission measure per instance: three state buffers of 6 Np K values
    plus the geometry and connectivity records, for tetrahedra (27 values and
    8 ints per element) and triangles (15 values and 6 ints)."""
    import meshgen as MG

    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    K = E.shape[0]
    for order, prec in ((3, 32), (5, 64)):
        s = prec // 8
        Np = (order + 1) * (order + 2) * (order + 3) // 6
        assert footprint(vx, E, fbc, order, prec) == 3 * 6 * Np * K * s + K * 27 * s + K * 8 * 4
    vx, E, fbc = MG.square_mesh(3, bc="pec")
    K = E.shape[0]
    for order, prec in ((1, 32), (4, 64)):
        s = prec // 8
        Np = (order + 1) * (order + 2) // 2
        assert footprint(vx, E, fbc, order, prec) == 3 * 6 * Np * K * s + K * 15 * s + K * 6 * 4


def test_batch_refuses_quadrilaterals_and_hexahedra():
    """The batch advances hesthaven-operator instances; quadrilaterals and
    hexahedra run the Global operator only, so the admission measure
    refuses them (the same check guards dgtd_batch_add)."""
    for vx, E, fbc in (D.cartesian_2d(3, 2, "quad"), D.cartesian_3d(2, 2, 1, "hex")):
        with pytest.raises(D.DgtdError, match="Global operator only") as ei:
            D.footprint(vx, E, fbc, 2)
        assert ei.value.code == -4
