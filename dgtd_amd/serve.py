"""Solve serving over the batch engine: requests arrive over time, are set up
on host worker threads (mesh, materials, initial fields), admitted
first-come-first-served while their device footprint fits the HBM budget
(geometry, connectivity and upload run inside the admission while the
previously queued steps execute on the GPU), and advance together with everything already
resident (continuous batching: each graph-replayed step moves every running
solve by one of its own time steps).  A finished solve is retired by the
device, detected at the next poll, timed and released.

Latency = completion time - arrival time.  Throughput = sum over completed
solves of DOF * RK stages / wall time; executed throughput also counts the
steps made by solves still running when the stream is cut off.  Neither has a
counterpart in the reference,
which runs one case per process (src/launcher/launcher.cpp).
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .batch import Batch, CapacityError, footprint
from .hesthaven import cube_mesh, gaussian, num_nodes, project_initial_fields


@dataclass
class Request:
    """One cavity or SMA-box solve: n^3 hexes x 6 tets on [0, L]^3."""
    rid: int
    arrival: float          # seconds from the start of the stream
    n: int
    order: int
    final_time: float
    bc: str = "pec"         # "pec" cavity or "sma" open box
    L: float = 1.0
    eps_range: tuple = (1.0, 4.0)
    seed: int = 0

    @property
    def K(self):
        return 6 * self.n ** 3

    @property
    def dof(self):
        return 6 * num_nodes(self.order) * self.K


@dataclass
class Record:
    rid: int
    arrival: float
    setup_done: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0
    steps: int = 0
    dof: int = 0
    order: int = 0
    K: int = 0
    status: str = ""        # "done", "failed" (stability test), "cancelled", "rejected" (never fits)
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def poisson_stream(rate, duration, seed=0, k_range=(1000, 400000), orders=(2, 3, 4), t_range=(0.5, 2.0),
                   bcs=("pec", "sma")):
    """Poisson arrivals; sizes log-uniform over k_range tets; order, final time
    and boundary drawn uniformly."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        K = float(np.exp(rng.uniform(np.log(k_range[0]), np.log(k_range[1]))))
        n = max(2, int(round((K / 6.0) ** (1.0 / 3.0))))
        out.append(Request(len(out), t, n, int(rng.choice(orders)), float(rng.uniform(*t_range)),
                           str(rng.choice(bcs)), seed=int(rng.integers(1 << 30))))


def build_request(req):
    """Host setup of one request (runs on a worker thread; the C calls release
    the GIL): mesh, per-element eps_r, seeded Gaussian E_z pulse."""
    vx, E, fbc = cube_mesh(req.n, req.L, req.bc)
    rng = np.random.default_rng(req.seed)
    eps = rng.uniform(*req.eps_range, E.shape[0])
    c = rng.uniform(0.3, 0.7, 3) * req.L
    x0 = project_initial_fields(vx, E, req.order, [gaussian("E", c, 0.05 * req.L, (0, 0, 1))])
    return vx, E, fbc, eps, x0


class Server:
    def __init__(self, precision=32, hbm_budget=int(32e9), max_instances=256, poll_every=4, workers=8,
                 device=0, batch=None):
        """batch: an object with Batch's add/step/poll/info/get_state/remove
        methods to serve on (default: a new Batch on `device`)."""
        self.batch = batch if batch is not None else Batch(precision=precision, max_instances=max_instances,
                                                           hbm_budget=hbm_budget, device=device)
        self.precision = precision
        self.hbm_budget = hbm_budget
        self.poll_every = poll_every
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.keep_fields = False

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()

    def serve(self, requests, max_wall=None):
        """Run a request stream in real time; returns {rid: Record}."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        records = {r.rid: Record(r.rid, r.arrival, dof=r.dof, order=r.order, K=r.K) for r in reqs}
        t0 = time.perf_counter()
        nxt = 0
        setups = deque()   # (req, future) in arrival order
        waiting = deque()  # (req, built) set up, not yet admitted
        inst_of = {}
        running = 0
        b = self.batch
        while True:
            now = time.perf_counter() - t0
            if max_wall is not None and now > max_wall:
                break
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                r = reqs[nxt]
                setups.append((r, self.pool.submit(build_request, r)))
                nxt += 1
            while setups and setups[0][1].done():
                r, fut = setups.popleft()
                rec = records[r.rid]
                rec.setup_done = time.perf_counter() - t0
                built = fut.result()
                if footprint(built[0], built[1], built[2], r.order, self.precision) > self.hbm_budget:
                    # could never be admitted: refuse it now rather than block the queue behind it
                    rec.status = "rejected"
                    rec.finished = rec.setup_done
                    continue
                waiting.append((r, built))
            if running:
                b.step(self.poll_every)  # asynchronous: admission below overlaps these steps
            admitted = 0
            while waiting:  # FIFO admission by footprint (host setup inside add overlaps the GPU)
                r, (vx, E, fbc, eps, x0) = waiting[0]
                try:
                    i = b.add(vx, E, fbc, r.order, x0=x0, dt=0.0, final_time=r.final_time, eps=eps)
                except CapacityError:
                    break
                waiting.popleft()
                inst_of[i] = r.rid
                records[r.rid].admitted = time.perf_counter() - t0
                admitted += 1
            if running or admitted:
                done, running = b.poll()
                tdone = time.perf_counter() - t0
                for i in done:
                    rid = inst_of.pop(i)
                    rec = records[rid]
                    rec.finished = tdone
                    info = b.info(i)
                    rec.steps = info["steps"]
                    rec.status = info["state"]  # done, or failed by the stability test
                    if self.keep_fields:
                        rec.extra["x"] = b.get_state(i)
                    b.remove(i)
            elif nxt >= len(reqs) and not setups and not waiting:
                break
            else:
                wake = reqs[nxt].arrival if nxt < len(reqs) else now + 1e-3
                time.sleep(max(0.0, min(1e-3, wake - now)))
        self.wall = time.perf_counter() - t0
        for i, rid in inst_of.items():  # cut off by max_wall: keep the steps they made
            records[rid].steps = b.info(i)["steps"]
        return records


def summarize(records, wall):
    """Latency percentiles over the solves that completed (status "done");
    failed and rejected requests are counted apart."""
    done = [r for r in records.values() if r.finished > 0 and r.status == "done"]
    lat = np.array([r.latency for r in done]) if done else np.zeros(1)
    work = sum(r.dof * 5 * r.steps for r in done)
    executed = sum(r.dof * 5 * r.steps for r in records.values())
    return {
        "requests": len(records),
        "completed": len(done),
        "failed": sum(r.status == "failed" for r in records.values()),
        "rejected": sum(r.status == "rejected" for r in records.values()),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "throughput_dof_stages_per_s": work / wall if wall > 0 else 0.0,
        "executed_dof_stages_per_s": executed / wall if wall > 0 else 0.0,
        "wall_s": wall,
    }


__all__ = ["Request", "Record", "Server", "poisson_stream", "build_request", "summarize", "_abi"]
