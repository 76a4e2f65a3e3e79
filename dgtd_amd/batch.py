"""Multi-instance batch over the C ABI (dgtd_batch_* in include/dgtd_hip.h):
independent meshes of uneven size and order advance together, one persistent
launch per LSERK4 stage, each time step replayed from a hipGraph, fields
resident in HBM."""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, dptr, lib
from .hesthaven import _Mesh, num_nodes

STATES = {_abi.DGTD_INST_FREE: "free", _abi.DGTD_INST_RUNNING: "running", _abi.DGTD_INST_DONE: "done",
          _abi.DGTD_INST_FAILED: "failed", _abi.DGTD_INST_CANCELLED: "cancelled"}


def footprint(vx, EToV, face_bc, order, precision=32, quads=None):
    m = _Mesh(vx, EToV, face_bc, order, quads=quads)
    b = ctypes.c_int64()
    check(lib().dgtd_batch_footprint(ctypes.byref(m.desc), int(precision), ctypes.byref(b)))
    return b.value


class CapacityError(_abi.DgtdError):
    pass


class Batch:
    def __init__(self, precision=32, alpha=1.0, max_instances=256, hbm_budget=0, device=0):
        opts = _abi.BatchOpts(float(alpha), int(precision), int(device), int(max_instances), 0,
                              int(hbm_budget))
        h = ctypes.c_void_p()
        check(lib().dgtd_batch_create(ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self.precision = precision
        self._sizes = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().dgtd_batch_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def add(self, vx, EToV, face_bc, order, x0=None, dt=0.0, final_time=1.0, eps=None, mu=None, planewaves=None):
        """Admit one mesh (tetrahedra or triangles); planewaves: dgtd_amd.planewave
        sources injected on its TF/SF faces (code 7).  Raises CapacityError when
        the HBM budget or the slots are full."""
        m = _Mesh(vx, EToV, face_bc, order, eps, mu)
        x = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        i = ctypes.c_int32()
        waves = list(planewaves or [])
        arr = (_abi.PlaneWave * max(1, len(waves)))(*waves)
        rc = lib().dgtd_batch_add_tfsf(self._h, ctypes.byref(m.desc), dptr(x), float(dt), float(final_time),
                                       arr, len(waves), ctypes.byref(i))
        if rc == _abi.DGTD_E_CAPACITY:
            raise CapacityError(rc, lib().dgtd_last_error().decode())
        check(rc)
        npe = num_nodes(order) if m.dim == 3 else (order + 1) * (order + 2) // 2
        self._sizes[i.value] = 6 * npe * m.EToV.shape[0]
        return i.value

    def step(self, nsteps=1):
        check(lib().dgtd_batch_step(self._h, int(nsteps)))

    def poll(self, cap=4096):
        """(ids that ended since the last poll -- done, failed or cancelled --,
        number still running)."""
        ids = np.zeros(cap, dtype=np.int32)
        nd, nr = ctypes.c_int32(), ctypes.c_int32()
        check(lib().dgtd_batch_poll(self._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                    ctypes.byref(nd), ctypes.byref(nr)))
        return ids[: nd.value].tolist(), nr.value

    def info(self, i):
        t, dt = ctypes.c_double(), ctypes.c_double()
        steps, state = ctypes.c_int32(), ctypes.c_int32()
        check(lib().dgtd_batch_info(self._h, int(i), ctypes.byref(t), ctypes.byref(steps),
                                    ctypes.byref(state), ctypes.byref(dt)))
        return {"t": t.value, "steps": steps.value, "state": STATES[state.value], "dt": dt.value}

    def get_state(self, i):
        x = np.empty(self._sizes[i])
        check(lib().dgtd_batch_get_state(self._h, int(i), dptr(x)))
        return x

    def cancel(self, i):
        """Stop a running instance; poll reports it once as ended ("cancelled")."""
        check(lib().dgtd_batch_cancel(self._h, int(i)))

    def remove(self, i):
        check(lib().dgtd_batch_remove(self._h, int(i)))
        self._sizes.pop(i, None)

    def stats(self):
        b = ctypes.c_int64()
        r, it, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().dgtd_batch_stats(self._h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(it),
                                     ctypes.byref(g)))
        return {"bytes_in_use": b.value, "running": r.value, "items": it.value, "grid": g.value}

    def run(self, poll_every=8):
        """Step until every instance is done; returns total steps launched."""
        n = 0
        while True:
            _, running = self.poll()
            if running == 0:
                return n
            self.step(poll_every)
            n += poll_every
