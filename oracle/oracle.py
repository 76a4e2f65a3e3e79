"""ctypes binding of the CPU FP64 oracle (oracle/dgtd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the dgtd_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_lib", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle library (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_np.restype = ctypes.c_int
        L.oracle_nfp.restype = ctypes.c_int
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip,
                                    _dp, _dp, ctypes.c_double]
        L.oracle_create_global.restype = ctypes.c_void_p
        L.oracle_create_global.argtypes = L.oracle_create.argtypes
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_ndofs.restype = ctypes.c_int64
        L.oracle_ndofs.argtypes = [ctypes.c_void_p]
        L.oracle_max_fscale.restype = ctypes.c_double
        L.oracle_max_fscale.argtypes = [ctypes.c_void_p]
        L.oracle_node_positions.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_connectivity.argtypes = [ctypes.c_void_p, _ip, _ip]
        L.oracle_bc_faces.restype = ctypes.c_int
        L.oracle_bc_faces.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.oracle_geometry.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.oracle_mult.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_rk4_step.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double]
        L.oracle_lserk4_step.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double,
                                         ctypes.c_double]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_run.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, _dp]
        L.oracle_energy.restype = ctypes.c_double
        L.oracle_energy.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_ref_element.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.oracle_gll01.argtypes = [ctypes.c_int, _dp]
        L.oracle_threads.restype = ctypes.c_int
        L.oracle_mult_t.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double]
        L.oracle_set_planewaves.restype = ctypes.c_int
        L.oracle_set_planewaves.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = None
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def np_of(order):
    return (order + 1) * (order + 2) * (order + 3) // 6


def nfp_of(order):
    return (order + 1) * (order + 2) // 2


def gll01(p):
    x = np.zeros(p + 1)
    lib().oracle_gll01(p, _d(x))
    return x


def ref_element(order):
    """Reference matrices of the oracle: dict of rst, M, Dr, Ds, Dt, Emat, LIFT, Fmask."""
    Np, Nfp = np_of(order), nfp_of(order)
    out = dict(rst=np.zeros((Np, 3)), M=np.zeros((Np, Np)), Dr=np.zeros((Np, Np)),
               Ds=np.zeros((Np, Np)), Dt=np.zeros((Np, Np)), Emat=np.zeros((Np, 4 * Nfp)),
               LIFT=np.zeros((Np, 4 * Nfp)), Fmask=np.zeros(4 * Nfp, dtype=np.int32))
    lib().oracle_ref_element(order, _d(out["rst"]), _d(out["M"]), _d(out["Dr"]), _d(out["Ds"]),
                             _d(out["Dt"]), _d(out["Emat"]), _d(out["LIFT"]), _i(out["Fmask"]))
    return out


class PlaneWave(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("pol", ctypes.c_double * 3), ("dir", ctypes.c_double * 3),
                ("spread", ctypes.c_double), ("mean", ctypes.c_double), ("freq", ctypes.c_double)]


class Operator:
    """The oracle's HesthavenEvolution over one tetrahedral mesh."""

    def __init__(self, vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None, global_sma=False):
        """global_sma: SMA faces (codes 3 and 6) at upwind weight 1 whatever
        alpha, as the reference's Global operator assembles them
        (DGOperatorFactory.h:21-35, 485-567, 575-665)."""
        self.vx = np.ascontiguousarray(vx, dtype=np.float64).reshape(-1, 3)
        self.EToV = np.ascontiguousarray(EToV, dtype=np.int32).reshape(-1, 4)
        self.face_bc = np.ascontiguousarray(face_bc, dtype=np.int32).reshape(-1, 4)
        self.K = self.EToV.shape[0]
        self.order = order
        self.Np, self.Nfp = np_of(order), nfp_of(order)
        self.eps = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
        self.mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        create = lib().oracle_create_global if global_sma else lib().oracle_create
        h = create(self.K, order, _d(self.vx), self.vx.shape[0], _i(self.EToV),
                                _i(self.face_bc), _d(self.eps), _d(self.mu), float(alpha))
        if not h:
            raise RuntimeError("oracle_create: " + lib().oracle_last_error().decode())
        self.h = h
        self.ndofs = int(lib().oracle_ndofs(h))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().oracle_destroy(h)
            self.h = None

    def positions(self):
        xyz = np.zeros(3 * self.Np * self.K)
        lib().oracle_node_positions(self.h, _d(xyz))
        return xyz.reshape(3, self.K, self.Np)

    def connectivity(self):
        n = 4 * self.Nfp * self.K
        M = np.zeros(n, dtype=np.int32)
        P = np.zeros(n, dtype=np.int32)
        lib().oracle_connectivity(self.h, _i(M), _i(P))
        return M, P

    def bc_faces(self, kind):
        nf = lib().oracle_bc_faces(self.h, kind, None)
        out = np.zeros(max(nf, 1) * self.Nfp, dtype=np.int32)
        lib().oracle_bc_faces(self.h, kind, _i(out))
        return out[: nf * self.Nfp].reshape(nf, self.Nfp)

    def geometry(self):
        vol = np.zeros(self.K)
        fs = np.zeros(4 * self.K)
        nr = np.zeros(12 * self.K)
        lib().oracle_geometry(self.h, _d(vol), _d(fs), _d(nr))
        return vol, fs.reshape(self.K, 4), nr.reshape(self.K, 4, 3)

    def max_fscale(self):
        return lib().oracle_max_fscale(self.h)

    def mult(self, x, t=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.ndofs)
        lib().oracle_mult_t(self.h, _d(x), _d(out), float(t))
        return out

    def set_planewaves(self, waves):
        """waves: dicts with kind (0 Gaussian, 1 modulated), pol, dir, spread, mean, freq."""
        arr = (PlaneWave * max(1, len(waves)))()
        for i, w in enumerate(waves):
            arr[i].kind = int(w.get("kind", 0))
            arr[i].pol[:] = list(w["pol"])
            arr[i].dir[:] = list(w["dir"])
            arr[i].spread = float(w["spread"])
            arr[i].mean = float(w.get("mean", 0.0))
            arr[i].freq = float(w.get("freq", 0.0))
        if lib().oracle_set_planewaves(self.h, arr, len(waves)) != 0:
            raise ValueError("bad plane wave")

    def rk4_step(self, x, t, dt):
        lib().oracle_rk4_step(self.h, _d(x), t, dt)

    def lserk4_step(self, q, res, t, dt):
        lib().oracle_lserk4_step(self.h, _d(q), _d(res), t, dt)

    def run(self, x, dt, final_time, integrator="rk4"):
        t = np.zeros(1)
        steps = lib().oracle_run(self.h, _d(x), dt, final_time, 0 if integrator == "rk4" else 1,
                                 _d(t))
        return steps, float(t[0])

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().oracle_energy(self.h, _d(x))


def threads():
    return lib().oracle_threads()


def set_threads(n):
    lib().oracle_set_threads(int(n))
