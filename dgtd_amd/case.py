"""Case files of the reference's JSON launcher (driver.cpp:531-706, 940-1035,
1242-1300) through the library's C reader: mesh file (Gmsh 2.2 / MFEM v1.0),
uniform refinement, boundary tags, materials, initial-field sources, point
probes and solver options.  `run_case` mirrors Solver::run for the case
(src/solver/Solver.cpp:79-149, 483-551) on the hesthaven operator.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import check, dptr, i32ptr, lib
from .hesthaven import HesthavenEvolution, num_nodes

OPERATORS = {_abi.DGTD_OPERATOR_MAXWELL: "maxwell", _abi.DGTD_OPERATOR_GLOBAL: "global",
             _abi.DGTD_OPERATOR_HESTHAVEN: "hesthaven"}


class Case:
    """A loaded case: `info` (dict), mesh arrays `vx`, `EToV`, `face_bc`, `eps`,
    `mu`, `sigma` (normalised conductivity, None if no material conducts),
    `sources` (InitialField list), `planewaves` (PlaneWave list, the TF/SF
    faces carry code 7), `probes` (n x 3) and `field_probes` (n x 3, with
    `field_probe_components` 0..5 = Ex..Hz)."""

    def __init__(self, path):
        self.path = os.path.abspath(path)
        h = ctypes.c_void_p()
        check(lib().dgtd_case_load(self.path.encode(), ctypes.byref(h)))
        self._h = h
        try:
            ci = _abi.CaseInfo()
            check(lib().dgtd_case_get_info(h, ctypes.byref(ci)))
            self.info = {k: getattr(ci, k) for k, _ in _abi.CaseInfo._fields_}
            self.info["evolution_operator"] = OPERATORS[ci.evolution_operator]
            d = _abi.MeshDesc()
            check(lib().dgtd_case_mesh(h, ctypes.byref(d)))
            K, nv = d.K, d.nv
            self.quads = d.dim == 2 and d.nvert == 4
            self.hexes = d.dim != 2 and d.nvert == 8
            # triangles, quadrilaterals, tetrahedra or hexahedra
            nvert = 8 if self.hexes else 3 if d.dim == 2 and not self.quads else 4
            nfaces = 6 if self.hexes else nvert
            self.dim = 2 if d.dim == 2 else 3
            self.vx = np.ctypeslib.as_array(d.vx, shape=(3 * nv,)).reshape(nv, 3).copy()
            self.EToV = np.ctypeslib.as_array(d.EToV, shape=(nvert * K,)).reshape(K, nvert).copy()
            self.face_bc = np.ctypeslib.as_array(d.face_bc, shape=(nfaces * K,)).reshape(K, nfaces).copy()
            self.eps = None if not d.eps_r else np.ctypeslib.as_array(d.eps_r, shape=(K,)).copy()
            self.mu = None if not d.mu_r else np.ctypeslib.as_array(d.mu_r, shape=(K,)).copy()
            self.sigma = None if not d.sigma else np.ctypeslib.as_array(d.sigma, shape=(K,)).copy()
            n = ci.nsources
            arr = (_abi.InitialField * max(n, 1))()
            check(lib().dgtd_case_sources(h, arr))
            self.sources = [arr[i] for i in range(n)]
            w = (_abi.PlaneWave * max(ci.nplanewaves, 1))()
            check(lib().dgtd_case_planewaves(h, w))
            self.planewaves = [w[i] for i in range(ci.nplanewaves)]
            self.probes = np.zeros((ci.nprobes, 3))
            self.probe_steps = np.zeros(ci.nprobes, dtype=np.int32)
            self.probe_saves = np.zeros(ci.nprobes, dtype=np.int32)
            check(lib().dgtd_case_probes(h, dptr(self.probes), i32ptr(self.probe_steps),
                                         i32ptr(self.probe_saves)))
            nf = ci.nfield_probes
            self.field_probes = np.zeros((nf, 3))
            self.field_probe_steps = np.zeros(nf, dtype=np.int32)
            self.field_probe_saves = np.zeros(nf, dtype=np.int32)
            self.field_probe_components = np.zeros(nf, dtype=np.int32)  # 0..5 = Ex..Hz
            check(lib().dgtd_case_field_probes(h, dptr(self.field_probes), i32ptr(self.field_probe_steps),
                                               i32ptr(self.field_probe_saves),
                                               i32ptr(self.field_probe_components)))
        except Exception:
            self.close()
            raise

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().dgtd_case_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def save_mesh(self, path):
        """The case's (refined) mesh in MFEM v1.0 format."""
        check(lib().dgtd_case_save_mesh(self._h, os.fspath(path).encode()))

    @property
    def order(self):
        return self.info["order"]

    @property
    def K(self):
        return self.EToV.shape[0]

    @property
    def Np(self):
        p = self.order
        if self.quads:
            return (p + 1) ** 2
        if self.hexes:
            return (p + 1) ** 3
        return num_nodes(p) if self.dim == 3 else (p + 1) * (p + 2) // 2

    @property
    def ndofs(self):
        return 6 * self.Np * self.K

    def initial_state(self):
        from .hesthaven import _Mesh

        m = _Mesh(self.vx, self.EToV, self.face_bc, self.order, quads=self.quads)
        x = np.zeros(self.ndofs)
        srcs = (_abi.InitialField * max(len(self.sources), 1))(*self.sources)
        check(lib().dgtd_project_initial_fields(ctypes.byref(m.desc), srcs, len(self.sources), dptr(x)))
        return x

    def needs_assembled_operator(self):
        """True when kernel "auto" runs the assembled operator ("csr"), as the
        launcher decides: conductive materials, SMA faces (boundary or
        interior) with alpha != 1 under the Global or Maxwell operator (which
        weight SMA faces as at alpha = 1), quadrilateral meshes.  Hexahedral
        meshes never: their matrix-free stage applies the Global operator's
        flux and conductivity."""
        if self.hexes:
            return False
        global_sma = (self.info["evolution_operator"] != "hesthaven" and self.info["alpha"] != 1.0
                      and bool(((self.face_bc == 3) | (self.face_bc == 6)).any()))
        return self.sigma is not None or global_sma or self.quads

    def operator(self, precision=64, integrator="rk4", kernel="auto", device=0):
        """The case's operator; kernel="auto" selects the assembled one ("csr")
        where needs_assembled_operator() says so."""
        if kernel == "auto" and self.needs_assembled_operator():
            kernel = "csr"
        op = HesthavenEvolution(self.vx, self.EToV, self.face_bc, self.order, alpha=self.info["alpha"],
                                eps=self.eps, mu=self.mu, precision=precision, integrator=integrator,
                                kernel=kernel, device=device, sigma=self.sigma, quads=self.quads)
        if self.planewaves:
            op.set_planewaves(self.planewaves)
        return op

    def time_step(self, op=None):
        """time_step, or when it is 0 the reference's estimate
        (estimateTimeStep, Solver.cpp:110-115, 302-353): cfl / (max fscale *
        p^2) for the hesthaven operator on a 3D mesh, else the mesh rule
        0.75 * (2/3) * rmin * min vol/(surface/2) * cfl, halved on
        quadrilaterals, / 6 on hexahedra, / 0.8 for the Global operator on
        tetrahedra.  (op is
        accepted for compatibility and not needed.)"""
        dt = ctypes.c_double()
        check(lib().dgtd_case_time_step(self._h, ctypes.byref(dt)))
        return dt.value


def save_fields(directory, order, x, t):
    """Ex.gf ... Hz.gf (MFEM GridFunctions, L2 GaussLobatto, order p) and time.txt."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    check(lib().dgtd_save_fields(os.fspath(directory).encode(), int(order), x.size, dptr(x), float(t)))


def run_case(path, precision=64, integrator="rk4", final_time=None, kernel="auto"):
    """Load a case, project its initial fields and run it to final time on the
    GPU.  Returns (fields, info) with info holding steps, dt and time."""
    c = Case(path)
    op = c.operator(precision=precision, integrator=integrator, kernel=kernel)
    try:
        dt = c.time_step(op)
        op.set_state(c.initial_state(), 0.0)
        T = c.info["final_time"] if final_time is None else final_time
        steps = op.run(dt, T)
        x = op.get_state()
        return x, {"steps": steps, "dt": dt, "time": op.GetTime(), "K": c.K, "order": c.order,
                   "ndofs": c.ndofs}
    finally:
        op.close()


__all__ = ["Case", "run_case", "save_fields", "OPERATORS"]
