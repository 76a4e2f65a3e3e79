"""Host-side mirror of the reference's operator and solver surfaces over the C ABI.

- HesthavenEvolution mirrors maxwell::HesthavenEvolution
  (src/evolution/HesthavenEvolution.h:14-50): an mfem::TimeDependentOperator
  with Mult(in, out), SetTime/GetTime and height 6*NDofs.
- Solver mirrors maxwell::Solver's run/step/getTimeStep for the hesthaven
  operator (src/solver/Solver.cpp:79-149, 302-354, 483-551), with the state
  resident on the GPU.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, dptr, i32ptr, i8ptr, lib

INTEGRATORS = {"rk4": _abi.DGTD_RK4, "lserk4": _abi.DGTD_LSERK4}
# the autotuner's candidates, numbered as in dgtd_tuning
CANDIDATES = ["valu"] + [f"mfma_c{c}_{o}" for c in range(2) for o in ("morton", "element")]
KERNELS = {"auto": _abi.DGTD_KERNEL_AUTO, "valu": _abi.DGTD_KERNEL_VALU, "mfma": _abi.DGTD_KERNEL_MFMA,
           "csr": _abi.DGTD_KERNEL_CSR}


def num_nodes(order):
    return (order + 1) * (order + 2) * (order + 3) // 6


def num_face_nodes(order):
    return (order + 1) * (order + 2) // 2


def cube_mesh(n, L=1.0, bc="pec"):
    """n^3 hexahedra on [0,L]^3, six tetrahedra each (library generator).

    bc: one of none/pec/pmc/sma for all sides or a list of six codes
    (x=0, x=L, y=0, y=L, z=0, z=L).  Returns (vx, EToV, face_bc)."""
    names = {"none": 0, "pec": 1, "pmc": 2, "sma": 3}
    sides = [bc] * 6 if isinstance(bc, (str, int)) else list(bc)
    side = np.array([names[s] if isinstance(s, str) else int(s) for s in sides], dtype=np.int8)
    K = 6 * n ** 3
    vx = np.zeros(3 * (n + 1) ** 3)
    EToV = np.zeros(4 * K, dtype=np.int32)
    face_bc = np.zeros(4 * K, dtype=np.int8)
    check(lib().dgtd_cube_mesh(n, float(L), i8ptr(side), dptr(vx), i32ptr(EToV), i8ptr(face_bc)))
    return vx.reshape(-1, 3), EToV.reshape(-1, 4), face_bc.reshape(-1, 4)


def square_mesh(n, L=1.0, bc="pec"):
    """n^2 squares on [0,L]^2, two counter-clockwise triangles each, split
    along the (0,0)-(1,1) diagonal of every square (library generator for
    the 2D operator).  bc: one of none/pec/pmc/sma for all sides or a list of
    four (x=0, x=L, y=0, y=L).  Returns (vx (nv,3) with z = 0, EToV (K,3),
    face_bc (K,3)), edge f joining local vertices f and f+1 mod 3."""
    names = {"none": 0, "pec": 1, "pmc": 2, "sma": 3}
    sides = [bc] * 4 if isinstance(bc, (str, int)) else list(bc)
    side = [names[s] if isinstance(s, str) else int(s) for s in sides]
    g = np.linspace(0.0, L, n + 1)
    X, Y = np.meshgrid(g, g, indexing="xy")
    vx = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    a = (i + (n + 1) * j).ravel()
    b, c, d = a + 1, a + n + 2, a + n + 1
    EToV = np.empty((2 * a.size, 3), dtype=np.int32)
    EToV[0::2] = np.stack([a, b, c], axis=1)  # lower-right triangle
    EToV[1::2] = np.stack([a, c, d], axis=1)  # upper-left triangle
    face_bc = np.zeros(EToV.shape, dtype=np.int8)
    ii, jj = i.ravel(), j.ravel()
    face_bc[0::2, 0] = np.where(jj == 0, side[2], 0)          # a-b on y = 0
    face_bc[0::2, 1] = np.where(ii == n - 1, side[1], 0)      # b-c on x = L
    face_bc[1::2, 1] = np.where(jj == n - 1, side[3], 0)      # c-d on y = L
    face_bc[1::2, 2] = np.where(ii == 0, side[0], 0)          # d-a on x = 0
    return vx, EToV, face_bc


def cartesian_2d(nx, ny, shape="triangle", sx=1.0, sy=1.0, bc="pec"):
    """nx x ny cells on [0,sx] x [0,sy], as mfem::Mesh::MakeCartesian2D: per
    cell (a, b, c, d) = (i,j), (i+1,j), (i+1,j+1), (i,j+1) either one
    counter-clockwise quadrilateral (shape "quad") or the triangles (a, b, c)
    and (a, c, d) (shape "triangle").  bc: one code for all sides or four
    (bottom y=0, right x=sx, top y=sy, left x=0 -- MakeCartesian2D's boundary
    attributes 1..4).  Returns (vx (nv,3) with z = 0, EToV, face_bc), edge f
    joining local vertices f and f+1."""
    names = {"none": 0, "pec": 1, "pmc": 2, "sma": 3}
    sides = [bc] * 4 if isinstance(bc, (str, int)) else list(bc)
    side = [names[b] if isinstance(b, str) else int(b) for b in sides]
    X, Y = np.meshgrid(np.linspace(0.0, sx, nx + 1), np.linspace(0.0, sy, ny + 1), indexing="xy")
    vx = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    i, j = (g.ravel() for g in np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy"))
    a = i + (nx + 1) * j
    b, c, d = a + 1, a + nx + 2, a + nx + 1
    bottom = np.where(j == 0, side[0], 0)
    right = np.where(i == nx - 1, side[1], 0)
    top = np.where(j == ny - 1, side[2], 0)
    left = np.where(i == 0, side[3], 0)
    if shape == "quad":
        EToV = np.stack([a, b, c, d], axis=1).astype(np.int32)
        face_bc = np.stack([bottom, right, top, left], axis=1).astype(np.int8)
        return vx, EToV, face_bc
    if shape != "triangle":
        raise ValueError("shape must be 'triangle' or 'quad'")
    EToV = np.empty((2 * a.size, 3), dtype=np.int32)
    EToV[0::2] = np.stack([a, b, c], axis=1)
    EToV[1::2] = np.stack([a, c, d], axis=1)
    face_bc = np.zeros(EToV.shape, dtype=np.int8)
    face_bc[0::2, 0], face_bc[0::2, 1] = bottom, right  # a-b, b-c
    face_bc[1::2, 1], face_bc[1::2, 2] = top, left      # c-d, d-a
    return vx, EToV, face_bc


def cartesian_3d(nx, ny, nz, shape="hex", sx=1.0, sy=1.0, sz=1.0, bc="pec"):
    """nx x ny x nz cells on [0,sx] x [0,sy] x [0,sz], as
    mfem::Mesh::MakeCartesian3D: per cell one hexahedron with MFEM's vertex
    order (shape "hex"), or six positively oriented tetrahedra around the
    cell's (0,0,0)-(1,1,1) diagonal (shape "tet", a conforming split).  bc:
    one code for all sides or six for MakeCartesian3D's boundary attributes
    1..6: z=0, y=0, x=sx, y=sy, x=0, z=sz.  Returns (vx (nv,3), EToV,
    face_bc): hexahedron faces in MFEM's order (the same six sides), a
    tetrahedron's face f opposite its vertex f."""
    names = {"none": 0, "pec": 1, "pmc": 2, "sma": 3}
    sides = [bc] * 6 if isinstance(bc, (str, int)) else list(bc)
    side = [names[b] if isinstance(b, str) else int(b) for b in sides]
    X, Y, Z = np.meshgrid(np.linspace(0.0, sx, nx + 1), np.linspace(0.0, sy, ny + 1), np.linspace(0.0, sz, nz + 1),
                          indexing="ij")
    vx = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(), Z.transpose(2, 1, 0).ravel()], axis=1)
    i, j, k = (g.ravel() for g in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"))
    order = np.lexsort((i, j, k))  # x fastest, then y, then z
    i, j, k = i[order], j[order], k[order]
    vid = lambda a, b, c: a + (nx + 1) * (b + (ny + 1) * c)
    hexv = np.stack([vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
                     vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1), vid(i, j + 1, k + 1)], axis=1)
    planes = [(2, 0.0), (1, 0.0), (0, sx), (1, sy), (0, 0.0), (2, sz)]  # the six sides, attributes 1..6
    if shape == "hex":
        on = [k == 0, j == 0, i == nx - 1, j == ny - 1, i == 0, k == nz - 1]
        face_bc = np.stack([np.where(on[f], side[f], 0) for f in range(6)], axis=1).astype(np.int8)
        return vx, hexv.astype(np.int32), face_bc
    if shape != "tet":
        raise ValueError("shape must be 'hex' or 'tet'")
    kuhn = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    tets = np.concatenate([hexv[:, list(t)] for t in kuhn], axis=0)
    v = vx[tets]
    vol = np.einsum("ij,ij->i", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), v[:, 3] - v[:, 0])
    neg = vol < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    face_bc = np.zeros(tets.shape, dtype=np.int8)
    for f in range(4):
        fv = vx[np.delete(tets, f, axis=1)]  # (K, 3, 3): the face opposite vertex f
        for s_, (ax, val) in enumerate(planes):
            hit = np.all(np.abs(fv[:, :, ax] - val) < 1e-12, axis=1)
            face_bc[hit, f] = side[s_]
    return vx, tets.astype(np.int32), face_bc


class _Mesh:
    """Keeps the numpy arrays alive while the C descriptor points at them.

    Elements: triangles (EToV (K, 3)), tetrahedra (EToV (K, 4)),
    hexahedra (EToV (K, 8), MFEM's vertex order), or quadrilaterals: EToV
    (K, 4) with quads=True, or by default when the vertices are planar (vx
    (nv, 2), or z = 0 throughout, which no valid tetrahedron has)."""

    def __init__(self, vx, EToV, face_bc, order, eps=None, mu=None, sigma=None, quads=None):
        E = np.asarray(EToV)
        v = np.asarray(vx, dtype=np.float64)
        if quads is None:
            planar = v.ndim == 2 and (v.shape[1] == 2 or (v.shape[1] == 3 and not v[:, 2].any()))
            quads = E.ndim == 2 and E.shape[1] == 4 and planar
        self.quads = bool(quads)
        self.hexes = E.ndim == 2 and E.shape[1] == 8
        # triangles (K, 3) or quadrilaterals: the 2D operator; vertices may carry 2 or 3 coordinates
        self.dim = 2 if self.quads or (E.ndim == 2 and E.shape[1] == 3) else 3
        nvert = 8 if self.hexes else 4 if self.quads else self.dim + 1
        if self.dim == 2 and v.ndim == 2 and v.shape[1] == 2:
            v = np.hstack([v, np.zeros((v.shape[0], 1))])
        self.vx = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        self.EToV = np.ascontiguousarray(E, dtype=np.int32).reshape(-1, nvert)
        nfaces = 6 if self.hexes else nvert
        self.face_bc = None if face_bc is None else np.ascontiguousarray(face_bc, dtype=np.int8).reshape(-1, nfaces)
        self.eps = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
        self.mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        self.sigma = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        self.desc = _abi.MeshDesc(self.EToV.shape[0], order, dptr(self.vx), self.vx.shape[0],
                                  i32ptr(self.EToV), i8ptr(self.face_bc), dptr(self.eps),
                                  dptr(self.mu), self.dim, dptr(self.sigma),
                                  8 if self.hexes else 4 if self.quads else 0)
        self.Np = (order + 1) ** 3 if self.hexes else _nodes(order, self.dim, self.quads)


def gaussian(field="E", center=(0.5, 0.5, 0.5), spread=0.1, polarization=(0, 0, 1), dimension=3):
    """Gaussian initial field (the reference's "gaussian" source, driver.cpp:368-380)."""
    f = _abi.InitialField()
    f.kind, f.field = 0, 0 if field == "E" else 1
    f.center[:] = list(center)
    f.polarization[:] = list(polarization)
    f.spread, f.dimension = float(spread), int(dimension)
    return f


def resonant(modes=(1, 1), field="E", polarization=(0, 0, 1)):
    """Sinusoidal cavity mode prod sin(m_k pi x_k) (the "resonant" source, driver.cpp:382-391)."""
    f = _abi.InitialField()
    f.kind, f.field = 1, 0 if field == "E" else 1
    f.polarization[:] = list(polarization)
    f.nmodes = len(modes)
    f.modes[: len(modes)] = list(modes)
    return f


def bessel_j6(dim=2, field="E", polarization=(0, 0, 1)):
    """J6 cavity mode at t = 0, centred at the origin: J_6(a6 r) cos(6 phi) in
    2D ("besselj6_2D"), j_6(a6 r) Re Y_6^6 in 3D ("besselj6_3D"),
    driver.cpp:393-400, 474-481; a6 = 13.589290170541217."""
    f = _abi.InitialField()
    f.kind = 2 if dim == 2 else 3
    f.field = 0 if field == "E" else 1
    f.polarization[:] = list(polarization)
    return f


def planewave(polarization, propagation, spread, mean=0.0, frequency=None):
    """Gaussian (or, with a frequency, modulated Gaussian) plane wave for the
    TF/SF faces; mean is the projected 1D mean k.mean (driver.cpp:483-520)."""
    w = _abi.PlaneWave()
    w.kind = 0 if frequency is None else 1
    w.polarization[:] = list(polarization)
    w.propagation[:] = list(propagation)
    w.spread, w.mean = float(spread), float(mean)
    w.frequency = 0.0 if frequency is None else float(frequency)
    return w


def _nodes(order, dim, quads=False):
    if quads:
        return (order + 1) ** 2
    return num_nodes(order) if dim == 3 else (order + 1) * (order + 2) // 2


def project_initial_fields(vx, EToV, order, sources, quads=None):
    """Sum of the sources evaluated at the nodes: a 6*Np*K state vector."""
    m = _Mesh(vx, EToV, None, order, quads=quads)
    arr = (_abi.InitialField * len(sources))(*sources)
    x = np.empty(6 * m.Np * m.EToV.shape[0])
    check(lib().dgtd_project_initial_fields(ctypes.byref(m.desc), arr, len(sources), dptr(x)))
    return x


def ref_element(order):
    """The library's reference matrices (host FP64 setup)."""
    Np, Nfp = num_nodes(order), num_face_nodes(order)
    out = dict(rst=np.zeros((Np, 3)), M=np.zeros((Np, Np)), Dr=np.zeros((Np, Np)),
               Ds=np.zeros((Np, Np)), Dt=np.zeros((Np, Np)), LIFT=np.zeros((Np, 4 * Nfp)),
               Fmask=np.zeros(4 * Nfp, dtype=np.int32))
    check(lib().dgtd_ref_element(order, dptr(out["rst"]), dptr(out["M"]), dptr(out["Dr"]),
                                 dptr(out["Ds"]), dptr(out["Dt"]), dptr(out["LIFT"]),
                                 i32ptr(out["Fmask"])))
    return out


def mesh_vmaps(vx, EToV, face_bc, order, alpha=1.0):
    """(vmapM, vmapP, face_codes) exactly as the reference's Connectivities::global."""
    m = _Mesh(vx, EToV, face_bc, order)
    K = m.EToV.shape[0]
    n = 4 * num_face_nodes(order) * K
    M = np.zeros(n, dtype=np.int32)
    P = np.zeros(n, dtype=np.int32)
    codes = np.zeros(4 * K, dtype=np.int8)
    check(lib().dgtd_mesh_vmaps(ctypes.byref(m.desc), float(alpha), i32ptr(M), i32ptr(P),
                                i8ptr(codes)))
    return M, P, codes.reshape(K, 4)


def mesh_locate(vx, EToV, order, points, quads=None):
    """Host point location: (elem (n,), weights (n, Np)) as the probes use them."""
    m = _Mesh(vx, EToV, None, order, quads=quads)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    elem = np.zeros(pts.shape[0], dtype=np.int32)
    w = np.zeros((pts.shape[0], m.Np))
    check(lib().dgtd_mesh_locate(ctypes.byref(m.desc), pts.shape[0], dptr(pts), i32ptr(elem), dptr(w)))
    return elem, w


def mesh_geometry(vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None):
    m = _Mesh(vx, EToV, face_bc, order, eps, mu)
    geo = np.zeros(27 * m.EToV.shape[0])
    check(lib().dgtd_mesh_geometry(ctypes.byref(m.desc), float(alpha), dptr(geo)))
    return geo.reshape(-1, 27)


def global_assemble(vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None, sigma=None, quads=None):
    """The assembled operator of the "csr" kernel on the host, FP64:
    (rowptr, col, val) of the ndofs x ndofs CSR matrix (the Global operator's
    matrix, DGOperatorFactory::buildGlobalOperator, DGOperatorFactory.h:1428-1600)."""
    m = _Mesh(vx, EToV, face_bc, order, eps, mu, sigma, quads=quads)
    nnz = ctypes.c_int64()
    check(lib().dgtd_global_assemble(ctypes.byref(m.desc), float(alpha), ctypes.byref(nnz), None, None, None))
    n = 6 * m.Np * m.EToV.shape[0]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int32)
    val = np.zeros(nnz.value)
    check(lib().dgtd_global_assemble(ctypes.byref(m.desc), float(alpha), ctypes.byref(nnz),
                                     rowptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), i32ptr(col), dptr(val)))
    return rowptr, col, val


class HesthavenEvolution:
    """GPU hesthaven operator: Mult(in, out) over 6*Np*K FP64 host vectors.

    Construction errors raise DgtdError, as the reference constructor throws
    std::runtime_error (HesthavenEvolution.cpp:37,42,146)."""

    def __init__(self, vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None, precision=64,
                 integrator="lserk4", device=0, kernel="auto", sigma=None, quads=None):
        self.order = order
        self.precision = precision
        self.integrator = integrator
        self.alpha = alpha
        mesh = _Mesh(vx, EToV, face_bc, order, eps, mu, sigma, quads=quads)
        self.dim = mesh.dim
        self.quads = mesh.quads
        self.hexes = mesh.hexes
        self.K = mesh.EToV.shape[0]
        self.Np = mesh.Np
        self.Nfp = (order + 1) ** 2 if mesh.hexes else num_face_nodes(order) if mesh.dim == 3 else order + 1
        opts = _abi.Opts(float(alpha), int(precision), INTEGRATORS[integrator], int(device),
                         KERNELS[kernel])
        h = ctypes.c_void_p()
        check(lib().dgtd_op_create(ctypes.byref(mesh.desc), ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        check(lib().dgtd_op_ndofs(h, ctypes.byref(n)))
        self.ndofs = n.value
        self._t = 0.0

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().dgtd_op_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    # --- mfem::TimeDependentOperator surface ---
    def Height(self):
        return self.ndofs

    def Width(self):
        return self.ndofs

    def SetTime(self, t):
        self._t = float(t)

    def GetTime(self):
        return self._t

    def Mult(self, x, y):
        """y = L(x); y is overwritten (HesthavenEvolution.cpp:461-468,532-533)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.ndofs or y.size != self.ndofs or y.dtype != np.float64 or not y.flags.c_contiguous:
            raise ValueError("Mult expects contiguous float64 vectors of length Height()")
        check(lib().dgtd_op_mult(self._h, self._t, dptr(x), dptr(y)))

    def mult(self, x):
        y = np.empty(self.ndofs)
        self.Mult(x, y)
        return y

    def mult_device(self, d_in, d_out, stream=None):
        """Device pointers (ints) in the operator precision; asynchronous."""
        check(lib().dgtd_op_mult_device(self._h, self._t, ctypes.c_void_p(d_in),
                                        ctypes.c_void_p(d_out), ctypes.c_void_p(stream or 0)))

    # --- setup data ---
    def max_fscale(self):
        v = ctypes.c_double()
        check(lib().dgtd_op_max_fscale(self._h, ctypes.byref(v)))
        return v.value

    def node_positions(self):
        xyz = np.zeros(3 * self.Np * self.K)
        check(lib().dgtd_op_node_positions(self._h, dptr(xyz)))
        return xyz.reshape(3, self.K, self.Np)

    def kernel_variant(self):
        v = ctypes.c_int32()
        check(lib().dgtd_op_kernel_variant(self._h, ctypes.byref(v)))
        return {v_: k for k, v_ in KERNELS.items()}[v.value]

    def tuning(self):
        """The autotuner's choice (dgtd_op_tuning): variant, matrix-core
        configuration, waves per block, gather depth, residual prefetch, task
        order, whether it came from the process cache, and the timed
        candidates in microseconds per launch."""
        t = _abi.Tuning()
        check(lib().dgtd_op_tuning(self._h, ctypes.byref(t)))
        names = CANDIDATES
        return {"variant": {v_: k for k, v_ in KERNELS.items()}[t.variant], "config": t.config,
                "waves_per_block": t.waves_per_block, "gather_depth": t.gather_depth,
                "early_residual": bool(t.early_residual), "task_order": "morton" if t.task_order else "element",
                "from_cache": bool(t.from_cache),
                "candidates_us": {names[i]: round(t.candidate_us[i], 2) for i in range(min(t.ncandidates, 5))
                                  if t.candidate_us[i] > 0}}

    def set_candidate(self, candidate):
        """Pin the stage kernel to one autotuner candidate (dgtd_op_set_candidate):
        0 per-node, 1 + 2c + o matrix-core configuration c, Morton (o = 0) or
        element (o = 1) task order; the names of tuning()['candidates_us']."""
        if isinstance(candidate, str):
            candidate = CANDIDATES.index(candidate)
        check(lib().dgtd_op_set_candidate(self._h, int(candidate)))

    def work_per_element(self):
        b, f = ctypes.c_double(), ctypes.c_double()
        check(lib().dgtd_op_work_per_element(self._h, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value

    # --- resident state ---
    def set_state(self, x, t=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(lib().dgtd_op_set_state(self._h, dptr(x), float(t)))
        self._t = float(t)

    def get_state(self):
        x = np.empty(self.ndofs)
        t = ctypes.c_double()
        check(lib().dgtd_op_get_state(self._h, dptr(x), ctypes.byref(t)))
        self._t = t.value
        return x

    def advance(self, dt, nsteps):
        check(lib().dgtd_op_advance(self._h, float(dt), int(nsteps)))
        t = ctypes.c_double()
        check(lib().dgtd_op_time(self._h, ctypes.byref(t)))
        self._t = t.value

    def run(self, dt, final_time):
        n = ctypes.c_int32()
        check(lib().dgtd_op_run(self._h, float(dt), float(final_time), ctypes.byref(n)))
        t = ctypes.c_double()
        check(lib().dgtd_op_time(self._h, ctypes.byref(t)))
        self._t = t.value
        return n.value

    # --- TF/SF plane waves (HesthavenEvolution::evaluateTFSF, HesthavenEvolution.cpp:97-124) ---
    def set_planewaves(self, waves):
        """waves: iterable of planewave(...) structures (an empty list switches
        the injection off)."""
        waves = list(waves)
        arr = (_abi.PlaneWave * max(1, len(waves)))(*waves)
        check(lib().dgtd_op_set_planewaves(self._h, arr, len(waves)))

    # --- point probes (ProbesManager::updateProbe, ProbesManager.cpp:375-406) ---
    def probe_set(self, points):
        """Locate points (n x 3); returns the element of each (-1 outside)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        elem = np.zeros(pts.shape[0], dtype=np.int32)
        check(lib().dgtd_op_probe_set(self._h, pts.shape[0], dptr(pts), i32ptr(elem)))
        self._nprobes = pts.shape[0]
        return elem

    def probe_sample(self):
        """(n x 6) Ex, Ey, Ez, Hx, Hy, Hz of the resident state at the probes."""
        out = np.zeros((getattr(self, "_nprobes", 0), 6))
        check(lib().dgtd_op_probe_sample(self._h, dptr(out)))
        return out

    def norms(self):
        out = np.zeros(3)
        check(lib().dgtd_op_norms(self._h, dptr(out)))
        return {"norm2": out[0], "energy": out[1], "nonfinite": int(out[2])}

    def unstable(self):
        """True if a step since the last set_state stored a non-finite state
        entry or one above 1e20 (Solver::run's per-step stability test,
        Solver.cpp:499-516, which reports and goes on)."""
        f = ctypes.c_int32()
        check(lib().dgtd_op_unstable(self._h, ctypes.byref(f)))
        return bool(f.value)

    def set_graphs(self, on=True):
        """Replay time steps from a captured HIP graph, or (the default)
        launch every stage kernel directly."""
        check(lib().dgtd_op_set_graphs(self._h, 1 if on else 0))

    def set_timing(self, on=True):
        check(lib().dgtd_op_set_timing(self._h, 1 if on else 0))

    def last_kernel_us(self):
        us = ctypes.c_double()
        n = ctypes.c_int32()
        check(lib().dgtd_op_last_kernel_us(self._h, ctypes.byref(us), ctypes.byref(n)))
        return us.value, n.value


class GlobalEvolution(HesthavenEvolution):
    """The reference's Global operator (maxwell::GlobalEvolution,
    src/evolution/GlobalEvolution.cpp:628-1088) on the GPU: the operator
    assembled into one CSR matrix, applied by a fused SpMV + stage-update
    kernel, with per-element bulk conductivity sigma (normalised as sigma_SI
    * Z0, Material.h:30) adding -sigma/eps_r E (DGOperatorFactory.h:1351-1361)."""

    def __init__(self, vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None, sigma=None, precision=64,
                 integrator="lserk4", device=0, quads=None):
        super().__init__(vx, EToV, face_bc, order, alpha=alpha, eps=eps, mu=mu, precision=precision,
                         integrator=integrator, device=device, kernel="csr", sigma=sigma, quads=quads)


OPERATOR_TYPES = {"maxwell": _abi.DGTD_OPERATOR_MAXWELL, "global": _abi.DGTD_OPERATOR_GLOBAL,
                  "hesthaven": _abi.DGTD_OPERATOR_HESTHAVEN}


def mesh_time_step(vx, EToV, order, evolution_operator="global", cfl=1.0, quads=None):
    """estimate_time_step from the mesh alone, on the host."""
    m = _Mesh(vx, EToV, None, order, quads=quads)
    dt = ctypes.c_double()
    check(lib().dgtd_mesh_estimate_time_step(ctypes.byref(m.desc), OPERATOR_TYPES[evolution_operator], float(cfl),
                                             ctypes.byref(dt)))
    return dt.value


def estimate_time_step(op, cfl=1.0, evolution_operator=None):
    """Solver::estimateTimeStep (Solver.cpp:302-353) for op's mesh under the
    named evolution operator (default: "global" for a GlobalEvolution, else
    "hesthaven"): cfl / (max fscale * p^2) on tetrahedra under hesthaven,
    otherwise the mesh rule cfl * 0.75 * (2/3) * rmin * min vol/(surface/2),
    halved on quadrilaterals, / 0.8 on tetrahedra under Global."""
    if evolution_operator is None:
        evolution_operator = "global" if isinstance(op, GlobalEvolution) else "hesthaven"
    dt = ctypes.c_double()
    check(lib().dgtd_op_estimate_time_step(op._h, OPERATOR_TYPES[evolution_operator], float(cfl),
                                           ctypes.byref(dt)))
    return dt.value


class Solver:
    """Solver::run / step over the GPU operator with the state resident on device.

    time_step = 0 selects the automatic rule, as SolverOptions::time_step = 0
    does (Solver.cpp:111-116)."""

    def __init__(self, op, initial_state, final_time=2.0, time_step=0.0, cfl=1.0, evolution_operator=None):
        self.op = op
        self.final_time = final_time
        self.dt = estimate_time_step(op, cfl, evolution_operator) if time_step == 0.0 else time_step
        self.time = 0.0
        op.set_state(initial_state, 0.0)

    def step(self):
        truedt = min(self.dt, self.final_time - self.time)
        self.op.advance(truedt, 1)
        self.time += truedt

    def run(self):
        steps = self.op.run(self.dt, self.final_time)
        self.time = self.op.GetTime()
        return steps

    def getFields(self):
        return self.op.get_state()

    def getTimeStep(self):
        return self.dt
