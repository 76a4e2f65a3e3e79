"""Test-side mesh and input generators (numpy), independent of the product's C++
generator so that tests can check one against the other.

Structured cube: n^3 hexahedra on [0, L]^3, each split into 6 tetrahedra
around the main diagonal (Kuhn split), as the reference's Mesh::MakeCartesian3D
with Element::TETRAHEDRON does for the test meshes it builds
(test/hesthavenComparison/Hesthaven3DTest.cpp:18,30).  A tetrahedron with
negative orientation gets local vertices 2 and 3 swapped, MFEM's fix on load.
"""
import itertools

import numpy as np

PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]

BC = {"none": 0, "pec": 1, "pmc": 2, "sma": 3}


def cube_mesh(n, L=1.0, bc="pec"):
    """Returns vx (nv,3), EToV (K,4) int32, face_bc (K,4) int32.

    bc: one code for all six sides, or a list of six codes for the sides
    x=0, x=L, y=0, y=L, z=0, z=L.
    """
    n1 = n + 1
    g = np.arange(n1, dtype=np.float64) * (L / n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    vx = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                   Z.transpose(2, 1, 0).ravel()], axis=1)

    def vid(i, j, k):
        return i + n1 * (j + n1 * k)

    ks, js, is_ = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ks, js, is_ = ks.ravel(), js.ravel(), is_.ravel()
    tets = []
    for perm in PERMS:
        c = np.zeros((is_.size, 3), dtype=np.int64)
        verts = [vid(is_, js, ks)]
        cur = np.zeros(3, dtype=np.int64)
        for ax in perm:
            cur[ax] += 1
            verts.append(vid(is_ + cur[0], js + cur[1], ks + cur[2]))
        tets.append(np.stack(verts, axis=1))
    # hex-major, 6 tets per hex in PERMS order
    EToV = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    EToV = fix_orientation(vx, EToV)
    face_bc = boundary_codes(vx, EToV, L, bc)
    return vx, EToV, face_bc


def fix_orientation(vx, EToV):
    v = vx[EToV]
    d = np.einsum("ij,ij->i", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), v[:, 3] - v[:, 0])
    E = EToV.copy()
    neg = d < 0
    E[neg, 2], E[neg, 3] = EToV[neg, 3], EToV[neg, 2]
    return E


FACE_VERTS = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]


def face_keys(EToV):
    keys = []
    for f in range(4):
        fv = np.sort(EToV[:, FACE_VERTS[f]], axis=1)
        keys.append(fv)
    return np.stack(keys, axis=1)  # K,4,3


def boundary_faces(EToV):
    """Boolean (K,4): face has no neighbour."""
    keys = face_keys(EToV).reshape(-1, 3)
    _, inv, counts = np.unique(keys, axis=0, return_inverse=True, return_counts=True)
    return (counts[inv.ravel()] == 1).reshape(-1, 4)


def boundary_codes(vx, EToV, L, bc):
    codes = [BC[bc] if isinstance(bc, str) else bc] * 6 if not isinstance(bc, (list, tuple)) \
        else [BC[b] if isinstance(b, str) else b for b in bc]
    bnd = boundary_faces(EToV)
    face_bc = np.zeros(EToV.shape, dtype=np.int32)
    tol = 1e-9 * L
    for f in range(4):
        c = vx[EToV[:, FACE_VERTS[f]]].mean(axis=1)  # K,3 face centroid
        for side in range(6):
            ax, hi = side // 2, side % 2
            on = np.abs(c[:, ax] - (L if hi else 0.0)) < tol
            sel = bnd[:, f] & on
            face_bc[sel, f] = codes[side]
    return face_bc


def te110(pos):
    """Initial TE110 cavity mode: Ez = sin(pi x) sin(pi y), other components zero
    (the reference's "resonant" initial field with modes [1,1], polarization z:
    src/math/Function.h:145-168, src/driver/driver.cpp:382-391)."""
    x, y = pos[0], pos[1]
    K, Np = x.shape
    q = np.zeros((6, K, Np))
    q[2] = np.sin(np.pi * x) * np.sin(np.pi * y)
    return q.ravel()


def te110_exact(pos, t):
    """Analytic TE110 mode in the PEC unit cube, eps = mu = 1."""
    x, y = pos[0], pos[1]
    w = np.pi * np.sqrt(2.0)
    K, Np = x.shape
    q = np.zeros((6, K, Np))
    q[2] = np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(w * t)
    q[3] = -(np.pi / w) * np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(w * t)
    q[4] = (np.pi / w) * np.cos(np.pi * x) * np.sin(np.pi * y) * np.sin(w * t)
    return q.ravel()


def gaussian(pos, center, spread, comps=(2,), amplitude=1.0):
    """Gaussian initial field exp(-|x-c|^2/(2 s^2)) on the given components
    (src/math/Function.h:45-100 with dimension 3)."""
    x, y, z = pos
    K, Np = x.shape
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    g = amplitude * np.exp(-r2 / (2.0 * spread ** 2))
    q = np.zeros((6, K, Np))
    for c in comps:
        q[c] = g
    return q.ravel()


def read_gmsh22(path):
    """Minimal Gmsh 2.2 ASCII reader: vertices, tetrahedra, boundary triangles
    with their physical tag (first tag)."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    ids, xyz = [], []
    for ln in lines[i + 2: i + 2 + nn]:
        t = ln.split()
        ids.append(int(t[0]))
        xyz.append([float(v) for v in t[1:4]])
    idmap = {v: k for k, v in enumerate(ids)}
    i = lines.index("$Elements")
    ne = int(lines[i + 1])
    tets, tris = [], []
    for ln in lines[i + 2: i + 2 + ne]:
        t = [int(v) for v in ln.split()]
        etype, ntags = t[1], t[2]
        tags = t[3: 3 + ntags]
        nodes = [idmap[v] for v in t[3 + ntags:]]
        if etype == 4:
            tets.append(nodes)
        elif etype == 2:
            tris.append((tags[0], nodes))
    return np.array(xyz), np.array(tets, dtype=np.int32), tris


def face_bc_from_triangles(EToV, tris, tag_to_code):
    face_bc = np.zeros(EToV.shape, dtype=np.int32)
    lookup = {}
    for tag, nodes in tris:
        lookup[tuple(sorted(nodes))] = tag_to_code[tag]
    for e in range(EToV.shape[0]):
        for f in range(4):
            key = tuple(sorted(EToV[e, list(FACE_VERTS[f])]))
            if key in lookup:
                face_bc[e, f] = lookup[key]
    return face_bc


def sort_by_position(pos, q, decimals=9):
    """Order DOFs by rounded node position then component, for comparisons
    that must not depend on the tet split's vertex order."""
    x = np.round(pos[0].ravel(), decimals)
    y = np.round(pos[1].ravel(), decimals)
    z = np.round(pos[2].ravel(), decimals)
    order = np.lexsort((z, y, x))
    qq = q.reshape(6, -1)
    return qq[:, order]


def plane_faces(vx, EToV, axis, value, tol=1e-12):
    """(K, 4) mask of element faces lying in the plane x[axis] = value."""
    mask = np.zeros(EToV.shape, dtype=bool)
    for f in range(4):
        pts = vx[EToV[:, list(FACE_VERTS[f])]]
        mask[:, f] = (np.abs(pts[..., axis] - value) < tol).all(axis=1)
    return mask


def cut_along_plane(vx, EToV, axis, value, tol=1e-12):
    """Copy of the mesh in which the elements on the far side of the plane
    x[axis] = value use their own copies of the plane's vertices, so the plane
    becomes two boundary surfaces.  Element order and local vertex order (hence
    node order) are unchanged."""
    on = np.where(np.abs(vx[:, axis] - value) < tol)[0]
    copy = {int(v): vx.shape[0] + i for i, v in enumerate(on)}
    vx2 = np.vstack([vx, vx[on]])
    far = vx[EToV].mean(axis=1)[:, axis] > value
    E2 = EToV.copy()
    for e in np.where(far)[0]:
        E2[e] = [copy.get(int(v), int(v)) for v in EToV[e]]
    return vx2, E2


# ---- triangles (the 2D operator) ----

TRI_EDGES = ((0, 1), (1, 2), (2, 0))


def square_mesh(n, L=1.0, bc="pec", diagonal="alternate", jitter=0.0, seed=0):
    """n x n squares on [0, L]^2, each split into two triangles.

    Returns vx (nv, 3) with z = 0, EToV (K, 3) int32 (counter-clockwise),
    face_bc (K, 3) with edge f = (v_f, v_{f+1 mod 3}).  bc: one code for all
    four sides or a list for the sides x=0, x=L, y=0, y=L.  diagonal:
    "alternate" flips the split diagonal in a checkerboard; jitter moves the
    interior vertices by up to jitter * h (a non-uniform mesh)."""
    n1 = n + 1
    g = np.arange(n1, dtype=np.float64) * (L / n)
    X, Y = np.meshgrid(g, g, indexing="xy")
    vx = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        inner = (vx[:, 0] > 0) & (vx[:, 0] < L) & (vx[:, 1] > 0) & (vx[:, 1] < L)
        vx[inner, :2] += rng.uniform(-jitter, jitter, (inner.sum(), 2)) * (L / n)

    def vid(i, j):
        return i + n1 * j

    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            if diagonal == "alternate" and (i + j) % 2:
                tris += [(a, b, d), (b, c, d)]
            else:
                tris += [(a, b, c), (a, c, d)]
    EToV = np.array(tris, dtype=np.int32)
    codes = [BC[bc]] * 4 if isinstance(bc, str) else [BC[c] if isinstance(c, str) else int(c) for c in bc]
    face_bc = np.zeros(EToV.shape, dtype=np.int32)
    for e in range(EToV.shape[0]):
        for f, (p, q) in enumerate(TRI_EDGES):
            A, B = vx[EToV[e, p], :2], vx[EToV[e, q], :2]
            for side, (axis, val) in enumerate([(0, 0.0), (0, L), (1, 0.0), (1, L)]):
                if abs(A[axis] - val) < 1e-12 and abs(B[axis] - val) < 1e-12:
                    face_bc[e, f] = codes[side]
    return vx, EToV, face_bc


def tm_mode(pos, t, m=1, n=1, L=1.0):
    """TM_mn mode of the PEC square [0,L]^2 (eps = mu = 1): E_z, H_x, H_y at
    positions pos (..., 2) and time t; the other three components are zero."""
    x, y = pos[..., 0], pos[..., 1]
    kx, ky = m * np.pi / L, n * np.pi / L
    w = np.hypot(kx, ky)
    ez = np.sin(kx * x) * np.sin(ky * y) * np.cos(w * t)
    hx = -(ky / w) * np.sin(kx * x) * np.cos(ky * y) * np.sin(w * t)
    hy = (kx / w) * np.cos(kx * x) * np.sin(ky * y) * np.sin(w * t)
    z = np.zeros_like(ez)
    return np.stack([z, z, ez, hx, hy, z])


def read_gmsh22_2d(path):
    """Minimal Gmsh 2.2 reader for a 2D mesh: vertices, triangles (with the
    elementary tag) and lines as (node, node, elementary tag)."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    i = lines.index("$Nodes")
    nn = int(lines[i + 1])
    ids, xyz = [], []
    for ln in lines[i + 2: i + 2 + nn]:
        t = ln.split()
        ids.append(int(t[0]))
        xyz.append([float(v) for v in t[1:4]])
    idmap = {v: k for k, v in enumerate(ids)}
    i = lines.index("$Elements")
    ne = int(lines[i + 1])
    tris, segs = [], []
    for ln in lines[i + 2: i + 2 + ne]:
        t = [int(v) for v in ln.split()]
        etype, ntags = t[1], t[2]
        tag = t[3 + 1] if ntags >= 2 else t[3]
        nodes = [idmap[v] for v in t[3 + ntags:]]
        if etype == 2:
            tris.append(nodes + [tag])
        elif etype == 1:
            segs.append((nodes[0], nodes[1], tag))
    return np.array(xyz), np.array(tris, dtype=np.int64), segs


# ---- quadrilaterals (parallelograms, the assembled operator) ----

QUAD_EDGES = ((0, 1), (1, 2), (2, 3), (3, 0))


def quad_mesh(nx, ny, Lx=1.0, Ly=1.0, bc="pec", stretch=0.0, shear=0.0, seed=0):
    """nx x ny counter-clockwise quadrilaterals (a, b, c, d) = (i,j), (i+1,j),
    (i+1,j+1), (i,j+1) on [0,Lx] x [0,Ly]; stretch varies the column widths
    and row heights by up to that fraction (rectangles of unequal sizes), and
    shear maps x -> x + shear * y afterwards, so every element stays a
    parallelogram.  bc: one code for all sides or a list for x=0, x=Lx, y=0,
    y=Ly (the sides of the mesh before the shear).  Returns vx (nv, 3) with
    z = 0, EToV (K, 4), face_bc (K, 4), edge f = (v_f, v_{f+1 mod 4})."""
    rng = np.random.default_rng(seed)

    def ticks(n, L):
        w = 1.0 + stretch * rng.uniform(-1, 1, n)
        return np.concatenate([[0.0], np.cumsum(w) / w.sum() * L])

    gx, gy = ticks(nx, Lx), ticks(ny, Ly)
    X, Y = np.meshgrid(gx, gy, indexing="xy")
    vx = np.stack([X.ravel() + shear * Y.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    codes = [BC[bc]] * 4 if isinstance(bc, str) else [BC[c] if isinstance(c, str) else int(c) for c in bc]
    EToV, face_bc = [], []
    for j in range(ny):
        for i in range(nx):
            a = i + (nx + 1) * j
            EToV.append((a, a + 1, a + nx + 2, a + nx + 1))
            face_bc.append((codes[2] if j == 0 else 0, codes[1] if i == nx - 1 else 0,
                            codes[3] if j == ny - 1 else 0, codes[0] if i == 0 else 0))
    return vx, np.array(EToV, dtype=np.int32), np.array(face_bc, dtype=np.int32)
