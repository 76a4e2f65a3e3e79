"""Host side of the library on the CPU (no GPU needed): the C ABI loads and
exports every symbol include/dgtd_hip.h declares, the FP64 setup (reference
element, geometry, connectivity, mesh generator) matches the oracle, and
invalid inputs are refused with the documented error codes."""
import os
import re

import numpy as np
import pytest

import dgtd_amd as D
import meshgen as MG
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "dgtd_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(dgtd_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    lib = D.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding types every declared function
    from dgtd_amd._abi import SIGNATURES
    assert sorted(SIGNATURES) == names


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_reference_element_matches_oracle(p):
    a, b = D.ref_element(p), O.ref_element(p)
    for k in ("rst", "M", "Dr", "Ds", "Dt", "LIFT"):
        np.testing.assert_allclose(a[k], b[k], atol=1e-12 * max(1, np.abs(b[k]).max()), err_msg=k)
    np.testing.assert_array_equal(a["Fmask"], b["Fmask"])


def test_cube_generator_matches_test_generator():
    for n, bc in ((1, "pec"), (3, ["pec", "pmc", "sma", "none", "pec", "pmc"])):
        vx, E, fbc = D.cube_mesh(n, 2.0, bc)
        vx2, E2, fbc2 = MG.cube_mesh(n, 2.0, bc)
        np.testing.assert_allclose(vx, vx2, atol=1e-15)
        np.testing.assert_array_equal(E, E2)
        np.testing.assert_array_equal(fbc, fbc2)


def shuffled_mesh(seed):
    """Cube with randomly rotated local vertex orders and permuted vertex ids."""
    rng = np.random.default_rng(seed)
    vx, E, fbc = MG.cube_mesh(2, 1.0, ["pec", "pmc", "sma", "pec", "pmc", "sma"])
    perm = rng.permutation(vx.shape[0])
    inv = np.argsort(perm)
    vx = vx[perm]
    E = inv[E].astype(np.int32)
    rolls = rng.integers(0, 4, E.shape[0])
    E = np.array([np.roll(E[e], k) for e, k in enumerate(rolls)], dtype=np.int32)
    fbc = np.array([np.roll(fbc[e], k) for e, k in enumerate(rolls)], dtype=np.int32)
    return vx, E, fbc


@pytest.mark.parametrize("p", [1, 3, 5])
@pytest.mark.parametrize("mesh", ["cube", "shuffled", "gmsh"])
def test_connectivity_matches_oracle(p, mesh):
    if mesh == "cube":
        vx, E, fbc = MG.cube_mesh(3, 1.0, ["pec", "pmc", "sma", "pec", "none", "sma"])
    elif mesh == "shuffled":
        vx, E, fbc = shuffled_mesh(p)
    else:
        vx, E, tris = MG.read_gmsh22(os.path.join(ROOT, "tests", "golden", "3D_Bdr_Cube.msh"))
        E = MG.fix_orientation(vx, E)
        fbc = MG.face_bc_from_triangles(E, tris, {1: 1, 2: 1, 3: 2, 4: 2, 5: 3, 6: 3})
    ref = O.Operator(vx, E, fbc, p)
    M, P = ref.connectivity()
    M2, P2, codes = D.mesh_vmaps(vx, E, fbc, p)
    np.testing.assert_array_equal(M2, M)
    np.testing.assert_array_equal(P2, P)
    bnd = MG.boundary_faces(E)
    np.testing.assert_array_equal(codes, np.where(bnd, fbc, 0))


def test_geometry_matches_oracle():
    vx, E, fbc = shuffled_mesh(4)
    rng = np.random.default_rng(0)
    eps, mu = rng.uniform(1, 4, E.shape[0]), rng.uniform(1, 2, E.shape[0])
    geo = D.mesh_geometry(vx, E, fbc, 2, eps=eps, mu=mu)
    ref = O.Operator(vx, E, fbc, 2, eps=eps, mu=mu)
    vol, fs, nr = ref.geometry()
    f = geo[:, 9:25].reshape(-1, 4, 4)
    np.testing.assert_allclose(f[:, :, :3], nr, atol=1e-13)
    np.testing.assert_allclose(2 * f[:, :, 3], fs, rtol=1e-13)
    np.testing.assert_allclose(geo[:, 25], 1 / eps, rtol=1e-15)
    np.testing.assert_allclose(geo[:, 26], 1 / mu, rtol=1e-15)
    # metric: dr/dx etc. reproduce the affine map
    v = vx[E]
    J = 0.5 * np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)  # dx_d/dr_k
    G = geo[:, :9].reshape(-1, 3, 3)  # [k][d] = dr_k/dx_d
    np.testing.assert_allclose(np.einsum("ekd,edl->ekl", G, J), np.broadcast_to(np.eye(3), G.shape), atol=1e-12)


def test_create_rejects_invalid_input_before_touching_the_gpu():
    vx, E, fbc = MG.cube_mesh(1, 1.0, "sma")
    cases = [
        (dict(order=0), D._abi.DgtdError, -4),
        (dict(order=6), D._abi.DgtdError, -4),
        (dict(order=2, precision=16), D._abi.DgtdError, -1),
        (dict(order=2, alpha=0.0), D._abi.DgtdError, -1),  # SMA needs upwind
    ]
    for kw, exc, code in cases:
        order = kw.pop("order")
        with pytest.raises(exc) as ei:
            D.HesthavenEvolution(vx, E, fbc, order, **kw)
        assert ei.value.code == code
    bad = fbc.copy()
    bad[0, :] = 1  # boundary code on interior faces
    with pytest.raises(D.DgtdError, match="interior"):
        D.mesh_vmaps(vx, E, bad, 2)
    E2 = E.copy()
    E2[0, 0] = 10 ** 6
    with pytest.raises(D.DgtdError, match="out of range"):
        D.mesh_vmaps(vx, E2, fbc, 2)


def test_initial_field_projection_matches_reference_sources():
    """TE110 = "resonant" modes [1,1] polarised along z; a Gaussian on all six
    components = six unit-polarised sources; evaluated at the oracle's nodes."""
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    ref = O.Operator(vx, E, fbc, 3)
    pos = ref.positions()
    x = D.project_initial_fields(vx, E, 3, [D.resonant((1, 1))])
    np.testing.assert_allclose(x, MG.te110(pos), atol=1e-14)
    c = (0.4, 0.55, 0.5)
    srcs = [D.gaussian(f, c, 0.1, tuple(float(k == d) for k in range(3))) for f in "EH" for d in range(3)]
    x = D.project_initial_fields(vx, E, 3, srcs)
    np.testing.assert_allclose(x, MG.gaussian(pos, c, 0.1, comps=range(6)), atol=1e-14)
    # a 1D Gaussian uses x only (Function.h:73-78)
    x = D.project_initial_fields(vx, E, 3, [D.gaussian("H", c, 0.2, (0, 1, 0), dimension=1)])
    g = np.exp(-(pos[0] - c[0]) ** 2 / (2 * 0.04))
    np.testing.assert_allclose(x.reshape(6, -1)[4], g.ravel(), atol=1e-14)
    assert np.abs(np.delete(x.reshape(6, -1), 4, axis=0)).max() == 0


def test_empty_mesh_is_refused_before_any_device_call():
    import ctypes
    from dgtd_amd import _abi

    desc = _abi.MeshDesc()
    desc.K, desc.order, desc.nv, desc.dim = 0, 2, 0, 3
    opts = _abi.Opts(1.0, 64, _abi.DGTD_LSERK4, 0, _abi.DGTD_KERNEL_AUTO)
    h = ctypes.c_void_p()
    rc = _abi.lib().dgtd_op_create(ctypes.byref(desc), ctypes.byref(opts), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"at least one element" in _abi.lib().dgtd_last_error()
