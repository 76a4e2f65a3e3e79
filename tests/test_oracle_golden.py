"""The oracle against the golden vectors of the reference's own tests
(test/hesthavenComparison/Hesthaven3DTest.cpp), committed as fixtures under
tests/golden/."""
import json
import os

import numpy as np
import pytest

import meshgen as MG
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLD, "hesthaven3d_golden.json")) as fh:
        return json.load(fh)


def test_mass_matrix_p2(golden):
    """Hesthaven3DTest massMatrix: the p=2 reference mass in MFEM node order."""
    g = golden["massMatrix_p2"]
    M = O.ref_element(2)["M"]
    ref = np.array(g["matrix"])
    # Eigen isApprox: ||a - b|| <= prec * min(||a||, ||b||)
    assert np.linalg.norm(M - ref) <= g["tolerance_rel"] * min(np.linalg.norm(M), np.linalg.norm(ref))
    # the golden entries are 5-digit roundings of exact rationals: V/70 etc. with V = 4/3
    V = 4.0 / 3.0
    assert M[0, 0] == pytest.approx(V / 70, rel=1e-12)
    assert M[0, 2] == pytest.approx(V / 420, rel=1e-12)
    assert M[1, 1] == pytest.approx(32 * V / 420, rel=1e-12)


def test_derivative_matrices_p1(golden):
    """Hesthaven3DTest DerivativeOperators_onetetra: Dr, Ds, Dt at p=1."""
    g = golden["derivatives_p1"]
    R = O.ref_element(1)
    for k in ("Dr", "Ds", "Dt"):
        np.testing.assert_allclose(R[k], np.array(g[k]), atol=1e-13)


def read_mfem_mesh(path):
    """Minimal MFEM mesh v1.0 reader for the tetrahedral fixture."""
    with open(path) as fh:
        toks = [ln.strip() for ln in fh if ln.strip() and not ln.startswith("#")]
    i = toks.index("elements")
    ne = int(toks[i + 1])
    EToV = np.array([[int(v) for v in toks[i + 2 + k].split()[2:]] for k in range(ne)], dtype=np.int32)
    i = toks.index("vertices")
    nv = int(toks[i + 1])
    vx = np.array([[float(v) for v in toks[i + 3 + k].split()] for k in range(nv)])
    return vx, EToV


def test_physical_derivatives_onetetra(golden):
    """On onetetra.mesh (the unit tet) the reference's per-element derivative
    M^-1 S_d Vref/V_e (HesthavenEvolution.cpp:158) is D_y = 2 Ds, so with
    E = (0, 0, y), H = 0, PMC faces and centred flux (all jumps zero) the
    operator returns rhsH_x = -dEz/dy = -1 and zero elsewhere."""
    vx, EToV = read_mfem_mesh(os.path.join(GOLD, "onetetra.mesh"))
    op = O.Operator(vx, EToV, np.full((1, 4), 2, dtype=np.int32), 1, alpha=0.0)
    pos = op.positions()
    q = np.zeros((6, 1, 4))
    q[2] = pos[1]
    r = op.mult(q.ravel()).reshape(6, 4)
    np.testing.assert_allclose(r[3], -1.0, atol=1e-13)
    np.testing.assert_allclose(np.delete(r, 3, axis=0), 0.0, atol=1e-13)
    Ds = np.array(golden["derivatives_p1"]["Ds"])
    np.testing.assert_allclose(2 * Ds @ pos[1][0], np.ones(4), atol=1e-13)


def test_emat_p1(golden):
    """Hesthaven3DTest EmatO1: face f opposite vertex f, columns ascending."""
    g = golden["emat_p1"]
    E = O.ref_element(1)["Emat"]
    np.testing.assert_allclose(E, np.array(g["matrix"]), atol=g["tolerance_abs"])
    # exact values: face mass (A/12)[2 1 1; 1 2 1; 1 1 2] with A = 2
    np.testing.assert_allclose(E[1, :3], [1 / 3, 1 / 6, 1 / 6], atol=1e-14)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_boundary_tetra_information(golden, order):
    """Hesthaven3DTest bdrTetraInformation on 3D_Bdr_Cube.msh, p = 1..5."""
    g = golden["bdrTetraInformation"]
    vx, tets, tris = MG.read_gmsh22(os.path.join(GOLD, "3D_Bdr_Cube.msh"))
    tets = MG.fix_orientation(vx, tets)
    tag_to_bc = {int(k): v for k, v in g["tag_to_bc"].items()}
    fbc = MG.face_bc_from_triangles(tets, tris, tag_to_bc)
    op = O.Operator(vx, tets, fbc, order)
    nfp = (order + 1) * (order + 2) // 2
    pos = op.positions().reshape(3, -1)
    for kind, axis in ((1, 0), (2, 1), (3, 2)):
        nodes = op.bc_faces(kind)
        assert nodes.shape == (g["faces_per_bc"], nfp)
        vals = pos[axis][nodes.ravel()]
        assert np.all((np.abs(vals) <= g["position_tol"]) | (np.abs(vals - 1.0) <= g["position_tol"]))
