"""Case files of the reference's JSON launcher on the CPU: the library's JSON,
Gmsh 2.2 and MFEM mesh readers against independent Python readers of the
reference's own test-data files (tests/golden/cases: 3D_PEC and
3D_Resonant_Box_TM55_H1_P2 from testData/maxwellInputs), uniform refinement
against closed-form invariants, initial fields against their formulas, point
location against polynomial reproduction, and the documented refusals."""
import json
import os
import shutil

import numpy as np
import pytest

import dgtd_amd as D
import meshgen as MG
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = os.path.join(ROOT, "tests", "golden", "cases")
PEC = os.path.join(CASES, "3D_PEC.json")
TM55 = os.path.join(CASES, "3D_Resonant_Box_TM55_H1_P2.json")


def volumes(vx, E):
    v = vx[E]
    return np.einsum("ij,ij->i", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), v[:, 3] - v[:, 0]) / 6.0


def face_counts(E):
    keys = {}
    for f in range(4):
        for k in map(tuple, np.sort(E[:, list(MG.FACE_VERTS[f])], axis=1)):
            keys[k] = keys.get(k, 0) + 1
    return keys


def test_pec_case_options_sources_probes():
    c = D.Case(PEC)
    ref = json.load(open(PEC))
    so = ref["solver_options"]
    assert c.info["order"] == so["order"]
    assert c.info["alpha"] == so["upwind_alpha"]
    assert c.info["time_step"] == so["time_step"]
    assert c.info["final_time"] == so["final_time"]
    assert c.info["evolution_operator"] == so["evolution_operator"]
    assert c.info["exporter_steps"] == ref["probes"]["exporter"]["steps"]
    np.testing.assert_array_equal(c.probes, [p["position"] for p in ref["probes"]["point"]])
    np.testing.assert_array_equal(c.probe_steps, 0)
    (s,) = c.sources
    src = ref["sources"][0]
    assert (s.kind, s.field, s.dimension) == (0, 0, src["dimension"])
    assert s.spread == src["magnitude"]["spread"]
    assert list(s.center) == src["center"] and list(s.polarization) == src["polarization"]


def test_field_probes_parse(tmp_path):
    case = json.load(open(PEC))
    case["probes"]["field"] = [{"field_type": "electric", "polarization": "Z", "position": [0.1, 0.2, 0.3]},
                               {"field_type": "magnetic", "polarization": "X", "position": [0.5, 0.5, 0.5],
                                "steps": 3, "saves": 7}]
    shutil.copy(os.path.join(CASES, "3D_PEC.msh"), tmp_path / "3D_PEC.msh")
    p = tmp_path / "3D_PEC.json"
    p.write_text(json.dumps(case))
    c = D.Case(str(p))
    assert c.info["nfield_probes"] == 2
    np.testing.assert_array_equal(c.field_probes, [[0.1, 0.2, 0.3], [0.5, 0.5, 0.5]])
    np.testing.assert_array_equal(c.field_probe_components, [2, 3])
    np.testing.assert_array_equal(c.field_probe_steps, [0, 3])
    np.testing.assert_array_equal(c.field_probe_saves, [0, 7])
    assert c.probes.shape == (3, 3)  # point probes unchanged
    case["probes"]["field"][0]["polarization"] = "W"
    p.write_text(json.dumps(case))
    with pytest.raises(D.DgtdError, match="Wrong Field Polarization"):
        D.Case(str(p))


def test_gmsh_reader_matches_independent_reader():
    c = D.Case(PEC)
    vx, E, tris = MG.read_gmsh22(os.path.join(CASES, "3D_PEC.msh"))
    E = MG.fix_orientation(vx, E)
    np.testing.assert_array_equal(c.vx, vx)
    np.testing.assert_array_equal(c.EToV, E)
    assert (volumes(c.vx, c.EToV) > 0).all()
    codes = {1: 1, 2: 1, 5: 1, 6: 1, 3: 2, 4: 2}  # PEC tags 1,2,5,6; PMC 3,4
    np.testing.assert_array_equal(c.face_bc, MG.face_bc_from_triangles(E, tris, codes))
    assert np.bincount(c.face_bc.ravel()).tolist() == [4 * c.K - 704, 384, 320]
    assert c.eps is None and c.mu is None  # vacuum everywhere


def test_initial_fields_follow_their_formulas():
    c = D.Case(PEC)
    pos = O.Operator(c.vx, c.EToV, c.face_bc, c.order).positions()
    x = c.initial_state().reshape(6, -1)
    # 1D gaussian along x about 0.5, spread 0.07, on Ez (Function.h:45-100)
    np.testing.assert_allclose(x[2], np.exp(-(pos[0].ravel() - 0.5) ** 2 / (2 * 0.07 ** 2)), atol=1e-14)
    assert np.abs(np.delete(x, 2, axis=0)).max() == 0
    t = D.Case(TM55)
    pos = O.Operator(t.vx, t.EToV, t.face_bc, t.order).positions()
    x = t.initial_state().reshape(6, -1)
    # resonant modes [5, 5] on Ez: sin(5 pi x) sin(5 pi y) (Function.h:145-168)
    np.testing.assert_allclose(x[2], (np.sin(5 * np.pi * pos[0]) * np.sin(5 * np.pi * pos[1])).ravel(),
                               atol=1e-14)


def test_bessel_initial_fields_match_scipy():
    """besselj6_2D / besselj6_3D sources (Function.h:170-222) against scipy's
    Bessel and spherical-harmonic routines at the node positions."""
    import scipy.special as sp

    a6 = 13.589290170541217
    assert sp.jv(6, a6) == pytest.approx(0.0, abs=1e-14)  # the first zero of J_6
    vx, E, fbc = MG.cube_mesh(3, 1.0, "pec")
    vx = 2.0 * vx - 1.0  # [-1, 1]^3: r up to sqrt(3), a6 r up to 23.5 (series and recurrence)
    pos = O.Operator(vx, E, fbc, 3).positions().reshape(3, -1)
    x, y, z = pos
    r = np.sqrt(x * x + y * y + z * z)
    phi = np.arctan2(y, x)
    want3 = sp.spherical_jn(6, a6 * r) * np.real(sp.sph_harm_y(6, 6, np.arccos(np.clip(z / r, -1, 1)), phi))
    got3 = D.project_initial_fields(vx, E, 3, [D.bessel_j6(3, "H", (0, 1, 0))]).reshape(6, -1)
    np.testing.assert_allclose(got3[4], want3, rtol=1e-11, atol=1e-14 * np.abs(want3).max())
    assert np.abs(np.delete(got3, 4, axis=0)).max() == 0
    vt, Et, bt = MG.square_mesh(6, L=2.0, bc="pec", jitter=0.2, seed=3)
    vt = vt.copy()
    vt[:, :2] -= 1.0
    gt = D.project_initial_fields(vt, Et, 4, [D.bessel_j6(2)]).reshape(6, -1)
    from oracle import oracle2d as O2

    p2 = O2.Operator2D(vt, Et, bt, 4).pos.reshape(-1, 2)
    want2 = sp.jv(6, a6 * np.hypot(p2[:, 0], p2[:, 1])) * np.cos(6 * np.arctan2(p2[:, 1], p2[:, 0]))
    np.testing.assert_allclose(gt[2], want2, rtol=1e-12, atol=1e-15)
    assert np.abs(want2).max() > 0.1


def test_uniform_refinement_invariants():
    t = D.Case(TM55)  # "refinement": 1
    vx0, E0, tris0 = MG.read_gmsh22(os.path.join(CASES, "3D_Resonant_Box_TM55_H1_P2.msh"))
    E0 = MG.fix_orientation(vx0, E0)
    assert t.info["refinement"] == 1 and t.K == 8 * E0.shape[0]
    v0, v = volumes(vx0, E0), volumes(t.vx, t.EToV)
    assert (v > 0).all()
    np.testing.assert_allclose(v.reshape(-1, 8), np.repeat(v0[:, None] / 8, 8, axis=1), rtol=1e-12)
    np.testing.assert_allclose(v.sum(), 1.0, rtol=1e-12)
    edges = {tuple(sorted((a, b))) for tet in E0 for i, a in enumerate(tet) for b in tet[i + 1:]}
    assert t.vx.shape[0] == vx0.shape[0] + len(edges)
    # conforming: every face is shared by two children or lies on the boundary
    cnt = face_counts(t.EToV)
    assert set(cnt.values()) <= {1, 2}
    nb = sum(1 for k in cnt.values() if k == 1)
    assert nb == 4 * len(tris0)
    # every boundary face PEC, every interior face none
    bnd = MG.boundary_faces(t.EToV)
    np.testing.assert_array_equal(t.face_bc, np.where(bnd, 1, 0))


def write_case(tmp, mesh_name, boundaries, refinement=0):
    case = {"solver_options": {"order": 2, "final_time": 0.1},
            "model": {"filename": mesh_name, "refinement": refinement,
                      "materials": [{"tags": [1], "type": "vacuum"}], "boundaries": boundaries},
            "sources": []}
    p = os.path.join(tmp, "case.json")
    json.dump(case, open(p, "w"))
    return p


def test_mfem_mesh_reader_and_refinement(tmp_path):
    shutil.copy(os.path.join(ROOT, "tests", "golden", "onetetra.mesh"), tmp_path / "onetetra.mesh")
    bnd = [{"tags": [1], "type": "PEC"}, {"tags": [2], "type": "PMC"}, {"tags": [3], "type": "SMA"},
           {"tags": [4], "type": "PEC"}]
    c = D.Case(write_case(str(tmp_path), "onetetra.mesh", bnd))
    np.testing.assert_array_equal(c.vx, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    np.testing.assert_array_equal(c.EToV, [[0, 1, 2, 3]])
    # boundary element b lies opposite vertex b-1 (faces {1,2,3},{0,3,2},{0,1,3},{0,2,1})
    np.testing.assert_array_equal(c.face_bc, [[1, 2, 3, 1]])
    r = D.Case(write_case(str(tmp_path), "onetetra.mesh", bnd, refinement=2))
    assert r.K == 64
    np.testing.assert_allclose(volumes(r.vx, r.EToV), 1.0 / 6 / 64, rtol=1e-12)
    bf = MG.boundary_faces(r.EToV)
    assert bf.sum() == 4 * 16 and (r.face_bc[bf] > 0).all() and (r.face_bc[~bf] == 0).all()
    # children of a boundary face keep its attribute: 16 faces per parent face
    assert np.bincount(r.face_bc[bf]).tolist() == [0, 32, 16, 16]


def test_materials_by_tag(tmp_path):
    shutil.copy(os.path.join(CASES, "3D_PEC.msh"), tmp_path / "m.msh")
    p = write_case(str(tmp_path), "m.msh", [{"tags": [1, 2, 3, 4, 5, 6], "type": "SMA"}])
    case = json.load(open(p))
    case["model"]["materials"] = [{"tags": [1], "type": "dielectric", "relative_permittivity": 4.0,
                                   "relative_permeability": 2.0}]
    json.dump(case, open(p, "w"))
    c = D.Case(p)
    np.testing.assert_array_equal(c.eps, 4.0)
    np.testing.assert_array_equal(c.mu, 2.0)
    assert set(np.unique(c.face_bc)) == {0, 3}
    assert c.sigma is None


def test_bulk_conductivity_is_normalised_by_free_space_impedance(tmp_path):
    # driver.cpp:962-963: sigma = bulk_conductivity * Z0, Z0 = mu0 c0 (PhysicalConstants.h:9-15)
    shutil.copy(os.path.join(CASES, "3D_PEC.msh"), tmp_path / "m.msh")
    p = write_case(str(tmp_path), "m.msh", [{"tags": [1, 2, 3, 4, 5, 6], "type": "PEC"}])
    case = json.load(open(p))
    case["model"]["materials"] = [{"tags": [1], "type": "conductor", "bulk_conductivity": 0.01}]
    json.dump(case, open(p, "w"))
    c = D.Case(p)
    z0 = 4e-7 * np.pi * 299792458.0
    assert z0 == pytest.approx(376.730313, rel=1e-8)
    np.testing.assert_allclose(c.sigma, 0.01 * z0, rtol=1e-15)
    assert c.eps is None and c.mu is None


@pytest.mark.parametrize("mutate,code,msg", [
    (lambda c: c["sources"].append({"type": "dipole"}), -4, "dipole"),
    (lambda c: c["sources"].append({"type": "planewave", "polarization": [0, 0, 1], "propagation": [1, 0, 0],
                                    "tags": [1], "magnitude": {"type": "gaussian", "spread": 0.1}}),
     -1, "TF/SF attribute 1 lies on a boundary face"),
    (lambda c: c["model"]["boundaries"].__setitem__(0, {"tags": [1], "type": "XYZ"}), -1, "incorrect"),
    (lambda c: c["model"]["boundaries"].__setitem__(0, {"tags": [99], "type": "PEC"}), -1, "bdr_attribute"),
    (lambda c: c["model"]["boundaries"].__setitem__(0, {"tags": [1], "type": "SGBC"}), -4, "SGBC"),
    (lambda c: c["model"]["materials"][0].__setitem__("bulk_conductivity", -1.0), -1, "Conductivity under 0.0"),
    (lambda c: c["solver_options"].__setitem__("ode_type", 1), -4, "RK4"),
    (lambda c: c["solver_options"].__setitem__("evolution_operator", "foo"), -1, "Evolution Operator"),
    (lambda c: c["model"].__setitem__("filename", "missing.msh"), -1, "cannot open"),
    (lambda c: c["model"].__setitem__("filename", "m.vtk"), -1, "name.msh or name.mesh"),
    (lambda c: (c["solver_options"].__setitem__("upwind_alpha", 0.0),
                c["solver_options"].__setitem__("evolution_operator", "hesthaven"),
                c["model"]["boundaries"].__setitem__(0, {"tags": [1], "type": "SMA"})), -4, "Centered SMA"),
])
def test_case_refusals(tmp_path, mutate, code, msg):
    shutil.copy(os.path.join(CASES, "3D_PEC.msh"), tmp_path / "m.msh")
    p = write_case(str(tmp_path), "m.msh", [{"tags": [1, 2, 3, 4, 5, 6], "type": "PEC"}])
    case = json.load(open(p))
    mutate(case)
    json.dump(case, open(p, "w"))
    with pytest.raises(D.DgtdError, match=msg) as ei:
        D.Case(p)
    assert ei.value.code == code


@pytest.mark.parametrize("kind,code", [("PEC", 4), ("PMC", 5), ("SMA", 6)])
def test_interior_boundary_tags_become_interior_conditions(tmp_path, kind, code):
    # two tetrahedra sharing face {1,2,3}; a boundary element on the shared face
    txt = """MFEM mesh v1.0

dimension
3

elements
2
1 4 0 1 2 3
1 4 4 1 3 2

boundary
1
7 2 1 2 3

vertices
5
3
0 0 0
1 0 0
0 1 0
0 0 1
1 1 1
"""
    (tmp_path / "two.mesh").write_text(txt)
    c = D.Case(write_case(str(tmp_path), "two.mesh", [{"tags": [7], "type": kind}]))
    # the shared face is face 0 of element 0 and face 0 of element 1 (opposite vertex 4)
    np.testing.assert_array_equal(c.face_bc, [[code, 0, 0, 0], [code, 0, 0, 0]])


@pytest.mark.parametrize("p", [1, 3, 5])
def test_point_location_reproduces_polynomials(p):
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    pos = O.Operator(vx, E, fbc, p).positions().reshape(3, -1)
    rng = np.random.default_rng(p)
    pts = rng.uniform(0.02, 0.98, (40, 3))
    elem, w = D.mesh_locate(vx, E, p, pts)
    assert (elem >= 0).all()
    # random polynomial of total degree p, evaluated at nodes and at the points
    coef = rng.normal(size=(p + 1, p + 1, p + 1))

    def f(x, y, z):
        return sum(coef[a, b, c] * x ** a * y ** b * z ** c
                   for a in range(p + 1) for b in range(p + 1 - a) for c in range(p + 1 - a - b))

    u = f(*pos).reshape(E.shape[0], -1)
    got = np.einsum("pn,pn->p", w, u[elem])
    np.testing.assert_allclose(got, f(*pts.T), rtol=1e-11, atol=1e-11)
    # the containing element really contains the point
    v = vx[E[elem]]
    A = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)
    xi = np.linalg.solve(A, (pts - v[:, 0])[..., None])[..., 0]
    assert (xi > -1e-10).all() and (xi.sum(1) < 1 + 1e-10).all()


def test_point_location_nodes_outside_and_ties():
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    p = 4  # one interior node per element
    Np = D.num_nodes(p)
    pos = O.Operator(vx, E, fbc, p).positions().reshape(3, -1)
    v = vx[E[5]]
    A = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
    lam = np.linalg.solve(A, pos[:, 5 * Np: 6 * Np] - v[0][:, None])
    bary = np.vstack([1 - lam.sum(0), lam])
    k = int(np.argmax(bary.min(0)))  # the interior node
    assert bary[:, k].min() > 0.1
    elem, w = D.mesh_locate(vx, E, p, [pos[:, 5 * Np + k], [1.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    assert elem[0] == 5
    np.testing.assert_allclose(w[0], np.eye(Np)[k], atol=1e-11)
    assert elem[1] == -1 and np.abs(w[1]).max() == 0
    # a shared vertex goes to the lowest-index element holding it
    holds = [e for e in range(E.shape[0]) if (np.abs(vx[E[e]] - 0.5).sum(1) < 1e-14).any()]
    assert elem[2] == min(holds)


def read_gf(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "FiniteElementSpace"
    header = dict(ln.split(": ", 1) for ln in lines[1:4])
    vals = np.array([float(v) for v in lines[5:] if v.strip()])
    return header, vals


def test_snapshot_mesh_and_fields_round_trip(tmp_path):
    t = D.Case(TM55)  # refined once
    t.save_mesh(tmp_path / "snap.mesh")
    case = json.load(open(TM55))
    case["model"]["filename"] = "snap.mesh"
    case["model"]["refinement"] = 0
    json.dump(case, open(tmp_path / "again.json", "w"))
    r = D.Case(tmp_path / "again.json")
    np.testing.assert_allclose(r.vx, t.vx, rtol=0, atol=1e-15)  # 16 significant digits, as Mesh::Save(path, 16)
    np.testing.assert_array_equal(r.EToV, t.EToV)
    np.testing.assert_array_equal(r.face_bc, t.face_bc)
    x = t.initial_state()
    x[1::7] += 0.25  # all six components distinct from zero somewhere
    from dgtd_amd.case import save_fields
    save_fields(tmp_path, t.order, x, 0.125)
    n = x.size // 6
    for c, name in enumerate(["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]):
        header, vals = read_gf(tmp_path / f"{name}.gf")
        assert header == {"FiniteElementCollection": "L2_T1_3D_P2", "VDim": "1", "Ordering": "0"}
        np.testing.assert_allclose(vals, x[c * n:(c + 1) * n], rtol=1e-15, atol=0)  # 16 significant digits
    assert float((tmp_path / "time.txt").read_text()) == 0.125


def test_tfsf_case_plane_wave_and_sides():
    """3D_TFSF (testData/maxwellInputs): a Gaussian plane wave (E_z, +x) through
    the TF/SF plane x = 1 (tag 2), refined once."""
    c = D.Case(os.path.join(CASES, "3D_TFSF.json"))
    ref = json.load(open(os.path.join(CASES, "3D_TFSF.json")))
    (w,) = c.planewaves
    src = ref["sources"][0]
    assert w.kind == 0 and list(w.polarization) == src["polarization"]
    assert list(w.propagation) == src["propagation"]
    assert w.spread == src["magnitude"]["spread"]
    assert w.mean == pytest.approx(np.dot(src["magnitude"]["mean"], src["propagation"]))
    tf = c.face_bc == 7
    pts = c.vx[c.EToV[np.where(tf)[0]]]
    # every TF/SF face lies in the plane x = 1 (all three face vertices)
    faces = [c.EToV[e, list(MG.FACE_VERTS[f])] for e, f in zip(*np.where(tf))]
    assert np.allclose(c.vx[np.array(faces)][..., 0], 1.0)
    assert tf.sum() == 2 * 4 * 14  # 14 triangles, refined into 4 each, two sides
    assert pts.shape[0] == tf.sum()


def test_planewave_auto_mean(tmp_path):
    shutil.copy(os.path.join(CASES, "3D_TFSF.msh"), tmp_path / "3D_TFSF.msh")
    case = json.load(open(os.path.join(CASES, "3D_TFSF.json")))
    del case["sources"][0]["magnitude"]["mean"]
    json.dump(case, open(tmp_path / "c.json", "w"))
    (w,) = D.Case(tmp_path / "c.json").planewaves
    # the TF/SF plane is x = 1: min phase 1, minus 5 sqrt(2) spreads (driver.cpp:575-588)
    assert w.mean == pytest.approx(1.0 - 5.0 * 0.6 * np.sqrt(2.0))


PEC2D = os.path.join(CASES, "2D_PEC.json")


def test_2d_pec_case_triangles_segments_and_sources():
    """testData/maxwellInputs/2D_PEC: a Gmsh file without tetrahedra is a 2D
    mesh, triangles are the elements and lines the boundary; tags 2, 4 PEC
    and 1, 3 PMC; the Gaussian of dimension 1 varies along x only."""
    c = D.Case(PEC2D)
    assert c.dim == 2 and c.info["dim"] == 2
    vx, tris, lines = MG.read_gmsh22_2d(os.path.join(CASES, "2D_PEC.msh"))
    assert c.K == len(tris) == 246 and c.EToV.shape == (246, 3)
    # elements positively oriented, vertex sets as in the file
    v = c.vx[c.EToV][:, :, :2]
    area2 = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 2, 0] - v[:, 0, 0]) * (v[:, 1, 1] - v[:, 0, 1])
    assert (area2 > 0).all()
    assert sorted(map(tuple, np.sort(c.EToV, axis=1))) == sorted(map(tuple, np.sort(tris[:, :3], axis=1)))
    # boundary codes from the segment tags
    want = {}
    for a, b, tag in lines:
        want[tuple(sorted((a, b)))] = {2: 1, 4: 1, 1: 2, 3: 2}[tag]
    got = {}
    for e in range(c.K):
        for f, (p, q) in enumerate(MG.TRI_EDGES):
            if c.face_bc[e, f]:
                got[tuple(sorted((int(c.EToV[e, p]), int(c.EToV[e, q]))))] = int(c.face_bc[e, f])
    assert got == want and len(got) == 40
    # probes given as 2-vectors padded with z = 0
    np.testing.assert_array_equal(c.probes, [[0.0, 0.5, 0.0], [1.0, 0.5, 0.0], [0.5, 0.5, 0.0]])
    # initial Ez = exp(-(x - 0.5)^2 / (2 0.12^2)) at the triangle nodes
    from oracle import oracle2d as O2
    ref = O2.Operator2D(c.vx, c.EToV, c.face_bc, c.order)
    x = c.initial_state().reshape(6, c.K, -1)
    np.testing.assert_allclose(x[2], np.exp(-(ref.pos[..., 0] - 0.5) ** 2 / (2 * 0.12 ** 2)), atol=1e-14)
    assert np.abs(x[[0, 1, 3, 4, 5]]).max() == 0.0


def test_2d_snapshot_mesh_round_trip(tmp_path):
    c = D.Case(PEC2D)
    c.save_mesh(tmp_path / "mesh_rank0")
    text = (tmp_path / "mesh_rank0").read_text()
    assert "dimension\n2\n" in text
    # the written mesh reads back through the MFEM reader as the same triangles
    case = json.load(open(PEC2D))
    case["model"]["filename"] = str(tmp_path / "mesh.mesh")
    os.rename(tmp_path / "mesh_rank0", tmp_path / "mesh.mesh")
    (tmp_path / "c.json").write_text(json.dumps(case))
    c2 = D.Case(str(tmp_path / "c.json"))
    assert c2.dim == 2
    np.testing.assert_array_equal(c2.EToV, c.EToV)
    np.testing.assert_allclose(c2.vx, c.vx, atol=1e-15)
    np.testing.assert_array_equal(c2.face_bc, c.face_bc)


def test_2d_case_refusals(tmp_path):
    case = json.load(open(PEC2D))
    shutil.copy(os.path.join(CASES, "2D_PEC.msh"), tmp_path / "2D_PEC.msh")
    case["model"]["refinement"] = 1
    (tmp_path / "r.json").write_text(json.dumps(case))
    with pytest.raises(D.DgtdError, match="refinement of 2D meshes"):
        D.Case(str(tmp_path / "r.json"))
    # curved (second-order) triangles (type 9), prisms (type 6) and meshes
    # mixing triangles and quadrilaterals are refused by name
    for elems, msg in (("1\n1 9 2 1 1 1 2 3 4 5 6\n", "element type 9"),
                       ("1\n1 6 2 1 1 1 2 3 4 5 6\n", "element type 6"),
                       ("2\n1 2 2 1 1 1 2 3\n2 3 2 1 1 1 2 3 4\n", "mixing triangles and quadrilaterals")):
        msh = ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n6\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n"
               "5 0.5 0 0\n6 0.5 0.5 0\n$EndNodes\n$Elements\n%s$EndElements\n" % elems)
        (tmp_path / "q.msh").write_text(msh)
        case["model"]["refinement"] = 0
        case["model"]["filename"] = "q.msh"
        (tmp_path / "q.json").write_text(json.dumps(case))
        with pytest.raises(D.DgtdError, match=msg):
            D.Case(str(tmp_path / "q.json"))


QUADS2D = os.path.join(CASES, "2D_TFSF_Quads.json")


def test_2d_quad_case_elements_codes_and_time_step(tmp_path):
    """testData/maxwellInputs/2D_TFSF_Quads: Gmsh quadrilaterals (type 3) are
    the elements, lines the boundary; tags 1, 3, 5, 7 PEC, 4, 6 PMC and the
    plane wave's tag 2 the TF/SF line x = 1 (code 7 on both sides); the
    case's time step is taken as given."""
    c = D.Case(QUADS2D)
    assert c.dim == 2 and c.quads and c.K == 60 and c.EToV.shape == (60, 4) and c.Np == 16
    assert c.info["evolution_operator"] == "global" and c.time_step() == 0.01
    v = c.vx[c.EToV][:, :, :2]
    area2 = np.sum(v[:, :, 0] * np.roll(v[:, :, 1], -1, 1) - np.roll(v[:, :, 0], -1, 1) * v[:, :, 1], 1)
    assert (area2 > 0).all()
    np.testing.assert_allclose(area2.sum() / 2, 4.0, rtol=1e-14)  # [0,4] x [0,1]
    # element vertex sets as in the file
    with open(os.path.join(CASES, "2D_TFSF_Quads.msh")) as fh:
        lines = [ln.split() for ln in fh]
    i = next(k for k, ln in enumerate(lines) if ln == ["$Elements"])
    rows = [list(map(int, ln)) for ln in lines[i + 2: i + 2 + int(lines[i + 1][0])]]
    quads = sorted(tuple(sorted(r[5:9])) for r in rows if r[1] == 3)
    assert sorted(tuple(sorted(q + 1)) for q in c.EToV.tolist() for q in [np.array(q)]) == quads
    segs = {tuple(sorted((r[5] - 1, r[6] - 1))): r[4] for r in rows if r[1] == 1}
    code_of = {1: 1, 3: 1, 5: 1, 7: 1, 4: 2, 6: 2, 2: 7}
    seen = {}
    for e in range(c.K):
        for f in range(4):
            key = tuple(sorted((int(c.EToV[e, f]), int(c.EToV[e, (f + 1) % 4]))))
            if key in segs:
                assert c.face_bc[e, f] == code_of[segs[key]]
                seen[key] = seen.get(key, 0) + 1
            else:
                assert c.face_bc[e, f] == 0
    assert len(seen) == len(segs) == 42
    assert sum(n for k, n in seen.items() if segs[k] == 2) == 8  # 4 TF/SF edges, both sides
    assert np.linalg.norm(c.initial_state()) == 0.0
    w = c.planewaves[0]
    assert w.mean == pytest.approx(1.0 - 5 * 0.4 * np.sqrt(2), rel=1e-14)
    # a hesthaven variant of the case is refused (the reference disables it,
    # ExtensiveCasesTest.cpp:1159)
    case = json.load(open(QUADS2D))
    shutil.copy(os.path.join(CASES, "2D_TFSF_Quads.msh"), tmp_path / "2D_TFSF_Quads.msh")
    case["solver_options"]["evolution_operator"] = "hesthaven"
    (tmp_path / "h.json").write_text(json.dumps(case))
    with pytest.raises(D.DgtdError, match="Global operator only") as ei:
        D.Case(str(tmp_path / "h.json"))
    assert ei.value.code == -4


def test_2d_quad_snapshot_mesh_round_trip(tmp_path):
    c = D.Case(QUADS2D)
    c.save_mesh(tmp_path / "mesh.mesh")
    text = (tmp_path / "mesh.mesh").read_text()
    assert "dimension\n2\n" in text and "\n1 3 " in text  # attribute 1, geometry 3 (square)
    case = json.load(open(QUADS2D))
    case["model"]["filename"] = str(tmp_path / "mesh.mesh")
    (tmp_path / "c.json").write_text(json.dumps(case))
    c2 = D.Case(str(tmp_path / "c.json"))
    assert c2.quads
    np.testing.assert_array_equal(c2.EToV, c.EToV)
    np.testing.assert_allclose(c2.vx, c.vx, atol=1e-15)
    np.testing.assert_array_equal(c2.face_bc, c.face_bc)


def _mesh_rule(vx, EToV, order, cfl=1.0):
    """calcMeshTimeStep (Solver.cpp:290-300) restated: 0.75 * (2/3) * rmin *
    min vol / (surface / 2), rmin from the order's Gauss-Lobatto points."""
    from oracle import oracle2d as O2
    v = np.asarray(vx)[np.asarray(EToV)]
    if v.shape[1] == 4 and not np.asarray(vx)[:, 2].any():  # quadrilaterals
        d = v[:, :, :2]
        vol = 0.5 * np.abs(np.sum(d[:, :, 0] * np.roll(d[:, :, 1], -1, 1) - np.roll(d[:, :, 0], -1, 1) * d[:, :, 1], 1))
        surf = np.linalg.norm(np.roll(d, -1, 1) - d, axis=2).sum(1)
    elif v.shape[1] == 3:
        d = v[:, :, :2]
        e1, e2 = d[:, 1] - d[:, 0], d[:, 2] - d[:, 0]
        vol = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
        surf = np.linalg.norm(np.roll(d, -1, 1) - d, axis=2).sum(1)
    else:
        vol = np.abs(volumes(np.asarray(vx), np.asarray(EToV)))
        surf = sum(0.5 * np.linalg.norm(np.cross(v[:, b] - v[:, a], v[:, c] - v[:, a]), axis=1)
                   for a, b, c in MG.FACE_VERTS)
    g = O2.gll01(order)
    return cfl * 0.75 * (2 / 3) * 2 * (g[1] - g[0]) * np.min(vol / (surf / 2))


def test_automatic_time_step_follows_the_reference_rule():
    """estimateTimeStep (Solver.cpp:302-353): the mesh rule in 2D under any
    operator (halved on quadrilaterals) and in 3D under Maxwell (/0.8 under
    Global); the fscale rule only for hesthaven on tetrahedra."""
    vx, E, _ = D.cartesian_2d(14, 1, "quad")
    want = _mesh_rule(vx, E, 3) / 2
    for op in ("global", "hesthaven", "maxwell"):
        assert D.mesh_time_step(vx, E, 3, op) == pytest.approx(want, rel=1e-14)
    # hand check: cells 1/14 x 1, p = 3: rmin = 1 - 1/sqrt(5)
    h = (1 / 14) / ((2 / 14 + 2) / 2)
    assert want == pytest.approx(0.75 * (2 / 3) * (1 - 1 / np.sqrt(5)) * h / 2, rel=1e-14)
    vx, E, _ = D.cartesian_2d(5, 3, "triangle", sx=2.0)
    assert D.mesh_time_step(vx, E, 4, "global", cfl=0.5) == pytest.approx(_mesh_rule(vx, E, 4, 0.5), rel=1e-14)
    vx, E, _ = D.cube_mesh(2)
    base = _mesh_rule(vx, E, 2)
    assert D.mesh_time_step(vx, E, 2, "maxwell") == pytest.approx(base, rel=1e-14)
    assert D.mesh_time_step(vx, E, 2, "global") == pytest.approx(base / 0.8, rel=1e-14)
    # hesthaven: 1 / (max fscale p^2), fscale = (face area / 2) / (volume * 3 / 4)
    v = np.asarray(vx)[np.asarray(E)]
    fs = max(np.max(0.5 * np.linalg.norm(np.cross(v[:, b] - v[:, a], v[:, c] - v[:, a]), axis=1) / 2
                    / (np.abs(volumes(np.asarray(vx), np.asarray(E))) * 3 / 4)) for a, b, c in MG.FACE_VERTS)
    assert D.mesh_time_step(vx, E, 2, "hesthaven") == pytest.approx(1.0 / (fs * 4), rel=1e-13)
    # a case with time_step 0 takes the estimate of its mesh and operator
    c = D.Case(QUADS2D)
    assert c.time_step() == 0.01


def test_case_without_time_step_takes_the_estimate(tmp_path):
    for name in ("2D_TFSF_Quads", "2D_PEC", "3D_PEC"):
        case = json.load(open(os.path.join(CASES, name + ".json")))
        shutil.copy(os.path.join(CASES, name + ".msh"), tmp_path / (name + ".msh"))
        case["solver_options"]["time_step"] = 0.0
        case["solver_options"]["cfl"] = 0.7
        for op in ("global", "hesthaven"):
            if name.startswith("2D_TFSF_Quads") and op == "hesthaven":
                continue
            case["solver_options"]["evolution_operator"] = op
            (tmp_path / "c.json").write_text(json.dumps(case))
            c = D.Case(str(tmp_path / "c.json"))
            want = D.mesh_time_step(c.vx, c.EToV, c.order, op, cfl=0.7, quads=c.quads)
            assert c.time_step() == want > 0


def write_gmsh_hexes(path, nx, ny, nz, sx=1.0, sy=1.0, sz=1.0, mirror=(), x_planes=()):
    """A MakeCartesian3D-style hexahedral Gmsh 2.2 file: hexahedra (type 5,
    tag 1) and boundary quadrilaterals (type 3) tagged 1..6 as
    MakeCartesian3D's attributes (z=0, y=0, x=sx, y=sy, x=0, z=sz).
    Elements listed in `mirror` are written with a reversed (negative)
    orientation, which the reader must correct.  x_planes: (x, tag) pairs,
    interior quadrilaterals on the plane x tagged tag (written once, from the
    element on the plane's low side)."""
    vx, E, fbc = D.cartesian_3d(nx, ny, nz, "hex", sx=sx, sy=sy, sz=sz, bc=(1, 2, 3, 4, 5, 6))
    from oracle import oracle_hex as OH
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(vx))]
    lines += [f"{i + 1} {x:.17g} {y:.17g} {z:.17g}" for i, (x, y, z) in enumerate(vx)]
    lines += ["$EndNodes"]
    elems = []
    for e in range(E.shape[0]):
        for f in range(6):
            if fbc[e, f]:
                q = [int(E[e, j]) + 1 for j in OH.HEX_FACES[f]]
                elems.append(f"3 2 {fbc[e, f]} {fbc[e, f]} " + " ".join(map(str, q)))
            for xp, tag in x_planes:
                if f == 2 and np.allclose(vx[E[e, list(OH.HEX_FACES[f])], 0], xp):
                    q = [int(E[e, j]) + 1 for j in OH.HEX_FACES[f]]
                    elems.append(f"3 2 {tag} {tag} " + " ".join(map(str, q)))
    for e in range(E.shape[0]):
        v = [int(x) + 1 for x in E[e]]
        if e in mirror:
            v = [v[i] for i in (0, 3, 2, 1, 4, 7, 6, 5)]
        elems.append("5 2 1 1 " + " ".join(map(str, v)))
    lines += ["$Elements", str(len(elems))] + [f"{i + 1} {s}" for i, s in enumerate(elems)] + ["$EndElements"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return vx, E, fbc


def test_hexahedral_gmsh_case_and_mfem_round_trip(tmp_path):
    """Gmsh hexahedra (type 5) with quadrilateral boundary elements: elements
    in MFEM's vertex order (an inverted element is mirrored), face codes from
    the boundary quadrilaterals' tags on MFEM's hexahedron faces, the case's
    automatic time step (the mesh rule / 6 for hexahedra, / 0.8 for the
    Global operator), and the MFEM v1.0 snapshot mesh read back."""
    vx, E, fbc = write_gmsh_hexes(str(tmp_path / "h.msh"), 4, 2, 1, sx=2.0, mirror=(3,))
    bnd = [{"tags": [1, 6], "type": "PEC"}, {"tags": [2, 4], "type": "PMC"}, {"tags": [3, 5], "type": "SMA"}]
    p = write_case(str(tmp_path), "h.msh", bnd)
    c = D.Case(p)
    assert c.hexes and c.dim == 3 and c.K == 8 and c.EToV.shape == (8, 8) and c.Np == 27
    np.testing.assert_array_equal(c.EToV, E)  # the mirrored element restored to MFEM's order
    code = {1: 1, 6: 1, 2: 2, 4: 2, 3: 3, 5: 3}
    np.testing.assert_array_equal(c.face_bc, np.vectorize(lambda a: code.get(a, 0))(fbc))
    dt = c.time_step()
    assert dt == pytest.approx(D.mesh_time_step(c.vx, c.EToV, 2, "global", quads=False), rel=1e-14)
    # hand check: cells 0.5 x 0.5 x 1, p = 2 (rmin = 1): vol / (surface / 2) = 0.25 / 1.25
    assert dt == pytest.approx(0.75 * (2 / 3) * 1.0 * (0.25 / 1.25) / 6 / 0.8, rel=1e-14)
    c.save_mesh(tmp_path / "m.mesh")
    text = (tmp_path / "m.mesh").read_text()
    assert "\n1 5 " in text and "\n1 3 " in text  # cubes (geometry 5), squares (geometry 3)
    case = json.load(open(p))
    case["model"]["filename"] = str(tmp_path / "m.mesh")
    (tmp_path / "c2.json").write_text(json.dumps(case))
    c2 = D.Case(str(tmp_path / "c2.json"))
    np.testing.assert_array_equal(c2.EToV, c.EToV)
    np.testing.assert_array_equal(c2.face_bc, c.face_bc)
    # the hesthaven operator refuses hexahedra
    case["solver_options"]["evolution_operator"] = "hesthaven"
    (tmp_path / "c3.json").write_text(json.dumps(case))
    with pytest.raises(D.DgtdError, match="Global operator only") as ei:
        D.Case(str(tmp_path / "c3.json"))
    assert ei.value.code == -4


def test_hexahedral_case_uniform_refinement(tmp_path):
    """refinement: 1 splits each hexahedron 1:8 and each boundary
    quadrilateral 1:4 (MFEM's UniformRefinement): the 4 x 2 x 1 box becomes
    the 8 x 4 x 2 box, every child in MFEM's vertex order and with the
    boundary codes of the finer MakeCartesian3D mesh."""
    write_gmsh_hexes(str(tmp_path / "h.msh"), 4, 2, 1, sx=2.0, mirror=(3,))
    bnd = [{"tags": [1, 6], "type": "PEC"}, {"tags": [2, 4], "type": "PMC"}, {"tags": [3, 5], "type": "SMA"}]
    p = write_case(str(tmp_path), "h.msh", bnd)
    case = json.load(open(p))
    case["model"]["refinement"] = 1
    (tmp_path / "r.json").write_text(json.dumps(case))
    c = D.Case(str(tmp_path / "r.json"))
    assert c.hexes and c.K == 64
    vx, E, fbc = D.cartesian_3d(8, 4, 2, "hex", sx=2.0, bc=(1, 2, 3, 2, 3, 1))
    cent = lambda v, e: np.round(v[e].mean(axis=1), 12)
    want = {tuple(x): i for i, x in enumerate(cent(vx, E))}
    got = cent(c.vx, c.EToV)
    idx = np.array([want[tuple(x)] for x in got])
    assert sorted(idx) == list(range(64))
    # same corner coordinates in the same (MFEM) vertex order, same face codes
    np.testing.assert_allclose(c.vx[c.EToV], vx[E[idx]], atol=1e-14)
    np.testing.assert_array_equal(c.face_bc, fbc[idx])
    assert len(np.unique(np.round(c.vx, 12), axis=0)) == len(c.vx) == 9 * 5 * 3  # shared points merged


def test_hexahedral_tfsf_case(tmp_path):
    """A plane wave on hexahedra: the TF/SF surface is an interior
    quadrilateral surface (tag 7, plane x = 1) of a Gmsh hexahedral guide;
    both sides of each of its faces carry code 7, and the automatic delay
    starts the pulse 5 sqrt(2) spreads before the plane (driver.cpp:575-588)."""
    write_gmsh_hexes(str(tmp_path / "g.msh"), 16, 1, 1, sx=4.0, sy=0.5, sz=0.5, x_planes=[(1.0, 7)])
    case = {"solver_options": {"time_step": 0.0, "final_time": 2.2, "order": 3},
            "model": {"filename": "g.msh", "materials": [{"tags": [1], "type": "vacuum"}],
                      "boundaries": [{"tags": [1, 6], "type": "PEC"}, {"tags": [2, 4], "type": "PMC"},
                                     {"tags": [3, 5], "type": "SMA"}]},
            "sources": [{"type": "planewave", "polarization": [0.0, 0.0, 1.0], "propagation": [1.0, 0.0, 0.0],
                         "tags": [7], "magnitude": {"type": "gaussian", "spread": 0.3}}]}
    (tmp_path / "t.json").write_text(json.dumps(case))
    c = D.Case(str(tmp_path / "t.json"))
    assert c.hexes and c.K == 16
    (w,) = c.planewaves
    assert w.mean == pytest.approx(1.0 - 5.0 * 0.3 * np.sqrt(2.0), rel=1e-14)
    tf = np.argwhere(c.face_bc == 7)
    assert len(tf) == 2
    from oracle import oracle_hex as OH
    for e, f in tf:
        np.testing.assert_allclose(c.vx[c.EToV[e, list(OH.HEX_FACES[f])], 0], 1.0)
    np.testing.assert_array_equal(c.face_bc[:, 2][c.face_bc[:, 2] != 7], [0] * 14 + [3])


def test_case_kernel_choice_counts_interior_sma(tmp_path):
    """kernel "auto" takes the assembled operator where the Global operator's
    SMA weighting differs from the matrix-free one: SMA faces at alpha != 1,
    interior SMA (code 6) as well as boundary SMA, as the launcher decides."""
    shutil.copy(os.path.join(CASES, "2D_TFSF_IntBoundary.msh"), tmp_path / "2D_TFSF_IntBoundary.msh")
    case = json.load(open(os.path.join(CASES, "2D_TFSF_IntBoundary.json")))

    def variant(name, interior, alpha, op="global"):
        case["model"]["boundaries"] = [{"tags": [18], "type": interior},
                                       {"tags": [13, 15, 16, 17, 19, 20, 21, 22], "type": "PEC"}]
        case["solver_options"].update(upwind_alpha=alpha, evolution_operator=op)
        (tmp_path / name).write_text(json.dumps(case))
        return D.Case(str(tmp_path / name))

    c = variant("a.json", "SMA", 0.5)
    assert (c.face_bc == 6).any() and not (c.face_bc == 3).any()
    assert c.needs_assembled_operator()
    assert not variant("b.json", "PEC", 0.5).needs_assembled_operator()
    assert not variant("c.json", "SMA", 1.0).needs_assembled_operator()
    assert not variant("d.json", "SMA", 0.5, "hesthaven").needs_assembled_operator()
    q = D.Case(os.path.join(CASES, "2D_TFSF_Quads.json"))
    assert q.quads and q.needs_assembled_operator()
    # hexahedra: the matrix-free stage carries the Global operator's flux
    write_gmsh_hexes(str(tmp_path / "h.msh"), 2, 1, 1)
    h = json.loads(json.dumps(case))
    h["model"].update(filename="h.msh", boundaries=[{"tags": [1, 2, 3, 4, 5, 6], "type": "SMA"}])
    h["sources"] = []
    h["solver_options"].update(evolution_operator="global", upwind_alpha=0.5)
    (tmp_path / "h.json").write_text(json.dumps(h))
    hc = D.Case(str(tmp_path / "h.json"))
    assert hc.hexes and not hc.needs_assembled_operator()
