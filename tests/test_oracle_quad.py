"""Pins of the quadrilateral oracle (oracle/oracle2d.py RefQuad and
Operator2D(quads=True)) and the product's host set-up on quadrilaterals.

The reference runs quadrilateral meshes on its Global operator, which
assembles the DG operator of the L2 Gauss-Lobatto space
(ExtensiveCasesTest.cpp:1003,1082; Solver2DTest.cpp:143-246); its own tests
hold no quadrilateral matrices.  The oracle is pinned by closed forms: the
Gauss-Lobatto points of orders 2-5, MFEM's lexicographic node order,
polynomial exactness of the derivatives, the mass and the edge mass, the
divergence theorem per element, the exact curl of polynomial fields on a
sheared mesh, energy conservation with the centred flux and dissipation with
the upwind one, and convergence to the TM_11 cavity mode.  A transposed
metric, a wrong normal or fscale, a reversed edge or a dropped flux term fails
one of them.
"""
import numpy as np
import pytest

import dgtd_amd as D
from oracle import oracle2d as O

import meshgen as G


def _m1(n):
    """int_{-1}^{1} t^n dt."""
    return 0.0 if n % 2 else 2.0 / (n + 1)


def test_gll_points_and_lexicographic_nodes():
    s5 = np.sqrt(1 / 3 - 2 * np.sqrt(7) / 21), np.sqrt(1 / 3 + 2 * np.sqrt(7) / 21)
    closed = {1: [-1, 1], 2: [-1, 0, 1], 3: [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1],
              4: [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1],
              5: [-1, -s5[1], -s5[0], s5[0], s5[1], 1]}
    for p, x in closed.items():
        R = O.RefQuad(p)
        r = 2 * R.xi - 1
        n1 = p + 1
        for j in range(n1):
            for i in range(n1):
                np.testing.assert_allclose(r[i + n1 * j], (x[i], x[j]), atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_reference_matrices_exact(p):
    """Dr, Ds exact on Q_p; M and the edge masses integrate products exactly;
    LIFT = M^-1 E."""
    R = O.RefQuad(p)
    r, s = 2 * R.xi[:, 0] - 1, 2 * R.xi[:, 1] - 1
    for a in range(p + 1):
        for b in range(p + 1):
            u = r ** a * s ** b
            np.testing.assert_allclose(R.Dr @ u, a * r ** max(a - 1, 0) * s ** b if a else 0 * u, atol=1e-11)
            np.testing.assert_allclose(R.Ds @ u, b * r ** a * s ** max(b - 1, 0) if b else 0 * u, atol=1e-11)
    mono = [(a, b) for a in range(p + 1) for b in range(p + 1)]
    U = np.stack([r ** a * s ** b for a, b in mono], axis=1)
    want = np.array([[_m1(a + c) * _m1(b + d) for c, d in mono] for a, b in mono])
    np.testing.assert_allclose(U.T @ R.M @ U, want, atol=1e-12)
    g = 2 * O.gll01(p) - 1
    T = np.stack([g ** k for k in range(p + 1)], axis=1)
    w1 = np.array([[_m1(k + l) for l in range(p + 1)] for k in range(p + 1)])
    for f in range(4):
        Eb = R.E[np.ix_(R.Fmask[f], f * R.Nfp + np.arange(R.Nfp))]
        np.testing.assert_allclose(T.T @ Eb @ T, w1, atol=1e-12)
    np.testing.assert_allclose(R.M @ R.LIFT, R.E, atol=1e-12)


def test_edge_node_lists_follow_edges():
    """Fmask[f] runs from vertex f to vertex f+1 of (-1,-1), (1,-1), (1,1), (-1,1)."""
    corners = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1.0]])
    for p in (1, 3):
        R = O.RefQuad(p)
        r = 2 * R.xi - 1
        for f, (a, b) in enumerate(O.RefQuad.EDGES):
            pts = r[R.Fmask[f]]
            np.testing.assert_allclose(pts[0], corners[a], atol=1e-14)
            np.testing.assert_allclose(pts[-1], corners[b], atol=1e-14)
            t = (pts - corners[a]) @ (corners[b] - corners[a]) / 4.0
            assert np.all(np.diff(t) > 0)


@pytest.mark.parametrize("p", [2, 4])
def test_divergence_theorem_per_element(p):
    """int_K du/dx dA = sum_f int_f u n_x ds for any nodal field: the metric,
    J, the normals and fscale together."""
    vx, E, bc = G.quad_mesh(3, 2, Lx=2.0, bc="pec", stretch=0.4, shear=0.35, seed=3)
    op = O.Operator2D(vx, E, bc, p, quads=True)
    R = op.R
    u = np.sin(np.arange(op.K * R.Np) * 0.7).reshape(op.K, R.Np)
    ur = np.einsum("ij,kj->ki", R.Dr, u)
    us = np.einsum("ij,kj->ki", R.Ds, u)
    for comp, (gr, gs) in enumerate([(op.rx, op.sx), (op.ry, op.sy)]):
        du = gr[:, None] * ur + gs[:, None] * us
        vol = np.abs(op.J) * np.einsum("i,ij,kj->k", np.ones(R.Np), R.M, du)
        surf = np.zeros(op.K)
        for f in range(4):
            Eb = R.E[np.ix_(R.Fmask[f], f * R.Nfp + np.arange(R.Nfp))]
            surf += (op.fscale[:, f] * np.abs(op.J) * op.nrm[:, f, comp]
                     * np.einsum("i,ij,kj->k", np.ones(R.Nfp), Eb, u[:, R.Fmask[f]]))
        np.testing.assert_allclose(vol, surf, atol=1e-11)
    # the parallelogram's area is 4 |J| and its edges 2 fscale |J|
    area = np.abs(op.J) * 4
    v = np.asarray(vx)[:, :2][E]
    shoelace = 0.5 * np.abs(np.sum(v[:, :, 0] * np.roll(v[:, :, 1], -1, 1) - np.roll(v[:, :, 0], -1, 1) * v[:, :, 1], 1))
    np.testing.assert_allclose(area, shoelace, rtol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3, 5])
def test_polynomial_fields_give_exact_curl(p):
    """Fields of total degree <= p are continuous across edges, so with no
    boundary condition ("none": zero jump) the operator is the exact curl:
    dE/dt = curl H / eps, dH/dt = -curl E / mu with d/dz = 0."""
    vx, E, bc = G.quad_mesh(3, 3, bc="none", stretch=0.3, shear=-0.4, seed=p)
    rng = np.random.default_rng(p)
    eps, mu = rng.uniform(1, 3, E.shape[0]), rng.uniform(1, 2, E.shape[0])
    for alpha in (0.0, 1.0):
        op = O.Operator2D(vx, E, bc, p, alpha=alpha, eps=eps, mu=mu, quads=True)
        x, y = op.pos[..., 0], op.pos[..., 1]
        c = rng.standard_normal((6, 3))
        f = [c[k, 0] * x ** p + c[k, 1] * x * y ** (p - 1) + c[k, 2] * y for k in range(6)]
        dfx = [c[k, 0] * p * x ** (p - 1) + c[k, 1] * y ** (p - 1) for k in range(6)]
        dfy = [c[k, 1] * (p - 1) * x * y ** max(p - 2, 0) * (p > 1) + c[k, 2] for k in range(6)]
        out = op.mult(np.stack(f).reshape(-1)).reshape(6, op.K, op.Np)
        Ex, Ey, Ez, Hx, Hy, Hz = range(6)
        curlH = [dfy[Hz], -dfx[Hz], dfx[Hy] - dfy[Hx]]
        curlE = [dfy[Ez], -dfx[Ez], dfx[Ey] - dfy[Ex]]
        for k in range(3):
            np.testing.assert_allclose(out[k], curlH[k] / eps[:, None], atol=1e-9)
            np.testing.assert_allclose(out[3 + k], -curlE[k] / mu[:, None], atol=1e-9)


def test_centred_flux_conserves_energy_upwind_dissipates():
    vx, E, bc = G.quad_mesh(3, 3, bc=["pec", "pmc", "pec", "pmc"], stretch=0.2, shear=0.2, seed=1)
    x0, energies = None, {}
    for alpha in (0.0, 1.0):
        op = O.Operator2D(vx, E, bc, 3, alpha=alpha, quads=True)
        if x0 is None:
            r2 = (op.pos[..., 0] - 0.5) ** 2 + (op.pos[..., 1] - 0.55) ** 2
            g = np.exp(-r2 / (2 * 0.15 ** 2))
            x0 = np.zeros((6,) + g.shape)
            x0[2], x0[3], x0[4] = g, 0.5 * g, -0.3 * g
            x0 = x0.reshape(-1)
        x, _, _ = op.run(x0, 0.2 / (op.max_fscale() * 9), 0.3, "lserk4")
        energies[alpha] = (op.energy(x0), op.energy(x))
    e0, e1 = energies[0.0]
    assert abs(e1 - e0) / e0 < 1e-7
    e0, e1 = energies[1.0]
    assert e1 < e0 * (1 - 1e-5)


def test_tm11_mode_converges():
    errs = []
    for n in (2, 4):
        vx, E, bc = G.quad_mesh(n, n, bc="pec")
        op = O.Operator2D(vx, E, bc, 3, quads=True)
        x0 = G.tm_mode(op.pos, 0.0).reshape(-1)
        x, _, t = op.run(x0, 0.25 / (op.max_fscale() * 9), 0.3, "lserk4")
        ex = G.tm_mode(op.pos, t).reshape(-1)
        errs.append(np.linalg.norm(x - ex) / np.linalg.norm(ex))
    assert errs[0] < 0.02 and errs[1] < errs[0] / 12  # order p + 1 = 4 in h


def test_product_host_setup_matches_oracle():
    """Node positions (through the initial-field projection) and point
    location with interpolation weights, on a sheared stretched mesh."""
    vx, E, bc = G.quad_mesh(3, 2, bc="pec", stretch=0.3, shear=0.25, seed=5)
    for p in (1, 3, 5):
        op = O.Operator2D(vx, E, bc, p, quads=True)
        src = [D.resonant((2, 1), field="E", polarization=(0, 0, 1))]
        x = D.project_initial_fields(vx, E, p, src, quads=True).reshape(6, op.K, op.Np)
        want = np.sin(2 * np.pi * op.pos[..., 0]) * np.sin(np.pi * op.pos[..., 1])
        np.testing.assert_allclose(x[2], want, atol=1e-13)
        assert np.abs(x[[0, 1, 3, 4, 5]]).max() == 0.0
        rng = np.random.default_rng(p)
        yy = rng.uniform(0, 1, 20)
        pts = np.column_stack([rng.uniform(0, 1, 20) + 0.25 * yy, yy, np.zeros(20)])
        elem, w = D.mesh_locate(vx, E, p, pts, quads=True)
        assert (elem >= 0).all()
        u = op.pos[..., 0] ** p + 2 * op.pos[..., 0] * op.pos[..., 1] ** (p - 1)
        np.testing.assert_allclose(np.einsum("pn,pn->p", w, u[elem]),
                                   pts[:, 0] ** p + 2 * pts[:, 0] * pts[:, 1] ** (p - 1), atol=1e-11)
        # the oracle's own interpolation weights agree
        for k in range(3):
            e = elem[k]
            v0, v1, v3 = (np.asarray(vx)[E[e, j], :2] for j in (0, 1, 3))
            xi, eta = np.linalg.solve(np.column_stack([v1 - v0, v3 - v0]), pts[k, :2] - v0)
            np.testing.assert_allclose(w[k], op.R.interp_weights(xi, eta), atol=1e-12)


def test_non_parallelogram_refused():
    vx, E, bc = G.quad_mesh(2, 1, bc="pec")
    vx = vx.copy()
    vx[E[0, 2], 1] += 0.1  # moves a corner shared by both elements off the parallelogram
    with pytest.raises(ValueError):
        O.Operator2D(vx, E, bc, 2, quads=True)
    with pytest.raises(D.DgtdError):
        D.global_assemble(vx, E, bc, 2, quads=True)
