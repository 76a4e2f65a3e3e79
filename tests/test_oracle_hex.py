"""Pins of the hexahedron oracle (oracle/oracle_hex.py) and the product's host
set-up on hexahedra.

The reference runs hexahedral meshes on its Global operator
(Solver3DTest.cpp:57-200); its tests hold no hexahedral matrices.  The oracle
is pinned by closed forms: MFEM's lexicographic node order on Gauss-Lobatto
points, polynomial exactness of the derivatives, the mass and the face
masses, face node lists that follow MFEM's face cycles, the divergence
theorem per element, the exact curl of polynomial fields on a sheared mesh,
centred energy conservation and upwind dissipation, and convergence to the
TE_110 cavity mode.  The product's host matrix equals the oracle's Mult
(tests/test_global_assembly.py).
"""
import numpy as np
import pytest

import dgtd_amd as D
from oracle import oracle_hex as OH

import meshgen as G


def _m1(n):
    return 0.0 if n % 2 else 2.0 / (n + 1)


def sheared_hexes(nx, ny, nz, bc="pec", seed=0, sx=1.0):
    """A MakeCartesian3D-style hexahedral mesh mapped by a fixed shear (every
    element stays a parallelepiped)."""
    vx, E, fbc = D.cartesian_3d(nx, ny, nz, "hex", sx=sx, bc=bc)
    vx = vx.copy()
    vx[:, 0] += 0.3 * vx[:, 1] + 0.1 * vx[:, 2]
    vx[:, 1] += 0.2 * vx[:, 2]
    return vx, E, fbc


def test_lexicographic_gauss_lobatto_nodes():
    for p in (1, 2, 3, 4):
        R = OH.RefHex(p)
        g = 2 * OH.gll01(p) - 1
        n1 = p + 1
        r = 2 * R.xi - 1
        for k in range(n1):
            for j in range(n1):
                for i in range(n1):
                    np.testing.assert_allclose(r[i + n1 * (j + n1 * k)], (g[i], g[j], g[k]), atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_reference_matrices_exact(p):
    R = OH.RefHex(p)
    r, s, t = (2 * R.xi[:, d] - 1 for d in range(3))
    mono = [(a, b, c) for a in range(p + 1) for b in range(p + 1) for c in range(p + 1)]
    for a, b, c in mono:
        u = r ** a * s ** b * t ** c
        np.testing.assert_allclose(R.Dr @ u, a * r ** max(a - 1, 0) * s ** b * t ** c if a else 0 * u, atol=1e-10)
        np.testing.assert_allclose(R.Ds @ u, b * r ** a * s ** max(b - 1, 0) * t ** c if b else 0 * u, atol=1e-10)
        np.testing.assert_allclose(R.Dt @ u, c * r ** a * s ** b * t ** max(c - 1, 0) if c else 0 * u, atol=1e-10)
    U = np.stack([r ** a * s ** b * t ** c for a, b, c in mono], axis=1)
    want = np.array([[_m1(a + d) * _m1(b + e) * _m1(c + f) for d, e, f in mono] for a, b, c in mono])
    np.testing.assert_allclose(U.T @ R.M @ U, want, atol=1e-11)
    g = 2 * OH.gll01(p) - 1
    # a face's nodes (m + (p+1) n) carry (u_m, u_n): face mass integrates products
    n1 = p + 1
    fm = [(a, b) for a in range(n1) for b in range(n1)]
    T = np.stack([np.repeat(g[None, :], n1, 0).ravel() ** a * np.repeat(g[:, None], n1, 1).ravel() ** b
                  for a, b in fm], axis=1)
    want2 = np.array([[_m1(a + c) * _m1(b + d) for c, d in fm] for a, b in fm])
    for f in range(6):
        Eb = R.E[np.ix_(R.Fmask[f], f * R.Nfp + np.arange(R.Nfp))]
        np.testing.assert_allclose(T.T @ Eb @ T, want2, atol=1e-11)
    np.testing.assert_allclose(R.M @ R.LIFT, R.E, atol=1e-11)


def test_face_node_lists_follow_the_face_cycles():
    for p in (1, 3):
        R = OH.RefHex(p)
        n1 = p + 1
        for f, cyc in enumerate(OH.HEX_FACES):
            pts = R.xi[R.Fmask[f]]
            A, B, Dv = (OH.HEX_CORNERS[cyc[q]] for q in (0, 1, 3))
            np.testing.assert_allclose(pts[0], A, atol=1e-14)
            np.testing.assert_allclose(pts[p], B, atol=1e-14)          # m runs from a to b
            np.testing.assert_allclose(pts[n1 * p], Dv, atol=1e-14)     # n runs from a to d
            # outward normal (b - a) x (d - a) on the unit cube
            c = np.cross(B - A, Dv - A)
            centre = pts.mean(axis=0) - 0.5
            assert c @ centre > 0


@pytest.mark.parametrize("p", [2, 3])
def test_divergence_theorem_per_element(p):
    vx, E, fbc = sheared_hexes(2, 2, 1, seed=p)
    op = OH.OperatorHex(vx, E, fbc, p)
    R = op.R
    u = np.sin(np.arange(op.K * R.Np) * 0.37).reshape(op.K, R.Np)
    ur, us, ut = (np.einsum("ij,kj->ki", Dk, u) for Dk in (R.Dr, R.Ds, R.Dt))
    for d in range(3):
        du = op.G[:, 0, d, None] * ur + op.G[:, 1, d, None] * us + op.G[:, 2, d, None] * ut
        vol = op.J * np.einsum("i,ij,kj->k", np.ones(R.Np), R.M, du)
        surf = np.zeros(op.K)
        for f in range(6):
            Eb = R.E[np.ix_(R.Fmask[f], f * R.Nfp + np.arange(R.Nfp))]
            surf += (op.fscale[:, f] * op.J * op.nrm[:, f, d]
                     * np.einsum("i,ij,kj->k", np.ones(R.Nfp), Eb, u[:, R.Fmask[f]]))
        np.testing.assert_allclose(vol, surf, atol=1e-11)
    v = np.asarray(vx)[E]
    a = np.stack([v[:, 1] - v[:, 0], v[:, 3] - v[:, 0], v[:, 4] - v[:, 0]], axis=1)
    np.testing.assert_allclose(op.J * 8, np.abs(np.linalg.det(a)), rtol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_polynomial_fields_give_exact_curl(p):
    """Fields of total degree <= p: zero jumps with no boundary condition
    ("none"), so the operator is the exact curl (random eps, mu)."""
    vx, E, fbc = sheared_hexes(2, 2, 2, bc="none")
    rng = np.random.default_rng(p)
    eps, mu = rng.uniform(1, 3, E.shape[0]), rng.uniform(1, 2, E.shape[0])
    for alpha in (0.0, 1.0):
        op = OH.OperatorHex(vx, E, fbc, p, alpha=alpha, eps=eps, mu=mu)
        x, y, z = (op.pos[..., d] for d in range(3))
        c = rng.standard_normal((6, 4))
        f = [c[k, 0] * x ** p + c[k, 1] * y ** (p - 1) * z + c[k, 2] * z + c[k, 3] * x * y ** (p - 1) for k in range(6)]
        d1 = lambda k: [c[k, 0] * p * x ** (p - 1) + c[k, 3] * y ** (p - 1),
                        c[k, 1] * (p - 1) * y ** max(p - 2, 0) * z * (p > 1)
                        + c[k, 3] * (p - 1) * x * y ** max(p - 2, 0) * (p > 1),
                        c[k, 1] * y ** (p - 1) + c[k, 2]]
        grads = [d1(k) for k in range(6)]
        out = op.mult(np.stack(f).reshape(-1)).reshape(6, op.K, op.Np)
        for xx in range(3):
            yy, zz = (xx + 1) % 3, (xx + 2) % 3
            curlH = grads[3 + zz][yy] - grads[3 + yy][zz]
            curlE = grads[zz][yy] - grads[yy][zz]
            np.testing.assert_allclose(out[xx], curlH / eps[:, None], atol=1e-9)
            np.testing.assert_allclose(out[3 + xx], -curlE / mu[:, None], atol=1e-9)


def test_centred_flux_conserves_energy_upwind_dissipates():
    vx, E, fbc = sheared_hexes(2, 2, 2, bc=("pec", "pmc", "pec", "pmc", "pec", "pec"))
    x0, en = None, {}
    for alpha in (0.0, 1.0):
        op = OH.OperatorHex(vx, E, fbc, 2, alpha=alpha)
        if x0 is None:
            c = op.pos - op.pos.reshape(-1, 3).mean(axis=0)
            g = np.exp(-(c ** 2).sum(-1) / (2 * 0.3 ** 2))
            x0 = np.zeros((6,) + g.shape)
            x0[2], x0[3], x0[1] = g, 0.4 * g, -0.2 * g
            x0 = x0.reshape(-1)
        x, _, _ = op.run(x0, 0.2 / (op.max_fscale() * 4), 0.2, "lserk4")
        en[alpha] = (op.energy(x0), op.energy(x))
    assert abs(en[0.0][1] - en[0.0][0]) / en[0.0][0] < 1e-7
    assert en[1.0][1] < en[1.0][0] * (1 - 1e-5)


def test_te110_mode_converges():
    errs = []
    for n in (1, 2):
        vx, E, fbc = D.cartesian_3d(n, n, n, "hex", bc="pec")
        op = OH.OperatorHex(vx, E, fbc, 3)
        pos = np.moveaxis(op.pos, -1, 0)
        x0 = G.te110(pos)
        x, _, t = op.run(x0, 0.25 / (op.max_fscale() * 9), 0.2, "lserk4")
        ex = G.te110_exact(pos, t)
        errs.append(np.linalg.norm(x - ex) / np.linalg.norm(ex))
    assert errs[0] < 0.15 and errs[1] < errs[0] / 12  # order p + 1 = 4 in h (0.105 -> 0.0067)


def test_product_host_setup_matches_oracle():
    vx, E, fbc = sheared_hexes(2, 2, 1)
    for p in (1, 3):
        op = OH.OperatorHex(vx, E, fbc, p)
        src = [D.resonant((2, 1, 1), field="E", polarization=(0, 0, 1))]
        x = D.project_initial_fields(vx, E, p, src).reshape(6, op.K, op.Np)
        want = np.prod([np.sin(m * np.pi * op.pos[..., d]) for d, m in enumerate((2, 1, 1))], axis=0)
        np.testing.assert_allclose(x[2], want, atol=1e-13)
        rng = np.random.default_rng(p)
        q = rng.uniform(0.05, 0.95, (15, 3))  # unit-cube points mapped by the mesh's affine map
        pts = np.stack([q[:, 0] + 0.3 * q[:, 1] + 0.1 * q[:, 2], q[:, 1] + 0.2 * q[:, 2], q[:, 2]], axis=1)
        pts[:, 0] *= 1.0
        elem, w = D.mesh_locate(vx, E, p, pts)
        assert (elem >= 0).all()
        u = op.pos[..., 0] ** p + 2 * op.pos[..., 1] * op.pos[..., 2] ** (p - 1)
        np.testing.assert_allclose(np.einsum("pn,pn->p", w, u[elem]), pts[:, 0] ** p + 2 * pts[:, 1] * pts[:, 2] ** (p - 1),
                                   atol=1e-11)


def test_non_parallelepiped_refused():
    vx, E, fbc = D.cartesian_3d(2, 1, 1, "hex")
    vx = vx.copy()
    vx[E[0, 6], 2] += 0.1
    with pytest.raises(ValueError):
        OH.OperatorHex(vx, E, fbc, 2)
    with pytest.raises(D.DgtdError, match="parallelepipeds"):
        D.global_assemble(vx, E, fbc, 2)
