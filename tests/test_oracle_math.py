"""Closed forms, invariants and an independently assembled operator that pin
the oracle where the reference's golden vectors do not reach (orders >= 3,
the full Mult, the integrators).

A plausible mistake -- a dropped term, a flipped sign, a wrong index or a
transposed operand -- breaks at least one of: exact polynomial integration
and differentiation, the face-mass closed form, the brute-force operator,
semi-discrete energy conservation (centred flux) / dissipation (upwind),
invariance under the local vertex order, and the convergence orders.
"""
import math

import numpy as np
import pytest

import meshgen as MG
from oracle import oracle as O

# ---------------------------------------------------------------- helpers


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def tet_quadrature(n):
    """Collapsed (Duffy) Gauss rule on the unit tetrahedron."""
    u, wu = gauss01(n)
    pts, wts = [], []
    for a, wa in zip(u, wu):
        for b, wb in zip(u, wu):
            for c, wc in zip(u, wu):
                pts.append((a, (1 - a) * b, (1 - a) * (1 - b) * c))
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.array(pts), np.array(wts)


def tri_quadrature(n):
    u, wu = gauss01(n)
    pts, wts = [], []
    for a, wa in zip(u, wu):
        for b, wb in zip(u, wu):
            pts.append((a, (1 - a) * b))
            wts.append(wa * wb * (1 - a))
    return np.array(pts), np.array(wts)


def monomials(p):
    return [(a, b, c) for a in range(p + 1) for b in range(p + 1 - a) for c in range(p + 1 - a - b)]


def unit_tet_monomial_integral(a, b, c):
    """int over the unit tetrahedron of xi^a eta^b zeta^c = a! b! c! / (a+b+c+3)!"""
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


# ---------------------------------------------------------------- 1D nodes


@pytest.mark.parametrize("p", range(1, 9))
def test_gll_points_are_legendre_derivative_roots(p):
    x = O.gll01(p)
    assert x[0] == 0.0 and x[-1] == 1.0
    interior = np.sort(np.polynomial.legendre.Legendre.basis(p).deriv().roots().real)
    np.testing.assert_allclose(2 * x[1:-1] - 1, interior, atol=1e-14)


@pytest.mark.parametrize("p", range(1, 6))
def test_node_set_symmetry_and_edges(p):
    """The MFEM GaussLobatto tet node set is invariant under vertex
    permutations, and its edge nodes are the 1D GLL points."""
    R = O.ref_element(p)
    xi = (R["rst"] + 1) / 2
    lam = np.column_stack([1 - xi.sum(1), xi])
    key = lambda L: {tuple(np.round(r, 12)) for r in L}
    base = key(lam)
    for perm in [(1, 0, 2, 3), (0, 2, 1, 3), (3, 1, 2, 0), (2, 3, 0, 1)]:
        assert key(lam[:, perm]) == base
    on_edge = lam[(lam[:, 2] < 1e-14) & (lam[:, 3] < 1e-14)]
    np.testing.assert_allclose(np.sort(on_edge[:, 1]), O.gll01(p), atol=1e-14)


# ---------------------------------------------------------------- reference matrices


@pytest.mark.parametrize("p", range(1, 6))
def test_mass_integrates_polynomials_exactly(p):
    """u^T M v = int_T u v over the Hesthaven tetrahedron (8 x the unit tet)
    for monomials of degree <= p."""
    R = O.ref_element(p)
    xi = (R["rst"] + 1) / 2
    mons = monomials(p)
    rng = np.random.default_rng(p)
    for _ in range(4):
        m1, m2 = mons[rng.integers(len(mons))], mons[rng.integers(len(mons))]
        u = np.prod(xi ** np.array(m1), axis=1)
        v = np.prod(xi ** np.array(m2), axis=1)
        exact = 8.0 * unit_tet_monomial_integral(*(np.array(m1) + np.array(m2)))
        assert u @ R["M"] @ v == pytest.approx(exact, rel=1e-11)
    assert R["M"].sum() == pytest.approx(4.0 / 3.0, rel=1e-13)
    assert np.all(np.linalg.eigvalsh(R["M"]) > 0)


@pytest.mark.parametrize("p", range(1, 6))
def test_derivatives_exact_on_polynomials(p):
    R = O.ref_element(p)
    r, s, t = R["rst"].T
    for (a, b, c) in monomials(p):
        f = r ** a * s ** b * t ** c
        df = [a * r ** max(a - 1, 0) * s ** b * t ** c if a else 0 * r,
              b * r ** a * s ** max(b - 1, 0) * t ** c if b else 0 * r,
              c * r ** a * s ** b * t ** max(c - 1, 0) if c else 0 * r]
        for D, d in zip((R["Dr"], R["Ds"], R["Dt"]), df):
            np.testing.assert_allclose(D @ f, d, atol=1e-11 * max(1.0, np.abs(d).max()))


@pytest.mark.parametrize("p", range(1, 6))
def test_emat_is_exact_face_mass_on_area_two(p):
    """Emat block f: u^T E_f v = (2/A_f) int_{face f} u v dA (every face mapped
    to a reference face of area 2, HesthavenEvolution.cpp:56-81) and LIFT = M^-1 E."""
    R = O.ref_element(p)
    Nfp = (p + 1) * (p + 2) // 2
    xi = (R["rst"] + 1) / 2
    V = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [0, 0, 2]], float)
    X = -1 + 2 * xi  # nodes on the Hesthaven tet shifted to the origin corner
    fv = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]
    qp, qw = tri_quadrature(p + 2)
    mons = monomials(p)
    rng = np.random.default_rng(10 + p)
    for f in range(4):
        a, b, c = (V[k] - 1 for k in fv[f])
        area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        fm = R["Fmask"][f * Nfp:(f + 1) * Nfp]
        E = R["Emat"][:, f * Nfp:(f + 1) * Nfp]
        # the columns of the face block are zero off the face rows
        off = np.setdiff1d(np.arange(len(xi)), fm)
        assert np.abs(E[off]).max() < 1e-13
        for _ in range(3):
            m1, m2 = mons[rng.integers(len(mons))], mons[rng.integers(len(mons))]
            u = np.prod((X[fm] + 1) ** np.array(m1), axis=1)
            v = np.prod((X[fm] + 1) ** np.array(m2), axis=1)
            pts = a + np.outer(qp[:, 0], b - a) + np.outer(qp[:, 1], c - a)
            integ = 2 * area * np.sum(qw * np.prod((pts + 1) ** np.array(m1), 1) * np.prod((pts + 1) ** np.array(m2), 1))
            assert u @ E[fm] @ v == pytest.approx(2.0 / area * integ, rel=1e-11, abs=1e-13)
    np.testing.assert_allclose(R["M"] @ R["LIFT"], R["Emat"], atol=1e-12)


# ---------------------------------------------------------------- geometry


def test_geometry_closed_forms():
    rng = np.random.default_rng(3)
    vx = rng.standard_normal((4, 3))
    E = np.array([[0, 1, 2, 3]], dtype=np.int32)
    E = MG.fix_orientation(vx, E)
    op = O.Operator(vx, E, np.ones((1, 4), dtype=np.int32), 2)
    vol, fs, nr = op.geometry()
    v = vx[E[0]]
    V = abs(np.linalg.det(np.array([v[1] - v[0], v[2] - v[0], v[3] - v[0]]))) / 6
    assert vol[0] == pytest.approx(V, rel=1e-13)
    fv = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]
    s = np.zeros(3)
    for f in range(4):
        a, b, c = v[list(fv[f])]
        cr = np.cross(b - a, c - a)
        A = 0.5 * np.linalg.norm(cr)
        n = cr / np.linalg.norm(cr)
        if np.dot(n, a - v[f]) < 0:
            n = -n  # outward: away from the opposite vertex
        np.testing.assert_allclose(nr[0, f], n, atol=1e-13)
        assert fs[0, f] == pytest.approx(2 * A / (3 * V), rel=1e-13)
        s += A * nr[0, f]
    np.testing.assert_allclose(s, 0, atol=1e-13)


def test_max_fscale_on_kuhn_cube():
    """Kuhn split of n^3 hexes: max fscale = 2 sqrt(2) n (the time-step rule's input)."""
    for n in (2, 3):
        vx, E, fbc = MG.cube_mesh(n)
        op = O.Operator(vx, E, fbc, 1)
        assert op.max_fscale() == pytest.approx(2 * math.sqrt(2) * n, rel=1e-12)


def test_connectivity_pairs_colocated_nodes():
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    op = O.Operator(vx, E, fbc, 3)
    M, P = op.connectivity()
    pos = op.positions().reshape(3, -1)
    np.testing.assert_allclose(pos[:, M], pos[:, P], atol=1e-13)
    interior = M != P
    # involution: (M, P) appears as (P, M)
    pairs = set(zip(M[interior].tolist(), P[interior].tolist()))
    assert all((p, m) in pairs for m, p in pairs)
    # boundary slots are exactly the faces with a code
    nb = sum(op.bc_faces(k).shape[0] for k in (1, 2, 3))
    assert (~interior).sum() == nb * op.Nfp


# ---------------------------------------------------------------- the operator


class BruteForceDG:
    """Independent nodal strong-form DG Maxwell operator: monomial Lagrange
    bases in physical coordinates, Duffy quadrature for volume and face
    integrals, normals from cross products, neighbours paired by position.
    rhs = curl term + 1/2 M_e^-1 sum_f int_f phi flux_f dA, with the
    reference's jump convention and flux (HesthavenEvolution.cpp:474-534)."""

    def __init__(self, vx, EToV, fbc, p, pos, alpha, eps=None, mu=None):
        self.K, self.Np = EToV.shape[0], pos.shape[2]
        self.alpha = alpha
        K, Np = self.K, self.Np
        self.eps = np.ones(K) if eps is None else eps
        self.mu = np.ones(K) if mu is None else mu
        mons = monomials(p)
        qv, qvw = tet_quadrature(p + 3)
        qf, qfw = tri_quadrature(p + 3)
        fv = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]
        self.D = np.zeros((K, 3, Np, Np))
        self.Minv = np.zeros((K, Np, Np))
        self.face = []  # per element, per face: (face node ids, face mass Np x nf, normal, nbr info)
        xyz = pos.transpose(1, 2, 0)  # K, Np, 3
        keymap = {}
        for e in range(K):
            for f in range(4):
                keymap.setdefault(tuple(sorted(EToV[e, list(fv[f])])), []).append((e, f))
        for e in range(K):
            v = vx[EToV[e]]
            x0 = v[0]
            h = max(np.linalg.norm(v[i] - v[j]) for i in range(4) for j in range(i))
            loc = lambda X: (X - x0) / h

            def evalm(X):
                L = loc(X)
                return np.stack([L[:, 0] ** a * L[:, 1] ** b * L[:, 2] ** c for a, b, c in mons], 1)

            def evald(X, d):
                L = loc(X)
                cols = []
                for m in mons:
                    if m[d] == 0:
                        cols.append(0 * L[:, 0])
                        continue
                    mm = list(m)
                    mm[d] -= 1
                    cols.append(m[d] * L[:, 0] ** mm[0] * L[:, 1] ** mm[1] * L[:, 2] ** mm[2] / h)
                return np.stack(cols, 1)

            C = np.linalg.inv(evalm(xyz[e]))  # phi_j = sum_m C[m, j] mono_m
            for d in range(3):
                self.D[e, d] = evald(xyz[e], d) @ C
            A = np.array([v[1] - v[0], v[2] - v[0], v[3] - v[0]]).T
            Xq = v[0] + qv @ A.T
            phi = evalm(Xq) @ C
            M = (phi * (qvw * abs(np.linalg.det(A)))[:, None]).T @ phi
            self.Minv[e] = np.linalg.inv(M)
            faces = []
            for f in range(4):
                a, b, c = v[list(fv[f])]
                cr = np.cross(b - a, c - a)
                n = cr / np.linalg.norm(cr)
                if np.dot(n, a - v[f]) < 0:
                    n = -n
                # nodes on the face plane
                dist = np.abs((xyz[e] - a) @ n)
                fn = np.where(dist < 1e-10 * h)[0]
                Xf = a + np.outer(qf[:, 0], b - a) + np.outer(qf[:, 1], c - a)
                ph = evalm(Xf) @ C
                Mf = (ph * (qfw * np.linalg.norm(cr))[:, None]).T @ ph[:, fn]
                others = [ef for ef in keymap[tuple(sorted(EToV[e, list(fv[f])]))] if ef != (e, f)]
                if others:
                    e2, f2 = others[0]
                    nbr = []
                    for i in fn:
                        d2 = np.linalg.norm(xyz[e2] - xyz[e, i], axis=1)
                        nbr.append(e2 * Np + int(np.argmin(d2)))
                    faces.append((fn, Mf, n, np.array(nbr), 0))
                else:
                    faces.append((fn, Mf, n, None, int(fbc[e, f])))
            self.face.append(faces)

    def mult(self, q):
        K, Np, al = self.K, self.Np, self.alpha
        Q = q.reshape(6, K * Np)
        out = np.zeros((6, K, Np))
        for e in range(K):
            u = Q[:, e * Np:(e + 1) * Np]
            Dq = np.einsum("dij,cj->cdi", self.D[e], u)  # c, d, i
            curlH = np.array([Dq[5, 1] - Dq[4, 2], Dq[3, 2] - Dq[5, 0], Dq[4, 0] - Dq[3, 1]])
            curlE = np.array([Dq[2, 1] - Dq[1, 2], Dq[0, 2] - Dq[2, 0], Dq[1, 0] - Dq[0, 1]])
            lift = np.zeros((6, Np))
            for fn, Mf, n, nbr, bc in self.face[e]:
                uM = u[:, fn]
                if nbr is not None:
                    dq = Q[:, nbr] - uM
                else:
                    ce, ch = {0: (0, 0), 1: (-2, 0), 2: (0, -2)}[bc] if bc != 3 else (-1 / al, -1 / al)
                    dq = np.vstack([ce * uM[:3], ch * uM[3:]])
                dE, dH = dq[:3], dq[3:]
                ndH, ndE = n @ dH, n @ dE
                fH = -np.cross(n[:, None], dE, axis=0) + al * (dH - np.outer(n, ndH))
                fE = np.cross(n[:, None], dH, axis=0) + al * (dE - np.outer(n, ndE))
                lift += np.vstack([fE, fH]) @ Mf.T
            lift = 0.5 * (self.Minv[e] @ lift.T).T
            out[:3, e] = (curlH + lift[:3]) / self.eps[e]
            out[3:, e] = (-curlE + lift[3:]) / self.mu[e]
        return out.ravel()


@pytest.mark.parametrize("p,n,alpha,media", [(1, 1, 1.0, False), (2, 1, 0.5, True), (3, 1, 1.0, True),
                                             (2, 2, 0.0, False)])
def test_mult_matches_brute_force_operator(p, n, alpha, media):
    bc = ["pec", "pmc", "sma", "pec", "pmc", "sma"] if alpha else ["pec", "pmc", "pec", "pmc", "pmc", "pec"]
    vx, E, fbc = MG.cube_mesh(n, 1.0, bc)
    rng = np.random.default_rng(p + 10 * n)
    K = E.shape[0]
    eps = rng.uniform(1, 4, K) if media else None
    mu = rng.uniform(1, 2, K) if media else None
    op = O.Operator(vx, E, fbc, p, alpha=alpha, eps=eps, mu=mu)
    bf = BruteForceDG(vx, E, fbc, p, op.positions(), alpha, eps, mu)
    x = rng.standard_normal(op.ndofs)
    ref = bf.mult(x)
    got = op.mult(x)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_mult_single_distorted_tet():
    rng = np.random.default_rng(5)
    vx = np.array([[0, 0, 0], [1.3, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.3, 1.1]])
    E = np.array([[0, 1, 2, 3]], dtype=np.int32)
    fbc = np.array([[1, 2, 3, 1]], dtype=np.int32)
    op = O.Operator(vx, E, fbc, 3, alpha=1.0)
    bf = BruteForceDG(vx, E, fbc, 3, op.positions(), 1.0)
    x = rng.standard_normal(op.ndofs)
    ref = bf.mult(x)
    assert np.abs(op.mult(x) - ref).max() <= 1e-9 * np.abs(ref).max()


def energy_rate(op, q):
    """q^T Mbb L(q) by polarisation of the quadratic energy."""
    r = op.mult(q)
    return op.energy(q + r) - op.energy(q) - op.energy(r), r


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_centred_flux_conserves_energy(p):
    """alpha = 0 with PEC/PMC walls: the semi-discrete operator is skew-adjoint
    in the energy inner product, q^T Mbb L q = 0 (a flipped sign in the curl,
    the flux or a boundary coefficient breaks this)."""
    vx, E, fbc = MG.cube_mesh(2, 1.0, ["pec", "pmc", "pec", "pmc", "pmc", "pec"])
    rng = np.random.default_rng(p)
    K = E.shape[0]
    op = O.Operator(vx, E, fbc, p, alpha=0.0, eps=rng.uniform(1, 4, K), mu=rng.uniform(1, 2, K))
    q = rng.standard_normal(op.ndofs)
    rate, r = energy_rate(op, q)
    scale = math.sqrt(op.energy(q) * op.energy(r))
    assert abs(rate) <= 1e-12 * scale


@pytest.mark.parametrize("bc", ["pec", "pmc", "sma"])
def test_upwind_flux_dissipates(bc):
    vx, E, fbc = MG.cube_mesh(2, 1.0, bc)
    op = O.Operator(vx, E, fbc, 2, alpha=1.0)
    rng = np.random.default_rng(1)
    for _ in range(3):
        q = rng.standard_normal(op.ndofs)
        rate, r = energy_rate(op, q)
        assert rate < 0


def test_linearity_and_constant_fields():
    vx, E, fbc = MG.cube_mesh(3, 1.0, "pec")
    op = O.Operator(vx, E, fbc, 2)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(op.ndofs), rng.standard_normal(op.ndofs)
    np.testing.assert_allclose(op.mult(2 * x - 3 * y), 2 * op.mult(x) - 3 * op.mult(y), atol=1e-10)
    c = np.repeat(rng.standard_normal(6), op.K * op.Np)
    r = op.mult(c).reshape(6, op.K, op.Np)
    interior = np.where((fbc == 0).all(axis=1))[0]
    assert interior.size > 0
    np.testing.assert_allclose(r[:, interior], 0, atol=1e-11)


def test_invariant_under_local_vertex_order():
    """Rotating each element's local vertex order changes node and face
    numbering but not the operator: results agree node by node (by position)."""
    vx, E, fbc = MG.cube_mesh(2, 1.0, ["pec", "pmc", "sma", "pec", "pmc", "sma"])
    rng = np.random.default_rng(8)
    perms = [np.roll(np.arange(4), k) for k in range(4)]
    choice = rng.integers(0, 4, E.shape[0])
    E2 = np.array([E[e][perms[k]] for e, k in enumerate(choice)], dtype=np.int32)
    fbc2 = np.array([fbc[e][perms[k]] for e, k in enumerate(choice)], dtype=np.int32)
    op1 = O.Operator(vx, E, fbc, 3)
    op2 = O.Operator(vx, E2, fbc2, 3)
    f = lambda pos: np.concatenate([np.sin(pos[0] + 2 * pos[1] + c) * np.cos(pos[2] - c) for c in range(6)]).ravel()
    p1, p2 = op1.positions(), op2.positions()
    r1 = MG.sort_by_position(p1, op1.mult(f(p1)))
    r2 = MG.sort_by_position(p2, op2.mult(f(p2)))
    np.testing.assert_allclose(r1, r2, atol=1e-11)


# ---------------------------------------------------------------- errors


def test_setup_errors():
    vx, E, fbc = MG.cube_mesh(1, 1.0, "sma")
    with pytest.raises(RuntimeError, match="SMA"):
        O.Operator(vx, E, fbc, 2, alpha=0.0)
    bad = fbc.copy()
    bad[0, :] = 1  # element 0 has interior faces
    with pytest.raises(RuntimeError, match="interior"):
        O.Operator(vx, E, bad, 2)


# ---------------------------------------------------------------- time integration


@pytest.mark.parametrize("integrator", ["rk4", "lserk4"])
def test_fourth_order_in_time(integrator):
    vx, E, fbc = MG.cube_mesh(2, 1.0, "pec")
    op = O.Operator(vx, E, fbc, 2)
    q0 = MG.te110(op.positions())
    T = 0.2

    def run(dt):
        x = q0.copy()
        op.run(x, dt, T, integrator)
        return x

    ref = run(T / 256)
    errs = [np.linalg.norm(run(T / m) - ref) for m in (8, 16)]
    rate = math.log2(errs[0] / errs[1])
    assert 3.6 < rate < 4.6


def test_te110_spatial_convergence():
    errs = []
    for n in (2, 4):
        vx, E, fbc = MG.cube_mesh(n, 1.0, "pec")
        op = O.Operator(vx, E, fbc, 2)
        pos = op.positions()
        x = MG.te110(pos)
        _, t = op.run(x, 0.05 / n, 0.3, "lserk4")
        ex = MG.te110_exact(pos, t)
        errs.append(np.linalg.norm(x - ex) / np.linalg.norm(ex))
    assert math.log2(errs[0] / errs[1]) > 2.5  # order p+1 = 3 asymptotically
