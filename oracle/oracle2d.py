"""CPU FP64 restatement (numpy) of the reference's Hesthaven DGTD operator on
triangles, the 2D variant of HesthavenEvolution::Mult
(src/evolution/HesthavenEvolution.cpp:450-542).

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the dgtd_amd
package.

The reference's operator is dimension-generic: the directional matrices are
built for d = X, Y, Z and `buildDerivativeSubOperator` returns an empty
(zero) operator for d >= mesh dimension (src/components/DGOperatorFactory.h:
428-437), and 2D face normals have no z component; faces are the element's
edges with order+1 nodes (HesthavenEvolution.cpp:178-179).  So on triangles
the six field components evolve with the 3D flux and curl formulas with
d/dz = 0 and n_z = 0.

Reference element: MFEM's L2 Gauss-Lobatto triangle nodes, the orthonormal
Proriol-Koornwinder-Dubiner basis (Hesthaven & Warburton 2008, ch. 6),
Dr, Ds = V_r V^-1, V_s V^-1, M = (V V^T)^-1, LIFT = M^-1 E with E the edge
mass matrices (reference edge length 2); fscale = (edge length / 2) / J with
J = area / 2 (reference area 2, HesthavenEvolution.cpp:133-147).

Quadrilaterals (RefQuad, Operator2D(quads=True)): the reference runs them on
its Global operator, which assembles the same semi-discrete DG operator over
the L2 Gauss-Lobatto space of the quadrilateral (DGOperatorFactory.h:
1428-1600, ExtensiveCasesTest.cpp:1003,1082).  The element is the tensor
product of the p+1 Gauss-Lobatto points in MFEM's lexicographic order with
the orthonormal Legendre tensor basis; elements are parallelograms (affine
maps, J = area / 4), edges (v_f, v_{f+1 mod 4}).  Pinned in
tests/test_oracle_quad.py by closed forms (no quadrilateral matrices exist
in the reference's tests).
"""
import numpy as np

# Carpenter & Kennedy (1994) low-storage RK4(5) coefficients
RK4A = np.array([0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                 -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0])
RK4B = np.array([1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                 1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                 2277821191437.0 / 14882151754819.0])
RK4C = np.array([0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                 2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0])

# boundary codes (same as the 3D operator): 0 interior/none, 1 PEC, 2 PMC,
# 3 SMA, 4/5/6 interior PEC/PMC/SMA
EDGES = ((0, 1), (1, 2), (2, 0))  # MFEM triangle edge vertices


def gll01(p):
    """Gauss-Lobatto-Legendre points on [0, 1] (ends and roots of P_p')."""
    if p == 1:
        return np.array([0.0, 1.0])
    c = np.zeros(p + 1)
    c[p] = 1.0
    inner = np.sort(np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c)).real)
    x = np.concatenate([[-1.0], inner, [1.0]])
    x = 0.5 * (x - x[::-1])  # exact symmetry
    return 0.5 * (x + 1.0)


def jacobi_p(x, a, b, n):
    """Orthonormal Jacobi polynomial P_n^{(a,b)} (Hesthaven & Warburton, JacobiP)."""
    x = np.asarray(x, dtype=float)
    from math import gamma
    g0 = 2 ** (a + b + 1) / (a + b + 1) * gamma(a + 1) * gamma(b + 1) / gamma(a + b + 1)
    p0 = np.full_like(x, 1.0 / np.sqrt(g0))
    if n == 0:
        return p0
    g1 = (a + 1) * (b + 1) / (a + b + 3) * g0
    p1 = ((a + b + 2) * x / 2 + (a - b) / 2) / np.sqrt(g1)
    if n == 1:
        return p1
    aold = 2 / (2 + a + b) * np.sqrt((a + 1) * (b + 1) / (a + b + 3))
    for i in range(1, n):
        h1 = 2 * i + a + b
        anew = 2 / (h1 + 2) * np.sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3))
        bnew = -(a * a - b * b) / h1 / (h1 + 2)
        p0, p1 = p1, 1 / anew * (-aold * p0 + (x - bnew) * p1)
        aold = anew
    return p1


def grad_jacobi_p(x, a, b, n):
    if n == 0:
        return np.zeros_like(np.asarray(x, dtype=float))
    return np.sqrt(n * (n + a + b + 1)) * jacobi_p(x, a + 1, b + 1, n - 1)


def _rs_to_ab(r, s):
    a = np.where(np.abs(s - 1) > 1e-14, 2 * (1 + r) / np.where(np.abs(s - 1) > 1e-14, 1 - s, 1) - 1, -1.0)
    return a, s


def simplex2d(r, s, i, j):
    """Orthonormal basis psi_ij on the reference triangle (Simplex2DP)."""
    a, b = _rs_to_ab(r, s)
    return np.sqrt(2.0) * jacobi_p(a, 0, 0, i) * jacobi_p(b, 2 * i + 1, 0, j) * (1 - b) ** i


def grad_simplex2d(r, s, i, j):
    """d psi_ij / dr, d psi_ij / ds (GradSimplex2DP)."""
    a, b = _rs_to_ab(r, s)
    fa, dfa = jacobi_p(a, 0, 0, i), grad_jacobi_p(a, 0, 0, i)
    gb, dgb = jacobi_p(b, 2 * i + 1, 0, j), grad_jacobi_p(b, 2 * i + 1, 0, j)
    dr = dfa * gb
    if i > 0:
        dr = dr * (0.5 * (1 - b)) ** (i - 1)
    ds = dfa * (gb * 0.5 * (1 + a))
    if i > 0:
        ds = ds * (0.5 * (1 - b)) ** (i - 1)
    tmp = dgb * (0.5 * (1 - b)) ** i
    if i > 0:
        tmp = tmp - 0.5 * i * gb * (0.5 * (1 - b)) ** (i - 1)
    ds = ds + fa * tmp
    sc = 2 ** (i + 0.5)
    return dr * sc, ds * sc


class RefTri:
    """Reference triangle of order p in MFEM's L2 Gauss-Lobatto node order:
    j (eta) outer, i (xi) inner, node = (g_i, g_j) / (g_i + g_j + g_k),
    k = p - i - j; r = 2 xi - 1, s = 2 eta - 1."""

    def __init__(self, p):
        self.p = p
        g = gll01(p)
        lat, xi = [], []
        for j in range(p + 1):
            for i in range(p + 1 - j):
                k = p - i - j
                w = g[i] + g[j] + g[k]
                lat.append((i, j, k))
                xi.append((g[i] / w, g[j] / w))
        self.lattice = np.array(lat)
        self.xi = np.array(xi)
        self.Np = len(xi)
        self.Nfp = p + 1
        r, s = 2 * self.xi[:, 0] - 1, 2 * self.xi[:, 1] - 1
        modes = [(i, j) for i in range(p + 1) for j in range(p + 1 - i)]
        self.V = np.stack([simplex2d(r, s, i, j) for i, j in modes], axis=1)
        gr = [grad_simplex2d(r, s, i, j) for i, j in modes]
        Vr = np.stack([t[0] for t in gr], axis=1)
        Vs = np.stack([t[1] for t in gr], axis=1)
        self.modes = modes
        self.Vinv = np.linalg.inv(self.V)
        self.Dr = Vr @ self.Vinv
        self.Ds = Vs @ self.Vinv
        self.Minv = self.V @ self.V.T
        self.M = np.linalg.inv(self.Minv)
        # edge node lists, ordered from the edge's first vertex to its second
        # (vertex weights: v0 <- k, v1 <- i, v2 <- j)
        L = self.lattice
        f0 = [n for n in np.argsort(L[:, 0]) if L[n, 1] == 0]          # v0 -> v1: j = 0, i ascending
        f1 = [n for n in np.argsort(L[:, 1]) if L[n, 2] == 0]          # v1 -> v2: k = 0, j ascending
        f2 = [n for n in np.argsort(-L[:, 1]) if L[n, 0] == 0]         # v2 -> v0: i = 0, j descending
        self.Fmask = np.array([f0, f1, f2])
        # edge mass (reference edge length 2) from the 1D orthonormal basis
        E = np.zeros((self.Np, 3 * self.Nfp))
        lat_v = np.stack([L[:, 2], L[:, 0], L[:, 1]], axis=1)  # lattice index of v0, v1, v2
        for f, (va, vb) in enumerate(EDGES):
            # on an edge the node's weight of vertex vb is the GLL point g[l_vb]:
            # edge coordinate -1 at va, +1 at vb
            tg = 2 * g[lat_v[self.Fmask[f], vb]] - 1
            V1 = np.stack([jacobi_p(tg, 0, 0, n) for n in range(self.Nfp)], axis=1)
            M1 = np.linalg.inv(V1 @ V1.T)
            E[np.ix_(self.Fmask[f], f * self.Nfp + np.arange(self.Nfp))] = M1
        self.E = E
        self.LIFT = self.Minv @ E

    def interp_weights(self, xi, eta):
        """w with u(xi, eta) = sum_n w[n] u_n."""
        r, s = np.array([2 * xi - 1.0]), np.array([2 * eta - 1.0])
        P = np.array([simplex2d(r, s, i, j)[0] for i, j in self.modes])
        return self.Vinv.T @ P


class RefQuad:
    """Reference quadrilateral [-1, 1]^2 of order p: the tensor product of the
    p+1 Gauss-Lobatto-Legendre points in MFEM's lexicographic L2 order (node
    i + (p+1) j at (x_i, x_j)), the basis of products of orthonormal Legendre
    polynomials, Dr, Ds = V_r V^-1, V_s V^-1, M = (V V^T)^-1 (exact mass),
    edges (v0,v1), (v1,v2), (v2,v3), (v3,v0) of MFEM's quadrilateral with
    vertices (-1,-1), (1,-1), (1,1), (-1,1), LIFT = M^-1 E with E the edge
    masses (reference edge length 2).  xi = (r + 1) / 2, (s + 1) / 2."""

    EDGES = ((0, 1), (1, 2), (2, 3), (3, 0))
    AXES = (1, 3)  # x = v0 + xi (v1 - v0) + eta (v3 - v0)

    def __init__(self, p):
        self.p = p
        n1 = p + 1
        g = gll01(p)
        self.xi = np.array([(g[i], g[j]) for j in range(n1) for i in range(n1)])
        self.Np, self.Nfp = n1 * n1, n1
        r, s = 2 * self.xi[:, 0] - 1, 2 * self.xi[:, 1] - 1
        self.modes = [(a, b) for b in range(n1) for a in range(n1)]
        self.V = np.stack([jacobi_p(r, 0, 0, a) * jacobi_p(s, 0, 0, b) for a, b in self.modes], axis=1)
        Vr = np.stack([grad_jacobi_p(r, 0, 0, a) * jacobi_p(s, 0, 0, b) for a, b in self.modes], axis=1)
        Vs = np.stack([jacobi_p(r, 0, 0, a) * grad_jacobi_p(s, 0, 0, b) for a, b in self.modes], axis=1)
        self.Vinv = np.linalg.inv(self.V)
        self.Dr = Vr @ self.Vinv
        self.Ds = Vs @ self.Vinv
        self.Minv = self.V @ self.V.T
        self.M = np.linalg.inv(self.Minv)
        idx = lambda i, j: i + n1 * j
        self.Fmask = np.array([[idx(i, 0) for i in range(n1)],              # v0 -> v1: s = -1
                               [idx(p, j) for j in range(n1)],              # v1 -> v2: r = 1
                               [idx(i, p) for i in reversed(range(n1))],    # v2 -> v3: s = 1
                               [idx(0, j) for j in reversed(range(n1))]])   # v3 -> v0: r = -1
        tg = 2 * g - 1  # edge coordinate of the k-th node from the edge's first vertex
        V1 = np.stack([jacobi_p(tg, 0, 0, n) for n in range(n1)], axis=1)
        M1 = np.linalg.inv(V1 @ V1.T)
        E = np.zeros((self.Np, 4 * n1))
        for f in range(4):
            E[np.ix_(self.Fmask[f], f * n1 + np.arange(n1))] = M1
        self.E = E
        self.LIFT = self.Minv @ E

    def interp_weights(self, xi, eta):
        """w with u(xi, eta) = sum_n w[n] u_n."""
        r, s = np.array([2 * xi - 1.0]), np.array([2 * eta - 1.0])
        P = np.array([(jacobi_p(r, 0, 0, a) * jacobi_p(s, 0, 0, b))[0] for a, b in self.modes])
        return self.Vinv.T @ P


RefTri.EDGES = EDGES
RefTri.AXES = (1, 2)  # x = v0 + xi (v1 - v0) + eta (v2 - v0)


class Operator2D:
    """The 2D operator on a triangle mesh, or with quads=True on a mesh of
    parallelograms (the Global operator's quadrilateral cases: the same
    semi-discrete operator on the tensor-product element).

    vx: (nv, 2) or (nv, 3) vertex coordinates (z ignored); EToV: (K, n);
    face_bc: (K, n) codes per edge f = (v_f, v_{f+1 mod n}), n = 3 or 4;
    eps, mu: (K,)."""

    def __init__(self, vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None, global_sma=False, quads=False):
        # global_sma: SMA edges (codes 3 and 6) at upwind weight 1 whatever
        # alpha, as the reference's Global operator assembles them
        # (DGOperatorFactory.h:21-35, 485-567, 575-665)
        self.global_sma = bool(global_sma)
        self.R = R = RefQuad(order) if quads else RefTri(order)
        self.nf = nf = len(R.EDGES)
        vx = np.asarray(vx, dtype=float)[:, :2]
        EToV = np.asarray(EToV, dtype=np.int64).reshape(-1, nf)
        self.K = K = EToV.shape[0]
        self.Np, self.Nfp = R.Np, R.Nfp
        self.alpha = float(alpha)
        self.eps = np.ones(K) if eps is None else np.asarray(eps, dtype=float)
        self.mu = np.ones(K) if mu is None else np.asarray(mu, dtype=float)
        verts = [vx[EToV[:, j]] for j in range(nf)]
        v0, va, vb = verts[0], verts[R.AXES[0]], verts[R.AXES[1]]
        xr, xs = 0.5 * (va - v0), 0.5 * (vb - v0)
        J = xr[:, 0] * xs[:, 1] - xs[:, 0] * xr[:, 1]
        if np.any(np.abs(J) < 1e-300):
            raise ValueError("degenerate element")
        if quads:
            gap = np.abs(verts[2] - (va + vb - v0)).max(axis=1)
            if np.any(gap > 1e-10 * np.sqrt(np.abs(J))):
                raise ValueError("quadrilateral is not a parallelogram")
        self.J = J
        # d/dx = rx d/dr + sx d/ds, d/dy = ry d/dr + sy d/ds
        self.rx, self.ry = xs[:, 1] / J, -xs[:, 0] / J
        self.sx, self.sy = -xr[:, 1] / J, xr[:, 0] / J
        self.pos = (v0[:, None, :] + R.xi[None, :, 0:1] * (va - v0)[:, None, :]
                    + R.xi[None, :, 1:2] * (vb - v0)[:, None, :])  # (K, Np, 2)
        self.nrm = np.zeros((K, nf, 2))
        self.fscale = np.zeros((K, nf))
        sgn = np.sign(J)
        for f, (a, b) in enumerate(R.EDGES):
            t = verts[b] - verts[a]
            Lf = np.hypot(t[:, 0], t[:, 1])
            self.nrm[:, f, 0] = sgn * t[:, 1] / Lf
            self.nrm[:, f, 1] = -sgn * t[:, 0] / Lf
            self.fscale[:, f] = (Lf / 2) / np.abs(J)
        # connectivity: shared edges by vertex pair, node matching by position
        bc = np.asarray(face_bc, dtype=np.int64).reshape(K, nf).copy()
        owner = {}
        nbr = -np.ones((K, nf), dtype=np.int64)
        nbrf = -np.ones((K, nf), dtype=np.int64)
        for e in range(K):
            for f, (a, b) in enumerate(R.EDGES):
                key = tuple(sorted((int(EToV[e, a]), int(EToV[e, b]))))
                if key in owner:
                    e2, f2 = owner.pop(key)
                    nbr[e, f], nbrf[e, f], nbr[e2, f2], nbrf[e2, f2] = e2, f2, e, f
                else:
                    owner[key] = (e, f)
        # an interior condition or TF/SF code on one side applies to both
        for e in range(K):
            for f in range(nf):
                if nbr[e, f] >= 0 and bc[e, f] >= 4:
                    e2, f2 = nbr[e, f], nbrf[e, f]
                    if bc[e2, f2] not in (0, bc[e, f]):
                        raise ValueError("conflicting interior conditions on the two sides of an edge")
                    bc[e2, f2] = bc[e, f]
        for e in range(K):
            for f in range(nf):
                if nbr[e, f] >= 0 and 1 <= bc[e, f] <= 3:
                    raise ValueError("PEC/PMC/SMA codes 1-3 on an interior edge (use 4-6)")
                if nbr[e, f] < 0 and bc[e, f] >= 4:
                    raise ValueError("interior-boundary codes on a boundary edge")
                if bc[e, f] == 3 and self.alpha == 0.0 and not self.global_sma:
                    raise ValueError("SMA boundaries need upwind flux (alpha != 0)")
        Nfp = R.Nfp
        h = np.min(np.abs(J)) ** 0.5
        self.vmapM = np.zeros((K, nf, Nfp), dtype=np.int64)
        self.vmapP = np.zeros((K, nf, Nfp), dtype=np.int64)
        for e in range(K):
            for f in range(nf):
                own = e * R.Np + R.Fmask[f]
                self.vmapM[e, f] = own
                e2 = nbr[e, f]
                if e2 < 0:
                    self.vmapP[e, f] = own
                    continue
                f2 = nbrf[e, f]
                pm = self.pos[e, R.Fmask[f]]
                pp = self.pos[e2, R.Fmask[f2]]
                d = np.linalg.norm(pm[:, None, :] - pp[None, :, :], axis=2)
                b2 = np.argmin(d, axis=1)
                if np.any(d[np.arange(Nfp), b2] > 1e-7 * h):
                    raise ValueError("non-conforming edge: node not matched")
                self.vmapP[e, f] = e2 * R.Np + R.Fmask[f2][b2]
        # TF/SF edges (code 7): normal neighbour jumps plus the incident field,
        # -inc on the scattered side (the lower-index element, MFEM's first
        # element of the face) and +inc on the total side
        # (HesthavenEvolution::evaluateTFSF, HesthavenEvolution.cpp:97-124)
        self.tfsf = np.where(bc == 7, np.where(np.arange(K)[:, None] < nbr, -1, 1), 0)
        bc = np.where(bc == 7, 0, bc)
        self.bc = bc
        self.nbr = nbr
        self.waves = []

    def ndofs(self):
        return 6 * self.Np * self.K

    def max_fscale(self):
        return float(self.fscale.max())

    def node_positions(self):
        return self.pos

    def set_planewaves(self, waves):
        """waves: list of dicts kind ('gaussian' | 'modulated'), polarization,
        propagation, spread, mean (projected, k.mean), frequency."""
        out = []
        for w in waves:
            pol = np.asarray(w["polarization"], dtype=float)
            k = np.asarray(w["propagation"], dtype=float)
            pol, k = pol / np.linalg.norm(pol), k / np.linalg.norm(k)
            out.append(dict(pol=pol, k=k, kxp=np.cross(k, pol), spread=float(w["spread"]),
                            mean=float(w.get("mean", 0.0)), freq=float(w.get("frequency", 0.0)),
                            modulated=w.get("kind", "gaussian") == "modulated"))
        self.waves = out

    def incident(self, pos, t):
        """E_inc, H_inc (each (3, ...)) of the plane waves at positions pos
        (..., 2): g(u) pol and g(u) k x pol with u = k.x - t - mean
        (Planewave, src/math/Function.h:331-374)."""
        E = np.zeros((3,) + pos.shape[:-1])
        H = np.zeros_like(E)
        for w in self.waves:
            u = w["k"][0] * pos[..., 0] + w["k"][1] * pos[..., 1] - t - w["mean"]
            g = np.exp(-u * u / (2 * w["spread"] ** 2))
            if w["modulated"]:
                g = g * np.cos(2 * np.pi * w["freq"] * u)
            E += np.multiply.outer(w["pol"], g)
            H += np.multiply.outer(w["kxp"], g)
        return E, H

    def mult(self, x, t=0.0):
        """HesthavenEvolution::Mult with d/dz = 0 and n_z = 0, at time t."""
        R, K, Np, Nfp = self.R, self.K, self.Np, self.Nfp
        q = np.asarray(x, dtype=float).reshape(6, K * Np)
        M, P = self.vmapM.reshape(K, -1), self.vmapP.reshape(K, -1)
        nf = self.nf
        dF = q[:, P] - q[:, M]  # (6, K, nf Nfp): jump = exterior - interior
        own = q[:, M]
        code = np.repeat(self.bc, Nfp, axis=1)  # (K, nf Nfp)
        # per edge-node upwind weight: alpha, or 1 on SMA edges (global_sma)
        a = np.where((code == 3) | (code == 6), 1.0, self.alpha) if self.global_sma else np.full(code.shape, self.alpha)
        sma = np.divide(-1.0, a, out=np.zeros_like(a), where=a != 0.0)
        ce = np.select([code == 1, code == 3, code == 4, code == 6], [-2.0, sma, -1.0, -0.5], 0.0)
        ch = np.select([code == 2, code == 3, code == 5, code == 6], [-2.0, sma, -1.0, -0.5], 0.0)
        ovr = code > 0
        for c in range(3):
            dF[c] = np.where(ovr, ce * own[c], dF[c])
            dF[3 + c] = np.where(ovr, ch * own[3 + c], dF[3 + c])
        if self.waves and np.any(self.tfsf):
            sg = np.repeat(self.tfsf, Nfp, axis=1)  # (K, nf Nfp)
            pM = self.pos.reshape(-1, 2)[M.reshape(-1)].reshape(K, nf * Nfp, 2)
            Ei, Hi = self.incident(pM, t)
            for c in range(3):
                dF[c] = dF[c] + sg * Ei[c]
                dF[3 + c] = dF[3 + c] + sg * Hi[c]
        n = np.repeat(self.nrm, Nfp, axis=1)  # (K, nf Nfp, 2)
        nv = [n[..., 0], n[..., 1], np.zeros_like(n[..., 0])]
        ndE = nv[0] * dF[0] + nv[1] * dF[1] + nv[2] * dF[2]
        ndH = nv[0] * dF[3] + nv[1] * dF[4] + nv[2] * dF[5]
        fs = np.repeat(self.fscale, Nfp, axis=1) / 2.0
        out = np.zeros((6, K, Np))
        u = q.reshape(6, K, Np)
        ur = np.einsum("ij,ckj->cki", R.Dr, u)
        us = np.einsum("ij,ckj->cki", R.Ds, u)
        dx = self.rx[None, :, None] * ur + self.sx[None, :, None] * us
        dy = self.ry[None, :, None] * ur + self.sy[None, :, None] * us
        D = [dx, dy, np.zeros_like(dx)]  # D[d][c]: d/dx_d of component c
        for x in range(3):
            y, z = (x + 1) % 3, (x + 2) % 3
            fH = -nv[y] * dF[z] + nv[z] * dF[y] + a * (dF[3 + x] - ndH * nv[x])
            fE = nv[y] * dF[3 + z] - nv[z] * dF[3 + y] + a * (dF[x] - ndE * nv[x])
            lH = np.einsum("im,km->ki", R.LIFT, fH * fs)
            lE = np.einsum("im,km->ki", R.LIFT, fE * fs)
            hres = -D[y][z] + D[z][y] + lH
            eres = D[y][3 + z] - D[z][3 + y] + lE
            out[3 + x] = hres / self.mu[:, None]
            out[x] = eres / self.eps[:, None]
        return out.reshape(-1)

    def rk4_step(self, x, dt, t=0.0):
        """mfem::RK4Solver stage order and stage times (Solver.cpp:535-551)."""
        k = self.mult(x, t)
        y, z = x + dt / 2 * k, x + dt / 6 * k
        k = self.mult(y, t + dt / 2)
        y, z = x + dt / 2 * k, z + dt / 3 * k
        k = self.mult(y, t + dt / 2)
        y, z = x + dt * k, z + dt / 3 * k
        k = self.mult(y, t + dt)
        return z + dt / 6 * k

    def lserk4_step(self, q, res, dt, t=0.0):
        for s in range(5):
            k = self.mult(q, t + RK4C[s] * dt)
            res = RK4A[s] * res + dt * k
            q = q + RK4B[s] * res
        return q, res

    def run(self, x, dt, final_time, integrator="lserk4"):
        """Solver::run loop guard and last-step truncation (Solver.cpp:497, 537)."""
        x = np.array(x, dtype=float)
        res = np.zeros_like(x)
        t, steps = 0.0, 0
        while t <= final_time - 1e-8 * dt:
            h = min(dt, final_time - t)
            if integrator == "rk4":
                x = self.rk4_step(x, h, t)
            else:
                x, res = self.lserk4_step(x, res, h, t)
            t += h
            steps += 1
        return x, steps, t

    def energy(self, x):
        """1/2 sum_e J_e (eps E^T M E + mu H^T M H) over all six components."""
        u = np.asarray(x, dtype=float).reshape(6, self.K, self.Np)
        s = np.einsum("cki,ij,ckj->ck", u, self.R.M, u)
        J = np.abs(self.J)
        return float(0.5 * np.sum(J * (self.eps * s[:3].sum(0) + self.mu * s[3:].sum(0))))
