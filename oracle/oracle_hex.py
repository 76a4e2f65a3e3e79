"""CPU FP64 restatement (numpy) of the reference's Global-operator semi-discrete
DG Maxwell operator on hexahedral meshes (parallelepipeds).

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the dgtd_amd
package.

The reference runs hexahedral meshes on its Global operator, which assembles
the DG operator of the L2 Gauss-Lobatto space (DGOperatorFactory.h:
1428-1600; Solver3DTest.cpp:57-200 run it): the same strong-form curl, upwind
/ centred face flux and boundary jumps as HesthavenEvolution::Mult
(HesthavenEvolution.cpp:450-542, with the Global operator's SMA weighting),
on the tensor-product element.  Reference element [-1, 1]^3: the (p+1)^3
Gauss-Lobatto points in MFEM's lexicographic order (node i + (p+1) j +
(p+1)^2 k at (x_i, x_j, x_k)), the orthonormal Legendre tensor basis, Dr, Ds,
Dt = V_r V^-1 ..., M = (V V^T)^-1, faces of MFEM's hexahedron (vertex cycles
(3,2,1,0), (0,1,5,4), (1,2,6,5), (2,3,7,6), (3,0,4,7), (4,5,6,7)), LIFT =
M^-1 E with E the exact face masses (reference face area 4).  Elements are
affine (J = volume / 8, fscale = (face area / 4) / J); neighbours and face
nodes are matched by position.  Pinned in tests/test_oracle_hex.py by closed
forms (the reference's tests hold no hexahedral matrices).
"""
import numpy as np

from oracle.oracle2d import RK4A, RK4B, RK4C, gll01, grad_jacobi_p, jacobi_p

HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
HEX_FACES = ((3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7))


class RefHex:
    def __init__(self, p):
        self.p = p
        n1 = p + 1
        g = gll01(p)
        self.g = g
        self.xi = np.array([(g[i], g[j], g[k]) for k in range(n1) for j in range(n1) for i in range(n1)])
        self.Np, self.Nfp = n1 ** 3, n1 ** 2
        r, s, t = (2 * self.xi[:, d] - 1 for d in range(3))
        self.modes = [(a, b, c) for c in range(n1) for b in range(n1) for a in range(n1)]
        P = lambda x, n: jacobi_p(x, 0, 0, n)
        dP = lambda x, n: grad_jacobi_p(x, 0, 0, n)
        self.V = np.stack([P(r, a) * P(s, b) * P(t, c) for a, b, c in self.modes], axis=1)
        Vr = np.stack([dP(r, a) * P(s, b) * P(t, c) for a, b, c in self.modes], axis=1)
        Vs = np.stack([P(r, a) * dP(s, b) * P(t, c) for a, b, c in self.modes], axis=1)
        Vt = np.stack([P(r, a) * P(s, b) * dP(t, c) for a, b, c in self.modes], axis=1)
        self.Vinv = np.linalg.inv(self.V)
        self.Dr, self.Ds, self.Dt = Vr @ self.Vinv, Vs @ self.Vinv, Vt @ self.Vinv
        self.Minv = self.V @ self.V.T
        self.M = np.linalg.inv(self.Minv)
        # face node (m, n) at corner a + u_m (b - a) + u_n (d - a) of the face cycle
        lattice = {tuple(np.rint(x * p).astype(int)): q for q, x in
                   enumerate(np.array([(i, j, k) for k in range(n1) for j in range(n1) for i in range(n1)]) / p)}
        Fm = []
        for f in HEX_FACES:
            A, B, Dv = HEX_CORNERS[f[0]], HEX_CORNERS[f[1]], HEX_CORNERS[f[3]]
            Fm.append([lattice[tuple(p * A + m * (B - A) + n * (Dv - A))] for n in range(n1) for m in range(n1)])
        self.Fmask = np.array(Fm)
        tg = 2 * g - 1
        V1 = np.stack([P(tg, n) for n in range(n1)], axis=1)
        M1 = np.linalg.inv(V1 @ V1.T)
        M2 = np.kron(M1, M1)  # on (n, m) pairs with m fastest: (m + n1 n)
        E = np.zeros((self.Np, 6 * self.Nfp))
        for f in range(6):
            E[np.ix_(self.Fmask[f], f * self.Nfp + np.arange(self.Nfp))] = M2
        self.E = E
        self.LIFT = self.Minv @ E

    def interp_weights(self, xi, eta, zeta):
        r, s, t = (np.array([2 * v - 1.0]) for v in (xi, eta, zeta))
        P = np.array([(jacobi_p(r, 0, 0, a) * jacobi_p(s, 0, 0, b) * jacobi_p(t, 0, 0, c))[0]
                      for a, b, c in self.modes])
        return self.Vinv.T @ P


class OperatorHex:
    """The Global operator's semi-discrete DG Maxwell operator on a mesh of
    parallelepipeds.  vx: (nv, 3); EToV: (K, 8) in MFEM's vertex order;
    face_bc: (K, 6) codes per face (0 none/interior, 1-3 PEC/PMC/SMA boundary,
    4-6 interior PEC/PMC/SMA, 7 TF/SF); eps, mu: (K,).  global_sma: SMA
    faces at upwind weight 1 whatever alpha (DGOperatorFactory.h:485-567,
    587-591, 655-659)."""

    def __init__(self, vx, EToV, face_bc, order, alpha=1.0, eps=None, mu=None, global_sma=True):
        self.R = R = RefHex(order)
        self.global_sma = bool(global_sma)
        vx = np.asarray(vx, dtype=float)
        E = np.asarray(EToV, dtype=np.int64).reshape(-1, 8)
        self.K = K = E.shape[0]
        self.Np, self.Nfp = R.Np, R.Nfp
        self.alpha = float(alpha)
        self.eps = np.ones(K) if eps is None else np.asarray(eps, dtype=float)
        self.mu = np.ones(K) if mu is None else np.asarray(mu, dtype=float)
        v = vx[E]  # (K, 8, 3)
        a = np.stack([v[:, 1] - v[:, 0], v[:, 3] - v[:, 0], v[:, 4] - v[:, 0]], axis=1)  # (K, 3 axes, 3)
        want = v[:, :1] + np.einsum("vk,ekd->evd", HEX_CORNERS.astype(float), a)
        if np.abs(want - v).max() > 1e-10 * np.abs(a).max():
            raise ValueError("hexahedra must be parallelepipeds")
        det = np.linalg.det(a)
        if np.any(det <= 0):
            raise ValueError("inverted hexahedron")
        self.J = det / 8.0
        # dx/dr_k = a_k / 2: metric dr_k/dx_d = inverse of [a_0 a_1 a_2] / 2 (columns)
        Jm = 0.5 * np.transpose(a, (0, 2, 1))  # (K, d, k)
        self.G = np.linalg.inv(Jm)             # (K, k, d)
        self.pos = v[:, :1] + np.einsum("nk,ekd->end", R.xi, a)  # (K, Np, 3)
        self.nrm = np.zeros((K, 6, 3))
        self.fscale = np.zeros((K, 6))
        for f, cyc in enumerate(HEX_FACES):
            u = v[:, cyc[1]] - v[:, cyc[0]]
            w = v[:, cyc[3]] - v[:, cyc[0]]
            c = np.cross(u, w)
            area = np.linalg.norm(c, axis=1)
            self.nrm[:, f] = c / area[:, None]
            self.fscale[:, f] = (area / 4.0) / self.J
        bc = np.asarray(face_bc, dtype=np.int64).reshape(K, 6).copy()
        owner, nbr, nbrf = {}, -np.ones((K, 6), dtype=np.int64), -np.ones((K, 6), dtype=np.int64)
        for e in range(K):
            for f, cyc in enumerate(HEX_FACES):
                key = tuple(sorted(int(E[e, q]) for q in cyc))
                if key in owner:
                    e2, f2 = owner.pop(key)
                    nbr[e, f], nbrf[e, f], nbr[e2, f2], nbrf[e2, f2] = e2, f2, e, f
                else:
                    owner[key] = (e, f)
        for e in range(K):
            for f in range(6):
                if nbr[e, f] >= 0 and bc[e, f] >= 4:
                    e2, f2 = nbr[e, f], nbrf[e, f]
                    if bc[e2, f2] not in (0, bc[e, f]):
                        raise ValueError("conflicting interior conditions")
                    bc[e2, f2] = bc[e, f]
        for e in range(K):
            for f in range(6):
                if nbr[e, f] >= 0 and 1 <= bc[e, f] <= 3:
                    raise ValueError("PEC/PMC/SMA codes 1-3 on an interior face (use 4-6)")
                if nbr[e, f] < 0 and bc[e, f] >= 4:
                    raise ValueError("interior-boundary codes on a boundary face")
        Nfp = R.Nfp
        h = np.min(self.J) ** (1 / 3)
        self.vmapM = np.zeros((K, 6, Nfp), dtype=np.int64)
        self.vmapP = np.zeros((K, 6, Nfp), dtype=np.int64)
        for e in range(K):
            for f in range(6):
                own = e * R.Np + R.Fmask[f]
                self.vmapM[e, f] = own
                e2 = nbr[e, f]
                if e2 < 0:
                    self.vmapP[e, f] = own
                    continue
                f2 = nbrf[e, f]
                pm, pp = self.pos[e, R.Fmask[f]], self.pos[e2, R.Fmask[f2]]
                d = np.linalg.norm(pm[:, None, :] - pp[None, :, :], axis=2)
                b2 = np.argmin(d, axis=1)
                if np.any(d[np.arange(Nfp), b2] > 1e-7 * h):
                    raise ValueError("non-conforming face: node not matched")
                self.vmapP[e, f] = e2 * R.Np + R.Fmask[f2][b2]
        # TF/SF faces (code 7): -inc on the scattered side (lower index), +inc on the total side
        self.tfsf = np.where(bc == 7, np.where(np.arange(K)[:, None] < nbr, -1, 1), 0)
        self.bc = np.where(bc == 7, 0, bc)
        self.nbr = nbr
        self.waves = []

    def ndofs(self):
        return 6 * self.Np * self.K

    def max_fscale(self):
        return float(self.fscale.max())

    def set_planewaves(self, waves):
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
        """E_inc, H_inc of the plane waves at positions pos (..., 3)
        (Planewave, src/math/Function.h:331-374)."""
        E = np.zeros((3,) + pos.shape[:-1])
        H = np.zeros_like(E)
        for w in self.waves:
            u = pos @ w["k"] - t - w["mean"]
            g = np.exp(-u * u / (2 * w["spread"] ** 2))
            if w["modulated"]:
                g = g * np.cos(2 * np.pi * w["freq"] * u)
            E += np.multiply.outer(w["pol"], g)
            H += np.multiply.outer(w["kxp"], g)
        return E, H

    def mult(self, x, t=0.0):
        R, K, Np, Nfp = self.R, self.K, self.Np, self.Nfp
        q = np.asarray(x, dtype=float).reshape(6, K * Np)
        M, P = self.vmapM.reshape(K, -1), self.vmapP.reshape(K, -1)
        dF = q[:, P] - q[:, M]
        own = q[:, M]
        code = np.repeat(self.bc, Nfp, axis=1)
        a = np.where((code == 3) | (code == 6), 1.0, self.alpha) if self.global_sma else np.full(code.shape, self.alpha)
        sma = np.divide(-1.0, a, out=np.zeros_like(a), where=a != 0.0)
        ce = np.select([code == 1, code == 3, code == 4, code == 6], [-2.0, sma, -1.0, -0.5], 0.0)
        ch = np.select([code == 2, code == 3, code == 5, code == 6], [-2.0, sma, -1.0, -0.5], 0.0)
        ovr = code > 0
        for c in range(3):
            dF[c] = np.where(ovr, ce * own[c], dF[c])
            dF[3 + c] = np.where(ovr, ch * own[3 + c], dF[3 + c])
        if self.waves and np.any(self.tfsf):
            sg = np.repeat(self.tfsf, Nfp, axis=1)
            pM = self.pos.reshape(-1, 3)[M.reshape(-1)].reshape(K, 6 * Nfp, 3)
            Ei, Hi = self.incident(pM, t)
            for c in range(3):
                dF[c] = dF[c] + sg * Ei[c]
                dF[3 + c] = dF[3 + c] + sg * Hi[c]
        n = np.repeat(self.nrm, Nfp, axis=1)  # (K, 6Nfp, 3)
        nv = [n[..., 0], n[..., 1], n[..., 2]]
        ndE = nv[0] * dF[0] + nv[1] * dF[1] + nv[2] * dF[2]
        ndH = nv[0] * dF[3] + nv[1] * dF[4] + nv[2] * dF[5]
        fs = np.repeat(self.fscale, Nfp, axis=1) / 2.0
        u = q.reshape(6, K, Np)
        ur = np.einsum("ij,ckj->cki", R.Dr, u)
        us = np.einsum("ij,ckj->cki", R.Ds, u)
        ut = np.einsum("ij,ckj->cki", R.Dt, u)
        G = self.G  # (K, k, d)
        D = [G[None, :, 0, d, None] * ur + G[None, :, 1, d, None] * us + G[None, :, 2, d, None] * ut for d in range(3)]
        out = np.zeros((6, K, Np))
        for x in range(3):
            y, z = (x + 1) % 3, (x + 2) % 3
            fH = -nv[y] * dF[z] + nv[z] * dF[y] + a * (dF[3 + x] - ndH * nv[x])
            fE = nv[y] * dF[3 + z] - nv[z] * dF[3 + y] + a * (dF[x] - ndE * nv[x])
            lH = np.einsum("im,km->ki", R.LIFT, fH * fs)
            lE = np.einsum("im,km->ki", R.LIFT, fE * fs)
            out[3 + x] = (-D[y][z] + D[z][y] + lH) / self.mu[:, None]
            out[x] = (D[y][3 + z] - D[z][3 + y] + lE) / self.eps[:, None]
        return out.reshape(-1)

    def rk4_step(self, x, dt, t=0.0):
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
        u = np.asarray(x, dtype=float).reshape(6, self.K, self.Np)
        s = np.einsum("cki,ij,ckj->ck", u, self.R.M, u)
        return float(0.5 * np.sum(self.J * (self.eps * s[:3].sum(0) + self.mu * s[3:].sum(0))))
