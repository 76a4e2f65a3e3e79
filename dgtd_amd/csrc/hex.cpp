// Reference hexahedron (tensor product of Gauss-Lobatto-Legendre points,
// MFEM's L2 Gauss-Lobatto hexahedron: node i + (p+1) j + (p+1)^2 k at
// (x_i, x_j, x_k), lexicographic) and hexahedral-mesh setup for the assembled
// operator.
#include "hex.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>

#include "quad.hpp"

namespace dgtd {

const int kHexFaces[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};

namespace {

// unit-cube corner coordinates of MFEM's hexahedron vertices
const int kHexCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// corners of the unit square in cycle order, as (s, t)
const int kSquareCorner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

}  // namespace

RefHex build_refhex(int p)
{
    if (p < 1 || p > 5) throw std::runtime_error("hexahedron order must be in 1..5");
    RefHex R;
    R.p = p;
    const int n1 = p + 1;
    R.Nfp = n1 * n1;
    R.Np = n1 * n1 * n1;
    const int Np = R.Np, Nfp = R.Nfp;
    const Gll1D g1 = build_gll1d(p);
    R.x1 = g1.x;
    auto node = [&](int i, int j, int k) { return i + n1 * (j + n1 * k); };
    R.xi.resize(3 * (size_t)Np);
    R.M.assign((size_t)Np * Np, 0.0);
    R.Dr.assign((size_t)Np * Np, 0.0);
    R.Ds.assign((size_t)Np * Np, 0.0);
    R.Dt.assign((size_t)Np * Np, 0.0);
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) {
                const int a = node(i, j, k);
                R.xi[3 * a] = 0.5 * (R.x1[i] + 1.0);
                R.xi[3 * a + 1] = 0.5 * (R.x1[j] + 1.0);
                R.xi[3 * a + 2] = 0.5 * (R.x1[k] + 1.0);
                for (int kk = 0; kk < n1; ++kk)
                    for (int jj = 0; jj < n1; ++jj)
                        for (int ii = 0; ii < n1; ++ii) {
                            const size_t q = (size_t)a * Np + node(ii, jj, kk);
                            R.M[q] = g1.M[i * n1 + ii] * g1.M[j * n1 + jj] * g1.M[k * n1 + kk];
                            if (jj == j && kk == k) R.Dr[q] = g1.D[i * n1 + ii];
                            if (ii == i && kk == k) R.Ds[q] = g1.D[j * n1 + jj];
                            if (ii == i && jj == j) R.Dt[q] = g1.D[k * n1 + kk];
                        }
            }
    // face nodes: (m, n) at corner a + s (b - a) + t (d - a), on the lattice
    // of 1D indices 0..p (index p - m is the mirror of m: the points are
    // symmetric)
    R.Fmask.resize(6 * (size_t)Nfp);
    for (int f = 0; f < 6; ++f) {
        const int* A = kHexCorner[kHexFaces[f][0]];
        const int* B = kHexCorner[kHexFaces[f][1]];
        const int* Dv = kHexCorner[kHexFaces[f][3]];
        for (int n = 0; n < n1; ++n)
            for (int m = 0; m < n1; ++m) {
                int ix[3];
                for (int d = 0; d < 3; ++d) ix[d] = p * A[d] + m * (B[d] - A[d]) + n * (Dv[d] - A[d]);
                R.Fmask[(size_t)f * Nfp + m + n1 * n] = node(ix[0], ix[1], ix[2]);
            }
    }
    // LIFT = M^-1 E, E the face masses (reference face area 4): M1 x M1 on the
    // face's (m, n).  With the tensor mass M1 x M1 x M1 the tangential factors
    // cancel (M1^-1 M1 = I; the Gauss-Lobatto points are symmetric, so a face
    // traversed downwards gives the same identity), which leaves the closed
    // form: face node q lifts onto the normal line through it, weight
    // M1^-1[i, b] at line index i (b = 0 or p, the face's index on that line),
    // and exactly zero elsewhere -- what the assembled matrix stores.
    R.LIFT.assign((size_t)Np * 6 * Nfp, 0.0);
    for (int f = 0; f < 6; ++f) {
        const int* A = kHexCorner[kHexFaces[f][0]];
        const int* B = kHexCorner[kHexFaces[f][1]];
        const int* Dv = kHexCorner[kHexFaces[f][3]];
        int nd = 0;
        while (!(A[nd] == B[nd] && A[nd] == Dv[nd])) ++nd;
        const int b = p * A[nd];
        const int stride = nd == 0 ? 1 : nd == 1 ? n1 : n1 * n1;
        for (int q = 0; q < Nfp; ++q) {
            const int c = R.Fmask[(size_t)f * Nfp + q];
            const int base = c - b * stride;  // the line's node with index 0
            for (int i = 0; i < n1; ++i)
                R.LIFT[(size_t)(base + i * stride) * 6 * Nfp + f * Nfp + q] = g1.Minv[(size_t)i * n1 + b];
        }
    }
    // neighbour table: own face node (m, n), own corners A, B, D at the
    // neighbour cycle's corners r, r + 1 (s = 0) or r - 1 (s = 1), and r - 1
    // or r + 1: the neighbour's face parameters are affine in (m, n)
    R.face_perm.assign(48 * (size_t)Nfp, -1);
    for (int g = 0; g < 6; ++g)
        for (int o = 0; o < 8; ++o) {
            const int r = o / 2, s = o % 2;
            const int* cA = kSquareCorner[r];
            const int* cB = kSquareCorner[s == 0 ? (r + 1) % 4 : (r + 3) % 4];
            const int* cD = kSquareCorner[s == 0 ? (r + 3) % 4 : (r + 1) % 4];
            for (int n = 0; n < n1; ++n)
                for (int m = 0; m < n1; ++m) {
                    const int m2 = p * cA[0] + m * (cB[0] - cA[0]) + n * (cD[0] - cA[0]);
                    const int n2 = p * cA[1] + m * (cB[1] - cA[1]) + n * (cD[1] - cA[1]);
                    R.face_perm[((size_t)g * 8 + o) * Nfp + m + n1 * n] = R.Fmask[(size_t)g * Nfp + m2 + n1 * n2];
                }
        }
    return R;
}

void hex_interp_weights(const RefHex& R, double xi, double eta, double zeta, double* w)
{
    const std::vector<double> lr = gll_lagrange(R.x1, 2 * xi - 1), ls = gll_lagrange(R.x1, 2 * eta - 1),
                              lt = gll_lagrange(R.x1, 2 * zeta - 1);
    const int n1 = R.p + 1;
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) w[i + n1 * (j + n1 * k)] = lr[i] * ls[j] * lt[k];
}

namespace {

// Affine frame of a hexahedron: origin v0 and edge vectors v1 - v0, v3 - v0,
// v4 - v0 (the unit cube's x, y, z axes); checks that the other four
// vertices are where a parallelepiped puts them.
struct HexFrame {
    double o[3], a[3][3];  // a[k] = edge vector of unit-cube axis k
    double det;            // a[0] . (a[1] x a[2])
};

HexFrame hex_frame(const double* vx, const int32_t* ev)
{
    HexFrame F;
    const double* v[8];
    for (int k = 0; k < 8; ++k) v[k] = vx + 3 * (size_t)ev[k];
    const int axis_vertex[3] = {1, 3, 4};
    for (int d = 0; d < 3; ++d) F.o[d] = v[0][d];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) F.a[k][d] = v[axis_vertex[k]][d] - v[0][d];
    F.det = F.a[0][0] * (F.a[1][1] * F.a[2][2] - F.a[1][2] * F.a[2][1]) -
            F.a[0][1] * (F.a[1][0] * F.a[2][2] - F.a[1][2] * F.a[2][0]) +
            F.a[0][2] * (F.a[1][0] * F.a[2][1] - F.a[1][1] * F.a[2][0]);
    double scale = 0.0;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) scale = std::max(scale, std::fabs(F.a[k][d]));
    for (int vv = 0; vv < 8; ++vv)
        for (int d = 0; d < 3; ++d) {
            double want = F.o[d];
            for (int k = 0; k < 3; ++k) want += kHexCorner[vv][k] * F.a[k][d];
            if (std::fabs(v[vv][d] - want) > 1e-12 * scale)
                throw std::runtime_error("hexahedra must be parallelepipeds (affine elements)");
        }
    return F;
}

}  // namespace

HexMeshData build_hex_mesh(const MeshInput& in, const RefHex& R, double alpha)
{
    (void)R;
    const int K = in.K;
    if (K <= 0 || in.nv <= 0 || !in.vx || !in.EToV) throw std::runtime_error("empty or incomplete hexahedral mesh");
    for (size_t i = 0; i < 8 * (size_t)K; ++i)
        if (in.EToV[i] < 0 || in.EToV[i] >= in.nv) throw std::runtime_error("EToV vertex id out of range");
    HexMeshData md;
    md.K = K;
    md.geo.assign((size_t)kHexGeoStride * K, 0.0);
    md.conn.assign((size_t)kHexConnStride * K, 0);
    md.jac.resize(K);
    md.eps.assign(K, 1.0);
    md.mu.assign(K, 1.0);
    md.face_code.assign(6 * (size_t)K, 0);
    for (int e = 0; e < K; ++e) {
        if (in.eps_r) md.eps[e] = in.eps_r[e];
        if (in.mu_r) md.mu[e] = in.mu_r[e];
        if (!(md.eps[e] > 0) || !(md.mu[e] > 0)) throw std::runtime_error("eps_r and mu_r must be positive");
        const int32_t* ev = in.EToV + 8 * (size_t)e;
        const HexFrame F = hex_frame(in.vx, ev);
        if (!(F.det > 0)) throw std::runtime_error("inverted or degenerate hexahedron (MFEM vertex order, positive volume)");
        // x = o + (r+1)/2 a0 + (s+1)/2 a1 + (t+1)/2 a2: dx/dr_k = a_k / 2, J = det / 8
        double Jm[3][3];  // Jm[d][k] = dx_d / dr_k
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 3; ++k) Jm[d][k] = 0.5 * F.a[k][d];
        const double J = F.det / 8.0;
        md.jac[e] = J;
        double* g = &md.geo[(size_t)kHexGeoStride * e];
        // dr_k/dx_d = (Jm^-1)[k][d]: cofactors over J
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) {
                const int d1 = (d + 1) % 3, d2 = (d + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                g[3 * k + d] = (Jm[d1][k1] * Jm[d2][k2] - Jm[d1][k2] * Jm[d2][k1]) / J;
            }
        for (int f = 0; f < 6; ++f) {
            const double* A = in.vx + 3 * (size_t)ev[kHexFaces[f][0]];
            const double* B = in.vx + 3 * (size_t)ev[kHexFaces[f][1]];
            const double* Dv = in.vx + 3 * (size_t)ev[kHexFaces[f][3]];
            double u[3], w[3], c[3];
            for (int d = 0; d < 3; ++d) {
                u[d] = B[d] - A[d];
                w[d] = Dv[d] - A[d];
            }
            c[0] = u[1] * w[2] - u[2] * w[1];
            c[1] = u[2] * w[0] - u[0] * w[2];
            c[2] = u[0] * w[1] - u[1] * w[0];
            const double area = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            for (int d = 0; d < 3; ++d) g[9 + 4 * f + d] = c[d] / area;
            const double fs = (area / 4.0) / J;  // surface Jacobian over volume Jacobian
            g[12 + 4 * f] = 0.5 * fs;
            md.max_fscale = std::max(md.max_fscale, fs);
        }
        g[33] = 1.0 / md.eps[e];
        g[34] = 1.0 / md.mu[e];
    }
    // faces by sorted vertex quadruple
    std::map<std::array<int, 4>, std::pair<int, int>> open;
    std::vector<int> nbr(6 * (size_t)K, -1), code(6 * (size_t)K, 0);
    auto orientation = [&](int e, int f, int e2, int f2) {
        // own cycle's first two vertices within the neighbour's cycle
        const int32_t* ev = in.EToV + 8 * (size_t)e;
        const int32_t* ev2 = in.EToV + 8 * (size_t)e2;
        const int a = ev[kHexFaces[f][0]], b = ev[kHexFaces[f][1]];
        int r = -1;
        for (int k = 0; k < 4; ++k)
            if (ev2[kHexFaces[f2][k]] == a) r = k;
        if (r < 0) throw std::runtime_error("hexahedral faces do not match");
        if (ev2[kHexFaces[f2][(r + 1) % 4]] == b) return 2 * r;
        if (ev2[kHexFaces[f2][(r + 3) % 4]] == b) return 2 * r + 1;
        throw std::runtime_error("hexahedral faces do not match");
    };
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 6; ++f) {
            std::array<int, 4> key;
            for (int k = 0; k < 4; ++k) key[k] = in.EToV[8 * (size_t)e + kHexFaces[f][k]];
            std::sort(key.begin(), key.end());
            if (key[0] == key[1] || key[1] == key[2] || key[2] == key[3])
                throw std::runtime_error("hexahedron with repeated vertex");
            auto it = open.find(key);
            if (it == open.end()) {
                open.emplace(key, std::make_pair(e, f));
                continue;
            }
            const int e2 = it->second.first, f2 = it->second.second;
            open.erase(it);
            if (e2 == e) throw std::runtime_error("hexahedron shares a face with itself");
            nbr[6 * (size_t)e + f] = e2;
            code[6 * (size_t)e + f] = f2 * 8 + orientation(e, f, e2, f2);
            nbr[6 * (size_t)e2 + f2] = e;
            code[6 * (size_t)e2 + f2] = f * 8 + orientation(e2, f2, e, f);
        }
    // effective codes: an interior condition or TF/SF code given on one side
    // of an interior face applies to both sides
    std::vector<int> eff(6 * (size_t)K, 0);
    for (size_t i = 0; i < eff.size(); ++i) {
        eff[i] = in.face_bc ? in.face_bc[i] : 0;
        if (eff[i] < 0 || eff[i] > 7) throw std::runtime_error("face_bc code out of range");
    }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 6; ++f) {
            const size_t i = 6 * (size_t)e + f;
            if (nbr[i] < 0 || eff[i] < 4) continue;
            const size_t j = 6 * (size_t)nbr[i] + code[i] / 8;
            if (eff[j] != 0 && eff[j] != eff[i])
                throw std::runtime_error("conflicting interior conditions on the two sides of a face");
            eff[j] = eff[i];
        }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 6; ++f) {
            const size_t i = 6 * (size_t)e + f;
            const int bc = eff[i];
            int32_t* cr = &md.conn[(size_t)kHexConnStride * e];
            if (nbr[i] < 0) {
                if (bc >= 4) throw std::runtime_error("interior-boundary codes 4-7 on a boundary face");
                if (bc == 3 && alpha == 0.0) throw std::runtime_error("SMA boundaries need upwind flux (alpha != 0)");
                cr[f] = -(bc + 1);
                cr[6 + f] = 0;
            } else if (bc == 7) {
                // TF/SF face: normal neighbour gather, plus the injected plane
                // waves; the lower-index element is the scattered-field side
                cr[f] = nbr[i];
                cr[6 + f] = code[i];
                md.tfsf.push_back({e, f, e < nbr[i] ? -1 : 1});
            } else if (bc >= 4) {
                cr[f] = -(bc + 1);  // interior condition: both sides use the own trace
                cr[6 + f] = 0;
            } else if (bc >= 1) {
                throw std::runtime_error("PEC/PMC/SMA codes 1-3 on an interior face (use 4-6)");
            } else {
                cr[f] = nbr[i];
                cr[6 + f] = code[i];
            }
            md.face_code[i] = (int8_t)bc;
        }
    return md;
}

void hex_node_positions(const MeshInput& in, const RefHex& R, double* xyz)
{
    const size_t N = (size_t)R.Np * in.K;
    for (int e = 0; e < in.K; ++e) {
        const HexFrame F = hex_frame(in.vx, in.EToV + 8 * (size_t)e);
        for (int n = 0; n < R.Np; ++n) {
            const size_t o = (size_t)e * R.Np + n;
            for (int d = 0; d < 3; ++d) {
                double x = F.o[d];
                for (int k = 0; k < 3; ++k) x += R.xi[3 * n + k] * F.a[k][d];
                xyz[d * N + o] = x;
            }
        }
    }
}

void hex_locate_points(const MeshInput& in, const RefHex& R, int n, const double* xyz, int32_t* elem, double* w)
{
    std::vector<HexFrame> frames(in.K);
    for (int e = 0; e < in.K; ++e) frames[e] = hex_frame(in.vx, in.EToV + 8 * (size_t)e);
    for (int p = 0; p < n; ++p) {
        elem[p] = -1;
        for (int e = 0; e < in.K && elem[p] < 0; ++e) {
            const HexFrame& F = frames[e];
            double rhs[3], xi[3];
            for (int d = 0; d < 3; ++d) rhs[d] = xyz[3 * (size_t)p + d] - F.o[d];
            // Cramer's rule on [a0 a1 a2] xi = rhs
            for (int k = 0; k < 3; ++k) {
                double m[3][3];
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d) m[c][d] = c == k ? rhs[d] : F.a[c][d];
                xi[k] = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) /
                        F.det;
            }
            const double tol = 1e-10;
            if (xi[0] >= -tol && xi[1] >= -tol && xi[2] >= -tol && xi[0] <= 1 + tol && xi[1] <= 1 + tol &&
                xi[2] <= 1 + tol) {
                elem[p] = e;
                hex_interp_weights(R, std::min(1.0, std::max(0.0, xi[0])), std::min(1.0, std::max(0.0, xi[1])),
                                   std::min(1.0, std::max(0.0, xi[2])), w + (size_t)p * R.Np);
            }
        }
    }
}

}  // namespace dgtd
