// Host setup: affine geometric factors and face connectivity.
//
// Geometry follows the reference's per-element data (HesthavenEvolution.cpp
// :150-205): physical derivative = metric-weighted reference derivatives
// (the matrices D_d = M^-1 S_d Vref/V_e of :158 reduce to this on affine
// tetrahedra), unit outward face normals, fscale = 2 sJ/J = 2 A_f/(3 V).
// Connectivity restates Connectivities::global (HesthavenEvolutionMethods.cpp
// :721-791) in compact form: neighbour id plus a vertex-permutation code
// instead of one (vmapM, vmapP) pair per face node.
#include "mesh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>

namespace dgtd {

namespace {

struct FaceKey {
    int32_t v[3];
    int32_t ef;  // e*4 + f
};

}  // namespace

MeshData build_mesh(const MeshInput& in, const RefTet& R, double alpha)
{
    if (in.K <= 0) throw std::runtime_error("mesh has no elements");
    if (!in.vx || !in.EToV) throw std::runtime_error("mesh vertex or element array is NULL");
    const int K = in.K;
    MeshData m;
    m.K = K;
    m.geo.assign((size_t)kGeoStride * K, 0.0);
    m.conn.assign((size_t)kConnStride * K, 0);
    m.face_code.assign((size_t)4 * K, 0);
    m.jac.resize(K);
    m.eps.resize(K);
    m.mu.resize(K);

    for (int e = 0; e < K; ++e) {
        double v[4][3];
        for (int a = 0; a < 4; ++a) {
            const int id = in.EToV[4 * (size_t)e + a];
            if (id < 0 || id >= in.nv)
                throw std::runtime_error("EToV entry out of range at element " + std::to_string(e));
            for (int d = 0; d < 3; ++d) v[a][d] = in.vx[3 * (size_t)id + d];
        }
        // A[d][k] = dx_d/dxi_k for x = v0 + sum_k (v_{k+1} - v0) xi_k
        double A[3][3];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 3; ++k) A[d][k] = v[k + 1][d] - v[0][d];
        const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
        const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
        const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
        const double vol = std::fabs(det) / 6.0;
        if (!(vol > 0.0) || !std::isfinite(det))
            throw std::runtime_error("degenerate element " + std::to_string(e));
        // gradient of xi_k: row k of A^-1
        double G[3][3];
        G[0][0] = c00 / det;
        G[1][0] = c01 / det;
        G[2][0] = c02 / det;
        G[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
        G[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
        G[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
        G[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
        G[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
        G[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
        double* g = &m.geo[(size_t)kGeoStride * e];
        // r = 2 xi - 1  =>  dr/dx = 2 dxi/dx
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) g[3 * k + d] = 2.0 * G[k][d];
        // barycentric gradients; lambda_0 = 1 - sum xi_k
        double gl[4][3];
        for (int d = 0; d < 3; ++d) {
            gl[1][d] = G[0][d];
            gl[2][d] = G[1][d];
            gl[3][d] = G[2][d];
            gl[0][d] = -(G[0][d] + G[1][d] + G[2][d]);
        }
        for (int f = 0; f < 4; ++f) {
            const double nrm = std::sqrt(gl[f][0] * gl[f][0] + gl[f][1] * gl[f][1] + gl[f][2] * gl[f][2]);
            g[9 + 4 * f + 0] = -gl[f][0] / nrm;
            g[9 + 4 * f + 1] = -gl[f][1] / nrm;
            g[9 + 4 * f + 2] = -gl[f][2] / nrm;
            g[9 + 4 * f + 3] = nrm;  // fscale/2 = |grad lambda_f| = A_f/(3V)
            m.max_fscale = std::max(m.max_fscale, 2.0 * nrm);
        }
        const double er = in.eps_r ? in.eps_r[e] : 1.0;
        const double mr = in.mu_r ? in.mu_r[e] : 1.0;
        if (!(er > 0.0) || !(mr > 0.0) || !std::isfinite(er) || !std::isfinite(mr))
            throw std::runtime_error("eps_r and mu_r must be positive and finite");
        g[25] = 1.0 / er;
        g[26] = 1.0 / mr;
        m.eps[e] = er;
        m.mu[e] = mr;
        m.jac[e] = vol / (4.0 / 3.0);
    }

    // Face matching by sorted vertex triple.
    std::vector<FaceKey> keys((size_t)4 * K);
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 4; ++f) {
            FaceKey& k = keys[4 * (size_t)e + f];
            for (int a = 0; a < 3; ++a) k.v[a] = in.EToV[4 * (size_t)e + kFaceVerts[f][a]];
            std::sort(k.v, k.v + 3);
            k.ef = 4 * e + f;
        }
    std::sort(keys.begin(), keys.end(), [](const FaceKey& a, const FaceKey& b) {
        if (a.v[0] != b.v[0]) return a.v[0] < b.v[0];
        if (a.v[1] != b.v[1]) return a.v[1] < b.v[1];
        if (a.v[2] != b.v[2]) return a.v[2] < b.v[2];
        return a.ef < b.ef;
    });
    auto same = [](const FaceKey& a, const FaceKey& b) {
        return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2];
    };
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && same(keys[i], keys[j])) ++j;
        if (j - i > 2) throw std::runtime_error("non-manifold mesh: a face is shared by more than two elements");
        if (j - i == 2) {
            // interior face: 0, an interior condition 4-6 or the TF/SF code 7,
            // given on either side
            int ic = 0;
            if (in.face_bc) {
                const int c0 = in.face_bc[keys[i].ef], c1 = in.face_bc[keys[i + 1].ef];
                if (c0 < 0 || c0 > 7 || c1 < 0 || c1 > 7) throw std::runtime_error("face_bc code must be 0..7");
                if (c0 != 0 && c1 != 0 && c0 != c1)
                    throw std::runtime_error("the two sides of an interior face carry different codes");
                ic = c0 != 0 ? c0 : c1;
                if (ic >= 1 && ic <= 3)
                    throw std::runtime_error("PEC/PMC/SMA codes 1-3 on an interior face (interior conditions are 4-6)");
            }
            for (int side = 0; side < 2; ++side) {
                const int ef = keys[i + side].ef, og = keys[i + 1 - side].ef;
                const int e = ef / 4, f = ef % 4, e2 = og / 4, g = og % 4;
                m.face_code[ef] = (int8_t)ic;
                if (ic == 7) {  // TF/SF: the lower-index element is the scattered-field side
                    m.tfsf.push_back({e, f, e < e2 ? -1 : 1});
                    m.n_boundary_faces[7]++;
                }
                if (ic >= 4 && ic <= 6) {  // the jump uses the own trace only: no neighbour gather
                    m.conn[(size_t)kConnStride * e + f] = -(ic + 1);
                    m.conn[(size_t)kConnStride * e + 4 + f] = 0;
                    m.n_boundary_faces[ic]++;
                    continue;
                }
                int pi[3];
                for (int k = 0; k < 3; ++k) {
                    const int vid = in.EToV[4 * (size_t)e + kFaceVerts[f][k]];
                    pi[k] = -1;
                    for (int k2 = 0; k2 < 3; ++k2)
                        if (in.EToV[4 * (size_t)e2 + kFaceVerts[g][k2]] == vid) pi[k] = k2;
                }
                const int pc = perm_code(pi);
                if (pc < 0) throw std::runtime_error("repeated vertex in an element");
                m.conn[(size_t)kConnStride * e + f] = e2;
                m.conn[(size_t)kConnStride * e + 4 + f] = (f * 4 + g) * 6 + pc;
            }
        } else {
            const int ef = keys[i].ef, e = ef / 4, f = ef % 4;
            const int bc = in.face_bc ? in.face_bc[ef] : 0;
            if (bc < 0 || bc > 7) throw std::runtime_error("face_bc code must be 0..7");
            if (bc >= 4) throw std::runtime_error("interior-boundary codes 4-7 on a boundary face");
            m.face_code[ef] = (int8_t)bc;
            if (bc == 3 && alpha == 0.0)
                throw std::runtime_error("SMA boundaries require upwind flux (alpha != 0)");
            m.conn[(size_t)kConnStride * e + f] = -(bc + 1);
            m.conn[(size_t)kConnStride * e + 4 + f] = 0;
            m.n_boundary_faces[bc]++;
        }
        i = j;
    }
    (void)R;
    return m;
}

void node_positions(const MeshInput& in, const RefTet& R, double* xyz)
{
    const size_t N = (size_t)R.Np * in.K;
    for (int e = 0; e < in.K; ++e) {
        const int32_t* ev = &in.EToV[4 * (size_t)e];
        for (int n = 0; n < R.Np; ++n) {
            const double a = R.xi[3 * n], b = R.xi[3 * n + 1], c = R.xi[3 * n + 2];
            for (int d = 0; d < 3; ++d) {
                const double x0 = in.vx[3 * (size_t)ev[0] + d];
                xyz[d * N + (size_t)e * R.Np + n] = x0 + (in.vx[3 * (size_t)ev[1] + d] - x0) * a +
                                                    (in.vx[3 * (size_t)ev[2] + d] - x0) * b +
                                                    (in.vx[3 * (size_t)ev[3] + d] - x0) * c;
            }
        }
    }
}

void locate_points(const MeshInput& in, const RefTet& R, int n, const double* xyz, int32_t* elem, double* w)
{
    auto det3 = [](const double A[3][3]) {
        return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    };
    const double tol = 1e-10;
    for (int p = 0; p < n; ++p) {
        const double* x = xyz + 3 * (size_t)p;
        elem[p] = -1;
        std::fill(w + (size_t)p * R.Np, w + (size_t)(p + 1) * R.Np, 0.0);
        for (int e = 0; e < in.K; ++e) {
            const int32_t* v = &in.EToV[4 * (size_t)e];
            const double* x0 = &in.vx[3 * (size_t)v[0]];
            double A[3][3], b[3];
            for (int d = 0; d < 3; ++d) {
                for (int k = 0; k < 3; ++k) A[d][k] = in.vx[3 * (size_t)v[k + 1] + d] - x0[d];
                b[d] = x[d] - x0[d];
            }
            const double det = det3(A);
            double xi[3];
            for (int k = 0; k < 3; ++k) {  // Cramer's rule for x = x0 + A xi
                double M[3][3];
                for (int d = 0; d < 3; ++d)
                    for (int j = 0; j < 3; ++j) M[d][j] = j == k ? b[d] : A[d][j];
                xi[k] = det3(M) / det;
            }
            if (xi[0] >= -tol && xi[1] >= -tol && xi[2] >= -tol && xi[0] + xi[1] + xi[2] <= 1 + tol) {
                elem[p] = e;
                interp_weights(R, xi, w + (size_t)p * R.Np);
                break;
            }
        }
    }
}

void cube_mesh(int n, double L, const int8_t side_bc[6], double* vx, int32_t* EToV, int8_t* face_bc)
{
    const int n1 = n + 1;
    auto vid = [n1](int i, int j, int k) { return i + n1 * (j + n1 * k); };
    if (vx)
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) {
                    double* p = &vx[3 * (size_t)vid(i, j, k)];
                    p[0] = i * (L / n);
                    p[1] = j * (L / n);
                    p[2] = k * (L / n);
                }
    if (!EToV && !face_bc) return;
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    size_t e = 0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int t = 0; t < 6; ++t, ++e) {
                    int c[4][3];
                    int cur[3] = {0, 0, 0};
                    for (int d = 0; d < 3; ++d) c[0][d] = 0;
                    for (int s = 0; s < 3; ++s) {
                        cur[perms[t][s]] += 1;
                        for (int d = 0; d < 3; ++d) c[s + 1][d] = cur[d];
                    }
                    // orientation of (c1-c0, c2-c0, c3-c0); swap local 2 and 3 if negative
                    int a[3][3];
                    for (int s = 0; s < 3; ++s)
                        for (int d = 0; d < 3; ++d) a[s][d] = c[s + 1][d] - c[0][d];
                    const int det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                                    a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                                    a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
                    if (det < 0)
                        for (int d = 0; d < 3; ++d) std::swap(c[2][d], c[3][d]);
                    if (EToV)
                        for (int s = 0; s < 4; ++s)
                            EToV[4 * e + s] = vid(i + c[s][0], j + c[s][1], k + c[s][2]);
                    if (face_bc)
                        for (int f = 0; f < 4; ++f) {
                            int8_t code = 0;
                            for (int ax = 0; ax < 3; ++ax) {
                                const int base = ax == 0 ? i : (ax == 1 ? j : k);
                                bool lo = true, hi = true;
                                for (int q = 0; q < 3; ++q) {
                                    const int cc = base + c[kFaceVerts[f][q]][ax];
                                    lo = lo && cc == 0;
                                    hi = hi && cc == n;
                                }
                                if (lo) code = side_bc ? side_bc[2 * ax] : 0;
                                if (hi) code = side_bc ? side_bc[2 * ax + 1] : 0;
                            }
                            face_bc[4 * e + f] = code;
                        }
                }
}

}  // namespace dgtd
