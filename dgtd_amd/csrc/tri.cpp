// Reference triangle (MFEM L2 Gauss-Lobatto nodes, orthonormal PKD basis,
// Hesthaven & Warburton 2008 ch. 6) and triangle-mesh setup for the 2D
// hesthaven operator.  Neighbour nodes are matched through integer lattice
// indices, never through coordinates.
#include "tri.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>

#include "reftet.hpp"

namespace dgtd {

const int kTriEdges[3][2] = {{0, 1}, {1, 2}, {2, 0}};
const int kQuadEdgeTable[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

namespace {

using Mat = std::vector<double>;  // row-major n x n

Mat inverse(Mat A, int n)
{
    Mat X((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (std::fabs(A[(size_t)piv * n + c]) < 1e-300) throw std::runtime_error("reference triangle: singular matrix");
        for (int k = 0; k < n; ++k) {
            std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
            std::swap(X[(size_t)c * n + k], X[(size_t)piv * n + k]);
        }
        const double d = A[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) {
            A[(size_t)c * n + k] /= d;
            X[(size_t)c * n + k] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                X[(size_t)r * n + k] -= f * X[(size_t)c * n + k];
            }
        }
    }
    return X;
}

Mat matmul(const Mat& A, const Mat& B, int n, int m, int k)  // (n x k) (k x m)
{
    Mat C((size_t)n * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int l = 0; l < k; ++l) {
            const double a = A[(size_t)i * k + l];
            for (int j = 0; j < m; ++j) C[(size_t)i * m + j] += a * B[(size_t)l * m + j];
        }
    return C;
}

// psi_ij(r,s) = sqrt(2) P_i(a) P_j^{(2i+1,0)}(b) (1-b)^i with the collapsed
// coordinates a = 2(1+r)/(1-s) - 1, b = s, and its r, s derivatives.
void pkd_tri(double r, double s, int i, int j, double& v, double& vr, double& vs)
{
    const double a = std::fabs(s - 1.0) > 1e-14 ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
    const double b = s;
    const double fa = jac(a, 0, 0, i), dfa = djac(a, 0, 0, i);
    const double gb = jac(b, 2 * i + 1, 0, j), dgb = djac(b, 2 * i + 1, 0, j);
    const double hb = 0.5 * (1.0 - b);
    v = std::sqrt(2.0) * fa * gb * std::pow(1.0 - b, i);
    double dr = dfa * gb;
    if (i > 0) dr *= std::pow(hb, i - 1);
    double ds = dfa * gb * 0.5 * (1.0 + a);
    if (i > 0) ds *= std::pow(hb, i - 1);
    double tb = dgb * std::pow(hb, i);
    if (i > 0) tb -= 0.5 * i * gb * std::pow(hb, i - 1);
    ds += fa * tb;
    const double sc = std::pow(2.0, i + 0.5);
    vr = dr * sc;
    vs = ds * sc;
}

}  // namespace

RefTri build_reftri(int p)
{
    if (p < 1 || p > 5) throw std::runtime_error("order must be in 1..5");
    RefTri R;
    R.p = p;
    R.Np = tri_num_nodes(p);
    R.Nfp = p + 1;
    const int Np = R.Np, Nfp = R.Nfp;
    // MFEM L2 GaussLobatto triangle: j ascending, i fastest, node =
    // (g_i, g_j) / (g_i + g_j + g_k), k = p - i - j (vertex 0 <- k, 1 <- i, 2 <- j)
    const std::vector<double> g = gll01(p);
    for (int j = 0; j <= p; ++j)
        for (int i = 0; i + j <= p; ++i) {
            const int k = p - i - j;
            const double w = g[i] + g[j] + g[k];
            R.xi.insert(R.xi.end(), {g[i] / w, g[j] / w});
            R.lattice.insert(R.lattice.end(), {k, i, j});
        }
    Mat V((size_t)Np * Np), Vr((size_t)Np * Np), Vs((size_t)Np * Np);
    for (int n = 0; n < Np; ++n) {
        const double r = 2 * R.xi[2 * n] - 1, s = 2 * R.xi[2 * n + 1] - 1;
        int m = 0;
        for (int i = 0; i <= p; ++i)
            for (int j = 0; i + j <= p; ++j, ++m)
                pkd_tri(r, s, i, j, V[(size_t)n * Np + m], Vr[(size_t)n * Np + m], Vs[(size_t)n * Np + m]);
    }
    const Mat Vi = inverse(V, Np);
    R.Vinv = Vi;
    Mat Minv((size_t)Np * Np, 0.0);  // V V^T
    for (int a = 0; a < Np; ++a)
        for (int b = 0; b < Np; ++b) {
            double s = 0;
            for (int m = 0; m < Np; ++m) s += V[(size_t)a * Np + m] * V[(size_t)b * Np + m];
            Minv[(size_t)a * Np + b] = s;
        }
    R.M = inverse(Minv, Np);
    R.Dr = matmul(Vr, Vi, Np, Np, Np);
    R.Ds = matmul(Vs, Vi, Np, Np, Np);

    // Edge nodes: lattice index of the opposite vertex is 0; ordered by the
    // lattice index of the edge's second vertex (first vertex -> second).
    R.Fmask.resize(3 * Nfp);
    for (int f = 0; f < 3; ++f) {
        const int va = kTriEdges[f][0], vb = kTriEdges[f][1], vo = 3 - va - vb;
        for (int n = 0; n < Np; ++n)
            if (R.lattice[3 * n + vo] == 0) R.Fmask[f * Nfp + R.lattice[3 * n + vb]] = n;
    }
    // Edge mass on [-1,1] from the 1D orthonormal Legendre basis at the edge
    // GLL points; LIFT = M^-1 E.
    Mat E((size_t)Np * 3 * Nfp, 0.0);
    {
        Mat V1((size_t)Nfp * Nfp);
        for (int b = 0; b < Nfp; ++b)
            for (int m = 0; m < Nfp; ++m) V1[(size_t)b * Nfp + m] = jac(2 * g[b] - 1, 0, 0, m);
        Mat M1inv((size_t)Nfp * Nfp, 0.0);
        for (int a = 0; a < Nfp; ++a)
            for (int b = 0; b < Nfp; ++b) {
                double s = 0;
                for (int m = 0; m < Nfp; ++m) s += V1[(size_t)a * Nfp + m] * V1[(size_t)b * Nfp + m];
                M1inv[(size_t)a * Nfp + b] = s;
            }
        const Mat M1 = inverse(M1inv, Nfp);
        for (int f = 0; f < 3; ++f)
            for (int a = 0; a < Nfp; ++a)
                for (int b = 0; b < Nfp; ++b)
                    E[(size_t)R.Fmask[f * Nfp + a] * 3 * Nfp + f * Nfp + b] = M1[(size_t)a * Nfp + b];
    }
    R.LIFT = matmul(Minv, E, Np, 3 * Nfp, Np);

    // Neighbour table: own edge node b sits at lattice position b along
    // (va -> vb); the neighbour edge g with orientation c holds it at position
    // b (same order) or Nfp-1-b (reversed).
    R.edge_perm.assign(18 * Nfp, -1);
    for (int f = 0; f < 3; ++f)
        for (int gg = 0; gg < 3; ++gg)
            for (int c = 0; c < 2; ++c)
                for (int b = 0; b < Nfp; ++b)
                    R.edge_perm[((f * 3 + gg) * 2 + c) * Nfp + b] = R.Fmask[gg * Nfp + (c == 0 ? b : Nfp - 1 - b)];
    return R;
}

void tri_interp_weights(const RefTri& R, double xi, double eta, double* w)
{
    const int p = R.p, Np = R.Np;
    std::vector<double> P(Np);
    int m = 0;
    for (int i = 0; i <= p; ++i)
        for (int j = 0; i + j <= p; ++j, ++m) {
            double vr, vs;
            pkd_tri(2 * xi - 1, 2 * eta - 1, i, j, P[m], vr, vs);
        }
    for (int n = 0; n < Np; ++n) {
        double acc = 0.0;
        for (int mm = 0; mm < Np; ++mm) acc += R.Vinv[(size_t)mm * Np + n] * P[mm];
        w[n] = acc;
    }
}

TriMeshData build_tri_mesh(const TriMeshInput& in, const RefTri& R, double alpha)
{
    (void)R;
    return build_2d_mesh(in, 3, alpha);
}

TriMeshData build_2d_mesh(const TriMeshInput& in, int NV, double alpha)
{
    const int K = in.K;
    const bool quad = NV == 4;
    const int (*edges)[2] = quad ? kQuadEdgeTable : kTriEdges;
    const int stride = quad ? 18 : kTriGeoStride, cstride = quad ? 8 : kTriConnStride;
    const char* what = quad ? "quadrilateral" : "triangle";
    if (K <= 0 || in.nv <= 0 || !in.vx || !in.EToV)
        throw std::runtime_error(std::string("empty or incomplete ") + what + " mesh");
    for (size_t i = 0; i < (size_t)NV * K; ++i)
        if (in.EToV[i] < 0 || in.EToV[i] >= in.nv) throw std::runtime_error("EToV vertex id out of range");
    TriMeshData md;
    md.K = K;
    md.geo.assign((size_t)stride * K, 0.0);
    md.conn.assign((size_t)cstride * K, 0);
    md.jac.resize(K);
    md.eps.assign(K, 1.0);
    md.mu.assign(K, 1.0);
    md.face_code.assign((size_t)NV * K, 0);
    for (int e = 0; e < K; ++e) {
        if (in.eps_r) md.eps[e] = in.eps_r[e];
        if (in.mu_r) md.mu[e] = in.mu_r[e];
        if (!(md.eps[e] > 0) || !(md.mu[e] > 0)) throw std::runtime_error("eps_r and mu_r must be positive");
        const double* v[4];
        for (int k = 0; k < NV; ++k) v[k] = in.vx + 3 * (size_t)in.EToV[(size_t)NV * e + k];
        // affine map from [-1,1]^2 (quadrilateral, vertices 1 and 3 on the r and
        // s axes) or from the reference triangle (vertices 1 and 2)
        const double* vs = quad ? v[3] : v[2];
        const double xr = 0.5 * (v[1][0] - v[0][0]), yr = 0.5 * (v[1][1] - v[0][1]);
        const double xs = 0.5 * (vs[0] - v[0][0]), ys = 0.5 * (vs[1] - v[0][1]);
        const double J = xr * ys - xs * yr;
        if (!(std::fabs(J) > 0)) throw std::runtime_error(std::string("degenerate ") + what);
        if (quad) {  // a parallelogram: v2 = v1 + v3 - v0
            const double scale = std::fabs(xr) + std::fabs(yr) + std::fabs(xs) + std::fabs(ys);
            for (int d = 0; d < 2; ++d)
                if (std::fabs(v[2][d] - (v[1][d] + v[3][d] - v[0][d])) > 1e-12 * scale)
                    throw std::runtime_error("quadrilaterals must be parallelograms (affine elements)");
        }
        md.jac[e] = std::fabs(J);
        double* g = &md.geo[(size_t)stride * e];
        g[0] = ys / J;   // dr/dx
        g[1] = -xs / J;  // dr/dy
        g[2] = -yr / J;  // ds/dx
        g[3] = xr / J;   // ds/dy
        const double sg = J > 0 ? 1.0 : -1.0;
        for (int f = 0; f < NV; ++f) {
            const double* a = v[edges[f][0]];
            const double* b = v[edges[f][1]];
            const double tx = b[0] - a[0], ty = b[1] - a[1];
            const double L = std::hypot(tx, ty);
            g[4 + 3 * f] = sg * ty / L;
            g[5 + 3 * f] = -sg * tx / L;
            const double fs = (0.5 * L) / std::fabs(J);
            g[6 + 3 * f] = 0.5 * fs;
            md.max_fscale = std::max(md.max_fscale, fs);
        }
        g[4 + 3 * NV] = 1.0 / md.eps[e];
        g[5 + 3 * NV] = 1.0 / md.mu[e];
    }
    // edges by sorted vertex pair
    std::map<std::pair<int, int>, std::pair<int, int>> open;
    std::vector<int> nbr((size_t)NV * K, -1), code((size_t)NV * K, 0);
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < NV; ++f) {
            const int a = in.EToV[(size_t)NV * e + edges[f][0]], b = in.EToV[(size_t)NV * e + edges[f][1]];
            if (a == b) throw std::runtime_error(std::string(what) + " with repeated vertex");
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = open.find(key);
            if (it == open.end()) {
                open.emplace(key, std::make_pair(e, f));
                continue;
            }
            const int e2 = it->second.first, f2 = it->second.second;
            open.erase(it);
            if (e2 == e) throw std::runtime_error(std::string(what) + " shares an edge with itself");
            const int a2 = in.EToV[(size_t)NV * e2 + edges[f2][0]];
            const int c = (a2 == a) ? 0 : 1;  // same or opposite traversal
            nbr[(size_t)NV * e + f] = e2;
            code[(size_t)NV * e + f] = (f * NV + f2) * 2 + c;
            nbr[(size_t)NV * e2 + f2] = e;
            code[(size_t)NV * e2 + f2] = (f2 * NV + f) * 2 + c;
        }
    // effective codes: an interior condition or TF/SF code given on one side
    // of an interior edge applies to both sides
    std::vector<int> eff((size_t)NV * K, 0);
    for (size_t i = 0; i < eff.size(); ++i) {
        eff[i] = in.face_bc ? in.face_bc[i] : 0;
        if (eff[i] < 0 || eff[i] > 7) throw std::runtime_error("face_bc code out of range");
    }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < NV; ++f) {
            const size_t i = (size_t)NV * e + f;
            if (nbr[i] < 0 || eff[i] < 4) continue;
            const size_t j = (size_t)NV * nbr[i] + (code[i] / 2) % NV;
            if (eff[j] != 0 && eff[j] != eff[i])
                throw std::runtime_error("conflicting interior conditions on the two sides of an edge");
            eff[j] = eff[i];
        }
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < NV; ++f) {
            const size_t i = (size_t)NV * e + f;
            const int bc = eff[i];
            int32_t* cr = &md.conn[(size_t)cstride * e];
            if (nbr[i] < 0) {
                if (bc >= 4) throw std::runtime_error("interior-boundary codes 4-7 on a boundary edge");
                if (bc == 3 && alpha == 0.0) throw std::runtime_error("SMA boundaries need upwind flux (alpha != 0)");
                cr[f] = -(bc + 1);
                cr[NV + f] = 0;
            } else if (bc == 7) {
                // TF/SF edge: normal neighbour gather, plus the injected plane
                // waves; the lower-index element is the scattered-field side
                cr[f] = nbr[i];
                cr[NV + f] = code[i];
                md.tfsf.push_back({e, f, e < nbr[i] ? -1 : 1});
            } else if (bc >= 4) {
                // interior condition: both sides use the own trace (no gather)
                cr[f] = -(bc + 1);
                cr[NV + f] = 0;
            } else if (bc >= 1) {
                throw std::runtime_error("PEC/PMC/SMA codes 1-3 on an interior edge (use 4-6)");
            } else {
                cr[f] = nbr[i];
                cr[NV + f] = code[i];
            }
            md.face_code[i] = (int8_t)bc;
        }
    return md;
}

void tri_node_positions(const TriMeshInput& in, const RefTri& R, double* xyz)
{
    const size_t N = (size_t)R.Np * in.K;
    for (int e = 0; e < in.K; ++e) {
        const double* v[3];
        for (int k = 0; k < 3; ++k) v[k] = in.vx + 3 * (size_t)in.EToV[3 * (size_t)e + k];
        for (int n = 0; n < R.Np; ++n) {
            const double a = R.xi[2 * n], b = R.xi[2 * n + 1];
            const size_t o = (size_t)e * R.Np + n;
            for (int d = 0; d < 2; ++d) xyz[d * N + o] = v[0][d] + a * (v[1][d] - v[0][d]) + b * (v[2][d] - v[0][d]);
            xyz[2 * N + o] = 0.0;
        }
    }
}

void tri_locate_points(const TriMeshInput& in, const RefTri& R, int n, const double* xyz, int32_t* elem, double* w)
{
    for (int p = 0; p < n; ++p) {
        elem[p] = -1;
        const double x = xyz[3 * (size_t)p], y = xyz[3 * (size_t)p + 1];
        for (int e = 0; e < in.K && elem[p] < 0; ++e) {
            const double* v[3];
            for (int k = 0; k < 3; ++k) v[k] = in.vx + 3 * (size_t)in.EToV[3 * (size_t)e + k];
            const double ax = v[1][0] - v[0][0], ay = v[1][1] - v[0][1];
            const double bx = v[2][0] - v[0][0], by = v[2][1] - v[0][1];
            const double det = ax * by - bx * ay;
            const double px = x - v[0][0], py = y - v[0][1];
            const double xi = (px * by - bx * py) / det, eta = (ax * py - px * ay) / det;
            const double tol = 1e-10;
            if (xi >= -tol && eta >= -tol && xi + eta <= 1 + tol) {
                elem[p] = e;
                tri_interp_weights(R, xi, eta, w + (size_t)p * R.Np);
            }
        }
    }
}

}  // namespace dgtd
