// Reference quadrilateral (tensor product of Gauss-Lobatto-Legendre points,
// MFEM's L2 Gauss-Lobatto quadrilateral: node i + (p+1) j at (x_i, x_j),
// lexicographic) and quadrilateral-mesh setup for the assembled operator.
#include "quad.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace dgtd {

const int kQuadEdges[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

namespace {

// Legendre P_0..P_n at x (three-term recurrence).
std::vector<double> legendre(int n, double x)
{
    std::vector<double> P(n + 1);
    P[0] = 1.0;
    if (n >= 1) P[1] = x;
    for (int k = 2; k <= n; ++k) P[k] = ((2 * k - 1) * x * P[k - 1] - (k - 1) * P[k - 2]) / k;
    return P;
}

// Gauss-Lobatto-Legendre points of degree p (p+1 points, ascending): the
// endpoints and the roots of P_p', by Newton from the Chebyshev-Gauss-Lobatto
// points with P_p' = p (P_{p-1} - x P_p) / (1 - x^2) and, from Legendre's
// equation, P_p'' = (2 x P_p' - p (p+1) P_p) / (1 - x^2).
std::vector<double> gll_points(int p)
{
    std::vector<double> x(p + 1);
    for (int k = 0; k <= p; ++k) x[k] = -std::cos(M_PI * k / p);
    for (int k = 1; k < p; ++k) {
        double t = x[k];
        for (int it = 0; it < 100; ++it) {
            const std::vector<double> P = legendre(p, t);
            const double d1 = p * (P[p - 1] - t * P[p]) / (1 - t * t);
            const double d2 = (2 * t * d1 - p * (p + 1) * P[p]) / (1 - t * t);
            const double dt = d1 / d2;
            t -= dt;
            if (std::fabs(dt) < 1e-16) break;
        }
        x[k] = t;
    }
    x[0] = -1.0;
    x[p] = 1.0;
    return x;
}

// Gauss-Legendre points and weights (n points): Newton on P_n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.resize(n);
    w.resize(n);
    for (int k = 0; k < n; ++k) {
        double t = std::cos(M_PI * (k + 0.75) / (n + 0.5));
        double dP = 1.0;
        for (int it = 0; it < 100; ++it) {
            const std::vector<double> P = legendre(n, t);
            dP = n * (P[n - 1] - t * P[n]) / (1 - t * t);
            const double dt = P[n] / dP;
            t -= dt;
            if (std::fabs(dt) < 1e-16) break;
        }
        const std::vector<double> P = legendre(n, t);
        dP = n * (P[n - 1] - t * P[n]) / (1 - t * t);
        x[k] = t;
        w[k] = 2.0 / ((1 - t * t) * dP * dP);
    }
}

// Lagrange basis on the nodes x: values at y (barycentric form).
std::vector<double> lagrange_values(const std::vector<double>& x, double y)
{
    const int n = (int)x.size();
    std::vector<double> l(n, 1.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
            if (k != j) l[j] *= (y - x[k]) / (x[j] - x[k]);
    return l;
}

std::vector<double> inverse(std::vector<double> A, int n)
{
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (std::fabs(A[(size_t)piv * n + c]) < 1e-300) throw std::runtime_error("reference quadrilateral: singular matrix");
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

}  // namespace

Gll1D build_gll1d(int p)
{
    Gll1D g;
    g.p = p;
    const int n1 = p + 1;
    g.x = gll_points(p);
    // differentiation D[i][j] = l_j'(x_i) (barycentric weights) and the exact
    // mass M[i][j] = int l_i l_j over [-1, 1] (Gauss-Legendre, p+2 points)
    std::vector<double> bw(n1, 1.0);
    for (int j = 0; j < n1; ++j)
        for (int k = 0; k < n1; ++k)
            if (k != j) bw[j] /= (g.x[j] - g.x[k]);
    g.D.assign((size_t)n1 * n1, 0.0);
    g.M.assign((size_t)n1 * n1, 0.0);
    for (int i = 0; i < n1; ++i) {
        double diag = 0.0;
        for (int j = 0; j < n1; ++j)
            if (j != i) {
                g.D[(size_t)i * n1 + j] = (bw[j] / bw[i]) / (g.x[i] - g.x[j]);
                diag -= g.D[(size_t)i * n1 + j];
            }
        g.D[(size_t)i * n1 + i] = diag;
    }
    std::vector<double> gx, gw;
    gauss_legendre(p + 2, gx, gw);
    for (size_t q = 0; q < gx.size(); ++q) {
        const std::vector<double> l = lagrange_values(g.x, gx[q]);
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n1; ++j) g.M[(size_t)i * n1 + j] += gw[q] * l[i] * l[j];
    }
    g.Minv = inverse(g.M, n1);
    return g;
}

std::vector<double> gll_lagrange(const std::vector<double>& x, double y) { return lagrange_values(x, y); }

RefQuad build_refquad(int p)
{
    if (p < 1 || p > 5) throw std::runtime_error("quadrilateral order must be in 1..5");
    RefQuad R;
    R.p = p;
    const int n1 = p + 1;
    R.Nfp = n1;
    R.Np = n1 * n1;
    const int Np = R.Np, Nfp = R.Nfp;
    const Gll1D g1 = build_gll1d(p);
    R.x1 = g1.x;
    const std::vector<double>& D1 = g1.D;
    const std::vector<double>& M1 = g1.M;
    const std::vector<double>& M1inv = g1.Minv;
    // 2D tensor products, node a = i + n1 j
    R.xi.resize(2 * (size_t)Np);
    R.M.assign((size_t)Np * Np, 0.0);
    R.Dr.assign((size_t)Np * Np, 0.0);
    R.Ds.assign((size_t)Np * Np, 0.0);
    for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n1; ++i) {
            const int a = i + n1 * j;
            R.xi[2 * a] = 0.5 * (R.x1[i] + 1.0);
            R.xi[2 * a + 1] = 0.5 * (R.x1[j] + 1.0);
            for (int l = 0; l < n1; ++l)
                for (int k = 0; k < n1; ++k) {
                    const int b = k + n1 * l;
                    R.M[(size_t)a * Np + b] = M1[(size_t)i * n1 + k] * M1[(size_t)j * n1 + l];
                    if (l == j) R.Dr[(size_t)a * Np + b] = D1[(size_t)i * n1 + k];
                    if (k == i) R.Ds[(size_t)a * Np + b] = D1[(size_t)j * n1 + l];
                }
        }
    // edges from their first vertex to their second: vertices 0 (-1,-1),
    // 1 (1,-1), 2 (1,1), 3 (-1,1)
    R.Fmask.resize(4 * (size_t)Nfp);
    for (int b = 0; b < Nfp; ++b) {
        R.Fmask[0 * Nfp + b] = b;                         // s = -1, i up
        R.Fmask[1 * Nfp + b] = p + n1 * b;                // r = 1, j up
        R.Fmask[2 * Nfp + b] = (p - b) + n1 * p;          // s = 1, i down
        R.Fmask[3 * Nfp + b] = n1 * (p - b);              // r = -1, j down
    }
    // LIFT = M^-1 E with E the edge mass (each edge of reference length 2).
    // With the tensor mass M1 x M1 the tangential factor cancels (M1^-1 M1 =
    // I; the points are symmetric, so a reversed edge gives the same
    // identity): edge node q lifts onto the normal line through it, weight
    // M1^-1[i, b] at line index i (b = 0 or p), exactly zero elsewhere.
    R.LIFT.assign((size_t)Np * 4 * Nfp, 0.0);
    for (int f = 0; f < 4; ++f) {
        const bool along_r = f == 1 || f == 3;  // edges r = +-1: the normal line runs along r
        const int b = (f == 1 || f == 2) ? p : 0;
        const int stride = along_r ? 1 : n1;
        for (int q = 0; q < Nfp; ++q) {
            const int c = R.Fmask[f * Nfp + q];
            const int base = c - b * stride;
            for (int i = 0; i < n1; ++i)
                R.LIFT[(size_t)(base + i * stride) * 4 * Nfp + f * Nfp + q] = M1inv[(size_t)i * n1 + b];
        }
    }
    // neighbour table: the neighbour edge g traverses the shared edge in the
    // same order (c = 0) or reversed (c = 1)
    R.edge_perm.assign(32 * (size_t)Nfp, -1);
    for (int f = 0; f < 4; ++f)
        for (int g = 0; g < 4; ++g)
            for (int c = 0; c < 2; ++c)
                for (int b = 0; b < Nfp; ++b)
                    R.edge_perm[((f * 4 + g) * 2 + c) * Nfp + b] = R.Fmask[g * Nfp + (c == 0 ? b : Nfp - 1 - b)];
    return R;
}

void quad_interp_weights(const RefQuad& R, double xi, double eta, double* w)
{
    const std::vector<double> lr = lagrange_values(R.x1, 2 * xi - 1), ls = lagrange_values(R.x1, 2 * eta - 1);
    const int n1 = R.p + 1;
    for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n1; ++i) w[i + n1 * j] = lr[i] * ls[j];
}

TriMeshData build_quad_mesh(const TriMeshInput& in, const RefQuad& R, double alpha)
{
    (void)R;
    return build_2d_mesh(in, 4, alpha);
}

void quad_node_positions(const TriMeshInput& in, const RefQuad& R, double* xyz)
{
    const size_t N = (size_t)R.Np * in.K;
    for (int e = 0; e < in.K; ++e) {
        const double* v[4];
        for (int k = 0; k < 4; ++k) v[k] = in.vx + 3 * (size_t)in.EToV[4 * (size_t)e + k];
        for (int n = 0; n < R.Np; ++n) {
            const double a = R.xi[2 * n], b = R.xi[2 * n + 1];
            const size_t o = (size_t)e * R.Np + n;
            for (int d = 0; d < 2; ++d) xyz[d * N + o] = v[0][d] + a * (v[1][d] - v[0][d]) + b * (v[3][d] - v[0][d]);
            xyz[2 * N + o] = 0.0;
        }
    }
}

void quad_locate_points(const TriMeshInput& in, const RefQuad& R, int n, const double* xyz, int32_t* elem, double* w)
{
    for (int p = 0; p < n; ++p) {
        elem[p] = -1;
        const double x = xyz[3 * (size_t)p], y = xyz[3 * (size_t)p + 1];
        for (int e = 0; e < in.K && elem[p] < 0; ++e) {
            const double* v[4];
            for (int k = 0; k < 4; ++k) v[k] = in.vx + 3 * (size_t)in.EToV[4 * (size_t)e + k];
            const double ax = v[1][0] - v[0][0], ay = v[1][1] - v[0][1];
            const double bx = v[3][0] - v[0][0], by = v[3][1] - v[0][1];
            const double det = ax * by - bx * ay;
            const double px = x - v[0][0], py = y - v[0][1];
            const double xi = (px * by - bx * py) / det, eta = (ax * py - px * ay) / det;
            const double tol = 1e-10;
            if (xi >= -tol && eta >= -tol && xi <= 1 + tol && eta <= 1 + tol) {
                elem[p] = e;
                quad_interp_weights(R, xi, eta, w + (size_t)p * R.Np);
            }
        }
    }
}

}  // namespace dgtd
