// Reference tetrahedron: MFEM's L2 Gauss-Lobatto node set, the orthonormal
// Proriol-Koornwinder-Dubiner basis (Hesthaven & Warburton 2008, ch. 6 and 10),
// Vandermonde-based mass, derivative and LIFT matrices, and the face-node
// permutation table used for neighbour gathers.
#include "reftet.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace dgtd {

const int kFaceVerts[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

namespace {

// Small dense row-major matrix.
struct Dense {
    int r = 0, c = 0;
    std::vector<double> a;
    Dense() = default;
    Dense(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
    double& operator()(int i, int j) { return a[(size_t)i * c + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

Dense mul(const Dense& A, const Dense& B)
{
    Dense C(A.r, B.c);
    for (int i = 0; i < A.r; ++i)
        for (int k = 0; k < A.c; ++k) {
            const double v = A(i, k);
            if (v == 0.0) continue;
            for (int j = 0; j < B.c; ++j) C(i, j) += v * B(k, j);
        }
    return C;
}

Dense mul_bt(const Dense& A, const Dense& B)  // A * B^T
{
    Dense C(A.r, B.r);
    for (int i = 0; i < A.r; ++i)
        for (int j = 0; j < B.r; ++j) {
            double s = 0;
            for (int k = 0; k < A.c; ++k) s += A(i, k) * B(j, k);
            C(i, j) = s;
        }
    return C;
}

// LU with partial pivoting, then solve for the identity.
Dense inv(const Dense& A)
{
    const int n = A.r;
    Dense LU = A;
    std::vector<int> piv(n);
    for (int i = 0; i < n; ++i) piv[i] = i;
    for (int k = 0; k < n; ++k) {
        int m = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(LU(i, k)) > std::fabs(LU(m, k))) m = i;
        if (LU(m, k) == 0.0) throw std::runtime_error("reference element: singular Vandermonde");
        if (m != k) {
            for (int j = 0; j < n; ++j) std::swap(LU(k, j), LU(m, j));
            std::swap(piv[k], piv[m]);
        }
        for (int i = k + 1; i < n; ++i) {
            LU(i, k) /= LU(k, k);
            const double l = LU(i, k);
            for (int j = k + 1; j < n; ++j) LU(i, j) -= l * LU(k, j);
        }
    }
    Dense X(n, n);
    std::vector<double> y(n);
    for (int col = 0; col < n; ++col) {
        for (int i = 0; i < n; ++i) {
            double s = (piv[i] == col) ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) s -= LU(i, j) * y[j];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < n; ++j) s -= LU(i, j) * X(j, col);
            X(i, col) = s / LU(i, i);
        }
    }
    return X;
}

}  // namespace

// Orthonormal Jacobi polynomial P_n^{(a,b)} on [-1,1] (weight (1-x)^a (1+x)^b).
double jac(double x, double a, double b, int n)
{
    const double g0 = std::pow(2.0, a + b + 1.0) / (a + b + 1.0) * std::tgamma(a + 1.0) *
                      std::tgamma(b + 1.0) / std::tgamma(a + b + 1.0);
    double pm = 1.0 / std::sqrt(g0);
    if (n == 0) return pm;
    const double g1 = (a + 1.0) * (b + 1.0) / (a + b + 3.0) * g0;
    double pc = ((a + b + 2.0) * 0.5 * x + 0.5 * (a - b)) / std::sqrt(g1);
    double ao = 2.0 / (2.0 + a + b) * std::sqrt((a + 1.0) * (b + 1.0) / (a + b + 3.0));
    for (int i = 1; i < n; ++i) {
        const double h = 2.0 * i + a + b;
        const double an = 2.0 / (h + 2.0) *
                          std::sqrt((i + 1.0) * (i + 1.0 + a + b) * (i + 1.0 + a) * (i + 1.0 + b) /
                                    ((h + 1.0) * (h + 3.0)));
        const double bn = -(a * a - b * b) / (h * (h + 2.0));
        const double pn = ((x - bn) * pc - ao * pm) / an;
        pm = pc;
        pc = pn;
        ao = an;
    }
    return pc;
}

double djac(double x, double a, double b, int n)
{
    return n == 0 ? 0.0 : std::sqrt(n * (n + a + b + 1.0)) * jac(x, a + 1.0, b + 1.0, n - 1);
}

// Gauss-Lobatto-Legendre points on [0,1]: the ends and the roots of P_p',
// by Newton's method on P_p' seeded with Chebyshev-Lobatto points.
std::vector<double> gll01(int p)
{
    std::vector<double> x(p + 1);
    x[0] = -1.0;
    x[p] = 1.0;
    for (int i = 1; i < p; ++i) {
        double t = -std::cos(M_PI * i / p);
        for (int it = 0; it < 60; ++it) {
            double q0 = 1.0, q1 = t;
            for (int k = 2; k <= p; ++k) {
                const double q2 = ((2.0 * k - 1.0) * t * q1 - (k - 1.0) * q0) / k;
                q0 = q1;
                q1 = q2;
            }
            const double dP = p * (t * q1 - q0) / (t * t - 1.0);
            const double d2P = (2.0 * t * dP - p * (p + 1.0) * q1) / (1.0 - t * t);
            const double step = dP / d2P;
            t -= step;
            if (std::fabs(step) < 1e-16) break;
        }
        x[i] = t;
    }
    for (int i = 0; i <= p / 2; ++i) {
        const double m = 0.5 * (x[p - i] - x[i]);
        x[i] = -m;
        x[p - i] = m;
    }
    if (p % 2 == 0) x[p / 2] = 0.0;
    for (auto& v : x) v = 0.5 * (v + 1.0);
    return x;
}

namespace {

// Basis psi_{ijk}(r,s,t) and its gradient in collapsed coordinates (a,b,c).
void pkd3(double r, double s, double t, int i, int j, int k, double& v, double& vr, double& vs,
          double& vt)
{
    const double a = (std::fabs(s + t) > 1e-14) ? 2.0 * (1.0 + r) / (-s - t) - 1.0 : -1.0;
    const double b = (std::fabs(t - 1.0) > 1e-14) ? 2.0 * (1.0 + s) / (1.0 - t) - 1.0 : -1.0;
    const double c = t;
    const double fa = jac(a, 0, 0, i), dfa = djac(a, 0, 0, i);
    const double gb = jac(b, 2 * i + 1, 0, j), dgb = djac(b, 2 * i + 1, 0, j);
    const double hc = jac(c, 2 * (i + j) + 2, 0, k), dhc = djac(c, 2 * (i + j) + 2, 0, k);
    const double ob = 0.5 * (1.0 - b), oc = 0.5 * (1.0 - c);
    v = 2.0 * std::sqrt(2.0) * fa * gb * std::pow(1.0 - b, i) * hc * std::pow(1.0 - c, i + j);

    double dr = dfa * gb * hc;
    if (i > 0) dr *= std::pow(ob, i - 1);
    if (i + j > 0) dr *= std::pow(oc, i + j - 1);
    double tb = dgb * std::pow(ob, i);
    if (i > 0) tb -= 0.5 * i * gb * std::pow(ob, i - 1);
    if (i + j > 0) tb *= std::pow(oc, i + j - 1);
    tb *= fa * hc;
    const double ds = 0.5 * (1.0 + a) * dr + tb;
    double tc = dhc * std::pow(oc, i + j);
    if (i + j > 0) tc -= 0.5 * (i + j) * hc * std::pow(oc, i + j - 1);
    tc *= fa * gb * std::pow(ob, i);
    const double dt = 0.5 * (1.0 + a) * dr + 0.5 * (1.0 + b) * tb + tc;
    const double sc = std::pow(2.0, 2 * i + j + 1.5);
    vr = dr * sc;
    vs = ds * sc;
    vt = dt * sc;
}

double pkd2(double r, double s, int i, int j)
{
    const double a = (std::fabs(s - 1.0) > 1e-14) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
    return std::sqrt(2.0) * jac(a, 0, 0, i) * jac(s, 2 * i + 1, 0, j) * std::pow(1.0 - s, i);
}

const int kPerms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

}  // namespace

void interp_weights(const RefTet& R, const double xi[3], double* w)
{
    const int p = R.p, Np = R.Np;
    const double r = 2 * xi[0] - 1, s = 2 * xi[1] - 1, t = 2 * xi[2] - 1;
    std::vector<double> P(Np);
    int m = 0;
    for (int i = 0; i <= p; ++i)
        for (int j = 0; i + j <= p; ++j)
            for (int k = 0; i + j + k <= p; ++k, ++m) {
                double vr, vs, vt;
                pkd3(r, s, t, i, j, k, P[m], vr, vs, vt);
            }
    for (int n = 0; n < Np; ++n) {
        double acc = 0.0;
        for (int mm = 0; mm < Np; ++mm) acc += R.Vinv[(size_t)mm * Np + n] * P[mm];
        w[n] = acc;
    }
}

int perm_code(const int pi[3])
{
    for (int c = 0; c < 6; ++c)
        if (kPerms[c][0] == pi[0] && kPerms[c][1] == pi[1] && kPerms[c][2] == pi[2]) return c;
    return -1;
}

void perm_from_code(int code, int pi[3])
{
    for (int k = 0; k < 3; ++k) pi[k] = kPerms[code][k];
}

RefTet build_reftet(int p)
{
    if (p < 1 || p > kMaxOrder) throw std::runtime_error("order must be in 1..5");
    RefTet R;
    R.p = p;
    R.Np = num_nodes(p);
    R.Nfp = num_face_nodes(p);
    const int Np = R.Np, Nfp = R.Nfp;

    // MFEM L2 GaussLobatto tetrahedron nodes: k, j, i ascending (i fastest),
    // node = (g_i, g_j, g_k) / (g_i + g_j + g_k + g_l), l = p - i - j - k.
    const std::vector<double> g = gll01(p);
    for (int k = 0; k <= p; ++k)
        for (int j = 0; j + k <= p; ++j)
            for (int i = 0; i + j + k <= p; ++i) {
                const int l = p - i - j - k;
                const double w = g[i] + g[j] + g[k] + g[l];
                R.xi.insert(R.xi.end(), {g[i] / w, g[j] / w, g[k] / w});
                R.lattice.insert(R.lattice.end(), {l, i, j, k});
            }

    Dense V(Np, Np), Vr(Np, Np), Vs(Np, Np), Vt(Np, Np);
    for (int n = 0; n < Np; ++n) {
        const double r = 2 * R.xi[3 * n] - 1, s = 2 * R.xi[3 * n + 1] - 1, t = 2 * R.xi[3 * n + 2] - 1;
        int m = 0;
        for (int i = 0; i <= p; ++i)
            for (int j = 0; i + j <= p; ++j)
                for (int k = 0; i + j + k <= p; ++k, ++m)
                    pkd3(r, s, t, i, j, k, V(n, m), Vr(n, m), Vs(n, m), Vt(n, m));
    }
    const Dense Vi = inv(V);
    R.Vinv = Vi.a;
    const Dense Minv = mul_bt(V, V);
    R.M = inv(Minv).a;
    R.Dr = mul(Vr, Vi).a;
    R.Ds = mul(Vs, Vi).a;
    R.Dt = mul(Vt, Vi).a;

    // Face masks: nodes whose lattice index for the opposite vertex is 0.
    R.Fmask.resize(4 * Nfp);
    for (int f = 0; f < 4; ++f) {
        int b = 0;
        for (int n = 0; n < Np; ++n)
            if (R.lattice[4 * n + f] == 0) R.Fmask[f * Nfp + b++] = n;
        if (b != Nfp) throw std::runtime_error("reference element: face mask size");
    }

    // Face mass on a triangle of area 2 (the same for all four faces), placed
    // through the lattice coordinates of two face vertices; LIFT = M^-1 E.
    Dense E(Np, 4 * Nfp);
    for (int f = 0; f < 4; ++f) {
        const int va = kFaceVerts[f][0], vb = kFaceVerts[f][1];
        Dense V2(Nfp, Nfp);
        for (int q = 0; q < Nfp; ++q) {
            const int n = R.Fmask[f * Nfp + q];
            // barycentric of vertex v at node n
            auto bary = [&](int v) {
                return v == 0 ? 1.0 - R.xi[3 * n] - R.xi[3 * n + 1] - R.xi[3 * n + 2] : R.xi[3 * n + v - 1];
            };
            const double r = 2 * bary(va) - 1, s = 2 * bary(vb) - 1;
            int m = 0;
            for (int i = 0; i <= p; ++i)
                for (int j = 0; i + j <= p; ++j, ++m) V2(q, m) = pkd2(r, s, i, j);
        }
        const Dense Mf = inv(mul_bt(V2, V2));
        for (int q = 0; q < Nfp; ++q)
            for (int w = 0; w < Nfp; ++w) E(R.Fmask[f * Nfp + q], f * Nfp + w) = Mf(q, w);
    }
    R.LIFT = mul(Minv, E).a;

    // Neighbour-node table from lattice indices (exact integer matching).
    R.face_perm.assign(96 * Nfp, -1);
    for (int f = 0; f < 4; ++f)
        for (int gf = 0; gf < 4; ++gf)
            for (int c = 0; c < 6; ++c) {
                int pi[3];
                perm_from_code(c, pi);
                for (int b = 0; b < Nfp; ++b) {
                    const int n = R.Fmask[f * Nfp + b];
                    int want[3];
                    for (int k = 0; k < 3; ++k) want[pi[k]] = R.lattice[4 * n + kFaceVerts[f][k]];
                    int hit = -1;
                    for (int b2 = 0; b2 < Nfp && hit < 0; ++b2) {
                        const int m = R.Fmask[gf * Nfp + b2];
                        bool same = true;
                        for (int k = 0; k < 3; ++k)
                            same = same && R.lattice[4 * m + kFaceVerts[gf][k]] == want[k];
                        if (same) hit = m;
                    }
                    if (hit < 0) throw std::runtime_error("reference element: face permutation");
                    R.face_perm[((f * 4 + gf) * 6 + c) * Nfp + b] = hit;
                }
            }
    return R;
}

}  // namespace dgtd
