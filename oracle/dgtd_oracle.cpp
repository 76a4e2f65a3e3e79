/*
 * dgtd_oracle.cpp -- CPU FP64 restatement of OpenSEMBA/dgtd's "hesthaven"
 * DGTD Maxwell evolution operator (src/evolution/HesthavenEvolution.cpp) and
 * of the Runge-Kutta drivers that call it (src/solver/Solver.cpp).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation.
 * Nothing in the product (dgtd_amd/) links, loads or calls this file.
 *
 * The reference builds its matrices with MFEM (absent here).  This file
 * restates them from the mathematics MFEM and Hesthaven & Warburton (2008)
 * publish: the orthonormal simplex (PKD) basis, Vandermonde matrices on the
 * MFEM L2 Gauss-Lobatto node set, exact face mass matrices.  Each function
 * cites the reference code it follows.  What pins each part is listed in
 * tests/test_oracle_golden.py and tests/test_oracle_math.py.
 */
#include "dgtd_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

using Mat = std::vector<double>; /* row-major, dims carried alongside */

/* ------------------------------------------------------------------ */
/* Dense linear algebra (small matrices, Np <= 84)                     */
/* ------------------------------------------------------------------ */

Mat matmul(const Mat& A, const Mat& B, int n, int k, int m)
{
    Mat C((size_t)n * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int l = 0; l < k; ++l) {
            const double a = A[(size_t)i * k + l];
            for (int j = 0; j < m; ++j) C[(size_t)i * m + j] += a * B[(size_t)l * m + j];
        }
    return C;
}

Mat transpose(const Mat& A, int n, int m)
{
    Mat T((size_t)n * m);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) T[(size_t)j * n + i] = A[(size_t)i * m + j];
    return T;
}

/* Gauss-Jordan inverse with partial pivoting. */
Mat inverse(const Mat& A, int n)
{
    Mat a = A, inv((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
        if (a[(size_t)piv * n + c] == 0.0) throw std::runtime_error("singular matrix");
        if (piv != c)
            for (int j = 0; j < n; ++j) {
                std::swap(a[(size_t)piv * n + j], a[(size_t)c * n + j]);
                std::swap(inv[(size_t)piv * n + j], inv[(size_t)c * n + j]);
            }
        const double d = a[(size_t)c * n + c];
        for (int j = 0; j < n; ++j) {
            a[(size_t)c * n + j] /= d;
            inv[(size_t)c * n + j] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
                inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j];
            }
        }
    }
    return inv;
}

/* ------------------------------------------------------------------ */
/* 1D: orthonormal Jacobi polynomials and Gauss-Lobatto-Legendre nodes */
/* (Hesthaven & Warburton 2008, JacobiP / GradJacobiP / JacobiGL)      */
/* ------------------------------------------------------------------ */

double jacobiP(double x, double alpha, double beta, int n)
{
    const double g0 = std::pow(2.0, alpha + beta + 1.0) / (alpha + beta + 1.0) *
                      std::tgamma(alpha + 1.0) * std::tgamma(beta + 1.0) /
                      std::tgamma(alpha + beta + 1.0);
    double p0 = 1.0 / std::sqrt(g0);
    if (n == 0) return p0;
    const double g1 = (alpha + 1.0) * (beta + 1.0) / (alpha + beta + 3.0) * g0;
    double p1 = ((alpha + beta + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / std::sqrt(g1);
    if (n == 1) return p1;
    double aold = 2.0 / (2.0 + alpha + beta) *
                  std::sqrt((alpha + 1.0) * (beta + 1.0) / (alpha + beta + 3.0));
    for (int i = 1; i < n; ++i) {
        const double h1 = 2.0 * i + alpha + beta;
        const double anew = 2.0 / (h1 + 2.0) *
                            std::sqrt((i + 1.0) * (i + 1.0 + alpha + beta) * (i + 1.0 + alpha) *
                                      (i + 1.0 + beta) / (h1 + 1.0) / (h1 + 3.0));
        const double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
        const double p2 = 1.0 / anew * (-aold * p0 + (x - bnew) * p1);
        p0 = p1;
        p1 = p2;
        aold = anew;
    }
    return p1;
}

double gradJacobiP(double x, double alpha, double beta, int n)
{
    if (n == 0) return 0.0;
    return std::sqrt(n * (n + alpha + beta + 1.0)) * jacobiP(x, alpha + 1.0, beta + 1.0, n - 1);
}

/* Legendre P_n and P_n' by the three-term recurrence. */
void legendre(int n, double x, double& p, double& dp)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) { p = 1.0; dp = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    p = p1;
    dp = n * (x * p1 - p0) / (x * x - 1.0);
}

/* GLL points on [-1,1]: endpoints and roots of P_p'.  Newton iteration on
 * (1-x^2) P_p'(x) from Chebyshev-Gauss-Lobatto guesses, the standard method. */
std::vector<double> gll(int p)
{
    std::vector<double> x(p + 1);
    if (p == 0) { x[0] = 0.0; return x; }
    x[0] = -1.0;
    x[p] = 1.0;
    for (int i = 1; i < p; ++i) {
        double xi = -std::cos(M_PI * i / p);
        for (int it = 0; it < 100; ++it) {
            /* f = P_p'(x); f' = P_p''(x) = (2x P_p' - p(p+1) P_p) / (1-x^2) */
            double P, dP;
            legendre(p, xi, P, dP);
            const double d2P = (2.0 * xi * dP - p * (p + 1.0) * P) / (1.0 - xi * xi);
            const double dx = dP / d2P;
            xi -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        x[i] = xi;
    }
    /* symmetrise */
    for (int i = 0; i <= p / 2; ++i) {
        const double s = 0.5 * (x[p - i] - x[i]);
        x[i] = -s;
        x[p - i] = s;
    }
    if (p % 2 == 0) x[p / 2] = 0.0;
    return x;
}

/* ------------------------------------------------------------------ */
/* Orthonormal PKD bases on the triangle and tetrahedron               */
/* (Hesthaven & Warburton, Simplex2DP / Simplex3DP / GradSimplex3DP)   */
/* ------------------------------------------------------------------ */

void rst_to_abc(double r, double s, double t, double& a, double& b, double& c)
{
    a = (std::fabs(s + t) > 1e-14) ? 2.0 * (1.0 + r) / (-s - t) - 1.0 : -1.0;
    b = (std::fabs(t - 1.0) > 1e-14) ? 2.0 * (1.0 + s) / (1.0 - t) - 1.0 : -1.0;
    c = t;
}

double simplex3DP(double a, double b, double c, int i, int j, int k)
{
    const double h1 = jacobiP(a, 0.0, 0.0, i);
    const double h2 = jacobiP(b, 2.0 * i + 1.0, 0.0, j);
    const double h3 = jacobiP(c, 2.0 * (i + j) + 2.0, 0.0, k);
    return 2.0 * std::sqrt(2.0) * h1 * h2 * std::pow(1.0 - b, i) * h3 * std::pow(1.0 - c, i + j);
}

void gradSimplex3DP(double a, double b, double c, int id, int jd, int kd, double& dr,
                    double& ds, double& dt)
{
    const double fa = jacobiP(a, 0, 0, id), dfa = gradJacobiP(a, 0, 0, id);
    const double gb = jacobiP(b, 2 * id + 1, 0, jd), dgb = gradJacobiP(b, 2 * id + 1, 0, jd);
    const double hc = jacobiP(c, 2 * (id + jd) + 2, 0, kd),
                 dhc = gradJacobiP(c, 2 * (id + jd) + 2, 0, kd);
    const double hb = 0.5 * (1.0 - b), hcc = 0.5 * (1.0 - c);

    double vr = dfa * (gb * hc);
    if (id > 0) vr *= std::pow(hb, id - 1);
    if (id + jd > 0) vr *= std::pow(hcc, id + jd - 1);

    double vs = 0.5 * (1.0 + a) * vr;
    double tmp = dgb * std::pow(hb, id);
    if (id > 0) tmp += (-0.5 * id) * (gb * std::pow(hb, id - 1));
    if (id + jd > 0) tmp *= std::pow(hcc, id + jd - 1);
    tmp = fa * (tmp * hc);
    vs += tmp;

    double vt = 0.5 * (1.0 + a) * vr + 0.5 * (1.0 + b) * tmp;
    tmp = dhc * std::pow(hcc, id + jd);
    if (id + jd > 0) tmp -= 0.5 * (id + jd) * (hc * std::pow(hcc, id + jd - 1));
    tmp = fa * (gb * tmp);
    tmp *= std::pow(hb, id);
    vt += tmp;

    const double sc = std::pow(2.0, 2 * id + jd + 1.5);
    dr = vr * sc;
    ds = vs * sc;
    dt = vt * sc;
}

double simplex2DP(double r, double s, int i, int j)
{
    const double a = (std::fabs(s - 1.0) > 1e-14) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
    const double b = s;
    const double h1 = jacobiP(a, 0, 0, i);
    const double h2 = jacobiP(b, 2 * i + 1, 0, j);
    return std::sqrt(2.0) * h1 * h2 * std::pow(1.0 - b, i);
}

/* ------------------------------------------------------------------ */
/* Reference element                                                   */
/* ------------------------------------------------------------------ */

struct RefElem {
    int p = 0, Np = 0, Nfp = 0;
    std::vector<double> xi, eta, zeta;     /* MFEM unit-tet coordinates */
    Mat M, Minv, Dr, Ds, Dt, Emat, LIFT;   /* Np x Np, Np x 4Nfp */
    std::vector<int> Fmask;                /* 4*Nfp */
};

/* MFEM L2_TetrahedronElement node set for BasisType::GaussLobatto: for
 * k, j, i ascending (i fastest), node = (x_i, x_j, x_k)/(x_i+x_j+x_k+x_l),
 * x = GLL points on [0,1], l = p-i-j-k.  Used by the reference through
 * L2_FECollection(order, 3, GaussLobatto) (HesthavenEvolution.cpp:49,
 * Solver.cpp:86). */
void build_nodes(int p, std::vector<double>& xi, std::vector<double>& eta,
                 std::vector<double>& zeta)
{
    std::vector<double> g = gll(p);
    for (auto& v : g) v = 0.5 * (v + 1.0);
    xi.clear(); eta.clear(); zeta.clear();
    if (p == 0) { xi = {0.25}; eta = {0.25}; zeta = {0.25}; return; }
    for (int k = 0; k <= p; ++k)
        for (int j = 0; j + k <= p; ++j)
            for (int i = 0; i + j + k <= p; ++i) {
                const int l = p - i - j - k;
                const double w = g[i] + g[j] + g[k] + g[l];
                xi.push_back(g[i] / w);
                eta.push_back(g[j] / w);
                zeta.push_back(g[k] / w);
            }
}

/* Reference matrices on the Hesthaven tetrahedron (0,0,0),(2,0,0),(0,2,0),
 * (0,0,2) (HesthavenEvolutionMethods.cpp:257-268), i.e. r = 2*xi-1 etc.
 *  - M: reference mass (HesthavenEvolution.cpp:13-25,46-54), = (V V^T)^-1
 *  - Dr,Ds,Dt: reference strong derivatives, = V_r V^-1
 *  - Emat: face mass matrices on a reference face of area 2 with face f
 *    opposite vertex f, columns in ascending local node index
 *    (HesthavenEvolution.cpp:56-81, HesthavenEvolutionMethods.cpp:223-242,
 *    289-300; the integrator BilinearIntegrators.cpp:671-714 integrates with
 *    reference-face weights times 2, and the caller doubles again for tets)
 *  - LIFT = M^-1 Emat (HesthavenEvolution.cpp:365). */
RefElem build_ref(int p)
{
    RefElem R;
    R.p = p;
    R.Np = (p + 1) * (p + 2) * (p + 3) / 6;
    R.Nfp = (p + 1) * (p + 2) / 2;
    const int Np = R.Np, Nfp = R.Nfp;
    build_nodes(p, R.xi, R.eta, R.zeta);

    Mat V((size_t)Np * Np), Vr((size_t)Np * Np), Vs((size_t)Np * Np), Vt((size_t)Np * Np);
    for (int n = 0; n < Np; ++n) {
        const double r = 2 * R.xi[n] - 1, s = 2 * R.eta[n] - 1, t = 2 * R.zeta[n] - 1;
        double a, b, c;
        rst_to_abc(r, s, t, a, b, c);
        int m = 0;
        for (int i = 0; i <= p; ++i)
            for (int j = 0; i + j <= p; ++j)
                for (int k = 0; i + j + k <= p; ++k) {
                    V[(size_t)n * Np + m] = simplex3DP(a, b, c, i, j, k);
                    double dr, ds, dt;
                    gradSimplex3DP(a, b, c, i, j, k, dr, ds, dt);
                    Vr[(size_t)n * Np + m] = dr;
                    Vs[(size_t)n * Np + m] = ds;
                    Vt[(size_t)n * Np + m] = dt;
                    ++m;
                }
    }
    const Mat Vinv = inverse(V, Np);
    R.Minv = matmul(V, transpose(V, Np, Np), Np, Np, Np);
    R.M = inverse(R.Minv, Np);
    R.Dr = matmul(Vr, Vinv, Np, Np, Np);
    R.Ds = matmul(Vs, Vinv, Np, Np, Np);
    R.Dt = matmul(Vt, Vinv, Np, Np, Np);

    /* Face f is opposite vertex f: barycentric lambda_f = 0 with
     * lambda_0 = 1-xi-eta-zeta, lambda_1 = xi, lambda_2 = eta, lambda_3 = zeta
     * (MFEM tetrahedron FaceVert {1,2,3},{0,3,2},{0,1,3},{0,2,1}). */
    R.Fmask.assign(4 * Nfp, -1);
    const double tol = 1e-10;
    for (int f = 0; f < 4; ++f) {
        int b = 0;
        for (int n = 0; n < Np; ++n) {
            const double lam[4] = {1.0 - R.xi[n] - R.eta[n] - R.zeta[n], R.xi[n], R.eta[n],
                                   R.zeta[n]};
            if (std::fabs(lam[f]) < tol) {
                if (b >= Nfp) throw std::runtime_error("face mask overflow");
                R.Fmask[f * Nfp + b++] = n;
            }
        }
        if (b != Nfp) throw std::runtime_error("face mask underflow");
    }

    /* Face mass on a triangle of area 2: (V2 V2^T)^-1 with V2 the orthonormal
     * 2D basis on (-1,-1),(1,-1),(-1,1) at the face nodes.  Face nodes are
     * placed on that triangle by an affine map through two barycentrics
     * of the face (the mass of a Lagrange basis depends only on the node
     * positions up to the area factor, which is fixed to 2 for all faces,
     * Hesthaven's convention). */
    R.Emat.assign((size_t)Np * 4 * Nfp, 0.0);
    for (int f = 0; f < 4; ++f) {
        /* two face barycentrics that vanish at no common face vertex */
        int ba, bb;
        switch (f) {
            case 0: ba = 1; bb = 2; break;
            case 1: ba = 2; bb = 3; break;
            case 2: ba = 1; bb = 3; break;
            default: ba = 1; bb = 2; break;
        }
        Mat V2((size_t)Nfp * Nfp);
        for (int q = 0; q < Nfp; ++q) {
            const int n = R.Fmask[f * Nfp + q];
            const double lam[4] = {1.0 - R.xi[n] - R.eta[n] - R.zeta[n], R.xi[n], R.eta[n],
                                   R.zeta[n]};
            const double r = 2.0 * lam[ba] - 1.0, s = 2.0 * lam[bb] - 1.0;
            int m = 0;
            for (int i = 0; i <= p; ++i)
                for (int j = 0; i + j <= p; ++j) V2[(size_t)q * Nfp + m++] = simplex2DP(r, s, i, j);
        }
        const Mat Mf = inverse(matmul(V2, transpose(V2, Nfp, Nfp), Nfp, Nfp, Nfp), Nfp);
        for (int q = 0; q < Nfp; ++q)
            for (int w = 0; w < Nfp; ++w)
                R.Emat[(size_t)R.Fmask[f * Nfp + q] * 4 * Nfp + f * Nfp + w] = Mf[(size_t)q * Nfp + w];
    }
    R.LIFT = matmul(R.Minv, R.Emat, Np, Np, 4 * Nfp);
    return R;
}

/* LSERK4 coefficients: Carpenter & Kennedy (1994), five-stage fourth-order
 * 2N-storage scheme, as tabulated in Hesthaven & Warburton (2008). */
const double rk4a[5] = {0.0, -567301805773.0 / 1357537059087.0,
                        -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0,
                        -1275806237668.0 / 842570457699.0};
const double rk4b[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double rk4c[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                        2526269341429.0 / 6820363962896.0, 2006345519317.0 / 3224310063776.0,
                        2802321613138.0 / 2924317926251.0};

} // namespace

/* ------------------------------------------------------------------ */
/* Operator                                                            */
/* ------------------------------------------------------------------ */

struct oracle_op {
    int K = 0, p = 0, Np = 0, Nfp = 0;
    double alpha = 1.0;
    bool global_sma = false; /* SMA faces at upwind weight 1 (Global operator) */
    RefElem R;
    std::vector<double> metric;   /* 9K: rx ry rz sx sy sz tx ty tz */
    std::vector<double> vol;      /* K */
    std::vector<double> nrm;      /* 12K */
    std::vector<double> fscale;   /* 4K */
    std::vector<double> inv_eps, inv_mu, eps, mu; /* K */
    std::vector<int> vmapM, vmapP;                /* 4*Nfp*K */
    std::vector<int> bc;                          /* 4K: BC code of each face (0 if interior) */
    std::vector<double> pos;                      /* 3*Np*K */
    std::vector<int> tfsf;                        /* 4K: -1 scattered-field side, +1 total-field side, 0 */
    std::vector<oracle_planewave> waves;          /* plane waves injected on the TF/SF faces */
};

/* Planewave::eval (Function.h:361-406) with a Gaussian or modulated Gaussian
 * magnitude: u = k.x / c - t, g = exp(-(u - mean)^2 / (2 spread^2))
 * [* cos(2 pi freq (u - mean))]; E = g pol, H = g (k x pol), c = 1. */
static void planewave_eval(const oracle_planewave& w, const double* x, double t, double E[3], double H[3])
{
    const double u = w.dir[0] * x[0] + w.dir[1] * x[1] + w.dir[2] * x[2] - t;
    const double a = u - w.mean;
    double g = std::exp(-a * a / (2.0 * w.spread * w.spread));
    if (w.kind == 1) g *= std::cos(2.0 * M_PI * w.freq * a);
    const double kxp[3] = {w.dir[1] * w.pol[2] - w.dir[2] * w.pol[1], w.dir[2] * w.pol[0] - w.dir[0] * w.pol[2],
                           w.dir[0] * w.pol[1] - w.dir[1] * w.pol[0]};
    for (int d = 0; d < 3; ++d) {
        E[d] = g * w.pol[d];
        H[d] = g * kxp[d];
    }
}

extern "C" {

int oracle_np(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }
int oracle_nfp(int order) { return (order + 1) * (order + 2) / 2; }
const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void oracle_gll01(int p, double* x)
{
    std::vector<double> g = gll(p);
    for (int i = 0; i <= p; ++i) x[i] = 0.5 * (g[i] + 1.0);
}

void oracle_ref_element(int order, double* rst, double* M, double* Dr, double* Ds, double* Dt,
                        double* emat, double* lift, int* fmask)
{
    const RefElem R = build_ref(order);
    const int Np = R.Np, Nfp = R.Nfp;
    if (rst)
        for (int n = 0; n < Np; ++n) {
            rst[3 * n + 0] = 2 * R.xi[n] - 1;
            rst[3 * n + 1] = 2 * R.eta[n] - 1;
            rst[3 * n + 2] = 2 * R.zeta[n] - 1;
        }
    if (M) std::copy(R.M.begin(), R.M.end(), M);
    if (Dr) std::copy(R.Dr.begin(), R.Dr.end(), Dr);
    if (Ds) std::copy(R.Ds.begin(), R.Ds.end(), Ds);
    if (Dt) std::copy(R.Dt.begin(), R.Dt.end(), Dt);
    if (emat) std::copy(R.Emat.begin(), R.Emat.end(), emat);
    if (lift) std::copy(R.LIFT.begin(), R.LIFT.end(), lift);
    if (fmask) std::copy(R.Fmask.begin(), R.Fmask.end(), fmask);
    (void)Nfp;
}

namespace {
oracle_op* create_op(int K, int order, const double* vx, int nv, const int* EToV,
                     const int* face_bc, const double* eps, const double* mu, double alpha,
                     bool global_sma)
{
    try {
        if (K <= 0 || order < 1 || order > 8 || !vx || !EToV || !face_bc)
            throw std::runtime_error("bad arguments");
        auto* op = new oracle_op;
        op->K = K;
        op->p = order;
        op->alpha = alpha;
        op->global_sma = global_sma;
        op->R = build_ref(order);
        const int Np = op->R.Np, Nfp = op->R.Nfp;
        op->Np = Np;
        op->Nfp = Nfp;
        op->metric.resize(9 * (size_t)K);
        op->vol.resize(K);
        op->nrm.resize(12 * (size_t)K);
        op->fscale.resize(4 * (size_t)K);
        op->eps.resize(K);
        op->mu.resize(K);
        op->inv_eps.resize(K);
        op->inv_mu.resize(K);
        op->pos.resize(3 * (size_t)Np * K);
        const size_t N = (size_t)Np * K;

        for (int e = 0; e < K; ++e) {
            double v[4][3];
            for (int a = 0; a < 4; ++a) {
                const int id = EToV[4 * e + a];
                if (id < 0 || id >= nv) throw std::runtime_error("EToV out of range");
                for (int d = 0; d < 3; ++d) v[a][d] = vx[3 * id + d];
            }
            /* x = v0 + (v1-v0) xi + (v2-v0) eta + (v3-v0) zeta, r = 2 xi - 1 */
            double J[3][3]; /* J[d][k] = dx_d / dr_k */
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 3; ++k) J[d][k] = 0.5 * (v[k + 1][d] - v[0][d]);
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            if (std::fabs(det) < 1e-300) throw std::runtime_error("degenerate element");
            double Ji[3][3]; /* Ji[k][d] = dr_k / dx_d */
            Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) op->metric[9 * (size_t)e + 3 * k + d] = Ji[k][d];
            /* volume = |det(dx/dr)| * Vref with Vref = 4/3 */
            op->vol[e] = std::fabs(det) * 4.0 / 3.0;

            /* Faces (HesthavenEvolution.cpp:172-205): unit outward normal and
             * fscale = 2 sJ / J = 2 A / (3 V).  With lambda_f the barycentric
             * of vertex f, grad(lambda_f) is normal to face f pointing
             * inward, |grad lambda_f| = A_f / (3 V). */
            double gl[4][3];
            for (int d = 0; d < 3; ++d) {
                /* dxi/dx = 0.5 * dr/dx ... lambda_1 = xi = (r+1)/2 */
                gl[1][d] = 0.5 * Ji[0][d];
                gl[2][d] = 0.5 * Ji[1][d];
                gl[3][d] = 0.5 * Ji[2][d];
                gl[0][d] = -(gl[1][d] + gl[2][d] + gl[3][d]);
            }
            for (int f = 0; f < 4; ++f) {
                const double g = std::sqrt(gl[f][0] * gl[f][0] + gl[f][1] * gl[f][1] + gl[f][2] * gl[f][2]);
                for (int d = 0; d < 3; ++d) op->nrm[12 * (size_t)e + 3 * f + d] = -gl[f][d] / g;
                op->fscale[4 * (size_t)e + f] = 2.0 * g;
            }
            const double ee = eps ? eps[e] : 1.0, mm = mu ? mu[e] : 1.0;
            if (!(ee > 0) || !(mm > 0)) throw std::runtime_error("eps and mu must be positive");
            op->eps[e] = ee;
            op->mu[e] = mm;
            op->inv_eps[e] = 1.0 / ee;
            op->inv_mu[e] = 1.0 / mm;
            for (int n = 0; n < Np; ++n) {
                const double a = op->R.xi[n], b = op->R.eta[n], c = op->R.zeta[n];
                for (int d = 0; d < 3; ++d)
                    op->pos[d * N + (size_t)e * Np + n] =
                        v[0][d] + (v[1][d] - v[0][d]) * a + (v[2][d] - v[0][d]) * b + (v[3][d] - v[0][d]) * c;
            }
        }

        /* Connectivity (HesthavenEvolutionMethods.cpp:721-791): for each element,
         * each local face in order, each face node in ascending local index:
         * (M, P) with P the co-located node of the neighbour across the face,
         * or P = M on a boundary face.  The reference pairs nodes by equal
         * face-mass entries (:59-75), which is node co-location for conforming
         * meshes; here the pairing is by node position. */
        static const int fv[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
        std::map<std::array<int, 3>, std::vector<std::pair<int, int>>> faces;
        for (int e = 0; e < K; ++e)
            for (int f = 0; f < 4; ++f) {
                std::array<int, 3> key = {EToV[4 * e + fv[f][0]], EToV[4 * e + fv[f][1]], EToV[4 * e + fv[f][2]]};
                std::sort(key.begin(), key.end());
                faces[key].push_back({e, f});
            }
        op->vmapM.resize(4 * (size_t)Nfp * K);
        op->vmapP.resize(4 * (size_t)Nfp * K);
        op->bc.assign(4 * (size_t)K, 0);
        op->tfsf.assign(4 * (size_t)K, 0);
        double hmin = 1e300;
        for (int e = 0; e < K; ++e) hmin = std::min(hmin, std::cbrt(op->vol[e]));
        const double ptol = 1e-8 * hmin;
        for (int e = 0; e < K; ++e)
            for (int f = 0; f < 4; ++f) {
                std::array<int, 3> key = {EToV[4 * e + fv[f][0]], EToV[4 * e + fv[f][1]], EToV[4 * e + fv[f][2]]};
                std::sort(key.begin(), key.end());
                const auto& lst = faces[key];
                if (lst.size() > 2) throw std::runtime_error("non-manifold face");
                int e2 = -1, f2 = -1;
                for (auto& pr : lst)
                    if (pr.first != e || pr.second != f) { e2 = pr.first; f2 = pr.second; }
                /* codes 1-3 on boundary faces; 4-6 (interior PEC/PMC/SMA) on
                 * interior faces, given on either side (the condition belongs
                 * to the face, :294-311 applies it to both elements) */
                int code = face_bc[4 * e + f];
                if (code < 0 || code > 7) throw std::runtime_error("bad face_bc code");
                if (e2 >= 0) {
                    const int other = face_bc[4 * e2 + f2];
                    if (other < 0 || other > 7) throw std::runtime_error("bad face_bc code");
                    if (code != 0 && other != 0 && code != other)
                        throw std::runtime_error("the two sides of an interior face carry different codes");
                    if (code == 0) code = other;
                    if (code >= 1 && code <= 3)
                        throw std::runtime_error("PEC/PMC/SMA codes 1-3 on an interior face (use 4-6)");
                } else if (code >= 4) {
                    throw std::runtime_error("interior-boundary codes 4-7 on a boundary face");
                }
                if (code == 7) {
                    /* TF/SF face: the scattered-field side is the face's first
                     * element (the lower index), as MFEM orders a face's two
                     * elements (HesthavenEvolutionMethods.cpp:562-574) */
                    op->tfsf[4 * (size_t)e + f] = e < e2 ? -1 : 1;
                    code = 0;
                }
                if (code == 3 && alpha == 0.0 && !global_sma)
                    throw std::runtime_error("SMA boundaries need upwind flux (alpha != 0)");
                op->bc[4 * (size_t)e + f] = code;
                for (int b = 0; b < Nfp; ++b) {
                    const size_t n = (size_t)e * 4 * Nfp + (size_t)f * Nfp + b;
                    const int iM = op->R.Fmask[f * Nfp + b];
                    const size_t gM = (size_t)e * Np + iM;
                    op->vmapM[n] = (int)gM;
                    if (e2 < 0) { op->vmapP[n] = (int)gM; continue; }
                    int found = -1;
                    for (int b2 = 0; b2 < Nfp; ++b2) {
                        const size_t gP = (size_t)e2 * Np + op->R.Fmask[f2 * Nfp + b2];
                        double d2 = 0;
                        for (int d = 0; d < 3; ++d) {
                            const double dd = op->pos[d * N + gM] - op->pos[d * N + gP];
                            d2 += dd * dd;
                        }
                        if (std::sqrt(d2) < ptol) { found = (int)gP; break; }
                    }
                    if (found < 0) throw std::runtime_error("non-conforming face: node not matched");
                    op->vmapP[n] = found;
                }
            }
        return op;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}
}  // namespace

oracle_op* oracle_create(int K, int order, const double* vx, int nv, const int* EToV,
                         const int* face_bc, const double* eps, const double* mu, double alpha)
{
    return create_op(K, order, vx, nv, EToV, face_bc, eps, mu, alpha, false);
}

oracle_op* oracle_create_global(int K, int order, const double* vx, int nv, const int* EToV,
                                const int* face_bc, const double* eps, const double* mu, double alpha)
{
    return create_op(K, order, vx, nv, EToV, face_bc, eps, mu, alpha, true);
}

void oracle_destroy(oracle_op* op) { delete op; }

int64_t oracle_ndofs(const oracle_op* op) { return 6LL * op->Np * op->K; }

double oracle_max_fscale(const oracle_op* op)
{
    /* Solver.cpp:334-346 reads fscale.maxCoeff() over all elements. */
    double m = 0;
    for (double v : op->fscale) m = std::max(m, v);
    return m;
}

void oracle_node_positions(const oracle_op* op, double* xyz)
{
    std::copy(op->pos.begin(), op->pos.end(), xyz);
}

void oracle_connectivity(const oracle_op* op, int* vmapM, int* vmapP)
{
    std::copy(op->vmapM.begin(), op->vmapM.end(), vmapM);
    std::copy(op->vmapP.begin(), op->vmapP.end(), vmapP);
}

int oracle_bc_faces(const oracle_op* op, int kind, int* out_nodes)
{
    int nf = 0;
    for (int e = 0; e < op->K; ++e)
        for (int f = 0; f < 4; ++f)
            if (op->bc[4 * (size_t)e + f] == kind) {
                if (out_nodes)
                    for (int b = 0; b < op->Nfp; ++b)
                        out_nodes[(size_t)nf * op->Nfp + b] =
                            op->vmapM[(size_t)e * 4 * op->Nfp + (size_t)f * op->Nfp + b];
                ++nf;
            }
    return nf;
}

void oracle_geometry(const oracle_op* op, double* vol, double* fscale, double* normals)
{
    if (vol) std::copy(op->vol.begin(), op->vol.end(), vol);
    if (fscale) std::copy(op->fscale.begin(), op->fscale.end(), fscale);
    if (normals) std::copy(op->nrm.begin(), op->nrm.end(), normals);
}

int oracle_set_planewaves(oracle_op* op, const oracle_planewave* w, int n)
{
    if (!op || n < 0 || (n > 0 && !w)) return -1;
    op->waves.assign(w, w + n);
    for (auto& v : op->waves) {  /* Planewave normalises both directions (Function.h:331-340) */
        const double np = std::sqrt(v.pol[0] * v.pol[0] + v.pol[1] * v.pol[1] + v.pol[2] * v.pol[2]);
        const double nd = std::sqrt(v.dir[0] * v.dir[0] + v.dir[1] * v.dir[1] + v.dir[2] * v.dir[2]);
        if (!(np > 0) || !(nd > 0) || !(v.spread > 0)) return -1;
        for (int d = 0; d < 3; ++d) {
            v.pol[d] /= np;
            v.dir[d] /= nd;
        }
    }
    return 0;
}

void oracle_mult(const oracle_op* op, const double* in, double* out) { oracle_mult_t(op, in, out, 0.0); }

/* HesthavenEvolution::Mult (HesthavenEvolution.cpp:450-542) at time t. */
void oracle_mult_t(const oracle_op* op, const double* in, double* out, double t)
{
    const int K = op->K, Np = op->Np, Nfp = op->Nfp, Nf = 4 * Nfp;
    const size_t N = (size_t)Np * K;
    const double alpha = op->alpha;
    const RefElem& R = op->R;

#pragma omp parallel
    {
        std::vector<double> dF[6], flux[6], ur(Np), us(Np), ut(Np), dq[6][3];
        for (int c = 0; c < 6; ++c) {
            dF[c].resize(Nf);
            flux[c].resize(Nf);
            for (int k = 0; k < 3; ++k) dq[c][k].resize(Np);
        }
#pragma omp for schedule(static)
        for (int e = 0; e < K; ++e) {
            /* jumps dQ = Q[P] - Q[M] (:474-479), boundary overrides (:275-313) */
            for (int f = 0; f < 4; ++f) {
                const int code = op->bc[4 * (size_t)e + f];
                const double af = op->global_sma && (code == 3 || code == 6) ? 1.0 : alpha;
                for (int b = 0; b < Nfp; ++b) {
                    const size_t n = (size_t)e * Nf + (size_t)f * Nfp + b;
                    const size_t M = op->vmapM[n], P = op->vmapP[n];
                    for (int c = 0; c < 6; ++c) dF[c][f * Nfp + b] = in[c * N + P] - in[c * N + M];
                    if (code == 1) { /* PEC: dE = -2 E, dH = 0 */
                        for (int c = 0; c < 3; ++c) dF[c][f * Nfp + b] = -2.0 * in[c * N + M];
                        for (int c = 3; c < 6; ++c) dF[c][f * Nfp + b] = 0.0 * in[c * N + M];
                    } else if (code == 2) { /* PMC: dE = 0, dH = -2 H */
                        for (int c = 0; c < 3; ++c) dF[c][f * Nfp + b] = 0.0 * in[c * N + M];
                        for (int c = 3; c < 6; ++c) dF[c][f * Nfp + b] = -2.0 * in[c * N + M];
                    } else if (code == 3) { /* SMA: dE = -E/alpha, dH = -H/alpha */
                        for (int c = 0; c < 6; ++c) dF[c][f * Nfp + b] = (-1.0 / af) * in[c * N + M];
                    } else if (code >= 4) { /* interior PEC (-1, 0), PMC (0, -1), SMA (-1/2, -1/2) (:309-311) */
                        const double ce = code == 4 ? -1.0 : (code == 6 ? -0.5 : 0.0);
                        const double ch = code == 5 ? -1.0 : (code == 6 ? -0.5 : 0.0);
                        for (int c = 0; c < 3; ++c) dF[c][f * Nfp + b] = ce * in[c * N + M];
                        for (int c = 3; c < 6; ++c) dF[c][f * Nfp + b] = ch * in[c * N + M];
                    }
                    /* TF/SF (evaluateTFSF, :97-124): the incident field at the
                     * face node, subtracted on the scattered-field side and
                     * added on the total-field side */
                    const int sg = op->tfsf[4 * (size_t)e + f];
                    if (sg != 0) {
                        double xM[3];
                        for (int d = 0; d < 3; ++d) xM[d] = op->pos[d * N + M];
                        for (const auto& w : op->waves) {
                            double Ei[3], Hi[3];
                            planewave_eval(w, xM, t, Ei, Hi);
                            for (int d = 0; d < 3; ++d) {
                                dF[d][f * Nfp + b] += sg * Ei[d];
                                dF[3 + d][f * Nfp + b] += sg * Hi[d];
                            }
                        }
                    }
                }
            }
            /* fluxes (:499-520), scaled by fscale/2 (applyLIFT :126-132) */
            for (int f = 0; f < 4; ++f) {
                const double* nv = &op->nrm[12 * (size_t)e + 3 * f];
                const double fs = op->fscale[4 * (size_t)e + f] / 2.0;
                const int code = op->bc[4 * (size_t)e + f];
                const double af = op->global_sma && (code == 3 || code == 6) ? 1.0 : alpha;
                for (int b = 0; b < Nfp; ++b) {
                    const int m = f * Nfp + b;
                    double ndotdH = 0.0, ndotdE = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        ndotdH += nv[d] * dF[3 + d][m];
                        ndotdE += nv[d] * dF[d][m];
                    }
                    for (int x = 0; x < 3; ++x) {
                        const int y = (x + 1) % 3, z = (x + 2) % 3;
                        const double fH = -1.0 * nv[y] * dF[z][m] + nv[z] * dF[y][m] +
                                          af * (dF[3 + x][m] - ndotdH * nv[x]);
                        const double fE = nv[y] * dF[3 + z][m] - 1.0 * nv[z] * dF[3 + y][m] +
                                          af * (dF[x][m] - ndotdE * nv[x]);
                        flux[3 + x][m] = fH * fs;
                        flux[x][m] = fE * fs;
                    }
                }
            }
            /* physical derivatives D_d u = r_d Dr u + s_d Ds u + t_d Dt u; the
             * reference stores D_d = M^-1 S_d Vref / V_e per element (:150-170),
             * the same matrix for affine elements */
            const double* G = &op->metric[9 * (size_t)e];
            for (int c = 0; c < 6; ++c) {
                const double* u = in + c * N + (size_t)e * Np;
                for (int i = 0; i < Np; ++i) {
                    double a = 0, b = 0, cc = 0;
                    for (int j = 0; j < Np; ++j) {
                        a += R.Dr[(size_t)i * Np + j] * u[j];
                        b += R.Ds[(size_t)i * Np + j] * u[j];
                        cc += R.Dt[(size_t)i * Np + j] * u[j];
                    }
                    ur[i] = a; us[i] = b; ut[i] = cc;
                }
                for (int d = 0; d < 3; ++d)
                    for (int i = 0; i < Np; ++i)
                        dq[c][d][i] = G[0 + d] * ur[i] + G[3 + d] * us[i] + G[6 + d] * ut[i];
            }
            /* rhsH_x = -D_y E_z + D_z E_y + LIFT g(fluxH_x)
             * rhsE_x =  D_y H_z - D_z H_y + LIFT g(fluxE_x)  (:522-534) */
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                for (int i = 0; i < Np; ++i) {
                    double lH = 0, lE = 0;
                    const double* L = &R.LIFT[(size_t)i * Nf];
                    for (int m = 0; m < Nf; ++m) {
                        lH += L[m] * flux[3 + x][m];
                        lE += L[m] * flux[x][m];
                    }
                    const double hres = -1.0 * dq[z][y][i] + dq[y][z][i] + lH;
                    const double eres = dq[3 + z][y][i] - 1.0 * dq[3 + y][z][i] + lE;
                    /* heterogeneous media: per-element 1/eps, 1/mu (the Global
                     * operator's M_eps^-1, M_mu^-1 semantics; DGOperatorFactory.h:411-424) */
                    out[(3 + x) * N + (size_t)e * Np + i] = hres * op->inv_mu[e];
                    out[x * N + (size_t)e * Np + i] = eres * op->inv_eps[e];
                }
            }
        }
    }
}

/* mfem::RK4Solver::Step as called from Solver::step (Solver.cpp:535-551):
 *   k1 = f(x); y = x + dt/2 k1; z = x + dt/6 k1
 *   k2 = f(y); y = x + dt/2 k2; z += dt/3 k2
 *   k3 = f(y); y = x + dt k3;   z += dt/3 k3
 *   k4 = f(y); x = z + dt/6 k4 */
void oracle_rk4_step(const oracle_op* op, double* x, double t, double dt)
{
    /* stage times t, t + dt/2, t + dt/2, t + dt (the SetTime calls of RK4Solver) */
    const size_t n = (size_t)oracle_ndofs(op);
    std::vector<double> k(n), y(n), z(n);
    oracle_mult_t(op, x, k.data(), t);
    for (size_t i = 0; i < n; ++i) { y[i] = x[i] + dt / 2 * k[i]; z[i] = x[i] + dt / 6 * k[i]; }
    oracle_mult_t(op, y.data(), k.data(), t + dt / 2);
    for (size_t i = 0; i < n; ++i) { y[i] = x[i] + dt / 2 * k[i]; z[i] += dt / 3 * k[i]; }
    oracle_mult_t(op, y.data(), k.data(), t + dt / 2);
    for (size_t i = 0; i < n; ++i) { y[i] = x[i] + dt * k[i]; z[i] += dt / 3 * k[i]; }
    oracle_mult_t(op, y.data(), k.data(), t + dt);
    for (size_t i = 0; i < n; ++i) x[i] = z[i] + dt / 6 * k[i];
}

/* LSERK4: res = a_s res + dt L(q); q = q + b_s res, s = 1..5. */
void oracle_lserk4_step(const oracle_op* op, double* q, double* res, double t, double dt)
{
    const size_t n = (size_t)oracle_ndofs(op);
    std::vector<double> k(n);
    for (int s = 0; s < 5; ++s) {
        oracle_mult_t(op, q, k.data(), t + rk4c[s] * dt);
        for (size_t i = 0; i < n; ++i) {
            res[i] = rk4a[s] * res[i] + dt * k[i];
            q[i] = q[i] + rk4b[s] * res[i];
        }
    }
}

/* Solver::run loop (Solver.cpp:497): while t <= T - 1e-8 dt, step with
 * truedt = min(dt, T - t) (:537). */
int oracle_run(const oracle_op* op, double* x, double dt, double final_time, int integrator,
               double* t_out)
{
    double t = 0.0;
    int steps = 0;
    std::vector<double> res;
    if (integrator == 1) res.assign((size_t)oracle_ndofs(op), 0.0);
    while (t <= final_time - 1e-8 * dt) {
        const double truedt = std::min(dt, final_time - t);
        if (integrator == 0) oracle_rk4_step(op, x, t, truedt);
        else oracle_lserk4_step(op, x, res.data(), t, truedt);
        t += truedt;
        ++steps;
    }
    if (t_out) *t_out = t;
    return steps;
}

double oracle_energy(const oracle_op* op, const double* x)
{
    const int K = op->K, Np = op->Np;
    const size_t N = (size_t)Np * K;
    double total = 0.0;
    for (int e = 0; e < K; ++e) {
        const double J = op->vol[e] / (4.0 / 3.0);
        double eE = 0, eH = 0;
        for (int c = 0; c < 6; ++c) {
            const double* u = x + c * N + (size_t)e * Np;
            double s = 0;
            for (int i = 0; i < Np; ++i) {
                double r = 0;
                for (int j = 0; j < Np; ++j) r += op->R.M[(size_t)i * Np + j] * u[j];
                s += u[i] * r;
            }
            if (c < 3) eE += s; else eH += s;
        }
        total += 0.5 * J * (op->eps[e] * eE + op->mu[e] * eH);
    }
    return total;
}

} // extern "C"
