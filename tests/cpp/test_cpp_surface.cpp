// The C++ TimeDependentOperator surface (include/dgtd_hip.hpp) against the
// CPU oracle: Mult and an RK4 Solver::run on a 2^3 PEC cavity at p=3, FP64;
// the assembled (CSR) variant against the oracle's Mult and, with a uniform
// conductivity sigma (eps = 1), against A x - sigma x_E (the conductive term
// is -sigma/eps on the E DOFs); on a triangle mesh the per-node, matrix-core
// and assembled variants against each other.  Exit status 0 and "OK".
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "dgtd_hip.hpp"
#include "../../oracle/dgtd_oracle.h"

static double rel(const std::vector<double>& a, const std::vector<double>& b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / den);
}

int main()
{
    const int8_t pec[6] = {1, 1, 1, 1, 1, 1};
    dgtd::Mesh mesh = dgtd::Mesh::cube(2, 1.0, pec);
    dgtd::Options opts;
    opts.order = 3;
    opts.precision = 64;
    opts.integrator = DGTD_RK4;
    dgtd::HesthavenEvolutionHip op(mesh, opts);
    const int K = (int)mesh.EToV.size() / 4;
    std::vector<int> fbc(mesh.face_bc.begin(), mesh.face_bc.end());
    oracle_op* ref = oracle_create(K, 3, mesh.vx.data(), (int)mesh.vx.size() / 3, mesh.EToV.data(), fbc.data(),
                                   nullptr, nullptr, 1.0);
    if (!ref) { std::printf("oracle: %s\n", oracle_last_error()); return 2; }
    const int64_t n = op.Height();
    if (n != oracle_ndofs(ref)) { std::printf("height mismatch\n"); return 1; }
    std::mt19937_64 rng(7);
    std::normal_distribution<double> g;
    dgtd::Vector x(n), y, yr(n);
    for (auto& v : x) v = g(rng);
    op.SetTime(0.25);
    op.Mult(x, y);
    oracle_mult(ref, x.data(), yr.data());
    const double e1 = rel(y, yr);
    if (op.MaxFscale() != oracle_max_fscale(ref) && std::fabs(op.MaxFscale() - oracle_max_fscale(ref)) > 1e-12) {
        std::printf("max fscale mismatch\n");
        return 1;
    }
    dgtd::Solver solver(op, x, 0.05);
    const int steps = solver.run();
    dgtd::Vector xs = solver.getFields();
    dgtd::Vector xr = x;
    double tr = 0;
    const int steps_r = oracle_run(ref, xr.data(), solver.getTimeStep(), 0.05, 0, &tr);
    const double e2 = rel(xs, xr);
    std::printf("mult rel %.3e, run steps %d/%d rel %.3e, t %.17g/%.17g\n", e1, steps, steps_r, e2,
                solver.getTime(), tr);
    bool ok = e1 < 1e-12 && e2 < 1e-12 && steps == steps_r && std::fabs(solver.getTime() - tr) < 1e-14;
    try {  // errors cross the boundary as dgtd::Error
        dgtd::Options bad = opts;
        bad.order = 9;
        dgtd::HesthavenEvolutionHip b2(mesh, bad);
        ok = false;
    } catch (const dgtd::Error& e) {
        ok = ok && e.code == DGTD_E_UNSUPPORTED;
    }
    {  // assembled operator, same mesh
        dgtd::Options oc = opts;
        oc.kernel = DGTD_KERNEL_CSR;
        dgtd::HesthavenEvolutionHip csr(mesh, oc);
        dgtd::Vector yc;
        csr.Mult(x, yc);
        const double e3 = rel(yc, yr);
        dgtd::Mesh ms = mesh;
        const double sigma = 0.75;
        ms.sigma.assign(K, sigma);
        dgtd::HesthavenEvolutionHip cond(ms, oc);
        dgtd::Vector ys;
        cond.Mult(x, ys);
        dgtd::Vector want = yc;
        for (int64_t i = 0; i < n / 2; ++i) want[i] -= sigma * x[i];
        const double e4 = rel(ys, want);
        std::printf("csr rel %.3e, conductive rel %.3e\n", e3, e4);
        ok = ok && e3 < 1e-12 && e4 < 1e-13;
    }
    {  // triangles: 4 x 4 squares of the unit square, two triangles each, PEC
        const int m = 4;
        dgtd::Mesh tm;
        tm.dim = 2;
        for (int j = 0; j <= m; ++j)
            for (int i = 0; i <= m; ++i) tm.vx.insert(tm.vx.end(), {(double)i / m, (double)j / m, 0.0});
        auto vid = [&](int i, int j) { return j * (m + 1) + i; };
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                tm.EToV.insert(tm.EToV.end(), {vid(i, j), vid(i + 1, j), vid(i + 1, j + 1)});
                tm.face_bc.insert(tm.face_bc.end(), {(int8_t)(j == 0), (int8_t)(i == m - 1), 0});
                tm.EToV.insert(tm.EToV.end(), {vid(i, j), vid(i + 1, j + 1), vid(i, j + 1)});
                tm.face_bc.insert(tm.face_bc.end(), {0, (int8_t)(j == m - 1), (int8_t)(i == 0)});
            }
        dgtd::Vector out[3];
        const int kernels[3] = {DGTD_KERNEL_VALU, DGTD_KERNEL_MFMA, DGTD_KERNEL_CSR};
        dgtd::Vector xt;
        for (int k = 0; k < 3; ++k) {
            dgtd::Options ot = opts;
            ot.kernel = kernels[k];
            dgtd::HesthavenEvolutionHip tri(tm, ot);
            if (k == 0) {
                xt.resize(tri.Height());
                for (auto& v : xt) v = g(rng);
            }
            tri.Mult(xt, out[k]);
        }
        const double e5 = rel(out[1], out[0]), e6 = rel(out[2], out[0]);
        std::printf("triangles: mfma vs valu %.3e, csr vs valu %.3e\n", e5, e6);
        ok = ok && e5 < 1e-12 && e6 < 1e-12;
    }
    oracle_destroy(ref);
    std::printf(ok ? "OK\n" : "FAIL\n");
    return ok ? 0 : 1;
}
