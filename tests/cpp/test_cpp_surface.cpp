// The C++ TimeDependentOperator surface (include/dgtd_hip.hpp) against the
// CPU oracle: Mult and an RK4 Solver::run on a 2^3 PEC cavity at p=3, FP64.
// Exit status 0 and "OK" on success.
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
    oracle_destroy(ref);
    std::printf(ok ? "OK\n" : "FAIL\n");
    return ok ? 0 : 1;
}
