// dgtd_hip.hpp -- C++ surface over the C ABI in dgtd_hip.h.
//
// Mirrors the reference's operator surface: maxwell::HesthavenEvolution is an
// mfem::TimeDependentOperator with Mult(const Vector&, Vector&) const
// (src/evolution/HesthavenEvolution.h:14-50), driven by mfem::RK4Solver from
// Solver::step / Solver::run (src/solver/Solver.cpp:483-551).
//
//   dgtd::TimeDependentOperator    -- the same virtual interface, on plain
//                                     host vectors (dgtd::Vector)
//   dgtd::HesthavenEvolutionHip    -- the GPU operator behind it
//   dgtd::Solver                   -- run/step with the state resident on
//                                     the GPU (fused RK stages)
//   maxwell::HesthavenEvolutionHip -- mfem::TimeDependentOperator adapter,
//                                     compiled only when <mfem.hpp> exists
//
// Errors from the ABI become dgtd::Error (std::runtime_error), as the
// reference constructor throws std::runtime_error.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgtd_hip.h"

namespace dgtd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
    if (rc != DGTD_OK) throw Error(rc, dgtd_last_error());
}

// Host vector (the reference uses mfem::Vector).
using Vector = std::vector<double>;

class TimeDependentOperator {
public:
    explicit TimeDependentOperator(int64_t n) : height_(n) {}
    virtual ~TimeDependentOperator() = default;
    virtual void Mult(const Vector& in, Vector& out) const = 0;
    virtual void SetTime(double t) { t_ = t; }
    double GetTime() const { return t_; }
    int64_t Height() const { return height_; }
    int64_t Width() const { return height_; }

protected:
    int64_t height_;
    double t_ = 0.0;
};

// Mesh arrays as the reference's Model + ParFiniteElementSpace would provide
// them: vertex coordinates, tetrahedra, a boundary code per face (face f
// opposite local vertex f), optional eps_r/mu_r per element.
struct Mesh {
    int dim = 3;                  // 3: tetrahedra; 2: triangles (3 vertices and edges per element)
    std::vector<double> vx;       // 3 per vertex
    std::vector<int32_t> EToV;    // dim+1 per element
    std::vector<int8_t> face_bc;  // dim+1 per element, DGTD_BC_*
    std::vector<double> eps_r, mu_r;
    std::vector<double> sigma;    // bulk conductivity per element (DGTD_KERNEL_CSR only)

    dgtd_mesh_desc desc(int order) const
    {
        dgtd_mesh_desc d{};
        d.dim = dim;
        d.K = (int32_t)(EToV.size() / (dim + 1));
        d.order = order;
        d.vx = vx.data();
        d.nv = (int32_t)(vx.size() / 3);
        d.EToV = EToV.data();
        d.face_bc = face_bc.empty() ? nullptr : face_bc.data();
        d.eps_r = eps_r.empty() ? nullptr : eps_r.data();
        d.mu_r = mu_r.empty() ? nullptr : mu_r.data();
        d.sigma = sigma.empty() ? nullptr : sigma.data();
        return d;
    }

    static Mesh cube(int n, double L, const int8_t side_bc[6])
    {
        Mesh m;
        const size_t K = 6 * (size_t)n * n * n;
        m.vx.resize(3 * (size_t)(n + 1) * (n + 1) * (n + 1));
        m.EToV.resize(4 * K);
        m.face_bc.resize(4 * K);
        check(dgtd_cube_mesh(n, L, side_bc, m.vx.data(), m.EToV.data(), m.face_bc.data()));
        return m;
    }
};

struct Options {  // EvolutionOptions (src/evolution/EvolutionOptions.h:13-20) + implementation knobs
    double alpha = 1.0;
    int order = 2;
    int precision = 64;
    int integrator = DGTD_RK4;  // the reference's mfem::RK4Solver
    int device = 0;
    int kernel = DGTD_KERNEL_AUTO;
};

class HesthavenEvolutionHip : public TimeDependentOperator {
public:
    HesthavenEvolutionHip(const Mesh& mesh, const Options& opts) : TimeDependentOperator(0), opts_(opts)
    {
        const dgtd_mesh_desc d = mesh.desc(opts.order);
        dgtd_opts o{};
        o.alpha = opts.alpha;
        o.precision = opts.precision;
        o.integrator = opts.integrator;
        o.device = opts.device;
        o.kernel = opts.kernel;
        check(dgtd_op_create(&d, &o, &op_));
        check(dgtd_op_ndofs(op_, &height_));
    }
    ~HesthavenEvolutionHip() override { dgtd_op_destroy(op_); }
    HesthavenEvolutionHip(const HesthavenEvolutionHip&) = delete;
    HesthavenEvolutionHip& operator=(const HesthavenEvolutionHip&) = delete;

    void Mult(const Vector& in, Vector& out) const override
    {
        if ((int64_t)in.size() != height_) throw Error(DGTD_E_ARG, "Mult: input size != Height()");
        out.resize(height_);
        check(dgtd_op_mult(op_, t_, in.data(), out.data()));
    }

    // Device vectors in the operator's precision; asynchronous on `stream`.
    void MultDevice(const void* d_in, void* d_out, void* stream = nullptr) const
    {
        check(dgtd_op_mult_device(op_, t_, d_in, d_out, stream));
    }

    // getHesthavenElement(e).fscale.maxCoeff() over e (Solver.cpp:337-340).
    double MaxFscale() const
    {
        double v = 0;
        check(dgtd_op_max_fscale(op_, &v));
        return v;
    }

    const Options& options() const { return opts_; }
    dgtd_op* handle() const { return op_; }

private:
    Options opts_;
    dgtd_op* op_ = nullptr;
};

// Solver::run / step (Solver.cpp:483-551) with the state on the GPU.
class Solver {
public:
    Solver(HesthavenEvolutionHip& op, const Vector& x0, double final_time, double time_step = 0.0,
           double cfl = 1.0)
        : op_(op), final_time_(final_time)
    {
        const int p = op.options().order;
        dt_ = time_step == 0.0 ? cfl / (op.MaxFscale() * p * p) : time_step;  // Solver.cpp:334-346
        check(dgtd_op_set_state(op.handle(), x0.data(), 0.0));
    }
    void step()
    {
        double t = 0;
        check(dgtd_op_time(op_.handle(), &t));
        check(dgtd_op_advance(op_.handle(), std::min(dt_, final_time_ - t), 1));
    }
    int run()
    {
        int32_t n = 0;
        check(dgtd_op_run(op_.handle(), dt_, final_time_, &n));
        return n;
    }
    Vector getFields() const
    {
        Vector x(op_.Height());
        double t = 0;
        check(dgtd_op_get_state(op_.handle(), x.data(), &t));
        return x;
    }
    double getTime() const
    {
        double t = 0;
        check(dgtd_op_time(op_.handle(), &t));
        return t;
    }
    double getTimeStep() const { return dt_; }

private:
    HesthavenEvolutionHip& op_;
    double final_time_, dt_ = 0.0;
};

}  // namespace dgtd

#if __has_include(<mfem.hpp>)
#include <mfem.hpp>
namespace maxwell {
// Drop-in for maxwell::HesthavenEvolution in Solver::assignEvolutionOperator
// (Solver.cpp:24-26): same base class, same Mult signature.
class HesthavenEvolutionHip : public mfem::TimeDependentOperator {
public:
    HesthavenEvolutionHip(const dgtd::Mesh& mesh, const dgtd::Options& opts)
        : impl_(mesh, opts)
    {
        height = width = (int)impl_.Height();
    }
    // HostRead/HostWrite: under an MFEM device backend (the reference
    // launcher's -d option) the valid copy of `in` may be on the device, so
    // it is brought to the host first, and `out` is marked valid on the host
    // (GetData() would read a stale host copy and leave the result unseen).
    void Mult(const mfem::Vector& in, mfem::Vector& out) const override
    {
        const double* x = in.HostRead();
        double* y = out.HostWrite();
        dgtd::check(dgtd_op_mult(impl_.handle(), GetTime(), x, y));
    }
    double MaxFscale() const { return impl_.MaxFscale(); }

private:
    dgtd::HesthavenEvolutionHip impl_;
};
}  // namespace maxwell
#endif
