// dgtd_launch: runs a case file of the reference's JSON launcher
// (src/launcher/launcher.cpp, src/driver/driver.cpp) on the GPU hesthaven
// operator.  Mirrors Solver::run (src/solver/Solver.cpp:483-551) with point
// probes written as ProbesManager writes them (ProbesManager.cpp:183-218,
// 375-406): Exports/<case>/PointProbes/PointProbe<ID>.dat, one line per
// sample with time / c (seconds) and Ex Ey Ez Hx Hy Hz; field probes as
// FieldProbes/FieldProbe<ID>.dat with one component (:222-259, 344-371); and domain snapshots
// as ProbesManager writes them (:457-495): DomainSnapshotProbes/meshes/mesh_rank0
// and DomainSnapshotProbes/rank_0/cycle_NNNNNN/{Ex,Ey,Ez,Hx,Hy,Hz}.gf, time.txt.
//
//   dgtd_launch CASE.json [--precision 64|32] [--integrator rk4|lserk4]
//               [--kernel auto|valu|mfma|csr] [--final-time T] [--out DIR]
// A case with conductive materials runs on the assembled operator (csr) under
// --kernel auto.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "dgtd_hip.h"

namespace {

constexpr double kSpeedOfLightSI = 299792458.0;  // physicalConstants::speedOfLight_SI

void die(const char* what)
{
    std::fprintf(stderr, "dgtd_launch: %s: %s\n", what, dgtd_last_error());
    std::exit(1);
}

void check(int rc, const char* what)
{
    if (rc != DGTD_OK) die(what);
}

std::string stem(const std::string& path)
{
    std::string s = std::filesystem::path(path).filename().string();
    const size_t dot = s.rfind('.');
    return dot == std::string::npos ? s : s.substr(0, dot);
}

void usage()
{
    std::fprintf(stderr,
                 "usage: dgtd_launch CASE.json [--precision 64|32] [--integrator rk4|lserk4]\n"
                 "                   [--kernel auto|valu|mfma|csr] [--final-time T] [--out DIR]\n");
    std::exit(2);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) usage();
    const std::string case_path = argv[1];
    int precision = 64, integrator = DGTD_RK4, kernel = DGTD_KERNEL_AUTO;
    double final_time = -1.0;
    std::string out_dir;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) usage();
            return argv[++i];
        };
        if (a == "--precision") precision = std::atoi(val().c_str());
        else if (a == "--integrator") {
            const std::string v = val();
            integrator = v == "lserk4" ? DGTD_LSERK4 : v == "rk4" ? DGTD_RK4 : -1;
            if (integrator < 0) usage();
        } else if (a == "--kernel") {
            const std::string v = val();
            kernel = v == "auto"   ? DGTD_KERNEL_AUTO
                     : v == "valu" ? DGTD_KERNEL_VALU
                     : v == "mfma" ? DGTD_KERNEL_MFMA
                     : v == "csr"  ? DGTD_KERNEL_CSR
                                   : -1;
            if (kernel < 0) usage();
        } else if (a == "--final-time") final_time = std::atof(val().c_str());
        else if (a == "--out") out_dir = val();
        else usage();
    }

    const auto t_setup = std::chrono::steady_clock::now();
    dgtd_case* cs = nullptr;
    check(dgtd_case_load(case_path.c_str(), &cs), "case");
    dgtd_case_info info{};
    check(dgtd_case_get_info(cs, &info), "case info");
    if (final_time < 0) final_time = info.final_time;
    dgtd_mesh_desc mesh{};
    check(dgtd_case_mesh(cs, &mesh), "case mesh");

    dgtd_opts o{};
    o.alpha = info.alpha;
    o.precision = precision;
    o.integrator = integrator;
    // conductivity, or the Global operator's own SMA weighting (alpha != 1
    // with SMA faces), needs the assembled operator
    bool sma = false;
    const int nfaces = mesh.nvert == 8 ? 6 : mesh.dim == 2 && mesh.nvert != 4 ? 3 : 4;
    for (int64_t i = 0; i < (int64_t)mesh.K * nfaces; ++i)
        sma = sma || mesh.face_bc[i] == DGTD_BC_SMA || mesh.face_bc[i] == DGTD_BC_INT_SMA;
    const bool global_sma = info.evolution_operator != DGTD_OPERATOR_HESTHAVEN && sma && info.alpha != 1.0;
    // (hexahedra: AUTO takes their matrix-free stage, which applies the Global
    // operator's flux and conductivity)
    o.kernel = kernel == DGTD_KERNEL_AUTO && mesh.nvert != 8 && (mesh.sigma || global_sma) ? DGTD_KERNEL_CSR : kernel;
    dgtd_op* op = nullptr;
    check(dgtd_op_create(&mesh, &o, &op), "operator");

    int64_t ndofs = 0;
    check(dgtd_op_ndofs(op, &ndofs), "ndofs");
    std::vector<dgtd_initial_field> src(info.nsources > 0 ? info.nsources : 1);
    check(dgtd_case_sources(cs, src.data()), "sources");
    std::vector<double> x((size_t)ndofs, 0.0);
    check(dgtd_project_initial_fields(&mesh, src.data(), info.nsources, x.data()), "initial fields");
    check(dgtd_op_set_state(op, x.data(), 0.0), "set state");
    if (info.nplanewaves > 0) {  // TF/SF plane waves
        std::vector<dgtd_planewave> waves(info.nplanewaves);
        check(dgtd_case_planewaves(cs, waves.data()), "plane waves");
        check(dgtd_op_set_planewaves(op, waves.data(), info.nplanewaves), "set plane waves");
    }

    // dt: time_step, or the reference's automatic estimate (Solver.cpp:110-115, 302-353)
    double dt = 0.0;
    check(dgtd_case_time_step(cs, &dt), "time step");

    // point probes, then field probes, in one device probe set, and their
    // sampling intervals (driver.cpp:717-733, 764-778)
    const int npp = info.nprobes, nfp = info.nfield_probes, np = npp + nfp;
    std::vector<double> pxyz(3 * (size_t)np);
    std::vector<int32_t> psteps(np), psaves(np), pelem(np), interval(np, 1), comp(np, -1);
    check(dgtd_case_probes(cs, pxyz.data(), psteps.data(), psaves.data()), "probes");
    check(dgtd_case_field_probes(cs, pxyz.data() + 3 * (size_t)npp, psteps.data() + npp, psaves.data() + npp,
                                 comp.data() + npp),
          "field probes");
    if (np > 0) check(dgtd_op_probe_set(op, np, pxyz.data(), pelem.data()), "probe set");
    const int total_steps = (int)std::ceil(final_time / dt);
    for (int p = 0; p < np; ++p)
        interval[p] = psteps[p] > 0 ? psteps[p] : psaves[p] > 0 ? std::max(1, total_steps / psaves[p]) : 1;

    if (out_dir.empty()) out_dir = "Exports/" + stem(case_path);
    std::vector<std::ofstream> files(np);
    static const char* const kComp[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    for (int kind = 0; kind < 2; ++kind) {  // PointProbes (:186-218), FieldProbes (:222-259)
        const int first = kind == 0 ? 0 : npp, count = kind == 0 ? npp : nfp;
        if (count == 0) continue;
        const std::string name = kind == 0 ? "PointProbe" : "FieldProbe";
        const std::string base = out_dir + "/" + name + "s/";
        std::filesystem::create_directories(base);
        for (int p = first; p < first + count; ++p) {
            if (pelem[p] < 0) continue;  // outside the mesh: no file, as elementId -2
            const int id = p - first;
            files[p].open(base + name + std::to_string(id) + ".dat", std::ios::trunc);
            files[p] << name << " ID " << id << "\n";
            files[p] << "Spatial Position (X, Y, Z) \n";
            files[p] << std::scientific << std::setprecision(5);
            files[p] << std::to_string(pxyz[3 * p]) + " " + std::to_string(pxyz[3 * p + 1]) + " " +
                            std::to_string(pxyz[3 * p + 2])
                     << "\n";
            if (kind == 0) files[p] << "Time (s) // Ex // Ey // Ez // Hx // Hy // Hz \n";
            else files[p] << "Time (s) // " << kComp[comp[p]] << "\n";
        }
    }
    // domain snapshots (ProbesManager.cpp:457-495): the mesh once, then
    // Ex..Hz GridFunctions and time.txt per sampled cycle
    const bool snapshots = info.snapshot_steps >= 0;
    const int snap_every = info.snapshot_steps > 0 ? info.snapshot_steps
                           : info.snapshot_saves > 0 ? std::max(1, total_steps / info.snapshot_saves) : 1;
    const std::string snap_base = out_dir + "/DomainSnapshotProbes/";
    if (snapshots) {
        std::filesystem::create_directories(snap_base + "meshes");
        check(dgtd_case_save_mesh(cs, (snap_base + "meshes/mesh_rank0").c_str()), "snapshot mesh");
    }
    long nsnap = 0;
    auto snapshot = [&](long cycle, double t, bool final) {
        if (!snapshots || !(final || cycle % snap_every == 0)) return;
        char name[32];
        std::snprintf(name, sizeof(name), "cycle_%06ld", cycle);
        const std::string dir = snap_base + "rank_0/" + name;
        std::filesystem::create_directories(dir);
        double tt = 0.0;
        check(dgtd_op_get_state(op, x.data(), &tt), "snapshot state");
        check(dgtd_case_save_fields(cs, dir.c_str(), ndofs, x.data(), t), "snapshot fields");
        ++nsnap;
    };
    std::vector<double> vals(6 * (size_t)np);
    auto sample = [&](long cycle, double t, bool final) {
        bool any = false;
        for (int p = 0; p < np; ++p) any = any || (pelem[p] >= 0 && (final || cycle % interval[p] == 0));
        if (!any) return;
        check(dgtd_op_probe_sample(op, vals.data()), "probe sample");
        for (int p = 0; p < np; ++p) {
            if (pelem[p] < 0 || !(final || cycle % interval[p] == 0)) continue;
            files[p] << t / kSpeedOfLightSI;
            if (comp[p] < 0)
                for (int c = 0; c < 6; ++c) files[p] << " " << vals[6 * p + c];
            else
                files[p] << " " << vals[6 * p + comp[p]];
            files[p] << "\n";
        }
    };
    const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_setup).count();

    const auto t_run = std::chrono::steady_clock::now();
    double t = 0.0;
    long steps = 0;
    if (np == 0 && !snapshots) {
        int32_t n = 0;
        check(dgtd_op_run(op, dt, final_time, &n), "run");
        steps = n;
        check(dgtd_op_time(op, &t), "time");
    } else {
        long cycle = 0;
        auto update = [&]() {  // ProbesManager::updateProbes (ProbesManager.cpp:498-528)
            const bool final = std::fabs(t - final_time) < 1e-8;
            sample(cycle, t, final);
            snapshot(cycle, t, final);
            ++cycle;
        };
        update();
        while (t <= final_time - 1e-8 * dt) {  // Solver::run loop (Solver.cpp:497)
            const double truedt = std::min(dt, final_time - t);  // Solver::step (:537)
            check(dgtd_op_advance(op, truedt, 1), "advance");
            check(dgtd_op_time(op, &t), "time");
            ++steps;
            update();
        }
    }
    double norms[3] = {0, 0, 0};
    check(dgtd_op_norms(op, norms), "norms");
    int32_t unstable = 0;
    check(dgtd_op_unstable(op, &unstable), "stability");
    if (unstable)  // the warning of Solver::run (Solver.cpp:499-516); the run went on, as there
        std::fprintf(stderr,
                     "WARNING: Simulation is potentially unstable (a step stored a non-finite field value or one "
                     "above 1e20; final field norm = %g).\n  Time: %g / %g,  dt: %g\n"
                     "  Verify your setup and consider lowering the time step.\n",
                     norms[0], t, final_time, dt);
    const double run_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_run).count();
    const int stages = integrator == DGTD_LSERK4 ? 5 : 4;
    std::printf("{\"case\": \"%s\", \"K\": %d, \"order\": %d, \"ndofs\": %lld, \"precision\": %d, "
                "\"integrator\": \"%s\", \"dt\": %.10e, \"steps\": %ld, \"time\": %.10e, \"norm2\": %.10e, "
                "\"energy\": %.10e, \"nonfinite\": %d, \"unstable\": %d, \"probes\": %d, \"setup_s\": %.3f, \"run_s\": %.3f, "
                "\"dof_stages_per_s\": %.4e, \"snapshots\": %ld, \"out\": \"%s\"}\n",
                stem(case_path).c_str(), info.K, info.order, (long long)ndofs, precision,
                integrator == DGTD_LSERK4 ? "lserk4" : "rk4", dt, steps, t, norms[0], norms[1], (int)norms[2], (int)unstable, np,
                setup_s, run_s, run_s > 0 ? (double)ndofs * stages * steps / run_s : 0.0, nsnap, out_dir.c_str());
    for (auto& f : files)
        if (f.is_open()) f.close();
    dgtd_op_destroy(op);
    dgtd_case_destroy(cs);
    return norms[2] != 0.0 ? 3 : 0;
}
