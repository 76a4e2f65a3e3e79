// Case files: the solver_options / model / sources / probes subset of the
// reference's JSON launcher (driver.cpp:531-706, 940-1035, 1242-1300) that the
// hesthaven path can run, read into plain arrays for dgtd_op_create.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgtd_hip.h"
#include "json.hpp"
#include "meshio.hpp"

namespace dgtd {
int api_fail(int code, const std::string& msg);  // dgtd_api.cpp
}

struct dgtd_case {
    dgtd::TetMesh mesh;
    std::vector<int32_t> EToV;
    std::vector<int8_t> face_bc;
    std::vector<double> eps, mu, sigma;  // sigma: normalised conductivity, empty if none
    dgtd_case_info info{};
    std::vector<dgtd_initial_field> sources;
    std::vector<double> probes;  // 3 per point probe
    std::vector<int32_t> probe_steps, probe_saves;
    std::vector<double> fprobes;  // field probes: 3 per probe
    std::vector<int32_t> fprobe_steps, fprobe_saves, fprobe_comp;
    std::vector<dgtd_planewave> waves;
    std::vector<int> tfsf_tags;
    std::vector<bool> wave_auto_mean;  // mean from the TF/SF surface (no "mean" in the case)
};

namespace {

struct Unsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};

std::string dir_of(const std::string& p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

// assemble3DVector (driver.cpp:346-353): up to three entries, the rest zero.
void vec3(const dgtd::Json& j, double out[3])
{
    const std::vector<double> v = j.numbers();
    if (v.size() > 3) throw std::runtime_error("vector with more than 3 entries");
    for (int d = 0; d < 3; ++d) out[d] = d < (int)v.size() ? v[d] : 0.0;
}

int field_type(const dgtd::Json& j)
{
    const std::string& s = j.string();
    if (s == "electric" || s == "E") return 0;
    if (s == "magnetic" || s == "H") return 1;
    throw std::runtime_error("unknown field_type '" + s + "'");
}

void read_solver_options(const dgtd::Json& c, dgtd_case_info& in)
{
    // SolverOptions defaults (SolverOptions.h:20-23, EvolutionOptions.h:13-17)
    in.order = 2;
    in.alpha = 1.0;
    in.time_step = 0.0;
    in.final_time = 2.0;
    in.cfl = 1.0;
    in.evolution_operator = DGTD_OPERATOR_GLOBAL;
    if (!c.contains("solver_options")) return;
    const dgtd::Json& s = c.at("solver_options");
    if (s.contains("upwind_alpha")) in.alpha = s.at("upwind_alpha").number();
    if (s.contains("time_step")) in.time_step = s.at("time_step").number();
    if (s.contains("final_time")) in.final_time = s.at("final_time").number();
    if (s.contains("cfl")) in.cfl = s.at("cfl").number();
    if (s.contains("order")) in.order = s.at("order").integer();
    if (s.contains("evolution_operator")) {
        const std::string& op = s.at("evolution_operator").string();
        if (op == "hesthaven") in.evolution_operator = DGTD_OPERATOR_HESTHAVEN;
        else if (op == "global") in.evolution_operator = DGTD_OPERATOR_GLOBAL;
        else if (op == "maxwell") in.evolution_operator = DGTD_OPERATOR_MAXWELL;
        else
            throw std::runtime_error("Wrong type of Evolution Operator defined, please choose: 'maxwell', 'global' or "
                                     "'hesthaven'.");
    }
    if (s.contains("basis_type") && s.at("basis_type").integer() != 1)
        throw Unsupported("only GaussLobatto nodes (basis_type 1) are supported");
    if (s.contains("ode_type") && s.at("ode_type").integer() != 0)
        throw Unsupported("only the explicit RK4 integrator (ode_type 0) is supported");
    if (in.order < 1 || in.order > 5) throw Unsupported("order must be in 1..5");
    if (!(in.final_time >= 0) || !(in.time_step >= 0) || !(in.cfl > 0))
        throw std::runtime_error("final_time, time_step must be >= 0 and cfl > 0");
}

void read_sources(const dgtd::Json& c, dgtd_case& cs)
{
    if (!c.contains("sources")) return;
    const dgtd::Json& ss = c.at("sources");
    for (size_t i = 0; i < ss.size(); ++i) {
        const dgtd::Json& s = ss.at(i);
        const std::string& type = s.at("type").string();
        if (type == "planewave") {  // driver.cpp:565-609
            const dgtd::Json& mag = s.at("magnitude");
            if (mag.contains("type") && mag.at("type").string() != "gaussian")
                throw Unsupported("plane-wave magnitude '" + mag.at("type").string() + "' is not supported");
            dgtd_planewave w{};
            vec3(s.at("polarization"), w.polarization);
            vec3(s.at("propagation"), w.propagation);
            w.spread = mag.at("spread").number();
            w.kind = DGTD_PLANEWAVE_GAUSSIAN;
            if (mag.contains("frequency")) {
                w.kind = DGTD_PLANEWAVE_MODULATED;
                w.frequency = mag.at("frequency").number() / 299792458.0;  // f / c_SI (driver.cpp:515)
            }
            const double nd = std::sqrt(w.propagation[0] * w.propagation[0] + w.propagation[1] * w.propagation[1] +
                                        w.propagation[2] * w.propagation[2]);
            if (!(nd > 0)) throw std::runtime_error("plane-wave propagation must be nonzero");
            const bool auto_mean = !mag.contains("mean");
            if (!auto_mean) {  // projected mean k.mean (buildGaussianPlanewave, driver.cpp:491-496)
                double m[3];
                vec3(mag.at("mean"), m);
                w.mean = (m[0] * w.propagation[0] + m[1] * w.propagation[1] + m[2] * w.propagation[2]) / nd;
            }
            if (s.contains("tags")) {
                const dgtd::Json& tags = s.at("tags");
                for (size_t t = 0; t < tags.size(); ++t) cs.tfsf_tags.push_back(tags.at(t).integer());
            }
            cs.waves.push_back(w);
            cs.wave_auto_mean.push_back(auto_mean);
            continue;
        }
        if (type != "initial") throw Unsupported("source type '" + type + "' (dipole) is not supported");
        const dgtd::Json& mag = s.at("magnitude");
        const std::string& mt = mag.at("type").string();
        dgtd_initial_field f{};
        f.field = field_type(s.at("field_type"));
        vec3(s.at("polarization"), f.polarization);
        if (mt == "gaussian") {  // buildGaussianInitialField, driver.cpp:368-380
            f.kind = DGTD_FIELD_GAUSSIAN;
            f.spread = mag.at("spread").number();
            f.dimension = s.at("dimension").integer();
            vec3(s.at("center"), f.center);
        } else if (mt == "resonant") {  // buildResonantModeInitialField, driver.cpp:382-391
            f.kind = DGTD_FIELD_RESONANT;
            const std::vector<double> modes = mag.at("modes").numbers();
            if (modes.empty() || modes.size() > 3) throw std::runtime_error("resonant modes must have 1..3 entries");
            f.nmodes = (int32_t)modes.size();
            for (size_t k = 0; k < modes.size(); ++k) f.modes[k] = (int32_t)modes[k];
        } else if (mt == "besselj6_2D") {  // buildBesselJ6InitialField, driver.cpp:393-400, centre 0
            f.kind = DGTD_FIELD_BESSELJ6_2D;
        } else if (mt == "besselj6_3D") {  // buildSphericalBesselJ6InitialField, driver.cpp:474-481, centre 0
            f.kind = DGTD_FIELD_BESSELJ6_3D;
        } else {
            throw Unsupported("initial field magnitude '" + mt + "' is not supported");
        }
        cs.sources.push_back(f);
    }
}

void read_probes(const dgtd::Json& c, dgtd_case& cs)
{
    cs.info.exporter_steps = 0;
    cs.info.snapshot_steps = -1;
    cs.info.snapshot_saves = 0;
    if (!c.contains("probes")) return;
    const dgtd::Json& p = c.at("probes");
    if (p.contains("exporter") && p.at("exporter").contains("steps"))
        cs.info.exporter_steps = p.at("exporter").at("steps").integer();
    if (p.contains("domain_snapshot")) {  // driver.cpp:808-820
        const dgtd::Json& d = p.at("domain_snapshot");
        cs.info.snapshot_steps = d.contains("steps") ? d.at("steps").integer() : 0;
        cs.info.snapshot_saves = d.contains("saves") ? d.at("saves").integer() : 0;
    }
    if (p.contains("field")) {  // driver.cpp:764-778
        const dgtd::Json& fs = p.at("field");
        for (size_t i = 0; i < fs.size(); ++i) {
            const dgtd::Json& fj = fs.at(i);
            const std::string& pol = fj.at("polarization").string();
            const int d = pol == "X" ? 0 : pol == "Y" ? 1 : pol == "Z" ? 2 : -1;
            if (d < 0) throw std::runtime_error("Wrong Field Polarization in Point Probe assignation.");
            double x[3];
            vec3(fj.at("position"), x);
            cs.fprobes.insert(cs.fprobes.end(), x, x + 3);
            cs.fprobe_comp.push_back(3 * field_type(fj.at("field_type")) + d);
            cs.fprobe_steps.push_back(fj.contains("steps") ? fj.at("steps").integer() : 0);
            cs.fprobe_saves.push_back(fj.contains("saves") ? fj.at("saves").integer() : 0);
        }
    }
    if (p.contains("point")) {
        const dgtd::Json& pts = p.at("point");
        for (size_t i = 0; i < pts.size(); ++i) {
            const dgtd::Json& pj = pts.at(i);
            double x[3];
            vec3(pj.at("position"), x);
            cs.probes.insert(cs.probes.end(), x, x + 3);
            cs.probe_steps.push_back(pj.contains("steps") ? pj.at("steps").integer() : 0);
            cs.probe_saves.push_back(pj.contains("saves") ? pj.at("saves").integer() : 0);
        }
    }
}

void read_model(const dgtd::Json& c, const std::string& case_path, dgtd_case& cs)
{
    if (!c.contains("model")) throw std::runtime_error("JSON data does not include 'model'.");
    const dgtd::Json& m = c.at("model");
    std::string file = m.at("filename").string();
    if (file.empty() || file[0] != '/') file = dir_of(case_path) + "/" + file;
    cs.mesh = dgtd::read_mesh_file(file);
    dgtd::orient_positive(cs.mesh);
    cs.info.refinement = m.contains("refinement") ? m.at("refinement").integer() : 0;
    if (cs.info.refinement < 0 || cs.info.refinement > 6) throw std::runtime_error("refinement must be in 0..6");
    for (int r = 0; r < cs.info.refinement; ++r) cs.mesh = dgtd::refine_uniform(cs.mesh);

    // materials by element attribute (driver.cpp:946-966); the operator
    // applies eps_r, mu_r per element as 1/eps, 1/mu on the rhs, and the bulk
    // conductivity, normalised sigma_SI * Z0 (driver.cpp:962-963,
    // PhysicalConstants.h:9-15), as -sigma/eps E (assembled operator only)
    constexpr double kZ0 = 4.0e-7 * M_PI * 299792458.0;
    std::map<int, std::array<double, 3>> mat;
    if (!m.contains("materials")) throw std::runtime_error("JSON data does not include 'materials'.");
    const dgtd::Json& mats = m.at("materials");
    for (size_t i = 0; i < mats.size(); ++i) {
        const dgtd::Json& mj = mats.at(i);
        double eps = 1.0, mu = 1.0;
        if (mj.contains("relative_permittivity")) eps = mj.at("relative_permittivity").number();
        if (mj.contains("relative_permeability")) mu = mj.at("relative_permeability").number();
        const double sig = mj.contains("bulk_conductivity") ? mj.at("bulk_conductivity").number() * kZ0 : 0.0;
        if (!(eps > 0) || !(mu > 0)) throw std::runtime_error("relative permittivity/permeability must be > 0");
        if (!(sig >= 0.0)) throw std::runtime_error("Conductivity under 0.0 not allowed.");
        const dgtd::Json& tags = mj.at("tags");
        for (size_t t = 0; t < tags.size(); ++t) mat[tags.at(t).integer()] = {eps, mu, sig};
    }
    const int32_t K = cs.mesh.K();
    cs.eps.assign(K, 1.0);
    cs.mu.assign(K, 1.0);
    cs.sigma.assign(K, 0.0);
    bool any = false, conductive = false;
    for (int32_t e = 0; e < K; ++e) {
        auto it = mat.find(cs.mesh.elem_attr()[e]);
        if (it == mat.end()) continue;
        cs.eps[e] = it->second[0];
        cs.mu[e] = it->second[1];
        cs.sigma[e] = it->second[2];
        any = any || it->second[0] != 1.0 || it->second[1] != 1.0;
        conductive = conductive || it->second[2] != 0.0;
    }
    if (!any) {
        cs.eps.clear();
        cs.mu.clear();
    }
    if (!conductive) cs.sigma.clear();

    // boundaries by triangle attribute (driver.cpp:870-887, 994-1035)
    if (!m.contains("boundaries")) throw std::runtime_error("JSON data does not include 'boundaries'.");
    int max_attr = 0;
    for (int a : cs.mesh.bdr_attr()) max_attr = std::max(max_attr, a);
    std::vector<int> attr_bc(max_attr + 1, -1);
    const dgtd::Json& bs = m.at("boundaries");
    for (size_t i = 0; i < bs.size(); ++i) {
        const dgtd::Json& b = bs.at(i);
        const std::string& type = b.at("type").string();
        int code;
        if (type == "PEC") code = DGTD_BC_PEC;
        else if (type == "PMC") code = DGTD_BC_PMC;
        else if (type == "SMA") code = DGTD_BC_SMA;
        else if (type == "SGBC") throw Unsupported("SGBC boundaries are not supported");
        else throw std::runtime_error("The defined Boundary Type " + type + " is incorrect.");
        const dgtd::Json& tags = b.at("tags");
        if (tags.size() == 0) throw std::runtime_error("boundary without tags");
        for (size_t t = 0; t < tags.size(); ++t) {
            const int a = tags.at(t).integer();
            if (a < 0 || a > max_attr)
                throw std::runtime_error("There is no bdr_attribute " + std::to_string(a) +
                                         " defined in the mesh, but it is defined in the JSON.");
            attr_bc[a] = code;
        }
    }
    for (int a : cs.tfsf_tags) {
        if (a < 0 || a > max_attr)
            throw std::runtime_error("There is no bdr_attribute " + std::to_string(a) +
                                     " defined in the mesh, but it is defined in the JSON.");
        attr_bc[a] = DGTD_BC_TFSF;
    }
    cs.face_bc = dgtd::face_codes(cs.mesh, attr_bc);
    // automatic plane-wave delay: the pulse starts 5 spreads (x sqrt 2) before
    // the upstream point of the TF/SF surface (driver.cpp:408, 455-472, 575-588)
    for (size_t i = 0; i < cs.waves.size(); ++i) {
        if (!cs.wave_auto_mean[i]) continue;
        dgtd_planewave& w = cs.waves[i];
        const double nd = std::sqrt(w.propagation[0] * w.propagation[0] + w.propagation[1] * w.propagation[1] +
                                    w.propagation[2] * w.propagation[2]);
        double min_phase = 0.0;
        bool any = false;
        // TF/SF boundary elements: segments in 2D, triangles on tetrahedra,
        // quadrilaterals on hexahedra
        const bool two_d = cs.mesh.dim == 2, hexm = cs.mesh.hex();
        const size_t nb = two_d ? cs.mesh.segs.size() : hexm ? cs.mesh.quads.size() : cs.mesh.tris.size();
        for (size_t t = 0; t < nb; ++t) {
            const int a = two_d ? cs.mesh.seg_attr[t] : hexm ? cs.mesh.quad_attr[t] : cs.mesh.tri_attr[t];
            if (std::find(cs.tfsf_tags.begin(), cs.tfsf_tags.end(), a) == cs.tfsf_tags.end()) continue;
            const int nvt = two_d ? 2 : hexm ? 4 : 3;
            auto vert = [&](int k) {
                return two_d ? cs.mesh.segs[t][k] : hexm ? cs.mesh.quads[t][k] : cs.mesh.tris[t][k];
            };
            double ph = 0.0;
            for (int d = 0; d < 3; ++d) {
                double c = 0.0;
                for (int k = 0; k < nvt; ++k) c += cs.mesh.vx[3 * (size_t)vert(k) + d];
                ph += c / nvt * w.propagation[d] / nd;
            }
            min_phase = any ? std::min(min_phase, ph) : ph;
            any = true;
        }
        w.mean = min_phase - 5.0 * w.spread * std::sqrt(2.0);
    }
    const int nvert = cs.mesh.nvert();
    cs.EToV.resize((size_t)nvert * K);
    for (int32_t e = 0; e < K; ++e)
        for (int j = 0; j < nvert; ++j) cs.EToV[(size_t)nvert * e + j] = cs.mesh.vertex(e, j);
    cs.info.dim = cs.mesh.dim;
    cs.info.K = K;
    cs.info.nv = cs.mesh.nv();
}

}  // namespace

extern "C" int dgtd_case_load(const char* json_path, dgtd_case** out)
{
    if (!json_path || !out) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    *out = nullptr;
    try {
        auto cs = std::make_unique<dgtd_case>();
        const dgtd::Json c = dgtd::Json::parse_file(json_path);
        if (!c.is_object()) throw std::runtime_error("case file is not a JSON object");
        read_solver_options(c, cs->info);
        read_sources(c, *cs);  // plane-wave tags mark TF/SF faces in the model
        read_model(c, json_path, *cs);
        read_probes(c, *cs);
        cs->info.nsources = (int32_t)cs->sources.size();
        cs->info.nplanewaves = (int32_t)cs->waves.size();
        cs->info.nprobes = (int32_t)(cs->probes.size() / 3);
        cs->info.nfield_probes = (int32_t)cs->fprobe_comp.size();
        if (cs->mesh.quad() && cs->info.evolution_operator == DGTD_OPERATOR_HESTHAVEN)
            // the reference's hesthaven quadrilateral cases are disabled
            // (ExtensiveCasesTest.cpp:1159 DISABLED_2D_TFSF_Centered_Quads_Hesthaven)
            throw Unsupported("quadrilateral meshes run on the Global operator only, not on hesthaven");
        if (cs->mesh.hex() && cs->info.evolution_operator == DGTD_OPERATOR_HESTHAVEN)
            // the hesthaven operator builds a tetrahedral reference element in 3D
            // (HesthavenEvolution.cpp:39-40)
            throw Unsupported("hexahedral meshes run on the Global operator only, not on hesthaven");
        if (cs->info.alpha == 0.0 && cs->info.evolution_operator == DGTD_OPERATOR_HESTHAVEN)
            for (int8_t b : cs->face_bc)
                if (b == DGTD_BC_SMA)  // driver.cpp:1596-1598 (the Global operator runs it)
                    throw Unsupported("Centered SMA with Hesthaven Evolution Operator not supported yet.");
        *out = cs.release();
    } catch (const Unsupported& e) {
        return dgtd::api_fail(DGTD_E_UNSUPPORTED, e.what());
    } catch (const dgtd::MeshUnsupported& e) {
        return dgtd::api_fail(DGTD_E_UNSUPPORTED, e.what());
    } catch (const std::exception& e) {
        return dgtd::api_fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

extern "C" void dgtd_case_destroy(dgtd_case* c) { delete c; }

extern "C" int dgtd_case_get_info(const dgtd_case* c, dgtd_case_info* info)
{
    if (!c || !info) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    *info = c->info;
    return DGTD_OK;
}

extern "C" int dgtd_case_mesh(const dgtd_case* c, dgtd_mesh_desc* mesh)
{
    if (!c || !mesh) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    std::memset(mesh, 0, sizeof(*mesh));
    mesh->K = c->info.K;
    mesh->order = c->info.order;
    mesh->vx = c->mesh.vx.data();
    mesh->nv = c->info.nv;
    mesh->EToV = c->EToV.data();
    mesh->face_bc = c->face_bc.data();
    mesh->eps_r = c->eps.empty() ? nullptr : c->eps.data();
    mesh->mu_r = c->mu.empty() ? nullptr : c->mu.data();
    mesh->dim = c->mesh.dim;
    mesh->sigma = c->sigma.empty() ? nullptr : c->sigma.data();
    mesh->nvert = c->mesh.quad() ? 4 : c->mesh.hex() ? 8 : 0;
    return DGTD_OK;
}

extern "C" int dgtd_case_time_step(const dgtd_case* c, double* dt)
{
    if (!c || !dt) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    if (c->info.time_step > 0.0) {
        *dt = c->info.time_step;
        return DGTD_OK;
    }
    dgtd_mesh_desc mesh{};
    if (int rc = dgtd_case_mesh(c, &mesh)) return rc;
    return dgtd_mesh_estimate_time_step(&mesh, c->info.evolution_operator, c->info.cfl, dt);
}

extern "C" int dgtd_case_save_mesh(const dgtd_case* c, const char* path)
{
    if (!c || !path) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    try {
        dgtd::write_mfem_mesh(c->mesh, path);
    } catch (const std::exception& e) {
        return dgtd::api_fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

namespace {

int save_fields(const char* dir, int dim, int32_t order, int64_t ndofs, const double* x, double t)
{
    if (!dir || !x) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    if (order < 1 || order > 5) return dgtd::api_fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    if (ndofs <= 0 || ndofs % 6) return dgtd::api_fail(DGTD_E_ARG, "ndofs must be a positive multiple of 6");
    static const char* names[6] = {"Ex.gf", "Ey.gf", "Ez.gf", "Hx.gf", "Hy.gf", "Hz.gf"};
    const int64_t n = ndofs / 6;
    const std::string base(dir);
    const std::string fec = std::string(dim == 2 ? "L2_T1_2D_P" : "L2_T1_3D_P") + std::to_string(order);  // L2, GaussLobatto (type 1)
    for (int c = 0; c < 6; ++c) {
        const std::string path = base + "/" + names[c];
        FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return dgtd::api_fail(DGTD_E_ARG, "cannot write " + path);
        std::fprintf(f, "FiniteElementSpace\nFiniteElementCollection: %s\nVDim: 1\nOrdering: 0\n\n", fec.c_str());
        for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%.16g\n", x[c * n + i]);
        if (std::fclose(f) != 0) return dgtd::api_fail(DGTD_E_ARG, "write failed: " + path);
    }
    FILE* f = std::fopen((base + "/time.txt").c_str(), "w");
    if (!f) return dgtd::api_fail(DGTD_E_ARG, "cannot write " + base + "/time.txt");
    std::fprintf(f, "%.17g", t);
    std::fclose(f);
    return DGTD_OK;
}

}  // namespace

extern "C" int dgtd_save_fields(const char* dir, int32_t order, int64_t ndofs, const double* x, double t)
{
    return save_fields(dir, 3, order, ndofs, x, t);
}

extern "C" int dgtd_case_save_fields(const dgtd_case* c, const char* dir, int64_t ndofs, const double* x, double t)
{
    if (!c) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    return save_fields(dir, c->mesh.dim, c->info.order, ndofs, x, t);
}

extern "C" int dgtd_case_planewaves(const dgtd_case* c, dgtd_planewave* out)
{
    if (!c || (!out && !c->waves.empty())) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    for (size_t i = 0; i < c->waves.size(); ++i) out[i] = c->waves[i];
    return DGTD_OK;
}

extern "C" int dgtd_case_sources(const dgtd_case* c, dgtd_initial_field* out)
{
    if (!c || (!out && !c->sources.empty())) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    for (size_t i = 0; i < c->sources.size(); ++i) out[i] = c->sources[i];
    return DGTD_OK;
}

extern "C" int dgtd_case_field_probes(const dgtd_case* c, double* xyz, int32_t* steps, int32_t* saves,
                                      int32_t* component)
{
    if (!c) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    for (size_t i = 0; i < c->fprobe_comp.size(); ++i) {
        if (xyz)
            for (int d = 0; d < 3; ++d) xyz[3 * i + d] = c->fprobes[3 * i + d];
        if (steps) steps[i] = c->fprobe_steps[i];
        if (saves) saves[i] = c->fprobe_saves[i];
        if (component) component[i] = c->fprobe_comp[i];
    }
    return DGTD_OK;
}

extern "C" int dgtd_case_probes(const dgtd_case* c, double* xyz, int32_t* steps, int32_t* saves)
{
    if (!c || (!xyz && !c->probes.empty())) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    for (size_t i = 0; i < c->probes.size(); ++i) xyz[i] = c->probes[i];
    for (size_t i = 0; i < c->probe_steps.size(); ++i) {
        if (steps) steps[i] = c->probe_steps[i];
        if (saves) saves[i] = c->probe_saves[i];
    }
    return DGTD_OK;
}
