// C ABI of the operator (include/dgtd_hip.h): setup on the host in FP64,
// device-resident buffers, fused stage launches.
#include "dgtd_hip.h"
#include "global.hpp"
#include "tri.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "fields.hpp"
#include "kernels.hpp"
#include "trace.hpp"
#include "hex_dev.hpp"
#include "tfsf_host.hpp"
#include "mesh.hpp"
#include "packing.hpp"
#include "reftet.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// a request the implementation does not cover (DGTD_E_UNSUPPORTED)
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// LSERK4: Carpenter & Kennedy (1994), 5-stage 4th-order 2N-storage
// coefficients as published (also tabulated in Hesthaven & Warburton 2008).
const double kA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                      -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double kB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                      1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                      2277821191437.0 / 14882151754819.0};

// LSERK4 stage times t + c_s dt (Carpenter & Kennedy 1994).
const double kC[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                      2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

template <typename T>
void* upload(const std::vector<T>& h)
{
    void* d = nullptr;
    hip_check(hipMalloc(&d, sizeof(T) * h.size()), "hipMalloc");
    hip_check(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}

}  // namespace

namespace {
// Process-wide autotuning cache: the chosen candidate per (element shape,
// order, precision, stage mode, size class = floor(log2 K)).
struct TuneKey {
    int dim, order, precision, mode, size_class;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(dim, order, precision, mode, size_class) < std::tie(o.dim, o.order, o.precision, o.mode, o.size_class);
    }
};
// The fastest candidate overall, the fastest matrix-core candidate (-1:
// none) and every candidate's time in microseconds (0: did not fit).
struct TuneChoice {
    int best = 0, best_mfma = -1;
    double us[8] = {};
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tune;

int size_class(int64_t K)
{
    int c = 0;
    while (K > 1) {
        K >>= 1;
        ++c;
    }
    return c;
}

bool tune_cache_get(int dim, int order, int precision, int mode, int64_t K, TuneChoice& choice)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tune.find(TuneKey{dim, order, precision, mode, size_class(K)});
    if (it == g_tune.end()) return false;
    choice = it->second;
    return true;
}

void tune_cache_put(int dim, int order, int precision, int mode, int64_t K, TuneChoice choice)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[TuneKey{dim, order, precision, mode, size_class(K)}] = choice;
}
}  // namespace

struct dgtd_op {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = 64;
    int integrator = DGTD_LSERK4;
    double alpha = 1.0;
    dgtd::RefTet R;
    int dim = 3;       // 2: triangles (the 2D operator, RefTri T2) or quadrilaterals (quad, RefQuad Q2)
    dgtd::RefTri T2;
    bool quad = false;
    dgtd::RefQuad Q2;
    bool hex = false;  // hexahedra (dim 3, RefHex H3; the Global operator: assembled or matrix-free)
    dgtd::RefHex H3;
    int K = 0, Np = 0, Nfp = 0;
    size_t N = 0;
    int64_t ndofs = 0;
    double max_fscale = 0.0;
    std::vector<double> vx;
    std::vector<int32_t> EToV;
    int nv = 0;

    void* d_geo = nullptr;
    int* d_conn = nullptr;
    void* d_mats = nullptr;
    void* d_mats_mfma = nullptr;
    int* d_order = nullptr;
    int variant = DGTD_KERNEL_VALU;
    int grid_mfma = 0;                            // grid of the configuration in use
    int mfma_cfg = 0;                             // matrix-core tuning configuration in use
    bool morton = true;                           // matrix-core tasks in Morton order (else element order)
    int grid_cfg[dgtd::kMfmaConfigs] = {};        // grid per configuration, 0 = does not fit
    // autotuning record: candidates timed at creation (or taken from the cache)
    dgtd_tuning tuning{};
    // assembled operator (DGTD_KERNEL_CSR)
    int64_t* d_rowptr = nullptr;
    uint16_t* d_col = nullptr;
    int* d_etab = nullptr;
    void* d_val = nullptr;
    int64_t nnz = 0;
    int* d_tabs = nullptr;
    double* d_M = nullptr;
    double* d_jac = nullptr;
    double* d_eps = nullptr;
    double* d_mu = nullptr;
    double* d_partials = nullptr;
    double* d_out3 = nullptr;
    int* d_unstable = nullptr;  // per-step stability test (report_unstable), since set_state
    int nparts = 0;
    void* d_buf[4] = {nullptr, nullptr, nullptr, nullptr};
    int cur = 0;  // LSERK: buffer holding the state
    double t = 0.0;
    int grid = 1, nblk = 1;

    // TF/SF injection: elements owning TF/SF faces and the plane waves
    int ntfsf = 0, nwaves = 0;
    int* d_tfsf_elem = nullptr;
    int* d_tfsf_sign = nullptr;
    double* d_tfsf_xyz = nullptr;
    double* d_tfsf_geo = nullptr;
    double* d_lift = nullptr;
    dgtd::TfsfWave* d_waves = nullptr;

    int nprobes = 0;
    int* d_probe_elem = nullptr;
    double* d_probe_w = nullptr;
    double* d_probe_out = nullptr;

    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // One time step replayed from a hipGraph (per LSERK buffer parity),
    // captured for the current dt; not used while plane waves are set (the
    // TF/SF correction takes the stage times as arguments).  Off by default:
    // measured on the 8^3 and 32^3 cavities, replaying the graph is no faster
    // than the five asynchronous launches (the stage kernels dominate).
    bool graphs = false;
    hipGraphExec_t step_graph[2] = {nullptr, nullptr};
    double graph_dt = 0.0;
    double last_us = 0.0;
    int last_launches = 0;

    size_t real_size() const { return precision == 64 ? sizeof(double) : sizeof(float); }
    void* state() const { return integrator == DGTD_LSERK4 ? d_buf[cur] : d_buf[0]; }

    ~dgtd_op()
    {
        (void)hipSetDevice(device);
        for (void* p : {d_geo, (void*)d_conn, d_mats, d_mats_mfma, (void*)d_order, (void*)d_tabs, (void*)d_M, (void*)d_jac,
                        (void*)d_eps, (void*)d_mu, (void*)d_partials, (void*)d_out3, (void*)d_unstable, (void*)d_probe_elem,
                        (void*)d_tfsf_elem, (void*)d_tfsf_sign, (void*)d_tfsf_xyz, (void*)d_tfsf_geo,
                        (void*)d_lift, (void*)d_waves, (void*)d_rowptr, (void*)d_col, (void*)d_etab, d_val,
                        (void*)d_probe_w, (void*)d_probe_out, d_buf[0],
                        d_buf[1], d_buf[2], d_buf[3]})
            if (p) (void)hipFree(p);
        drop_graphs();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    dgtd::StageArgs<T> base_args() const
    {
        dgtd::StageArgs<T> a{};
        a.K = K;
        a.nblk = nblk;
        a.alpha = (T)alpha;
        a.geo = static_cast<const T*>(d_geo);
        a.conn = d_conn;
        a.mats = static_cast<const T*>(d_mats);
        a.tabs = d_tabs;
        a.rowptr = d_rowptr;
        a.col16 = d_col;
        a.etab = d_etab;
        a.np = Np;
        a.val = static_cast<const T*>(d_val);
        a.nrows = ndofs;
        a.nnz = nnz;
        return a;
    }

    template <typename T>
    hipError_t launch_on(const dgtd::StageArgs<T>& a0, int mode, int var, hipStream_t s) const
    {
        if (var == DGTD_KERNEL_CSR) return dgtd::launch_csr_stage<T>(a0, mode, s);
        if (hex) return dgtd::launch_hex_stage<T>(H3.p, a0, mode, s);
        if (var == DGTD_KERNEL_MFMA) {
            dgtd::StageArgs<T> a = a0;
            a.mats = static_cast<const T*>(d_mats_mfma);
            a.order = morton ? d_order : nullptr;
            if (dim == 2) return dgtd::launch_tri_stage_mfma<T>(T2.p, a, mode, grid_cfg[mfma_cfg], s, mfma_cfg);
            return dgtd::launch_stage_mfma<T>(R.p, a, mode, grid_cfg[mfma_cfg], s, mfma_cfg);
        }
        if (dim == 2) return dgtd::launch_tri_stage<T>(T2.p, a0, mode, grid, s);
        return dgtd::launch_stage<T>(R.p, a0, mode, grid, s);
    }

    template <typename T>
    void launch(const dgtd::StageArgs<T>& a, int mode)
    {
        dgtd::TraceRange tr("dgtd stage launch");
        hip_check(launch_on<T>(a, mode, variant, stream), "stage kernel launch");
    }

    // Candidate stage kernels of the autotuner: the per-node variant, then
    // the matrix-core variant in each tuning configuration that fits (waves
    // per block, gather depth, residual prefetch), with its tasks in Morton
    // order or in element order.
    static constexpr int kCandidates = 1 + 2 * dgtd::kMfmaConfigs;
    static_assert(kCandidates <= (int)(sizeof(dgtd_tuning::candidate_us) / sizeof(double)),
                  "dgtd_tuning holds every candidate's time");
    bool candidate(int c, int& var, int& cfg, bool& mort) const
    {
        var = c == 0 ? DGTD_KERNEL_VALU : DGTD_KERNEL_MFMA;
        cfg = c == 0 ? 0 : (c - 1) / 2;
        mort = c == 0 || (c - 1) % 2 == 0;
        return c == 0 || (grid_cfg[cfg] > 0 && (mort || d_order));
    }

    void apply_candidate(int c)
    {
        int var, cfg;
        bool mort;
        candidate(c, var, cfg, mort);
        variant = var;
        if (var == DGTD_KERNEL_MFMA) {
            mfma_cfg = cfg;
            morton = mort;
            grid_mfma = grid_cfg[cfg];
        }
        const dgtd::MfmaInfo mi = precision == 64 ? dgtd::mfma_info<double>(dim == 2 ? T2.p : R.p, dim, mfma_cfg)
                                                  : dgtd::mfma_info<float>(dim == 2 ? T2.p : R.p, dim, mfma_cfg);
        tuning.variant = variant;
        tuning.config = mfma_cfg;
        tuning.waves_per_block = variant == DGTD_KERNEL_MFMA ? mi.waves : 4;
        tuning.gather_depth = variant == DGTD_KERNEL_MFMA ? mi.gather_depth : 0;
        tuning.early_residual = variant == DGTD_KERNEL_MFMA && mi.early_residual;
        tuning.task_order = variant == DGTD_KERNEL_MFMA && morton ? 1 : 0;
    }

    // Times every candidate on the zeroed state buffers (the fused LSERK
    // stage for LSERK4 operators -- the launch that steps run; with
    // ca = cb = dt = 0 the buffers stay zero -- the operator application
    // otherwise) and keeps the fastest.  The choice is cached per (element
    // shape, order, precision, stage mode, size class) for the process, so
    // later operators of the same kind skip the timing.
    template <typename T>
    void autotune()
    {
        dgtd::TraceRange tr("dgtd autotune");
        const int order = dim == 2 ? T2.p : R.p;
        const int mode_key = integrator == DGTD_LSERK4 ? dgtd::kModeLserk : dgtd::kModeMult;
        TuneChoice cached;
        if (tune_cache_get(dim, order, precision, mode_key, K, cached)) {
            apply_candidate(cached.best);
            tuning.from_cache = 1;
            tuning.ncandidates = kCandidates;
            for (int c = 0; c < kCandidates; ++c) tuning.candidate_us[c] = cached.us[c];
            return;
        }
        auto a = base_args<T>();
        a.qin = static_cast<const T*>(d_buf[0]);
        a.out = static_cast<T*>(d_buf[1]);
        int mode = dgtd::kModeMult;
        if (integrator == DGTD_LSERK4) {
            mode = dgtd::kModeLserk;
            a.res = static_cast<T*>(d_buf[2]);
            a.ca = T(0);
            a.cb = T(0);
            a.dt = T(0);
        }
        // Two rounds over the candidates: three launches each, then enough
        // launches for about 1 ms of kernel time (3..64); a candidate's time
        // is the lower of its two per-launch means (interference only adds).
        auto time_launches = [&](int var, int reps) {
            hip_check(launch_on<T>(a, mode, var, stream), "autotune launch");  // warm-up
            hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
            for (int r = 0; r < reps; ++r) hip_check(launch_on<T>(a, mode, var, stream), "autotune launch");
            hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
            hip_check(hipEventSynchronize(ev1), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
            return 1000.0 * ms / reps;
        };
        for (int c = 0; c < kCandidates; ++c) tuning.candidate_us[c] = 0.0;
        for (int round = 0; round < 2; ++round)
            for (int c = 0; c < kCandidates; ++c) {
                int var, cfg;
                bool mort;
                if (!candidate(c, var, cfg, mort)) continue;
                apply_candidate(c);
                double& us = tuning.candidate_us[c];
                const int reps = round == 0 ? 3 : std::max(3, std::min(64, (int)std::ceil(1000.0 / std::max(us, 1.0))));
                const double t = time_launches(var, reps);
                us = round == 0 ? t : std::min(us, t);
            }
        double best = 1e30, best_m = 1e30;
        int best_c = 0, best_mc = -1;
        for (int c = 0; c < kCandidates; ++c) {
            int var, cfg;
            bool mort;
            if (!candidate(c, var, cfg, mort)) continue;
            const double us = tuning.candidate_us[c];
            if (us < best) { best = us; best_c = c; }
            if (var == DGTD_KERNEL_MFMA && us < best_m) { best_m = us; best_mc = c; }
        }
        apply_candidate(best_c);
        tuning.from_cache = 0;
        tuning.ncandidates = kCandidates;
        TuneChoice made{best_c, best_mc, {}};
        for (int c = 0; c < kCandidates; ++c) made.us[c] = tuning.candidate_us[c];
        tune_cache_put(dim, order, precision, mode_key, K, made);
    }

    // Adds the TF/SF right-hand side at time ts: c_out * S to out and
    // c_res * S to res (res may be null); no-op without TF/SF faces or waves.
    template <typename T>
    void tfsf(double ts, T* out, double c_out, T* res, double c_res, hipStream_t s)
    {
        if (ntfsf == 0 || nwaves == 0) return;
        dgtd::TfsfArgs<T> a{};
        a.dim = hex ? 5 : quad ? 4 : dim;
        a.nelem = ntfsf;
        a.elem = d_tfsf_elem;
        a.sign = d_tfsf_sign;
        a.xyz = d_tfsf_xyz;
        a.geo = d_tfsf_geo;
        a.lift = d_lift;
        a.waves = d_waves;
        a.nwaves = nwaves;
        a.Np = Np;
        a.Nfp = Nfp;
        a.N = N;
        a.alpha = alpha;
        a.t = ts;
        a.out = out;
        a.c_out = c_out;
        a.res = res;
        a.c_res = c_res;
        hip_check(dgtd::launch_tfsf<T>(a, s), "TF/SF kernel launch");
    }

    // One step from time t (the stage kernels take the TF/SF right-hand side
    // at the stage times: LSERK t + c_s dt, RK4 t, t + dt/2, t + dt/2, t + dt).
    template <typename T>
    int one_step(double dt)
    {
        dgtd::TraceRange tr("dgtd step");
        auto a = base_args<T>();
        a.dt = (T)dt;
        if (integrator == DGTD_LSERK4) {
            for (int s = 0; s < 5; ++s) {
                a.qin = static_cast<const T*>(d_buf[cur]);
                a.out = static_cast<T*>(d_buf[cur ^ 1]);
                a.res = static_cast<T*>(d_buf[2]);
                a.ca = (T)kA[s];
                a.cb = (T)kB[s];
                a.unstable = s == 4 ? d_unstable : nullptr;  // the step's new state
                launch<T>(a, dgtd::kModeLserk);
                // res' = res + dt S, q' = q + b res'  =>  q' += b dt S
                tfsf<T>(t + kC[s] * dt, a.out, kB[s] * dt, a.res, dt, stream);
                cur ^= 1;
            }
            return 5;
        }
        // mfem::RK4Solver stage order: buffers x = 0, y ping-pong 1/2, z = 3
        T* x = static_cast<T*>(d_buf[0]);
        T* y1 = static_cast<T*>(d_buf[1]);
        T* y2 = static_cast<T*>(d_buf[2]);
        T* z = static_cast<T*>(d_buf[3]);
        a.res = z;
        a.x = x;
        a.qin = x; a.out = y1; a.ca = (T)(dt / 2); a.cb = (T)(dt / 6);
        launch<T>(a, dgtd::kModeRk1);
        tfsf<T>(t, y1, dt / 2, z, dt / 6, stream);
        a.qin = y1; a.out = y2; a.ca = (T)(dt / 2); a.cb = (T)(dt / 3);
        launch<T>(a, dgtd::kModeRk2);
        tfsf<T>(t + dt / 2, y2, dt / 2, z, dt / 3, stream);
        a.qin = y2; a.out = y1; a.ca = (T)dt; a.cb = (T)(dt / 3);
        launch<T>(a, dgtd::kModeRk3);
        tfsf<T>(t + dt / 2, y1, dt, z, dt / 3, stream);
        a.qin = y1; a.out = x; a.ca = (T)(dt / 6);
        a.unstable = d_unstable;  // the step's new state
        launch<T>(a, dgtd::kModeRk4);
        tfsf<T>(t + dt, x, dt / 6, nullptr, 0.0, stream);
        return 4;
    }

    int step_eager(double dt) { return precision == 64 ? one_step<double>(dt) : one_step<float>(dt); }

    void drop_graphs()
    {
        for (auto& g : step_graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    // A step: replayed from the captured graph of this dt and buffer parity
    // (captured on first use), or launched directly.
    int step(double dt)
    {
        if (!graphs || (ntfsf > 0 && nwaves > 0)) return step_eager(dt);
        if (dt != graph_dt) {
            drop_graphs();
            graph_dt = dt;
        }
        const int par = integrator == DGTD_LSERK4 ? cur : 0;
        int launches = integrator == DGTD_LSERK4 ? 5 : 4;
        if (!step_graph[par]) {
            const int cur0 = cur;
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                launches = step_eager(dt);
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                cur = cur0;
                throw;
            }
            hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&step_graph[par], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(e, "hipGraphInstantiate");
            cur = cur0;  // captured, not run
        }
        hip_check(hipGraphLaunch(step_graph[par], stream), "hipGraphLaunch");
        if (integrator == DGTD_LSERK4) cur ^= 1;  // five buffer flips
        return launches;
    }

    template <typename T>
    void mult_dev(const void* in, void* out, hipStream_t s, double tm)
    {
        auto a = base_args<T>();
        a.qin = static_cast<const T*>(in);
        a.out = static_cast<T*>(out);
        hip_check(launch_on<T>(a, dgtd::kModeMult, variant, s), "stage kernel launch");
        tfsf<T>(tm, static_cast<T*>(out), 1.0, nullptr, 0.0, s);
    }

    void to_device(const double* h, void* d)
    {
        if (precision == 64) {
            hip_check(hipMemcpyAsync(d, h, sizeof(double) * ndofs, hipMemcpyHostToDevice, stream), "H2D");
        } else {
            std::vector<float> tmp(h, h + ndofs);
            hip_check(hipMemcpyAsync(d, tmp.data(), sizeof(float) * ndofs, hipMemcpyHostToDevice, stream),
                      "H2D");
            hip_check(hipStreamSynchronize(stream), "sync");
        }
    }

    void to_host(const void* d, double* h)
    {
        if (precision == 64) {
            hip_check(hipMemcpyAsync(h, d, sizeof(double) * ndofs, hipMemcpyDeviceToHost, stream), "D2H");
            hip_check(hipStreamSynchronize(stream), "sync");
        } else {
            std::vector<float> tmp(ndofs);
            hip_check(hipMemcpyAsync(tmp.data(), d, sizeof(float) * ndofs, hipMemcpyDeviceToHost, stream),
                      "D2H");
            hip_check(hipStreamSynchronize(stream), "sync");
            std::copy(tmp.begin(), tmp.end(), h);
        }
    }
};

namespace {

// Processing order of the wave-tasks of the MFMA variant: Morton
// (Z-order) of the task centroids, so that tasks in flight together (and the
// contiguous ranges the kernel gives each XCD) are compact in space and their
// neighbour traces are reused from L2.
std::vector<int> morton_task_order(const dgtd_mesh_desc* mesh)
{
    const int TE = dgtd::kMfmaTaskElems;
    const int nvert = mesh->dim == 2 ? 3 : 4;
    const int K = mesh->K, nt = (K + TE - 1) / TE;
    std::vector<double> c(3 * (size_t)nt, 0.0);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int t = 0; t < nt; ++t) {
        const int e0 = TE * t, e1 = std::min(K, e0 + TE);
        for (int e = e0; e < e1; ++e)
            for (int v = 0; v < nvert; ++v)
                for (int d = 0; d < 3; ++d)
                    c[3 * (size_t)t + d] += mesh->vx[3 * (size_t)mesh->EToV[nvert * (size_t)e + v] + d];
        for (int d = 0; d < 3; ++d) {
            c[3 * (size_t)t + d] /= (double)nvert * (e1 - e0);
            lo[d] = std::min(lo[d], c[3 * (size_t)t + d]);
            hi[d] = std::max(hi[d], c[3 * (size_t)t + d]);
        }
    }
    auto spread = [](uint64_t v) {  // 21 bits -> every third bit
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    std::vector<std::pair<uint64_t, int>> key(nt);
    for (int t = 0; t < nt; ++t) {
        uint64_t code = 0;
        for (int d = 0; d < 3; ++d) {
            const double w = hi[d] > lo[d] ? (c[3 * (size_t)t + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
            code |= spread((uint64_t)(w * ((1 << 20) - 1))) << d;
        }
        key[t] = {code, t};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(nt);
    for (int t = 0; t < nt; ++t) order[t] = key[t].second;
    return order;
}


}  // namespace

namespace {

dgtd::TriMeshInput tri_input(const dgtd_mesh_desc* mesh)
{
    dgtd::TriMeshInput in;
    in.K = mesh->K;
    in.nv = mesh->nv;
    in.vx = mesh->vx;
    in.EToV = mesh->EToV;
    in.face_bc = mesh->face_bc;
    in.eps_r = mesh->eps_r;
    in.mu_r = mesh->mu_r;
    return in;
}

dgtd::TriMeshInput tri_input(const dgtd_op* op)
{
    dgtd::TriMeshInput in;
    in.K = op->K;
    in.nv = op->nv;
    in.vx = op->vx.data();
    in.EToV = op->EToV.data();
    return in;
}

// Set-up of a triangle-mesh operator (mesh->dim == 2): reference triangle,
// geometry and connectivity records, the per-node stage kernel only.
void upload_csr(dgtd_op* op, const dgtd::Csr& A)
{
    op->nnz = (int64_t)A.val.size();
    const dgtd::CsrLocal16 L = dgtd::csr_local16(A, (int64_t)op->N, op->Np);
    op->d_rowptr = static_cast<int64_t*>(upload(A.rowptr));
    op->d_col = static_cast<uint16_t*>(upload(L.col));
    op->d_etab = static_cast<int*>(upload(L.etab));
    op->d_val = op->precision == 64 ? upload(A.val) : upload(dgtd::cast_vec<float>(A.val));
}

// Persistent grid of the matrix-core stage: enough blocks for the tasks, at
// most the resident ones, in whole XCD groups (the kernel splits the tasks
// per XCD).
int mfma_grid(int64_t K, int waves, int resident)
{
    const int64_t ntasks = (K + dgtd::kMfmaTaskElems - 1) / dgtd::kMfmaTaskElems;
    int g = (int)std::min<int64_t>((ntasks + waves - 1) / waves, resident);
    if (g >= 8) g -= g % 8;
    return std::max(1, g);
}

// The stage variant: assembled (CSR), autotuned, or as requested (matrix
// cores in configuration 0 with Morton-ordered tasks).
template <typename CsrSetup>
void select_variant(dgtd_op* op, int kernel, CsrSetup csr_setup)
{
    if (kernel == DGTD_KERNEL_CSR) {
        csr_setup();
        op->variant = DGTD_KERNEL_CSR;
        op->tuning.variant = DGTD_KERNEL_CSR;
    } else if (kernel == DGTD_KERNEL_AUTO) {
        if (op->precision == 64) op->autotune<double>();
        else op->autotune<float>();
    } else {
        op->apply_candidate(kernel == DGTD_KERNEL_MFMA ? 1 : 0);
    }
}

void create_tri(dgtd_op* op, const dgtd_mesh_desc* mesh, int kernel)
{
    op->dim = 2;
    op->T2 = dgtd::build_reftri(mesh->order);
    const dgtd::RefTri& R = op->T2;
    op->K = mesh->K;
    op->Np = R.Np;
    op->Nfp = R.Nfp;
    op->N = (size_t)op->Np * mesh->K;
    op->ndofs = 6 * (int64_t)op->N;
    // (alpha only guards SMA faces against the centred flux, which the
    // assembled operator handles: its SMA faces are weighted as at alpha = 1)
    const dgtd::TriMeshData md =
        dgtd::build_tri_mesh(tri_input(mesh), R, kernel == DGTD_KERNEL_CSR ? 1.0 : op->alpha);
    op->max_fscale = md.max_fscale;
    op->nv = mesh->nv;
    op->vx.assign(mesh->vx, mesh->vx + 3 * (size_t)mesh->nv);
    op->EToV.assign(mesh->EToV, mesh->EToV + 3 * (size_t)mesh->K);
    hip_check(hipSetDevice(op->device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&op->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&op->ev1), "hipEventCreate");
    std::vector<int> tabs(R.Fmask.begin(), R.Fmask.end());
    tabs.insert(tabs.end(), R.edge_perm.begin(), R.edge_perm.end());
    op->d_tabs = static_cast<int*>(upload(tabs));
    op->d_conn = static_cast<int*>(upload(md.conn));
    // Dr^T, Ds^T ([j][i]), then LIFT^T ([m][i])
    const int Np = R.Np, NF = 3 * R.Nfp;
    std::vector<double> m((size_t)2 * Np * Np + (size_t)NF * Np);
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
            m[(size_t)j * Np + i] = R.Dr[(size_t)i * Np + j];
            m[(size_t)Np * Np + (size_t)j * Np + i] = R.Ds[(size_t)i * Np + j];
        }
    for (int i = 0; i < Np; ++i)
        for (int k = 0; k < NF; ++k) m[2 * (size_t)Np * Np + (size_t)k * Np + i] = R.LIFT[(size_t)i * NF + k];
    if (op->precision == 64) {
        op->d_geo = upload(md.geo);
        op->d_mats = upload(m);
    } else {
        op->d_geo = upload(dgtd::cast_vec<float>(md.geo));
        op->d_mats = upload(dgtd::cast_vec<float>(m));
    }
    op->d_M = static_cast<double*>(upload(R.M));
    op->d_jac = static_cast<double*>(upload(md.jac));
    op->d_eps = static_cast<double*>(upload(md.eps));
    op->d_mu = static_cast<double*>(upload(md.mu));
    const int te = dgtd::norms_tile(op->Np);
    op->nparts = std::max(1, std::min(dgtd::kNormsMaxParts, (op->K + te - 1) / te));
    hip_check(hipMalloc(&op->d_partials, sizeof(double) * 3 * op->nparts), "hipMalloc");
    double* scratch_dev = nullptr;  // resolves the allocation-free entry point's scratch on this device
    hip_check(dgtd::norms_scratch(op->device, &scratch_dev), "norms scratch");
    hip_check(hipMalloc(&op->d_out3, sizeof(double) * 3), "hipMalloc");
    hip_check(hipMalloc(&op->d_unstable, sizeof(int)), "hipMalloc");
    hip_check(hipMemsetAsync(op->d_unstable, 0, sizeof(int), op->stream), "memset");
    const int nbuf = op->integrator == DGTD_LSERK4 ? 3 : 4;
    for (int b = 0; b < nbuf; ++b) {
        hip_check(hipMalloc(&op->d_buf[b], op->real_size() * op->ndofs), "hipMalloc state");
        hip_check(hipMemsetAsync(op->d_buf[b], 0, op->real_size() * op->ndofs, op->stream), "memset");
    }
    if (!md.tfsf.empty()) {  // triangles owning TF/SF edges: edge signs, edge-node positions, records
        const dgtd::TfsfRecords r = dgtd::tfsf_records(mesh, R, md.tfsf, md.geo);
        op->ntfsf = (int)r.elems.size();
        op->d_tfsf_elem = static_cast<int*>(upload(r.elems));
        op->d_tfsf_sign = static_cast<int*>(upload(r.signs));
        op->d_tfsf_xyz = static_cast<double*>(upload(r.xyz));
        op->d_tfsf_geo = static_cast<double*>(upload(r.geo));
        op->d_lift = static_cast<double*>(upload(R.LIFT));
    }
    hip_check(hipStreamSynchronize(op->stream), "sync");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, op->device), "hipGetDeviceProperties");
    const int epb = dgtd::tri_elements_per_block(R.p);
    op->nblk = (op->K + epb - 1) / epb;
    op->grid = std::max(1, std::min(op->nblk, prop.multiProcessorCount * 4));
    op->variant = DGTD_KERNEL_VALU;
    if (op->precision == 64) dgtd::tri_prepare<double>(R.p);
    else dgtd::tri_prepare<float>(R.p);
    // matrix-core variant: 8-triangle wave tasks (Morton or element order), per configuration
    bool any_mfma = false;
    for (int c = 0; c < dgtd::kMfmaConfigs; ++c) {
        const int mbpc = op->precision == 64 ? dgtd::tri_mfma_blocks_per_cu<double>(R.p, c)
                                             : dgtd::tri_mfma_blocks_per_cu<float>(R.p, c);
        if (mbpc <= 0 || !dgtd::mfma_fits_buffers(6 * (int64_t)op->Np * op->K, op->precision / 8)) continue;
        const int mwaves = op->precision == 64 ? dgtd::tri_mfma_waves_per_block<double>(R.p, c)
                                               : dgtd::tri_mfma_waves_per_block<float>(R.p, c);
        op->grid_cfg[c] = mfma_grid(op->K, mwaves, prop.multiProcessorCount * mbpc);
        any_mfma = true;
    }
    if (any_mfma) {
        op->d_mats_mfma = op->precision == 64 ? upload(dgtd::pack_mats_mfma<double>(R))
                                              : upload(dgtd::pack_mats_mfma<float>(R));
        op->d_order = static_cast<int*>(upload(morton_task_order(mesh)));
    }
    if (kernel == DGTD_KERNEL_MFMA && op->grid_cfg[0] <= 0)
        throw std::runtime_error("the MFMA kernel does not fit this order and precision, or a state buffer reaches 4 GiB");
    select_variant(op, kernel, [&] { upload_csr(op, dgtd::assemble_global_tri(md, R, mesh->sigma, op->alpha)); });
}

// Set-up of a quadrilateral-mesh operator (dim = 2, nvert = 4): the
// assembled operator only.
void create_quad(dgtd_op* op, const dgtd_mesh_desc* mesh, int kernel)
{
    if (kernel != DGTD_KERNEL_CSR && kernel != DGTD_KERNEL_AUTO)
        throw UnsupportedError("quadrilaterals run on the assembled operator (DGTD_KERNEL_CSR)");
    op->dim = 2;
    op->quad = true;
    op->Q2 = dgtd::build_refquad(mesh->order);
    const dgtd::RefQuad& R = op->Q2;
    op->K = mesh->K;
    op->Np = R.Np;
    op->Nfp = R.Nfp;
    op->N = (size_t)op->Np * mesh->K;
    op->ndofs = 6 * (int64_t)op->N;
    const dgtd::TriMeshData md = dgtd::build_quad_mesh(tri_input(mesh), R, 1.0);
    op->max_fscale = md.max_fscale;
    op->nv = mesh->nv;
    op->vx.assign(mesh->vx, mesh->vx + 3 * (size_t)mesh->nv);
    op->EToV.assign(mesh->EToV, mesh->EToV + 4 * (size_t)mesh->K);
    hip_check(hipSetDevice(op->device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&op->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&op->ev1), "hipEventCreate");
    op->d_M = static_cast<double*>(upload(R.M));
    op->d_jac = static_cast<double*>(upload(md.jac));
    op->d_eps = static_cast<double*>(upload(md.eps));
    op->d_mu = static_cast<double*>(upload(md.mu));
    const int te = dgtd::norms_tile(op->Np);
    op->nparts = std::max(1, std::min(dgtd::kNormsMaxParts, (op->K + te - 1) / te));
    hip_check(hipMalloc(&op->d_partials, sizeof(double) * 3 * op->nparts), "hipMalloc");
    double* scratch_dev = nullptr;  // resolves the allocation-free entry point's scratch on this device
    hip_check(dgtd::norms_scratch(op->device, &scratch_dev), "norms scratch");
    hip_check(hipMalloc(&op->d_out3, sizeof(double) * 3), "hipMalloc");
    hip_check(hipMalloc(&op->d_unstable, sizeof(int)), "hipMalloc");
    hip_check(hipMemsetAsync(op->d_unstable, 0, sizeof(int), op->stream), "memset");
    const int nbuf = op->integrator == DGTD_LSERK4 ? 3 : 4;
    for (int b = 0; b < nbuf; ++b) {
        hip_check(hipMalloc(&op->d_buf[b], op->real_size() * op->ndofs), "hipMalloc state");
        hip_check(hipMemsetAsync(op->d_buf[b], 0, op->real_size() * op->ndofs, op->stream), "memset");
    }
    if (!md.tfsf.empty()) {  // quadrilaterals owning TF/SF edges: edge signs, edge-node positions, records
        const dgtd::TfsfRecords r = dgtd::tfsf_records(mesh, R, md.tfsf, md.geo);
        op->ntfsf = (int)r.elems.size();
        op->d_tfsf_elem = static_cast<int*>(upload(r.elems));
        op->d_tfsf_sign = static_cast<int*>(upload(r.signs));
        op->d_tfsf_xyz = static_cast<double*>(upload(r.xyz));
        op->d_tfsf_geo = static_cast<double*>(upload(r.geo));
        op->d_lift = static_cast<double*>(upload(R.LIFT));
    }
    hip_check(hipStreamSynchronize(op->stream), "sync");
    select_variant(op, DGTD_KERNEL_CSR, [&] { upload_csr(op, dgtd::assemble_global_quad(md, R, mesh->sigma, op->alpha)); });
}

// Solver::estimateTimeStep (Solver.cpp:302-353).  On tetrahedra under
// hesthaven: cfl / (max fscale * p^2).  Otherwise calcMeshTimeStep
// (:290-300): cfl * 0.75 * (2/3) * rmin * min_e vol_e / (surface_e / 2)
// (getTimeStepScale, :207-252: areas and edge lengths in 2D, volumes and
// face areas in 3D; getJacobiGQ_RMin, :254-263: the first Gauss-Lobatto
// spacing of order p on [-1, 1]; speed of light 1), halved when the mesh holds
// quadrilaterals, divided by 0.8 on tetrahedra under the Global operator.
int estimate_time_step(int dim, bool quad, bool hex, int order, int K, const double* vx, const int32_t* EToV,
                       double max_fscale, int evolution_operator, double cfl, double* dt)
{
    if (evolution_operator < DGTD_OPERATOR_MAXWELL || evolution_operator > DGTD_OPERATOR_HESTHAVEN)
        return fail(DGTD_E_ARG, "evolution_operator must be DGTD_OPERATOR_MAXWELL, _GLOBAL or _HESTHAVEN");
    if (!(cfl > 0.0)) return fail(DGTD_E_ARG, "cfl must be > 0");
    if (dim == 3 && !hex && evolution_operator == DGTD_OPERATOR_HESTHAVEN) {
        *dt = cfl / (max_fscale * order * order);
        return DGTD_OK;
    }
    const int nvert = hex ? 8 : dim == 3 || quad ? 4 : 3;
    auto P = [&](int e, int j) { return vx + 3 * (size_t)EToV[(size_t)nvert * e + j]; };
    double scale = INFINITY;
    for (int e = 0; e < K; ++e) {
        double vol = 0.0, surf = 0.0;
        if (dim == 2) {
            for (int j = 0; j < nvert; ++j) {  // shoelace area, edge lengths
                const double *a = P(e, j), *b = P(e, (j + 1) % nvert);
                vol += 0.5 * (a[0] * b[1] - b[0] * a[1]);
                surf += std::hypot(b[0] - a[0], b[1] - a[1]);
            }
            vol = std::fabs(vol);
        } else if (hex) {  // parallelepiped: |det| of the edge vectors, parallelogram faces
            double u[3][3];
            const int axis_vertex[3] = {1, 3, 4};
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) u[k][d] = P(e, axis_vertex[k])[d] - P(e, 0)[d];
            vol = std::fabs(u[0][0] * (u[1][1] * u[2][2] - u[1][2] * u[2][1]) -
                            u[0][1] * (u[1][0] * u[2][2] - u[1][2] * u[2][0]) +
                            u[0][2] * (u[1][0] * u[2][1] - u[1][1] * u[2][0]));
            for (const auto& f : dgtd::kHexFaces) {
                double a[3], b[3];
                for (int d = 0; d < 3; ++d) {
                    a[d] = P(e, f[1])[d] - P(e, f[0])[d];
                    b[d] = P(e, f[3])[d] - P(e, f[0])[d];
                }
                const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2],
                             cz = a[0] * b[1] - a[1] * b[0];
                surf += std::sqrt(cx * cx + cy * cy + cz * cz);
            }
        } else {
            static const int kFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
            double u[3][3];
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) u[k][d] = P(e, k + 1)[d] - P(e, 0)[d];
            vol = std::fabs(u[0][0] * (u[1][1] * u[2][2] - u[1][2] * u[2][1]) -
                            u[0][1] * (u[1][0] * u[2][2] - u[1][2] * u[2][0]) +
                            u[0][2] * (u[1][0] * u[2][1] - u[1][1] * u[2][0])) / 6.0;
            for (const auto& f : kFace) {
                double a[3], b[3];
                for (int d = 0; d < 3; ++d) {
                    a[d] = P(e, f[1])[d] - P(e, f[0])[d];
                    b[d] = P(e, f[2])[d] - P(e, f[0])[d];
                }
                const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2],
                             cz = a[0] * b[1] - a[1] * b[0];
                surf += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            }
        }
        scale = std::min(scale, vol / (surf / 2.0));
    }
    const std::vector<double> g = dgtd::gll01(order);  // on [0, 1]
    double base = scale * (2.0 * (g[1] - g[0])) * 2.0 / 3.0 * 0.75;
    if (quad) base /= 2.0;  // the reference's divisors for meshes holding quadrilaterals
    if (hex) base /= 6.0;   // or hexahedra (Solver.cpp:324-328)
    base *= cfl;
    if (dim == 3 && evolution_operator == DGTD_OPERATOR_GLOBAL) base /= 0.8;
    *dt = base;
    return DGTD_OK;
}

dgtd::MeshInput mesh_input(const dgtd_mesh_desc* mesh)
{
    dgtd::MeshInput in;
    in.K = mesh->K;
    in.nv = mesh->nv;
    in.vx = mesh->vx;
    in.EToV = mesh->EToV;
    in.face_bc = mesh->face_bc;
    in.eps_r = mesh->eps_r;
    in.mu_r = mesh->mu_r;
    return in;
}

// Set-up of a hexahedral-mesh operator (dim 3, nvert = 8): the assembled
// operator only.
// Hexahedra run the Global operator's flux, either assembled (DGTD_KERNEL_CSR)
// or matrix-free (DGTD_KERNEL_VALU, which DGTD_KERNEL_AUTO selects:
// hex_stage.hip reads no matrix, and the set-up assembles none).
void create_hex(dgtd_op* op, const dgtd_mesh_desc* mesh, int kernel)
{
    if (kernel != DGTD_KERNEL_CSR && kernel != DGTD_KERNEL_AUTO && kernel != DGTD_KERNEL_VALU)
        throw UnsupportedError("hexahedra run the Global operator: assembled (DGTD_KERNEL_CSR) or matrix-free "
                               "(DGTD_KERNEL_VALU), not on the matrix cores");
    op->dim = 3;
    op->hex = true;
    op->H3 = dgtd::build_refhex(mesh->order);
    const dgtd::RefHex& R = op->H3;
    op->K = mesh->K;
    op->Np = R.Np;
    op->Nfp = R.Nfp;
    op->N = (size_t)op->Np * mesh->K;
    op->ndofs = 6 * (int64_t)op->N;
    const dgtd::HexMeshData md = dgtd::build_hex_mesh(mesh_input(mesh), R, 1.0);
    op->max_fscale = md.max_fscale;
    op->nv = mesh->nv;
    op->vx.assign(mesh->vx, mesh->vx + 3 * (size_t)mesh->nv);
    op->EToV.assign(mesh->EToV, mesh->EToV + 8 * (size_t)mesh->K);
    hip_check(hipSetDevice(op->device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&op->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&op->ev1), "hipEventCreate");
    op->d_M = static_cast<double*>(upload(R.M));
    op->d_jac = static_cast<double*>(upload(md.jac));
    op->d_eps = static_cast<double*>(upload(md.eps));
    op->d_mu = static_cast<double*>(upload(md.mu));
    const int te = dgtd::norms_tile(op->Np);
    op->nparts = std::max(1, std::min(dgtd::kNormsMaxParts, (op->K + te - 1) / te));
    hip_check(hipMalloc(&op->d_partials, sizeof(double) * 3 * op->nparts), "hipMalloc");
    double* scratch_dev = nullptr;  // resolves the allocation-free entry point's scratch on this device
    hip_check(dgtd::norms_scratch(op->device, &scratch_dev), "norms scratch");
    hip_check(hipMalloc(&op->d_out3, sizeof(double) * 3), "hipMalloc");
    hip_check(hipMalloc(&op->d_unstable, sizeof(int)), "hipMalloc");
    hip_check(hipMemsetAsync(op->d_unstable, 0, sizeof(int), op->stream), "memset");
    const int nbuf = op->integrator == DGTD_LSERK4 ? 3 : 4;
    for (int b = 0; b < nbuf; ++b) {
        hip_check(hipMalloc(&op->d_buf[b], op->real_size() * op->ndofs), "hipMalloc state");
        hip_check(hipMemsetAsync(op->d_buf[b], 0, op->real_size() * op->ndofs, op->stream), "memset");
    }
    if (!md.tfsf.empty()) {  // hexahedra owning TF/SF faces: face signs, face-node positions, records
        const dgtd::TfsfRecords r = dgtd::tfsf_records(mesh, R, md.tfsf, md.geo);
        op->ntfsf = (int)r.elems.size();
        op->d_tfsf_elem = static_cast<int*>(upload(r.elems));
        op->d_tfsf_sign = static_cast<int*>(upload(r.signs));
        op->d_tfsf_xyz = static_cast<double*>(upload(r.xyz));
        op->d_tfsf_geo = static_cast<double*>(upload(r.geo));
        op->d_lift = static_cast<double*>(upload(R.LIFT));
    }
    if (kernel != DGTD_KERNEL_CSR) {
        // matrix-free records: geometry with sigma/eps, connectivity, D1 and
        // the two used columns of M1^-1, face tables
        const int K = mesh->K, n1 = R.p + 1, Np = R.Np, Nfp = R.Nfp;
        std::vector<double> geo((size_t)dgtd::kHexGeoDev * K);
        for (int e = 0; e < K; ++e) {
            for (int v = 0; v < dgtd::kHexGeoStride; ++v)
                geo[(size_t)dgtd::kHexGeoDev * e + v] = md.geo[(size_t)dgtd::kHexGeoStride * e + v];
            const double sg = mesh->sigma ? mesh->sigma[e] : 0.0;
            if (!(sg >= 0.0) || !std::isfinite(sg)) throw std::runtime_error("conductivity must be finite and >= 0");
            geo[(size_t)dgtd::kHexGeoDev * e + 35] = sg * md.geo[(size_t)dgtd::kHexGeoStride * e + 33];
        }
        const dgtd::Gll1D g1 = dgtd::build_gll1d(R.p);
        std::vector<double> m(g1.D.begin(), g1.D.end());
        for (int i = 0; i < n1; ++i) {
            m.push_back(g1.Minv[(size_t)i * n1 + 0]);
            m.push_back(g1.Minv[(size_t)i * n1 + R.p]);
        }
        std::vector<int> tabs(R.Fmask.begin(), R.Fmask.end());
        tabs.insert(tabs.end(), R.face_perm.begin(), R.face_perm.end());
        std::vector<int> fq((size_t)6 * Np, -1);
        for (int f = 0; f < 6; ++f) {
            const int axis = f == 0 || f == 5 ? 2 : f == 1 || f == 3 ? 1 : 0;
            const int stride = axis == 0 ? 1 : axis == 1 ? n1 : n1 * n1;
            const int b = f == 2 || f == 3 || f == 5 ? R.p : 0;
            for (int q = 0; q < Nfp; ++q) {
                const int c = R.Fmask[(size_t)f * Nfp + q];
                if ((c / stride) % n1 != b) throw std::logic_error("hexahedron face table");
                for (int i = 0; i < n1; ++i) fq[(size_t)f * Np + c + (i - b) * stride] = q;
            }
        }
        for (int v : fq)
            if (v < 0) throw std::logic_error("hexahedron normal-line table");
        tabs.insert(tabs.end(), fq.begin(), fq.end());
        op->d_tabs = static_cast<int*>(upload(tabs));
        op->d_conn = static_cast<int*>(upload(md.conn));
        if (op->precision == 64) {
            op->d_geo = upload(geo);
            op->d_mats = upload(m);
        } else {
            op->d_geo = upload(dgtd::cast_vec<float>(geo));
            op->d_mats = upload(dgtd::cast_vec<float>(m));
        }
        hip_check(hipStreamSynchronize(op->stream), "sync");
        op->variant = DGTD_KERNEL_VALU;
        op->tuning.variant = DGTD_KERNEL_VALU;
        return;
    }
    hip_check(hipStreamSynchronize(op->stream), "sync");
    select_variant(op, DGTD_KERNEL_CSR, [&] { upload_csr(op, dgtd::assemble_global_hex(md, R, mesh->sigma, op->alpha)); });
}


}  // namespace

extern "C" {

const char* dgtd_last_error(void) { return g_last_error.c_str(); }
const char* dgtd_version(void) { return "dgtd_hip 0.1 (gfx950)"; }

int dgtd_op_create(const dgtd_mesh_desc* mesh, const dgtd_opts* opts, dgtd_op** out)
{
    dgtd::TraceRange tr("dgtd_op_create");
    if (!mesh || !opts || !out) return fail(DGTD_E_ARG, "NULL argument");
    *out = nullptr;
    if (opts->precision != 64 && opts->precision != 32)
        return fail(DGTD_E_ARG, "precision must be 64 or 32");
    if (opts->integrator != DGTD_RK4 && opts->integrator != DGTD_LSERK4)
        return fail(DGTD_E_ARG, "integrator must be DGTD_RK4 or DGTD_LSERK4");
    if (mesh->order < 1 || mesh->order > dgtd::kMaxOrder)
        return fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    if (!std::isfinite(opts->alpha)) return fail(DGTD_E_ARG, "alpha must be finite");
    if (opts->kernel < DGTD_KERNEL_AUTO || opts->kernel > DGTD_KERNEL_CSR)
        return fail(DGTD_E_ARG, "kernel must be DGTD_KERNEL_AUTO, _VALU, _MFMA or _CSR");
    const bool hexes = mesh->dim != 2 && mesh->nvert == 8;  // their matrix-free stage applies sigma too
    if (mesh->sigma && opts->kernel != DGTD_KERNEL_CSR && !hexes)
        return fail(DGTD_E_UNSUPPORTED, "bulk conductivity needs the assembled operator (DGTD_KERNEL_CSR)");
    if (mesh->dim != 0 && mesh->dim != 2 && mesh->dim != 3) return fail(DGTD_E_ARG, "dim must be 0, 2 or 3");
    if (mesh->nvert != 0 && !(mesh->dim == 2 && (mesh->nvert == 3 || mesh->nvert == 4)) &&
        !(mesh->dim != 2 && (mesh->nvert == 4 || mesh->nvert == 8)))
        return fail(DGTD_E_ARG, "nvert must be 0, 3 or 4 with dim = 2, or 4 or 8 with dim = 0 or 3");
    if (mesh->K < 1 || mesh->nv < 1 || !mesh->vx || !mesh->EToV)
        return fail(DGTD_E_ARG, "the mesh needs at least one element, its vertices and EToV");
    dgtd_op* op = nullptr;
    try {
        op = new dgtd_op;
        op->device = opts->device;
        op->precision = opts->precision;
        op->integrator = opts->integrator;
        op->alpha = opts->alpha;
        if (mesh->dim == 2 && mesh->nvert == 4) {
            create_quad(op, mesh, opts->kernel);
            *out = op;
            return DGTD_OK;
        }
        if (mesh->dim != 2 && mesh->nvert == 8) {
            create_hex(op, mesh, opts->kernel);
            *out = op;
            return DGTD_OK;
        }
        if (mesh->dim == 2) {
            create_tri(op, mesh, opts->kernel);
            *out = op;
            return DGTD_OK;
        }
        op->R = dgtd::build_reftet(mesh->order);
        op->K = mesh->K;
        op->Np = op->R.Np;
        op->Nfp = op->R.Nfp;
        op->N = (size_t)op->Np * mesh->K;
        op->ndofs = 6 * (int64_t)op->N;
        dgtd::MeshInput in;
        in.K = mesh->K;
        in.nv = mesh->nv;
        in.vx = mesh->vx;
        in.EToV = mesh->EToV;
        in.face_bc = mesh->face_bc;
        in.eps_r = mesh->eps_r;
        in.mu_r = mesh->mu_r;
        dgtd::MeshData md = dgtd::build_mesh(in, op->R, opts->kernel == DGTD_KERNEL_CSR ? 1.0 : op->alpha);
        op->max_fscale = md.max_fscale;
        op->nv = mesh->nv;
        op->vx.assign(mesh->vx, mesh->vx + 3 * (size_t)mesh->nv);
        op->EToV.assign(mesh->EToV, mesh->EToV + 4 * (size_t)mesh->K);

        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&op->ev0), "hipEventCreate");
        hip_check(hipEventCreate(&op->ev1), "hipEventCreate");
        std::vector<int> tabs(op->R.Fmask.begin(), op->R.Fmask.end());
        tabs.insert(tabs.end(), op->R.face_perm.begin(), op->R.face_perm.end());
        op->d_tabs = static_cast<int*>(upload(tabs));
        op->d_conn = static_cast<int*>(upload(md.conn));
        int bpc;
        if (op->precision == 64) {
            op->d_geo = upload(md.geo);
            op->d_mats = upload(dgtd::pack_mats<double>(op->R));
            op->d_mats_mfma = upload(dgtd::pack_mats_mfma<double>(op->R));
            bpc = dgtd::stage_blocks_per_cu<double>(op->R.p);
        } else {
            op->d_geo = upload(dgtd::cast_vec<float>(md.geo));
            op->d_mats = upload(dgtd::pack_mats<float>(op->R));
            op->d_mats_mfma = upload(dgtd::pack_mats_mfma<float>(op->R));
            bpc = dgtd::stage_blocks_per_cu<float>(op->R.p);
        }
        op->d_M = static_cast<double*>(upload(op->R.M));
        op->d_jac = static_cast<double*>(upload(md.jac));
        op->d_eps = static_cast<double*>(upload(md.eps));
        op->d_mu = static_cast<double*>(upload(md.mu));
        const int te = dgtd::norms_tile(op->Np);
        op->nparts = std::max(1, std::min(dgtd::kNormsMaxParts, (op->K + te - 1) / te));
        hip_check(hipMalloc(&op->d_partials, sizeof(double) * 3 * op->nparts), "hipMalloc");
        double* scratch_dev = nullptr;  // resolves the allocation-free entry point's scratch on this device
        hip_check(dgtd::norms_scratch(op->device, &scratch_dev), "norms scratch");
        hip_check(hipMalloc(&op->d_out3, sizeof(double) * 3), "hipMalloc");
        hip_check(hipMalloc(&op->d_unstable, sizeof(int)), "hipMalloc");
        hip_check(hipMemsetAsync(op->d_unstable, 0, sizeof(int), op->stream), "memset");
        const int nbuf = op->integrator == DGTD_LSERK4 ? 3 : 4;
        for (int b = 0; b < nbuf; ++b) {
            hip_check(hipMalloc(&op->d_buf[b], op->real_size() * op->ndofs), "hipMalloc state");
            hip_check(hipMemsetAsync(op->d_buf[b], 0, op->real_size() * op->ndofs, op->stream), "memset");
        }
        if (!md.tfsf.empty()) {  // elements owning TF/SF faces, their face signs, nodes and records
            const dgtd::TfsfRecords r = dgtd::tfsf_records(mesh, op->R, md.tfsf, md.geo);
            op->ntfsf = (int)r.elems.size();
            op->d_tfsf_elem = static_cast<int*>(upload(r.elems));
            op->d_tfsf_sign = static_cast<int*>(upload(r.signs));
            op->d_tfsf_xyz = static_cast<double*>(upload(r.xyz));
            op->d_tfsf_geo = static_cast<double*>(upload(r.geo));
            op->d_lift = static_cast<double*>(upload(op->R.LIFT));
        }
        hip_check(hipStreamSynchronize(op->stream), "sync");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, op->device), "hipGetDeviceProperties");
        const int epb = dgtd::elements_per_block(op->R.p);
        op->nblk = (op->K + epb - 1) / epb;
        op->grid = std::max(1, std::min(op->nblk, prop.multiProcessorCount * bpc));
        bool any_mfma = false;
        for (int c = 0; c < dgtd::kMfmaConfigs; ++c) {
            const int mbpc = op->precision == 64 ? dgtd::mfma_blocks_per_cu<double>(op->R.p, c)
                                                 : dgtd::mfma_blocks_per_cu<float>(op->R.p, c);
            if (mbpc <= 0 || !dgtd::mfma_fits_buffers(6 * (int64_t)op->Np * op->K, op->precision / 8)) continue;
            const int mwaves = op->precision == 64 ? dgtd::mfma_waves_per_block<double>(op->R.p, c)
                                                   : dgtd::mfma_waves_per_block<float>(op->R.p, c);
            op->grid_cfg[c] = mfma_grid(op->K, mwaves, prop.multiProcessorCount * mbpc);
            any_mfma = true;
        }
        if (any_mfma) op->d_order = static_cast<int*>(upload(morton_task_order(mesh)));
        if (opts->kernel == DGTD_KERNEL_MFMA && op->grid_cfg[0] <= 0)
            throw std::runtime_error("the MFMA kernel does not fit this order and precision, or a state buffer reaches 4 GiB");
        select_variant(op, opts->kernel,
                       [&] { upload_csr(op, dgtd::assemble_global_tet(md, op->R, mesh->sigma, op->alpha)); });
    } catch (const HipError& e) {
        delete op;
        return fail(DGTD_E_HIP, e.what());
    } catch (const UnsupportedError& e) {
        delete op;
        return fail(DGTD_E_UNSUPPORTED, e.what());
    } catch (const std::bad_alloc&) {
        delete op;
        return fail(DGTD_E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        delete op;
        return fail(DGTD_E_ARG, e.what());
    }
    *out = op;
    return DGTD_OK;
}

void dgtd_op_destroy(dgtd_op* op) { delete op; }

int dgtd_op_ndofs(const dgtd_op* op, int64_t* n)
{
    if (!op || !n) return fail(DGTD_E_ARG, "NULL argument");
    *n = op->ndofs;
    return DGTD_OK;
}

int dgtd_op_np(const dgtd_op* op, int32_t* Np)
{
    if (!op || !Np) return fail(DGTD_E_ARG, "NULL argument");
    *Np = op->Np;
    return DGTD_OK;
}

int dgtd_op_max_fscale(const dgtd_op* op, double* out)
{
    if (!op || !out) return fail(DGTD_E_ARG, "NULL argument");
    *out = op->max_fscale;
    return DGTD_OK;
}

int dgtd_op_estimate_time_step(const dgtd_op* op, int32_t evolution_operator, double cfl, double* dt)
{
    if (!op || !dt) return fail(DGTD_E_ARG, "NULL argument");
    const int order = op->hex ? op->H3.p : op->dim == 3 ? op->R.p : op->quad ? op->Q2.p : op->T2.p;
    return estimate_time_step(op->dim, op->quad, op->hex, order, op->K, op->vx.data(), op->EToV.data(),
                              op->max_fscale, evolution_operator, cfl, dt);
}

int dgtd_mesh_estimate_time_step(const dgtd_mesh_desc* mesh, int32_t evolution_operator, double cfl, double* dt)
{
    if (!mesh || !dt || !mesh->vx || !mesh->EToV || mesh->K < 1) return fail(DGTD_E_ARG, "NULL argument or empty mesh");
    if (mesh->order < 1 || mesh->order > dgtd::kMaxOrder) return fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    try {
        const bool quad = mesh->dim == 2 && mesh->nvert == 4;
        const bool hex = mesh->dim != 2 && mesh->nvert == 8;
        double max_fscale = 0.0;
        if (mesh->dim != 2 && !hex && evolution_operator == DGTD_OPERATOR_HESTHAVEN)
            max_fscale = dgtd::build_mesh(mesh_input(mesh), dgtd::build_reftet(mesh->order), 1.0).max_fscale;
        return estimate_time_step(mesh->dim == 2 ? 2 : 3, quad, hex, mesh->order, mesh->K, mesh->vx, mesh->EToV,
                                  max_fscale, evolution_operator, cfl, dt);
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
}

int dgtd_op_node_positions(const dgtd_op* op, double* xyz)
{
    if (!op || !xyz) return fail(DGTD_E_ARG, "NULL argument");
    if (op->quad) {
        dgtd::quad_node_positions(tri_input(op), op->Q2, xyz);
        return DGTD_OK;
    }
    if (op->hex) {
        dgtd::MeshInput in;
        in.K = op->K;
        in.nv = op->nv;
        in.vx = op->vx.data();
        in.EToV = op->EToV.data();
        dgtd::hex_node_positions(in, op->H3, xyz);
        return DGTD_OK;
    }
    if (op->dim == 2) {
        dgtd::tri_node_positions(tri_input(op), op->T2, xyz);
        return DGTD_OK;
    }
    dgtd::MeshInput in;
    in.K = op->K;
    in.nv = op->nv;
    in.vx = op->vx.data();
    in.EToV = op->EToV.data();
    dgtd::node_positions(in, op->R, xyz);
    return DGTD_OK;
}

int dgtd_op_mult_device(dgtd_op* op, double t, const void* d_in, void* d_out, void* stream)
{
    dgtd::TraceRange tr("dgtd_op_mult_device");
    if (!op || !d_in || !d_out) return fail(DGTD_E_ARG, "NULL argument");
    if (d_in == d_out) return fail(DGTD_E_ARG, "in and out must not alias");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : op->stream;
        if (op->precision == 64) op->mult_dev<double>(d_in, d_out, s, t);
        else op->mult_dev<float>(d_in, d_out, s, t);
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_mult(dgtd_op* op, double t, const double* in, double* out)
{
    dgtd::TraceRange tr("dgtd_op_mult");
    if (!op || !in || !out) return fail(DGTD_E_ARG, "NULL argument");
    void *d_in = nullptr, *d_out = nullptr;
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hip_check(hipMalloc(&d_in, op->real_size() * op->ndofs), "hipMalloc");
        hip_check(hipMalloc(&d_out, op->real_size() * op->ndofs), "hipMalloc");
        op->to_device(in, d_in);
        if (op->precision == 64) op->mult_dev<double>(d_in, d_out, op->stream, t);
        else op->mult_dev<float>(d_in, d_out, op->stream, t);
        op->to_host(d_out, out);
    } catch (const std::exception& e) {
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        return fail(DGTD_E_HIP, e.what());
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return DGTD_OK;
}

int dgtd_op_set_state(dgtd_op* op, const double* x, double t)
{
    if (!op || !x) return fail(DGTD_E_ARG, "NULL argument");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        op->cur = 0;
        op->to_device(x, op->d_buf[0]);
        // LSERK residual starts at zero (rk4a[0] = 0 makes the first stage ignore it)
        if (op->integrator == DGTD_LSERK4)
            hip_check(hipMemsetAsync(op->d_buf[2], 0, op->real_size() * op->ndofs, op->stream), "memset");
        hip_check(hipMemsetAsync(op->d_unstable, 0, sizeof(int), op->stream), "memset");
        hip_check(hipStreamSynchronize(op->stream), "sync");
        op->t = t;
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_get_state(dgtd_op* op, double* x, double* t)
{
    if (!op || !x) return fail(DGTD_E_ARG, "NULL argument");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        op->to_host(op->state(), x);
        if (t) *t = op->t;
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_advance(dgtd_op* op, double dt, int32_t nsteps)
{
    dgtd::TraceRange tr("dgtd_op_advance");
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    if (nsteps < 0 || !std::isfinite(dt)) return fail(DGTD_E_ARG, "bad dt or nsteps");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        if (op->timing) hip_check(hipEventRecord(op->ev0, op->stream), "hipEventRecord");
        int launches = 0;
        for (int s = 0; s < nsteps; ++s) {
            launches += op->step(dt);
            op->t += dt;
        }
        if (op->timing) {
            hip_check(hipEventRecord(op->ev1, op->stream), "hipEventRecord");
            hip_check(hipEventSynchronize(op->ev1), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, op->ev0, op->ev1), "hipEventElapsedTime");
            op->last_launches = launches;
            op->last_us = launches ? 1000.0 * ms / launches : 0.0;
        }
        hip_check(hipStreamSynchronize(op->stream), "sync");
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_run(dgtd_op* op, double dt, double final_time, int32_t* steps)
{
    dgtd::TraceRange tr("dgtd_op_run");
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    if (!(dt > 0.0) || !std::isfinite(final_time)) return fail(DGTD_E_ARG, "bad dt or final_time");
    int n = 0;
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        if (op->timing) hip_check(hipEventRecord(op->ev0, op->stream), "hipEventRecord");
        int launches = 0;
        while (op->t <= final_time - 1e-8 * dt) {
            const double truedt = std::min(dt, final_time - op->t);
            launches += op->step(truedt);
            op->t += truedt;
            ++n;
        }
        if (op->timing) {
            hip_check(hipEventRecord(op->ev1, op->stream), "hipEventRecord");
            hip_check(hipEventSynchronize(op->ev1), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, op->ev0, op->ev1), "hipEventElapsedTime");
            op->last_launches = launches;
            op->last_us = launches ? 1000.0 * ms / launches : 0.0;
        }
        hip_check(hipStreamSynchronize(op->stream), "sync");
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    if (steps) *steps = n;
    return DGTD_OK;
}

int dgtd_op_time(const dgtd_op* op, double* t)
{
    if (!op || !t) return fail(DGTD_E_ARG, "NULL argument");
    *t = op->t;
    return DGTD_OK;
}

int dgtd_op_step(dgtd_op* op, double* x, double* t, double dt, int32_t nsteps)
{
    if (!op || !x || !t) return fail(DGTD_E_ARG, "NULL argument");
    int rc = dgtd_op_set_state(op, x, *t);
    if (rc) return rc;
    rc = dgtd_op_advance(op, dt, nsteps);
    if (rc) return rc;
    return dgtd_op_get_state(op, x, t);
}

int dgtd_op_unstable(dgtd_op* op, int32_t* flag)
{
    if (!op || !flag) return fail(DGTD_E_ARG, "NULL argument");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        int h = 0;
        hip_check(hipMemcpyAsync(&h, op->d_unstable, sizeof(int), hipMemcpyDeviceToHost, op->stream), "D2H");
        hip_check(hipStreamSynchronize(op->stream), "sync");
        *flag = h ? 1 : 0;
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_norms_noalloc(dgtd_op* op, const void* d_state, double* d_out3, void* stream)
{
    if (!op || !d_out3) return fail(DGTD_E_ARG, "NULL argument");
    if (op->nparts > dgtd::kNormsMaxParts)
        return fail(DGTD_E_ARG, "reduction needs " + std::to_string(op->nparts) + " partial blocks, the scratch holds " +
                                    std::to_string(dgtd::kNormsMaxParts));
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        double* part = nullptr;
        hip_check(dgtd::norms_scratch(op->device, &part), "norms scratch");
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : op->stream;
        const void* q = d_state ? d_state : op->state();
        hipError_t e;
        if (op->precision == 64)
            e = dgtd::launch_norms<double>(static_cast<const double*>(q), op->K, op->Np, op->d_M, op->d_jac,
                                           op->d_eps, op->d_mu, part, op->nparts, d_out3, s);
        else
            e = dgtd::launch_norms<float>(static_cast<const float*>(q), op->K, op->Np, op->d_M, op->d_jac,
                                          op->d_eps, op->d_mu, part, op->nparts, d_out3, s);
        hip_check(e, "norms kernel launch");
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_norms(dgtd_op* op, double* out3)
{
    if (!op || !out3) return fail(DGTD_E_ARG, "NULL argument");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hipError_t e;
        if (op->precision == 64)
            e = dgtd::launch_norms<double>(static_cast<const double*>(op->state()), op->K, op->Np, op->d_M,
                                           op->d_jac, op->d_eps, op->d_mu, op->d_partials, op->nparts,
                                           op->d_out3, op->stream);
        else
            e = dgtd::launch_norms<float>(static_cast<const float*>(op->state()), op->K, op->Np, op->d_M,
                                          op->d_jac, op->d_eps, op->d_mu, op->d_partials, op->nparts,
                                          op->d_out3, op->stream);
        hip_check(e, "norms kernel launch");
        hip_check(hipMemcpyAsync(out3, op->d_out3, 3 * sizeof(double), hipMemcpyDeviceToHost, op->stream), "D2H");
        hip_check(hipStreamSynchronize(op->stream), "sync");
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_set_planewaves(dgtd_op* op, const dgtd_planewave* w, int32_t n)
{
    if (!op || n < 0 || (n > 0 && !w)) return fail(DGTD_E_ARG, "NULL argument or negative count");
    std::vector<dgtd::TfsfWave> dev;
    std::string err;
    if (!dgtd::planewaves_device(w, n, dev, err)) return fail(DGTD_E_ARG, err);
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(op->stream), "sync");
        if (op->d_waves) hip_check(hipFree(op->d_waves), "hipFree");
        op->d_waves = nullptr;
        op->nwaves = n;
        if (n > 0) op->d_waves = static_cast<dgtd::TfsfWave*>(upload(dev));
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_probe_set(dgtd_op* op, int32_t n, const double* xyz, int32_t* elem)
{
    if (!op || n < 0 || (n > 0 && !xyz)) return fail(DGTD_E_ARG, "NULL argument or negative count");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        std::vector<int32_t> el(n, -1);
        std::vector<double> w((size_t)n * op->Np, 0.0);
        if (op->quad) {
            dgtd::quad_locate_points(tri_input(op), op->Q2, n, xyz, el.data(), w.data());
        } else if (op->hex) {
            dgtd::MeshInput in;
            in.K = op->K;
            in.nv = op->nv;
            in.vx = op->vx.data();
            in.EToV = op->EToV.data();
            dgtd::hex_locate_points(in, op->H3, n, xyz, el.data(), w.data());
        } else if (op->dim == 2) {
            dgtd::tri_locate_points(tri_input(op), op->T2, n, xyz, el.data(), w.data());
        } else {
            dgtd::MeshInput in;
            in.K = op->K;
            in.nv = op->nv;
            in.vx = op->vx.data();
            in.EToV = op->EToV.data();
            dgtd::locate_points(in, op->R, n, xyz, el.data(), w.data());
        }
        for (void* q : {(void*)op->d_probe_elem, (void*)op->d_probe_w, (void*)op->d_probe_out})
            if (q) hip_check(hipFree(q), "hipFree");
        op->d_probe_elem = nullptr;
        op->d_probe_w = nullptr;
        op->d_probe_out = nullptr;
        op->nprobes = n;
        if (n > 0) {
            op->d_probe_elem = static_cast<int*>(upload(el));
            op->d_probe_w = static_cast<double*>(upload(w));
            hip_check(hipMalloc(&op->d_probe_out, 6 * sizeof(double) * (size_t)n), "hipMalloc");
        }
        if (elem)
            for (int p = 0; p < n; ++p) elem[p] = el[p];
    } catch (const std::bad_alloc&) {
        return fail(DGTD_E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_probe_sample(dgtd_op* op, double* out)
{
    if (!op || (op->nprobes > 0 && !out)) return fail(DGTD_E_ARG, "NULL argument");
    if (op->nprobes == 0) return DGTD_OK;
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hipError_t e;
        if (op->precision == 64)
            e = dgtd::launch_probe_sample<double>(static_cast<const double*>(op->state()), op->N, op->Np,
                                                  op->nprobes, op->d_probe_elem, op->d_probe_w, op->d_probe_out,
                                                  op->stream);
        else
            e = dgtd::launch_probe_sample<float>(static_cast<const float*>(op->state()), op->N, op->Np,
                                                 op->nprobes, op->d_probe_elem, op->d_probe_w, op->d_probe_out,
                                                 op->stream);
        hip_check(e, "probe kernel launch");
        hip_check(hipMemcpyAsync(out, op->d_probe_out, 6 * sizeof(double) * (size_t)op->nprobes,
                                 hipMemcpyDeviceToHost, op->stream),
                  "D2H");
        hip_check(hipStreamSynchronize(op->stream), "sync");
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_set_timing(dgtd_op* op, int32_t on)
{
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    op->timing = on != 0;
    return DGTD_OK;
}

int dgtd_op_set_graphs(dgtd_op* op, int32_t on)
{
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(op->stream), "sync");
        op->graphs = on != 0;
        op->drop_graphs();
    } catch (const std::exception& e) {
        return fail(DGTD_E_HIP, e.what());
    }
    return DGTD_OK;
}

int dgtd_op_last_kernel_us(const dgtd_op* op, double* avg_us, int32_t* launches)
{
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    if (avg_us) *avg_us = op->last_us;
    if (launches) *launches = op->last_launches;
    return DGTD_OK;
}

}  // extern "C"

namespace dgtd {
// The fastest matrix-core configuration of the LSERK stage for elements of
// this shape (dim 3 tetrahedra, 2 triangles), order and precision at about K
// elements: from the process cache, else by tuning an operator on a PEC cube
// or square of about K elements, which fills the cache.  Used by the batch engine, whose instances
// cannot be timed in place; it keeps its own rule for the per-node or
// matrix-core choice (that choice changes the arithmetic, the configuration
// does not), so an instance's result is the same alone and in any mix.
// Returns false when no matrix-core configuration was timed.
bool tuned_mfma_config(int dim, int order, int precision, int64_t K, int* cfg)
{
    TuneChoice choice;
    if (!tune_cache_get(dim, order, precision, kModeLserk, K, choice)) {
        // a PEC cube of n^3 x 6 tetrahedra (n in 4..32), or a PEC square of
        // n^2 x 2 triangles (n in 8..724)
        std::vector<double> vx;
        std::vector<int32_t> EToV;
        std::vector<int8_t> fbc;
        size_t nk = 0;
        if (dim == 2) {
            const int n = std::min(724, std::max(8, (int)std::lround(std::sqrt((double)K / 2.0))));
            nk = 2 * (size_t)n * n;
            vx.resize(3 * (size_t)(n + 1) * (n + 1));
            for (int j = 0; j <= n; ++j)
                for (int i = 0; i <= n; ++i) {
                    const size_t v = (size_t)j * (n + 1) + i;
                    vx[3 * v] = (double)i / n;
                    vx[3 * v + 1] = (double)j / n;
                    vx[3 * v + 2] = 0.0;
                }
            EToV.resize(3 * nk);
            fbc.assign(3 * nk, 0);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const int32_t a = j * (n + 1) + i, b = a + 1, c = a + n + 2, d = a + n + 1;
                    const size_t t = 2 * ((size_t)j * n + i);  // two counter-clockwise triangles
                    const int32_t lo[3] = {a, b, c}, up[3] = {a, c, d};
                    for (int k = 0; k < 3; ++k) {
                        EToV[3 * t + k] = lo[k];
                        EToV[3 * t + 3 + k] = up[k];
                    }
                    if (j == 0) fbc[3 * t] = DGTD_BC_PEC;              // a-b on y = 0
                    if (i == n - 1) fbc[3 * t + 1] = DGTD_BC_PEC;      // b-c on x = 1
                    if (j == n - 1) fbc[3 * t + 3 + 1] = DGTD_BC_PEC;  // c-d on y = 1
                    if (i == 0) fbc[3 * t + 3 + 2] = DGTD_BC_PEC;      // d-a on x = 0
                }
        } else {
            const int n = std::min(32, std::max(4, (int)std::lround(std::cbrt((double)K / 6.0))));
            const int8_t pec[6] = {DGTD_BC_PEC, DGTD_BC_PEC, DGTD_BC_PEC, DGTD_BC_PEC, DGTD_BC_PEC, DGTD_BC_PEC};
            nk = 6 * (size_t)n * n * n;
            vx.resize(3 * (size_t)(n + 1) * (n + 1) * (n + 1));
            EToV.resize(4 * nk);
            fbc.resize(4 * nk);
            if (dgtd_cube_mesh(n, 1.0, pec, vx.data(), EToV.data(), fbc.data()) != DGTD_OK) return false;
        }
        dgtd_mesh_desc m{};
        m.K = (int32_t)nk;
        m.order = order;
        m.vx = vx.data();
        m.nv = (int32_t)(vx.size() / 3);
        m.EToV = EToV.data();
        m.face_bc = fbc.data();
        m.dim = dim;
        dgtd_opts o{};
        o.alpha = 1.0;
        o.precision = precision;
        o.integrator = DGTD_LSERK4;
        o.kernel = DGTD_KERNEL_AUTO;
        int dev = 0;
        (void)hipGetDevice(&dev);
        o.device = dev;
        dgtd_op* op = nullptr;
        if (dgtd_op_create(&m, &o, &op) != DGTD_OK) return false;
        const bool got = tune_cache_get(dim, order, precision, kModeLserk, (int64_t)nk, choice);
        dgtd_op_destroy(op);
        if (!got) return false;
        tune_cache_put(dim, order, precision, kModeLserk, K, choice);
    }
    if (choice.best_mfma < 1) return false;
    *cfg = (choice.best_mfma - 1) / 2;
    return true;
}
}  // namespace dgtd

extern "C" {

int dgtd_op_tuning(const dgtd_op* op, dgtd_tuning* out)
{
    if (!op || !out) return fail(DGTD_E_ARG, "NULL argument");
    *out = op->tuning;
    return DGTD_OK;
}

int dgtd_op_set_candidate(dgtd_op* op, int32_t candidate)
{
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    if (op->variant == DGTD_KERNEL_CSR) return fail(DGTD_E_ARG, "assembled operator: no stage candidates");
    if (op->hex) return fail(DGTD_E_ARG, "hexahedra: one matrix-free stage, no candidates");
    if (candidate < 0 || candidate >= dgtd_op::kCandidates) return fail(DGTD_E_ARG, "candidate out of range");
    int var, cfg;
    bool mort;
    if (!op->candidate(candidate, var, cfg, mort)) return fail(DGTD_E_ARG, "candidate does not fit this operator");
    try {
        hip_check(hipSetDevice(op->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(op->stream), "hipStreamSynchronize");
        op->apply_candidate(candidate);
        op->tuning.from_cache = 0;
        op->drop_graphs();  // captured with the previous kernel
    } catch (const std::exception& ex) {
        return fail(DGTD_E_HIP, ex.what());
    }
    return DGTD_OK;
}

int dgtd_op_kernel_variant(const dgtd_op* op, int32_t* variant)
{
    if (!op || !variant) return fail(DGTD_E_ARG, "NULL argument");
    *variant = op->variant;
    return DGTD_OK;
}

int dgtd_op_work_per_element(const dgtd_op* op, double* bytes, double* flops)
{
    if (!op) return fail(DGTD_E_ARG, "NULL argument");
    const double s = (double)op->real_size();
    const double Np = op->Np, Nfp = op->Nfp;
    if (op->variant == DGTD_KERNEL_CSR) {
        // per entry its value and 16-bit column code; per row its pointer,
        // its share of the block's element table (32 ints per 16 rows) and
        // the LSERK vectors (state, residual in/out, out)
        const double rows = 6.0 * Np, nnz_e = (double)op->nnz / op->K;
        if (bytes) *bytes = nnz_e * (s + 2.0) + rows * (8.0 + 8.0) + rows * 4.0 * s;
        if (flops) *flops = 2.0 * nnz_e + 4.0 * rows;
        return DGTD_OK;
    }
    if (op->hex) {  // matrix-free hexahedra: state vectors, 36-value record, 12 connectivity ints
        const double n1 = op->H3.p + 1.0;
        if (bytes) *bytes = 24.0 * Np * s + 36.0 * s + 48.0;
        if (flops) *flops = Np * (36.0 * n1 + 198.0) + 240.0 * Nfp;
        return DGTD_OK;
    }
    if (op->dim == 2) {  // triangles: 2 derivative directions, 3 edges, 15-value record
        if (bytes) *bytes = 24.0 * Np * s + 15.0 * s + 24.0;
        if (flops) *flops = 24.0 * Np * Np + 36.0 * Np * Nfp + 150.0 * Nfp + 96.0 * Np;
        return DGTD_OK;
    }
    if (bytes) *bytes = 24.0 * Np * s + 27.0 * s + 32.0;
    if (flops) *flops = 36.0 * Np * Np + 48.0 * Np * Nfp + 240.0 * Nfp + 96.0 * Np;
    return DGTD_OK;
}

int dgtd_cube_mesh(int32_t n, double L, const int8_t* side_bc, double* vx, int32_t* EToV, int8_t* face_bc)
{
    if (n < 1 || !(L > 0.0)) return fail(DGTD_E_ARG, "n must be >= 1 and L > 0");
    dgtd::cube_mesh(n, L, side_bc, vx, EToV, face_bc);
    return DGTD_OK;
}

int dgtd_ref_element(int32_t order, double* rst, double* M, double* Dr, double* Ds, double* Dt, double* lift,
                     int32_t* fmask)
{
    try {
        const dgtd::RefTet R = dgtd::build_reftet(order);
        if (rst)
            for (size_t i = 0; i < R.xi.size(); ++i) rst[i] = 2.0 * R.xi[i] - 1.0;
        if (M) std::copy(R.M.begin(), R.M.end(), M);
        if (Dr) std::copy(R.Dr.begin(), R.Dr.end(), Dr);
        if (Ds) std::copy(R.Ds.begin(), R.Ds.end(), Ds);
        if (Dt) std::copy(R.Dt.begin(), R.Dt.end(), Dt);
        if (lift) std::copy(R.LIFT.begin(), R.LIFT.end(), lift);
        if (fmask) std::copy(R.Fmask.begin(), R.Fmask.end(), fmask);
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

int dgtd_mesh_vmaps(const dgtd_mesh_desc* mesh, double alpha, int32_t* vmapM, int32_t* vmapP, int8_t* face_codes)
{
    if (!mesh || !vmapM || !vmapP) return fail(DGTD_E_ARG, "NULL argument");
    if (mesh->dim == 2 || mesh->nvert == 8) return fail(DGTD_E_UNSUPPORTED, "tetrahedral meshes only");
    try {
        const dgtd::RefTet R = dgtd::build_reftet(mesh->order);
        const dgtd::MeshData md = dgtd::build_mesh(mesh_input(mesh), R, alpha);
        const int Np = R.Np, Nfp = R.Nfp;
        for (int e = 0; e < mesh->K; ++e)
            for (int f = 0; f < 4; ++f) {
                const int nb = md.conn[(size_t)dgtd::kConnStride * e + f];
                const int code = md.conn[(size_t)dgtd::kConnStride * e + 4 + f];
                if (face_codes) face_codes[4 * (size_t)e + f] = md.face_code[4 * (size_t)e + f];
                for (int b = 0; b < Nfp; ++b) {
                    const size_t n = ((size_t)e * 4 + f) * Nfp + b;
                    vmapM[n] = e * Np + R.Fmask[f * Nfp + b];
                    vmapP[n] = nb >= 0 ? nb * Np + R.face_perm[(size_t)code * Nfp + b] : vmapM[n];
                }
            }
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

int dgtd_mesh_locate(const dgtd_mesh_desc* mesh, int32_t n, const double* xyz, int32_t* elem, double* weights)
{
    if (!mesh || n < 0 || (n > 0 && (!xyz || !elem || !weights))) return fail(DGTD_E_ARG, "NULL argument");
    if (mesh->order < 1 || mesh->order > dgtd::kMaxOrder) return fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    try {
        if (mesh->dim == 2 && mesh->nvert == 4) {
            dgtd::quad_locate_points(tri_input(mesh), dgtd::build_refquad(mesh->order), n, xyz, elem, weights);
            return DGTD_OK;
        }
        if (mesh->dim == 2) {
            dgtd::tri_locate_points(tri_input(mesh), dgtd::build_reftri(mesh->order), n, xyz, elem, weights);
            return DGTD_OK;
        }
        if (mesh->nvert == 8) {
            dgtd::hex_locate_points(mesh_input(mesh), dgtd::build_refhex(mesh->order), n, xyz, elem, weights);
            return DGTD_OK;
        }
        const dgtd::RefTet R = dgtd::build_reftet(mesh->order);
        dgtd::locate_points(mesh_input(mesh), R, n, xyz, elem, weights);
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

int dgtd_mesh_geometry(const dgtd_mesh_desc* mesh, double alpha, double* geo)
{
    if (!mesh || !geo) return fail(DGTD_E_ARG, "NULL argument");
    if (mesh->dim == 2 || mesh->nvert == 8) return fail(DGTD_E_UNSUPPORTED, "tetrahedral meshes only");
    try {
        const dgtd::RefTet R = dgtd::build_reftet(mesh->order);
        const dgtd::MeshData md = dgtd::build_mesh(mesh_input(mesh), R, alpha);
        std::copy(md.geo.begin(), md.geo.end(), geo);
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

int dgtd_global_assemble(const dgtd_mesh_desc* mesh, double alpha, int64_t* nnz, int64_t* rowptr, int32_t* col,
                         double* val)
{
    if (!mesh || !nnz) return fail(DGTD_E_ARG, "NULL argument");
    const bool fill = rowptr || col || val;
    if (fill && !(rowptr && col && val)) return fail(DGTD_E_ARG, "rowptr, col and val: all or none");
    if (mesh->order < 1 || mesh->order > dgtd::kMaxOrder) return fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    if (!std::isfinite(alpha)) return fail(DGTD_E_ARG, "alpha must be finite");
    try {
        dgtd::Csr A;
        if (mesh->dim == 2 && mesh->nvert == 4) {
            const dgtd::RefQuad R = dgtd::build_refquad(mesh->order);
            A = dgtd::assemble_global_quad(dgtd::build_quad_mesh(tri_input(mesh), R, 1.0), R, mesh->sigma, alpha);
        } else if (mesh->dim == 2) {
            const dgtd::RefTri R = dgtd::build_reftri(mesh->order);
            A = dgtd::assemble_global_tri(dgtd::build_tri_mesh(tri_input(mesh), R, 1.0), R, mesh->sigma, alpha);
        } else if (mesh->nvert == 8) {
            const dgtd::RefHex R = dgtd::build_refhex(mesh->order);
            A = dgtd::assemble_global_hex(dgtd::build_hex_mesh(mesh_input(mesh), R, 1.0), R, mesh->sigma, alpha);
        } else {
            const dgtd::RefTet R = dgtd::build_reftet(mesh->order);
            A = dgtd::assemble_global_tet(dgtd::build_mesh(mesh_input(mesh), R, 1.0), R, mesh->sigma, alpha);
        }
        if (fill && (int64_t)A.val.size() != *nnz) throw std::runtime_error("*nnz does not match the matrix");
        *nnz = (int64_t)A.val.size();
        if (fill) {
            std::copy(A.rowptr.begin(), A.rowptr.end(), rowptr);
            std::copy(A.col.begin(), A.col.end(), col);
            std::copy(A.val.begin(), A.val.end(), val);
        }
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}

}  // extern "C"

namespace dgtd {
// Shared error reporting for the other ABI sources.
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace dgtd

extern "C" int dgtd_project_initial_fields(const dgtd_mesh_desc* mesh, const dgtd_initial_field* src,
                                           int32_t nsrc, double* x)
{
    if (!mesh || !x || (nsrc > 0 && !src)) return fail(DGTD_E_ARG, "NULL argument");
    try {
        if (mesh->order < 1 || mesh->order > dgtd::kMaxOrder) throw std::runtime_error("order must be in 1..5");
        const bool quad = mesh->dim == 2 && mesh->nvert == 4;
        const bool hex = mesh->dim != 2 && mesh->nvert == 8;
        const int Np = hex    ? dgtd::hex_num_nodes(mesh->order)
                       : quad ? dgtd::quad_num_nodes(mesh->order)
                              : mesh->dim == 2 ? dgtd::tri_num_nodes(mesh->order) : dgtd::num_nodes(mesh->order);
        const size_t N = (size_t)Np * mesh->K;
        std::vector<double> xyz(3 * N);
        if (hex) dgtd::hex_node_positions(mesh_input(mesh), dgtd::build_refhex(mesh->order), xyz.data());
        else if (quad) dgtd::quad_node_positions(tri_input(mesh), dgtd::build_refquad(mesh->order), xyz.data());
        else if (mesh->dim == 2) dgtd::tri_node_positions(tri_input(mesh), dgtd::build_reftri(mesh->order), xyz.data());
        else dgtd::node_positions(mesh_input(mesh), dgtd::build_reftet(mesh->order), xyz.data());
        std::fill(x, x + 6 * N, 0.0);
        for (int s = 0; s < nsrc; ++s) {
            const dgtd_initial_field& f = src[s];
            if (f.field != 0 && f.field != 1) throw std::runtime_error("field must be 0 (E) or 1 (H)");
            if (f.kind == DGTD_FIELD_GAUSSIAN && (f.dimension < 1 || f.dimension > 3 || !(f.spread > 0)))
                throw std::runtime_error("gaussian needs dimension 1..3 and spread > 0");
            if (f.kind == DGTD_FIELD_RESONANT && (f.nmodes < 1 || f.nmodes > 3))
                throw std::runtime_error("resonant mode needs 1..3 modes");
            if (f.kind < DGTD_FIELD_GAUSSIAN || f.kind > DGTD_FIELD_BESSELJ6_3D)
                throw std::runtime_error("unknown initial field kind");
            for (size_t n = 0; n < N; ++n) {
                double y[3];
                for (int d = 0; d < 3; ++d) y[d] = xyz[d * N + n] - f.center[d];
                double v;
                if (f.kind == DGTD_FIELD_GAUSSIAN) {
                    double r2 = 0.0;
                    for (int d = 0; d < f.dimension; ++d) r2 += y[d] * y[d];
                    v = std::exp(-r2 / (2.0 * f.spread * f.spread));
                } else if (f.kind == DGTD_FIELD_BESSELJ6_2D) {
                    v = dgtd::besselj6_2d(y[0], y[1]);
                } else if (f.kind == DGTD_FIELD_BESSELJ6_3D) {
                    v = dgtd::besselj6_3d(y[0], y[1], y[2]);
                } else {
                    v = 1.0;
                    for (int d = 0; d < f.nmodes; ++d) v *= std::sin(f.modes[d] * M_PI * y[d]);
                }
                for (int d = 0; d < 3; ++d) x[(3 * f.field + d) * N + n] += v * f.polarization[d];
            }
        }
    } catch (const std::exception& e) {
        return fail(DGTD_E_ARG, e.what());
    }
    return DGTD_OK;
}
