// Host side of the multi-instance batch (include/dgtd_hip.h, dgtd_batch_*):
// admission by device footprint, resident instance storage, the device
// instance table and work list, and the hipGraph of one time step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgtd_hip.h"
#include "kernels.hpp"
#include "trace.hpp"
#include "mesh.hpp"
#include "packing.hpp"
#include "reftet.hpp"
#include "tri.hpp"
#include "tfsf_host.hpp"

namespace dgtd {
int api_fail(int code, const std::string& msg);  // dgtd_api.cpp
bool tuned_mfma_config(int dim, int order, int precision, int64_t K, int* cfg);  // dgtd_api.cpp
}

namespace {

struct HipErr : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hcheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw HipErr(std::string(what) + ": " + hipGetErrorString(e));
}

// LSERK4 (Carpenter & Kennedy 1994) as published.
const double kA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                      -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double kB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                      1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                      2277821191437.0 / 14882151754819.0};
const double kC[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                      2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

int nodes(int dim, int order) { return dim == 2 ? dgtd::tri_num_nodes(order) : dgtd::num_nodes(order); }

// Device bytes of an instance: three state buffers, geometry and connectivity
// records, and the records of its ntf elements owning TF/SF faces (id, four
// face signs, face-node positions, FP64 geometry).
int64_t footprint(int dim, int K, int order, int precision, int64_t ntf = 0)
{
    const int64_t s = precision == 64 ? 8 : 4;
    const int64_t Np = nodes(dim, order);
    const int64_t geo = dim == 2 ? dgtd::kTriGeoStride : dgtd::kGeoStride;
    const int64_t conn = dim == 2 ? dgtd::kTriConnStride : dgtd::kConnStride;
    const int64_t nfn = dim == 2 ? 3 * (order + 1) : 4 * (order + 1) * (order + 2) / 2;
    return 3 * 6 * Np * K * s + (int64_t)K * geo * s + (int64_t)K * conn * 4 +
           ntf * ((1 + dgtd::kTfsfSignStride) * 4 + nfn * 3 * 8 + geo * 8);
}

// Elements with a TF/SF face code.
int64_t tfsf_elements(const dgtd_mesh_desc* mesh, int dim)
{
    int64_t n = 0;
    if (mesh->face_bc)
        for (int e = 0; e < mesh->K; ++e)
            for (int f = 0; f <= dim; ++f)
                if (mesh->face_bc[(size_t)(dim + 1) * e + f] == DGTD_BC_TFSF) {
                    ++n;
                    break;
                }
    return n;
}

}  // namespace

struct dgtd_batch {
    // Every entry point holds `mu` while it touches the tables or the stream;
    // dgtd_batch_add does its host setup and upload outside it, so callers may
    // admit from worker threads while another thread steps the batch.
    std::mutex mu;
    int device = 0, precision = 32;
    double alpha = 1.0;
    int max_slots = 0;
    int64_t budget = 0, used = 0;
    hipStream_t stream = nullptr;
    void* d_inst = nullptr;  // BatchInst<T>[max_slots]
    dgtd::BatchItem* d_items = nullptr;
    int item_cap = 0;
    int* d_count = nullptr;
    // matrix-core item ranges: tetrahedra of order o at [2o, 2o+1], triangles
    // of order o at [kTriRanges + 2o, kTriRanges + 2o + 1]
    static constexpr int kTriRanges = 2 * (dgtd::kMaxBatchOrder + 1);
    int* d_ranges = nullptr;
    void* d_mats[dgtd::kMaxBatchOrder + 1] = {};
    void* d_mats_mfma[dgtd::kMaxBatchOrder + 1] = {};
    int grid_mfma[dgtd::kMaxBatchOrder + 1] = {};
    // per order, decided by the autotuner when its first instance is admitted
    // (the process-wide cache, else a tuning run on a cube of that size):
    // matrix-core stage or per-node stage, and the matrix-core configuration
    bool tuned[dgtd::kMaxBatchOrder + 1] = {};
    bool mfma[dgtd::kMaxBatchOrder + 1] = {};
    int cfg[dgtd::kMaxBatchOrder + 1] = {};
    int* d_tabs[dgtd::kMaxBatchOrder + 1] = {};
    // triangles (every order on the matrix-core stage), per order: matrices,
    // node tables, launch grid, tuned configuration
    void* d_mats_tri[dgtd::kMaxBatchOrder + 1] = {};
    int* d_tabs_tri[dgtd::kMaxBatchOrder + 1] = {};
    int grid_tri[dgtd::kMaxBatchOrder + 1] = {};
    // TF/SF injection: LIFT per shape and order (FP64), the work list of
    // (instance, TF/SF element record) items and its device count
    double* d_lift[dgtd::kMaxBatchOrder + 1] = {};
    double* d_lift_tri[dgtd::kMaxBatchOrder + 1] = {};
    dgtd::BatchItem* d_tf_items = nullptr;
    int tf_cap = 0;
    int* d_tf_count = nullptr;
    bool cap_tf = false;  // the captured step launches the injection
    bool tuned_tri[dgtd::kMaxBatchOrder + 1] = {};
    int cfg_tri[dgtd::kMaxBatchOrder + 1] = {};
    int num_cu = 1;

    static constexpr int kSlotReserved = 100;  // host only: being set up by dgtd_batch_add (not a BatchState)
    static_assert(kSlotReserved > dgtd::kSlotCancelled, "the host-only state must differ from every BatchState");

    struct Slot {
        int state = dgtd::kSlotFree;
        int dim = 3, order = 0, K = 0, nblk = 0;
        int64_t bytes = 0;
        void* buf[5] = {};  // q0, q1, res, geo, conn
        void* tf_buf[5] = {};  // TF/SF element ids, face signs, node positions, geometry, waves
        int tf_n = 0, tf_nwaves = 0;
        bool reported = false;
        // host copy of the device record (authoritative only while dirty)
        double t = 0, dt = 0, dt_step = 0, final_time = 0;
        int cur = 0, steps = 0;
    };
    std::vector<Slot> slots;
    bool dirty = true;
    int nitems = 0, grid = 1;
    int graph_order = 0;
    size_t smem = 0;
    hipGraphExec_t graph = nullptr;

    size_t rs() const { return precision == 64 ? 8 : 4; }

    // Orders 2-5 run on the matrix-core stage (FP64 order 5 too, with its
    // matrices in global memory: the 64-cavity batch's FP64 rate rose from
    // 2.03e10 to 2.68e10 DOF*stages/s against the per-node stage); order 1
    // (4 nodes, 16-column tiles mostly padding) runs on the per-node stage.
    // The matrix-core configuration comes from the autotuner (the same
    // arithmetic in every configuration, so results do not depend on it).
    bool use_mfma(int order) const { return tuned[order] && mfma[order]; }

    // Tunes the order once per batch.  Called without `mu` (it may build and
    // time an operator); `tune_mu` serialises tuning, `mu` guards the result.
    std::mutex tune_mu;
    static constexpr int64_t kTuneElements = 131072;
    void tune(int dim, int order, int64_t K)
    {
        bool* done = dim == 2 ? tuned_tri : tuned;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (done[order]) return;
        }
        std::lock_guard<std::mutex> tl(tune_mu);
        {
            std::lock_guard<std::mutex> lk(mu);
            if (done[order]) return;
        }
        int c = 0;
        auto available = [&](int cf) {
            if (dim == 2)
                return (precision == 64 ? dgtd::tri_mfma_blocks_per_cu<double>(order, cf)
                                        : dgtd::tri_mfma_blocks_per_cu<float>(order, cf)) > 0;
            return precision == 64 ? dgtd::mfma_available<double>(order, cf) : dgtd::mfma_available<float>(order, cf);
        };
        const bool can = (dim == 2 || order >= 2) && available(0);
        // tuned at the size of a launch that fills the GPU (the order's
        // launch carries every instance of the order, not this one alone)
        if (can && !dgtd::tuned_mfma_config(dim, order, precision, std::max<int64_t>(K, kTuneElements), &c)) c = 0;
        const bool fits = available(c);
        if (dim == 2) {
            std::lock_guard<std::mutex> lk(mu);
            cfg_tri[order] = fits ? c : 0;
            tuned_tri[order] = true;
            return;
        }
        // Leaves `dirty` alone: marking the host table dirty without pulling
        // the device table first would push stale step counters over steps in
        // flight.  The admission that follows pulls, publishes and marks it;
        // the flush after that sees the order's new stage choice and
        // recaptures the step.
        std::lock_guard<std::mutex> lk(mu);
        mfma[order] = can;
        cfg[order] = fits ? c : 0;
        tuned[order] = true;
    }
    bool cap_mfma[dgtd::kMaxBatchOrder + 1] = {};
    int cap_cfg[dgtd::kMaxBatchOrder + 1] = {};
    bool cap_tri[dgtd::kMaxBatchOrder + 1] = {};
    int cap_cfg_tri[dgtd::kMaxBatchOrder + 1] = {};

    ~dgtd_batch()
    {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (auto& s : slots) {
            for (void* p : s.buf)
                if (p) (void)hipFree(p);
            for (void* p : s.tf_buf)
                if (p) (void)hipFree(p);
        }
        for (void* p : {(void*)d_tf_items, (void*)d_tf_count}) if (p) (void)hipFree(p);
        for (void* p : {d_inst, (void*)d_items, (void*)d_count, (void*)d_ranges}) if (p) (void)hipFree(p);
        for (int o = 0; o <= dgtd::kMaxBatchOrder; ++o) {
            if (d_mats[o]) (void)hipFree(d_mats[o]);
            if (d_mats_mfma[o]) (void)hipFree(d_mats_mfma[o]);
            if (d_tabs[o]) (void)hipFree(d_tabs[o]);
            if (d_mats_tri[o]) (void)hipFree(d_mats_tri[o]);
            if (d_tabs_tri[o]) (void)hipFree(d_tabs_tri[o]);
            if (d_lift[o]) (void)hipFree(d_lift[o]);
            if (d_lift_tri[o]) (void)hipFree(d_lift_tri[o]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Device table <-> host mirror.  The device owns t/cur/steps/state of
    // running instances between flushes.
    template <typename T>
    void pull()
    {
        std::vector<dgtd::BatchInst<T>> h(max_slots);
        hcheck(hipMemcpyAsync(h.data(), d_inst, sizeof(h[0]) * max_slots, hipMemcpyDeviceToHost, stream), "D2H table");
        hcheck(hipStreamSynchronize(stream), "sync");
        for (int i = 0; i < max_slots; ++i) {
            Slot& s = slots[i];
            if (s.state == dgtd::kSlotFree || s.state == kSlotReserved) continue;
            s.state = h[i].state;
            s.t = h[i].t;
            s.cur = h[i].cur;
            s.steps = h[i].steps;
            s.dt_step = h[i].dt_step;
        }
    }

    template <typename T>
    void push()
    {
        std::vector<dgtd::BatchInst<T>> h(max_slots);
        for (int i = 0; i < max_slots; ++i) {
            const Slot& s = slots[i];
            dgtd::BatchInst<T>& r = h[i];
            r = dgtd::BatchInst<T>{};
            r.state = s.state == kSlotReserved ? dgtd::kSlotFree : s.state;
            if (r.state == dgtd::kSlotFree) continue;
            r.order = s.order;
            r.K = s.K;
            r.nblk = s.nblk;
            r.cur = s.cur;
            r.steps = s.steps;
            r.q[0] = static_cast<T*>(s.buf[0]);
            r.q[1] = static_cast<T*>(s.buf[1]);
            r.res = static_cast<T*>(s.buf[2]);
            r.geo = static_cast<const T*>(s.buf[3]);
            r.conn = static_cast<const int*>(s.buf[4]);
            r.alpha = alpha;
            r.t = s.t;
            r.dt = s.dt;
            r.dt_step = s.dt_step;
            r.final_time = s.final_time;
            r.dim = s.dim;
            r.tf_elem = static_cast<const int*>(s.tf_buf[0]);
            r.tf_sign = static_cast<const int*>(s.tf_buf[1]);
            r.tf_xyz = static_cast<const double*>(s.tf_buf[2]);
            r.tf_geo = static_cast<const double*>(s.tf_buf[3]);
            r.tf_waves = static_cast<const dgtd::TfsfWave*>(s.tf_buf[4]);
            r.tf_n = s.tf_n;
            r.tf_nwaves = s.tf_nwaves;
        }
        hcheck(hipMemcpyAsync(d_inst, h.data(), sizeof(h[0]) * max_slots, hipMemcpyHostToDevice, stream), "H2D table");
        // work list: running instances' element blocks for the per-node stage
        // (sorted by order), then per matrix-core order its (instance, task) items
        std::vector<dgtd::BatchItem> items;
        int max_order = 1;
        for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
            if (use_mfma(o)) continue;
            for (int i = 0; i < max_slots; ++i)
                if (slots[i].state == dgtd::kSlotRunning && slots[i].dim == 3 && slots[i].order == o) {
                    max_order = std::max(max_order, o);
                    for (int blk = 0; blk < slots[i].nblk; ++blk) items.push_back({i, blk});
                }
        }
        const int nvalu = (int)items.size();
        int ranges[2 * kTriRanges] = {};
        auto add_tasks = [&](int dim, int o) {
            for (int i = 0; i < max_slots; ++i)
                if (slots[i].state == dgtd::kSlotRunning && slots[i].dim == dim && slots[i].order == o) {
                    const int ntask = (slots[i].K + dgtd::kMfmaTaskElems - 1) / dgtd::kMfmaTaskElems;
                    for (int t = 0; t < ntask; ++t) items.push_back({i, t});
                }
        };
        for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
            ranges[2 * o] = (int)items.size();
            if (use_mfma(o)) add_tasks(3, o);
            ranges[2 * o + 1] = (int)items.size();
        }
        for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
            ranges[kTriRanges + 2 * o] = (int)items.size();
            if (tuned_tri[o]) add_tasks(2, o);
            ranges[kTriRanges + 2 * o + 1] = (int)items.size();
        }
        if ((int)items.size() > item_cap) {
            if (d_items) hcheck(hipFree(d_items), "hipFree");
            item_cap = std::max<int>((int)items.size(), 2 * item_cap);
            hcheck(hipMalloc(&d_items, sizeof(dgtd::BatchItem) * item_cap), "hipMalloc items");
            if (graph) {
                hcheck(hipGraphExecDestroy(graph), "hipGraphExecDestroy");
                graph = nullptr;
            }
        }
        nitems = (int)items.size();
        if (nitems)
            hcheck(hipMemcpyAsync(d_items, items.data(), sizeof(items[0]) * nitems, hipMemcpyHostToDevice, stream),
                   "H2D items");
        hcheck(hipMemcpyAsync(d_count, &nvalu, sizeof(int), hipMemcpyHostToDevice, stream), "H2D count");
        hcheck(hipMemcpyAsync(d_ranges, ranges, sizeof(ranges), hipMemcpyHostToDevice, stream), "H2D ranges");
        // TF/SF work list: the running instances' elements owning TF/SF faces
        std::vector<dgtd::BatchItem> tf;
        for (int i = 0; i < max_slots; ++i)
            if (slots[i].state == dgtd::kSlotRunning && slots[i].tf_nwaves > 0)
                for (int k = 0; k < slots[i].tf_n; ++k) tf.push_back({i, k});
        const int ntf = (int)tf.size();
        if (ntf > tf_cap) {
            if (d_tf_items) hcheck(hipFree(d_tf_items), "hipFree");
            tf_cap = std::max(ntf, 2 * tf_cap);
            hcheck(hipMalloc(&d_tf_items, sizeof(dgtd::BatchItem) * tf_cap), "hipMalloc TF/SF items");
            if (graph) {
                hcheck(hipGraphExecDestroy(graph), "hipGraphExecDestroy");
                graph = nullptr;
            }
        }
        if (ntf)
            hcheck(hipMemcpyAsync(d_tf_items, tf.data(), sizeof(tf[0]) * ntf, hipMemcpyHostToDevice, stream),
                   "H2D TF/SF items");
        hcheck(hipMemcpyAsync(d_tf_count, &ntf, sizeof(int), hipMemcpyHostToDevice, stream), "H2D TF/SF count");
        hcheck(hipStreamSynchronize(stream), "sync");  // host vectors die here
        bool recapture = max_order != graph_order || !graph || (ntf > 0) != cap_tf;
        for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o)
            recapture = recapture || use_mfma(o) != cap_mfma[o] || cfg[o] != cap_cfg[o] ||
                        tuned_tri[o] != cap_tri[o] || cfg_tri[o] != cap_cfg_tri[o];
        if (recapture) capture<T>(max_order, ntf > 0);
    }

    template <typename T>
    void capture(int max_order, bool with_tf)
    {
        dgtd::TraceRange tr("dgtd batch capture");
        if (graph) {
            hcheck(hipGraphExecDestroy(graph), "hipGraphExecDestroy");
            graph = nullptr;
        }
        smem = dgtd::batch_smem_bytes<T>(max_order);
        const int bpc = dgtd::batch_blocks_per_cu<T>(smem);
        grid = std::max(1, num_cu * bpc);
        dgtd::BatchArgs<T> a{};
        a.inst = static_cast<dgtd::BatchInst<T>*>(d_inst);
        a.items = d_items;
        a.count = d_count;
        for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
            a.mats[o] = static_cast<const T*>(d_mats[o]);
            a.mats_mfma[o] = static_cast<const T*>(d_mats_mfma[o]);
            a.tabs[o] = d_tabs[o];
            a.mats_mfma_tri[o] = static_cast<const T*>(d_mats_tri[o]);
            a.tabs_tri[o] = d_tabs_tri[o];
            if (use_mfma(o)) {
                const int bpc = dgtd::mfma_blocks_per_cu<T>(o, cfg[o]);
                grid_mfma[o] = std::max(1, bpc) * num_cu;
            }
            if (tuned_tri[o]) {
                const int bpc = dgtd::tri_mfma_blocks_per_cu<T>(o, cfg_tri[o]);
                grid_tri[o] = std::max(1, bpc) * num_cu;
            }
        }
        hipGraph_t g = nullptr;
        hcheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipError_t err = hipSuccess;
        for (int s = 0; s < 5 && err == hipSuccess; ++s) {
            a.stage = s;
            a.ca = (T)kA[s];
            a.cb = (T)kB[s];
            err = dgtd::launch_batch_stage<T>(a, grid, smem, stream);
            for (int o = 1; o <= dgtd::kMaxBatchOrder && err == hipSuccess; ++o)
                if (use_mfma(o))
                    err = dgtd::launch_batch_stage_mfma<T>(o, 3, a, d_ranges + 2 * o, grid_mfma[o], stream, cfg[o]);
            for (int o = 1; o <= dgtd::kMaxBatchOrder && err == hipSuccess; ++o)
                if (tuned_tri[o])
                    err = dgtd::launch_batch_stage_mfma<T>(o, 2, a, d_ranges + kTriRanges + 2 * o, grid_tri[o], stream,
                                                           cfg_tri[o]);
            if (with_tf && err == hipSuccess) {
                // res += dt S, q += b dt S at the stage time t + c dt (every instance its own t, dt)
                dgtd::BatchTfsfArgs<T> ta{};
                ta.inst = a.inst;
                ta.items = d_tf_items;
                ta.count = d_tf_count;
                for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
                    ta.lift[o] = d_lift[o];
                    ta.lift_tri[o] = d_lift_tri[o];
                }
                ta.stage = s;
                ta.c = kC[s];
                ta.b = kB[s];
                err = dgtd::launch_batch_tfsf<T>(ta, 4 * num_cu, stream);
            }
        }
        if (err == hipSuccess) err = dgtd::launch_batch_advance<T>(a.inst, max_slots, stream);
        hipError_t e2 = hipStreamEndCapture(stream, &g);
        hcheck(err, "batch launch during capture");
        hcheck(e2, "hipStreamEndCapture");
        hipError_t e3 = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hcheck(e3, "hipGraphInstantiate");
        graph_order = max_order;
        cap_tf = with_tf;
        for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
            cap_mfma[o] = use_mfma(o);
            cap_cfg[o] = cfg[o];
            cap_tri[o] = tuned_tri[o];
            cap_cfg_tri[o] = cfg_tri[o];
        }
    }

    void flush()
    {
        if (!dirty) return;
        if (precision == 64) push<double>();
        else push<float>();
        dirty = false;
    }

    void sync_pull()
    {
        hcheck(hipStreamSynchronize(stream), "sync");
        if (precision == 64) pull<double>();
        else pull<float>();
    }
};

namespace {

// Runs without the batch lock: synchronous copies on the null stream (the
// batch stream is non-blocking, so a step in flight is not waited for).
template <typename T, typename MD>
void upload_instance(dgtd_batch* b, dgtd_batch::Slot& s, const MD& md, const double* x0, size_t ndofs)
{
    (void)b;
    const size_t bytes_q = sizeof(T) * ndofs;
    for (int k = 0; k < 3; ++k) hcheck(hipMalloc(&s.buf[k], bytes_q), "hipMalloc instance");
    hcheck(hipMalloc(&s.buf[3], sizeof(T) * md.geo.size()), "hipMalloc geo");
    hcheck(hipMalloc(&s.buf[4], sizeof(int) * md.conn.size()), "hipMalloc conn");
    std::vector<T> tmp(ndofs, T(0));
    if (x0)
        for (size_t i = 0; i < ndofs; ++i) tmp[i] = (T)x0[i];
    hcheck(hipMemcpy(s.buf[0], tmp.data(), bytes_q, hipMemcpyHostToDevice), "H2D x0");
    std::fill(tmp.begin(), tmp.end(), T(0));
    hcheck(hipMemcpy(s.buf[2], tmp.data(), bytes_q, hipMemcpyHostToDevice), "H2D res");
    const std::vector<T> g = dgtd::cast_vec<T>(md.geo);
    hcheck(hipMemcpy(s.buf[3], g.data(), sizeof(T) * g.size(), hipMemcpyHostToDevice), "H2D geo");
    hcheck(hipMemcpy(s.buf[4], md.conn.data(), sizeof(int) * md.conn.size(), hipMemcpyHostToDevice), "H2D conn");
}

// The instance's TF/SF element records and plane waves (device copies).
void upload_tfsf(dgtd_batch::Slot& s, const dgtd::TfsfRecords& r, const std::vector<dgtd::TfsfWave>& waves)
{
    auto put = [](void*& dst, const void* src, size_t bytes) {
        hcheck(hipMalloc(&dst, bytes), "hipMalloc TF/SF");
        hcheck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "H2D TF/SF");
    };
    put(s.tf_buf[0], r.elems.data(), sizeof(int) * r.elems.size());
    put(s.tf_buf[1], r.signs.data(), sizeof(int) * r.signs.size());
    put(s.tf_buf[2], r.xyz.data(), sizeof(double) * r.xyz.size());
    put(s.tf_buf[3], r.geo.data(), sizeof(double) * r.geo.size());
    put(s.tf_buf[4], waves.data(), sizeof(dgtd::TfsfWave) * waves.size());
    s.tf_n = (int)r.elems.size();
    s.tf_nwaves = (int)waves.size();
}

template <typename T>
void upload_refs(dgtd_batch* b)
{
    for (int o = 1; o <= dgtd::kMaxBatchOrder; ++o) {
        const dgtd::RefTet R = dgtd::build_reftet(o);
        const std::vector<T> m = dgtd::pack_mats<T>(R);
        std::vector<int> tabs(R.Fmask.begin(), R.Fmask.end());
        tabs.insert(tabs.end(), R.face_perm.begin(), R.face_perm.end());
        hcheck(hipMalloc(&b->d_mats[o], sizeof(T) * m.size()), "hipMalloc mats");
        hcheck(hipMemcpy(b->d_mats[o], m.data(), sizeof(T) * m.size(), hipMemcpyHostToDevice), "H2D mats");
        if (dgtd::mfma_available<T>(o, 0) || dgtd::mfma_available<T>(o, 1)) {
            const std::vector<T> mm = dgtd::pack_mats_mfma<T>(R);
            hcheck(hipMalloc(&b->d_mats_mfma[o], sizeof(T) * mm.size()), "hipMalloc mats");
            hcheck(hipMemcpy(b->d_mats_mfma[o], mm.data(), sizeof(T) * mm.size(), hipMemcpyHostToDevice), "H2D mats");
        }
        hcheck(hipMalloc(&b->d_tabs[o], sizeof(int) * tabs.size()), "hipMalloc tabs");
        hcheck(hipMemcpy(b->d_tabs[o], tabs.data(), sizeof(int) * tabs.size(), hipMemcpyHostToDevice), "H2D tabs");
        hcheck(hipMalloc(&b->d_lift[o], sizeof(double) * R.LIFT.size()), "hipMalloc LIFT");
        hcheck(hipMemcpy(b->d_lift[o], R.LIFT.data(), sizeof(double) * R.LIFT.size(), hipMemcpyHostToDevice),
               "H2D LIFT");
        if (dgtd::tri_mfma_blocks_per_cu<T>(o, 0) > 0) {  // triangles
            const dgtd::RefTri Rt = dgtd::build_reftri(o);
            const std::vector<T> mt = dgtd::pack_mats_mfma<T>(Rt);
            std::vector<int> tt(Rt.Fmask.begin(), Rt.Fmask.end());
            tt.insert(tt.end(), Rt.edge_perm.begin(), Rt.edge_perm.end());
            hcheck(hipMalloc(&b->d_mats_tri[o], sizeof(T) * mt.size()), "hipMalloc mats");
            hcheck(hipMemcpy(b->d_mats_tri[o], mt.data(), sizeof(T) * mt.size(), hipMemcpyHostToDevice), "H2D mats");
            hcheck(hipMalloc(&b->d_tabs_tri[o], sizeof(int) * tt.size()), "hipMalloc tabs");
            hcheck(hipMemcpy(b->d_tabs_tri[o], tt.data(), sizeof(int) * tt.size(), hipMemcpyHostToDevice), "H2D tabs");
            hcheck(hipMalloc(&b->d_lift_tri[o], sizeof(double) * Rt.LIFT.size()), "hipMalloc LIFT");
            hcheck(hipMemcpy(b->d_lift_tri[o], Rt.LIFT.data(), sizeof(double) * Rt.LIFT.size(),
                             hipMemcpyHostToDevice),
                   "H2D LIFT");
        }
    }
}

int batch_error(const std::exception& e)
{
    if (dynamic_cast<const HipErr*>(&e)) return dgtd::api_fail(DGTD_E_HIP, e.what());
    return dgtd::api_fail(DGTD_E_ARG, e.what());
}

}  // namespace

// Quadrilaterals and hexahedra run on the assembled (Global) operator only;
// the batch advances hesthaven-operator instances.
constexpr const char* kBatchNoTensorElements =
    "quadrilaterals and hexahedra run the Global operator only (dgtd_op_create: DGTD_KERNEL_CSR, or for hexahedra "
    "DGTD_KERNEL_AUTO / _VALU), "
    "not in a batch";
inline bool tensor_elements(const dgtd_mesh_desc* mesh)
{
    return (mesh->dim == 2 && mesh->nvert == 4) || (mesh->dim != 2 && mesh->nvert == 8);
}

extern "C" {

int dgtd_batch_footprint(const dgtd_mesh_desc* mesh, int32_t precision, int64_t* bytes)
{
    if (!mesh || !bytes) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    if (mesh->order < 1 || mesh->order > dgtd::kMaxOrder) return dgtd::api_fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    if (tensor_elements(mesh)) return dgtd::api_fail(DGTD_E_UNSUPPORTED, kBatchNoTensorElements);
    const int dim = mesh->dim == 2 ? 2 : 3;
    *bytes = footprint(dim, mesh->K, mesh->order, precision, tfsf_elements(mesh, dim));
    return DGTD_OK;
}

int dgtd_batch_create(const dgtd_batch_opts* opts, dgtd_batch** out)
{
    if (!opts || !out) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    *out = nullptr;
    if (opts->precision != 32 && opts->precision != 64) return dgtd::api_fail(DGTD_E_ARG, "precision must be 64 or 32");
    if (opts->max_instances < 1 || opts->max_instances > 4096)
        return dgtd::api_fail(DGTD_E_ARG, "max_instances must be in 1..4096");
    if (!std::isfinite(opts->alpha)) return dgtd::api_fail(DGTD_E_ARG, "alpha must be finite");
    dgtd_batch* b = nullptr;
    try {
        b = new dgtd_batch;
        b->device = opts->device;
        b->precision = opts->precision;
        b->alpha = opts->alpha;
        b->max_slots = opts->max_instances;
        b->slots.resize(b->max_slots);
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        hcheck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        hipDeviceProp_t prop;
        hcheck(hipGetDeviceProperties(&prop, b->device), "hipGetDeviceProperties");
        b->num_cu = prop.multiProcessorCount;
        size_t free_b = 0, total_b = 0;
        hcheck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        b->budget = opts->hbm_budget > 0 ? opts->hbm_budget : (int64_t)(0.8 * (double)free_b);
        const size_t rec = b->precision == 64 ? sizeof(dgtd::BatchInst<double>) : sizeof(dgtd::BatchInst<float>);
        hcheck(hipMalloc(&b->d_inst, rec * b->max_slots), "hipMalloc table");
        hcheck(hipMalloc(&b->d_count, sizeof(int)), "hipMalloc count");
        hcheck(hipMalloc(&b->d_tf_count, sizeof(int)), "hipMalloc TF/SF count");
        hcheck(hipMalloc(&b->d_ranges, sizeof(int) * 2 * dgtd_batch::kTriRanges), "hipMalloc ranges");
        b->item_cap = 1 << 16;
        hcheck(hipMalloc(&b->d_items, sizeof(dgtd::BatchItem) * b->item_cap), "hipMalloc items");
        if (b->precision == 64) upload_refs<double>(b);
        else upload_refs<float>(b);
        b->flush();
    } catch (const std::exception& e) {
        delete b;
        return batch_error(e);
    }
    *out = b;
    return DGTD_OK;
}

void dgtd_batch_destroy(dgtd_batch* b) { delete b; }

int dgtd_batch_add(dgtd_batch* b, const dgtd_mesh_desc* mesh, const double* x0, double dt, double final_time,
                   int32_t* id)
{
    return dgtd_batch_add_tfsf(b, mesh, x0, dt, final_time, nullptr, 0, id);
}

int dgtd_batch_add_tfsf(dgtd_batch* b, const dgtd_mesh_desc* mesh, const double* x0, double dt, double final_time,
                        const dgtd_planewave* waves, int32_t nwaves, int32_t* id)
{
    dgtd::TraceRange tr("dgtd_batch_add");
    if (!b || !mesh || !id || nwaves < 0 || (nwaves > 0 && !waves))
        return dgtd::api_fail(DGTD_E_ARG, "NULL argument or negative wave count");
    std::vector<dgtd::TfsfWave> dwaves;
    {
        std::string err;
        if (!dgtd::planewaves_device(waves, nwaves, dwaves, err)) return dgtd::api_fail(DGTD_E_ARG, err);
    }
    const int dim = mesh->dim == 2 ? 2 : 3;
    if (mesh->nvert != 0 && !(mesh->dim == 2 && (mesh->nvert == 3 || mesh->nvert == 4)) &&
        !(mesh->dim != 2 && (mesh->nvert == 4 || mesh->nvert == 8)))
        return dgtd::api_fail(DGTD_E_ARG, "nvert must be 0, 3 or 4 with dim = 2, or 4 or 8 with dim = 0 or 3");
    if (tensor_elements(mesh)) return dgtd::api_fail(DGTD_E_UNSUPPORTED, kBatchNoTensorElements);
    if (mesh->K < 1 || mesh->nv < 1 || !mesh->vx || !mesh->EToV)
        return dgtd::api_fail(DGTD_E_ARG, "the mesh needs at least one element, its vertices and EToV");
    if (mesh->sigma)
        return dgtd::api_fail(DGTD_E_UNSUPPORTED, "bulk conductivity needs the assembled operator (DGTD_KERNEL_CSR)");
    if (mesh->order < 1 || mesh->order > dgtd::kMaxBatchOrder)
        return dgtd::api_fail(DGTD_E_UNSUPPORTED, "order must be in 1..5");
    if (!std::isfinite(final_time) || !std::isfinite(dt)) return dgtd::api_fail(DGTD_E_ARG, "bad dt or final_time");
    const int64_t bytes = footprint(dim, mesh->K, mesh->order, b->precision, tfsf_elements(mesh, dim));
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        b->tune(dim, mesh->order, mesh->K);
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    {
        std::lock_guard<std::mutex> lk(b->mu);
        if (dim == 2 && !b->tuned_tri[mesh->order])
            return dgtd::api_fail(DGTD_E_UNSUPPORTED, "no matrix-core stage for triangles of this order and precision");
        if ((dim == 2 || b->use_mfma(mesh->order)) &&
            !dgtd::mfma_fits_buffers(6 * (int64_t)nodes(dim, mesh->order) * mesh->K, b->precision / 8))
            return dgtd::api_fail(DGTD_E_UNSUPPORTED,
                                  "instance state reaches 4 GiB per buffer: too large for the batched "
                                  "matrix-core stage (run it as a single operator)");
    }
    int slot = -1;
    {  // reserve a slot and the bytes
        std::lock_guard<std::mutex> lk(b->mu);
        for (int i = 0; i < b->max_slots && slot < 0; ++i)
            if (b->slots[i].state == dgtd::kSlotFree) slot = i;
        if (slot < 0) return dgtd::api_fail(DGTD_E_CAPACITY, "no free instance slot");
        if (b->used + bytes > b->budget) return dgtd::api_fail(DGTD_E_CAPACITY, "HBM budget exceeded");
        b->slots[slot] = dgtd_batch::Slot{};
        b->slots[slot].state = dgtd_batch::kSlotReserved;
        b->slots[slot].bytes = bytes;
        b->used += bytes;
    }
    dgtd_batch::Slot s;  // built outside the lock
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        const size_t ndofs = 6 * (size_t)nodes(dim, mesh->order) * mesh->K;
        double max_fscale = 0.0;
        if (dim == 2) {
            const dgtd::RefTri R = dgtd::build_reftri(mesh->order);
            dgtd::TriMeshInput in;
            in.K = mesh->K;
            in.nv = mesh->nv;
            in.vx = mesh->vx;
            in.EToV = mesh->EToV;
            in.face_bc = mesh->face_bc;
            in.eps_r = mesh->eps_r;
            in.mu_r = mesh->mu_r;
            const dgtd::TriMeshData md = dgtd::build_tri_mesh(in, R, b->alpha);
            max_fscale = md.max_fscale;
            if (b->precision == 64) upload_instance<double>(b, s, md, x0, ndofs);
            else upload_instance<float>(b, s, md, x0, ndofs);
            if (!md.tfsf.empty() && !dwaves.empty()) upload_tfsf(s, dgtd::tfsf_records(mesh, R, md.tfsf, md.geo), dwaves);
        } else {
            const dgtd::RefTet R = dgtd::build_reftet(mesh->order);
            dgtd::MeshInput in;
            in.K = mesh->K;
            in.nv = mesh->nv;
            in.vx = mesh->vx;
            in.EToV = mesh->EToV;
            in.face_bc = mesh->face_bc;
            in.eps_r = mesh->eps_r;
            in.mu_r = mesh->mu_r;
            const dgtd::MeshData md = dgtd::build_mesh(in, R, b->alpha);
            max_fscale = md.max_fscale;
            if (b->precision == 64) upload_instance<double>(b, s, md, x0, ndofs);
            else upload_instance<float>(b, s, md, x0, ndofs);
            if (!md.tfsf.empty() && !dwaves.empty()) upload_tfsf(s, dgtd::tfsf_records(mesh, R, md.tfsf, md.geo), dwaves);
        }
        s.dim = dim;
        s.order = mesh->order;
        s.K = mesh->K;
        s.nblk = dim == 2 ? 0
                          : (mesh->K + dgtd::elements_per_block(mesh->order) - 1) / dgtd::elements_per_block(mesh->order);
        s.bytes = bytes;
        s.dt = dt > 0.0 ? dt : 1.0 / (max_fscale * mesh->order * mesh->order);
        s.final_time = final_time;
        if (0.0 <= final_time - 1e-8 * s.dt) {
            s.state = dgtd::kSlotRunning;
            s.dt_step = std::min(s.dt, final_time);
        } else {
            s.state = dgtd::kSlotDone;
        }
    } catch (const std::exception& e) {
        for (void*& p : s.buf)
            if (p) { (void)hipFree(p); p = nullptr; }
        for (void*& p : s.tf_buf)
            if (p) { (void)hipFree(p); p = nullptr; }
        std::lock_guard<std::mutex> lk(b->mu);
        b->slots[slot] = dgtd_batch::Slot{};
        b->used -= bytes;
        return batch_error(e);
    }
    {  // publish
        std::lock_guard<std::mutex> lk(b->mu);
        try {
            if (!b->dirty) b->sync_pull();  // device values are authoritative: fetch before editing
        } catch (const std::exception& e) {
            return batch_error(e);
        }
        b->slots[slot] = s;
        b->dirty = true;
    }
    *id = slot;
    return DGTD_OK;
}

int dgtd_batch_step(dgtd_batch* b, int32_t nsteps)
{
    dgtd::TraceRange tr("dgtd_batch_step");
    if (!b || nsteps < 0) return dgtd::api_fail(DGTD_E_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        b->flush();
        for (int s = 0; s < nsteps; ++s) hcheck(hipGraphLaunch(b->graph, b->stream), "hipGraphLaunch");
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    return DGTD_OK;
}

int dgtd_batch_poll(dgtd_batch* b, int32_t* done_ids, int32_t cap, int32_t* ndone, int32_t* nrunning)
{
    if (!b) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        b->flush();
        b->sync_pull();
        int nd = 0, nr = 0;
        for (int i = 0; i < b->max_slots; ++i) {
            auto& s = b->slots[i];
            if (s.state == dgtd::kSlotRunning) ++nr;
            const bool ended = s.state == dgtd::kSlotDone || s.state == dgtd::kSlotFailed ||
                               s.state == dgtd::kSlotCancelled;
            if (ended && !s.reported && nd < cap) {
                if (done_ids) done_ids[nd] = i;
                s.reported = true;
                ++nd;
            }
        }
        if (ndone) *ndone = nd;
        if (nrunning) *nrunning = nr;
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    return DGTD_OK;
}

int dgtd_batch_info(dgtd_batch* b, int32_t id, double* t, int32_t* steps, int32_t* state, double* dt)
{
    if (!b || id < 0 || id >= b->max_slots) return dgtd::api_fail(DGTD_E_ARG, "bad instance id");
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        b->flush();
        b->sync_pull();
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    const auto& s = b->slots[id];
    if (t) *t = s.t;
    if (steps) *steps = s.steps;
    if (state) *state = s.state;
    if (dt) *dt = s.dt;
    return DGTD_OK;
}

int dgtd_batch_get_state(dgtd_batch* b, int32_t id, double* x)
{
    if (!b || !x || id < 0 || id >= b->max_slots) return dgtd::api_fail(DGTD_E_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        b->flush();
        b->sync_pull();
        const auto& s = b->slots[id];
        if (s.state == dgtd::kSlotFree || s.state == dgtd_batch::kSlotReserved)
            return dgtd::api_fail(DGTD_E_STATE, "instance slot is free or still being set up");
        const size_t ndofs = 6 * (size_t)nodes(s.dim, s.order) * s.K;
        if (b->precision == 64) {
            hcheck(hipMemcpy(x, s.buf[s.cur], sizeof(double) * ndofs, hipMemcpyDeviceToHost), "D2H state");
        } else {
            std::vector<float> tmp(ndofs);
            hcheck(hipMemcpy(tmp.data(), s.buf[s.cur], sizeof(float) * ndofs, hipMemcpyDeviceToHost), "D2H state");
            std::copy(tmp.begin(), tmp.end(), x);
        }
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    return DGTD_OK;
}

int dgtd_batch_cancel(dgtd_batch* b, int32_t id)
{
    if (!b || id < 0 || id >= b->max_slots) return dgtd::api_fail(DGTD_E_ARG, "bad instance id");
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        if (!b->dirty) b->sync_pull();  // device values are authoritative: fetch before editing
        else hcheck(hipStreamSynchronize(b->stream), "sync");
        auto& s = b->slots[id];
        if (s.state == dgtd::kSlotFree || s.state == dgtd_batch::kSlotReserved)
            return dgtd::api_fail(DGTD_E_STATE, "instance slot is free or still being set up");
        if (s.state == dgtd::kSlotRunning) {
            s.state = dgtd::kSlotCancelled;
            b->dirty = true;  // leaves the work list at the next step
        }
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    return DGTD_OK;
}

int dgtd_batch_remove(dgtd_batch* b, int32_t id)
{
    if (!b || id < 0 || id >= b->max_slots) return dgtd::api_fail(DGTD_E_ARG, "bad instance id");
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hcheck(hipSetDevice(b->device), "hipSetDevice");
        if (!b->dirty) b->sync_pull();
        else hcheck(hipStreamSynchronize(b->stream), "sync");
        auto& s = b->slots[id];
        if (s.state == dgtd::kSlotFree || s.state == dgtd_batch::kSlotReserved)
            return dgtd::api_fail(DGTD_E_STATE, "instance slot is free or still being set up");
        for (void*& p : s.buf)
            if (p) { hcheck(hipFree(p), "hipFree"); p = nullptr; }
        for (void*& p : s.tf_buf)
            if (p) { hcheck(hipFree(p), "hipFree"); p = nullptr; }
        b->used -= s.bytes;
        s = dgtd_batch::Slot{};
        b->dirty = true;
    } catch (const std::exception& e) {
        return batch_error(e);
    }
    return DGTD_OK;
}

int dgtd_batch_stats(dgtd_batch* b, int64_t* bytes_in_use, int32_t* running, int32_t* items, int32_t* grid)
{
    if (!b) return dgtd::api_fail(DGTD_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(b->mu);
    int nr = 0;
    for (const auto& s : b->slots) nr += s.state == dgtd::kSlotRunning;
    if (bytes_in_use) *bytes_in_use = b->used;
    if (running) *running = nr;
    if (items) *items = b->nitems;
    if (grid) *grid = b->grid;
    return DGTD_OK;
}

}  // extern "C"
