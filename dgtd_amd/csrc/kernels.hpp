// Device-kernel interface shared by the host code and the HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dgtd {

constexpr int kGeoStrideDev = 27;  // matches mesh.hpp kGeoStride
constexpr int kConnStrideDev = 8;  // matches mesh.hpp kConnStride
constexpr int kTriGeoStrideDev = 15;  // matches tri.hpp kTriGeoStride
constexpr int kTriConnStrideDev = 6;  // matches tri.hpp kTriConnStride

// Stage variants of the fused kernel.
enum StageMode : int {
    kModeMult = 0,   // out = L(qin)
    kModeLserk = 1,  // res = ca*res + dt*L(qin); out = qin + cb*res
    kModeRk1 = 2,    // k = L(x):  out(y) = x + ca k, res(z) = x + cb k   (qin = x)
    kModeRk2 = 3,    // k = L(y):  out(y') = x + ca k, res(z) += cb k
    kModeRk3 = 4,    // same as kModeRk2 with other coefficients
    kModeRk4 = 5,    // k = L(y):  out(x) = res(z) + ca k
};

template <typename T>
struct StageArgs {
    int K;
    int nblk;            // number of element blocks
    const T* qin;        // operator input, 6 blocks of Np*K
    T* out;              // see StageMode
    T* res;              // LSERK residual / RK4 accumulator z
    const T* x;          // RK4 base state (stages 2 and 3)
    T ca, cb, dt, alpha;
    const T* geo;        // kGeoStrideDev per element
    const int* conn;     // kConnStrideDev per element
    const T* mats;       // Dr^T, Ds^T, Dt^T ([j][i] each), then LIFT^T ([m][i])
    const int* tabs;     // Fmask[4*Nfp], then face permutation table [96*Nfp]
    const int* order;    // MFMA variant: processing order of the 16-element tasks (NULL = identity)
    // assembled (Global) operator, rows and columns in the state layout
    const int64_t* rowptr;
    const uint16_t* col16;  // element-local column codes (csr_local16)
    const int* etab;        // kCsrTab element ids per 16-DOF block
    const T* val;
    int64_t nrows, nnz;
    int np;  // nodes per element
    int* unstable;  // non-null: set to 1 when a stored state entry is unstable (unstable_entry)
};

// Stability test of the reference's time loop (Solver::run, Solver.cpp:499-516,
// flags a non-finite state norm or one above 1e20), applied to the entries a
// step's last stage stores (stability.hpp).
constexpr double kUnstableBound = 1e20;

// Device form of the assembled operator: the rows of DOFs [16 b, 16 b + 16)
// (all six components) reference at most kCsrTab elements, listed in
// etab[b * kCsrTab ...]; a column (component k, element etab[slot], node j)
// is stored as the 16-bit code slot << kCsrSlotShift | k << kCsrCompShift | j
// (slot < 32, nodes < 256: hexahedra up to order 5).
constexpr int kCsrRowsPerBlock = 16;
constexpr int kCsrTab = 32;
constexpr int kCsrCompShift = 8;
constexpr int kCsrSlotShift = 11;

// Stage over the assembled operator (csr.hip): row r of the product is the
// right-hand side of DOF r; the stage update as in the matrix-free kernels.
template <typename T>
hipError_t launch_csr_stage(const StageArgs<T>& a, int mode, hipStream_t s);

// Matrix-free hexahedron stage (hex_stage.hip): geo kHexGeoDev and conn
// kHexConnDev per element, mats = D1 (row-major (p+1)^2) then M1^-1 columns 0
// and p ((p+1) x 2), tabs = Fmask [6 Nfp], face permutations [48 Nfp], and per
// face the face node on each node's normal line [6 Np].
template <typename T>
hipError_t launch_hex_stage(int order, const StageArgs<T>& a, int mode, hipStream_t s);

template <typename T>
hipError_t launch_stage(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s);

// Sets the kernels' LDS limits and returns resident blocks per CU.
template <typename T>
int stage_blocks_per_cu(int order);

int elements_per_block(int order);

// Triangle (2D) stage kernel, dg_tri.hip: StageArgs with the triangle records
// (geo: 15 values, conn: 6 ints per element; mats: Dr^T, Ds^T, LIFT^T; tabs:
// Fmask[3*Nfp] then the edge table [18*Nfp]).
template <typename T>
hipError_t launch_tri_stage(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s);
int tri_elements_per_block(int order);
// Raises the dynamic-LDS limit of the triangle kernels of this order (at set-up,
// outside any stream capture).
template <typename T>
void tri_prepare(int order);
size_t tri_smem_bytes(int order, int precision);

// MFMA variant (mfma_stage.hpp, dg_mfma*.hip): wave-tasks of kMfmaTaskElems
// elements, waves_per_block waves, in one of kMfmaConfigs tuning
// configurations (waves per block, gather depth, residual prefetch).
constexpr int kMfmaTaskElems = 8;
// The matrix-core stage addresses each state buffer (6 N values) through
// 32-bit byte offsets: it runs while a buffer stays below 4 GiB (FP32: up to
// 178,956,970 nodes, e.g. 5.1 million tetrahedra at order 4).
inline bool mfma_fits_buffers(int64_t ndofs, int value_bytes)
{
    return ndofs * value_bytes < (int64_t(1) << 32);
}
constexpr int kMfmaConfigs = 2;
struct MfmaInfo {
    bool ok;              // the configuration's LDS tiles fit
    int waves;            // waves per block
    int gather_depth;     // flux k-steps of exterior traces in flight
    bool early_residual;  // first residual pass read before the volume products
};
template <int CFG, typename T>
MfmaInfo mfma_info_cfg(int order, int dim);
template <int CFG, typename T>
int mfma_prepare_cfg(int order, int dim);
template <int CFG, typename T>
hipError_t mfma_launch_cfg(int order, int dim, const StageArgs<T>& a, int mode, int grid, hipStream_t s);
template <typename T>
struct BatchArgs;
template <int CFG, typename T>
hipError_t batch_mfma_launch_cfg(int order, int dim, const BatchArgs<T>& b, const int* range, int grid,
                                 hipStream_t s);

template <typename T>
MfmaInfo mfma_info(int order, int dim, int cfg);
// mats layout: Bv [KP][3*NPG] then Bl [4*Nfp][NPG] (see pack in dgtd_api.cpp).
template <typename T>
hipError_t launch_stage_mfma(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s, int cfg);
// Sets LDS limits; returns resident blocks per CU, 0 if the variant does not fit.
template <typename T>
int mfma_blocks_per_cu(int order, int cfg);
template <typename T>
int mfma_waves_per_block(int order, int cfg);
// The same matrix-core stage on triangles (the 2D operator; records as in
// tri.hpp, mats Bv [KP][2*NPG] then Bl [3*NFPP][NPG]).
template <typename T>
hipError_t launch_tri_stage_mfma(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s, int cfg);
template <typename T>
int tri_mfma_blocks_per_cu(int order, int cfg);
template <typename T>
int tri_mfma_waves_per_block(int order, int cfg);

// ---------------------------------------------------------------------------
// Multi-instance batch (dg_batch.hip): many meshes of different size and order
// advance together; one persistent launch per LSERK stage walks a work list
// of (instance, element block) items sorted by order.
constexpr int kMaxBatchOrder = 5;

// kSlotFailed: retired by the stability test (unstable_entry) of a step's
// last stage; kSlotCancelled: stopped by the host (dgtd_batch_cancel).
enum BatchState : int { kSlotFree = 0, kSlotRunning = 1, kSlotDone = 2, kSlotFailed = 3, kSlotCancelled = 4 };

struct TfsfWave;

template <typename T>
struct BatchInst {
    int order, K, nblk, state;  // state: BatchState
    int cur, steps;             // cur: which q buffer holds the state
    int unstable;               // set by the last stage of a step (report_unstable)
    int dim;                    // 3 tetrahedra, 2 triangles
    T* q[2];
    T* res;
    const T* geo;
    const int* conn;
    double alpha, t, dt, dt_step, final_time;
    // elements owning TF/SF faces (TfsfArgs layout) and the plane waves
    const int* tf_elem;
    const int* tf_sign;
    const double* tf_xyz;
    const double* tf_geo;
    const TfsfWave* tf_waves;
    int tf_n, tf_nwaves;
};

struct BatchItem {
    int inst, blk;
};

template <typename T>
struct BatchArgs {
    BatchInst<T>* inst;
    const BatchItem* items;
    const int* count;  // number of valid items of the per-node kernel (device)
    const T* mats[kMaxBatchOrder + 1];       // per-node kernel layout
    const T* mats_mfma[kMaxBatchOrder + 1];  // matrix-core kernel layout
    const int* tabs[kMaxBatchOrder + 1];
    const T* mats_mfma_tri[kMaxBatchOrder + 1];  // the same for triangles
    const int* tabs_tri[kMaxBatchOrder + 1];
    int stage;
    T ca, cb;
};

template <typename T>
hipError_t launch_batch_stage(const BatchArgs<T>& b, int grid, size_t smem, hipStream_t s);
// Matrix-core batch stage for the instances of one element shape (dim 3
// tetrahedra, 2 triangles) and order: items (instance, 8-element task) in
// [range[0], range[1]) of b.items (range in device memory).
template <typename T>
hipError_t launch_batch_stage_mfma(int order, int dim, const BatchArgs<T>& b, const int* range, int grid,
                                   hipStream_t s, int cfg);
// Whether the matrix-core variant exists for this order, precision and
// configuration (its LDS tiles fit).
template <typename T>
bool mfma_available(int order, int cfg);
template <typename T>
hipError_t launch_batch_advance(BatchInst<T>* inst, int nslots, hipStream_t s);
// LDS bytes of the batch kernel when orders up to max_order are present,
// and resident blocks per CU at that size.
template <typename T>
size_t batch_smem_bytes(int max_order);
template <typename T>
int batch_blocks_per_cu(size_t smem);

// FP64 reductions over a state in T: per-block partials then a fixed-order
// final pass (deterministic, no atomics).
//   out[0] = sum of squares, out[1] = energy, out[2] = non-finite count.
// TF/SF plane-wave injection (tfsf.hip): plane wave with normalised
// polarisation and propagation, k x pol precomputed, Gaussian (kind 0) or
// modulated Gaussian (kind 1) magnitude of u = k.x - t.
struct TfsfWave {
    double pol[3], kxp[3], dir[3];
    double spread, mean, freq;
    int kind;
};

// Face-side slots per element in the TF/SF sign arrays (the hexahedron's six faces).
constexpr int kTfsfSignStride = 6;

template <typename T>
struct TfsfArgs {
    int dim;             // 3: tetrahedra (4 faces, 27-value records); 2: triangles (3 edges, 15 values);
                         // 4: quadrilaterals (4 edges, 18 values); 5: hexahedra (6 faces, 35 values)
    int nelem;           // elements owning TF/SF faces
    const int* elem;     // [nelem] element ids
    const int* sign;     // [kTfsfSignStride*nelem] per face: -1 scattered side, +1 total side, 0
    const double* xyz;   // [nelem][nfaces*Nfp][3] face node positions
    const double* geo;   // [nelem][record] FP64 geometry records
    const double* lift;  // [Np][nfaces*Nfp]
    const TfsfWave* waves;
    int nwaves;
    int Np, Nfp;
    size_t N;
    double alpha, t;
    T* out;
    double c_out;
    T* res;  // may be null
    double c_res;
};

template <typename T>
hipError_t launch_tfsf(const TfsfArgs<T>& a, hipStream_t s);

// TF/SF injection of a batch stage: items (instance, TF/SF element record)
// of the running instances, *count of them (device).
template <typename T>
struct BatchTfsfArgs {
    BatchInst<T>* inst;
    const BatchItem* items;
    const int* count;
    const double* lift[kMaxBatchOrder + 1];      // tetrahedra: LIFT [Np][4 Nfp]
    const double* lift_tri[kMaxBatchOrder + 1];  // triangles: LIFT [Np][3 Nfp]
    int stage;
    double c, b;  // the stage's time fraction and b coefficient
};
template <typename T>
hipError_t launch_batch_tfsf(const BatchTfsfArgs<T>& a, int grid, hipStream_t s);

// Elements per tile of the norms reduction: mass (Np^2), tile (6*TE*Np) and
// reduction scratch (3*256) doubles stay within 60 KiB of LDS.
inline int norms_tile(int Np)
{
    const int te = (61440 - 8 * Np * Np - 8 * 3 * 256) / (6 * 8 * Np);
    return te < 1 ? 1 : (te > 32 ? 32 : te);
}

// Partial-sum blocks of the norms reduction (the operators use at most this
// many), and the device address of the allocation-free entry point's static
// partials array on `device` (resolved once per device, with that device
// current).
constexpr int kNormsMaxParts = 1024;
hipError_t norms_scratch(int device, double** ptr);

template <typename T>
hipError_t launch_norms(const T* q, int K, int Np, const double* Mref, const double* jac,
                        const double* eps, const double* mu, double* partials, int nparts,
                        double* out3, hipStream_t s);

// Point-probe values: out[6*p + c] = sum_n w[p*Np + n] q[c*N + elem[p]*Np + n]
// (NaN where elem[p] < 0).  elem[p] < K and w hold Np weights per probe.
template <typename T>
hipError_t launch_probe_sample(const T* q, size_t N, int Np, int nprobes, const int* elem, const double* w,
                               double* out, hipStream_t s);

}  // namespace dgtd
