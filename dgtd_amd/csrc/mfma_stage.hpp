// MFMA (matrix-core) variant of the fused DGTD stage kernel for gfx950:
// the device code, templated on order, precision, element dimension and
// tuning configuration; dg_mfma_c0.hip and dg_mfma_c1.hip instantiate one
// configuration each, dg_mfma.hip dispatches at run time.
//
// Per element the stage needs, for the six right-hand sides r_c,
//     r_c = sum_k D_k w_{c,k}  +  LIFT f_c
// where D_k (k = r,s,t) are the reference derivative matrices, w_{c,k} the
// metric-weighted curl combinations of two field components
// (e.g. r_Hx: w_k = (g_kz E_y - g_ky E_z)/mu, g_kd = dr_k/dx_d) and f_c the
// fscale/2-scaled face flux (with 1/eps or 1/mu folded in).  Both terms are
// products with matrices shared by all elements, so a wave-task of 8 elements
// evaluates, for x = 0..2,
//     [r_Ex ; r_Hx]^T (16 x Np) = [W_x F_x]^T (16 x (3Np + 4Nfp)) * [D_r D_s D_t LIFT]^T
// with v_mfma_{f32,f64}_16x16x4: A rows 0-7 carry the E_x operands and rows
// 8-15 the H_x operands of the same 8 elements, columns = output nodes in
// groups of 16, K = input nodes (for D_k) and face nodes (for LIFT).
//   - A operands (W, F) are formed in registers from the LDS-staged fields:
//     lane (row l%16, slot l/16); for the LIFT a k-step covers four
//     consecutive nodes of one face, each lane gathering the neighbour trace
//     of its node and forming the flux there; fluxes never touch LDS or HBM.
//   - B operands (the reference matrices, zero-padded to 16-column groups and
//     to whole k-steps per face) stay in LDS for the life of the block.
//   - The accumulators are the finished right-hand sides; the stage update is
//     applied straight from them.
// Waves run independent tasks (only the initial matrix load is block-wide);
// tasks run in Morton order of their centroids, in contiguous ranges per XCD
// so neighbour re-reads stay in that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bc_coeffs.hpp"
#include "kernels.hpp"
#include "lds_copy.hpp"
#include "stability.hpp"

namespace dgtd {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<float> {
    using acc_t = f32x4;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // element row held by (slot q, register r): C[4q + r][l%16]
    static __device__ __forceinline__ int row(int q, int r) { return 4 * q + r; }
};

template <>
struct Mfma<double> {
    using acc_t = f64x4;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // C[q + 4r][l%16]
    static __device__ __forceinline__ int row(int q, int r) { return q + 4 * r; }
};

// Field accesses go through buffer descriptors with 32-bit byte offsets
// (one descriptor per state buffer, uniform across the wave): the per-lane
// address is a 32-bit offset, and the component and task bases ride in the
// scalar offset, so no per-lane 64-bit address arithmetic is issued.  Each
// state buffer (6 N values) must stay below 4 GiB (mfma_fits_buffers).
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kBufferFlags = 0x00020000;  // raw (untyped) 32-bit dword access on gfx9
constexpr int kAuxNontemporal = 2;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferFlags);
}

// AUX: the cache-policy bits (0, or kAuxNontemporal for data used once)
template <typename T, int AUX = 0>
__device__ __forceinline__ T bload(rsrc_t r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 4) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, AUX));
    } else {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX);
        return __longlong_as_double(((long long)v.y << 32) | v.x);
    }
}

template <int AUX>
__device__ __forceinline__ void bstore(float v, rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, AUX);
}
template <int AUX>
__device__ __forceinline__ void bstore(double v, rsrc_t r, uint32_t voff, uint32_t soff)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const u32x2 w = {(unsigned)b, (unsigned)(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)voff, (int)soff, AUX);
}

// A value known to be the same in every lane of the wave, moved to a scalar
// register (the first active lane's copy).
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename P>
__device__ __forceinline__ P* uniform(P* p)
{
    const unsigned long long b = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return reinterpret_cast<P*>(((unsigned long long)hi << 32) | lo);
}

// Element shape: tetrahedra (DIM 3) or triangles (DIM 2, the 2D operator:
// d/dz = 0, n_z = 0, edges as faces; records as tri.hpp lays them out).
template <int P, int DIM = 3>
struct MShape {
    static constexpr int NFACE = DIM == 3 ? 4 : 3;
    static constexpr int NDIR = DIM;  // reference derivative directions
    static constexpr int NP = DIM == 3 ? (P + 1) * (P + 2) * (P + 3) / 6 : (P + 1) * (P + 2) / 2;
    static constexpr int NFP = DIM == 3 ? (P + 1) * (P + 2) / 2 : P + 1;
    static constexpr int NF = NFACE * NFP;
    static constexpr int NPERM = DIM == 3 ? 96 : 18;  // rows of the neighbour-node table
    static constexpr int GEO = DIM == 3 ? kGeoStrideDev : kTriGeoStrideDev;
    static constexpr int CONN = DIM == 3 ? kConnStrideDev : kTriConnStrideDev;
    static constexpr int KP = (NP + 3) / 4 * 4;  // K padding of the volume product
    static constexpr int S = KP / 4;             // volume k-steps
    static constexpr int NPG = (NP + 15) / 16 * 16;
    static constexpr int G = NPG / 16;           // output-node groups
    static constexpr int US = KP + 1;            // LDS row stride of a task's fields (odd: KP % 4 == 0)
    static constexpr int NFPP = (NFP + 3) / 4 * 4;  // face nodes padded to whole k-steps
    static constexpr int SF = NFPP / 4;             // LIFT k-steps per face
};

constexpr int kTaskElems = kMfmaTaskElems;  // MFMA rows: 8 elements x {E, H} outputs

// Tuning configurations of the stage (chosen per operator by the autotuner).
// A task issues the exterior-trace gathers of its first flux k-steps, runs
// the volume products (which hide their latency), then the flux loop, which
// keeps the same number of k-steps in flight:
//   0 "wide": as many waves as LDS allows (at most 12 FP32 / 8 FP64, i.e. 3
//     or 2 waves per SIMD sharing its registers), 2 k-steps of traces in
//     flight, the residual read when the update starts;
//   1 "deep": at most 8 FP32 / 4 FP64 waves (twice the registers per wave),
//     16 FP32 / 4 FP64 k-steps in flight (all of a face set up to order 4 in
//     FP32), the first residual pass read during the flux loop.
template <int CFG>
struct MCfg {
    static constexpr int kMaxWaves32 = CFG == 0 ? 12 : 8;
    static constexpr int kMaxWaves64 = CFG == 0 ? 8 : 4;
    static constexpr int kGatherDepth32 = CFG == 0 ? 2 : 16;  // flux k-steps whose exterior traces are in flight
    static constexpr int kGatherDepth64 = CFG == 0 ? 2 : 4;
    static constexpr bool kEarlyResidual = CFG == 1;
    template <typename T>
    static constexpr int gather_depth() { return sizeof(T) == 4 ? kGatherDepth32 : kGatherDepth64; }
};

// Output nodes on the VALU instead of a mostly padded last MFMA column group:
// when the last group of 16 would hold only NT < 8 valid columns (order 3: 4,
// order 4: 3), summing those nodes per lane on the VALU costs 3 NT FMAs per
// lane and k-step against three MFMAs of 32 (FP32) or 64 (FP64) cycles.
// Not for FP64 at order 3, where the partial sums' registers make the stage
// spill (measured 840 -> 912 us on the 663,552-element order-3 box).
template <int P, typename T, int DIM>
struct MTail {
    using S = MShape<P, DIM>;
    static constexpr int NT = S::NP - 16 * (S::G - 1);
    static constexpr bool on = S::G >= 2 && NT < 8 && !(sizeof(T) == 8 && P == 3);
};

// Waves per block, chosen so the block's LDS fits in 160 KiB.  When the
// reference matrices and kMinLdsWaves waves' tiles do not fit together (FP64
// at order 5: 135 KiB of matrices), the matrices stay in global memory (read
// through the caches; 135 KiB is L2-resident) and LDS holds the node tables
// and the tiles only.
template <int P, typename T, int DIM = 3, int CFG = 0>
struct MWaves {
    // the VALU tail's partial sums need the registers of at most 2 waves per
    // SIMD (at 3 the FP32 order 3 and 4 stages spill)
    static constexpr int kMaxWaves = sizeof(T) == 4 ? (MTail<P, T, DIM>::on ? 8 : MCfg<CFG>::kMaxWaves32)
                                                    : MCfg<CFG>::kMaxWaves64;
    static constexpr int kMinLdsWaves = 4;
    using S = MShape<P, DIM>;
    static constexpr size_t bmats =
        sizeof(T) * ((size_t)S::KP * S::NDIR * S::NPG + (size_t)S::NFACE * S::NFPP * S::NPG);
    static constexpr size_t tables = sizeof(int) * (size_t)(S::NFACE * S::NFP + S::NPERM * S::NFP);
    static constexpr size_t per_wave = sizeof(T) * (6 * kTaskElems * (size_t)S::US + 2 * kTaskElems * (size_t)S::NP);
    static constexpr size_t budget = 160 * 1024;
    static constexpr bool bglobal = bmats + tables + kMinLdsWaves * per_wave > budget;
    static constexpr size_t mats = (bglobal ? 0 : bmats) + tables;
    static constexpr int fit = (int)((budget - mats) / per_wave);
    static constexpr int W = fit >= kMaxWaves ? kMaxWaves : (fit < 1 ? 1 : fit);
    static constexpr size_t bytes = mats + per_wave * W;
    static constexpr bool ok = bytes <= budget;
};

// Value of the lane 8 away within the 16-lane row (DPP row_ror:8 swaps the
// two halves of every row: rows l%16 < 8 and >= 8 of an MFMA A operand).
// (All rows and banks enabled and row_ror never leaves the row, so no lane
// keeps its old value: mov_dpp needs no initialised destination.)
__device__ __forceinline__ float dpp_swap8(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xF, 0xF, false));
}

__device__ __forceinline__ double dpp_swap8(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(bits & 0xffffffffLL), 0x128, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), 0x128, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-block LDS carve-up of the MFMA stage (reference matrices shared by the
// block, per-wave field and staging tiles).
template <int P, typename T, int DIM = 3>
struct MfmaLds {
    const T* sBv;  // [KP][3*NPG] (LDS, or global memory when MWaves::bglobal)
    const T* sBl;  // [4*NFPP][NPG], row f*NFPP + b
    T* sU;    // [WAVES][6][TE][US]
    T* sS;    // [WAVES][2][TE*NP] output staging
    int* sFm; // [4*NFP]
    int* sPt; // [96*NFP]
};

// Loads the reference matrices and node tables of order P into LDS and zeroes
// the field tiles (their padding columns j >= NP stay 0); block-wide.
template <int P, typename T, int DIM, int CFG>
__device__ MfmaLds<P, T, DIM> mfma_load_ref(const T* mats, const int* tabs, unsigned char* smem_raw)
{
    using S = MShape<P, DIM>;
    constexpr int NP = S::NP, NFP = S::NFP, KP = S::KP, NPG = S::NPG, US = S::US, NFPP = S::NFPP;
    constexpr int NDIR = S::NDIR, NFACE = S::NFACE;
    constexpr int TE = kTaskElems;
    constexpr int WAVES = MWaves<P, T, DIM, CFG>::W;
    constexpr bool kBGlobal = MWaves<P, T, DIM, CFG>::bglobal;
    MfmaLds<P, T, DIM> L;
    T* sB = reinterpret_cast<T*>(smem_raw);
    L.sU = kBGlobal ? sB : sB + KP * NDIR * NPG + NFACE * NFPP * NPG;
    L.sS = L.sU + WAVES * 6 * TE * US;
    L.sFm = reinterpret_cast<int*>(L.sS + WAVES * 2 * TE * NP);
    L.sPt = L.sFm + NFACE * NFP;
    const int tid = threadIdx.x;
    const int nthreads = WAVES * 64;
    if constexpr (kBGlobal) {
        L.sBv = mats;
    } else {
        lds_copy<8>(sB, mats, KP * NDIR * NPG + NFACE * NFPP * NPG, tid, nthreads);
        L.sBv = sB;
    }
    L.sBl = L.sBv + KP * NDIR * NPG;
    lds_copy<8>(L.sFm, tabs, NFACE * NFP + S::NPERM * NFP, tid, nthreads);
    for (int i = tid; i < WAVES * 6 * TE * US; i += nthreads) L.sU[i] = T(0);
    __syncthreads();
    return L;
}

// One wave-task: the fused stage of elements [task*8, task*8 + 8) of the mesh
// described by `a`.  Only wave-level synchronisation inside.
template <int P, typename T, int MODE, int DIM, int CFG>
__device__ __forceinline__ void mfma_task(const StageArgs<T>& a, const MfmaLds<P, T, DIM>& L, const int task)
{
    using S = MShape<P, DIM>;
    constexpr int NFACE = S::NFACE, NDIR = S::NDIR;
    using MM = Mfma<T>;
    using acc_t = typename MM::acc_t;
    constexpr int NP = S::NP, NFP = S::NFP, NPG = S::NPG, G = S::G, US = S::US;
    constexpr int NFPP = S::NFPP, SF = S::SF;
    constexpr int TE = kTaskElems;
    const T* sBv = L.sBv;
    const T* sBl = L.sBl;
    const int* sFm = L.sFm;
    const int* sPt = L.sPt;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int q = lane >> 4;    // k slot within a step
    const int li = lane & 15;   // A row (= output column within a node group)
    const int el = li & 7;      // element of the A row
    const int half = li >> 3;   // A rows 0-7: E outputs, 8-15: H outputs of the same elements
    T* U = L.sU + wave * 6 * TE * US;
    T* Sg = L.sS + wave * 2 * TE * NP;
    const int K = a.K;
    const size_t N = (size_t)NP * K;
    const int e0 = task * TE;
    const int ne = min(TE, K - e0);
    // byte offsets below 4 GiB (mfma_fits_buffers): component c of node n at
    // (c*N + n) * sizeof(T)
    constexpr uint32_t SZ = sizeof(T);
    const uint32_t bytes = (uint32_t)(6 * N * SZ);
    const uint32_t nb_bytes = (uint32_t)(N * SZ);  // one component
    const rsrc_t rq = make_rsrc(a.qin, bytes);

    // Every global load and store of a task is unconditional (out-of-range
    // lanes use a clamped, in-bounds index): with no branch around memory
    // instructions the compiler counts outstanding operations exactly, so a
    // wait for one load does not also wait for every store and prefetch
    // issued before it.

    // ---- stage the task's fields in LDS: all loads in flight (fields and the
    // lane's element record), then the LDS writes ----
    wave_sync();  // previous task's reads of U are done
    const bool valid = el < ne;
    const int e = e0 + (valid ? el : 0);
    const T* ge = a.geo + (size_t)e * S::GEO;
    T gm[9], gface[NFACE][4];  // metric terms dr_k/dx_d [k][d]; per face (nx, ny, nz, fscale/2)
    T scale;
    int fnb[NFACE], fcode[NFACE];
    {
        constexpr int NL = (TE * NP + 63) / 64;
        const int cnt = ne * NP;
        T v[6][NL];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const uint32_t so = c * nb_bytes + (uint32_t)e0 * NP * SZ;
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int t = lane + 64 * k;
                v[c][k] = bload<T>(rq, (uint32_t)(t < cnt ? t : cnt - 1) * SZ, so);
            }
        }
        if constexpr (DIM == 3) {
#pragma unroll
            for (int k = 0; k < 9; ++k) gm[k] = ge[k];
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int k = 0; k < 4; ++k) gface[f][k] = ge[9 + 4 * f + k];
        } else {  // [0..3] dr/dx dr/dy ds/dx ds/dy; edge f: [4+3f..6+3f] nx ny fscale/2
#pragma unroll
            for (int k = 0; k < 9; ++k) gm[k] = T(0);
            gm[0] = ge[0];
            gm[1] = ge[1];
            gm[3] = ge[2];
            gm[4] = ge[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                gface[f][0] = ge[4 + 3 * f];
                gface[f][1] = ge[5 + 3 * f];
                gface[f][2] = T(0);
                gface[f][3] = ge[6 + 3 * f];
            }
        }
        // r_E = (curl H + LIFT f_E)/eps, r_H = (-curl E + LIFT f_H)/mu
        const T sc = ge[S::GEO - 2 + half];
        scale = valid ? sc : T(0);
#pragma unroll
        for (int f = 0; f < NFACE; ++f) {
            fnb[f] = a.conn[(size_t)e * S::CONN + f];
            fcode[f] = a.conn[(size_t)e * S::CONN + NFACE + f];
        }
        // rows ne..TE-1 of a partial task receive copies of the last valid
        // value (only the rows of the padding elements read them); lanes past
        // TE*NP write the never-read last column of row 0 (US - 1 >= KP), so
        // the writes need no branch into which the compiler could sink loads
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const int t = lane + 64 * k;
                const int r = t / NP;
                U[t < TE * NP ? (c * TE + r) * US + (t - r * NP) : c * TE * US + US - 1] = v[c][k];
            }
    }
    wave_sync();

    // half 0 forms E-output operands, half 1 H-output
    const T sign = half ? T(-1) : T(1);
    const T sscale = sign * scale;
    T gms[9];  // the metric terms times the lane's sign and 1/eps or 1/mu
#pragma unroll
    for (int k = 0; k < 9; ++k) gms[k] = gm[k] * sscale;

    // Output nodes: GM groups of 16 on the matrix cores and, with MTail, the
    // last NT nodes summed per lane on the VALU (each lane over its own
    // k-slots, the four slots added at the end).
    constexpr bool kVTail = MTail<P, T, DIM>::on;
    constexpr int GM = kVTail ? G - 1 : G;
    constexpr int NT = kVTail ? MTail<P, T, DIM>::NT : 1;
    acc_t R[3][GM];
    T Rt[3][NT];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
#pragma unroll
        for (int g = 0; g < GM; ++g) R[x][g] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < NT; ++n) Rt[x][n] = T(0);
    }
    // B values of the tail nodes in row `row` of a reference matrix whose
    // rows are NPG long (the same for the lanes of a k-slot: a broadcast read)
    auto tail_b = [&](const T* row, T (&bt)[NT]) {
#pragma unroll
        for (int n = 0; n < NT; ++n) bt[n] = row[16 * GM + n];
    };

    // ---- surface -> LIFT (after the volume products, below) ----
    // Face by face; k-step t of face f covers nodes b = 4t + q, so the
    // lanes of an element read the same neighbour element.  A lane forms
    // the jumps of its own field only (E in rows 0-7, H in rows 8-15) and
    // takes the other field's jumps from its partner row (lane ^ 8, one
    // DPP row rotation).  Exterior traces are gathered from memory D k-steps
    // ahead, the first D before the volume products (a boundary face reads
    // the element's own trace, weighted by the
    // boundary coefficient minus one; HesthavenEvolution.cpp:290-292,
    // 309-311).  Only neighbours in another task are gathered, the others
    // read the task's fields in LDS; except for FP64 tetrahedra at orders 4
    // and 5, where every exterior trace is gathered (fewer selects; measured
    // 1.5 % faster at order 4, while at order 3 and on triangles the
    // remote-only form is 9-10 % faster in FP64).
    constexpr bool kRemoteOnly = sizeof(T) == 4 || DIM == 2 || P <= 3;
    const T* Ue = U + el * US;
    const int cbase = half ? 3 : 0;  // own field's first component
    auto face_remote = [&](int f) { return fnb[f] >= 0 && (unsigned)(fnb[f] - e0) >= (unsigned)ne; };
    auto face_tab = [&](int f) { return fnb[f] >= 0 ? sPt + fcode[f] * NFP : sFm + f * NFP; };
    // exterior trace of k-step s (own field's 3 components).  Every lane
    // issues the load (no branch); with kRemoteOnly, lanes without a remote
    // face all read the task's first value, one cache line, and discard it.
    auto gather = [&](int s, T (&v)[3]) {
        const int f = s / SF, b = 4 * (s % SF) + q;
        const int idx = face_tab(f)[b < NFP ? b : NFP - 1];
        uint32_t off;
        if constexpr (kRemoteOnly)
            off = face_remote(f) ? (uint32_t)fnb[f] * NP + idx : (uint32_t)e0 * NP;
        else
            off = (uint32_t)(fnb[f] >= 0 ? fnb[f] : e) * NP + idx;
        const uint32_t vo = (half ? 3 * nb_bytes : 0u) + off * SZ;
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = bload<T>(rq, vo, c * nb_bytes);
    };
    constexpr int kGatherDepth = MCfg<CFG>::template gather_depth<T>();
    constexpr int D = kGatherDepth < NFACE * SF ? kGatherDepth : NFACE * SF;
    T pf[D][3];
#pragma unroll
    for (int d = 0; d < D; ++d) gather(d, pf[d]);
    // The residual and base-state reads and the stage's stores are used once
    // per stage: they stream non-temporally, so L2 keeps the fields that
    // neighbouring tasks gather.
    // ---- stage update: per x, the E_x and H_x right-hand sides go through
    // the wave's LDS staging buffer and the update runs as a linear pass
    // over the task's contiguous chunk (whole sectors, coalesced).  The
    // residual / base state of the pair x+1 (linear index t over
    // E_x|H_x, t < 2*TE*NP) is read while pass x runs. ----
    constexpr int NLIN = (2 * TE * NP + 63) / 64;
    constexpr bool kNeedRes = MODE == kModeLserk || MODE == kModeRk2 || MODE == kModeRk3 || MODE == kModeRk4;
    constexpr bool kNeedX = MODE == kModeRk2 || MODE == kModeRk3;
    // Linear slot t (E_x | H_x halves) clamped to the last valid node of its
    // half: lanes past the end repeat that slot's load, arithmetic and store
    // (identical values to the same address).
    auto clamp_slot = [&](int t) {
        const int tc = t < 2 * TE * NP ? t : 2 * TE * NP - 1;
        const int h = tc >= TE * NP;
        const int u = tc - h * TE * NP;
        return h * TE * NP + (u < ne * NP ? u : ne * NP - 1);
    };
    // byte offset of linear slot tc within component x's pass: the lane part
    // (E_x or H_x half, node within the task) and the wave-uniform part
    auto lin_voff = [&](int tc) {
        const int h = tc >= TE * NP;
        return (h ? 3 * nb_bytes : 0u) + (uint32_t)(tc - h * TE * NP) * SZ;
    };
    auto lin_soff = [&](int x) { return x * nb_bytes + (uint32_t)e0 * NP * SZ; };
    const rsrc_t rres = make_rsrc(a.res, bytes);
    const rsrc_t rout = make_rsrc(a.out, bytes);
    const rsrc_t rx = make_rsrc(a.x, bytes);
    // the lane's linear slots, the same for every x: byte offsets (lane part)
    uint32_t lvo[NLIN];
#pragma unroll
    for (int k = 0; k < NLIN; ++k) lvo[k] = lin_voff(clamp_slot(lane + 64 * k));
    T rl[2][NLIN], xl[2][kNeedX ? NLIN : 1];
    auto fetch = [&](int x, int slot) {
#pragma unroll
        for (int k = 0; k < NLIN; ++k) {
            if constexpr (kNeedRes) rl[slot][k] = bload<T, kAuxNontemporal>(rres, lvo[k], lin_soff(x));
            if constexpr (kNeedX) xl[slot][k] = bload<T, kAuxNontemporal>(rx, lvo[k], lin_soff(x));
        }
    };
    // ---- volume: sum_k D_k w_k with w_k the metric-weighted curl terms ----
    // E_x out: (g_ky H_z - g_kz H_y)/eps;  H_x out: -(g_ky E_z - g_kz E_y)/mu
    // three k-steps per iteration; one with the VALU tail at order 3 and with
    // the matrices in global memory (loads hoisted over three k-steps spill)
    constexpr int kVolUnroll = (kVTail && P <= 3) || MWaves<P, T, DIM, CFG>::bglobal ? 1 : 3;
    // the lane's k-slot q covers input nodes j = q*S .. q*S + S-1 of the
    // other field (H for E outputs, E for H outputs)
    const T* usrc = Ue + (half ? 0 : 3) * TE * US + q * S::S;
    const T* bvq = sBv + q * S::S * NDIR * NPG;
#pragma unroll kVolUnroll
    for (int ks = 0; ks < S::S; ++ks) {
        T src[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) src[c] = usrc[c * TE * US + ks];
        const T* brow = bvq + ks * NDIR * NPG;
        const T* bv = brow + li;
#pragma unroll
        for (int kd = 0; kd < NDIR; ++kd) {
            const T* gk = gms + 3 * kd;
            T w[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                w[x] = gk[y] * src[z] - gk[z] * src[y];
            }
#pragma unroll
            for (int g = 0; g < GM; ++g) {
                const T bval = bv[kd * NPG + 16 * g];
#pragma unroll
                for (int x = 0; x < 3; ++x) R[x][g] = MM::mma(w[x], bval, R[x][g]);
            }
            if constexpr (kVTail) {
                T bt[NT];
                tail_b(brow + kd * NPG, bt);
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int n = 0; n < NT; ++n) Rt[x][n] += w[x] * bt[n];
            }
        }
    }

    if constexpr (MCfg<CFG>::kEarlyResidual) fetch(0, 0);  // in flight during the flux loop

#pragma unroll
    for (int f = 0; f < NFACE; ++f) {
        const int nb = fnb[f];
        const bool interior = nb >= 0;
        const bool remote = face_remote(f);
        const int* tab = face_tab(f);
        const T* Ub = U + (interior && !remote ? nb - e0 : el) * US + cbase * TE * US;
        T cown;
        {
            T ce, ch;
            bc_coeffs<T>(-nb - 1, a.alpha, ce, ch);
            cown = interior ? T(-1) : (half ? ch : ce) - T(1);
        }
        const T n3[3] = {gface[f][0], gface[f][1], gface[f][2]};
        const T fsc = gface[f][3] * scale;
        // the face's flux coefficients, fscale/2 and 1/eps or 1/mu folded in:
        // F_x = sn_y A_z - sn_z A_y + aw B_x - an_x (n.B)
        T sn[3], an[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sn[d] = sign * fsc * n3[d];
            an[d] = a.alpha * fsc * n3[d];
        }
        const T aw = a.alpha * fsc;
#pragma unroll
        for (int t = 0; t < SF; ++t) {
            const int s = f * SF + t;
            const int b = 4 * t + q;
            const int bb = b < NFP ? b : NFP - 1;
            const int iM = sFm[f * NFP + bb];
            T cur[3];
            if constexpr (kRemoteOnly) {
                // the empty asm keeps the LDS values register operands, so the
                // compiler cannot fold the select into one load from a generic
                // pointer here (which would undo the prefetch); it is not
                // volatile, so it does not count as a memory side effect
                const int ip = tab[bb];
                T lv[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) lv[c] = Ub[c * TE * US + ip];
                asm("" : "+v"(lv[0]), "+v"(lv[1]), "+v"(lv[2]));
#pragma unroll
                for (int c = 0; c < 3; ++c) cur[c] = remote ? pf[s % D][c] : lv[c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) cur[c] = pf[s % D][c];
            }
            if (s + D < NFACE * SF) gather(s + D, pf[s % D]);
            __builtin_amdgcn_sched_barrier(0);  // the prefetch stays D k-steps ahead of its use
            T B3[3], A3[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                B3[c] = cur[c] + cown * Ue[(cbase + c) * TE * US + iM];
                A3[c] = dpp_swap8(B3[c]);  // partner row's jump: the other field
            }
            const T nB = n3[0] * B3[0] + n3[1] * B3[1] + n3[2] * B3[2];
            // E out: n x dH + alpha (dE - n (n.dE));  H out: -n x dE + alpha (dH - n (n.dH)),
            // times fscale/2 and 1/eps or 1/mu.  Padded face slots (b >= NFP) meet
            // zero LIFT rows, so their (finite, clamped-node) flux adds nothing.
            T F[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                F[x] = sn[y] * A3[z] - sn[z] * A3[y] + aw * B3[x] - an[x] * nB;
            }
            const T* bl = sBl + (f * NFPP + b) * NPG + li;
#pragma unroll
            for (int g = 0; g < GM; ++g) {
                const T bv = bl[16 * g];
#pragma unroll
                for (int x = 0; x < 3; ++x) R[x][g] = MM::mma(F[x], bv, R[x][g]);
            }
            if constexpr (kVTail) {
                T bt[NT];
                tail_b(sBl + (f * NFPP + b) * NPG, bt);
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int n = 0; n < NT; ++n) Rt[x][n] += F[x] * bt[n];
            }
        }
    }

    if constexpr (!MCfg<CFG>::kEarlyResidual) fetch(0, 0);
    // and their staging-buffer index and field-tile index of component x = 0
    // (E_x or H_x row of the slot's element; component x adds x * TE * US)
    int lst[NLIN], lqv[NLIN];
#pragma unroll
    for (int k = 0; k < NLIN; ++k) {
        const int t = clamp_slot(lane + 64 * k);
        const int h = t >= TE * NP;
        const unsigned u = (unsigned)(t - h * TE * NP);
        const unsigned r_el = u / NP;
        lst[k] = t;
        lqv[k] = (int)(((h ? 3u : 0u) * TE + r_el) * US + (u - r_el * NP));
    }
    bool bad = false;  // an unstable state entry stored by this lane (a.unstable set)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const int cur = x & 1;
#pragma unroll
        for (int g = 0; g < GM; ++g) {
            const int i = 16 * g + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = MM::row(q, r);
                if (i < NP) Sg[(row >> 3) * TE * NP + (row & 7) * NP + i] = R[x][g][r];
            }
        }
        if constexpr (kVTail) {
            // row li's tail nodes: the four k-slot partial sums added (the
            // lanes li, li+16, li+32, li+48), node n stored by slot n % 4
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                T v = Rt[x][n];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (n % 4 == q) Sg[(li >> 3) * TE * NP + (li & 7) * NP + 16 * GM + n] = v;
            }
        }
        wave_sync();
        if (x + 1 < 3) fetch(x + 1, cur ^ 1);
#pragma unroll
        for (int k = 0; k < NLIN; ++k) {
            const uint32_t vo = lvo[k];
            const uint32_t so = lin_soff(x);
            const T rhs = Sg[lst[k]];
            const T qv = U[lqv[k] + x * TE * US];
            constexpr int nt = kAuxNontemporal;
            if constexpr (MODE == kModeMult) {
                bstore<nt>(rhs, rout, vo, so);
            } else if constexpr (MODE == kModeLserk) {
                const T rr = a.ca * rl[cur][k] + a.dt * rhs;
                const T qn = qv + a.cb * rr;
                bstore<nt>(rr, rres, vo, so);
                bstore<nt>(qn, rout, vo, so);
                bad = bad || unstable_entry(qn);
            } else if constexpr (MODE == kModeRk1) {
                bstore<nt>(qv + a.ca * rhs, rout, vo, so);
                bstore<nt>(qv + a.cb * rhs, rres, vo, so);
            } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
                bstore<nt>(xl[cur][k] + a.ca * rhs, rout, vo, so);
                bstore<nt>(rl[cur][k] + a.cb * rhs, rres, vo, so);
            } else {
                const T qn = rl[cur][k] + a.ca * rhs;
                bstore<nt>(qn, rout, vo, so);
                bad = bad || unstable_entry(qn);
            }
        }
        wave_sync();  // staging buffer free for the next x
    }
    if constexpr (MODE == kModeLserk || MODE == kModeRk4) report_unstable(a.unstable, bad);
}

template <int P, typename T, int MODE, int DIM, int CFG>
__global__ __launch_bounds__((MWaves<P, T, DIM, CFG>::W * 64)) void dg_stage_mfma(StageArgs<T> a)
{
    constexpr int TE = kTaskElems;
    constexpr int WAVES = MWaves<P, T, DIM, CFG>::W;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const MfmaLds<P, T, DIM> L = mfma_load_ref<P, T, DIM, CFG>(a.mats, a.tabs, smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int ntasks = (a.K + TE - 1) / TE;
    // XCD-contiguous task ranges: block b runs on XCD b % 8.
    const int nxcd = (gridDim.x >= 8 && gridDim.x % 8 == 0) ? 8 : 1;
    const int xcd = blockIdx.x % nxcd;
    const int bpx = gridDim.x / nxcd;
    const int bin = blockIdx.x / nxcd;
    const int per_xcd = (ntasks + nxcd - 1) / nxcd;
    const int t_begin = xcd * per_xcd;
    const int t_end = min(ntasks, t_begin + per_xcd);
    for (int tix = t_begin + bin * WAVES + wave; tix < t_end; tix += bpx * WAVES)
        mfma_task<P, T, MODE, DIM, CFG>(a, L, a.order ? a.order[tix] : tix);
}

// Batch variant: the running instances of element shape DIM and order P, one
// (instance, task) work item per wave-task, items [range[0], range[1]) of the
// list; stage `b.stage` of LSERK4 with the instance's own dt, buffers and
// geometry.
template <int P, typename T, int DIM, int CFG>
__global__ __launch_bounds__((MWaves<P, T, DIM, CFG>::W * 64)) void dg_batch_stage_mfma(BatchArgs<T> b,
                                                                                        const int* range)
{
    constexpr int WAVES = MWaves<P, T, DIM, CFG>::W;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const MfmaLds<P, T, DIM> L = mfma_load_ref<P, T, DIM, CFG>(DIM == 3 ? b.mats_mfma[P] : b.mats_mfma_tri[P],
                                                               DIM == 3 ? b.tabs[P] : b.tabs_tri[P], smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int begin = range[0], end = range[1];
    for (int it = begin + blockIdx.x * WAVES + wave; it < end; it += gridDim.x * WAVES) {
        // The item and the instance record are the same for the whole wave
        // but come through vector loads (the records are written on the
        // device); made wave-uniform (scalar registers) so that the task's
        // buffer descriptors and offsets are scalar, without waterfall loops.
        const BatchItem w = b.items[it];
        const int inst = uniform(w.inst);
        const int blk = uniform(w.blk);
        const BatchInst<T>& I = b.inst[inst];
        if (uniform(I.state) != kSlotRunning) continue;
        StageArgs<T> a{};
        a.K = uniform(I.K);
        const int src = uniform(I.cur) ^ (b.stage & 1);
        a.qin = uniform(I.q[src]);
        a.out = uniform(I.q[src ^ 1]);
        a.res = uniform(I.res);
        a.ca = b.ca;
        a.cb = b.cb;
        a.dt = (T)uniform(I.dt_step);
        a.alpha = (T)uniform(I.alpha);
        a.geo = uniform(I.geo);
        a.conn = uniform(I.conn);
        a.unstable = b.stage == 4 ? &b.inst[inst].unstable : nullptr;  // the step's last stage
        mfma_task<P, T, kModeLserk, DIM, CFG>(a, L, blk);
    }
}

template <int P, typename T, int MODE, int DIM, int CFG>
void mfma_set_attr()
{
    if constexpr (!MWaves<P, T, DIM, CFG>::ok) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_stage_mfma<P, T, MODE, DIM, CFG>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)MWaves<P, T, DIM, CFG>::bytes);
}

// Sets the LDS attribute of every mode (set-up time, outside any capture)
// and returns the resident blocks per CU (0: no variant for this order).
template <int P, typename T, int DIM, int CFG>
int mfma_prepare()
{
    using W = MWaves<P, T, DIM, CFG>;
    if constexpr (!W::ok) return 0;
    mfma_set_attr<P, T, kModeMult, DIM, CFG>();
    mfma_set_attr<P, T, kModeLserk, DIM, CFG>();
    mfma_set_attr<P, T, kModeRk1, DIM, CFG>();
    mfma_set_attr<P, T, kModeRk2, DIM, CFG>();
    mfma_set_attr<P, T, kModeRk3, DIM, CFG>();
    mfma_set_attr<P, T, kModeRk4, DIM, CFG>();
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_batch_stage_mfma<P, T, DIM, CFG>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)W::bytes);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dg_stage_mfma<P, T, kModeLserk, DIM, CFG>, W::W * 64,
                                                     W::bytes) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int P, typename T, int DIM, int CFG>
hipError_t mfma_launch(const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    using W = MWaves<P, T, DIM, CFG>;
    constexpr int threads = W::W * 64;
    constexpr size_t bytes = W::bytes;
    if constexpr (!W::ok) return hipErrorInvalidConfiguration;
    switch (mode) {
        case kModeMult: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeMult, DIM, CFG>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeLserk: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeLserk, DIM, CFG>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk1: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk1, DIM, CFG>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk2: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk2, DIM, CFG>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk3: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk3, DIM, CFG>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk4: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk4, DIM, CFG>), dim3(grid), dim3(threads), bytes, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int P, typename T, int DIM, int CFG>
hipError_t batch_mfma_launch(const BatchArgs<T>& b, const int* range, int grid, hipStream_t s)
{
    using W = MWaves<P, T, DIM, CFG>;
    if constexpr (!W::ok) {
        return hipErrorInvalidConfiguration;
    } else {
        hipLaunchKernelGGL((dg_batch_stage_mfma<P, T, DIM, CFG>), dim3(grid), dim3(W::W * 64), W::bytes, s, b,
                           range);
        return hipGetLastError();
    }
}

}  // namespace

// Per-configuration entry points (one translation unit per CFG); the
// order switch maps a run-time order onto the template.
#define DGTD_ORDER_SWITCH(ORDER, EXPR_OF_P, FALLBACK) \
    switch (ORDER) {                                  \
        case 1: return EXPR_OF_P(1);                  \
        case 2: return EXPR_OF_P(2);                  \
        case 3: return EXPR_OF_P(3);                  \
        case 4: return EXPR_OF_P(4);                  \
        case 5: return EXPR_OF_P(5);                  \
        default: return FALLBACK;                     \
    }

template <int CFG, typename T>
MfmaInfo mfma_info_cfg(int order, int dim)
{
#define DGTD_INFO(P)                                                                                       \
    (dim == 2 ? MfmaInfo{MWaves<P, T, 2, CFG>::ok, MWaves<P, T, 2, CFG>::W, MCfg<CFG>::template gather_depth<T>(),        \
                         MCfg<CFG>::kEarlyResidual}                                                          \
              : MfmaInfo{MWaves<P, T, 3, CFG>::ok, MWaves<P, T, 3, CFG>::W, MCfg<CFG>::template gather_depth<T>(),        \
                         MCfg<CFG>::kEarlyResidual})
    DGTD_ORDER_SWITCH(order, DGTD_INFO, (MfmaInfo{false, 1, 0, false}))
#undef DGTD_INFO
}

template <int CFG, typename T>
int mfma_prepare_cfg(int order, int dim)
{
#define DGTD_PREP(P) (dim == 2 ? mfma_prepare<P, T, 2, CFG>() : mfma_prepare<P, T, 3, CFG>())
    DGTD_ORDER_SWITCH(order, DGTD_PREP, 0)
#undef DGTD_PREP
}

template <int CFG, typename T>
hipError_t mfma_launch_cfg(int order, int dim, const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
#define DGTD_LAUNCH(P) (dim == 2 ? mfma_launch<P, T, 2, CFG>(a, mode, grid, s) : mfma_launch<P, T, 3, CFG>(a, mode, grid, s))
    DGTD_ORDER_SWITCH(order, DGTD_LAUNCH, hipErrorInvalidValue)
#undef DGTD_LAUNCH
}

template <int CFG, typename T>
hipError_t batch_mfma_launch_cfg(int order, int dim, const BatchArgs<T>& b, const int* range, int grid,
                                 hipStream_t s)
{
#define DGTD_BLAUNCH(P) \
    (dim == 2 ? batch_mfma_launch<P, T, 2, CFG>(b, range, grid, s) : batch_mfma_launch<P, T, 3, CFG>(b, range, grid, s))
    DGTD_ORDER_SWITCH(order, DGTD_BLAUNCH, hipErrorInvalidValue)
#undef DGTD_BLAUNCH
}

#define DGTD_INSTANTIATE_CFG(CFG, T)                                                                          \
    template MfmaInfo mfma_info_cfg<CFG, T>(int, int);                                                       \
    template int mfma_prepare_cfg<CFG, T>(int, int);                                                         \
    template hipError_t mfma_launch_cfg<CFG, T>(int, int, const StageArgs<T>&, int, int, hipStream_t);       \
    template hipError_t batch_mfma_launch_cfg<CFG, T>(int, int, const BatchArgs<T>&, const int*, int, hipStream_t);

}  // namespace dgtd
