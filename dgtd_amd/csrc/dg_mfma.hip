// MFMA (matrix-core) variant of the fused DGTD stage kernel for gfx950.
//
// Per element the stage needs, for the six right-hand sides r_c,
//     r_c = sum_k D_k w_{c,k}  +  LIFT f_c
// where D_k (k = r,s,t) are the reference derivative matrices, w_{c,k} the
// metric-weighted curl combinations of two field components
// (e.g. r_Hx: w_k = (g_kz E_y - g_ky E_z)/mu, g_kd = dr_k/dx_d) and f_c the
// fscale/2-scaled face flux (with 1/eps or 1/mu folded in).  Both terms are
// products with matrices shared by all elements, so for a wave-task of 16
// elements the kernel evaluates
//     R^T (16 x Np) = [W_r W_s W_t F]^T (16 x (3Np + 4Nfp)) * [D_r D_s D_t LIFT]^T
// with v_mfma_{f32,f64}_16x16x4: rows = the 16 elements, columns = output
// nodes in groups of 16, K = input nodes (for D_k) and face nodes (for LIFT).
//   - A operands (W, F) are formed in registers from the LDS-staged fields:
//     lane (element l%16, slot l/16) -- for the LIFT the slot is the face, so
//     each lane gathers the neighbour trace and forms the flux of one face
//     node per k-step; fluxes never touch LDS or HBM.
//   - B operands (the reference matrices, zero-padded to 16-column groups)
//     stay in LDS for the life of the block.
//   - The accumulators are the finished right-hand sides; the stage update is
//     applied straight from them.
// Waves run independent tasks (only the initial matrix load is block-wide);
// blocks take contiguous task ranges per XCD so neighbour re-reads stay in
// that XCD's L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

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

template <int P>
struct MShape {
    static constexpr int NP = (P + 1) * (P + 2) * (P + 3) / 6;
    static constexpr int NFP = (P + 1) * (P + 2) / 2;
    static constexpr int NF = 4 * NFP;
    static constexpr int KP = (NP + 3) / 4 * 4;  // K padding of the volume product
    static constexpr int S = KP / 4;             // volume k-steps
    static constexpr int NPG = (NP + 15) / 16 * 16;
    static constexpr int G = NPG / 16;           // output-node groups
    static constexpr int US = (KP % 2 == 0) ? KP + 1 : KP;  // LDS row stride of a task's fields
};

// Waves per block, chosen so the block's LDS fits in 160 KiB.
template <int P, typename T>
struct MWaves {
    using S = MShape<P>;
    static constexpr size_t mats = sizeof(T) * ((size_t)S::KP * 3 * S::NPG + (size_t)S::NF * S::NPG) +
                                   sizeof(int) * (size_t)(4 * S::NFP + 96 * S::NFP);
    static constexpr size_t per_wave = sizeof(T) * 6 * 16 * (size_t)S::US;
    static constexpr size_t budget = 160 * 1024;
    static constexpr int fit = (int)((budget - mats) / per_wave);
    static constexpr int W = fit >= 8 ? 8 : (fit < 1 ? 1 : fit);
    static constexpr size_t bytes = mats + per_wave * W;
    static constexpr bool ok = bytes <= budget;  // FP64 at order 5 does not fit
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int P, typename T, int MODE>
__global__ __launch_bounds__((MWaves<P, T>::W * 64)) void dg_stage_mfma(StageArgs<T> a)
{
    using S = MShape<P>;
    using MM = Mfma<T>;
    using acc_t = typename MM::acc_t;
    constexpr int NP = S::NP, NFP = S::NFP, NF = S::NF, KP = S::KP, NPG = S::NPG, G = S::G, US = S::US;
    constexpr int WAVES = MWaves<P, T>::W;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sBv = reinterpret_cast<T*>(smem_raw);     // [KP][3*NPG]
    T* sBl = sBv + KP * 3 * NPG;                 // [NF][NPG]
    T* sU = sBl + NF * NPG;                      // [WAVES][6][16][US]
    int* sFm = reinterpret_cast<int*>(sU + WAVES * 6 * 16 * US);  // [4*NFP]
    int* sPt = sFm + 4 * NFP;                                     // [96*NFP]

    const int tid = threadIdx.x;
    const int nthreads = WAVES * 64;
    for (int i = tid; i < KP * 3 * NPG + NF * NPG; i += nthreads) sBv[i] = a.mats[i];
    for (int i = tid; i < 4 * NFP + 96 * NFP; i += nthreads) sFm[i] = a.tabs[i];
    for (int i = tid; i < WAVES * 6 * 16 * US; i += nthreads) sU[i] = T(0);  // padding j >= NP stays 0
    __syncthreads();

    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int q = lane >> 4;   // slot: k within a step / face for the LIFT
    const int li = lane & 15;  // A-row element / output node within a group
    T* U = sU + wave * 6 * 16 * US;

    const int K = a.K;
    const size_t N = (size_t)NP * K;
    const int ntasks = (K + 15) / 16;
    // XCD-contiguous task ranges: block b runs on XCD b % 8.
    const int nxcd = (gridDim.x >= 8 && gridDim.x % 8 == 0) ? 8 : 1;
    const int xcd = blockIdx.x % nxcd;
    const int bpx = gridDim.x / nxcd;  // blocks per XCD (grid is a multiple of nxcd)
    const int bin = blockIdx.x / nxcd;
    const int per_xcd = (ntasks + nxcd - 1) / nxcd;
    const int t_begin = xcd * per_xcd;
    const int t_end = min(ntasks, t_begin + per_xcd);

    for (int tix = t_begin + bin * WAVES + wave; tix < t_end; tix += bpx * WAVES) {
        const int task = a.order ? a.order[tix] : tix;
        const int e0 = task * 16;
        const int ne = min(16, K - e0);

        // ---- stage the 16 elements' fields in LDS (coalesced per component) ----
        // All loads of the task are issued before the first LDS write so the
        // wave has 6*ceil(16*Np/64) requests in flight.
        wave_sync();  // previous task's reads of U are done
        {
            constexpr int NL = (16 * NP + 63) / 64;
            const int cnt = ne * NP;
            T v[6][NL];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const T* src = a.qin + c * N + (size_t)e0 * NP;
#pragma unroll
                for (int k = 0; k < NL; ++k) {
                    const int t = lane + 64 * k;
                    v[c][k] = (t < cnt) ? src[t] : T(0);
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int k = 0; k < NL; ++k) {
                    const int t = lane + 64 * k;
                    if (t < cnt) {
                        const int el = t / NP;
                        U[(c * 16 + el) * US + (t - el * NP)] = v[c][k];
                    }
                }
        }
        wave_sync();

        // ---- per-lane element (li) and face (q) data ----
        const bool valid = li < ne;
        const int e = e0 + (valid ? li : 0);
        const T* ge = a.geo + (size_t)e * kGeoStrideDev;
        T gm[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) gm[k] = ge[k];
        const T ieps = ge[25], imu = ge[26];
        const T nx = ge[9 + 4 * q], ny = ge[10 + 4 * q], nz = ge[11 + 4 * q], fs = ge[12 + 4 * q];
        const int nb = a.conn[(size_t)e * kConnStrideDev + q];
        const int code = a.conn[(size_t)e * kConnStrideDev + 4 + q];
        T ce = T(0), ch = T(0);  // boundary jump coefficients (HesthavenEvolution.cpp:290-292)
        if (nb < 0) {
            const int bc = -nb - 1;
            if (bc == 1) ce = T(-2);
            else if (bc == 2) ch = T(-2);
            else if (bc == 3) { ce = T(-1) / a.alpha; ch = ce; }
        }
        const T sE = valid ? fs * ieps : T(0);  // fold 1/eps, 1/mu and fscale/2 into F
        const T sH = valid ? fs * imu : T(0);

        acc_t R[6][G];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int g = 0; g < G; ++g) R[c][g] = acc_t{0, 0, 0, 0};

        // ---- surface: flux of face q, node b of element li -> LIFT ----
        // Neighbour traces: from LDS when the neighbour is in this task, else
        // a global gather (prefetched two face nodes ahead); boundary faces
        // use the element's own trace with the boundary coefficients.
        const T* Ue = U + li * US;
        const bool interior = nb >= 0;
        const bool local = interior && (unsigned)(nb - e0) < (unsigned)ne;
        const bool remote = interior && !local;
        const T* qP = a.qin + (size_t)(remote ? nb : e) * NP;
        const T* Un = U + (local ? nb - e0 : li) * US;
        const int* pmap = interior ? sPt + code * NFP : sFm + q * NFP;
        const T cI = interior ? T(1) : T(0);  // weight of the neighbour trace
        const T cE = interior ? T(-1) : ce;   // weight of the own trace, E components
        const T cH = interior ? T(-1) : ch;   // weight of the own trace, H components
        T pf0[6], pf1[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) { pf0[c] = T(0); pf1[c] = T(0); }
        if (remote) {
            const int i0 = pmap[0], i1 = pmap[NFP > 1 ? 1 : 0];
#pragma unroll
            for (int c = 0; c < 6; ++c) { pf0[c] = qP[c * N + i0]; pf1[c] = qP[c * N + i1]; }
        }
#pragma unroll
        for (int b = 0; b < NFP; ++b) {
            const int ip = pmap[b];
            T cur[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                cur[c] = remote ? pf0[c] : Un[c * 16 * US + ip];
                pf0[c] = pf1[c];
            }
            if (remote && b + 2 < NFP) {
                const int i2 = pmap[b + 2];
#pragma unroll
                for (int c = 0; c < 6; ++c) pf1[c] = qP[c * N + i2];
            }
            const int iM = sFm[q * NFP + b];
            T d[6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                d[c] = cI * cur[c] + cE * Ue[c * 16 * US + iM];
                d[3 + c] = cI * cur[3 + c] + cH * Ue[(3 + c) * 16 * US + iM];
            }
            const T n3[3] = {nx, ny, nz};
            const T ndH = nx * d[3] + ny * d[4] + nz * d[5];
            const T ndE = nx * d[0] + ny * d[1] + nz * d[2];
            T F[6];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                F[3 + x] = (-n3[y] * d[z] + n3[z] * d[y] + a.alpha * (d[3 + x] - ndH * n3[x])) * sH;
                F[x] = (n3[y] * d[3 + z] - n3[z] * d[3 + y] + a.alpha * (d[x] - ndE * n3[x])) * sE;
            }
            const T* bl = sBl + (q * NFP + b) * NPG + li;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const T bv = bl[16 * g];
#pragma unroll
                for (int c = 0; c < 6; ++c) R[c][g] = MM::mma(F[c], bv, R[c][g]);
            }
        }

        // ---- volume: sum_k D_k w_{c,k} ----
        const T wE = valid ? ieps : T(0), wH = valid ? imu : T(0);
#pragma unroll 3
        for (int ks = 0; ks < S::S; ++ks) {
            const int j = q * S::S + ks;
            T u[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) u[c] = Ue[c * 16 * US + j];
            const T* bv = sBv + j * 3 * NPG + li;
#pragma unroll
            for (int kd = 0; kd < 3; ++kd) {
                const T* gk = gm + 3 * kd;  // dr_kd/dx_d, d = 0..2
                T w[6];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int y = (x + 1) % 3, z = (x + 2) % 3;
                    w[3 + x] = (gk[z] * u[y] - gk[y] * u[z]) * wH;          // H_x: -D_y E_z + D_z E_y
                    w[x] = (gk[y] * u[3 + z] - gk[z] * u[3 + y]) * wE;      // E_x:  D_y H_z - D_z H_y
                }
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const T bval = bv[kd * NPG + 16 * g];
#pragma unroll
                    for (int c = 0; c < 6; ++c) R[c][g] = MM::mma(w[c], bval, R[c][g]);
                }
            }
        }

        // ---- stage update straight from the accumulators ----
        // Per output-node group: every global read of the group is issued
        // before its first store (reads and stores may alias as far as the
        // compiler knows, which would otherwise serialise them pairwise).
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int i = 16 * g + li;
            size_t o[4];
            bool ok[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int el = MM::row(q, r);
                ok[r] = (i < NP) && (el < ne);
                o[r] = (size_t)(e0 + (ok[r] ? el : 0)) * NP + (ok[r] ? i : 0);
            }
            T rv[6][4], xv[6][4];
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if constexpr (MODE == kModeLserk || MODE == kModeRk2 || MODE == kModeRk3 || MODE == kModeRk4)
                        rv[c][r] = ok[r] ? a.res[c * N + o[r]] : T(0);
                    if constexpr (MODE == kModeRk2 || MODE == kModeRk3)
                        xv[c][r] = ok[r] ? a.x[c * N + o[r]] : T(0);
                }
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (!ok[r]) continue;
                    const T rhs = R[c][g][r];
                    const size_t oo = c * N + o[r];
                    const T qv = U[(c * 16 + MM::row(q, r)) * US + i];
                    if constexpr (MODE == kModeMult) {
                        a.out[oo] = rhs;
                    } else if constexpr (MODE == kModeLserk) {
                        const T rr = a.ca * rv[c][r] + a.dt * rhs;
                        a.res[oo] = rr;
                        a.out[oo] = qv + a.cb * rr;
                    } else if constexpr (MODE == kModeRk1) {
                        a.out[oo] = qv + a.ca * rhs;
                        a.res[oo] = qv + a.cb * rhs;
                    } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
                        a.out[oo] = xv[c][r] + a.ca * rhs;
                        a.res[oo] = rv[c][r] + a.cb * rhs;
                    } else {
                        a.out[oo] = rv[c][r] + a.ca * rhs;
                    }
                }
        }
    }
}

template <int P, typename T, int MODE>
void mfma_set_attr()
{
    if constexpr (!MWaves<P, T>::ok) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_stage_mfma<P, T, MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)MWaves<P, T>::bytes);
}

template <int P, typename T>
int mfma_prepare()
{
    if constexpr (!MWaves<P, T>::ok) return 0;
    mfma_set_attr<P, T, kModeMult>();
    mfma_set_attr<P, T, kModeLserk>();
    mfma_set_attr<P, T, kModeRk1>();
    mfma_set_attr<P, T, kModeRk2>();
    mfma_set_attr<P, T, kModeRk3>();
    mfma_set_attr<P, T, kModeRk4>();
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dg_stage_mfma<P, T, kModeLserk>,
                                                     MWaves<P, T>::W * 64, MWaves<P, T>::bytes) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int P, typename T>
hipError_t mfma_launch(const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    constexpr int threads = MWaves<P, T>::W * 64;
    constexpr size_t bytes = MWaves<P, T>::bytes;
    if constexpr (!MWaves<P, T>::ok) return hipErrorInvalidConfiguration;
    switch (mode) {
        case kModeMult: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeMult>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeLserk: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeLserk>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk1: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk1>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk2: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk2>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk3: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk3>), dim3(grid), dim3(threads), bytes, s, a); break;
        case kModeRk4: hipLaunchKernelGGL((dg_stage_mfma<P, T, kModeRk4>), dim3(grid), dim3(threads), bytes, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_stage_mfma(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (order) {
        case 1: return mfma_launch<1, T>(a, mode, grid, s);
        case 2: return mfma_launch<2, T>(a, mode, grid, s);
        case 3: return mfma_launch<3, T>(a, mode, grid, s);
        case 4: return mfma_launch<4, T>(a, mode, grid, s);
        case 5: return mfma_launch<5, T>(a, mode, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
int mfma_blocks_per_cu(int order)
{
    switch (order) {
        case 1: return mfma_prepare<1, T>();
        case 2: return mfma_prepare<2, T>();
        case 3: return mfma_prepare<3, T>();
        case 4: return mfma_prepare<4, T>();
        case 5: return mfma_prepare<5, T>();
        default: return 1;
    }
}

template <typename T>
int mfma_waves_per_block(int order)
{
    switch (order) {
        case 1: return MWaves<1, T>::W;
        case 2: return MWaves<2, T>::W;
        case 3: return MWaves<3, T>::W;
        case 4: return MWaves<4, T>::W;
        case 5: return MWaves<5, T>::W;
        default: return 1;
    }
}

template hipError_t launch_stage_mfma<float>(int, const StageArgs<float>&, int, int, hipStream_t);
template hipError_t launch_stage_mfma<double>(int, const StageArgs<double>&, int, int, hipStream_t);
template int mfma_blocks_per_cu<float>(int);
template int mfma_blocks_per_cu<double>(int);
template int mfma_waves_per_block<float>(int);
template int mfma_waves_per_block<double>(int);

}  // namespace dgtd
