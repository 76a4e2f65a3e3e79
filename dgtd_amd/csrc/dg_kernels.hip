// Fused DGTD Maxwell stage kernels for gfx950 (the "hesthaven" operator of
// OpenSEMBA/dgtd, HesthavenEvolution.cpp:450-542, with the RK stage update).
//
// One launch = one Runge-Kutta stage over all elements:
//   1. stage the block's fields (6 x Np per element), geometry and
//      connectivity records in LDS (coalesced: component-major state, one
//      contiguous chunk per component per element block);
//   2. per face node: gather the neighbour trace through (neighbour id,
//      permutation code), apply PEC/PMC/SMA, form the upwind/centred flux and
//      scale it by fscale/2 -- into LDS, never to HBM;
//   3. per volume node: reference derivatives [Dr;Ds;Dt] of the 6 fields and
//      LIFT * flux from LDS-resident reference matrices, metric combine into
//      the curl, 1/eps and 1/mu, then the stage update, writing each state
//      buffer once.
// Blocks are persistent (grid = resident capacity) and loop over element
// blocks, so the reference matrices are staged once per block.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace dgtd {

namespace {

template <int P>
struct Shape {
    static constexpr int NP = (P + 1) * (P + 2) * (P + 3) / 6;
    static constexpr int NFP = (P + 1) * (P + 2) / 2;
    static constexpr int NF = 4 * NFP;
    static constexpr int EB = (256 / NP) > 0 ? (256 / NP) : 1;  // elements per block
};

template <int P, typename T>
struct Smem {
    using S = Shape<P>;
    static constexpr int nD = 3 * S::NP * S::NP;
    static constexpr int nL = S::NF * S::NP;
    static constexpr int nQ = 6 * S::EB * S::NP;
    static constexpr int nF = 6 * S::EB * S::NF;
    static constexpr int nG = S::EB * kGeoStrideDev;
    static constexpr int nReal = nD + nL + nQ + nF + nG;
    static constexpr int nInt = S::EB * kConnStrideDev + 4 * S::NFP + 96 * S::NFP;
    static constexpr size_t bytes = sizeof(T) * (size_t)nReal + sizeof(int) * (size_t)nInt;
};

template <int P, typename T, int MODE>
__global__ __launch_bounds__(256) void dg_stage_kernel(StageArgs<T> a)
{
    using S = Shape<P>;
    using L = Smem<P, T>;
    constexpr int NP = S::NP, NFP = S::NFP, NF = S::NF, EB = S::EB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sD = reinterpret_cast<T*>(smem_raw);  // [3][j][i]
    T* sL = sD + L::nD;                      // [m][i]
    T* sQ = sL + L::nL;                      // [c][el*NP + j]
    T* sF = sQ + L::nQ;                      // [c][el*NF + m]
    T* sG = sF + L::nF;                      // [el][kGeoStride]
    int* sC = reinterpret_cast<int*>(sG + L::nG);  // [el][kConnStride]
    int* sFm = sC + EB * kConnStrideDev;            // [4*NFP]
    int* sPt = sFm + 4 * NFP;                       // [96*NFP]

    const int tid = threadIdx.x;
    for (int i = tid; i < L::nD + L::nL; i += blockDim.x) sD[i] = a.mats[i];
    for (int i = tid; i < 4 * NFP + 96 * NFP; i += blockDim.x) sFm[i] = a.tabs[i];

    const int K = a.K;
    const size_t N = (size_t)NP * K;

    for (int blk = blockIdx.x; blk < a.nblk; blk += gridDim.x) {
        const int e0 = blk * EB;
        const int ne = min(EB, K - e0);
        __syncthreads();  // previous iteration finished with sQ/sF/sG/sC (and tables loaded)
        {
            const size_t base = (size_t)e0 * NP;
            const int cnt = ne * NP;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                for (int i = tid; i < cnt; i += blockDim.x) sQ[c * EB * NP + i] = a.qin[c * N + base + i];
            for (int i = tid; i < ne * kGeoStrideDev; i += blockDim.x)
                sG[i] = a.geo[(size_t)e0 * kGeoStrideDev + i];
            for (int i = tid; i < ne * kConnStrideDev; i += blockDim.x)
                sC[i] = a.conn[(size_t)e0 * kConnStrideDev + i];
        }
        __syncthreads();

        // ---- surface: jumps, boundary conditions, flux, fscale/2 ----
        for (int s = tid; s < ne * NF; s += blockDim.x) {
            const int el = s / NF;
            const int m = s - el * NF;
            const int f = m / NFP;
            const int b = m - f * NFP;
            const T* g = sG + el * kGeoStrideDev;
            const int nb = sC[el * kConnStrideDev + f];
            const int iM = sFm[f * NFP + b];
            T uM[6], d[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) uM[c] = sQ[c * EB * NP + el * NP + iM];
            if (nb >= 0) {
                const int code = sC[el * kConnStrideDev + 4 + f];
                const size_t gP = (size_t)nb * NP + sPt[code * NFP + b];
#pragma unroll
                for (int c = 0; c < 6; ++c) d[c] = a.qin[c * N + gP] - uM[c];
            } else {
                const int bc = -nb - 1;
                // HesthavenEvolution.cpp:290-292: PEC (-2, 0), PMC (0, -2), SMA (-1/alpha, -1/alpha)
                T ce = T(0), ch = T(0);
                if (bc == 1) { ce = T(-2); ch = T(0); }
                else if (bc == 2) { ce = T(0); ch = T(-2); }
                else if (bc == 3) { ce = T(-1) / a.alpha; ch = ce; }
#pragma unroll
                for (int c = 0; c < 3; ++c) { d[c] = ce * uM[c]; d[3 + c] = ch * uM[3 + c]; }
            }
            const T nx = g[9 + 4 * f], ny = g[10 + 4 * f], nz = g[11 + 4 * f], fs = g[12 + 4 * f];
            const T n3[3] = {nx, ny, nz};
            const T ndH = nx * d[3] + ny * d[4] + nz * d[5];
            const T ndE = nx * d[0] + ny * d[1] + nz * d[2];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                const T fH = -n3[y] * d[z] + n3[z] * d[y] + a.alpha * (d[3 + x] - ndH * n3[x]);
                const T fE = n3[y] * d[3 + z] - n3[z] * d[3 + y] + a.alpha * (d[x] - ndE * n3[x]);
                sF[(3 + x) * EB * NF + el * NF + m] = fH * fs;
                sF[x * EB * NF + el * NF + m] = fE * fs;
            }
        }
        __syncthreads();

        // ---- volume: curl + LIFT, stage update ----
        for (int t = tid; t < ne * NP; t += blockDim.x) {
            const int el = t / NP;
            const int i = t - el * NP;
            T ar[6], as[6], at[6], lf[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) { ar[c] = T(0); as[c] = T(0); at[c] = T(0); lf[c] = T(0); }
#pragma unroll 5
            for (int j = 0; j < NP; ++j) {
                const T dr = sD[j * NP + i], ds = sD[NP * NP + j * NP + i], dt = sD[2 * NP * NP + j * NP + i];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const T u = sQ[c * EB * NP + el * NP + j];
                    ar[c] += dr * u;
                    as[c] += ds * u;
                    at[c] += dt * u;
                }
            }
#pragma unroll 4
            for (int m = 0; m < NF; ++m) {
                const T l = sL[m * NP + i];
#pragma unroll
                for (int c = 0; c < 6; ++c) lf[c] += l * sF[c * EB * NF + el * NF + m];
            }
            const T* g = sG + el * kGeoStrideDev;
            // D_d u = r_d ur + s_d us + t_d ut
            auto Dd = [&](int dd, int c) { return g[dd] * ar[c] + g[3 + dd] * as[c] + g[6 + dd] * at[c]; };
            T rhs[6];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                rhs[3 + x] = (-Dd(y, z) + Dd(z, y) + lf[3 + x]) * g[26];
                rhs[x] = (Dd(y, 3 + z) - Dd(z, 3 + y) + lf[x]) * g[25];
            }
            const size_t idx = (size_t)(e0 + el) * NP + i;
            // all reads before the first store (see the MFMA variant)
            T rv[6], xv[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                if constexpr (MODE == kModeLserk || MODE == kModeRk2 || MODE == kModeRk3 || MODE == kModeRk4)
                    rv[c] = a.res[c * N + idx];
                if constexpr (MODE == kModeRk2 || MODE == kModeRk3) xv[c] = a.x[c * N + idx];
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const size_t o = c * N + idx;
                const T qv = sQ[c * EB * NP + el * NP + i];
                if constexpr (MODE == kModeMult) {
                    a.out[o] = rhs[c];
                } else if constexpr (MODE == kModeLserk) {
                    const T r = a.ca * rv[c] + a.dt * rhs[c];
                    a.res[o] = r;
                    a.out[o] = qv + a.cb * r;
                } else if constexpr (MODE == kModeRk1) {
                    a.out[o] = qv + a.ca * rhs[c];
                    a.res[o] = qv + a.cb * rhs[c];
                } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
                    a.out[o] = xv[c] + a.ca * rhs[c];
                    a.res[o] = rv[c] + a.cb * rhs[c];
                } else {  // kModeRk4
                    a.out[o] = rv[c] + a.ca * rhs[c];
                }
            }
        }
    }
}

template <int P, typename T, int MODE>
hipError_t launch_one(const StageArgs<T>& a, int grid, hipStream_t s)
{
    constexpr size_t bytes = Smem<P, T>::bytes;
    auto kern = dg_stage_kernel<P, T, MODE>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), bytes, s, a);
    return hipGetLastError();
}

template <int P, typename T>
hipError_t launch_mode(const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (mode) {
        case kModeMult: return launch_one<P, T, kModeMult>(a, grid, s);
        case kModeLserk: return launch_one<P, T, kModeLserk>(a, grid, s);
        case kModeRk1: return launch_one<P, T, kModeRk1>(a, grid, s);
        case kModeRk2: return launch_one<P, T, kModeRk2>(a, grid, s);
        case kModeRk3: return launch_one<P, T, kModeRk3>(a, grid, s);
        case kModeRk4: return launch_one<P, T, kModeRk4>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <int P, typename T, int MODE>
void set_smem_attr()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_stage_kernel<P, T, MODE>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)Smem<P, T>::bytes);
}

// Raises the dynamic-LDS limit of every stage variant of this order (needed
// above 64 KiB) and returns resident blocks per CU.
template <int P, typename T>
int occupancy_one()
{
    int nb = 0;
    constexpr size_t bytes = Smem<P, T>::bytes;
    set_smem_attr<P, T, kModeMult>();
    set_smem_attr<P, T, kModeLserk>();
    set_smem_attr<P, T, kModeRk1>();
    set_smem_attr<P, T, kModeRk2>();
    set_smem_attr<P, T, kModeRk3>();
    set_smem_attr<P, T, kModeRk4>();
    auto kern = dg_stage_kernel<P, T, kModeLserk>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, bytes) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace

template <typename T>
hipError_t launch_stage(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (order) {
        case 1: return launch_mode<1, T>(a, mode, grid, s);
        case 2: return launch_mode<2, T>(a, mode, grid, s);
        case 3: return launch_mode<3, T>(a, mode, grid, s);
        case 4: return launch_mode<4, T>(a, mode, grid, s);
        case 5: return launch_mode<5, T>(a, mode, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
int stage_blocks_per_cu(int order)
{
    switch (order) {
        case 1: return occupancy_one<1, T>();
        case 2: return occupancy_one<2, T>();
        case 3: return occupancy_one<3, T>();
        case 4: return occupancy_one<4, T>();
        case 5: return occupancy_one<5, T>();
        default: return 1;
    }
}

int elements_per_block(int order)
{
    switch (order) {
        case 1: return Shape<1>::EB;
        case 2: return Shape<2>::EB;
        case 3: return Shape<3>::EB;
        case 4: return Shape<4>::EB;
        case 5: return Shape<5>::EB;
        default: return 1;
    }
}

template hipError_t launch_stage<double>(int, const StageArgs<double>&, int, int, hipStream_t);
template hipError_t launch_stage<float>(int, const StageArgs<float>&, int, int, hipStream_t);
template int stage_blocks_per_cu<double>(int);
template int stage_blocks_per_cu<float>(int);

}  // namespace dgtd
