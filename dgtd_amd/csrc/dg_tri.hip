// Fused DGTD stage kernel on triangles (the 2D hesthaven operator,
// HesthavenEvolution.cpp:450-542 on a 2D mesh: d/dz = 0, n_z = 0, edges as
// faces).  Same per-node structure as the tetrahedral VALU kernel: a block
// stages the fields of EB elements in LDS, one thread per edge node forms the
// jump (neighbour gather or boundary condition) and the flux into LDS, one
// thread per volume node applies [Dr; Ds] and LIFT from LDS-resident matrices,
// combines the curl with the metric, 1/eps and 1/mu, and applies the stage
// update.  Blocks are persistent over element blocks.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bc_coeffs.hpp"
#include "kernels.hpp"
#include "lds_copy.hpp"
#include "stability.hpp"

namespace dgtd {

namespace {

constexpr int kTriGeo = 15;   // tri.hpp kTriGeoStride
constexpr int kTriConn = 6;   // tri.hpp kTriConnStride

template <int P>
struct TShape {
    static constexpr int NP = (P + 1) * (P + 2) / 2;
    static constexpr int NFP = P + 1;
    static constexpr int NF = 3 * NFP;
    static constexpr int EB = 256 / NP;  // elements per block
};

template <int P, typename T>
struct TSmem {
    using S = TShape<P>;
    static constexpr int nD = 2 * S::NP * S::NP;
    static constexpr int nL = S::NF * S::NP;
    static constexpr int nQ = 6 * S::EB * S::NP;
    static constexpr int nF = 6 * S::EB * S::NF;
    static constexpr int nG = S::EB * kTriGeo;
    static constexpr int nReal = nD + nL + nQ + nF + nG;
    static constexpr int nInt = S::EB * kTriConn + 3 * S::NFP + 18 * S::NFP;
    static constexpr size_t bytes = sizeof(T) * (size_t)nReal + sizeof(int) * (size_t)nInt;
};

template <int P, typename T, int MODE>
__global__ __launch_bounds__(256) void dg_tri_stage(StageArgs<T> a)
{
    using S = TShape<P>;
    using L = TSmem<P, T>;
    constexpr int NP = S::NP, NFP = S::NFP, NF = S::NF, EB = S::EB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sD = reinterpret_cast<T*>(smem_raw);        // [2][j][i]
    T* sL = sD + L::nD;                            // [m][i]
    T* sQ = sL + L::nL;                            // [c][el*NP + j]
    T* sF = sQ + L::nQ;                            // [c][el*NF + m]
    T* sG = sF + L::nF;                            // [el][kTriGeo]
    int* sC = reinterpret_cast<int*>(sG + L::nG);  // [el][kTriConn]
    int* sFm = sC + EB * kTriConn;                 // [3*NFP]
    int* sPt = sFm + 3 * NFP;                      // [18*NFP]
    constexpr int kThreads = 256;  // the launch's block size
    static_assert(EB * NP <= kThreads, "one volume node per thread");
    const int tid = threadIdx.x;
    lds_copy<8>(sD, a.mats, L::nD + L::nL, tid, kThreads);
    lds_copy<8>(sFm, a.tabs, 21 * NFP, tid, kThreads);
    const int K = a.K;
    const size_t N = (size_t)NP * K;

    for (int blk = blockIdx.x; blk < a.nblk; blk += gridDim.x) {
        const int e0 = blk * EB;
        const int ne = min(EB, K - e0);
        __syncthreads();  // tables loaded; previous block done with the tiles
        {
            // every load of the block's fields and records in flight, then the
            // LDS writes (out-of-range threads load a clamped, valid address)
            constexpr int NG = (EB * kTriGeo + kThreads - 1) / kThreads;
            constexpr int NC = (EB * kTriConn + kThreads - 1) / kThreads;
            const size_t base = (size_t)e0 * NP;
            const int cnt = ne * NP, ng = ne * kTriGeo, nc = ne * kTriConn;
            const int iq = tid < cnt ? tid : cnt - 1;
            T q[6], gv[NG];
            int cv[NC];
#pragma unroll
            for (int c = 0; c < 6; ++c) q[c] = a.qin[c * N + base + iq];
#pragma unroll
            for (int k = 0; k < NG; ++k) {
                const int i = tid + k * kThreads;
                gv[k] = a.geo[(size_t)e0 * kTriGeo + (i < ng ? i : ng - 1)];
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const int i = tid + k * kThreads;
                cv[k] = a.conn[(size_t)e0 * kTriConn + (i < nc ? i : nc - 1)];
            }
            if (tid < cnt) {
#pragma unroll
                for (int c = 0; c < 6; ++c) sQ[c * EB * NP + tid] = q[c];
            }
#pragma unroll
            for (int k = 0; k < NG; ++k)
                if (tid + k * kThreads < ng) sG[tid + k * kThreads] = gv[k];
#pragma unroll
            for (int k = 0; k < NC; ++k)
                if (tid + k * kThreads < nc) sC[tid + k * kThreads] = cv[k];
        }
        __syncthreads();

        // ---- edges: jumps (exterior - interior), boundary conditions, flux ----
        // A thread's edge nodes (at most NS): every exterior trace gathered
        // first, then the fluxes formed.
        constexpr int NS = (EB * NF + kThreads - 1) / kThreads;
        const int nsf = ne * NF;
        T uMs[NS][6], ds[NS][6];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int s = tid + k * kThreads < nsf ? tid + k * kThreads : nsf - 1;
            const int el = s / NF;
            const int m = s - el * NF;
            const int f = m / NFP;
            const int b = m - f * NFP;
            const int nb = sC[el * kTriConn + f];
            const int iM = sFm[f * NFP + b];
#pragma unroll
            for (int c = 0; c < 6; ++c) uMs[k][c] = sQ[c * EB * NP + el * NP + iM];
            // a boundary edge reads the element's own node (replaced below)
            const int code = sC[el * kTriConn + 3 + f];
            const size_t gP = nb >= 0 ? (size_t)nb * NP + sPt[code * NFP + b] : (size_t)(e0 + el) * NP + iM;
#pragma unroll
            for (int c = 0; c < 6; ++c) ds[k][c] = a.qin[c * N + gP];
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int s = tid + k * kThreads;
            if (s >= nsf) break;
            const int el = s / NF;
            const int m = s - el * NF;
            const int f = m / NFP;
            const T* g = sG + el * kTriGeo;
            const int nb = sC[el * kTriConn + f];
            const T* uM = uMs[k];
            T* d = ds[k];
            if (nb >= 0) {
#pragma unroll
                for (int c = 0; c < 6; ++c) d[c] -= uM[c];
            } else {
                T ce, ch;
                bc_coeffs<T>(-nb - 1, a.alpha, ce, ch);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    d[c] = ce * uM[c];
                    d[3 + c] = ch * uM[3 + c];
                }
            }
            const T nx = g[4 + 3 * f], ny = g[5 + 3 * f], fs = g[6 + 3 * f];
            const T ndE = nx * d[0] + ny * d[1];
            const T ndH = nx * d[3] + ny * d[4];
            // n = (nx, ny, 0):  fH_x = -n_y dE_z + n_z dE_y + a (dH_x - ndH n_x), ...
            const T fHx = -ny * d[2] + a.alpha * (d[3] - ndH * nx);
            const T fHy = nx * d[2] + a.alpha * (d[4] - ndH * ny);
            const T fHz = -nx * d[1] + ny * d[0] + a.alpha * d[5];
            const T fEx = ny * d[5] + a.alpha * (d[0] - ndE * nx);
            const T fEy = -nx * d[5] + a.alpha * (d[1] - ndE * ny);
            const T fEz = nx * d[4] - ny * d[3] + a.alpha * d[2];
            T* o = sF + el * NF + m;
            o[0 * EB * NF] = fEx * fs;
            o[1 * EB * NF] = fEy * fs;
            o[2 * EB * NF] = fEz * fs;
            o[3 * EB * NF] = fHx * fs;
            o[4 * EB * NF] = fHy * fs;
            o[5 * EB * NF] = fHz * fs;
        }
        __syncthreads();

        // ---- volume: curl with d/dz = 0, LIFT, stage update ----
        bool bad = false;  // an unstable state entry stored by this thread (a.unstable set)
        if (tid < ne * NP) {
            const int t = tid;
            const int el = t / NP;
            const int i = t - el * NP;
            const size_t idx = (size_t)(e0 + el) * NP + i;
            // the stage's reads of the residual / base state, before the products
            T rv[6], xv[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                if constexpr (MODE == kModeLserk || MODE == kModeRk2 || MODE == kModeRk3 || MODE == kModeRk4)
                    rv[c] = a.res[c * N + idx];
                if constexpr (MODE == kModeRk2 || MODE == kModeRk3) xv[c] = a.x[c * N + idx];
            }
            T ar[6], as[6], lf[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) { ar[c] = T(0); as[c] = T(0); lf[c] = T(0); }
            for (int j = 0; j < NP; ++j) {
                const T dr = sD[j * NP + i], ds = sD[NP * NP + j * NP + i];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const T u = sQ[c * EB * NP + el * NP + j];
                    ar[c] += dr * u;
                    as[c] += ds * u;
                }
            }
            for (int m = 0; m < NF; ++m) {
                const T l = sL[m * NP + i];
#pragma unroll
                for (int c = 0; c < 6; ++c) lf[c] += l * sF[c * EB * NF + el * NF + m];
            }
            const T* g = sG + el * kTriGeo;
            auto Dx = [&](int c) { return g[0] * ar[c] + g[2] * as[c]; };
            auto Dy = [&](int c) { return g[1] * ar[c] + g[3] * as[c]; };
            // rhsE = curl H + LIFT fE (scaled 1/eps), rhsH = -curl E + LIFT fH (1/mu)
            T rhs[6];
            rhs[0] = (Dy(5) + lf[0]) * g[13];          // E_x:  dHz/dy
            rhs[1] = (-Dx(5) + lf[1]) * g[13];         // E_y: -dHz/dx
            rhs[2] = (Dx(4) - Dy(3) + lf[2]) * g[13];  // E_z:  dHy/dx - dHx/dy
            rhs[3] = (-Dy(2) + lf[3]) * g[14];         // H_x: -dEz/dy
            rhs[4] = (Dx(2) + lf[4]) * g[14];          // H_y:  dEz/dx
            rhs[5] = (-Dx(1) + Dy(0) + lf[5]) * g[14]; // H_z: -(dEy/dx - dEx/dy)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const size_t o = c * N + idx;
                const T qv = sQ[c * EB * NP + el * NP + i];
                if constexpr (MODE == kModeMult) {
                    a.out[o] = rhs[c];
                } else if constexpr (MODE == kModeLserk) {
                    const T r = a.ca * rv[c] + a.dt * rhs[c];
                    const T qn = qv + a.cb * r;
                    a.res[o] = r;
                    a.out[o] = qn;
                    bad = bad || unstable_entry(qn);
                } else if constexpr (MODE == kModeRk1) {
                    a.out[o] = qv + a.ca * rhs[c];
                    a.res[o] = qv + a.cb * rhs[c];
                } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
                    a.out[o] = xv[c] + a.ca * rhs[c];
                    a.res[o] = rv[c] + a.cb * rhs[c];
                } else {
                    const T qn = rv[c] + a.ca * rhs[c];
                    a.out[o] = qn;
                    bad = bad || unstable_entry(qn);
                }
            }
        }
        if constexpr (MODE == kModeLserk || MODE == kModeRk4) report_unstable(a.unstable, bad);
    }
}

template <int P, typename T, int MODE>
hipError_t tri_launch_one(const StageArgs<T>& a, int grid, hipStream_t s)
{
    constexpr size_t bytes = TSmem<P, T>::bytes;
    auto kern = dg_tri_stage<P, T, MODE>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), bytes, s, a);
    return hipGetLastError();
}

template <int P, typename T, int MODE>
void tri_set_attr()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_tri_stage<P, T, MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)TSmem<P, T>::bytes);
}

template <int P, typename T>
void tri_prepare_order()
{
    tri_set_attr<P, T, kModeMult>();
    tri_set_attr<P, T, kModeLserk>();
    tri_set_attr<P, T, kModeRk1>();
    tri_set_attr<P, T, kModeRk2>();
    tri_set_attr<P, T, kModeRk3>();
    tri_set_attr<P, T, kModeRk4>();
}

template <int P, typename T>
hipError_t tri_launch_mode(const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (mode) {
        case kModeMult: return tri_launch_one<P, T, kModeMult>(a, grid, s);
        case kModeLserk: return tri_launch_one<P, T, kModeLserk>(a, grid, s);
        case kModeRk1: return tri_launch_one<P, T, kModeRk1>(a, grid, s);
        case kModeRk2: return tri_launch_one<P, T, kModeRk2>(a, grid, s);
        case kModeRk3: return tri_launch_one<P, T, kModeRk3>(a, grid, s);
        case kModeRk4: return tri_launch_one<P, T, kModeRk4>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename T>
hipError_t launch_tri_stage(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (order) {
        case 1: return tri_launch_mode<1, T>(a, mode, grid, s);
        case 2: return tri_launch_mode<2, T>(a, mode, grid, s);
        case 3: return tri_launch_mode<3, T>(a, mode, grid, s);
        case 4: return tri_launch_mode<4, T>(a, mode, grid, s);
        case 5: return tri_launch_mode<5, T>(a, mode, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
void tri_prepare(int order)
{
    switch (order) {
        case 1: tri_prepare_order<1, T>(); break;
        case 2: tri_prepare_order<2, T>(); break;
        case 3: tri_prepare_order<3, T>(); break;
        case 4: tri_prepare_order<4, T>(); break;
        case 5: tri_prepare_order<5, T>(); break;
        default: break;
    }
}

int tri_elements_per_block(int order)
{
    switch (order) {
        case 1: return TShape<1>::EB;
        case 2: return TShape<2>::EB;
        case 3: return TShape<3>::EB;
        case 4: return TShape<4>::EB;
        case 5: return TShape<5>::EB;
        default: return 1;
    }
}

size_t tri_smem_bytes(int order, int precision)
{
    auto pick = [&](auto d) -> size_t {
        using T = decltype(d);
        switch (order) {
            case 1: return TSmem<1, T>::bytes;
            case 2: return TSmem<2, T>::bytes;
            case 3: return TSmem<3, T>::bytes;
            case 4: return TSmem<4, T>::bytes;
            case 5: return TSmem<5, T>::bytes;
            default: return 0;
        }
    };
    return precision == 64 ? pick(double{}) : pick(float{});
}

template void tri_prepare<double>(int);
template void tri_prepare<float>(int);
template hipError_t launch_tri_stage<double>(int, const StageArgs<double>&, int, int, hipStream_t);
template hipError_t launch_tri_stage<float>(int, const StageArgs<float>&, int, int, hipStream_t);

}  // namespace dgtd
