// Device body of the per-node (VALU) fused DGTD stage, shared by the
// single-mesh kernel (dg_kernels.hip) and the multi-instance batch kernel
// (dg_batch.hip).  See dg_kernels.hip for the algorithm.
#pragma once

#include <hip/hip_runtime.h>

#include "bc_coeffs.hpp"
#include "kernels.hpp"
#include "lds_copy.hpp"
#include "stability.hpp"

namespace dgtd {
namespace valu {

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

// Threads per block of the per-node kernels (single-mesh and batch).
constexpr int kThreads = 256;

// Loads the order-P reference matrices and tables into LDS (block-wide;
// the caller synchronises before use), several loads in flight per thread
// (up to 55 values per thread at order 5).
template <int P, typename T>
__device__ __forceinline__ void stage_load_ref(const T* __restrict__ mats, const int* __restrict__ tabs,
                                               unsigned char* smem_raw)
{
    using S = Shape<P>;
    using L = Smem<P, T>;
    constexpr int NFP = S::NFP, EB = S::EB;
    T* sD = reinterpret_cast<T*>(smem_raw);
    int* sFm = reinterpret_cast<int*>(sD + L::nReal) + EB * kConnStrideDev;
    lds_copy<8>(sD, mats, L::nD + L::nL, (int)threadIdx.x, kThreads);
    lds_copy<8>(sFm, tabs, 4 * NFP + 96 * NFP, (int)threadIdx.x, kThreads);
}

// One element block (EB elements from e0 = blk*EB) of one stage: gather,
// flux, curl, LIFT and the stage update.  Starts and ends with a block-wide
// barrier; the reference matrices must already be in LDS.
template <int P, typename T, int MODE>
__device__ __forceinline__ void stage_block(const StageArgs<T>& a, unsigned char* smem_raw, int blk)
{
    using S = Shape<P>;
    using L = Smem<P, T>;
    constexpr int NP = S::NP, NFP = S::NFP, NF = S::NF, EB = S::EB;
    T* sD = reinterpret_cast<T*>(smem_raw);  // [3][j][i]
    T* sL = sD + L::nD;                      // [m][i]
    T* sQ = sL + L::nL;                      // [c][el*NP + j]
    T* sF = sQ + L::nQ;                      // [c][el*NF + m]
    T* sG = sF + L::nF;                      // [el][kGeoStride]
    int* sC = reinterpret_cast<int*>(sG + L::nG);  // [el][kConnStride]
    int* sFm = sC + EB * kConnStrideDev;            // [4*NFP]
    int* sPt = sFm + 4 * NFP;                       // [96*NFP]

    const int tid = threadIdx.x;
    const int K = a.K;
    const size_t N = (size_t)NP * K;
    static_assert(EB * NP <= kThreads, "one volume node per thread");
    {
        const int e0 = blk * EB;
        const int ne = min(EB, K - e0);
        __syncthreads();  // previous iteration finished with sQ/sF/sG/sC (and tables loaded)
        {
            // every load of the block's fields and records in flight, then the
            // LDS writes (out-of-range threads load a clamped, valid address)
            constexpr int NG = (EB * kGeoStrideDev + kThreads - 1) / kThreads;
            constexpr int NC = (EB * kConnStrideDev + kThreads - 1) / kThreads;
            const size_t base = (size_t)e0 * NP;
            const int cnt = ne * NP, ng = ne * kGeoStrideDev, nc = ne * kConnStrideDev;
            const int iq = tid < cnt ? tid : cnt - 1;
            T q[6], gv[NG];
            int cv[NC];
#pragma unroll
            for (int c = 0; c < 6; ++c) q[c] = a.qin[c * N + base + iq];
#pragma unroll
            for (int k = 0; k < NG; ++k) {
                const int i = tid + k * kThreads;
                gv[k] = a.geo[(size_t)e0 * kGeoStrideDev + (i < ng ? i : ng - 1)];
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const int i = tid + k * kThreads;
                cv[k] = a.conn[(size_t)e0 * kConnStrideDev + (i < nc ? i : nc - 1)];
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

        // ---- surface: jumps, boundary conditions, flux, fscale/2 ----
        // A thread's face nodes (at most NS): first every exterior trace is
        // gathered, then the fluxes are formed.
        constexpr int NS = (EB * NF + kThreads - 1) / kThreads;
        const int nsf = ne * NF;
        T uM[NS][6], d[NS][6];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int s = tid + k * kThreads < nsf ? tid + k * kThreads : nsf - 1;
            const int el = s / NF;
            const int m = s - el * NF;
            const int f = m / NFP;
            const int b = m - f * NFP;
            const int nb = sC[el * kConnStrideDev + f];
            const int iM = sFm[f * NFP + b];
#pragma unroll
            for (int c = 0; c < 6; ++c) uM[k][c] = sQ[c * EB * NP + el * NP + iM];
            // a boundary face reads the element's own node (its value is
            // replaced below)
            const int code = sC[el * kConnStrideDev + 4 + f];
            const size_t gP = nb >= 0 ? (size_t)nb * NP + sPt[code * NFP + b] : (size_t)(e0 + el) * NP + iM;
#pragma unroll
            for (int c = 0; c < 6; ++c) d[k][c] = a.qin[c * N + gP];
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int s = tid + k * kThreads;
            if (s >= nsf) break;
            const int el = s / NF;
            const int m = s - el * NF;
            const int f = m / NFP;
            const T* g = sG + el * kGeoStrideDev;
            const int nb = sC[el * kConnStrideDev + f];
            T* dk = d[k];
            if (nb >= 0) {
#pragma unroll
                for (int c = 0; c < 6; ++c) dk[c] -= uM[k][c];
            } else {
                T ce, ch;
                bc_coeffs<T>(-nb - 1, a.alpha, ce, ch);
#pragma unroll
                for (int c = 0; c < 3; ++c) { dk[c] = ce * uM[k][c]; dk[3 + c] = ch * uM[k][3 + c]; }
            }
            const T nx = g[9 + 4 * f], ny = g[10 + 4 * f], nz = g[11 + 4 * f], fs = g[12 + 4 * f];
            const T n3[3] = {nx, ny, nz};
            const T ndH = nx * dk[3] + ny * dk[4] + nz * dk[5];
            const T ndE = nx * dk[0] + ny * dk[1] + nz * dk[2];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                const T fH = -n3[y] * dk[z] + n3[z] * dk[y] + a.alpha * (dk[3 + x] - ndH * n3[x]);
                const T fE = n3[y] * dk[3 + z] - n3[z] * dk[3 + y] + a.alpha * (dk[x] - ndE * n3[x]);
                sF[(3 + x) * EB * NF + el * NF + m] = fH * fs;
                sF[x * EB * NF + el * NF + m] = fE * fs;
            }
        }
        __syncthreads();

        // ---- volume: curl + LIFT, stage update ----
        bool bad = false;  // an unstable state entry stored by this thread (a.unstable set)
        if (tid < ne * NP) {
            const int t = tid;
            const int el = t / NP;
            const int i = t - el * NP;
            const size_t idx = (size_t)(e0 + el) * NP + i;
            // the stage's reads of the residual / base state, issued before
            // the products (all reads before the first store, see the MFMA variant)
            T rv[6], xv[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                if constexpr (MODE == kModeLserk || MODE == kModeRk2 || MODE == kModeRk3 || MODE == kModeRk4)
                    rv[c] = __builtin_nontemporal_load(a.res + c * N + idx);
                if constexpr (MODE == kModeRk2 || MODE == kModeRk3) xv[c] = a.x[c * N + idx];
            }
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
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const size_t o = c * N + idx;
                const T qv = sQ[c * EB * NP + el * NP + i];
                if constexpr (MODE == kModeMult) {
                    __builtin_nontemporal_store(rhs[c], a.out + o);
                } else if constexpr (MODE == kModeLserk) {
                    const T r = a.ca * rv[c] + a.dt * rhs[c];
                    const T qn = qv + a.cb * r;
                    __builtin_nontemporal_store(r, a.res + o);
                    __builtin_nontemporal_store(qn, a.out + o);
                    bad = bad || unstable_entry(qn);
                } else if constexpr (MODE == kModeRk1) {
                    __builtin_nontemporal_store(qv + a.ca * rhs[c], a.out + o);
                    __builtin_nontemporal_store(qv + a.cb * rhs[c], a.res + o);
                } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
                    __builtin_nontemporal_store(xv[c] + a.ca * rhs[c], a.out + o);
                    __builtin_nontemporal_store(rv[c] + a.cb * rhs[c], a.res + o);
                } else {  // kModeRk4
                    const T qn = rv[c] + a.ca * rhs[c];
                    __builtin_nontemporal_store(qn, a.out + o);
                    bad = bad || unstable_entry(qn);
                }
            }
        }
        if constexpr (MODE == kModeLserk || MODE == kModeRk4) report_unstable(a.unstable, bad);
    }
    __syncthreads();  // LDS tiles free for the next block
}

}  // namespace valu
}  // namespace dgtd
