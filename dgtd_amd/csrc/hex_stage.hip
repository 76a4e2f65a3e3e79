// Matrix-free stage kernel for hexahedra (parallelepipeds) under the Global
// operator's flux (GlobalEvolution::Mult, GlobalEvolution.cpp:628-1088): the
// same semi-discrete operator the assembled matrix of the "csr" variant
// holds, applied without the matrix.  On the Gauss-Lobatto tensor element a
// reference derivative touches the p + 1 nodes of one grid line, and the lift
// of a face node reaches only the normal line through it (weight
// M1^-1[i, b], b = 0 or p the face's index on that line), so per node the
// stage reads the element's own values once (LDS), one trace value per face
// and component, and no matrix: HBM traffic per DOF is the state vectors and
// the geometry record, where the matrix costs 18.5 entries per row at order 3.
//
// A block of 256 threads takes EB = 256 / Np elements (one at order 5):
//   1. the elements' six components into LDS (coalesced);
//   2. per face node: jump (neighbour's trace through the orientation table,
//      or the boundary / interior-condition weights of the own trace), flux
//      fluxE = a (dE - n (n . dE)) + n x dH, fluxH = a (dH - n (n . dH)) -
//      n x dE, scaled by fscale/2 and 1/eps or 1/mu, into LDS; a = 1 on SMA
//      faces whatever alpha, as the Global operator assembles them;
//   3. per node: metric-weighted tensor derivatives of the six components,
//      the curl, the six lifted face fluxes, -sigma/eps E, and the stage
//      update of its six DOFs.
#include <hip/hip_runtime.h>

#include "hex_dev.hpp"
#include "kernels.hpp"
#include "stability.hpp"

namespace dgtd {

namespace {

template <typename T, int P, int MODE>
__global__ __launch_bounds__(256) void hex_stage(StageArgs<T> a)
{
    constexpr int N1 = P + 1, NP = N1 * N1 * N1, NFP = N1 * N1;
    constexpr int EB = 256 / NP > 0 ? 256 / NP : 1;
    __shared__ T sq[EB * 6 * NP];
    __shared__ T sf[EB * 6 * NFP * 6];
    __shared__ T sD[N1 * N1];
    __shared__ T sMi[N1 * 2];
    const int t = threadIdx.x;
    const int64_t K = a.nrows / (6 * (int64_t)NP);
    const int64_t N = K * NP;
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int ne = (int)(K - e0 < EB ? K - e0 : EB);
    // reference tables: D1 row-major, then M1^-1 columns 0 and p
    if (t < N1 * N1) sD[t] = a.mats[t];
    if (t < N1 * 2) sMi[t] = a.mats[N1 * N1 + t];
    for (int c = 0; c < 6; ++c)
        for (int x = t; x < ne * NP; x += 256) sq[(x / NP) * 6 * NP + c * NP + x % NP] = a.qin[c * N + e0 * NP + x];
    __syncthreads();
    const int* fmask = a.tabs;                  // [6][NFP]
    const int* perm = a.tabs + 6 * NFP;         // [48][NFP]
    const int* fq = a.tabs + 54 * NFP;          // [6][NP]: face node on the node's normal line
    for (int x = t; x < ne * 6 * NFP; x += 256) {
        const int el = x / (6 * NFP), f = (x / NFP) % 6, b = x % NFP;
        const int64_t e = e0 + el;
        const T* g = a.geo + (int64_t)kHexGeoDev * e;
        const int nb = a.conn[kHexConnDev * e + f];
        const int m = fmask[f * NFP + b];
        const T* own = sq + el * 6 * NP;
        T d[6];
        T af = a.alpha;
        if (nb >= 0) {
            const int pn = perm[a.conn[kHexConnDev * e + 6 + f] * NFP + b];
#pragma unroll
            for (int c = 0; c < 6; ++c) d[c] = a.qin[c * N + (int64_t)nb * NP + pn] - own[c * NP + m];
        } else {
            const int bc = -nb - 1;
            if (bc == 3 || bc == 6) af = T(1);
            const T sma = af != T(0) ? T(-1) / af : T(0);
            const T ce = bc == 1 ? T(-2) : bc == 3 ? sma : bc == 4 ? T(-1) : bc == 6 ? T(-0.5) : T(0);
            const T ch = bc == 2 ? T(-2) : bc == 3 ? sma : bc == 5 ? T(-1) : bc == 6 ? T(-0.5) : T(0);
#pragma unroll
            for (int c = 0; c < 6; ++c) d[c] = (c < 3 ? ce : ch) * own[c * NP + m];
        }
        const T nx = g[9 + 4 * f], ny = g[10 + 4 * f], nz = g[11 + 4 * f], fs = g[12 + 4 * f];
        const T ndE = nx * d[0] + ny * d[1] + nz * d[2], ndH = nx * d[3] + ny * d[4] + nz * d[5];
        const T se = fs * g[33], sh = fs * g[34];
        T* out = sf + ((el * 6 + f) * NFP + b) * 6;
        out[0] = se * (af * (d[0] - nx * ndE) + (ny * d[5] - nz * d[4]));
        out[1] = se * (af * (d[1] - ny * ndE) + (nz * d[3] - nx * d[5]));
        out[2] = se * (af * (d[2] - nz * ndE) + (nx * d[4] - ny * d[3]));
        out[3] = sh * (af * (d[3] - nx * ndH) - (ny * d[2] - nz * d[1]));
        out[4] = sh * (af * (d[4] - ny * ndH) - (nz * d[0] - nx * d[2]));
        out[5] = sh * (af * (d[5] - nz * ndH) - (nx * d[1] - ny * d[0]));
    }
    __syncthreads();
    if (t >= ne * NP) return;
    const int el = t / NP, i = t % NP;
    const int ix = i % N1, iy = (i / N1) % N1, iz = i / NFP;
    const int64_t e = e0 + el;
    const T* g = a.geo + (int64_t)kHexGeoDev * e;
    const T* own = sq + el * 6 * NP;
    // physical derivatives of the six components: D_d u = sum_k G[k][d] d_k u
    T du[6][3];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const T* u = own + c * NP;
        T ur = T(0), us = T(0), ut = T(0);
#pragma unroll
        for (int m = 0; m < N1; ++m) {
            ur += sD[ix * N1 + m] * u[m + N1 * iy + NFP * iz];
            us += sD[iy * N1 + m] * u[ix + N1 * m + NFP * iz];
            ut += sD[iz * N1 + m] * u[ix + N1 * iy + NFP * m];
        }
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) du[c][dd] = g[dd] * ur + g[3 + dd] * us + g[6 + dd] * ut;
    }
    const T ieps = g[33], imu = g[34], sig = g[35];
    T rhs[6];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const int y = (x + 1) % 3, z = (x + 2) % 3;
        rhs[x] = ieps * (du[3 + z][y] - du[3 + y][z]) - sig * own[x * NP + i];
        rhs[3 + x] = imu * (du[y][z] - du[z][y]);
    }
    // lifted face fluxes: face f's node on this node's normal line
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int li = f == 0 || f == 5 ? iz : f == 1 || f == 3 ? iy : ix;
        const T w = sMi[li * 2 + (f == 2 || f == 3 || f == 5 ? 1 : 0)];
        const T* fl = sf + ((el * 6 + f) * NFP + fq[f * NP + i]) * 6;
#pragma unroll
        for (int c = 0; c < 6; ++c) rhs[c] += w * fl[c];
    }
    bool bad = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const int64_t r = c * N + e * NP + i;
        const T k = rhs[c];
        if constexpr (MODE == kModeMult) {
            a.out[r] = k;
        } else if constexpr (MODE == kModeLserk) {
            const T rr = a.ca * a.res[r] + a.dt * k;
            a.res[r] = rr;
            const T qn = a.qin[r] + a.cb * rr;
            a.out[r] = qn;
            bad = bad || unstable_entry(qn);
        } else if constexpr (MODE == kModeRk1) {
            const T q = a.qin[r];
            a.out[r] = q + a.ca * k;
            a.res[r] = q + a.cb * k;
        } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
            const T xv = a.x[r], rv = a.res[r];
            a.out[r] = xv + a.ca * k;
            a.res[r] = rv + a.cb * k;
        } else {
            const T qn = a.res[r] + a.ca * k;
            a.out[r] = qn;
            bad = bad || unstable_entry(qn);
        }
    }
    if constexpr (MODE == kModeLserk || MODE == kModeRk4) {
        if (a.unstable != nullptr && bad) atomicOr(a.unstable, 1);
    }
}

template <typename T, int P>
hipError_t hex_launch(const StageArgs<T>& a, int mode, hipStream_t s)
{
    constexpr int N1 = P + 1, NP = N1 * N1 * N1;
    constexpr int EB = 256 / NP > 0 ? 256 / NP : 1;
    const int64_t K = a.nrows / (6 * (int64_t)NP);
    const int64_t grid = (K + EB - 1) / EB;
    switch (mode) {
        case kModeMult: hipLaunchKernelGGL((hex_stage<T, P, kModeMult>), dim3((unsigned)grid), dim3(256), 0, s, a); break;
        case kModeLserk: hipLaunchKernelGGL((hex_stage<T, P, kModeLserk>), dim3((unsigned)grid), dim3(256), 0, s, a); break;
        case kModeRk1: hipLaunchKernelGGL((hex_stage<T, P, kModeRk1>), dim3((unsigned)grid), dim3(256), 0, s, a); break;
        case kModeRk2: hipLaunchKernelGGL((hex_stage<T, P, kModeRk2>), dim3((unsigned)grid), dim3(256), 0, s, a); break;
        case kModeRk3: hipLaunchKernelGGL((hex_stage<T, P, kModeRk3>), dim3((unsigned)grid), dim3(256), 0, s, a); break;
        case kModeRk4: hipLaunchKernelGGL((hex_stage<T, P, kModeRk4>), dim3((unsigned)grid), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_hex_stage(int order, const StageArgs<T>& a, int mode, hipStream_t s)
{
    if (a.nrows <= 0) return hipErrorInvalidValue;
    switch (order) {
        case 1: return hex_launch<T, 1>(a, mode, s);
        case 2: return hex_launch<T, 2>(a, mode, s);
        case 3: return hex_launch<T, 3>(a, mode, s);
        case 4: return hex_launch<T, 4>(a, mode, s);
        case 5: return hex_launch<T, 5>(a, mode, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_hex_stage<float>(int, const StageArgs<float>&, int, hipStream_t);
template hipError_t launch_hex_stage<double>(int, const StageArgs<double>&, int, hipStream_t);

}  // namespace dgtd
