// FP64 reductions over a state vector: DOF 2-norm (Fields::getNorml2,
// src/evolution/Fields.h:21, used by the stability check of Solver::run,
// Solver.cpp:500-516), electromagnetic energy and a non-finite count.
// Fixed-order two-pass tree reduction, no atomics: bitwise reproducible.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"

namespace dgtd {

namespace {

constexpr int kThreads = 256;

// Scratch of the allocation-free norms entry point (dgtd_op_norms_noalloc):
// the per-block partials of norms_partial, at most kNormsMaxParts blocks x 3
// doubles (24 KiB on every GPU the library runs on, held until the process
// ends).  In host code `scratch` names a host-side stand-in, never the device
// array: its device address is resolved per device by norms_scratch.  Every
// call on one GPU shares that GPU's copy, so calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[3 * kNormsMaxParts];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

__device__ void block_reduce3(double v[3], double* sh)
{
    const int t = threadIdx.x;
    for (int k = 0; k < 3; ++k) sh[k * kThreads + t] = v[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 3; ++k) sh[k * kThreads + t] += sh[k * kThreads + t + w];
        __syncthreads();
    }
    for (int k = 0; k < 3; ++k) v[k] = sh[k * kThreads];
}

// One block per tile of TE elements at a time: the tile's six components are
// read once, coalesced, into LDS (2-norm and non-finite count on the way),
// then each thread forms u^T M u for one (element, component) pair.
template <typename T>
__global__ __launch_bounds__(kThreads) void norms_partial(const T* __restrict__ q, int K, int Np, int TE,
                                                          const double* __restrict__ M,
                                                          const double* __restrict__ jac,
                                                          const double* __restrict__ eps,
                                                          const double* __restrict__ mu,
                                                          double* __restrict__ partials)
{
    extern __shared__ double sh[];  // Np*Np mass, 6*TE*Np tile, 3*kThreads scratch
    double* sM = sh;
    double* sU = sh + Np * Np;
    double* sR = sU + 6 * TE * Np;
    for (int i = threadIdx.x; i < Np * Np; i += kThreads) sM[i] = M[i];
    const size_t N = (size_t)Np * K;
    double acc[3] = {0.0, 0.0, 0.0};
    const int ntiles = (K + TE - 1) / TE;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int e0 = tile * TE;
        const int ne = min(TE, K - e0);
        const int cnt = ne * Np;
        __syncthreads();  // previous tile's reads done (and sM loaded)
        for (int c = 0; c < 6; ++c)
            for (int t = threadIdx.x; t < cnt; t += kThreads) {
                const double v = (double)q[c * N + (size_t)e0 * Np + t];
                acc[0] += v * v;
                if (!isfinite(v)) acc[2] += 1.0;
                sU[c * TE * Np + t] = v;
            }
        __syncthreads();
        for (int pr = threadIdx.x; pr < 6 * ne; pr += kThreads) {
            const int c = pr / ne, el = pr - c * ne;
            const double* u = sU + c * TE * Np + el * Np;
            double s = 0.0;
            for (int i = 0; i < Np; ++i) {
                double r = 0.0;
                for (int j = 0; j < Np; ++j) r += sM[i * Np + j] * u[j];
                s += u[i] * r;
            }
            const int e = e0 + el;
            acc[1] += 0.5 * jac[e] * (c < 3 ? eps[e] : mu[e]) * s;
        }
    }
    block_reduce3(acc, sR);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partials[3 * blockIdx.x + k] = acc[k];
}

__global__ __launch_bounds__(kThreads) void norms_final(const double* __restrict__ partials, int n,
                                                        double* __restrict__ out)
{
    __shared__ double sR[3 * kThreads];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < n; b += kThreads)
        for (int k = 0; k < 3; ++k) acc[k] += partials[3 * b + k];
    block_reduce3(acc, sR);
    if (threadIdx.x == 0) {
        out[0] = sqrt(acc[0]);
        out[1] = acc[1];
        out[2] = acc[2];
    }
}

}  // namespace

template <typename T>
hipError_t launch_norms(const T* q, int K, int Np, const double* Mref, const double* jac,
                        const double* eps, const double* mu, double* partials, int nparts,
                        double* out3, hipStream_t s)
{
    const int TE = norms_tile(Np);
    const size_t sh = sizeof(double) * ((size_t)Np * Np + 6 * (size_t)TE * Np + 3 * kThreads);
    hipLaunchKernelGGL(norms_partial<T>, dim3(nparts), dim3(kThreads), sh, s, q, K, Np, TE, Mref, jac, eps,
                       mu, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(norms_final, dim3(1), dim3(kThreads), 0, s, partials, nparts, out3);
    return hipGetLastError();
}

hipError_t norms_scratch(int device, double** ptr)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // resolved with the device current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *ptr = g_scratch[device];
    return hipSuccess;
}

template hipError_t launch_norms<double>(const double*, int, int, const double*, const double*,
                                         const double*, const double*, double*, int, double*,
                                         hipStream_t);
template hipError_t launch_norms<float>(const float*, int, int, const double*, const double*,
                                        const double*, const double*, double*, int, double*,
                                        hipStream_t);

// Point probes: one wave per (probe, component) forms
// sum_n w[probe][n] * q[c*N + elem[probe]*Np + n] in FP64; probes outside
// the mesh (elem < 0) give NaN.
template <typename T>
__global__ __launch_bounds__(64) void probe_sample(const T* __restrict__ q, size_t N, int Np,
                                                   const int* __restrict__ elem,
                                                   const double* __restrict__ w, double* __restrict__ out)
{
    const int p = blockIdx.x / 6, c = blockIdx.x % 6;
    const int e = elem[p];
    double s = 0.0;
    if (e >= 0)
        for (int n = threadIdx.x; n < Np; n += 64) s += w[(size_t)p * Np + n] * (double)q[c * N + (size_t)e * Np + n];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x == 0) out[blockIdx.x] = e >= 0 ? s : __builtin_nan("");
}

template <typename T>
hipError_t launch_probe_sample(const T* q, size_t N, int Np, int nprobes, const int* elem, const double* w,
                               double* out, hipStream_t s)
{
    if (nprobes <= 0) return hipSuccess;
    hipLaunchKernelGGL(probe_sample<T>, dim3(6 * nprobes), dim3(64), 0, s, q, N, Np, elem, w, out);
    return hipGetLastError();
}

template hipError_t launch_probe_sample<double>(const double*, size_t, int, int, const int*, const double*,
                                                double*, hipStream_t);
template hipError_t launch_probe_sample<float>(const float*, size_t, int, int, const int*, const double*,
                                               double*, hipStream_t);

}  // namespace dgtd
