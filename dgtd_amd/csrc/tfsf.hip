// Total-field / scattered-field plane-wave injection
// (HesthavenEvolution::evaluateTFSF, HesthavenEvolution.cpp:97-124).
//
// On a TF/SF face the jump gains -E_inc, -H_inc on the scattered-field side
// and +E_inc, +H_inc on the total-field side, the incident field of the plane
// waves evaluated at the face nodes at the stage time.  The flux is linear in
// the jump, so the injection is a state-independent right-hand side
//     S_e = LIFT (fscale/2 flux(+-inc)) / (eps or mu)
// on the elements that own TF/SF faces; this kernel adds c_out * S to the
// stage output and c_res * S to the residual after the fused stage kernel,
// with the coefficients the stage applies to the right-hand side.  One
// 64-thread block per such element, FP64 arithmetic.  On triangles (dim 2)
// the faces are the three edges and n_z = 0.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dgtd {

namespace {

constexpr int kTfsfMaxFaceNodes = 4 * 21;  // order 5

template <typename T>
__global__ __launch_bounds__(64) void tfsf_apply(TfsfArgs<T> a)
{
    __shared__ double flux[6][kTfsfMaxFaceNodes];
    const int k = blockIdx.x;
    const int e = a.elem[k];
    const bool tri = a.dim == 2;
    const int Np = a.Np, Nfp = a.Nfp, Nf = (tri ? 3 : 4) * Nfp;
    const double* g = a.geo + (size_t)k * (tri ? 15 : 27);
    for (int m = threadIdx.x; m < Nf; m += 64) {
        const int f = m / Nfp;
        const int sg = a.sign[4 * k + f];
        double dE[3] = {0.0, 0.0, 0.0}, dH[3] = {0.0, 0.0, 0.0};
        if (sg != 0) {
            const double* x = a.xyz + ((size_t)k * Nf + m) * 3;
            for (int w = 0; w < a.nwaves; ++w) {
                const TfsfWave& pw = a.waves[w];
                const double u = pw.dir[0] * x[0] + pw.dir[1] * x[1] + pw.dir[2] * x[2] - a.t;
                const double d = u - pw.mean;
                double mag = exp(-d * d / (2.0 * pw.spread * pw.spread));
                if (pw.kind == 1) mag *= cos(2.0 * M_PI * pw.freq * d);
                for (int c = 0; c < 3; ++c) {
                    dE[c] += sg * mag * pw.pol[c];
                    dH[c] += sg * mag * pw.kxp[c];
                }
            }
        }
        // records: tetrahedron [9+4f..12+4f] = nx ny nz fscale/2; triangle
        // [4+3f..6+3f] = nx ny fscale/2
        const double n[3] = {tri ? g[4 + 3 * f] : g[9 + 4 * f], tri ? g[5 + 3 * f] : g[10 + 4 * f],
                             tri ? 0.0 : g[11 + 4 * f]};
        const double fs = tri ? g[6 + 3 * f] : g[12 + 4 * f];
        const double ndE = n[0] * dE[0] + n[1] * dE[1] + n[2] * dE[2];
        const double ndH = n[0] * dH[0] + n[1] * dH[1] + n[2] * dH[2];
        for (int x = 0; x < 3; ++x) {
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            flux[x][m] = fs * (n[y] * dH[z] - n[z] * dH[y] + a.alpha * (dE[x] - ndE * n[x]));
            flux[3 + x][m] = fs * (-n[y] * dE[z] + n[z] * dE[y] + a.alpha * (dH[x] - ndH * n[x]));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Np; i += 64) {
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int m = 0; m < Nf; ++m) s += a.lift[(size_t)i * Nf + m] * flux[c][m];
            s *= c < 3 ? g[tri ? 13 : 25] : g[tri ? 14 : 26];
            const size_t idx = c * a.N + (size_t)e * Np + i;
            a.out[idx] = (T)((double)a.out[idx] + a.c_out * s);
            if (a.res) a.res[idx] = (T)((double)a.res[idx] + a.c_res * s);
        }
    }
}

}  // namespace

template <typename T>
hipError_t launch_tfsf(const TfsfArgs<T>& a, hipStream_t s)
{
    if (a.nelem <= 0 || a.nwaves <= 0) return hipSuccess;
    if (4 * a.Nfp > kTfsfMaxFaceNodes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tfsf_apply<T>, dim3(a.nelem), dim3(64), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_tfsf<float>(const TfsfArgs<float>&, hipStream_t);
template hipError_t launch_tfsf<double>(const TfsfArgs<double>&, hipStream_t);

}  // namespace dgtd
