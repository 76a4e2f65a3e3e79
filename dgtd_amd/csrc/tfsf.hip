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
// and quadrilaterals (dim 4 here) the faces are the edges and n_z = 0;
// hexahedra are dim 5 here (six faces).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dgtd {

namespace {

constexpr int kTfsfMaxFaceNodes = 6 * 36;  // hexahedra of order 5

// The injection of one element owning TF/SF faces (block-wide, 64 threads):
// record k of the element lists, element e of a mesh with Np nodes per
// element and N nodes per component.
template <typename T>
__device__ void tfsf_element(int dim, int k, int e, const int* sign, const double* xyz, const double* geo,
                             const double* lift, const TfsfWave* waves, int nwaves, int Np, int Nfp, size_t N,
                             double alpha, double t, T* out, double c_out, T* res, double c_res,
                             double (&flux)[6][kTfsfMaxFaceNodes])
{
    // dim 3 tetrahedra (27-value records), 2 triangles (15), 4 quadrilaterals
    // (18), 5 hexahedra (35)
    const bool planar = dim == 2 || dim == 4;
    const int stride = dim == 3 ? 27 : dim == 2 ? 15 : dim == 4 ? 18 : 35;
    const int Nf = (dim == 2 ? 3 : dim == 5 ? 6 : 4) * Nfp;
    const double* g = geo + (size_t)k * stride;
    for (int m = threadIdx.x; m < Nf; m += 64) {
        const int f = m / Nfp;
        const int sg = sign[kTfsfSignStride * k + f];
        double dE[3] = {0.0, 0.0, 0.0}, dH[3] = {0.0, 0.0, 0.0};
        if (sg != 0) {
            const double* x = xyz + ((size_t)k * Nf + m) * 3;
            for (int w = 0; w < nwaves; ++w) {
                const TfsfWave& pw = waves[w];
                const double u = pw.dir[0] * x[0] + pw.dir[1] * x[1] + pw.dir[2] * x[2] - t;
                const double d = u - pw.mean;
                double mag = exp(-d * d / (2.0 * pw.spread * pw.spread));
                if (pw.kind == 1) mag *= cos(2.0 * M_PI * pw.freq * d);
                for (int c = 0; c < 3; ++c) {
                    dE[c] += sg * mag * pw.pol[c];
                    dH[c] += sg * mag * pw.kxp[c];
                }
            }
        }
        // records: tetrahedron and hexahedron [9+4f..12+4f] = nx ny nz
        // fscale/2; triangle and quadrilateral [4+3f..6+3f] = nx ny fscale/2
        const double n[3] = {planar ? g[4 + 3 * f] : g[9 + 4 * f], planar ? g[5 + 3 * f] : g[10 + 4 * f],
                             planar ? 0.0 : g[11 + 4 * f]};
        const double fs = planar ? g[6 + 3 * f] : g[12 + 4 * f];
        const double ndE = n[0] * dE[0] + n[1] * dE[1] + n[2] * dE[2];
        const double ndH = n[0] * dH[0] + n[1] * dH[1] + n[2] * dH[2];
        for (int x = 0; x < 3; ++x) {
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            flux[x][m] = fs * (n[y] * dH[z] - n[z] * dH[y] + alpha * (dE[x] - ndE * n[x]));
            flux[3 + x][m] = fs * (-n[y] * dE[z] + n[z] * dE[y] + alpha * (dH[x] - ndH * n[x]));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Np; i += 64) {
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int m = 0; m < Nf; ++m) s += lift[(size_t)i * Nf + m] * flux[c][m];
            s *= c < 3 ? g[stride - 2] : g[stride - 1];  // 1/eps or 1/mu
            const size_t idx = c * N + (size_t)e * Np + i;
            out[idx] = (T)((double)out[idx] + c_out * s);
            if (res) res[idx] = (T)((double)res[idx] + c_res * s);
        }
    }
    __syncthreads();  // flux is reused by the block's next element
}

template <typename T>
__global__ __launch_bounds__(64) void tfsf_apply(TfsfArgs<T> a)
{
    __shared__ double flux[6][kTfsfMaxFaceNodes];
    const int k = blockIdx.x;
    tfsf_element<T>(a.dim, k, a.elem[k], a.sign, a.xyz, a.geo, a.lift, a.waves, a.nwaves, a.Np, a.Nfp, a.N,
                    a.alpha, a.t, a.out, a.c_out, a.res, a.c_res, flux);
}

// Batch variant: items (instance, record k) over the running instances'
// elements owning TF/SF faces, *count of them; stage `stage` of LSERK4 at the
// instance's own time t + c dt_step, adding dt_step b S to the stage output
// and dt_step S to the residual (as the single operator does).
template <typename T>
__global__ __launch_bounds__(64) void batch_tfsf_apply(BatchTfsfArgs<T> a)
{
    __shared__ double flux[6][kTfsfMaxFaceNodes];
    const int n = *a.count;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const BatchItem w = a.items[it];
        const BatchInst<T>& I = a.inst[w.inst];
        if (I.state != kSlotRunning) continue;  // uniform across the block
        const int dim = I.dim;
        const int Np = dim == 2 ? (I.order + 1) * (I.order + 2) / 2 : (I.order + 1) * (I.order + 2) * (I.order + 3) / 6;
        const int Nfp = dim == 2 ? I.order + 1 : (I.order + 1) * (I.order + 2) / 2;
        const int src = I.cur ^ (a.stage & 1);
        const double dt = I.dt_step;
        tfsf_element<T>(dim, w.blk, I.tf_elem[w.blk], I.tf_sign, I.tf_xyz, I.tf_geo,
                        dim == 2 ? a.lift_tri[I.order] : a.lift[I.order], I.tf_waves, I.tf_nwaves, Np, Nfp,
                        (size_t)Np * I.K, I.alpha, I.t + a.c * dt, I.q[src ^ 1], a.b * dt, I.res, dt, flux);
    }
}

}  // namespace

template <typename T>
hipError_t launch_tfsf(const TfsfArgs<T>& a, hipStream_t s)
{
    if (a.nelem <= 0 || a.nwaves <= 0) return hipSuccess;
    if ((a.dim == 5 ? 6 : 4) * a.Nfp > kTfsfMaxFaceNodes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tfsf_apply<T>, dim3(a.nelem), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_tfsf(const BatchTfsfArgs<T>& a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(batch_tfsf_apply<T>, dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_tfsf<float>(const TfsfArgs<float>&, hipStream_t);
template hipError_t launch_tfsf<double>(const TfsfArgs<double>&, hipStream_t);
template hipError_t launch_batch_tfsf<float>(const BatchTfsfArgs<float>&, int, hipStream_t);
template hipError_t launch_batch_tfsf<double>(const BatchTfsfArgs<double>&, int, hipStream_t);

}  // namespace dgtd
