// Stage kernel over the assembled Global operator (GlobalEvolution::Mult,
// GlobalEvolution.cpp:628-1088, out = A [E; H]) in CSR: 16 lanes per row
// stride over the row's entries (column-gathered state values), a 16-lane
// reduction, then the lane with the sum applies the stage update of that DOF
// (rows are in the state layout).  HBM-bound on the matrix: per entry one
// value and one 32-bit column index.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dgtd {

namespace {

constexpr int kLanesPerRow = 16;

template <typename T, int MODE>
__global__ __launch_bounds__(256) void csr_stage(StageArgs<T> a)
{
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanesPerRow;
    const int sub = threadIdx.x % kLanesPerRow;
    const int64_t r = row < a.nrows ? row : a.nrows - 1;  // tail rows repeat the last row (no branch)
    const int64_t b = a.rowptr[r], e = a.rowptr[r + 1];
    T acc = T(0);
    for (int64_t k = b + sub; k < e; k += kLanesPerRow) acc += a.val[k] * a.qin[a.col[k]];
#pragma unroll
    for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kLanesPerRow);
    if (sub != 0 || row >= a.nrows) return;
    const T rhs = acc;
    if constexpr (MODE == kModeMult) {
        a.out[r] = rhs;
    } else if constexpr (MODE == kModeLserk) {
        const T rr = a.ca * a.res[r] + a.dt * rhs;
        a.res[r] = rr;
        a.out[r] = a.qin[r] + a.cb * rr;
    } else if constexpr (MODE == kModeRk1) {
        const T q = a.qin[r];
        a.out[r] = q + a.ca * rhs;
        a.res[r] = q + a.cb * rhs;
    } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
        const T xv = a.x[r], rv = a.res[r];
        a.out[r] = xv + a.ca * rhs;
        a.res[r] = rv + a.cb * rhs;
    } else {
        a.out[r] = a.res[r] + a.ca * rhs;
    }
}

template <typename T, int MODE>
hipError_t csr_launch(const StageArgs<T>& a, hipStream_t s)
{
    const int64_t threads = a.nrows * kLanesPerRow;
    const int64_t grid = (threads + 255) / 256;
    hipLaunchKernelGGL((csr_stage<T, MODE>), dim3((unsigned)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_csr_stage(const StageArgs<T>& a, int mode, hipStream_t s)
{
    if (a.nrows <= 0) return hipErrorInvalidValue;
    switch (mode) {
        case kModeMult: return csr_launch<T, kModeMult>(a, s);
        case kModeLserk: return csr_launch<T, kModeLserk>(a, s);
        case kModeRk1: return csr_launch<T, kModeRk1>(a, s);
        case kModeRk2: return csr_launch<T, kModeRk2>(a, s);
        case kModeRk3: return csr_launch<T, kModeRk3>(a, s);
        case kModeRk4: return csr_launch<T, kModeRk4>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_csr_stage<float>(const StageArgs<float>&, int, hipStream_t);
template hipError_t launch_csr_stage<double>(const StageArgs<double>&, int, hipStream_t);

}  // namespace dgtd
