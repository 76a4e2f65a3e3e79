// Stage kernel over the assembled Global operator (GlobalEvolution::Mult,
// GlobalEvolution.cpp:628-1088, out = A [E; H]) in CSR: a block of 256
// threads takes 256 / LPR rows (LPR lanes per row: 16 for the long rows of
// tetrahedra, 8 or 4 for the short rows of tensor-product elements and low
// orders), streams their entries (value, 16-bit column code, gathered state
// value) into LDS products, the LPR lanes of a row sum them, and the lane with
// the sum applies the stage update of that DOF (rows are in the state layout).
// HBM-bound on the matrix: per entry one value and a 16-bit element-local
// column code, decoded through the element table of the entry's 16-DOF group
// (kCsrRowsPerBlock) in LDS.
//
// Block order: a block takes consecutive DOFs of one field component, and
// consecutive logical blocks cycle through the six components of the same
// DOFs, so a run of logical blocks needs the state of one element range (and
// its neighbours) only.  Hardware block b runs on XCD b % 8; logical block
// (b % 8) * (grid / 8) + b / 8 gives each XCD one contiguous run, whose
// gathered state then stays in that XCD's L2.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "stability.hpp"

namespace dgtd {

namespace {

constexpr int kXcds = 8;

// kUnroll: entries per thread and chunk (the chunk is 256 * kUnroll entries).
template <typename T, int MODE, int kUnroll, int LPR>
__global__ __launch_bounds__(256) void csr_stage(StageArgs<T> a)
{
    constexpr int kRows = 256 / LPR;
    constexpr int kGroups = kRows / kCsrRowsPerBlock;
    static_assert(kGroups >= 1 && kRows % kCsrRowsPerBlock == 0, "whole 16-DOF groups per block");
    constexpr int kChunk = 256 * kUnroll;
    __shared__ T prod[kChunk];
    __shared__ int tab[kGroups * kCsrTab];
    __shared__ int64_t gbound[kGroups + 1];
    const int t = threadIdx.x;
    const int64_t per = gridDim.x / kXcds;  // the grid is a multiple of kXcds
    const int64_t lb = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    const int64_t N = a.nrows / 6;
    const int64_t dof0 = (lb / 6) * kRows;
    const int64_t nrow = dof0 < N ? (N - dof0 < kRows ? N - dof0 : kRows) : 0;
    const int64_t row0 = (lb % 6) * N + (nrow > 0 ? dof0 : 0);
    const int64_t dof = dof0 + t / LPR;
    const bool valid = dof < N;  // padding blocks and the tail repeat the block's first row (no branch)
    const int64_t r = valid ? (lb % 6) * N + dof : row0;
    const int sub = t % LPR;
    // the element tables of the block's groups (clamped to the last group) and,
    // with several groups, the entry ranges where each group's rows start
    const int64_t ng = (N + kCsrRowsPerBlock - 1) / kCsrRowsPerBlock;
    if (t < kGroups * kCsrTab) {
        int64_t g = dof0 / kCsrRowsPerBlock + t / kCsrTab;
        g = g < ng ? g : ng - 1;
        tab[t] = a.etab[g * kCsrTab + t % kCsrTab];
    }
    if (kGroups > 1 && t <= kGroups) {
        const int64_t off = (int64_t)t * kCsrRowsPerBlock;
        gbound[t] = a.rowptr[row0 + (off < nrow ? off : nrow)];
    }
    __syncthreads();
    const int n32 = (int)N;  // 6 N < 2^31 (checked at assembly)
    // The block's rows are consecutive, so their entries are one contiguous
    // range [b0, b1): it streams through in chunks of kChunk entries, every
    // thread issuing kUnroll independent loads of (column, value) before their
    // gathers; the products go to LDS and each row sums its own part.
    const int64_t b0 = a.rowptr[row0], b1 = a.rowptr[row0 + nrow];
    const int64_t rb = a.rowptr[r], re = a.rowptr[r + 1];
    // The row's own stage operands do not depend on the product.  With short
    // rows and small chunks their loads are issued here and complete while the
    // entries stream (measured: order-3 quadrilaterals FP32 241 -> 219 us,
    // FP64 304 -> 286 us; hexahedra FP32 571 -> 493 us); with larger chunks
    // the two or four extra registers cost waves per SIMD and the loads stay
    // after the product.
    constexpr bool kEarly = LPR < 16 && kUnroll <= (sizeof(T) == 4 ? 5 : 4);
    T own0 = T(0), own1 = T(0);
    auto load_own = [&]() {
        if constexpr (MODE == kModeLserk) {
            own0 = a.res[r];
            own1 = a.qin[r];
        } else if constexpr (MODE == kModeRk1) {
            own0 = a.qin[r];
        } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
            own0 = a.x[r];
            own1 = a.res[r];
        } else if constexpr (MODE != kModeMult) {
            own0 = a.res[r];
        }
    };
    if constexpr (kEarly) load_own();
    T acc = T(0);
    for (int64_t c0 = b0; c0 < b1; c0 += kChunk) {  // block-uniform trip count
        unsigned c[kUnroll];
        T v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t k = c0 + u * 256 + t;
            const int64_t kk = k < b1 ? k : b1 - 1;
            c[u] = __builtin_nontemporal_load(a.col16 + kk);
            v[u] = __builtin_nontemporal_load(a.val + kk);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t k = c0 + u * 256 + t;
            const int64_t kk = k < b1 ? k : b1 - 1;  // the entry loaded above
            int g = 0;
#pragma unroll
            for (int j = 1; j < kGroups; ++j) g += kk >= gbound[j] ? 1 : 0;
            const int col = (int)(c[u] >> kCsrCompShift & 7) * n32 + tab[g * kCsrTab + (c[u] >> kCsrSlotShift)] * a.np +
                            (int)(c[u] & ((1u << kCsrCompShift) - 1));
            const T pr = v[u] * a.qin[col];
            prod[u * 256 + t] = k < b1 ? pr : T(0);
        }
        __syncthreads();
        const int64_t lo = rb > c0 ? rb : c0;
        const int64_t hi = re < c0 + kChunk ? re : c0 + kChunk;
        for (int64_t k = lo + sub; k < hi; k += LPR) acc += prod[k - c0];
        __syncthreads();
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (sub != 0 || !valid) return;
    if constexpr (!kEarly) load_own();
    const T rhs = acc;
    // stability test on the step's last stage (rare atomic; unstable_entry)
    auto check = [&](T qn) {
        if (a.unstable != nullptr && unstable_entry(qn)) atomicOr(a.unstable, 1);
    };
    if constexpr (MODE == kModeMult) {
        a.out[r] = rhs;
    } else if constexpr (MODE == kModeLserk) {
        const T rr = a.ca * own0 + a.dt * rhs;
        a.res[r] = rr;
        const T qn = own1 + a.cb * rr;
        a.out[r] = qn;
        check(qn);
    } else if constexpr (MODE == kModeRk1) {
        a.out[r] = own0 + a.ca * rhs;
        a.res[r] = own0 + a.cb * rhs;
    } else if constexpr (MODE == kModeRk2 || MODE == kModeRk3) {
        a.out[r] = own0 + a.ca * rhs;
        a.res[r] = own1 + a.cb * rhs;
    } else {
        const T qn = own0 + a.ca * rhs;
        a.out[r] = qn;
        check(qn);
    }
}

template <typename T, int MODE, int U, int LPR>
hipError_t csr_go(const StageArgs<T>& a, hipStream_t s)
{
    constexpr int kRows = 256 / LPR;
    const int64_t N = a.nrows / 6;
    const int64_t blocks = 6 * ((N + kRows - 1) / kRows);
    const int64_t grid = kXcds * ((blocks + kXcds - 1) / kXcds);
    hipLaunchKernelGGL((csr_stage<T, MODE, U, LPR>), dim3((unsigned)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T, int MODE>
hipError_t csr_launch(const StageArgs<T>& a, hipStream_t s)
{
    // Lanes per row by the average row length: short rows (tensor-product
    // elements: 13 entries per row on order-3 quadrilaterals, 18 on order-3
    // hexahedra; low orders) share a block among 64 or 32 rows, so a block
    // still streams a full chunk; long rows (156 on order-3 tetrahedra) take
    // 16 lanes.  One chunk per block where the rows allow: the smallest chunk
    // holding an average block's entries, at most 4096 entries.
    const double avg = (double)a.nnz / (double)a.nrows;
    if (avg <= 24.0) {
        const double per_block = avg * 64;
        if (per_block <= 256 * 2) return csr_go<T, MODE, 2, 4>(a, s);
        if (per_block <= 256 * 3) return csr_go<T, MODE, 3, 4>(a, s);
        if (per_block <= 256 * 4) return csr_go<T, MODE, 4, 4>(a, s);
        if (per_block <= 256 * 5) return csr_go<T, MODE, 5, 4>(a, s);
        return csr_go<T, MODE, 6, 4>(a, s);
    }
    if (avg <= 48.0) {
        const double per_block = avg * 32;
        if (per_block <= 256 * 4) return csr_go<T, MODE, 4, 8>(a, s);
        if (per_block <= 256 * 5) return csr_go<T, MODE, 5, 8>(a, s);
        return csr_go<T, MODE, 6, 8>(a, s);
    }
    const double per_block = avg * 16;
    if (per_block <= 256 * 5) return csr_go<T, MODE, 5, 16>(a, s);
    if (per_block <= 256 * 8) return csr_go<T, MODE, 8, 16>(a, s);
    if (sizeof(T) == 4 && per_block <= 256 * 10)  // FP64 measured faster with the 3072-entry chunk
        return csr_go<T, MODE, 10, 16>(a, s);
    if (per_block <= 256 * 12) return csr_go<T, MODE, 12, 16>(a, s);
    return csr_go<T, MODE, 16, 16>(a, s);
}

}  // namespace

template <typename T>
hipError_t launch_csr_stage(const StageArgs<T>& a, int mode, hipStream_t s)
{
    if (a.nrows <= 0 || a.nrows % 6 != 0) return hipErrorInvalidValue;
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
