// Block-wide copy of n values from global memory into LDS.  Each thread keeps
// CHUNK loads in flight before their LDS writes: a plain strided loop waits
// for each load before its write, so the memory latencies add up, one per
// value a thread copies.
#pragma once

#include <hip/hip_runtime.h>

namespace dgtd {

template <int CHUNK, typename V>
__device__ __forceinline__ void lds_copy(V* __restrict__ dst, const V* __restrict__ src, int n, int tid, int nthreads)
{
    for (int base = 0; base < n; base += CHUNK * nthreads) {
        V v[CHUNK];
#pragma unroll
        for (int k = 0; k < CHUNK; ++k) {
            const int i = base + k * nthreads + tid;
            v[k] = src[i < n ? i : n - 1];  // clamped: every load unconditional and in bounds
        }
#pragma unroll
        for (int k = 0; k < CHUNK; ++k) {
            const int i = base + k * nthreads + tid;
            if (i < n) dst[i] = v[k];
        }
    }
}

}  // namespace dgtd
