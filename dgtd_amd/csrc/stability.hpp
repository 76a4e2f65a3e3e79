// Device side of the per-step stability test (kernels.hpp kUnstableBound):
// the reference's time loop flags a non-finite state norm or one above 1e20
// (Solver::run, Solver.cpp:499-516); the stage kernels test the entries a
// step's last stage stores: non-finite, or above 1e20 in magnitude.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dgtd {

template <typename T>
__device__ __forceinline__ bool unstable_entry(T v)
{
    return !(fabs(v) <= T(kUnstableBound));  // NaN fails the comparison
}

// One wave-uniform test per call: when any active lane saw an unstable entry,
// the wave sets *flag (a rare event; stable steps issue no atomic).
__device__ __forceinline__ void report_unstable(int* flag, bool bad)
{
    if (flag != nullptr && __builtin_amdgcn_ballot_w64(bad) != 0) atomicOr(flag, 1);
}

}  // namespace dgtd
