// Boundary-condition weights of the own trace on a face with code bc
// (connectivity entry -(bc+1)): the jump is (ce * E-, ch * H-).
//   boundary (HesthavenEvolution.cpp:290-292): 1 PEC (-2, 0), 2 PMC (0, -2),
//     3 SMA (-1/alpha, -1/alpha), 0 none (0, 0);
//   interior conditions on both sides of the face (:309-311): 4 PEC (-1, 0),
//     5 PMC (0, -1), 6 SMA (-1/2, -1/2).
#pragma once

#include <hip/hip_runtime.h>

namespace dgtd {

template <typename T>
__device__ __forceinline__ void bc_coeffs(int bc, T alpha, T& ce, T& ch)
{
    const T sma = T(-1) / alpha;
    ce = bc == 1 ? T(-2) : bc == 3 ? sma : bc == 4 ? T(-1) : bc == 6 ? T(-0.5) : T(0);
    ch = bc == 2 ? T(-2) : bc == 3 ? sma : bc == 5 ? T(-1) : bc == 6 ? T(-0.5) : T(0);
}

}  // namespace dgtd
