// Device records of the matrix-free hexahedron stage (hex_stage.hip).
#pragma once

namespace dgtd {

// Per element, in the operator's precision: [0..8] dr_k/dx_d (k-major),
// face f: [9+4f..12+4f] nx ny nz fscale/2, [33] 1/eps_r, [34] 1/mu_r,
// [35] sigma/eps_r (the Global operator's conductive term).
constexpr int kHexGeoDev = 36;
// Per element: [0..5] neighbour across face f or -(code+1), [6..11] the
// neighbour's face and orientation g*8 + o (the host record as built).
constexpr int kHexConnDev = 12;

}  // namespace dgtd
