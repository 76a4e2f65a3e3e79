// Host FP64 setup of one tetrahedral mesh: per-element geometric factors and
// face connectivity in the compact form the device kernels read.
#pragma once

#include <cstdint>
#include <array>
#include <vector>

#include "reftet.hpp"

namespace dgtd {

// Per-element record of geometric factors, kGeoStride values:
//   [0..8]   dr/dx dr/dy dr/dz ds/dx ds/dy ds/dz dt/dx dt/dy dt/dz
//   [9..24]  face f: nx ny nz fscale/2   (f = 0..3)
//   [25]     1/eps_r   [26] 1/mu_r
constexpr int kGeoStride = 27;
// Per-element connectivity record, kConnStride int32:
//   [0..3]  neighbour element across face f, or -(bc+1) on a boundary face
//   [4..7]  neighbour code ((f*4 + g)*6 + perm), g = neighbour's local face
constexpr int kConnStride = 8;

struct MeshInput {
    int K = 0;
    int nv = 0;
    const double* vx = nullptr;
    const int32_t* EToV = nullptr;
    const int8_t* face_bc = nullptr;
    const double* eps_r = nullptr;
    const double* mu_r = nullptr;
};

struct MeshData {
    int K = 0;
    std::vector<double> geo;     // kGeoStride * K
    std::vector<int32_t> conn;   // kConnStride * K
    std::vector<double> jac;     // K: V_e / V_ref, for mass-weighted reductions
    std::vector<double> eps, mu; // K
    double max_fscale = 0.0;
    int n_boundary_faces[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per face code (interior codes count each side)
    std::vector<int8_t> face_code;              // 4K: the code of each face (0 plain interior)
    std::vector<std::array<int, 3>> tfsf;       // TF/SF face sides: (element, face, -1 scattered / +1 total)
};

// Throws std::runtime_error (message is user-facing) on invalid input.
MeshData build_mesh(const MeshInput& in, const RefTet& R, double alpha);

// Physical node positions (3 blocks of Np*K).
void node_positions(const MeshInput& in, const RefTet& R, double* xyz);

// Point location for probes: elem[p] = lowest-index element whose closure
// holds point p (-1 if none), w[p*Np..) = nodal interpolation weights there.
void locate_points(const MeshInput& in, const RefTet& R, int n, const double* xyz, int32_t* elem, double* w);

// n^3 hexahedra on [0,L]^3 split into six tetrahedra each (see dgtd_hip.h).
void cube_mesh(int n, double L, const int8_t side_bc[6], double* vx, int32_t* EToV,
               int8_t* face_bc);

}  // namespace dgtd
