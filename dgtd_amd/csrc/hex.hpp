// Hexahedra on the assembled (Global) operator.  The reference's default
// evolution operator assembles the DG operator of its L2 Gauss-Lobatto space
// (EvolutionOptions.h:14, Solver.cpp:27-28) on hexahedral meshes as on
// tetrahedra; its Solver3DTest cases run it (pec_global_1dot5D and
// sma_upwind_hexa_1dot5D, Solver3DTest.cpp:57-200).  Here the reference
// element is the tensor product of the p+1 Gauss-Lobatto-Legendre points,
// nodes in MFEM's lexicographic order, and the elements are parallelepipeds
// (affine maps).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "mesh.hpp"

namespace dgtd {

inline int hex_num_nodes(int p) { return (p + 1) * (p + 1) * (p + 1); }

// MFEM's hexahedron: vertices 0 (0,0,0), 1 (1,0,0), 2 (1,1,0), 3 (0,1,0),
// 4 (0,0,1), 5 (1,0,1), 6 (1,1,1), 7 (0,1,1) of the unit cube, faces as
// vertex cycles with outward normals (b - a) x (d - a).
extern const int kHexFaces[6][4];

// Records of a hexahedron, kHexGeoStride values:
//   [0..8] dr_k/dx_d (k-major), face f: [9+4f..12+4f] nx ny nz fscale/2,
//   [33] 1/eps_r, [34] 1/mu_r
constexpr int kHexGeoStride = 35;
// kHexConnStride int32: [0..5] neighbour across face f or -(bc+1), [6..11]
// neighbour code g*8 + o (g the neighbour's face, o the orientation) into
// RefHex::face_perm
constexpr int kHexConnStride = 12;

struct RefHex {
    int p = 0, Np = 0, Nfp = 0;
    std::vector<double> x1;      // p+1 Gauss-Lobatto-Legendre points on [-1, 1]
    std::vector<double> xi;      // 3*Np unit-cube coordinates of node i + (p+1) j + (p+1)^2 k
    std::vector<double> M;       // Np*Np mass on [-1, 1]^3
    std::vector<double> Dr, Ds, Dt;
    std::vector<double> LIFT;    // Np*6Nfp
    // face f node m + (p+1) n at vertex a + u_m (b - a) + u_n (d - a) of the
    // face's cycle (a, b, c, d), u = (x1 + 1)/2
    std::vector<int> Fmask;      // 6*Nfp
    // [(g*8 + o)*Nfp + b]: the neighbour's volume node at own face node b when
    // the neighbour's face g meets it in orientation o = 2 r + s: the own
    // face's first vertex is the neighbour cycle's vertex r, its second the
    // next (s = 0) or the previous (s = 1) vertex of that cycle
    std::vector<int> face_perm;  // 48*Nfp
};

RefHex build_refhex(int p);

// Nodal interpolation weights at unit-cube point (xi, eta, zeta).
void hex_interp_weights(const RefHex& R, double xi, double eta, double zeta, double* w);

struct HexMeshData {
    int K = 0;
    std::vector<double> geo;    // kHexGeoStride * K
    std::vector<int32_t> conn;  // kHexConnStride * K
    std::vector<double> jac;    // K: volume / 8 (reference volume 8)
    std::vector<double> eps, mu;
    double max_fscale = 0.0;
    std::vector<int8_t> face_code;         // 6K
    std::vector<std::array<int, 3>> tfsf;  // TF/SF face sides: (element, face, -1 scattered / +1 total)
};

// Mesh setup with the tetrahedra's rules (codes 1-3 on boundary faces, 4-7 on
// interior ones, TF/SF faces keep their neighbour, lower-index element on the
// scattered side).  MeshInput with EToV holding 8 vertex ids per element in
// MFEM's order.  Throws on a non-parallelepiped or inverted element.
HexMeshData build_hex_mesh(const MeshInput& in, const RefHex& R, double alpha);

void hex_node_positions(const MeshInput& in, const RefHex& R, double* xyz);

void hex_locate_points(const MeshInput& in, const RefHex& R, int n, const double* xyz, int32_t* elem, double* w);

}  // namespace dgtd
