// Quadrilaterals on the assembled (Global) operator.  The reference's default
// evolution operator assembles the DG operator of its L2 Gauss-Lobatto space
// (EvolutionOptions.h:14, Solver.cpp:27-28) on quadrilateral meshes as on
// triangles; its 2D_TFSF_Quads cases run it (ExtensiveCasesTest.cpp:1003,
// 1082).  Here the reference element is the tensor product of the p+1
// Gauss-Lobatto-Legendre points, nodes in MFEM's lexicographic order, and the
// elements are parallelograms (affine maps).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "tri.hpp"

namespace dgtd {

inline int quad_num_nodes(int p) { return (p + 1) * (p + 1); }

// Local vertices of edge f, MFEM's quadrilateral edge table.
extern const int kQuadEdges[4][2];

// Device-side records of a quadrilateral (the triangle's layout with a
// fourth edge), kQuadGeoStride values:
//   [0..3] dr/dx dr/dy ds/dx ds/dy, [4+3f..6+3f] edge f: nx ny fscale/2,
//   [16] 1/eps_r, [17] 1/mu_r
constexpr int kQuadGeoStride = 18;
// kQuadConnStride int32: [0..3] neighbour across edge f or -(bc+1), [4..7]
// neighbour code (f*4 + g)*2 + c into RefQuad::edge_perm
constexpr int kQuadConnStride = 8;

struct RefQuad {
    int p = 0, Np = 0, Nfp = 0;
    std::vector<double> x1;      // p+1 Gauss-Lobatto-Legendre points on [-1, 1]
    std::vector<double> xi;      // 2*Np unit-square coordinates of node i + (p+1) j
    std::vector<double> M;       // Np*Np mass on [-1, 1]^2
    std::vector<double> Dr, Ds;  // Np*Np derivatives in r, s
    std::vector<double> LIFT;    // Np*4Nfp
    std::vector<int> Fmask;      // 4*Nfp, edge nodes from the edge's first vertex to its second
    std::vector<int> edge_perm;  // 32*Nfp: [((f*4+g)*2+c)*Nfp + b] = neighbour node at own edge node b
};

RefQuad build_refquad(int p);

// The 1D Gauss-Lobatto-Legendre element of order p on [-1, 1]: points,
// differentiation D[i][j] = l_j'(x_i), exact mass and its inverse (row-major
// (p+1) x (p+1)); the factors of the tensor-product quadrilateral and
// hexahedron.
struct Gll1D {
    int p = 0;
    std::vector<double> x, D, M, Minv;
};
Gll1D build_gll1d(int p);
// Lagrange basis on the points x, evaluated at y.
std::vector<double> gll_lagrange(const std::vector<double>& x, double y);

// Nodal interpolation weights at unit-square point (xi, eta).
void quad_interp_weights(const RefQuad& R, double xi, double eta, double* w);

// Mesh setup with the triangle's rules (codes 1-3 on boundary edges, 4-7 on
// interior ones, TF/SF edges keep their neighbour, lower-index element on
// the scattered side); records in the quadrilateral layout above.  Throws on
// a non-parallelogram element.  TriMeshInput with EToV holding 4 vertex ids
// per element (counter-clockwise or clockwise, as MFEM reads them).
TriMeshData build_quad_mesh(const TriMeshInput& in, const RefQuad& R, double alpha);

void quad_node_positions(const TriMeshInput& in, const RefQuad& R, double* xyz);

void quad_locate_points(const TriMeshInput& in, const RefQuad& R, int n, const double* xyz, int32_t* elem,
                        double* w);

}  // namespace dgtd
