// Triangles: the 2D variant of the hesthaven operator.  The reference's
// operator is dimension-generic (HesthavenEvolution.cpp:450-542): on a 2D mesh
// the z-derivative operator is empty (DGOperatorFactory.h:428-437), face
// normals have no z component and faces are edges with p+1 nodes
// (HesthavenEvolution.cpp:178-179).  The six field components keep the 3D
// layout; only the element, its faces and its geometry change.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace dgtd {

inline int tri_num_nodes(int p) { return (p + 1) * (p + 2) / 2; }

// Local vertices of edge f, MFEM's triangle edge table: (0,1), (1,2), (2,0),
// and its quadrilateral table: (0,1), (1,2), (2,3), (3,0).
extern const int kTriEdges[3][2];
extern const int kQuadEdgeTable[4][2];

struct RefTri {
    int p = 0, Np = 0, Nfp = 0;
    std::vector<int> lattice;    // 3*Np: lattice index of vertices 0, 1, 2 (sum p)
    std::vector<double> xi;      // 2*Np unit-triangle coordinates (weights of vertices 1, 2)
    std::vector<double> Vinv;    // Np*Np inverse Vandermonde of the orthonormal basis
    std::vector<double> M;       // Np*Np mass on the reference triangle of area 2
    std::vector<double> Dr, Ds;  // Np*Np derivatives in r = 2 xi - 1, s = 2 eta - 1
    std::vector<double> LIFT;    // Np*3Nfp
    std::vector<int> Fmask;      // 3*Nfp, edge nodes from the edge's first vertex to its second
    // Neighbour-node table: own edge f, neighbour edge g, orientation c (0:
    // the neighbour traverses the shared edge in the same vertex order, 1:
    // reversed): entry [((f*3+g)*2+c)*Nfp + b] = neighbour node co-located with
    // own edge node b.
    std::vector<int> edge_perm;  // 18*Nfp
};

RefTri build_reftri(int p);

// Nodal interpolation weights at unit-triangle point (xi, eta).
void tri_interp_weights(const RefTri& R, double xi, double eta, double* w);

// Device records of a triangle, kTriGeoStride values:
//   [0..3]   dr/dx dr/dy ds/dx ds/dy
//   [4..12]  edge f: nx ny fscale/2   (f = 0..2)
//   [13]     1/eps_r   [14] 1/mu_r
constexpr int kTriGeoStride = 15;
// kTriConnStride int32: [0..2] neighbour across edge f or -(bc+1);
// [3..5] neighbour code (f*3 + g)*2 + c into RefTri::edge_perm
constexpr int kTriConnStride = 6;

struct TriMeshInput {
    int K = 0;
    int nv = 0;
    const double* vx = nullptr;      // 3*nv (z ignored)
    const int32_t* EToV = nullptr;   // 3*K
    const int8_t* face_bc = nullptr; // 3*K codes per edge, NULL = none
    const double* eps_r = nullptr;
    const double* mu_r = nullptr;
};

struct TriMeshData {
    int K = 0;
    std::vector<double> geo;    // kTriGeoStride * K
    std::vector<int32_t> conn;  // kTriConnStride * K
    std::vector<double> jac;    // K: area / 2 (reference area 2)
    std::vector<double> eps, mu;
    double max_fscale = 0.0;
    std::vector<int8_t> face_code;  // 3K
    std::vector<std::array<int, 3>> tfsf;  // TF/SF edge sides: (element, edge, -1 scattered / +1 total)
};

// Throws std::runtime_error on invalid input (same rules as the tetrahedral
// mesh: codes 1-3 only on boundary edges, 4-7 only on interior edges; a TF/SF
// edge keeps its neighbour and its lower-index element is the scattered side).
TriMeshData build_tri_mesh(const TriMeshInput& in, const RefTri& R, double alpha);
// The same for NV vertices per element: 3 triangles, 4 parallelograms (records
// in quad.hpp's layout: 18 geometry values, 8 connectivity ints).
TriMeshData build_2d_mesh(const TriMeshInput& in, int NV, double alpha);

// Physical node positions: 3 blocks of Np*K (z block zero).
void tri_node_positions(const TriMeshInput& in, const RefTri& R, double* xyz);

// Point location: lowest-index triangle whose closure holds the point (z
// ignored), -1 if none; w[p*Np..) = interpolation weights there.
void tri_locate_points(const TriMeshInput& in, const RefTri& R, int n, const double* xyz, int32_t* elem,
                       double* w);

}  // namespace dgtd
