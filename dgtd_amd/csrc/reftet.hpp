// Reference tetrahedron for the DG operator, built in FP64 on the host once per
// order and uploaded to the device (FP64 or FP32 copies).
#pragma once

#include <cstdint>
#include <vector>

namespace dgtd {

constexpr int kMaxOrder = 5;

inline int num_nodes(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
inline int num_face_nodes(int p) { return (p + 1) * (p + 2) / 2; }

// Local vertices of face f (face f is opposite vertex f), in MFEM's tetrahedron
// face order, which fixes the face numbering the reference iterates in
// (HesthavenEvolutionMethods.cpp:751-759) and the LIFT column blocks
// (HesthavenEvolution.cpp:56-81).
extern const int kFaceVerts[4][3];

struct RefTet {
    int p = 0, Np = 0, Nfp = 0;
    // Barycentric lattice index of each node w.r.t. vertices 0..3 (sum = p).
    std::vector<int> lattice;      // 4*Np
    std::vector<double> xi;        // 3*Np, unit-tetrahedron coordinates
    std::vector<double> Vinv;      // Np*Np inverse Vandermonde of the orthonormal basis
    std::vector<double> M;         // Np*Np mass on the reference tet of volume 4/3
    std::vector<double> Dr, Ds, Dt;// Np*Np strong derivatives in r = 2 xi - 1, ...
    std::vector<double> LIFT;      // Np*4Nfp
    std::vector<int> Fmask;        // 4*Nfp, ascending local node indices per face
    // Neighbour-node table: for own face f, neighbour face g and vertex
    // permutation code c (0..5), entry [((f*4+g)*6+c)*Nfp + b] = local node of
    // the neighbour co-located with own face node b.
    std::vector<int> face_perm;    // 96*Nfp
};

// Orthonormal Jacobi polynomial P_n^{(a,b)} on [-1,1] and its derivative;
// Gauss-Lobatto-Legendre points on [0,1] (p+1 points).
double jac(double x, double a, double b, int n);
double djac(double x, double a, double b, int n);
std::vector<double> gll01(int p);

// Builds the reference element of order p (1..kMaxOrder).  Throws
// std::runtime_error on failure.
RefTet build_reftet(int p);

// Weights w[0..Np) of the nodal interpolant at unit-tetrahedron point xi:
// u(xi) = sum_n w[n] u_n (orthonormal basis at xi times the inverse Vandermonde).
void interp_weights(const RefTet& R, const double xi[3], double* w);

// Index 0..5 of the permutation pi with own face vertex k equal to neighbour
// face vertex pi[k]; -1 if `pi` is not a permutation of {0,1,2}.
int perm_code(const int pi[3]);
void perm_from_code(int code, int pi[3]);

}  // namespace dgtd
