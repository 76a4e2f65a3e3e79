// Assembly of the Global operator (the same semi-discrete DG operator as the
// matrix-free kernels, plus bulk conductivity) into CSR, rows and columns in
// the state layout c*N + e*Np + i.
#pragma once

#include <cstdint>
#include <vector>

#include "hex.hpp"
#include "mesh.hpp"
#include "quad.hpp"
#include "tri.hpp"

namespace dgtd {

struct Csr {
    std::vector<int64_t> rowptr;  // 6N + 1
    std::vector<int32_t> col;
    std::vector<double> val;
};

// sigma: bulk conductivity per element (NULL = 0).  TF/SF faces are ordinary
// interior faces here: their injection is a separate, state-independent term.
Csr assemble_global_tet(const MeshData& md, const RefTet& R, const double* sigma, double alpha);
Csr assemble_global_tri(const TriMeshData& md, const RefTri& R, const double* sigma, double alpha);
Csr assemble_global_quad(const TriMeshData& md, const RefQuad& R, const double* sigma, double alpha);
Csr assemble_global_hex(const HexMeshData& md, const RefHex& R, const double* sigma, double alpha);

// Element-local 16-bit column codes of A for the device kernel (layout in
// kernels.hpp: kCsrRowsPerBlock, kCsrTab); N = Np * K.
struct CsrLocal16 {
    std::vector<uint16_t> col;   // one per entry of A
    std::vector<int32_t> etab;   // kCsrTab per block of kCsrRowsPerBlock DOFs
};
CsrLocal16 csr_local16(const Csr& A, int64_t N, int Np);

}  // namespace dgtd
