// Tetrahedral mesh files the reference's launcher reads (driver.cpp:1083-1188):
// Gmsh 2.2 ASCII (.msh) and MFEM v1.0 (.mesh), plus uniform refinement.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dgtd {

// Input the readers understand but the operator does not support (curved
// meshes, non-tetrahedral elements).
struct MeshUnsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// A mesh of the launcher.  dim 3: tetrahedra are the elements, triangles the
// boundary (and interior-boundary) elements, or hexahedra with
// quadrilateral boundary elements (never both).  dim 2: triangles, or
// quadrilaterals (never both), are the elements, segments the boundary
// elements.
struct TetMesh {
    int dim = 3;
    std::vector<double> vx;                 // 3 per vertex (z = 0 in 2D)
    std::vector<std::array<int32_t, 4>> tets;
    std::vector<int32_t> tet_attr;
    std::vector<std::array<int32_t, 3>> tris;
    std::vector<int32_t> tri_attr;
    std::vector<std::array<int32_t, 4>> quads;  // 2D elements (counter-clockwise), or 3D boundary of hexahedra
    std::vector<int32_t> quad_attr;
    std::vector<std::array<int32_t, 8>> hexes;  // 3D elements in MFEM's vertex order
    std::vector<int32_t> hex_attr;
    std::vector<std::array<int32_t, 2>> segs;  // 2D boundary elements
    std::vector<int32_t> seg_attr;

    bool quad() const { return dim == 2 && !quads.empty(); }
    bool hex() const { return dim == 3 && !hexes.empty(); }
    int nvert() const { return hex() ? 8 : dim == 2 && !quad() ? 3 : 4; }  // vertices per element
    int nfaces() const { return hex() ? 6 : nvert(); }
    int32_t nv() const { return (int32_t)(vx.size() / 3); }
    int32_t K() const
    {
        return dim == 2 ? (int32_t)(quad() ? quads.size() : tris.size()) : (int32_t)(hex() ? hexes.size() : tets.size());
    }
    const std::vector<int32_t>& elem_attr() const
    {
        return dim == 2 ? (quad() ? quad_attr : tri_attr) : hex() ? hex_attr : tet_attr;
    }
    const std::vector<int32_t>& bdr_attr() const { return dim == 2 ? seg_attr : hex() ? quad_attr : tri_attr; }
    // local vertex j of element e (j < nvert())
    int32_t vertex(int32_t e, int j) const
    {
        return dim == 3 ? (hex() ? hexes[e][j] : tets[e][j]) : quad() ? quads[e][j] : tris[e][j];
    }
};

// Gmsh 2.2 ASCII.  Tetrahedra (type 4) are elements, triangles (type 2)
// boundary elements, or hexahedra (type 5) with quadrilateral (type 3)
// boundary elements; points and lines are dropped, as the reference's
// fixGmshMesh does for 3D meshes (driver.cpp:1076-1080).  A file without
// tetrahedra is a 2D mesh: triangles, or quadrilaterals (type 3), are the
// elements and lines (type 1) the boundary elements.  The attribute is
// the elementary-entity tag (the reference swaps it into the physical-tag
// slot before MFEM reads the file, driver.cpp:1140-1142).
TetMesh read_gmsh22(const std::string& path);

// MFEM mesh v1.0 with tetrahedra (geometry 4) and triangle boundary elements
// (geometry 2), or hexahedra (geometry 5) and square boundary elements
// (geometry 3), or in 2D triangles or quadrilaterals (geometry 3) with
// segment boundary elements (geometry 1); straight-sided ("vertices" section; curved "nodes" refused).
TetMesh read_mfem_mesh(const std::string& path);

// Writes the mesh in MFEM v1.0 format (elements with attributes, boundary
// triangles with attributes, vertices), 16 significant digits, as
// Mesh::Save(path, 16) does for the reference's snapshot meshes
// (ProbesManager.cpp:481).
void write_mfem_mesh(const TetMesh& m, const std::string& path);

// Dispatch on the extension (.msh / .mesh), as driver.cpp:1276-1282.
TetMesh read_mesh_file(const std::string& path);

// Make every tetrahedron positively oriented (swap local vertices 2 and 3);
// every hexahedron (mirror: swap local vertices 1 and 3, 5 and 7); in 2D
// every triangle (swap local vertices 0 and 1) and every quadrilateral
// (reverse the cycle: swap local vertices 1 and 3).
void orient_positive(TetMesh& m);

// One level of uniform (1:8) refinement (tetrahedra only): edge midpoints, four corner
// children and the inner octahedron split along its shortest diagonal;
// boundary triangles split into four with the parent's attribute.  Children
// are positively oriented when the parent is.
TetMesh refine_uniform(const TetMesh& m);

// face_bc[4e+f] from the triangles' attributes via attr_to_bc (index =
// attribute, value DGTD_BC_PEC/PMC/SMA, DGTD_BC_TFSF or -1 for "no
// condition"); a PEC/PMC/SMA triangle on an interior face gives the interior
// code (value + 3) to both sides, as the reference marks such tags interior
// (driver.cpp:1024-1035); TF/SF triangles must be interior.  Throws if a
// triangle matches no element face.
// On hexahedra the same rules apply to boundary quadrilaterals and the six
// faces of MFEM's hexahedron, face_bc[6e+f].
// In 2D the same rules apply to segments and element edges, face_bc[n e+f]
// with edge f = (v_f, v_{f+1 mod n}), n = 3 for triangles, 4 for
// quadrilaterals.
std::vector<int8_t> face_codes(const TetMesh& m, const std::vector<int>& attr_to_bc);

}  // namespace dgtd
