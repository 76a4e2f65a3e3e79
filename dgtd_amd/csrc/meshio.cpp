// Mesh file readers and uniform refinement (see meshio.hpp).
#include "meshio.hpp"

#include "hex.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace dgtd {

namespace {

[[noreturn]] void bad(const std::string& path, const std::string& what)
{
    throw std::runtime_error(path + ": " + what);
}

double signed_volume6(const double* vx, const std::array<int32_t, 4>& t)
{
    const double* a = vx + 3 * t[0];
    double u[3], v[3], w[3];
    for (int d = 0; d < 3; ++d) {
        u[d] = vx[3 * t[1] + d] - a[d];
        v[d] = vx[3 * t[2] + d] - a[d];
        w[d] = vx[3 * t[3] + d] - a[d];
    }
    return u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
}

bool ends_with(const std::string& s, const std::string& suf)
{
    return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

}  // namespace

TetMesh read_gmsh22(const std::string& path)
{
    std::ifstream f(path);
    if (!f) bad(path, "cannot open");
    TetMesh m;
    std::unordered_map<long long, int32_t> node_index;
    std::string line;
    bool have_format = false, have_nodes = false, have_elements = false;
    while (std::getline(f, line)) {
        if (line.rfind("$MeshFormat", 0) == 0) {
            std::getline(f, line);
            std::istringstream is(line);
            double ver = 0;
            int ftype = -1;
            is >> ver >> ftype;
            if (ver < 2.0 || ver >= 3.0) bad(path, "only Gmsh format 2.x is supported");
            if (ftype != 0) bad(path, "binary Gmsh files are not supported");
            have_format = true;
        } else if (line.rfind("$Nodes", 0) == 0) {
            long long n = 0;
            f >> n;
            if (n <= 0) bad(path, "bad node count");
            m.vx.resize(3 * (size_t)n);
            for (long long i = 0; i < n; ++i) {
                long long id;
                double x, y, z;
                if (!(f >> id >> x >> y >> z)) bad(path, "truncated $Nodes");
                node_index[id] = (int32_t)i;
                m.vx[3 * i] = x;
                m.vx[3 * i + 1] = y;
                m.vx[3 * i + 2] = z;
            }
            have_nodes = true;
        } else if (line.rfind("$Elements", 0) == 0) {
            if (!have_nodes) bad(path, "$Elements before $Nodes");
            long long n = 0;
            f >> n;
            std::getline(f, line);
            for (long long i = 0; i < n; ++i) {
                if (!std::getline(f, line)) bad(path, "truncated $Elements");
                std::istringstream is(line);
                long long id;
                int type, ntags;
                if (!(is >> id >> type >> ntags) || ntags < 0) bad(path, "bad element line");
                std::vector<long long> tags(ntags);
                for (auto& t : tags) is >> t;
                const int attr = ntags >= 2 ? (int)tags[1] : (ntags == 1 ? (int)tags[0] : 1);
                int nn = 0;
                if (type == 5) nn = 8;
                else if (type == 4 || type == 3) nn = 4;
                else if (type == 2) nn = 3;
                else if (type == 1) nn = 2;
                else if (type == 15) continue;  // points dropped
                else
                    throw MeshUnsupported(path + ": element type " + std::to_string(type) +
                                          " is not supported (straight-sided tetrahedra, hexahedra, triangles "
                                          "and quadrilaterals only)");
                int32_t vs[8];
                for (int k = 0; k < nn; ++k) {
                    long long nid;
                    if (!(is >> nid)) bad(path, "bad element line");
                    auto it = node_index.find(nid);
                    if (it == node_index.end()) bad(path, "element references unknown node");
                    vs[k] = it->second;
                }
                if (type == 5) {  // Gmsh's hexahedron vertex order is MFEM's
                    m.hexes.push_back({vs[0], vs[1], vs[2], vs[3], vs[4], vs[5], vs[6], vs[7]});
                    m.hex_attr.push_back(attr);
                } else if (type == 3) {
                    m.quads.push_back({vs[0], vs[1], vs[2], vs[3]});
                    m.quad_attr.push_back(attr);
                } else if (nn == 4) {
                    m.tets.push_back({vs[0], vs[1], vs[2], vs[3]});
                    m.tet_attr.push_back(attr);
                } else if (nn == 3) {
                    m.tris.push_back({vs[0], vs[1], vs[2]});
                    m.tri_attr.push_back(attr);
                } else {
                    m.segs.push_back({vs[0], vs[1]});
                    m.seg_attr.push_back(attr);
                }
            }
            have_elements = true;
        }
    }
    if (!have_format || !have_nodes || !have_elements) bad(path, "missing $MeshFormat, $Nodes or $Elements");
    if (!m.hexes.empty()) {
        if (!m.tets.empty())
            throw MeshUnsupported(path + ": meshes mixing tetrahedra and hexahedra are not supported");
        if (!m.tris.empty())
            throw MeshUnsupported(path + ": triangle faces beside hexahedra (prisms, pyramids) are not supported");
        m.dim = 3;  // hexahedra are the elements, quadrilaterals the boundary
        m.segs.clear();
        m.seg_attr.clear();
    } else if (m.tets.empty()) {
        if (m.tris.empty() && m.quads.empty()) bad(path, "no tetrahedra, hexahedra, triangles or quadrilaterals");
        if (!m.tris.empty() && !m.quads.empty())
            throw MeshUnsupported(path + ": meshes mixing triangles and quadrilaterals are not supported");
        m.dim = 2;  // triangles or quadrilaterals are the elements, lines the boundary
    } else {
        if (!m.quads.empty())
            throw MeshUnsupported(path + ": quadrilateral faces beside tetrahedra (prisms, pyramids) are not supported");
        m.segs.clear();  // lines of a 3D mesh are dropped
        m.seg_attr.clear();
    }
    return m;
}

TetMesh read_mfem_mesh(const std::string& path)
{
    std::ifstream f(path);
    if (!f) bad(path, "cannot open");
    // tokens, with '#' comments removed
    std::vector<std::string> tok;
    std::string line;
    bool first = true;
    while (std::getline(f, line)) {
        if (first) {
            if (line.rfind("MFEM mesh v1.", 0) != 0) bad(path, "not an MFEM v1.x mesh");
            first = false;
            continue;
        }
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::string t;
        while (is >> t) tok.push_back(t);
    }
    size_t p = 0;
    auto next = [&]() -> const std::string& {
        if (p >= tok.size()) bad(path, "unexpected end of file");
        return tok[p++];
    };
    auto next_int = [&]() { return std::stol(next()); };
    TetMesh m;
    bool have_vertices = false, have_dim = false;
    while (p < tok.size()) {
        const std::string sec = next();
        if (sec == "dimension") {
            const long d = next_int();
            if (d != 3 && d != 2) bad(path, "only 2D and 3D meshes are supported");
            m.dim = (int)d;
            have_dim = true;
        } else if (sec == "elements") {
            if (!have_dim) bad(path, "elements before dimension");
            const long n = next_int();
            for (long i = 0; i < n; ++i) {
                const int attr = (int)next_int();
                const int geom = (int)next_int();
                // tetrahedra or cubes (hexahedra), or in 2D triangles or squares
                if (m.dim == 3 ? geom != 4 && geom != 5 : geom != 2 && geom != 3)
                    throw MeshUnsupported(path + ": element geometry " + std::to_string(geom) +
                                          " is not supported (tetrahedra, hexahedra, triangles or quadrilaterals "
                                          "only)");
                if (geom == 5) {
                    std::array<int32_t, 8> h;
                    for (auto& v : h) v = (int32_t)next_int();
                    m.hexes.push_back(h);
                    m.hex_attr.push_back(attr);
                } else if (geom == 3) {
                    std::array<int32_t, 4> q;
                    for (auto& v : q) v = (int32_t)next_int();
                    m.quads.push_back(q);
                    m.quad_attr.push_back(attr);
                } else if (m.dim == 3) {
                    std::array<int32_t, 4> t;
                    for (auto& v : t) v = (int32_t)next_int();
                    m.tets.push_back(t);
                    m.tet_attr.push_back(attr);
                } else {
                    std::array<int32_t, 3> t;
                    for (auto& v : t) v = (int32_t)next_int();
                    m.tris.push_back(t);
                    m.tri_attr.push_back(attr);
                }
            }
        } else if (sec == "boundary") {
            if (!have_dim) bad(path, "boundary before dimension");
            const long n = next_int();
            for (long i = 0; i < n; ++i) {
                const int attr = (int)next_int();
                const int geom = (int)next_int();
                if (m.dim == 3 && geom == 3) {  // squares: the boundary of hexahedra
                    std::array<int32_t, 4> q;
                    for (auto& v : q) v = (int32_t)next_int();
                    m.quads.push_back(q);
                    m.quad_attr.push_back(attr);
                } else if (m.dim == 3) {
                    if (geom != 2) bad(path, "boundary geometry must be triangles or squares");
                    std::array<int32_t, 3> t;
                    for (auto& v : t) v = (int32_t)next_int();
                    m.tris.push_back(t);
                    m.tri_attr.push_back(attr);
                } else {
                    if (geom != 1) bad(path, "boundary geometry must be segments");
                    std::array<int32_t, 2> t;
                    for (auto& v : t) v = (int32_t)next_int();
                    m.segs.push_back(t);
                    m.seg_attr.push_back(attr);
                }
            }
        } else if (sec == "vertices") {
            const long n = next_int();
            if (p < tok.size() && tok[p] == "nodes")
                throw MeshUnsupported(path + ": curved meshes (nodes) are not supported");
            const long vdim = next_int();
            if (vdim != 3 && vdim != 2) bad(path, "vertex dimension must be 2 or 3");
            m.vx.assign(3 * (size_t)n, 0.0);
            for (long i = 0; i < n; ++i)
                for (long d = 0; d < vdim; ++d) m.vx[3 * i + d] = std::stod(next());
            have_vertices = true;
        } else if (sec == "nodes") {
            throw MeshUnsupported(path + ": curved meshes (nodes) are not supported");
        } else if (sec == "mfem_mesh_end") {
            break;
        } else {
            bad(path, "unknown section '" + sec + "'");
        }
    }
    if (m.dim == 2 && !m.quads.empty() && !m.tris.empty())
        throw MeshUnsupported(path + ": meshes mixing triangles and quadrilaterals are not supported");
    if (m.dim == 3 && !m.hexes.empty() && (!m.tets.empty() || !m.tris.empty()))
        throw MeshUnsupported(path + ": meshes mixing hexahedra with tetrahedra or triangle faces are not supported");
    if (m.dim == 3 && m.hexes.empty() && !m.quads.empty())
        throw MeshUnsupported(path + ": square boundary elements need hexahedra");
    if (!have_vertices || m.K() == 0) bad(path, "missing vertices or elements");
    for (const auto& h : m.hexes)
        for (int v : h)
            if (v < 0 || v >= m.nv()) bad(path, "element vertex out of range");
    const int32_t nv = m.nv();
    for (const auto& q : m.quads)
        for (int v : q)
            if (v < 0 || v >= nv) bad(path, "element vertex out of range");
    for (const auto& t : m.tets)
        for (int v : t)
            if (v < 0 || v >= nv) bad(path, "element vertex out of range");
    for (const auto& t : m.tris)
        for (int v : t)
            if (v < 0 || v >= nv) bad(path, "element or boundary vertex out of range");
    for (const auto& t : m.segs)
        for (int v : t)
            if (v < 0 || v >= nv) bad(path, "boundary vertex out of range");
    return m;
}

void write_mfem_mesh(const TetMesh& m, const std::string& path)
{
    std::ofstream o(path);
    if (!o) throw std::runtime_error("cannot write " + path);
    o.precision(16);
    if (m.dim == 2) {  // triangles (geometry 2) or squares (3), segments (geometry 1), 2D vertices
        o << "MFEM mesh v1.0\n\ndimension\n2\n\nelements\n" << m.K() << "\n";
        for (int32_t e = 0; e < m.K(); ++e) {
            o << m.elem_attr()[e] << (m.quad() ? " 3" : " 2");
            for (int j = 0; j < m.nvert(); ++j) o << " " << m.vertex(e, j);
            o << "\n";
        }
        o << "\nboundary\n" << m.segs.size() << "\n";
        for (size_t i = 0; i < m.segs.size(); ++i) o << m.seg_attr[i] << " 1 " << m.segs[i][0] << " " << m.segs[i][1] << "\n";
        o << "\nvertices\n" << m.nv() << "\n2\n";
        for (int32_t v = 0; v < m.nv(); ++v) o << m.vx[3 * v] << " " << m.vx[3 * v + 1] << "\n";
        if (!o) throw std::runtime_error("write failed: " + path);
        return;
    }
    if (m.hex()) {  // cubes (geometry 5), squares (geometry 3)
        o << "MFEM mesh v1.0\n\ndimension\n3\n\nelements\n" << m.K() << "\n";
        for (int32_t e = 0; e < m.K(); ++e) {
            o << m.hex_attr[e] << " 5";
            for (int j = 0; j < 8; ++j) o << " " << m.hexes[e][j];
            o << "\n";
        }
        o << "\nboundary\n" << m.quads.size() << "\n";
        for (size_t i = 0; i < m.quads.size(); ++i)
            o << m.quad_attr[i] << " 3 " << m.quads[i][0] << " " << m.quads[i][1] << " " << m.quads[i][2] << " "
              << m.quads[i][3] << "\n";
        o << "\nvertices\n" << m.nv() << "\n3\n";
        for (int32_t v = 0; v < m.nv(); ++v)
            o << m.vx[3 * v] << " " << m.vx[3 * v + 1] << " " << m.vx[3 * v + 2] << "\n";
        if (!o) throw std::runtime_error("write failed: " + path);
        return;
    }
    o << "MFEM mesh v1.0\n\ndimension\n3\n\nelements\n" << m.tets.size() << "\n";
    for (size_t e = 0; e < m.tets.size(); ++e)
        o << m.tet_attr[e] << " 4 " << m.tets[e][0] << " " << m.tets[e][1] << " " << m.tets[e][2] << " "
          << m.tets[e][3] << "\n";
    o << "\nboundary\n" << m.tris.size() << "\n";
    for (size_t i = 0; i < m.tris.size(); ++i)
        o << m.tri_attr[i] << " 2 " << m.tris[i][0] << " " << m.tris[i][1] << " " << m.tris[i][2] << "\n";
    o << "\nvertices\n" << m.nv() << "\n3\n";
    for (int32_t v = 0; v < m.nv(); ++v)
        o << m.vx[3 * v] << " " << m.vx[3 * v + 1] << " " << m.vx[3 * v + 2] << "\n";
    if (!o) throw std::runtime_error("write failed: " + path);
}

TetMesh read_mesh_file(const std::string& path)
{
    if (ends_with(path, ".msh")) return read_gmsh22(path);
    if (ends_with(path, ".mesh")) return read_mfem_mesh(path);
    throw std::runtime_error("File format for mesh must be name.msh or name.mesh");
}

void orient_positive(TetMesh& m)
{
    for (auto& t : m.tris) {
        if (m.dim != 2) break;
        const double* a = &m.vx[3 * (size_t)t[0]];
        const double* b = &m.vx[3 * (size_t)t[1]];
        const double* c = &m.vx[3 * (size_t)t[2]];
        const double a2 = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
        if (a2 == 0.0) throw std::runtime_error("degenerate triangle in mesh");
        if (a2 < 0) std::swap(t[0], t[1]);
    }
    for (auto& h : m.hexes) {  // triple product of the edges from vertex 0 along x, y, z
        double a[3][3];
        const int axis[3] = {1, 3, 4};
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) a[k][d] = m.vx[3 * (size_t)h[axis[k]] + d] - m.vx[3 * (size_t)h[0] + d];
        const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        if (det == 0.0) throw std::runtime_error("degenerate hexahedron in mesh");
        if (det < 0) {  // mirror across the plane of vertices 0, 2, 4, 6
            std::swap(h[1], h[3]);
            std::swap(h[5], h[7]);
        }
    }
    for (auto& q : m.quads) {  // shoelace area; a reversed cycle keeps vertex 0
        if (m.dim != 2) break;
        double a2 = 0.0;
        for (int j = 0; j < 4; ++j) {
            const double* a = &m.vx[3 * (size_t)q[j]];
            const double* b = &m.vx[3 * (size_t)q[(j + 1) % 4]];
            a2 += a[0] * b[1] - b[0] * a[1];
        }
        if (a2 == 0.0) throw std::runtime_error("degenerate quadrilateral in mesh");
        if (a2 < 0) std::swap(q[1], q[3]);
    }
    for (auto& t : m.tets) {
        const double v6 = signed_volume6(m.vx.data(), t);
        if (v6 == 0.0) throw std::runtime_error("degenerate tetrahedron in mesh");
        if (v6 < 0) std::swap(t[2], t[3]);
    }
}

TetMesh refine_uniform(const TetMesh& m)
{
    if (m.dim == 2) throw MeshUnsupported("uniform refinement of 2D meshes is not supported");
    TetMesh r;
    r.vx = m.vx;
    const int64_t nv0 = m.nv();
    std::unordered_map<int64_t, int32_t> mid;
    auto midpoint = [&](int32_t a, int32_t b) {
        const int64_t key = (int64_t)std::min(a, b) * nv0 + std::max(a, b);
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        const int32_t id = (int32_t)(r.vx.size() / 3);
        for (int d = 0; d < 3; ++d) r.vx.push_back(0.5 * (m.vx[3 * a + d] + m.vx[3 * b + d]));
        mid.emplace(key, id);
        return id;
    };
    auto dist2 = [&](int32_t a, int32_t b) {
        double s = 0;
        for (int d = 0; d < 3; ++d) {
            const double t = r.vx[3 * a + d] - r.vx[3 * b + d];
            s += t * t;
        }
        return s;
    };
    r.tets.reserve(8 * m.tets.size());
    for (size_t e = 0; e < m.tets.size(); ++e) {
        const auto& t = m.tets[e];
        const int32_t v0 = t[0], v1 = t[1], v2 = t[2], v3 = t[3];
        const int32_t m01 = midpoint(v0, v1), m02 = midpoint(v0, v2), m03 = midpoint(v0, v3);
        const int32_t m12 = midpoint(v1, v2), m13 = midpoint(v1, v3), m23 = midpoint(v2, v3);
        std::array<int32_t, 4> kids[8] = {
            {v0, m01, m02, m03}, {m01, v1, m12, m13}, {m02, m12, v2, m23}, {m03, m13, m23, v3}};
        // inner octahedron: diagonal (a, b) and the cycle of the other four
        const double d0 = dist2(m01, m23), d1 = dist2(m02, m13), d2 = dist2(m03, m12);
        int32_t a, b, c[4];
        if (d0 <= d1 && d0 <= d2) {
            a = m01; b = m23; c[0] = m02; c[1] = m03; c[2] = m13; c[3] = m12;
        } else if (d1 <= d2) {
            a = m02; b = m13; c[0] = m01; c[1] = m03; c[2] = m23; c[3] = m12;
        } else {
            a = m03; b = m12; c[0] = m01; c[1] = m02; c[2] = m23; c[3] = m13;
        }
        for (int i = 0; i < 4; ++i) kids[4 + i] = {a, b, c[i], c[(i + 1) % 4]};
        for (auto& k : kids) {
            if (signed_volume6(r.vx.data(), k) < 0) std::swap(k[2], k[3]);
            r.tets.push_back(k);
            r.tet_attr.push_back(m.tet_attr[e]);
        }
    }
    for (size_t i = 0; i < m.tris.size(); ++i) {
        const auto& t = m.tris[i];
        const int32_t a = t[0], b = t[1], c = t[2];
        const int32_t mab = midpoint(a, b), mbc = midpoint(b, c), mca = midpoint(c, a);
        const std::array<int32_t, 3> kids[4] = {{a, mab, mca}, {mab, b, mbc}, {mca, mbc, c}, {mab, mbc, mca}};
        for (const auto& k : kids) {
            r.tris.push_back(k);
            r.tri_attr.push_back(m.tri_attr[i]);
        }
    }
    // Hexahedra 1:8 on the 3 x 3 x 3 lattice of corners, edge midpoints, face
    // centres and the centre; child (a, b, c) takes lattice points
    // (a + x, b + y, c + z) at MFEM's corner order, so children keep the
    // parent's orientation.  Boundary quadrilaterals split 1:4 on the same
    // edge midpoints and face centres.
    std::map<std::array<int32_t, 4>, int32_t> centre;
    auto face_centre = [&](int32_t a, int32_t b, int32_t c, int32_t d) {
        std::array<int32_t, 4> key{a, b, c, d};
        std::sort(key.begin(), key.end());
        auto it = centre.find(key);
        if (it != centre.end()) return it->second;
        const int32_t id = (int32_t)(r.vx.size() / 3);
        for (int k = 0; k < 3; ++k)
            r.vx.push_back(0.25 * (m.vx[3 * a + k] + m.vx[3 * b + k] + m.vx[3 * c + k] + m.vx[3 * d + k]));
        centre.emplace(key, id);
        return id;
    };
    static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    r.hexes.reserve(8 * m.hexes.size());
    for (size_t e = 0; e < m.hexes.size(); ++e) {
        const auto& h = m.hexes[e];
        auto vert = [&](int i, int j, int k) {  // corner at lattice (2i, 2j, 2k)
            for (int c = 0; c < 8; ++c)
                if (corner[c][0] == i && corner[c][1] == j && corner[c][2] == k) return h[c];
            return h[0];
        };
        int32_t L[3][3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const int mid = (i == 1) + (j == 1) + (k == 1);
                    if (mid == 0) {
                        L[i][j][k] = vert(i / 2, j / 2, k / 2);
                    } else if (mid == 1) {  // edge midpoint
                        const int a0[3] = {i == 1 ? 0 : i / 2, j == 1 ? 0 : j / 2, k == 1 ? 0 : k / 2};
                        const int a1[3] = {i == 1 ? 1 : i / 2, j == 1 ? 1 : j / 2, k == 1 ? 1 : k / 2};
                        L[i][j][k] = midpoint(vert(a0[0], a0[1], a0[2]), vert(a1[0], a1[1], a1[2]));
                    } else if (mid == 2) {  // face centre: the fixed coordinate spans no range
                        int32_t q[4];
                        int n = 0;
                        for (int x = 0; x < 2; ++x)
                            for (int y = 0; y < 2; ++y) {
                                const int ci = i == 1 ? x : i / 2;
                                const int cj = j == 1 ? (i == 1 ? y : x) : j / 2;
                                const int ck = k == 1 ? y : k / 2;
                                q[n++] = vert(ci, cj, ck);
                            }
                        L[i][j][k] = face_centre(q[0], q[1], q[2], q[3]);
                    } else {
                        const int32_t id = (int32_t)(r.vx.size() / 3);
                        for (int d = 0; d < 3; ++d) {
                            double c = 0.0;
                            for (int v = 0; v < 8; ++v) c += m.vx[3 * h[v] + d];
                            r.vx.push_back(0.125 * c);
                        }
                        L[1][1][1] = id;
                    }
                }
        for (int c = 0; c < 8; ++c) {
            const int a = corner[c][0], b = corner[c][1], z = corner[c][2];
            std::array<int32_t, 8> kid;
            for (int v = 0; v < 8; ++v) kid[v] = L[a + corner[v][0]][b + corner[v][1]][z + corner[v][2]];
            r.hexes.push_back(kid);
            r.hex_attr.push_back(m.hex_attr[e]);
        }
    }
    for (size_t i = 0; i < m.quads.size(); ++i) {
        const auto& q = m.quads[i];
        int32_t P[3][3];
        P[0][0] = q[0];
        P[2][0] = q[1];
        P[2][2] = q[2];
        P[0][2] = q[3];
        P[1][0] = midpoint(q[0], q[1]);
        P[2][1] = midpoint(q[1], q[2]);
        P[1][2] = midpoint(q[2], q[3]);
        P[0][1] = midpoint(q[3], q[0]);
        P[1][1] = face_centre(q[0], q[1], q[2], q[3]);
        for (int t = 0; t < 2; ++t)
            for (int u = 0; u < 2; ++u) {
                r.quads.push_back({P[u][t], P[u + 1][t], P[u + 1][t + 1], P[u][t + 1]});
                r.quad_attr.push_back(m.quad_attr[i]);
            }
    }
    return r;
}

namespace {

std::vector<int8_t> edge_codes(const TetMesh& m, const std::vector<int>& attr_to_bc)
{
    const int64_t nv = m.nv();
    const int n = m.nvert();  // edge f joins local vertices f and f+1 mod n
    const int32_t K = m.K();
    std::unordered_map<int64_t, std::pair<int64_t, int64_t>> edges;
    edges.reserve((size_t)n * K);
    for (int32_t e = 0; e < K; ++e)
        for (int f = 0; f < n; ++f) {
            const int64_t a = m.vertex(e, f), b = m.vertex(e, (f + 1) % n);
            const int64_t k = std::min(a, b) * nv + std::max(a, b);
            auto it = edges.find(k);
            if (it == edges.end()) edges.emplace(k, std::make_pair((int64_t)n * e + f, (int64_t)-1));
            else it->second.second = (int64_t)n * e + f;
        }
    std::vector<int8_t> bc((size_t)n * K, 0);
    for (size_t i = 0; i < m.segs.size(); ++i) {
        const int attr = m.seg_attr[i];
        const int code = (attr >= 0 && attr < (int)attr_to_bc.size()) ? attr_to_bc[attr] : -1;
        if (code < 0) continue;
        const int64_t a = m.segs[i][0], b = m.segs[i][1];
        auto it = edges.find(std::min(a, b) * nv + std::max(a, b));
        if (it == edges.end())
            throw std::runtime_error("boundary segment with attribute " + std::to_string(attr) + " is no element edge");
        if (code == 7) {  // TF/SF line: interior edges only
            if (it->second.second < 0)
                throw std::runtime_error("TF/SF attribute " + std::to_string(attr) + " lies on a boundary edge");
            bc[it->second.first] = 7;
            bc[it->second.second] = 7;
        } else if (it->second.second >= 0) {
            bc[it->second.first] = (int8_t)(code + 3);
            bc[it->second.second] = (int8_t)(code + 3);
        } else {
            bc[it->second.first] = (int8_t)code;
        }
    }
    return bc;
}

}  // namespace

namespace {

// Hexahedra: faces of MFEM's hexahedron keyed by their sorted vertices,
// boundary quadrilaterals matched by the same key.
std::vector<int8_t> hex_face_codes(const TetMesh& m, const std::vector<int>& attr_to_bc)
{
    std::map<std::array<int32_t, 4>, std::pair<int64_t, int64_t>> faces;  // key -> (slot, slot or -1)
    auto key_of = [](std::array<int32_t, 4> k) {
        std::sort(k.begin(), k.end());
        return k;
    };
    for (size_t e = 0; e < m.hexes.size(); ++e)
        for (int f = 0; f < 6; ++f) {
            std::array<int32_t, 4> k;
            for (int j = 0; j < 4; ++j) k[j] = m.hexes[e][kHexFaces[f][j]];
            const auto key = key_of(k);
            auto it = faces.find(key);
            if (it == faces.end())
                faces.emplace(key, std::make_pair((int64_t)(6 * e + f), (int64_t)-1));
            else
                it->second.second = (int64_t)(6 * e + f);
        }
    std::vector<int8_t> bc(6 * m.hexes.size(), 0);
    for (size_t i = 0; i < m.quads.size(); ++i) {
        const int attr = m.quad_attr[i];
        const int code = (attr >= 0 && attr < (int)attr_to_bc.size()) ? attr_to_bc[attr] : -1;
        if (code < 0) continue;
        auto it = faces.find(key_of(m.quads[i]));
        if (it == faces.end())
            throw std::runtime_error("boundary quadrilateral with attribute " + std::to_string(attr) +
                                     " is no element face");
        if (code == 7) {
            if (it->second.second < 0)
                throw std::runtime_error("TF/SF attribute " + std::to_string(attr) + " lies on a boundary face");
            bc[it->second.first] = 7;
            bc[it->second.second] = 7;
        } else if (it->second.second >= 0) {
            bc[it->second.first] = (int8_t)(code + 3);
            bc[it->second.second] = (int8_t)(code + 3);
        } else {
            bc[it->second.first] = (int8_t)code;
        }
    }
    return bc;
}

}  // namespace

std::vector<int8_t> face_codes(const TetMesh& m, const std::vector<int>& attr_to_bc)
{
    if (m.dim == 2) return edge_codes(m, attr_to_bc);
    if (m.hex()) return hex_face_codes(m, attr_to_bc);
    static const int kFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    const int64_t nv = m.nv();
    auto key_of = [&](int32_t a, int32_t b, int32_t c) {
        int32_t s[3] = {a, b, c};
        std::sort(s, s + 3);
        return ((int64_t)s[0] * nv + s[1]) * nv + s[2];  // nv^3 < 2^63 for nv < 2e6
    };
    if (nv >= 2000000) throw std::runtime_error("mesh too large for the face table (2e6 vertices)");
    std::unordered_map<int64_t, std::pair<int64_t, int64_t>> faces;  // key -> (slot, slot or -1)
    faces.reserve(4 * m.tets.size());
    for (size_t e = 0; e < m.tets.size(); ++e)
        for (int f = 0; f < 4; ++f) {
            const auto& t = m.tets[e];
            const int64_t k = key_of(t[kFace[f][0]], t[kFace[f][1]], t[kFace[f][2]]);
            auto it = faces.find(k);
            if (it == faces.end())
                faces.emplace(k, std::make_pair((int64_t)(4 * e + f), (int64_t)-1));
            else
                it->second.second = (int64_t)(4 * e + f);
        }
    std::vector<int8_t> bc(4 * m.tets.size(), 0);
    for (size_t i = 0; i < m.tris.size(); ++i) {
        const int attr = m.tri_attr[i];
        const int code = (attr >= 0 && attr < (int)attr_to_bc.size()) ? attr_to_bc[attr] : -1;
        if (code < 0) continue;
        const auto& t = m.tris[i];
        auto it = faces.find(key_of(t[0], t[1], t[2]));
        if (it == faces.end())
            throw std::runtime_error("boundary triangle with attribute " + std::to_string(attr) + " is no element face");
        if (code == 7) {  // TF/SF surface: interior faces only
            if (it->second.second < 0)
                throw std::runtime_error("TF/SF attribute " + std::to_string(attr) + " lies on a boundary face");
            bc[it->second.first] = 7;
            bc[it->second.second] = 7;
        } else if (it->second.second >= 0) {  // interior face: the interior condition on both sides
            bc[it->second.first] = (int8_t)(code + 3);
            bc[it->second.second] = (int8_t)(code + 3);
        } else {
            bc[it->second.first] = (int8_t)code;
        }
    }
    return bc;
}

}  // namespace dgtd
