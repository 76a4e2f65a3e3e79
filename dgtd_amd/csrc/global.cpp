// The Global operator as an assembled sparse matrix (GlobalEvolution::Mult,
// src/evolution/GlobalEvolution.cpp:628-1088: out = A [E; H] with A built
// from DGOperatorFactory's sub-operators, src/components/DGOperatorFactory.h).
//
// For straight-sided elements the weak-form DG operator the reference
// assembles with MFEM integrators and the strong-form hesthaven operator are
// the same semi-discrete operator (the reference's 2D tests check the Global
// sub-operators against Hesthaven's matrices, Hesthaven2DTest.cpp:234-553),
// so A is assembled here from the same element data as the matrix-free
// kernels, plus the bulk conductivity the Global operator adds to the E
// equations (-sigma/eps E, collectGlobalConductiveOperator,
// DGOperatorFactory.h:1554).  Rows follow the state layout (row c*N + e*Np + i),
// so a stage kernel can apply its update to row r of the product directly.
//
// Per element: rhs = volume curl (metric-weighted reference derivatives) +
// LIFT fscale/2 C(n) d over the face nodes, with d the jump (exterior -
// interior, or the boundary / interior-condition weights of the own trace)
// and C(n) = [[a(I - n n^T), [n x]], [-[n x], a(I - n n^T)]]; E rows scaled
// by 1/eps, H rows by 1/mu.  One difference from the hesthaven operator: on
// SMA boundary faces a = 1 whatever alpha, as the Global operator assembles
// them (with alpha = 1 the two agree).
#include "global.hpp"

#include "kernels.hpp"
#include "hex.hpp"
#include "quad.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <thread>

namespace dgtd {

namespace {

// BC weights of the own trace (same table as the device kernels).
void bc_weights(int bc, double alpha, double& ce, double& ch)
{
    const double sma = alpha != 0.0 ? -1.0 / alpha : 0.0;
    ce = bc == 1 ? -2.0 : bc == 3 ? sma : bc == 4 ? -1.0 : bc == 6 ? -0.5 : 0.0;
    ch = bc == 2 ? -2.0 : bc == 3 ? sma : bc == 5 ? -1.0 : bc == 6 ? -0.5 : 0.0;
}

// Element view shared by tetrahedra and triangles.
struct ElemView {
    int Np, Nfp, nf, nd;                  // nodes, face nodes, faces, reference directions
    const double* D[3];                   // Dr, Ds, Dt (row-major Np x Np)
    const double* lift;                   // Np x nf*Nfp
    const int* fmask;                     // nf*Nfp
    const int* perm;                      // neighbour-node table
    // per element e: metric G[k][d] = d r_k / d x_d, normals, fscale/2, 1/eps, 1/mu
    double G[3][3], n[6][3], fs[6], ieps, imu;
    int nb[6], code[6];
};

void load_tet(ElemView& v, const MeshData& md, int e)
{
    const double* g = &md.geo[(size_t)kGeoStride * e];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) v.G[k][d] = g[3 * k + d];
    for (int f = 0; f < 4; ++f) {
        for (int d = 0; d < 3; ++d) v.n[f][d] = g[9 + 4 * f + d];
        v.fs[f] = g[12 + 4 * f];
        v.nb[f] = md.conn[(size_t)kConnStride * e + f];
        v.code[f] = md.conn[(size_t)kConnStride * e + 4 + f];
    }
    v.ieps = g[25];
    v.imu = g[26];
}

void load_hex(ElemView& v, const HexMeshData& md, int e)
{
    const double* g = &md.geo[(size_t)kHexGeoStride * e];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) v.G[k][d] = g[3 * k + d];
    for (int f = 0; f < 6; ++f) {
        for (int d = 0; d < 3; ++d) v.n[f][d] = g[9 + 4 * f + d];
        v.fs[f] = g[12 + 4 * f];
        v.nb[f] = md.conn[(size_t)kHexConnStride * e + f];
        v.code[f] = md.conn[(size_t)kHexConnStride * e + 6 + f];
    }
    v.ieps = g[33];
    v.imu = g[34];
}

// triangles (nf = 3, 15-value records) and quadrilaterals (nf = 4, 18 values)
void load_2d(ElemView& v, const TriMeshData& md, int e, int nf)
{
    const int stride = nf == 4 ? kQuadGeoStride : kTriGeoStride, cstride = nf == 4 ? kQuadConnStride : kTriConnStride;
    const double* g = &md.geo[(size_t)stride * e];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) v.G[k][d] = 0.0;
    v.G[0][0] = g[0];
    v.G[0][1] = g[1];
    v.G[1][0] = g[2];
    v.G[1][1] = g[3];
    for (int f = 0; f < nf; ++f) {
        v.n[f][0] = g[4 + 3 * f];
        v.n[f][1] = g[5 + 3 * f];
        v.n[f][2] = 0.0;
        v.fs[f] = g[6 + 3 * f];
        v.nb[f] = md.conn[(size_t)cstride * e + f];
        v.code[f] = md.conn[(size_t)cstride * e + nf + f];
    }
    v.ieps = g[4 + 3 * nf];
    v.imu = g[5 + 3 * nf];
}

// One element's block of the matrix: rows (c, i) of the element, columns
// q = (slot * 6 + k) * Np + j with slot 0 = own element, 1 + f = neighbour
// across face f.  Dense storage with a list of the touched entries per row,
// so that filling, reading and clearing cost what the nonzeros cost.
struct BlockAcc {
    int rows = 0, ncol = 0;
    std::vector<double> B;
    std::vector<uint8_t> used;
    std::vector<std::vector<int>> touched;
    void init(int r, int c)
    {
        if (r == rows && c == ncol) return;
        rows = r;
        ncol = c;
        B.assign((size_t)r * c, 0.0);
        used.assign((size_t)r * c, 0);
        touched.assign(r, {});
    }
    void add(int row, int q, double w)
    {
        const size_t x = (size_t)row * ncol + q;
        if (!used[x]) {
            used[x] = 1;
            touched[row].push_back(q);
        }
        B[x] += w;
    }
    // the row's nonzero entries in column order; leaves the row cleared
    template <typename F>
    void drain(int row, F&& emit)
    {
        std::vector<int>& t = touched[row];
        std::sort(t.begin(), t.end());
        for (int q : t) {
            const size_t x = (size_t)row * ncol + q;
            if (B[x] != 0.0) emit(q, B[x]);
            B[x] = 0.0;
            used[x] = 0;
        }
        t.clear();
    }
};

void element_block(const ElemView& v, double alpha, double sigma, BlockAcc& acc)
{
    const int Np = v.Np, Nfp = v.Nfp, nf = v.nf;
    acc.init(6 * Np, (1 + nf) * 6 * Np);
    auto add = [&](int c, int i, int slot, int k, int j, double w) {
        acc.add(c * Np + i, (slot * 6 + k) * Np + j, w);
    };
    // volume: D_d = sum_k G[k][d] D_k
    std::vector<double> Dd[3];
    for (int d = 0; d < 3; ++d) {
        Dd[d].assign((size_t)Np * Np, 0.0);
        for (int k = 0; k < v.nd; ++k)
            for (size_t q = 0; q < (size_t)Np * Np; ++q) Dd[d][q] += v.G[k][d] * v.D[k][q];
    }
    for (int x = 0; x < 3; ++x) {
        const int y = (x + 1) % 3, z = (x + 2) % 3;
        for (int i = 0; i < Np; ++i)
            for (int j = 0; j < Np; ++j) {
                // E_x: (D_y H_z - D_z H_y)/eps;  H_x: (-D_y E_z + D_z E_y)/mu
                const double dy = Dd[y][(size_t)i * Np + j], dz = Dd[z][(size_t)i * Np + j];
                if (dy != 0.0) {
                    add(x, i, 0, 3 + z, j, dy * v.ieps);
                    add(3 + x, i, 0, z, j, -(dy * v.imu));
                }
                if (dz != 0.0) {
                    add(x, i, 0, 3 + y, j, -(dz * v.ieps));
                    add(3 + x, i, 0, y, j, dz * v.imu);
                }
            }
    }
    // surface
    for (int f = 0; f < nf; ++f) {
        const double* n = v.n[f];
        const bool interior = v.nb[f] >= 0;
        // The Global operator weights the upwind terms of SMA faces by 1
        // whatever alpha: boundary SMA (DGOperatorFactory.h:485-495, 521-531,
        // 557-567) and interior SMA (coeff 1.0 against c*alpha for PEC/PMC,
        // :587-591, 655-659), i.e. the face behaves as the hesthaven face with
        // alpha = 1.
        const int bcode = interior ? 0 : -v.nb[f] - 1;
        const double af = bcode == 3 || bcode == 6 ? 1.0 : alpha;
        double C[6][6] = {};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const double proj = af * ((a == b ? 1.0 : 0.0) - n[a] * n[b]);
                C[a][b] = proj;
                C[3 + a][3 + b] = proj;
            }
        for (int a = 0; a < 3; ++a) {  // [n x] v = n x v
            const int y = (a + 1) % 3, z = (a + 2) % 3;
            C[a][3 + z] += n[y];
            C[a][3 + y] -= n[z];
            C[3 + a][z] -= n[y];
            C[3 + a][y] += n[z];
        }
        double ce = 0.0, ch = 0.0;
        if (!interior) bc_weights(-v.nb[f] - 1, af, ce, ch);
        for (int b = 0; b < Nfp; ++b) {
            const int m = f * Nfp + b;
            const int M = v.fmask[m];
            const int P = interior ? v.perm[(size_t)v.code[f] * Nfp + b] : -1;
            for (int c = 0; c < 6; ++c) {
                const double sc = (c < 3 ? v.ieps : v.imu) * v.fs[f];
                for (int i = 0; i < Np; ++i) {
                    const double l = v.lift[(size_t)i * nf * Nfp + m] * sc;
                    if (l == 0.0) continue;
                    for (int k = 0; k < 6; ++k) {
                        const double w = l * C[c][k];
                        if (w == 0.0) continue;
                        if (interior) {
                            add(c, i, 1 + f, k, P, w);
                            add(c, i, 0, k, M, -w);
                        } else {
                            add(c, i, 0, k, M, w * (k < 3 ? ce : ch));
                        }
                    }
                }
            }
        }
    }
    // bulk conductivity on the E rows: -sigma/eps E
    if (sigma != 0.0)
        for (int c = 0; c < 3; ++c)
            for (int i = 0; i < Np; ++i) add(c, i, 0, c, i, -(sigma * v.ieps));
}

// Elements in contiguous ranges over host threads (each element writes its
// own rows only): a count pass sizes the rows, a second pass fills them.
template <typename View>
void assemble(int K, int Np, int nf, View&& view_of, const double* sigma, double alpha, Csr& out)
{
    if (sigma)  // Material.cpp:14 refuses negative conductivity
        for (int e = 0; e < K; ++e)
            if (!(sigma[e] >= 0.0) || !std::isfinite(sigma[e]))
                throw std::runtime_error("conductivity must be finite and >= 0");
    const size_t N = (size_t)Np * K;
    std::vector<int64_t> count(6 * N + 1, 0);
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nth = std::max(1, std::min({hw, 16, K / 64}));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            out.rowptr.assign(6 * N + 1, 0);
            for (size_t r = 0; r < 6 * N; ++r) out.rowptr[r + 1] = out.rowptr[r] + count[r];
            out.col.assign((size_t)out.rowptr[6 * N], 0);
            out.val.assign((size_t)out.rowptr[6 * N], 0.0);
            std::fill(count.begin(), count.end(), 0);
        }
        auto work = [&](int e0, int e1) {
            BlockAcc acc;
            for (int e = e0; e < e1; ++e) {
                ElemView v = view_of(e);
                element_block(v, alpha, sigma ? sigma[e] : 0.0, acc);
                for (int c = 0; c < 6; ++c)
                    for (int i = 0; i < Np; ++i) {
                        const size_t r = c * N + (size_t)e * Np + i;
                        acc.drain(c * Np + i, [&](int q, double val) {
                            if (pass == 1) {
                                const int slot = q / (6 * Np), k = (q / Np) % 6, j = q % Np;
                                const int elem = slot == 0 ? e : v.nb[slot - 1];
                                const size_t at = (size_t)out.rowptr[r] + count[r];
                                out.col[at] = (int32_t)(k * N + (size_t)elem * Np + j);
                                out.val[at] = val;
                            }
                            ++count[r];
                        });
                    }
            }
        };
        if (nth == 1) {
            work(0, K);
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < nth; ++t)
                pool.emplace_back(work, (int)((int64_t)K * t / nth), (int)((int64_t)K * (t + 1) / nth));
            for (std::thread& th : pool) th.join();
        }
    }
}

}  // namespace

CsrLocal16 csr_local16(const Csr& A, int64_t N, int Np)
{
    static_assert(kCsrTab <= (1 << (16 - kCsrSlotShift)), "slots fit the code");
    if (Np >= (1 << kCsrCompShift)) throw std::runtime_error("element-local columns need Np < 256");
    const int64_t nb = (N + kCsrRowsPerBlock - 1) / kCsrRowsPerBlock;
    CsrLocal16 out;
    out.col.resize(A.col.size());
    out.etab.assign((size_t)nb * kCsrTab, 0);
    for (int64_t b = 0; b < nb; ++b) {
        int32_t* tab = &out.etab[(size_t)b * kCsrTab];
        int used = 0;
        const int64_t d1 = std::min<int64_t>(N, (b + 1) * kCsrRowsPerBlock);
        for (int c = 0; c < 6; ++c)
            for (int64_t dof = b * kCsrRowsPerBlock; dof < d1; ++dof) {
                const int64_t r = c * N + dof;
                for (int64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
                    const int64_t col = A.col[k];
                    const int comp = (int)(col / N);
                    const int64_t rem = col % N;
                    const int32_t elem = (int32_t)(rem / Np);
                    const int j = (int)(rem % Np);
                    int slot = 0;
                    while (slot < used && tab[slot] != elem) ++slot;
                    if (slot == used) {
                        if (used == kCsrTab) throw std::runtime_error("a DOF block references too many elements");
                        tab[used++] = elem;
                    }
                    out.col[k] = (uint16_t)(slot << kCsrSlotShift | comp << kCsrCompShift | j);
                }
            }
    }
    return out;
}

Csr assemble_global_tet(const MeshData& md, const RefTet& R, const double* sigma, double alpha)
{
    if (6.0 * R.Np * (double)md.K > 2147483647.0) throw std::runtime_error("mesh too large for 32-bit column indices");
    Csr out;
    auto view = [&](int e) {
        ElemView v{};
        v.Np = R.Np;
        v.Nfp = R.Nfp;
        v.nf = 4;
        v.nd = 3;
        v.D[0] = R.Dr.data();
        v.D[1] = R.Ds.data();
        v.D[2] = R.Dt.data();
        v.lift = R.LIFT.data();
        v.fmask = R.Fmask.data();
        v.perm = R.face_perm.data();
        load_tet(v, md, e);
        return v;
    };
    assemble(md.K, R.Np, 4, view, sigma, alpha, out);
    return out;
}

Csr assemble_global_tri(const TriMeshData& md, const RefTri& R, const double* sigma, double alpha)
{
    if (6.0 * R.Np * (double)md.K > 2147483647.0) throw std::runtime_error("mesh too large for 32-bit column indices");
    Csr out;
    auto view = [&](int e) {
        ElemView v{};
        v.Np = R.Np;
        v.Nfp = R.Nfp;
        v.nf = 3;
        v.nd = 2;
        v.D[0] = R.Dr.data();
        v.D[1] = R.Ds.data();
        v.D[2] = R.Ds.data();  // unused (nd = 2)
        v.lift = R.LIFT.data();
        v.fmask = R.Fmask.data();
        v.perm = R.edge_perm.data();
        load_2d(v, md, e, 3);
        return v;
    };
    assemble(md.K, R.Np, 3, view, sigma, alpha, out);
    return out;
}

}  // namespace dgtd

namespace dgtd {

Csr assemble_global_quad(const TriMeshData& md, const RefQuad& R, const double* sigma, double alpha)
{
    if (6.0 * R.Np * (double)md.K > 2147483647.0) throw std::runtime_error("mesh too large for 32-bit column indices");
    Csr out;
    auto view = [&](int e) {
        ElemView v{};
        v.Np = R.Np;
        v.Nfp = R.Nfp;
        v.nf = 4;
        v.nd = 2;
        v.D[0] = R.Dr.data();
        v.D[1] = R.Ds.data();
        v.D[2] = R.Ds.data();  // unused (nd = 2)
        v.lift = R.LIFT.data();
        v.fmask = R.Fmask.data();
        v.perm = R.edge_perm.data();
        load_2d(v, md, e, 4);
        return v;
    };
    assemble(md.K, R.Np, 4, view, sigma, alpha, out);
    return out;
}

}  // namespace dgtd

namespace dgtd {

Csr assemble_global_hex(const HexMeshData& md, const RefHex& R, const double* sigma, double alpha)
{
    if (6.0 * R.Np * (double)md.K > 2147483647.0) throw std::runtime_error("mesh too large for 32-bit column indices");
    Csr out;
    auto view = [&](int e) {
        ElemView v{};
        v.Np = R.Np;
        v.Nfp = R.Nfp;
        v.nf = 6;
        v.nd = 3;
        v.D[0] = R.Dr.data();
        v.D[1] = R.Ds.data();
        v.D[2] = R.Dt.data();
        v.lift = R.LIFT.data();
        v.fmask = R.Fmask.data();
        v.perm = R.face_perm.data();
        load_hex(v, md, e);
        return v;
    };
    assemble(md.K, R.Np, 6, view, sigma, alpha, out);
    return out;
}

}  // namespace dgtd
