// Host set-up of the TF/SF plane-wave injection (tfsf_host.hpp).
#include "tfsf_host.hpp"

#include <algorithm>
#include <cmath>

#include "mesh.hpp"

namespace dgtd {

bool planewaves_device(const dgtd_planewave* w, int n, std::vector<TfsfWave>& out, std::string& err)
{
    out.assign((size_t)std::max(0, n), TfsfWave{});
    for (int i = 0; i < n; ++i) {
        const dgtd_planewave& p = w[i];
        if (p.kind != DGTD_PLANEWAVE_GAUSSIAN && p.kind != DGTD_PLANEWAVE_MODULATED) {
            err = "unknown plane-wave kind";
            return false;
        }
        const double np = std::sqrt(p.polarization[0] * p.polarization[0] + p.polarization[1] * p.polarization[1] +
                                    p.polarization[2] * p.polarization[2]);
        const double nd = std::sqrt(p.propagation[0] * p.propagation[0] + p.propagation[1] * p.propagation[1] +
                                    p.propagation[2] * p.propagation[2]);
        if (!(np > 0) || !(nd > 0) || !(p.spread > 0) || !std::isfinite(p.mean) || !std::isfinite(p.frequency)) {
            err = "plane wave needs nonzero polarization and propagation and spread > 0";
            return false;
        }
        TfsfWave& d = out[i];
        for (int c = 0; c < 3; ++c) {
            d.pol[c] = p.polarization[c] / np;
            d.dir[c] = p.propagation[c] / nd;
        }
        for (int c = 0; c < 3; ++c) {
            const int y = (c + 1) % 3, z = (c + 2) % 3;
            d.kxp[c] = d.dir[y] * d.pol[z] - d.dir[z] * d.pol[y];
        }
        d.spread = p.spread;
        d.mean = p.mean;
        d.freq = p.frequency;
        d.kind = p.kind;
    }
    return true;
}

namespace {

// Records of the elements owning TF/SF faces for an element of NV vertices
// (4 tetrahedron, 3 triangle); face node n sits at x = v0 + sum_k (v_{k+1} -
// v0) xi_k with the reference coordinates xi[(NV-1) n + k].
// (A parallelogram, NV = 4 with quad set, maps its unit square through
// vertices 1 and 3: x = v0 + (v1 - v0) xi_0 + (v3 - v0) xi_1.)
template <int NV>
TfsfRecords records(const dgtd_mesh_desc* mesh, const std::vector<int>& Fmask, const std::vector<double>& xi,
                    int nfaces, int Nfp, int stride, const std::vector<std::array<int, 3>>& tfsf,
                    const std::vector<double>& geo, bool quad = false)
{
    TfsfRecords r;
    std::vector<int> slot(mesh->K, -1);
    for (const auto& t : tfsf) {
        const int e = t[0];
        if (slot[e] < 0) {
            slot[e] = (int)r.elems.size();
            r.elems.push_back(e);
            r.signs.insert(r.signs.end(), kTfsfSignStride, 0);
            r.geo.insert(r.geo.end(), geo.begin() + (size_t)stride * e, geo.begin() + (size_t)stride * (e + 1));
            const int32_t* v = &mesh->EToV[(size_t)NV * e];
            for (int m = 0; m < nfaces * Nfp; ++m) {
                const int n = Fmask[m];
                for (int d = 0; d < 3; ++d) {
                    if ((NV < 4 || quad) && d == 2) {  // triangles and quadrilaterals lie in z = 0
                        r.xyz.push_back(0.0);
                        continue;
                    }
                    const double x0 = mesh->vx[3 * (size_t)v[0] + d];
                    double x = x0;
                    if (quad) {
                        x += (mesh->vx[3 * (size_t)v[1] + d] - x0) * xi[2 * (size_t)n] +
                             (mesh->vx[3 * (size_t)v[3] + d] - x0) * xi[2 * (size_t)n + 1];
                    } else if (NV == 8) {  // hexahedra: unit-cube axes along vertices 1, 3, 4
                        x += (mesh->vx[3 * (size_t)v[1] + d] - x0) * xi[3 * (size_t)n] +
                             (mesh->vx[3 * (size_t)v[3] + d] - x0) * xi[3 * (size_t)n + 1] +
                             (mesh->vx[3 * (size_t)v[4] + d] - x0) * xi[3 * (size_t)n + 2];
                    } else {
                        for (int k = 0; k < NV - 1; ++k)
                            x += (mesh->vx[3 * (size_t)v[k + 1] + d] - x0) * xi[(size_t)(NV - 1) * n + k];
                    }
                    r.xyz.push_back(x);
                }
            }
        }
        r.signs[kTfsfSignStride * (size_t)slot[e] + t[1]] = t[2];
    }
    return r;
}

}  // namespace

TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefTet& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo)
{
    return records<4>(mesh, R.Fmask, R.xi, 4, R.Nfp, kGeoStride, tfsf, geo);
}

TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefTri& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo)
{
    return records<3>(mesh, R.Fmask, R.xi, 3, R.Nfp, kTriGeoStride, tfsf, geo);
}

TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefQuad& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo)
{
    return records<4>(mesh, R.Fmask, R.xi, 4, R.Nfp, kQuadGeoStride, tfsf, geo, true);
}

}  // namespace dgtd

namespace dgtd {

TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefHex& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo)
{
    return records<8>(mesh, R.Fmask, R.xi, 6, R.Nfp, kHexGeoStride, tfsf, geo);
}

}  // namespace dgtd
