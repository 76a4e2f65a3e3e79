// Host set-up of the TF/SF plane-wave injection, shared by the single
// operator (dgtd_api.cpp) and the batch (batch.cpp).
#pragma once

#include <array>
#include <string>
#include <vector>

#include "dgtd_hip.h"
#include "kernels.hpp"
#include "hex.hpp"
#include "reftet.hpp"
#include "quad.hpp"
#include "tri.hpp"

namespace dgtd {

// The plane waves in device form: polarisation and propagation normalised as
// Planewave does (Function.h:331-340), H = k x pol (:365-374).  False, with a
// message, on an unknown kind, a zero direction or spread <= 0.
bool planewaves_device(const dgtd_planewave* w, int n, std::vector<TfsfWave>& out, std::string& err);

// The elements owning TF/SF faces, in TfsfArgs layout: element ids, per face
// (kTfsfSignStride slots per element) -1 (scattered side), +1 (total side)
// or 0, the face-node positions
// [nfaces*Nfp][3] and the element's FP64 geometry record.
struct TfsfRecords {
    std::vector<int> elems, signs;
    std::vector<double> xyz, geo;
};
// tfsf: (element, face, side) triples from build_mesh / build_tri_mesh;
// geo: their FP64 records of every element.
TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefTet& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo);
TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefTri& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo);
TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefQuad& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo);
TfsfRecords tfsf_records(const dgtd_mesh_desc* mesh, const RefHex& R, const std::vector<std::array<int, 3>>& tfsf,
                         const std::vector<double>& geo);

}  // namespace dgtd
