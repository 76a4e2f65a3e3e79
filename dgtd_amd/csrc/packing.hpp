// Host-side packing of the reference matrices into the layouts the device
// kernels read.
#pragma once

#include <vector>

#include "reftet.hpp"
#include "tri.hpp"

namespace dgtd {

// Per-node (VALU) layout: Dr^T, Ds^T, Dt^T ([j][i] each), then LIFT^T ([m][i]).
template <typename T>
inline std::vector<T> pack_mats(const RefTet& R)
{
    const int Np = R.Np, NF = 4 * R.Nfp;
    std::vector<T> m((size_t)3 * Np * Np + (size_t)NF * Np);
    const std::vector<double>* D[3] = {&R.Dr, &R.Ds, &R.Dt};
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < Np; ++i)
            for (int j = 0; j < Np; ++j) m[(size_t)d * Np * Np + (size_t)j * Np + i] = (T)(*D[d])[(size_t)i * Np + j];
    T* L = m.data() + 3 * (size_t)Np * Np;
    for (int i = 0; i < Np; ++i)
        for (int k = 0; k < NF; ++k) L[(size_t)k * Np + i] = (T)R.LIFT[(size_t)i * NF + k];
    return m;
}

// B operands of the MFMA variant: Bv[j][kd*NPG + i] = D_kd[i][j] (KP rows),
// then Bl[f*NFPP + b][i] = LIFT[i][f*Nfp + b] (4*NFPP rows, faces padded to a
// multiple of 4 nodes); zero outside i, j < Np and b < Nfp.
template <typename T>
inline std::vector<T> pack_mats_mfma(const RefTet& R)
{
    const int Np = R.Np, Nfp = R.Nfp, NF = 4 * Nfp;
    const int KP = (Np + 3) / 4 * 4, NPG = (Np + 15) / 16 * 16, NFPP = (Nfp + 3) / 4 * 4;
    std::vector<T> m((size_t)KP * 3 * NPG + (size_t)4 * NFPP * NPG, T(0));
    const std::vector<double>* D[3] = {&R.Dr, &R.Ds, &R.Dt};
    for (int j = 0; j < Np; ++j)
        for (int kd = 0; kd < 3; ++kd)
            for (int i = 0; i < Np; ++i) m[(size_t)j * 3 * NPG + kd * NPG + i] = (T)(*D[kd])[(size_t)i * Np + j];
    T* L = m.data() + (size_t)KP * 3 * NPG;
    for (int f = 0; f < 4; ++f)
        for (int b = 0; b < Nfp; ++b)
            for (int i = 0; i < Np; ++i)
                L[(size_t)(f * NFPP + b) * NPG + i] = (T)R.LIFT[(size_t)i * NF + f * Nfp + b];
    return m;
}

// Matrix-core layout on triangles: Bv [KP][2*NPG] (Dr, Ds transposed per input
// node j), then Bl [3*NFPP][NPG] (LIFT columns, each edge padded to whole k-steps).
template <typename T>
inline std::vector<T> pack_mats_mfma(const RefTri& R)
{
    const int Np = R.Np, Nfp = R.Nfp, NF = 3 * Nfp;
    const int KP = (Np + 3) / 4 * 4, NPG = (Np + 15) / 16 * 16, NFPP = (Nfp + 3) / 4 * 4;
    std::vector<T> m((size_t)KP * 2 * NPG + (size_t)3 * NFPP * NPG, T(0));
    const std::vector<double>* D[2] = {&R.Dr, &R.Ds};
    for (int j = 0; j < Np; ++j)
        for (int kd = 0; kd < 2; ++kd)
            for (int i = 0; i < Np; ++i) m[(size_t)j * 2 * NPG + kd * NPG + i] = (T)(*D[kd])[(size_t)i * Np + j];
    T* L = m.data() + (size_t)KP * 2 * NPG;
    for (int f = 0; f < 3; ++f)
        for (int b = 0; b < Nfp; ++b)
            for (int i = 0; i < Np; ++i)
                L[(size_t)(f * NFPP + b) * NPG + i] = (T)R.LIFT[(size_t)i * NF + f * Nfp + b];
    return m;
}

template <typename T>
inline std::vector<T> cast_vec(const std::vector<double>& v)
{
    return std::vector<T>(v.begin(), v.end());
}

}  // namespace dgtd
