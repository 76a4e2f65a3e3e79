// Run-time dispatch of the matrix-core stage over its tuning configurations
// (device code in mfma_stage.hpp, one translation unit per configuration).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dgtd {

template <typename T>
MfmaInfo mfma_info(int order, int dim, int cfg)
{
    return cfg == 1 ? mfma_info_cfg<1, T>(order, dim) : mfma_info_cfg<0, T>(order, dim);
}

template <typename T>
bool mfma_available(int order, int cfg)
{
    return mfma_info<T>(order, 3, cfg).ok;
}

template <typename T>
hipError_t launch_stage_mfma(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s, int cfg)
{
    return cfg == 1 ? mfma_launch_cfg<1, T>(order, 3, a, mode, grid, s) : mfma_launch_cfg<0, T>(order, 3, a, mode, grid, s);
}

template <typename T>
hipError_t launch_tri_stage_mfma(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s, int cfg)
{
    return cfg == 1 ? mfma_launch_cfg<1, T>(order, 2, a, mode, grid, s) : mfma_launch_cfg<0, T>(order, 2, a, mode, grid, s);
}

template <typename T>
hipError_t launch_batch_stage_mfma(int order, int dim, const BatchArgs<T>& b, const int* range, int grid,
                                   hipStream_t s, int cfg)
{
    return cfg == 1 ? batch_mfma_launch_cfg<1, T>(order, dim, b, range, grid, s)
                    : batch_mfma_launch_cfg<0, T>(order, dim, b, range, grid, s);
}

template <typename T>
int mfma_blocks_per_cu(int order, int cfg)
{
    return cfg == 1 ? mfma_prepare_cfg<1, T>(order, 3) : mfma_prepare_cfg<0, T>(order, 3);
}

template <typename T>
int tri_mfma_blocks_per_cu(int order, int cfg)
{
    return cfg == 1 ? mfma_prepare_cfg<1, T>(order, 2) : mfma_prepare_cfg<0, T>(order, 2);
}

template <typename T>
int mfma_waves_per_block(int order, int cfg)
{
    return mfma_info<T>(order, 3, cfg).waves;
}

template <typename T>
int tri_mfma_waves_per_block(int order, int cfg)
{
    return mfma_info<T>(order, 2, cfg).waves;
}

#define DGTD_DISPATCH(T)                                                                                   \
    template MfmaInfo mfma_info<T>(int, int, int);                                                        \
    template bool mfma_available<T>(int, int);                                                            \
    template hipError_t launch_stage_mfma<T>(int, const StageArgs<T>&, int, int, hipStream_t, int);       \
    template hipError_t launch_tri_stage_mfma<T>(int, const StageArgs<T>&, int, int, hipStream_t, int);   \
    template hipError_t launch_batch_stage_mfma<T>(int, int, const BatchArgs<T>&, const int*, int,         \
                                                   hipStream_t, int);                                     \
    template int mfma_blocks_per_cu<T>(int, int);                                                         \
    template int tri_mfma_blocks_per_cu<T>(int, int);                                                     \
    template int mfma_waves_per_block<T>(int, int);                                                       \
    template int tri_mfma_waves_per_block<T>(int, int);
DGTD_DISPATCH(float)
DGTD_DISPATCH(double)

}  // namespace dgtd
