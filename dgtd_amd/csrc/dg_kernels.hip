// Fused DGTD Maxwell stage kernels for gfx950 (the "hesthaven" operator of
// OpenSEMBA/dgtd, HesthavenEvolution.cpp:450-542, with the RK stage update).
//
// One launch = one Runge-Kutta stage over all elements:
//   1. stage the block's fields (6 x Np per element), geometry and
//      connectivity records in LDS (coalesced: component-major state, one
//      contiguous chunk per component per element block);
//   2. per face node: gather the neighbour trace through (neighbour id,
//      permutation code), apply PEC/PMC/SMA, form the upwind/centred flux and
//      scale it by fscale/2 -- into LDS, never to HBM;
//   3. per volume node: reference derivatives [Dr;Ds;Dt] of the 6 fields and
//      LIFT * flux from LDS-resident reference matrices, metric combine into
//      the curl, 1/eps and 1/mu, then the stage update, writing each state
//      buffer once.
// Blocks are persistent (grid = resident capacity) and loop over element
// blocks, so the reference matrices are staged once per block.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "stage_body.hpp"

namespace dgtd {

namespace {

using namespace valu;

template <int P, typename T, int MODE>
__global__ __launch_bounds__(256) void dg_stage_kernel(StageArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    stage_load_ref<P, T>(a.mats, a.tabs, smem_raw);
    __syncthreads();
    for (int blk = blockIdx.x; blk < a.nblk; blk += gridDim.x) stage_block<P, T, MODE>(a, smem_raw, blk);
}

template <int P, typename T, int MODE>
hipError_t launch_one(const StageArgs<T>& a, int grid, hipStream_t s)
{
    constexpr size_t bytes = Smem<P, T>::bytes;
    auto kern = dg_stage_kernel<P, T, MODE>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), bytes, s, a);
    return hipGetLastError();
}

template <int P, typename T>
hipError_t launch_mode(const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (mode) {
        case kModeMult: return launch_one<P, T, kModeMult>(a, grid, s);
        case kModeLserk: return launch_one<P, T, kModeLserk>(a, grid, s);
        case kModeRk1: return launch_one<P, T, kModeRk1>(a, grid, s);
        case kModeRk2: return launch_one<P, T, kModeRk2>(a, grid, s);
        case kModeRk3: return launch_one<P, T, kModeRk3>(a, grid, s);
        case kModeRk4: return launch_one<P, T, kModeRk4>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <int P, typename T, int MODE>
void set_smem_attr()
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_stage_kernel<P, T, MODE>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)Smem<P, T>::bytes);
}

// Raises the dynamic-LDS limit of every stage variant of this order (needed
// above 64 KiB) and returns resident blocks per CU.
template <int P, typename T>
int occupancy_one()
{
    int nb = 0;
    constexpr size_t bytes = Smem<P, T>::bytes;
    set_smem_attr<P, T, kModeMult>();
    set_smem_attr<P, T, kModeLserk>();
    set_smem_attr<P, T, kModeRk1>();
    set_smem_attr<P, T, kModeRk2>();
    set_smem_attr<P, T, kModeRk3>();
    set_smem_attr<P, T, kModeRk4>();
    auto kern = dg_stage_kernel<P, T, kModeLserk>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, bytes) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace

template <typename T>
hipError_t launch_stage(int order, const StageArgs<T>& a, int mode, int grid, hipStream_t s)
{
    switch (order) {
        case 1: return launch_mode<1, T>(a, mode, grid, s);
        case 2: return launch_mode<2, T>(a, mode, grid, s);
        case 3: return launch_mode<3, T>(a, mode, grid, s);
        case 4: return launch_mode<4, T>(a, mode, grid, s);
        case 5: return launch_mode<5, T>(a, mode, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
int stage_blocks_per_cu(int order)
{
    switch (order) {
        case 1: return occupancy_one<1, T>();
        case 2: return occupancy_one<2, T>();
        case 3: return occupancy_one<3, T>();
        case 4: return occupancy_one<4, T>();
        case 5: return occupancy_one<5, T>();
        default: return 1;
    }
}

int elements_per_block(int order)
{
    switch (order) {
        case 1: return Shape<1>::EB;
        case 2: return Shape<2>::EB;
        case 3: return Shape<3>::EB;
        case 4: return Shape<4>::EB;
        case 5: return Shape<5>::EB;
        default: return 1;
    }
}

template hipError_t launch_stage<double>(int, const StageArgs<double>&, int, int, hipStream_t);
template hipError_t launch_stage<float>(int, const StageArgs<float>&, int, int, hipStream_t);
template int stage_blocks_per_cu<double>(int);
template int stage_blocks_per_cu<float>(int);

}  // namespace dgtd
