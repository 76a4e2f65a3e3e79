// Multi-instance batch kernels: independent meshes of uneven size and order
// advance together.  One persistent launch per LSERK stage walks a
// device-resident work list of (instance, element block) items sorted by
// order; a block reloads the reference matrices only when the order of its
// next item changes.  A one-wave bookkeeping kernel then flips each running
// instance's buffers, advances its time and retires it at its final time, so
// a whole step (5 stages + bookkeeping) is a fixed launch sequence that is
// captured once in a hipGraph while instances come and go through the tables.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "stage_body.hpp"

namespace dgtd {

namespace {

using namespace valu;

template <typename T>
__global__ __launch_bounds__(256) void dg_batch_stage(BatchArgs<T> b)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int loaded = -1;
    const int n = *b.count;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const BatchItem w = b.items[it];
        const BatchInst<T> I = b.inst[w.inst];
        if (I.state != kSlotRunning) continue;  // uniform across the block
        const int order = I.order;
        if (order != loaded) {
            __syncthreads();
            switch (order) {
                case 1: stage_load_ref<1, T>(b.mats[1], b.tabs[1], smem_raw); break;
                case 2: stage_load_ref<2, T>(b.mats[2], b.tabs[2], smem_raw); break;
                case 3: stage_load_ref<3, T>(b.mats[3], b.tabs[3], smem_raw); break;
                case 4: stage_load_ref<4, T>(b.mats[4], b.tabs[4], smem_raw); break;
                default: stage_load_ref<5, T>(b.mats[5], b.tabs[5], smem_raw); break;
            }
            __syncthreads();
            loaded = order;
        }
        StageArgs<T> a{};
        a.K = I.K;
        a.nblk = I.nblk;
        const int src = I.cur ^ (b.stage & 1);
        a.qin = I.q[src];
        a.out = I.q[src ^ 1];
        a.res = I.res;
        a.ca = b.ca;
        a.cb = b.cb;
        a.dt = (T)I.dt_step;
        a.alpha = (T)I.alpha;
        a.geo = I.geo;
        a.conn = I.conn;
        a.unstable = b.stage == 4 ? &b.inst[w.inst].unstable : nullptr;  // the step's last stage
        switch (order) {
            case 1: stage_block<1, T, kModeLserk>(a, smem_raw, w.blk); break;
            case 2: stage_block<2, T, kModeLserk>(a, smem_raw, w.blk); break;
            case 3: stage_block<3, T, kModeLserk>(a, smem_raw, w.blk); break;
            case 4: stage_block<4, T, kModeLserk>(a, smem_raw, w.blk); break;
            default: stage_block<5, T, kModeLserk>(a, smem_raw, w.blk); break;
        }
    }
}

// After the 5 stages: the state is in the other buffer; time advances and an
// instance retires once t > final_time - 1e-8 dt (Solver::run's loop guard,
// Solver.cpp:497), its last step having been truncated to final_time - t (:537).
// An instance whose step stored an unstable entry (the per-step stability test
// of Solver::run, Solver.cpp:499-516, which the reference only reports) is
// retired as failed, so its slot and HBM are freed for others.
template <typename T>
__global__ void dg_batch_advance(BatchInst<T>* inst, int nslots)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    BatchInst<T>& I = inst[i];
    if (I.state != kSlotRunning) return;
    I.cur ^= 1;
    I.t += I.dt_step;
    I.steps += 1;
    if (I.unstable) {
        I.state = kSlotFailed;
    } else if (I.t <= I.final_time - 1e-8 * I.dt) {
        I.dt_step = fmin(I.dt, I.final_time - I.t);
    } else {
        I.state = kSlotDone;
    }
}

}  // namespace

template <typename T>
hipError_t launch_batch_stage(const BatchArgs<T>& b, int grid, size_t smem, hipStream_t s)
{
    hipLaunchKernelGGL(dg_batch_stage<T>, dim3(grid), dim3(256), smem, s, b);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_advance(BatchInst<T>* inst, int nslots, hipStream_t s)
{
    hipLaunchKernelGGL(dg_batch_advance<T>, dim3((nslots + 63) / 64), dim3(64), 0, s, inst, nslots);
    return hipGetLastError();
}

template <typename T>
size_t batch_smem_bytes(int max_order)
{
    size_t m = 0;
    if (max_order >= 1) m = Smem<1, T>::bytes > m ? Smem<1, T>::bytes : m;
    if (max_order >= 2) m = Smem<2, T>::bytes > m ? Smem<2, T>::bytes : m;
    if (max_order >= 3) m = Smem<3, T>::bytes > m ? Smem<3, T>::bytes : m;
    if (max_order >= 4) m = Smem<4, T>::bytes > m ? Smem<4, T>::bytes : m;
    if (max_order >= 5) m = Smem<5, T>::bytes > m ? Smem<5, T>::bytes : m;
    return m;
}

template <typename T>
int batch_blocks_per_cu(size_t smem)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(dg_batch_stage<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dg_batch_stage<T>, 256, smem) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

template hipError_t launch_batch_stage<float>(const BatchArgs<float>&, int, size_t, hipStream_t);
template hipError_t launch_batch_stage<double>(const BatchArgs<double>&, int, size_t, hipStream_t);
template hipError_t launch_batch_advance<float>(BatchInst<float>*, int, hipStream_t);
template hipError_t launch_batch_advance<double>(BatchInst<double>*, int, hipStream_t);
template size_t batch_smem_bytes<float>(int);
template size_t batch_smem_bytes<double>(int);
template int batch_blocks_per_cu<float>(size_t);
template int batch_blocks_per_cu<double>(size_t);

}  // namespace dgtd
