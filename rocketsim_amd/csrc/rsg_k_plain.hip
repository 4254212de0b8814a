// rsg_k_plain.hip -- the plain step kernel (one lane per arena, the arena's
// working copy in private memory) and the kickoff-reset kernel.
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"
#include "rsg_step.h"

using namespace rsg;

// One lane per arena; the arena's working copy lives in the lane's private
// (scratch) memory.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RSG_STEP_WAVES_PER_EU)))
rsg_step_arena_kernel(BatchDev d, ArenaGeom geom, int ticks) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= d.n) return;
    ArenaLocal A;
    StepSpace ws;  // the naive mapping: every lane's work space in its private memory
    step_arena(d, geom, a, ticks, A, &ws);
}

__global__ void __launch_bounds__(64) rsg_reset_kernel(BatchDev d, int only_arena, uint64_t seed) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= d.n) return;
    if (only_arena >= 0 && a != only_arena) return;
    ArenaLocal A;
    load_arena(d, a, A);
    arena_kickoff_reset(A, seed, (uint32_t)a);
    store_arena(d, a, A);
}

namespace rsg {

hipError_t launch_step_plain(const BatchDev& d, const ArenaGeom& geom, int ticks, int grid, int per_block,
                             hipStream_t s) {
    // A lane's private segment above the device's 64 KiB would abort the
    // queue at dispatch; refuse the launch instead.
    static int private_ok = -1;
    if (private_ok < 0) {
        hipFuncAttributes attr;
        if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(rsg_step_arena_kernel)) != hipSuccess)
            return hipErrorInvalidDeviceFunction;
        private_ok = attr.localSizeBytes <= 64 * 1024 ? 1 : 0;
    }
    if (!private_ok) return hipErrorLaunchOutOfResources;
    hipLaunchKernelGGL(rsg_step_arena_kernel, dim3(grid), dim3(per_block), 0, s, d, geom, ticks);
    return hipGetLastError();
}

hipError_t launch_reset(const BatchDev& d, int only_arena, uint64_t seed, hipStream_t s) {
    const int block = 64, grid = (d.n + block - 1) / block;
    hipLaunchKernelGGL(rsg_reset_kernel, dim3(grid), dim3(block), 0, s, d, only_arena, seed);
    return hipGetLastError();
}

}  // namespace rsg
