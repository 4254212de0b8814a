// rsg_k_lds.hip -- the step kernel with one lane per arena and the arena's
// working copy in LDS.
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"
#include "rsg_step.h"

using namespace rsg;

// One lane per arena; the arena's working copy lives in LDS, `stride` bytes
// per lane (the ArenaLocal prefix up to the batch's largest car count).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RSG_STEP_WAVES_PER_EU)))
rsg_step_arena_lds_kernel(BatchDev d, ArenaGeom geom, int ticks, int stride) {
    extern __shared__ __align__(16) unsigned char rsg_lds[];
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= d.n) return;
    ArenaLocal& A = *reinterpret_cast<ArenaLocal*>(rsg_lds + (size_t)threadIdx.x * (size_t)stride);
    step_arena(d, geom, a, ticks, A, arena_space(d, a));
}

namespace rsg {

hipError_t launch_step_lds(const BatchDev& d, const ArenaGeom& geom, int ticks, int grid, int per_block, int stride,
                           hipStream_t s) {
    const size_t lds = (size_t)per_block * (size_t)stride;
    hipLaunchKernelGGL(rsg_step_arena_lds_kernel, dim3(grid), dim3(per_block), lds, s, d, geom, ticks, stride);
    return hipGetLastError();
}

}  // namespace rsg
