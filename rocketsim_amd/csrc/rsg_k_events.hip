// rsg_k_events.hip -- the GameEventTracker kernel (rsg_game_events.h).
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"

using namespace rsg;

// GameEventTracker::Update for every arena (rsg_game_events.h), one lane each.
__global__ void __launch_bounds__(256) rsg_game_events_kernel(BatchDev d, GameEventDev g, rsg_game_event_config cfg) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= d.n) return;
    game_events_update(d, g, cfg, a);
}

namespace rsg {

hipError_t launch_game_events(const BatchDev& d, const GameEventDev& g, const rsg_game_event_config& cfg,
                              hipStream_t s) {
    const int block = 256, grid = (d.n + block - 1) / block;
    hipLaunchKernelGGL(rsg_game_events_kernel, dim3(grid), dim3(block), 0, s, d, g, cfg);
    return hipGetLastError();
}

}  // namespace rsg
