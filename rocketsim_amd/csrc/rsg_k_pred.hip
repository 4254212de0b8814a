// rsg_k_pred.hip -- the ball-prediction kernel (BallPredTracker updates,
// rsg_ball_pred.h).
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"
#include "rsg_step.h"

using namespace rsg;

// Ball prediction (rsg_ball_pred.h): one update of every tracker,
// `per_wave` trackers per 64-lane block (the rest of the wave idles, so small
// batches still reach every SIMD).  Tracker i's prediction arena (a
// ball-only arena of `pa`) is held in LDS, `stride` bytes per tracker.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RSG_STEP_WAVES_PER_EU)))
rsg_ball_pred_kernel(BatchDev pa, ArenaGeom geom, PredDev p, int kind, BatchDev src, int stride, int per_wave) {
    extern __shared__ __align__(16) unsigned char rsg_lds[];
    if ((int)threadIdx.x >= per_wave) return;
    const int i = blockIdx.x * per_wave + threadIdx.x;
    if (i >= pa.n) return;
    ArenaLocal& A = *reinterpret_cast<ArenaLocal*>(rsg_lds + (size_t)threadIdx.x * (size_t)stride);
    ball_pred_update(pa, mode_has_arena(pa.mode) ? &geom : nullptr, p, kind, src, i, A, arena_space(pa, i));
}

namespace rsg {

hipError_t launch_ball_pred(const BatchDev& pa, const ArenaGeom& geom, const PredDev& p, int kind, const BatchDev& src,
                            int stride, int per_wave, int grid, hipStream_t s) {
    hipLaunchKernelGGL(rsg_ball_pred_kernel, dim3(grid), dim3(64), (size_t)per_wave * (size_t)stride, s, pa, geom, p,
                       kind, src, stride, per_wave);
    return hipGetLastError();
}

}  // namespace rsg
