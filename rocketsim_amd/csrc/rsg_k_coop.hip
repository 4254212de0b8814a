// rsg_k_coop.hip -- the cooperative step kernel: 4 lanes per car slot of an
// arena share the tick's independent parts, the arena's working copy in LDS.
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"
#include "rsg_step.h"

using namespace rsg;

// Cooperative variant: `lanes` lanes per arena (4 per car slot), arena in
// LDS.  Lane 0 of each arena runs the serial parts of the tick; the lanes
// share:
//   * the wheel rays, one per lane, cast at the start of the pre-tick into
//     LDS (unless a car respawns that tick: then lane 0 casts them in the
//     loop as usual);
//   * the car pre-ticks, one car per lane, when they are independent of
//     their order (no wheel on a car);
//   * the narrowphase of the world step, one planned pair per lane (the
//     contact callbacks' effects on the arena recorded per pair and applied
//     by lane 0 in pair order; a pair that needs the penetration pass is
//     run again there);
//   * the world step's per-body parts, one body per lane: the proxy AABBs
//     (with the predicted pose) and the integration of the new poses;
//   * the car post-ticks, one car per lane, and the boost-pad tests, pads
//     spread over the lanes.
// Per-arena LDS block: ArenaLocal prefix, ray results, two flag words.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RSG_COOP_WAVES_PER_EU)))
rsg_step_arena_coop_kernel(BatchDev d, ArenaGeom geom, int ticks, int stride, int arena_bytes, int lanes) {
    extern __shared__ __align__(16) unsigned char rsg_lds[];
    const int slot = (int)threadIdx.x / lanes;
    const int k = (int)threadIdx.x - slot * lanes;
    const int a = (int)blockIdx.x * ((int)blockDim.x / lanes) + slot;
    const bool live = a < d.n;
    unsigned char* base = rsg_lds + (size_t)slot * (size_t)stride;
    ArenaLocal& A = *reinterpret_cast<ArenaLocal*>(base);
    RayHit* hits = reinterpret_cast<RayHit*>(base + arena_bytes);
    int* ahead = reinterpret_cast<int*>(base + arena_bytes + (size_t)(lanes) * sizeof(RayHit));
    TickParams P;
    P.mode = d.mode;
    P.dt = d.tick_time;
    P.seed = d.seed;
    P.arena_id = (uint32_t)a;
    P.synth_controls = d.synth_controls != 0;
    P.auto_reset = d.auto_reset != 0;
    const ArenaGeom* mp = mode_has_arena(d.mode) ? &geom : nullptr;
    StepSpace* ws = live ? arena_space(d, a) : nullptr;
    PairLog* log = live ? arena_log(d, a) : nullptr;
    int* concurrent = ahead + 1;
    if (live && k == 0) load_arena(d, a, A);
    __syncthreads();
    for (int t = 0; t < ticks; t++) {
        if (live && k == 0) *ahead = tick_begin(A, P) ? 1 : 0;
        __syncthreads();
        if (live && *ahead && k < 4 * A.n_cars) tick_cast_ray(A, P, mp, k >> 2, k & 3, hits[k]);
        __syncthreads();
        if (live && k == 0) {
            const RayHit* pre = *ahead ? hits : nullptr;
            *concurrent = car_pre_ticks_independent(A, pre) ? 1 : 0;
            if (!*concurrent) {
                World world(A, mp, mode_has_arena(d.mode));
                for (int oi = 0; oi < A.n_cars; oi++) car_pre_tick(A, P, world, A.order[oi], pre);
            }
        }
        __syncthreads();
        if (live && *concurrent && k < A.n_cars) {
            World world(A, mp, mode_has_arena(d.mode));
            car_pre_tick(A, P, world, k, hits);
        }
        __syncthreads();
        if (live && k == 0) {
            pads_pre_tick(A, P);
            tick_world_prologue(A, P, mp, ws);
        }
        __syncthreads();
        if (live) tick_world_aabbs(A, P, mp, ws, k, lanes);
        __syncthreads();
        if (live && k == 0) tick_world_broadphase(A, P, mp, ws);
        __syncthreads();
        if (live) tick_world_pairs(A, P, mp, ws, log, k, lanes);
        __syncthreads();
        if (live && k == 0) tick_world_solve(A, P, mp, ws, log);
        __syncthreads();
        if (live) tick_world_integrate(A, P, mp, ws, k, lanes);
        __syncthreads();
        if (live && k < A.n_cars) car_post_tick(A, P, k);
        __syncthreads();
        if (live && mode_has_arena(d.mode))
            for (int p = k; p < A.n_pads; p += lanes) pad_check_cars(A, p);
        __syncthreads();
        if (live && k == 0) tick_end(A, P);
        __syncthreads();
    }
    if (live && k == 0) store_arena(d, a, A);
}

namespace rsg {

hipError_t launch_step_coop(const BatchDev& d, const ArenaGeom& geom, int ticks, int grid, int per_block, int stride,
                            int arena_bytes, int lanes, hipStream_t s) {
    const size_t lds = (size_t)per_block * (size_t)stride;
    hipLaunchKernelGGL(rsg_step_arena_coop_kernel, dim3(grid), dim3(per_block * lanes), lds, s, d, geom, ticks,
                       stride, arena_bytes, lanes);
    return hipGetLastError();
}

}  // namespace rsg
