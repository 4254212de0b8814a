// rsg_kernels.h -- host-side launchers of the library's kernels.
//
// Each kernel lives in its own translation unit (rsg_k_*.hip), so the
// library's device code compiles in parallel; the C ABI in rsg_batch.hip
// enqueues work only through these functions.  Every launcher enqueues one
// kernel on `s` and returns hipGetLastError().
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rsg_state.h"
#include "rsg_layout.h"
#include "rsg_mesh.h"
#include "rsg_ball_pred.h"
#include "rsg_game_events.h"

namespace rsg {

// Step kernels (rsg_k_plain.hip, rsg_k_lds.hip, rsg_k_coop.hip): `grid`
// workgroups of `per_block` arenas advance `ticks` ticks.
hipError_t launch_step_plain(const BatchDev& d, const ArenaGeom& geom, int ticks, int grid, int per_block,
                             hipStream_t s);
hipError_t launch_step_lds(const BatchDev& d, const ArenaGeom& geom, int ticks, int grid, int per_block, int stride,
                           hipStream_t s);
hipError_t launch_step_coop(const BatchDev& d, const ArenaGeom& geom, int ticks, int grid, int per_block, int stride,
                            int arena_bytes, int lanes, hipStream_t s);
// Kickoff reset of arena `only_arena` (all arenas when < 0).
hipError_t launch_reset(const BatchDev& d, int only_arena, uint64_t seed, hipStream_t s);

// Differentiable rollouts (rsg_k_diff.hip).
constexpr int AIR_MAX_K = 32;
hipError_t launch_air_forward(int n, int ticks, int K, float dt, const float* s0, const float* ctrl,
                              const float* target, float* s_out, float* loss, float* ckpt, hipStream_t s);
hipError_t launch_air_backward(int n, int ticks, int K, float dt, const float* ctrl, const float* target,
                               const float* ckpt, const float* dloss, float* dctrl, float* ds0, hipStream_t s);
hipError_t launch_shot_forward(int n, int ticks, int K, float dt, const float* state0, const float* ctrl,
                               const float* target, float* state_out, float* loss, float* ckpt, uint32_t* flags,
                               hipStream_t s);
hipError_t launch_shot_backward(int n, int ticks, int K, float dt, const float* ctrl, const float* target,
                                const float* ckpt, const float* dloss, float* dctrl, float* ds0, hipStream_t s);

// Ball prediction (rsg_k_pred.hip): `grid` workgroups of `per_wave` trackers.
hipError_t launch_ball_pred(const BatchDev& pa, const ArenaGeom& geom, const PredDev& p, int kind, const BatchDev& src,
                            int stride, int per_wave, int grid, hipStream_t s);

// GameEventTracker update of every arena (rsg_k_events.hip).
hipError_t launch_game_events(const BatchDev& d, const GameEventDev& g, const rsg_game_event_config& cfg,
                              hipStream_t s);

// Status summary (rsg_k_status.hip): the OR of the n status words and the
// number of nonzero ones into out[0], out[1] (device memory), through
// 2 * status_blocks(n) words of partial results.
constexpr int STATUS_BLOCK = 256;
inline int status_blocks(int n) { return (n + STATUS_BLOCK - 1) / STATUS_BLOCK; }
hipError_t launch_status_summary(const uint32_t* status, int n, uint32_t* partials, uint32_t* out, hipStream_t s);

}  // namespace rsg
