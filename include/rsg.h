/*
 * rsg.h -- C ABI of the MI355X batched RocketSim tick step (librsg.so).
 *
 * The reference exposes a C++ class API only (src/Sim/Arena/Arena.h,
 * src/Sim/Car/Car.h, src/Sim/Ball/Ball.h); third-party bindings wrap that C++
 * (README.md:31-35).  This header is the drop-in boundary for its hot path,
 * Arena::Step(ticks) (src/Sim/Arena/Arena.cpp:689-756): a batch of N
 * independent arenas lives in GPU memory and is stepped by hand-written HIP
 * kernels.  Every entry point returns 0 on success and a negative code on
 * error (no exceptions cross the boundary); rsg_last_error() describes the
 * last failure on the calling thread.  The C++ facade in include/rsg.hpp
 * rethrows errors as std::runtime_error like the reference's RS_ERR_CLOSE
 * (src/Framework.h:67-72).
 *
 * Threading: one host thread per batch handle; all work is ordered on the
 * stream passed to rsg_batch_step (NULL = the default stream).
 */
#ifndef RSG_H
#define RSG_H

#include <stdint.h>
#include "rsg_state.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsg_batch rsg_batch;

/* Description of a batch (replaces Arena::Create + AddCar,
 * src/Sim/Arena/Arena.cpp:50-62, 416-540).  Cars of arena a are
 * car_team[off_a .. off_a + cars_per_arena[a]) with off_a the prefix sum. */
typedef struct rsg_batch_desc {
    int n_arenas;
    int mode;                     /* RSG_MODE_SOCCAR, RSG_MODE_HEATSEEKER or RSG_MODE_THE_VOID */
    float tick_rate;              /* ticks per second, 120 in every workload */
    uint64_t seed;                /* seeds kickoff shuffles, respawns and the synthetic controls */
    int auto_reset;               /* kickoff reset after a goal (Arena.cpp:748-752 + 113-223) */
    int synth_controls;           /* draw controls on device each tick (include/rsg_workload.h) */
    const int* cars_per_arena;    /* n_arenas entries (each 0..RSG_MAX_CARS) */
    const uint8_t* car_team;      /* RSG_TEAM_*, one per car */
    const uint8_t* car_preset;    /* RSG_PRESET_*, one per car (ignored with car_configs) */
    /* MutatorConfig of every arena, set before its cars are added (Arena::Create,
     * SetMutatorConfig, then AddCar: the ball is made with the config's radius
     * and mass, cars keep the car mass); NULL for the mode's defaults */
    const rsg_mutator_config* mutators;
    /* each car's CarConfig (AddCar(team, config), Arena.h:71), one per car;
     * NULL for the presets of car_preset */
    const rsg_car_config* car_configs;
    /* ArenaConfig::noBallRot == false (src/Sim/Arena/ArenaConfig/ArenaConfig.h,
     * Ball.cpp:109): the ball's pose rotates with its angular velocity; 0 (the
     * reference's default): it keeps its rotation */
    int ball_rotation;
    /* ArenaConfig::useCustomBoostPads / customBoostPads (ArenaConfig.h,
     * Arena.cpp:501-508): n_custom_pads (1..34) pads in place of the mode's,
     * in this order; NULL / 0 for the mode's pads.  Every car is tested
     * against every custom pad (Arena.cpp:724-728). */
    const rsg_boost_pad_config* custom_pads;
    int n_custom_pads;
} rsg_batch_desc;

/* Last error message of the calling thread ("" if none). */
const char* rsg_last_error(void);

/* Replaces RocketSim::InitFromMem (src/RocketSim.h:32, src/RocketSim.cpp:92-185)
 * for the Soccar arena (Soccar and Heatseeker batches): n collision mesh
 * files in the reference's .cmf format
 * (src/CollisionMeshFile/CollisionMeshFile.cpp:11-61), in the
 * order the reference's list holds them (1 <= n <= 26).  Each mesh becomes one static body
 * of every arena, in this order and before the floor, ceiling and side walls
 * (Arena::_SetupArenaCollisionShapes, src/Sim/Arena/Arena.cpp:990-1068).
 * Unlike the reference, a later call replaces the meshes for batches stepped
 * afterwards (it waits for the device first).  Must be called before
 * creating a SOCCAR or HEATSEEKER batch. */
int rsg_init_meshes(const void* const* cmf_data, const uint64_t* sizes, int n);
/* rsg_init_meshes with one mesh file. */
int rsg_init_mesh(const void* cmf_data, uint64_t size);
/* rsg_init_meshes for the arena of `mode`: RSG_MODE_SOCCAR or
 * RSG_MODE_HEATSEEKER (the Soccar arena: 4 planes) or RSG_MODE_HOOPS (the
 * Hoops arena: 6 planes, src/Sim/Arena/Arena.cpp:1036-1068; a mesh of 505
 * vertices is a net, which cars and their wheel rays pass through,
 * Arena.cpp:1005-1024).  1 <= n <= 26. */
int rsg_init_arena_meshes(int mode, const void* const* cmf_data, const uint64_t* sizes, int n);

/* Replaces Arena::Create + AddCar (src/Sim/Arena/Arena.h:71,111).  Cars start
 * at random kickoff positions (Arena::ResetToRandomKickoff). */
rsg_batch* rsg_batch_create(const rsg_batch_desc* desc);
void rsg_batch_destroy(rsg_batch* b);

int rsg_batch_num_arenas(const rsg_batch* b);
/* Total number of cars and index of arena a's first car in flattened arrays. */
int rsg_batch_num_cars(const rsg_batch* b);
int rsg_batch_car_offset(const rsg_batch* b, int arena);

/* Replaces Arena::Step(ticks) (src/Sim/Arena/Arena.h:132) for every arena of
 * the batch.  `stream` is a hipStream_t (NULL for the default stream). */
int rsg_batch_step(rsg_batch* b, int ticks, void* stream);
/* Waits for all work on the batch's last stream. */
int rsg_batch_sync(rsg_batch* b);

/* Full-state access (Car::Get/SetState Car.h:151-152, Ball::Get/SetState
 * Ball.h:68-69, BoostPad::Get/SetState BoostPad.h:45-46), including the
 * per-wheel values the reference keeps in btVehicleRL between ticks.
 * Flattened arrays: cars in (arena, car) order, pads 34 per arena.  The car
 * and control arrays may be NULL when the batch holds no cars. */
int rsg_batch_get_cars(rsg_batch* b, rsg_car_state* out);
int rsg_batch_set_cars(rsg_batch* b, const rsg_car_state* in);
int rsg_batch_get_balls(rsg_batch* b, rsg_ball_state* out);
int rsg_batch_set_balls(rsg_batch* b, const rsg_ball_state* in);
int rsg_batch_get_pads(rsg_batch* b, rsg_pad_state* out);
int rsg_batch_set_pads(rsg_batch* b, const rsg_pad_state* in);
/* Car::controls (Car.h:149), held for the following rsg_batch_step calls. */
int rsg_batch_set_controls(rsg_batch* b, const rsg_controls* in);
int rsg_batch_get_controls(rsg_batch* b, rsg_controls* out);
/* The order the per-car updates run in: the iteration order of the
 * reference's Arena::_cars, an std::unordered_set<Car*> (Arena.h:35) walked
 * by Arena::Step's pre-tick and post-tick/boost-pad loops (Arena.cpp:706-734)
 * and by ResetToRandomKickoff (Arena.cpp:150-152).  Its order follows the
 * cars' heap addresses; with libstdc++ and no bucket collision it is the
 * reverse of creation order, which is the default here.  `order` holds, per
 * arena (same flattening as the car arrays), a permutation of the arena's
 * car indices; anything else is rejected. */
int rsg_batch_set_car_order(rsg_batch* b, const int32_t* order);
int rsg_batch_get_car_order(rsg_batch* b, int32_t* order);
/* Launch geometry of rsg_batch_step (no counterpart in the reference, which
 * steps one arena per call): arenas per workgroup, 1..64, or 0 for the
 * default that spreads the batch over every SIMD of the device.  Results do
 * not depend on it.  rsg_batch_get_launch returns the value in use. */
int rsg_batch_set_launch(rsg_batch* b, int arenas_per_block);
int rsg_batch_get_launch(rsg_batch* b);
/* Step kernel variant (no counterpart in the reference):
 *   RSG_KERNEL_SCRATCH  one lane per arena, working copy in private memory;
 *   RSG_KERNEL_LDS      one lane per arena, working copy in LDS;
 *   RSG_KERNEL_COOP     4 lanes per car slot per arena, working copy in LDS,
 *                       the wheel rays of a tick cast by all lanes at once;
 *   RSG_KERNEL_AUTO     COOP wherever its per-arena block fits, else LDS
 *                       when its footprint covers the batch in one round
 *                       of waves, else SCRATCH.
 * Results do not depend on it; rsg_batch_get_kernel returns the kernel in
 * use.  (rsg_batch_set_launch's arenas_per_block counts arenas, not lanes.) */
enum { RSG_KERNEL_AUTO = 0, RSG_KERNEL_SCRATCH = 1, RSG_KERNEL_LDS = 2, RSG_KERNEL_COOP = 3 };
int rsg_batch_set_kernel(rsg_batch* b, int kernel);
int rsg_batch_get_kernel(rsg_batch* b);
/* Autotuner (no counterpart in the reference): times every feasible kernel
 * and arenas-per-workgroup choice for `ticks` ticks on a device copy of the
 * batch state (the batch itself is not advanced), keeps the fastest as the
 * batch's kernel and launch shape, and reports it (any out pointer may be
 * NULL; out_ms is per tick).  Allocates temporary device memory. */
int rsg_batch_autotune(rsg_batch* b, int ticks, void* stream, int* out_kernel, int* out_per_block, float* out_ms);
/* hipGraph of multi-tick chunks (no counterpart in the reference): captures
 * `launches` step launches of `ticks_per_launch` ticks each, with the batch's
 * current kernel and launch shape; rsg_batch_graph_launch replays it on a
 * stream (the same as that many rsg_batch_step calls), and
 * rsg_batch_graph_ticks returns the ticks one replay advances.  Re-create the
 * graph after changing the kernel or launch shape. */
int rsg_batch_graph_create(rsg_batch* b, int ticks_per_launch, int launches);
int rsg_batch_graph_launch(rsg_batch* b, void* stream);
int rsg_batch_graph_ticks(rsg_batch* b);
/* Arena::tickCount (Arena.h:61). */
int rsg_batch_get_tick_counts(rsg_batch* b, uint64_t* out);
int rsg_batch_set_tick_counts(rsg_batch* b, const uint64_t* in);

/* Per-arena goal events, replacing the goal-score callback
 * (Arena::SetGoalScoreCallback, Arena.h:24,102): goals scored by blue (ball
 * in the orange net), by orange, and the tick of the last goal. */
int rsg_batch_get_goals(rsg_batch* b, uint32_t* blue, uint32_t* orange, uint64_t* last_goal_tick);
/* Status of every arena (RSG_STATUS_* bits, rsg_state.h), in place of the
 * reference's exceptions (RS_ERR_CLOSE, src/Framework.h:67-72): writes one
 * word per arena to `per_arena` and their OR to `any` (either may be NULL)
 * and returns the number of arenas with a bit set.  Waits for the batch's
 * last step.  rsg_batch_clear_status zeroes the words.  With
 * rsg_batch_set_strict(b, 1), rsg_batch_step and rsg_batch_graph_launch wait
 * for their work and return -1, naming the first such arena in
 * rsg_last_error(), as the reference's Arena::Step would throw. */
int rsg_batch_get_status(rsg_batch* b, uint32_t* per_arena, uint32_t* any);
/* Device-side summary of the same words (no counterpart in the reference):
 * enqueues on `stream`, after the batch's last step, kernels that write the
 * OR of every arena's status word to out[0] and the number of arenas with a
 * bit set to out[1]; `out` is two words of device memory.  No host round
 * trip, so a stepping loop can test it on the device.  Allocates its
 * partial-result buffer (two words per 256 arenas) on the first call. */
int rsg_batch_status_summary(rsg_batch* b, uint32_t* out, void* stream);
/* rsg_batch_status_summary without any allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops).  Same result; the partial results go
 * to a file-scope device array sized for RSG_STATUS_NOALLOC_MAX_ARENAS
 * arenas, and a larger batch is refused with an error.  Every call on a GPU
 * shares that array: calls that overlap on one GPU race. */
enum { RSG_STATUS_NOALLOC_MAX_ARENAS = 1 << 17 };
int rsg_batch_status_summary_noalloc(rsg_batch* b, uint32_t* out, void* stream);
int rsg_batch_clear_status(rsg_batch* b);
int rsg_batch_set_strict(rsg_batch* b, int strict);

/* Arena::ResetToRandomKickoff (Arena.cpp:113-223) for one arena, or all
 * arenas when arena < 0; the kickoff order is drawn from `seed`. */
int rsg_batch_reset_kickoff(rsg_batch* b, int arena, uint64_t seed);

/* Game events (rsg_event, include/rsg_state.h) in place of the goal-score
 * and car-bump callbacks (Arena.h:24-25,99-108).  rsg_batch_enable_events
 * allocates a device ring of `capacity` events (0 frees it; events are off by
 * default).  rsg_batch_poll_events waits for the batch's stream, copies up to
 * max_out events in (arena, per-arena order) order into `out`, empties the
 * ring and returns the number copied; *n_lost (if not NULL) receives how many
 * were dropped because the ring or `out` was full. */
int rsg_batch_enable_events(rsg_batch* b, int capacity);
int rsg_batch_poll_events(rsg_batch* b, rsg_event* out, int max_out, int* n_lost);

/* Differentiable air control (no counterpart in the reference; its forward
 * model is the reference's tick of an airborne car: Car::_UpdateAirTorque
 * Car.cpp:584-667, Car::_UpdateBoost Car.cpp:486-530 with the boost input
 * relaxed to [0, 1], the world step's integration and the speed caps of
 * Car::_FinishPhysicsTick).  All pointers are device memory, structure of
 * arrays by word, n cars:
 *   state   [13][n]  position (uu), quaternion (x, y, z, w), velocity (uu/s),
 *                    angular velocity (rad/s)
 *   controls [ticks][4][n]  pitch, yaw, roll in [-1, 1], boost in [0, 1]
 *   target  [7][n]   quaternion, position (uu)
 *   checkpoints [ceil(ticks / ckpt_every)][13][n]  (written by the rollout,
 *                    read by the backward pass; ckpt_every <= 32)
 * rsg_air_rollout steps every car `ticks` ticks and writes the final state
 * and loss = 2 acos|<q, q_target>| + 1e-4 |p - p_target|^2 (uu) when the
 * pointers are not NULL.  rsg_air_rollout_backward recomputes each segment
 * from its checkpoint and writes d(sum dloss * loss)/dcontrols and, if not
 * NULL, /dstate0. */
int rsg_air_rollout(int n, int ticks, int ckpt_every, float tick_rate, const float* state0, const float* controls,
                    const float* target, float* state_out, float* loss, float* checkpoints, void* stream);
int rsg_air_rollout_backward(int n, int ticks, int ckpt_every, float tick_rate, const float* controls,
                             const float* target, const float* checkpoints, const float* dloss, float* dcontrols,
                             float* dstate0, void* stream);

/* Differentiable aerial shot: one airborne car and the ball per pair, through
 * the car-ball hit (no counterpart in the reference; the forward model is the
 * reference's tick for an airborne Octane and a flying ball with no world
 * contact: the air control and boost above, ball damping and gravity, the
 * car hitbox vs ball contact with Ball::_OnHit's extra impulse (Ball.cpp:
 * 236-290) and its sequential-impulse solve, integration and the speed caps;
 * rocketsim_amd/csrc/rsg_shot.h).  Device memory, structure of arrays, n pairs:
 *   state   [22][n]  car position (uu), quaternion, velocity (uu/s), angular
 *                    velocity; ball position (uu), velocity (uu/s), angular
 *                    velocity
 *   controls [ticks][4][n]  pitch, yaw, roll in [-1, 1], boost in [0, 1]
 *   target  [3][n]   where the ball should end (uu); loss = |ball - target|
 *   checkpoints [ceil(ticks / ckpt_every)][23][n]  (ckpt_every <= 32)
 *   flags   [n]      bit 0: the car touched the ball; bit 1: the ball centre
 *                    entered the hitbox core (not modelled); bit 2: the ball
 *                    or car came near the floor, ceiling or walls (not
 *                    modelled: the forward result there is not the reference's)
 * The backward pass recomputes each segment from its checkpoint and, per
 * tick, propagates one tangent per step input to form the exact
 * vector-Jacobian product for the contact events of the forward pass. */
int rsg_shot_rollout(int n, int ticks, int ckpt_every, float tick_rate, const float* state0, const float* controls,
                     const float* target, float* state_out, float* loss, float* checkpoints, uint32_t* flags,
                     void* stream);
int rsg_shot_rollout_backward(int n, int ticks, int ckpt_every, float tick_rate, const float* controls,
                              const float* target, const float* checkpoints, const float* dloss, float* dcontrols,
                              float* dstate0, void* stream);

/* Arena::IsBallScored (Arena.cpp:893-901) per arena, with the arena's
 * goalBaseThresholdY and ballRadius. */
int rsg_batch_is_ball_scored(rsg_batch* b, uint8_t* out);

/* The CarConfig presets CAR_CONFIG_OCTANE ... CAR_CONFIG_PSYCLOPS
 * (src/Sim/Car/CarConfig/CarConfig.cpp:93-100), by RSG_PRESET_*. */
int rsg_car_config_preset(int preset, rsg_car_config* out);

/* MutatorConfig(GameMode) (src/Sim/MutatorConfig/MutatorConfig.cpp:5-40):
 * the defaults of `mode` (RSG_MODE_*). */
int rsg_mutator_config_default(int mode, rsg_mutator_config* out);
/* Arena::SetMutatorConfig (src/Sim/Arena/Arena.h:53, Arena.cpp:17-48) on
 * arena `arena`, or on every arena of the batch (-1), with the reference's
 * side effects: the new gravity goes to the bodies that are awake (a
 * demolished car and a sleeping ball keep theirs); a new ball radius or mass
 * remakes the ball at rest at the centre, at the radius's height; a new car
 * mass gives every car the ball's mass and the car mass's inertia (the
 * reference passes ballMass to setMassProps).  A batch starts with the
 * mode's defaults in every arena.  Fails on non-finite fields, on a
 * non-positive ball radius, ball mass or car mass and on an unknown
 * demo_mode. */
int rsg_batch_set_mutator_config(rsg_batch* b, int arena, const rsg_mutator_config* cfg);
/* Arena::GetMutatorConfig (Arena.h:52). */
int rsg_batch_get_mutator_config(rsg_batch* b, int arena, rsg_mutator_config* out);

/* ---- Game-event tracking: one GameEventTracker
 * (src/Sim/GameEventTracker/GameEventTracker.cpp:51-166) per arena of a
 * SOCCAR batch, on the device.  Shots, goals (scorer and passer) and saves go
 * to the tracker's own event ring (RSG_EVENT_SHOT / _SCORER / _SAVE). */
typedef struct rsg_game_events rsg_game_events;

void rsg_game_event_config_default(rsg_game_event_config* out);

/* One tracker per arena of b (n arenas, SOCCAR only: the reference's
 * IsBallProbablyGoingIn supports no other mode here); `capacity` events are
 * kept between polls (later ones are counted as lost). */
rsg_game_events* rsg_game_events_create(rsg_batch* b, const rsg_game_event_config* cfg, int capacity);
void rsg_game_events_destroy(rsg_game_events* t);

/* GameEventTracker::Update (:51-158) for every arena of b, on `stream`.
 * Call it at the interval the reference's users call Update (every tick, or
 * at the tick-skip of an RL loop). */
int rsg_game_events_update(rsg_game_events* t, rsg_batch* b, void* stream);

/* Events since the last poll, sorted by arena and then by occurrence. */
int rsg_game_events_poll(rsg_game_events* t, rsg_event* out, int max_out, int* n_lost);

/* ResetPersistentInfo (:160-166) for one arena, or all with arena = -1. */
int rsg_game_events_reset(rsg_game_events* t, int arena);

/* ---- Snapshots in the reference's byte format (Arena::Serialize,
 * src/Sim/Arena/Arena.cpp:546-573; Arena::DeserializeNew, :575-644).
 *
 * rsg_snapshot_serialize writes the reference's bytes for one arena, then an
 * extension block the reference's reader ignores, holding the full state
 * records (hidden per-wheel values, flags, counters, broadphase records), so
 * the reference can load the file and this library restores it exactly.
 * Returns the byte count; writes only when cap is large enough (call with
 * out = NULL to size the buffer).  Host work only, no GPU. */
int64_t rsg_snapshot_serialize(const rsg_arena_snapshot* s, uint8_t* out, uint64_t cap);

/* Reads the reference's bytes (with or without the extension block).  Only
 * SOCCAR, HEATSEEKER and THE_VOID with the mode's default ArenaConfig, MutatorConfig and preset
 * car configs are accepted; anything else is an error.  Without the
 * extension, states are as Car/Ball::SetState leave them in a new arena
 * (tick counters and per-wheel values 0, broadphase records of newly created
 * proxies).  Host work only. */
int rsg_snapshot_deserialize(const uint8_t* data, uint64_t size, rsg_arena_snapshot* out);

/* One arena of a batch as a snapshot: cars in its car update order. */
int rsg_batch_get_snapshot(rsg_batch* b, int arena, rsg_arena_snapshot* out);

/* Loads a snapshot into an arena slot.  Mode, tick time and car count must
 * match the slot; its cars, taken in ascending id order (the reference's
 * creation order), must have the slots' teams and presets.  The snapshot's
 * car order becomes the arena's car update order and its ids the cars' ids. */
int rsg_batch_set_snapshot(rsg_batch* b, int arena, const rsg_arena_snapshot* s);

/* ---- Ball prediction: a batch of n BallPredTracker
 * (src/Sim/BallPredTracker/BallPredTracker.h:7-36).  Tracker i owns a
 * ball-only arena of `mode` (the reference's tracker makes its own car-less
 * copy of the arena, BallPredTracker.cpp:5-12) and num_pred_ticks predicted
 * ball states (predData), kept on the device.  Updates run on `stream`
 * (NULL = the default stream); the getters wait for the last update. */
typedef struct rsg_ball_pred rsg_ball_pred;

/* BallPredTracker::BallPredTracker (BallPredTracker.cpp:5-12) without the
 * first update: the prediction is empty until an update runs.  mode is
 * RSG_MODE_SOCCAR (needs rsg_init_mesh) or RSG_MODE_THE_VOID;
 * num_pred_ticks >= 1. */
rsg_ball_pred* rsg_ball_pred_create(int mode, int n, int num_pred_ticks, float tick_rate);
void rsg_ball_pred_destroy(rsg_ball_pred* p);

/* ForceUpdateAllPred (BallPredTracker.cpp:63-71) for every tracker from the
 * host states initial[n] (their bp_cell / bp_seq are not used). */
int rsg_ball_pred_force_update(rsg_ball_pred* p, const rsg_ball_state* initial, void* stream);

/* UpdatePredManual (BallPredTracker.cpp:25-61) for every tracker: cur[n] and
 * ticks_since_last_update[n] are host arrays.  A tracker whose prediction at
 * that tick Matches (Ball.cpp:12-17) the current state drops the consumed
 * ticks and predicts only the new tail; any other re-predicts in full. */
int rsg_ball_pred_update_manual(rsg_ball_pred* p, const rsg_ball_state* cur, const int32_t* ticks_since_last_update,
                                void* stream);

/* UpdatePredFromArena (BallPredTracker.cpp:18-23) with tracker i following
 * arena i of batch b (same n, mode and tick rate): the ball state and tick
 * count are read on the device, with no host round trip. */
int rsg_ball_pred_update_from_batch(rsg_ball_pred* p, rsg_batch* b, void* stream);

/* predData of one tracker: writes min(size, max_out) states, tick 0 first, and
 * returns the prediction's size (0 before the first update).  The states'
 * bp_cell / bp_seq are 0 (BallState carries no broadphase record). */
int rsg_ball_pred_get(rsg_ball_pred* p, int tracker, rsg_ball_state* out, int max_out);

/* Every tracker's prediction, tracker-major: out[i * num_pred_ticks + t]. */
int rsg_ball_pred_get_all(rsg_ball_pred* p, rsg_ball_state* out);

/* GetBallStateForTime (BallPredTracker.cpp:73-79): the state at index
 * clamp(pred_time / tick_time, 0, size - 1), truncated; an error while the
 * prediction is empty. */
int rsg_ball_pred_state_for_time(rsg_ball_pred* p, int tracker, float pred_time, rsg_ball_state* out);

/* Launch shape of the updates: trackers per 64-lane wave (1..64), or 0 for
 * the default, ceil(n / SIMDs) so that every SIMD holds a wave. */
int rsg_ball_pred_set_launch(rsg_ball_pred* p, int trackers_per_wave);

/* Device view of the predictions for consumers on the GPU (no copy). */
int rsg_ball_pred_view(rsg_ball_pred* p, rsg_ball_pred_view_t* out);

#ifdef __cplusplus
}
#endif

#endif /* RSG_H */
