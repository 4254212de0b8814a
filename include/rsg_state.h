/*
 * rsg_state.h -- plain-old-data state records exchanged across the C ABI.
 *
 * These records are the teacher-forcing surface of the batched tick step: one
 * record holds everything a car, the ball or a boost pad carries from one tick
 * to the next, so that a state written with rsg_batch_set_* and stepped once is
 * a full restatement of the reference tick on the same input.
 *
 * Field meanings follow the reference's public structs:
 *   CarState      src/Sim/Car/Car.h:17-120
 *   BallState     src/Sim/Ball/Ball.h:18-57
 *   BoostPadState src/Sim/BoostPad/BoostPad.h:23-31
 *   CarControls   src/Sim/CarControls.h:7-33
 *   BallHitInfo   src/Sim/BallHitInfo/BallHitInfo.h:9-27
 * plus the per-wheel vehicle values the reference keeps inside btVehicleRL
 * between ticks (src/Sim/btVehicleRL/btVehicleRL.h:9-27): engine force, brake,
 * steer angle, lateral/longitudinal friction and extra pushback.  The
 * reference's Car::SetState does not touch those, so a faithful one-tick
 * comparison has to carry them explicitly.
 *
 * Units are the reference's public units: positions in uu, velocities in uu/s,
 * angular velocity in rad/s.  Rotation is the column-major RotMat of the
 * reference (src/Math/MathTypes/MathTypes.h:169-170): rot[0..2] = forward,
 * rot[3..5] = right, rot[6..8] = up.
 */
#ifndef RSG_STATE_H
#define RSG_STATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Game mode ids, same order as src/Sim/GameMode.h:7-17. SOCCAR, HEATSEEKER
 * (the Soccar arena with the seeking ball) and THE_VOID are supported by the
 * batched step. */
enum { RSG_MODE_SOCCAR = 0, RSG_MODE_HOOPS = 1, RSG_MODE_HEATSEEKER = 2, RSG_MODE_THE_VOID = 5 };

/* Car config presets, same order as src/Sim/Car/CarConfig/CarConfig.cpp:20-100
 * (CAR_CONFIG_OCTANE ... CAR_CONFIG_PSYCLOPS).  RSG_PRESET_CUSTOM marks a car
 * whose config is none of them (rsg_car_config). */
enum {
    RSG_PRESET_OCTANE = 0, RSG_PRESET_DOMINUS = 1, RSG_PRESET_PLANK = 2,
    RSG_PRESET_BREAKOUT = 3, RSG_PRESET_HYBRID = 4, RSG_PRESET_MERC = 5,
    RSG_PRESET_PSYCLOPS = 6, RSG_PRESET_COUNT = 7, RSG_PRESET_CUSTOM = 255
};

/* CarConfig (src/Sim/Car/CarConfig/CarConfig.h:6-43), field for field, in uu:
 * the full hitbox size, its offset from the car's origin, the two wheel pairs
 * (connection point with y positive: the left wheel mirrors it), three-wheel
 * behaviour and the flip deadzone. */
typedef struct rsg_wheel_pair_config {
    float wheel_radius;
    float suspension_rest_length;
    float connection_point_offset[3];
} rsg_wheel_pair_config;

typedef struct rsg_car_config {
    float hitbox_size[3];
    float hitbox_pos_offset[3];
    rsg_wheel_pair_config front_wheels, back_wheels;
    uint8_t three_wheels;
    uint8_t _pad[3];
    float dodge_deadzone;
} rsg_car_config;

enum { RSG_TEAM_BLUE = 0, RSG_TEAM_ORANGE = 1 };

/* BoostPadConfig (src/Sim/BoostPad/BoostPad.h:13-19): a custom pad of
 * ArenaConfig::customBoostPads. */
typedef struct rsg_boost_pad_config {
    float pos[3];     /* uu */
    uint8_t is_big;
    uint8_t _pad[3];
} rsg_boost_pad_config;

typedef struct rsg_controls {
    float throttle, steer, pitch, yaw, roll;
    uint8_t jump, boost, handbrake, _pad;
} rsg_controls; /* 24 bytes */

/* Bit positions inside rsg_car_state.flags */
enum {
    RSG_CF_ON_GROUND = 1u << 0,
    RSG_CF_WHEEL0 = 1u << 1, RSG_CF_WHEEL1 = 1u << 2,
    RSG_CF_WHEEL2 = 1u << 3, RSG_CF_WHEEL3 = 1u << 4,
    RSG_CF_HAS_JUMPED = 1u << 5,
    RSG_CF_HAS_DOUBLE_JUMPED = 1u << 6,
    RSG_CF_HAS_FLIPPED = 1u << 7,
    RSG_CF_IS_FLIPPING = 1u << 8,
    RSG_CF_IS_JUMPING = 1u << 9,
    RSG_CF_IS_BOOSTING = 1u << 10,
    RSG_CF_SUPERSONIC = 1u << 11,
    RSG_CF_AUTO_FLIPPING = 1u << 12,
    RSG_CF_WORLD_CONTACT = 1u << 13,
    RSG_CF_DEMOED = 1u << 14,
    RSG_CF_HIT_VALID = 1u << 15,
    /* The car's rigid body has simulation disabled and no contact response
     * (btCollisionObject::CF_NO_CONTACT_RESPONSE + DISABLE_SIMULATION).  The
     * reference sets this in the pre-tick update of a demolished car and
     * clears it in that of a live one (src/Sim/Car/Car.cpp:68-84); it
     * persists on the body in between, so the wheel rays of cars updated
     * earlier in a tick see the value left by the previous tick. */
    RSG_CF_BODY_DISABLED = 1u << 16
};

typedef struct rsg_car_state {
    uint64_t tick_count_since_update;
    uint64_t hit_tick;              /* BallHitInfo.tickCountWhenHit */
    uint64_t hit_extra_tick;        /* BallHitInfo.tickCountWhenExtraImpulseApplied */
    float pos[3];
    float rot[9];
    float vel[3];
    float ang_vel[3];
    uint32_t flags;                 /* RSG_CF_* */
    float flip_rel_torque[3];
    float jump_time, flip_time;
    float air_time, air_time_since_jump;
    float boost, time_since_boosted, boosting_time;
    float supersonic_time, handbrake_val;
    float auto_flip_timer, auto_flip_torque_scale;
    float world_contact_normal[3];
    uint32_t car_contact_other_id;
    float car_contact_cooldown;
    float demo_respawn_timer;
    float hit_rel_pos[3], hit_ball_pos[3], hit_extra_vel[3];
    rsg_controls last_controls;
    /* per-wheel vehicle values carried between ticks (bullet units) */
    float wheel_engine_force, wheel_brake, wheel_steer;
    float wheel_lat_friction[4], wheel_long_friction[4];
    float wheel_extra_pushback[4];
    /* broadphase registration of the car's proxy (btRSBroadphase): the cell
     * index it is registered in and its registration order among the arena's
     * dynamic proxies; only the relative order of bp_seq values matters */
    int32_t bp_cell;
    uint32_t bp_seq;
} rsg_car_state;

typedef struct rsg_ball_state {
    uint64_t tick_count_since_update;
    float pos[3];
    float rot[9];
    float vel[3];
    float ang_vel[3];
    int32_t bp_cell;                /* as in rsg_car_state */
    uint32_t bp_seq;
    /* BallState::HeatseekerInfo (src/Sim/Ball/Ball.h:24-31): the net the
     * ball seeks (+1 the orange net, -1 the blue net, 0 none), its target
     * speed (uu/s) and the time since the last speed-up touch (s) */
    float hs_y_target_dir, hs_cur_target_speed, hs_time_since_hit;
} rsg_ball_state;

typedef struct rsg_pad_state {
    uint32_t is_active;
    float cooldown;
    uint32_t prev_locked_car_id;
} rsg_pad_state;

enum { RSG_NUM_PADS = 34, RSG_MAX_CARS = 8, RSG_SNAPSHOT_RAW_WORDS = 1024 };

/* Per-arena status bits (sticky until rsg_batch_clear_status).  The
 * reference grows its contact and solver arrays without bound and throws on
 * its own capacity errors (RS_ERR_CLOSE, src/Framework.h:67-72;
 * btRSBroadphase.cpp:206,229); the kernels hold fixed capacities per arena
 * and set a bit instead of dropping work silently:
 *   RSG_STATUS_NONFINITE         a car or the ball left a tick with a
 *                                non-finite position or velocity;
 *   RSG_STATUS_CONTACT_OVERFLOW  a tick needed more contact manifolds than
 *                                RSG_MAX_MANIFOLDS, or had more overlapping
 *                                pairs than its pair table holds (16 statics
 *                                per body on average plus every body pair;
 *                                the extra pairs were not collided);
 *   RSG_STATUS_SOLVER_OVERFLOW   a tick needed more than RSG_MAX_SOLVER_ROWS
 *                                contact or friction rows (the extra contacts
 *                                were not solved).
 * A tick that sets any of them has left the reference's results. */
enum { RSG_STATUS_NONFINITE = 1u, RSG_STATUS_CONTACT_OVERFLOW = 2u, RSG_STATUS_SOLVER_OVERFLOW = 4u };
enum { RSG_MAX_MANIFOLDS = 48, RSG_MAX_SOLVER_ROWS = 96 };

/* MutatorConfig (src/Sim/MutatorConfig/MutatorConfig.h:16-79), field for
 * field, in the reference's units (uu, uu/s, uu/s^2, seconds; masses in
 * Bullet units as the reference keeps them).  rsg_mutator_config_default
 * gives a mode's defaults (MutatorConfig(GameMode), MutatorConfig.cpp:5-40). */
enum { RSG_DEMO_NORMAL = 0, RSG_DEMO_ON_CONTACT = 1, RSG_DEMO_DISABLED = 2 };
typedef struct rsg_mutator_config {
    float gravity[3];
    float car_mass;
    float car_world_friction, car_world_restitution;
    float ball_mass, ball_max_speed, ball_drag;
    float ball_world_friction, ball_world_restitution;
    float jump_accel, jump_immediate_force;
    float boost_accel_ground, boost_accel_air, boost_used_per_second;
    float respawn_delay, bump_cooldown_time;
    float boost_pad_cooldown_big, boost_pad_cooldown_small;
    float car_spawn_boost_amount;
    float ball_hit_extra_force_scale, bump_force_scale;
    float ball_radius;
    uint8_t unlimited_flips, unlimited_double_jumps;
    uint8_t recharge_boost_enabled;
    uint8_t demo_mode;                 /* RSG_DEMO_* (DemoMode) */
    float recharge_boost_per_second, recharge_boost_delay;
    uint8_t enable_team_demos;
    uint8_t _pad[3];
    float goal_base_threshold_y;
} rsg_mutator_config;

/* A game event, replacing the reference's callbacks: the goal-score callback
 * (Arena::SetGoalScoreCallback, src/Sim/Arena/Arena.h:24,102; fired at the
 * end of the tick, Arena.cpp:748-752) and the car-bump callback
 * (Arena::SetCarBumpCallback, Arena.h:25,108; fired for bumps and
 * demolitions during collision handling, Arena.cpp:377-400). */
enum { RSG_EVENT_GOAL = 1, RSG_EVENT_BUMP = 2, RSG_EVENT_DEMO = 3 };
/* Game-event tracker events (GameEventTracker's shot, goal and save
 * callbacks, src/Sim/GameEventTracker/GameEventTracker.h:6-8): `a` is the
 * shooter / scorer / saver car id, `b` the passer's id or 0. */
enum { RSG_EVENT_SHOT = 4, RSG_EVENT_SCORER = 5, RSG_EVENT_SAVE = 6 };

/* GameEventTrackerConfig (GameEventTracker.h:10-41); the defaults are the
 * reference's (rsg_game_event_config_default). */
typedef struct rsg_game_event_config {
    float shot_min_speed;            /* uu/s toward the net, 1750 */
    float shot_touch_min_delay;      /* s since the shooter's touch, 0.3 */
    float pred_score_extra_margin;   /* uu, 0 */
    float shot_event_cooldown;       /* s between shots, 1 */
    float shot_min_score_time;       /* s, 2 */
    float goal_max_touch_time;       /* s, 4 */
    float pass_max_touch_time;       /* s, 2 */
    int32_t auto_state_set_detection;/* 1 */
} rsg_game_event_config;

/* Everything Arena::Serialize writes for one arena (src/Sim/Arena/Arena.cpp:
 * 546-573), plus the state records' hidden values: the unit of the
 * reference's snapshot format (rsg_snapshot_serialize / _deserialize). */
typedef struct rsg_arena_snapshot {
    int32_t mode;                       /* RSG_MODE_* (GameMode) */
    float tick_time;                    /* Arena::tickTime, seconds */
    uint64_t tick_count;                /* Arena::tickCount */
    uint32_t last_car_id;               /* Arena::_lastCarID */
    int32_t n_cars;
    /* cars in the arena's car update order (the reference's _cars iteration
     * order, which Serialize also follows) */
    uint32_t car_id[RSG_MAX_CARS];
    uint8_t car_team[RSG_MAX_CARS];
    uint8_t car_preset[RSG_MAX_CARS];  /* RSG_PRESET_*, RSG_PRESET_CUSTOM for other configs */
    uint8_t _pad[6];
    /* each car's CarConfig; an all-zero entry stands for car_preset's */
    rsg_car_config car_config[RSG_MAX_CARS];
    rsg_controls controls[RSG_MAX_CARS];   /* Car::controls */
    rsg_car_state cars[RSG_MAX_CARS];
    rsg_ball_state ball;
    rsg_pad_state pads[RSG_NUM_PADS];
    rsg_mutator_config mutators;       /* Arena::_mutatorConfig */
    /* This library's exact records of the arena, opaque: positions and
     * velocities in its internal units, the values the reference keeps
     * outside CarState / BallState (per-wheel vehicle values, contact flags,
     * tick counters, broadphase records) and the arena's counters.  Filled by
     * rsg_batch_get_snapshot and by reading bytes that carry the extension
     * block (raw_words = 0 otherwise).  When present, rsg_batch_set_snapshot
     * restores from them, so a snapshot resumes bit for bit. */
    uint32_t raw_layout;
    uint32_t raw_words;
    uint32_t raw[RSG_SNAPSHOT_RAW_WORDS];
} rsg_arena_snapshot;

/* Ball prediction records on the device (rsg_ball_pred_view): one row of
 * RSG_PRED_WORDS word planes per predicted tick, in the public units of
 * rsg_ball_state (BallState, src/Sim/Ball/Ball.h:18-57). */
enum {
    RSG_PRED_POS = 0,                 /* 3 floats, uu */
    RSG_PRED_ROT = 3,                 /* 9 floats, forward, right, up columns */
    RSG_PRED_VEL = 12,                /* 3 floats, uu/s */
    RSG_PRED_ANG_VEL = 15,            /* 3 floats, rad/s */
    RSG_PRED_TICKS_SINCE_UPDATE = 18, /* uint64 as two words, low first */
    RSG_PRED_HS = 20,                 /* 3 floats: Heatseeker target net, target speed, time since hit */
    RSG_PRED_WORDS = 23
};
typedef struct rsg_ball_pred_view_t {
    const float* traj;      /* [num_pred_ticks][RSG_PRED_WORDS][n]: word w of
                               tracker i's tick t is
                               traj[(row * RSG_PRED_WORDS + w) * n + i] with
                               row = (head[i] + t) % num_pred_ticks */
    const int32_t* head;    /* [n] first row of each tracker's ring */
    const int32_t* size;    /* [n] predicted ticks held (predData.size()) */
    int n;
    int num_pred_ticks;
} rsg_ball_pred_view_t;
typedef struct rsg_event {
    uint32_t arena;
    uint32_t type;      /* RSG_EVENT_* */
    uint64_t tick;      /* Arena::tickCount of the tick it happened in */
    uint32_t a;         /* GOAL: scoring team (RSG_TEAM_*); BUMP/DEMO: bumper car id */
    uint32_t b;         /* BUMP/DEMO: victim car id; GOAL: 0 */
    uint32_t seq;       /* per-arena sequence number (order of the callbacks) */
    uint32_t _pad;
} rsg_event;

#ifdef __cplusplus
}
#endif

#endif /* RSG_STATE_H */
