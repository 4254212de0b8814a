// rsg_layout.h -- struct-of-arrays layout of the batched arena state in HBM.
//
// Every entity is a fixed record of 32-bit words.  Word w of car slot s of
// arena a lives at  car[(w * max_cars + s) * n_arenas + a]:  consecutive arenas
// are consecutive words, so a wavefront that walks 64 arenas reads 256
// contiguous bytes per word and every access coalesces.  The ball, the 34 boost
// pads and the per-arena scalars use the same "word plane" layout.
//
// Internal units follow the reference's rigid bodies (Bullet metres, m/s):
// positions and velocities are converted to uu only at the API boundary,
// exactly where the reference's Car/Ball Get/SetState convert them
// (src/Sim/Car/Car.cpp:10-36, src/Sim/Ball/Ball.cpp:27-51).  Rotations are the
// Bullet row-major basis.
#pragma once

#include <stdint.h>

#include "rsg_state.h"

namespace rsg {

enum CarWord : int {
    CW_POS = 0,          // 3  (m)
    CW_ROT = 3,          // 9  basis rows
    CW_VEL = 12,         // 3  (m/s)
    CW_ANGVEL = 15,      // 3
    CW_FLAGS = 18,       // RSG_CF_*
    CW_FLIP_TORQUE = 19, // 3
    CW_JUMP_TIME = 22,
    CW_FLIP_TIME = 23,
    CW_AIR_TIME = 24,
    CW_AIR_TIME_SINCE_JUMP = 25,
    CW_BOOST = 26,
    CW_TIME_SINCE_BOOSTED = 27,
    CW_BOOSTING_TIME = 28,
    CW_SUPERSONIC_TIME = 29,
    CW_HANDBRAKE_VAL = 30,
    CW_AUTOFLIP_TIMER = 31,
    CW_AUTOFLIP_SCALE = 32,
    CW_WORLD_NORMAL = 33,  // 3
    CW_OTHER_CAR_ID = 36,
    CW_CAR_COOLDOWN = 37,
    CW_DEMO_TIMER = 38,
    CW_HIT_REL_POS = 39,   // 3
    CW_HIT_BALL_POS = 42,  // 3
    CW_HIT_EXTRA_VEL = 45, // 3
    CW_HIT_TICK = 48,      // 2 (u64 lo, hi)
    CW_HIT_EXTRA_TICK = 50,// 2
    CW_TICKS_SINCE_SET = 52, // 2
    CW_LAST_CTRL = 54,     // 5 floats + 1 bits word
    CW_ENGINE_FORCE = 60,
    CW_BRAKE = 61,
    CW_STEER = 62,
    CW_LAT_FRICTION = 63,  // 4
    CW_LONG_FRICTION = 67, // 4
    CW_PUSHBACK = 71,      // 4
    CW_CTRL = 75,          // 5 floats + 1 bits word
    CW_GEN_REDRAW = 81,
    CW_GEN_TICKS_LEFT = 82,
    CW_CONFIG = 83,        // config table index | team << 8
    CW_ID = 84,
    CW_BP_CELL = 85,       // broadphase cell the proxy is registered in
    CW_BP_SEQ = 86,        // registration order among the arena's dynamic proxies
    CW_DEMO_BODY_ROT = 87, // 9  rigid-body basis while demolished (CW_ROT then
                           //    holds the basis last reported, which the
                           //    reference freezes at demolition)
    // the rigid body's own mass properties and gravity, which the reference
    // sets when the body is made and changes only through
    // Arena::SetMutatorConfig (src/Sim/Arena/Arena.cpp:17-48); 0 in the mass
    // words means the car default
    CW_BODY_INV_MASS = 96,     // inverse mass
    CW_BODY_INERTIA_MASS = 97, // mass the box inertia is computed with
    CW_BODY_GRAVITY_ACC = 98,  // 3  btRigidBody::m_gravity_acceleration (m/s^2)
    CW_BODY_GRAVITY = 101,     // 3  btRigidBody::m_gravity, the force applied each tick
    CAR_WORDS = 104
};

enum BallWord : int {
    BW_POS = 0, BW_ROT = 3, BW_VEL = 12, BW_ANGVEL = 15, BW_TICKS_SINCE_SET = 18,
    BW_BP_CELL = 20, BW_BP_SEQ = 21,
    BW_HS = 22,          // 3  Heatseeker: target net, target speed, time since hit
    // the ball body's shape and mass as last made (Ball::_BulletSetup,
    // src/Sim/Ball/Ball.cpp:53-115), whether it is awake (not
    // ISLAND_SLEEPING) and its gravity, as for cars
    BW_RADIUS = 25,      // uu
    BW_MASS = 26,
    BW_AWAKE = 27,
    BW_GRAVITY_ACC = 28, // 3
    BW_GRAVITY = 31,     // 3
    BALL_WORDS = 34
};

enum PadWord : int { PW_ACTIVE = 0, PW_COOLDOWN = 1, PW_PREV_LOCKED = 2, PAD_WORDS = 3 };

enum ArenaWord : int {
    AW_TICK = 0,        // 2 (u64)
    AW_NUM_CARS = 2,
    AW_GOALS_BLUE = 3,  // goals scored into the orange net (blue team scores)
    AW_GOALS_ORANGE = 4,
    AW_LAST_GOAL_TICK = 5, // 2
    AW_RESETS = 7,
    AW_STATUS = 8,      // RSG_STATUS_* bits (sticky)
    AW_CAR_ORDER = 9,   // car update order: slot k = car index in bits 4k..4k+3
    AW_EVENT_SEQ = 10,  // number of events the arena has emitted
    ARENA_WORDS = 11
};

// Controls bit word
enum { CTRL_JUMP = 1u, CTRL_BOOST = 2u, CTRL_HANDBRAKE = 4u };

struct BatchDev {
    uint32_t* car;     // CAR_WORDS * max_cars * n
    uint32_t* ball;    // BALL_WORDS * n
    uint32_t* pad;     // PAD_WORDS * 34 * n
    uint32_t* arena;   // ARENA_WORDS * n
    int n;             // arenas
    int max_cars;      // car slots per arena
    int mode;          // RSG_MODE_*
    int auto_reset;    // kickoff reset after a goal
    int synth_controls;// draw controls from the seeded process each tick
    int ball_rot;      // the ball integrates its rotation (ArenaConfig::noBallRot == false)
    const rsg_boost_pad_config* custom_pads;  // ArenaConfig::customBoostPads, or null
    int n_custom_pads;
    uint64_t seed;
    float tick_time;
    // event ring (null when events are off): append index from *ev_count
    rsg_event* ev_buf;
    uint32_t* ev_count;
    uint32_t ev_cap;
    // per-arena world-step work space (ArenaSpace blocks, one per arena) for
    // the kernels that keep it in global memory; null when not allocated
    unsigned char* space;
    // per-arena MutatorConfig (n records); null: the mode's defaults
    const rsg_mutator_config* mut;
    // car config table: CW_CONFIG's low byte indexes it; null: the presets
    const rsg_car_config* cfgs;
};

}  // namespace rsg
