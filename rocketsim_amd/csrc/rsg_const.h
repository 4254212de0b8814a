// rsg_const.h -- game constants, curves, presets and spawn/pad tables of the
// Soccar tick step, restated as data.
//
// Values follow the reference exactly (they are facts of the game, not code):
//   src/RLConst.h:10-527            game constants, curves, pad and spawn tables
//   src/Sim/Car/CarConfig/CarConfig.cpp:20-100   hitbox / wheel presets
//   src/Sim/MutatorConfig/MutatorConfig.{h,cpp}  Soccar mutator defaults
//   src/BulletLink.h:12,15                       uu <-> Bullet metres (x50)
// Expressions are written in the same types as the reference so that constants
// computed in double and narrowed to float round identically.
#pragma once

#include "rsg_math.h"
#include "../../include/rsg_state.h"

namespace rsg {

constexpr float BT_TO_UU = 50.f;
constexpr float UU_TO_BT = 1.f / 50.f;

constexpr double PI_D = 3.14159265358979323846;
constexpr float GRAVITY_Z = -650.f;
constexpr float ARENA_EXTENT_X = 4096, ARENA_EXTENT_Y = 5120, ARENA_HEIGHT = 2048;
constexpr float ARENA_EXTENT_X_HOOPS = 8900 / 3.f, ARENA_EXTENT_Y_HOOPS = 3581, ARENA_HEIGHT_HOOPS = 1820;
// Hoops (src/RLConst.h:50-51,118-139): the ball's kickoff launch, radius,
// goal height and the stronger upward extra hit impulse of cars on the ground.
constexpr float BALL_HOOPS_LAUNCH_Z_VEL = 1000, BALL_HOOPS_LAUNCH_DELAY = 0.265f;
constexpr float BALL_COLLISION_RADIUS_HOOPS = 96.3831f;
constexpr float HOOPS_GOAL_SCORE_THRESHOLD_Z = 270.f;
constexpr float ARENA_COLLISION_BASE_FRICTION = 0.6f, ARENA_COLLISION_BASE_RESTITUTION = 0.3f;
constexpr float CAR_MASS_BT = 180.f;
constexpr float BALL_MASS_BT = CAR_MASS_BT / 6.f;
constexpr float CAR_COLLISION_FRICTION = 0.3f, CAR_COLLISION_RESTITUTION = 0.1f;
constexpr float CARBALL_COLLISION_FRICTION = 2.0f, CARBALL_COLLISION_RESTITUTION = 0.0f;
constexpr float CARWORLD_COLLISION_FRICTION = 0.3f, CARWORLD_COLLISION_RESTITUTION = 0.3f;
constexpr float CARCAR_COLLISION_FRICTION = 0.09f, CARCAR_COLLISION_RESTITUTION = 0.1f;
constexpr float BALL_REST_Z = 93.15f;
constexpr float BALL_MAX_ANG_SPEED = 6.f;
constexpr float BALL_DRAG = 0.03f;
constexpr float BALL_FRICTION = 0.35f;
constexpr float BALL_RESTITUTION = 0.6f;
constexpr float CAR_MAX_SPEED = 2300.f;
constexpr float BALL_MAX_SPEED = 6000.f;
constexpr float BOOST_MAX = 100.f;
constexpr float BOOST_USED_PER_SECOND = BOOST_MAX / 3;
constexpr float BOOST_MIN_TIME = 0.1f;
constexpr float BOOST_ACCEL_GROUND = 2975 / 3.f;
constexpr float BOOST_ACCEL_AIR = 3175 / 3.f;
constexpr float BOOST_SPAWN_AMOUNT = BOOST_MAX / 3;
constexpr float RECHARGE_BOOST_PER_SECOND = 10;
constexpr float RECHARGE_BOOST_DELAY = 0.25;
constexpr float CAR_MAX_ANG_SPEED = 5.5f;
constexpr float SUPERSONIC_START_SPEED = 2200.f;
constexpr float SUPERSONIC_MAINTAIN_MIN_SPEED = SUPERSONIC_START_SPEED - 100.f;
constexpr float SUPERSONIC_MAINTAIN_MAX_TIME = 1.f;
constexpr float POWERSLIDE_RISE_RATE = 5, POWERSLIDE_FALL_RATE = 2;
constexpr float THROTTLE_TORQUE_AMOUNT = CAR_MASS_BT * 400.f;
constexpr float BRAKE_TORQUE_AMOUNT = CAR_MASS_BT * (14.25f + (1.f / 3.f));
constexpr float STOPPING_FORWARD_VEL = 25.f;
constexpr float COASTING_BRAKE_FACTOR = 0.15f;
constexpr float BRAKING_NO_THROTTLE_SPEED_THRESH = 0.01f;
constexpr float THROTTLE_DEADZONE = 0.001f;
constexpr float THROTTLE_AIR_ACCEL = 200 / 3.f;
constexpr float JUMP_ACCEL = 4375.f / 3.f;
constexpr float JUMP_IMMEDIATE_FORCE = 875.f / 3.f;
constexpr float JUMP_MIN_TIME = 0.025f;
constexpr float JUMP_RESET_TIME_PAD = (1 / 40.f);
constexpr float JUMP_MAX_TIME = 0.2f;
constexpr float JUMP_PRE_MIN_ACCEL_SCALE = 0.62f;  // src/Sim/Car/Car.cpp:566
constexpr float DOUBLEJUMP_MAX_DELAY = 1.25f;
constexpr float FLIP_Z_DAMP_120 = 0.35f;
constexpr float FLIP_Z_DAMP_START = 0.15f;
constexpr float FLIP_Z_DAMP_END = 0.21f;
constexpr float FLIP_TORQUE_TIME = 0.65f;
constexpr float FLIP_PITCHLOCK_EXTRA_TIME = 0.3f;
constexpr float FLIP_INITIAL_VEL_SCALE = 500.f;
constexpr float FLIP_TORQUE_X = 260.f, FLIP_TORQUE_Y = 224.f;
constexpr float FLIP_FORWARD_IMPULSE_MAX_SPEED_SCALE = 1.f;
constexpr float FLIP_SIDE_IMPULSE_MAX_SPEED_SCALE = 1.9f;
constexpr float FLIP_BACKWARD_IMPULSE_MAX_SPEED_SCALE = 2.5f;
constexpr float FLIP_BACKWARD_IMPULSE_SCALE_X = 16.f / 15.f;
constexpr float BALL_COLLISION_RADIUS_SOCCAR = 91.25f;
constexpr float SOCCAR_GOAL_SCORE_BASE_THRESHOLD_Y = 5124.25f;
constexpr float CAR_TORQUE_SCALE = 2 * PI_D / (1 << 16) * 1000;
constexpr float CAR_AUTOFLIP_IMPULSE = 200;
constexpr float CAR_AUTOFLIP_TORQUE = 50;
constexpr float CAR_AUTOFLIP_TIME = 0.4f;
constexpr float CAR_AUTOFLIP_NORMZ_THRESH = 0.70710678118654752440;
constexpr float CAR_AUTOFLIP_ROLL_THRESH = 2.8f;
constexpr float CAR_AUTOROLL_FORCE = 100;
constexpr float CAR_AUTOROLL_TORQUE = 80;
constexpr float BALL_CAR_EXTRA_IMPULSE_Z_SCALE = 0.35f;
constexpr float BALL_CAR_EXTRA_IMPULSE_Z_SCALE_HOOPS_GROUND = BALL_CAR_EXTRA_IMPULSE_Z_SCALE * 1.55f;
constexpr float BALL_CAR_EXTRA_IMPULSE_Z_SCALE_HOOPS_NORMAL_Z_THRESH = 0.1f;
constexpr float BALL_CAR_EXTRA_IMPULSE_FORWARD_SCALE = 0.65f;
constexpr float BALL_CAR_EXTRA_IMPULSE_MAXDELTAVEL_UU = 4600.f;
constexpr float CAR_SPAWN_REST_Z = 17.f;
constexpr float CAR_RESPAWN_Z = 36.f;
constexpr float BUMP_COOLDOWN_TIME = 0.25f;
constexpr float BUMP_MIN_FORWARD_DIST = 64.5f;
constexpr float DEMO_RESPAWN_TIME = 3.f;

// CAR_AIR_CONTROL_TORQUE / _DAMPING, order pitch, yaw, roll.
constexpr float AIR_TORQUE_P = 130, AIR_TORQUE_Y = 95, AIR_TORQUE_R = 400;
constexpr float AIR_DAMP_P = 30, AIR_DAMP_Y = 20, AIR_DAMP_R = 50;

// btVehicleRL tuning (src/RLConst.h BTVehicle namespace)
constexpr float SUSPENSION_FORCE_SCALE_FRONT = 36.f - (1.f / 4.f);
constexpr float SUSPENSION_FORCE_SCALE_BACK = 54.f + (1.f / 4.f) + (1.5f / 100.f);
constexpr float SUSPENSION_STIFFNESS = 500.f;
constexpr float WHEELS_DAMPING_COMPRESSION = 25.f;
constexpr float WHEELS_DAMPING_RELAXATION = 40.f;
constexpr float MAX_SUSPENSION_TRAVEL = 12.f;
constexpr float SUSPENSION_SUBTRACTION = 0.05f;
constexpr float ROLLING_FRICTION_SCALE_MAGIC = 113.73963f;  // src/Sim/btVehicleRL/btVehicleRL.cpp:359

// Boost pads (src/RLConst.h BoostPads namespace)
constexpr float PAD_CYL_HEIGHT = 95, PAD_CYL_RAD_BIG = 208, PAD_CYL_RAD_SMALL = 144;
constexpr float PAD_BOX_HEIGHT = 64, PAD_BOX_RAD_BIG = 160, PAD_BOX_RAD_SMALL = 120;
constexpr float PAD_COOLDOWN_BIG = 10, PAD_COOLDOWN_SMALL = 4;
constexpr float PAD_BOOST_AMOUNT_BIG = 100, PAD_BOOST_AMOUNT_SMALL = 12;
constexpr int PAD_NUM_BIG = 6, PAD_NUM_SMALL = 28, PAD_NUM = 34;
constexpr int PAD_NUM_SMALL_HOOPS = 14;  // Hoops: 6 big + 14 small pads in the first 20 slots

// Boost pad grid (src/Sim/BoostPad/BoostPadGrid/BoostPadGrid.h)
constexpr float PAD_GRID_EXTENT_Z = PAD_CYL_HEIGHT + 250.f;
constexpr int PAD_GRID_CELLS_X = 8, PAD_GRID_CELLS_Y = 10;
constexpr int PAD_GRID_CELL_SIZE_X = (int)(4096.f / (8 / 2));
constexpr int PAD_GRID_CELL_SIZE_Y = (int)(5120.f / (10 / 2));

// Bullet solver / world settings used by the reference arena
// (src/Sim/Arena/Arena.cpp:473-475, BulletDynamics/ConstraintSolver/btContactSolverInfo.h:77-113)
constexpr int SOLVER_ITERATIONS = 10;
constexpr float SOLVER_ERP2 = 0.8f;
constexpr float SOLVER_SPLIT_TURN_ERP = 0.1f;
constexpr float SOLVER_RESTITUTION_VEL_THRESHOLD = 0.2f;
constexpr float FLT_EPS = 1.1920928955078125e-07f;
constexpr float ANGULAR_MOTION_THRESHOLD = 0.5f * 1.57079632679489661923f;
constexpr float CONTACT_BREAKING_THRESHOLD = 0.02f;  // gContactBreakingThreshold
constexpr float CONVEX_DISTANCE_MARGIN = 0.04f;

// Piecewise-linear curves (LinearPieceCurve, src/Math/Math.cpp:5-35).
struct Curve {
    int n;
    float x[6];
    float y[6];
};

RSG_FN float curve_eval(const Curve& c, float input, float default_out = 1.f) {
    if (c.n == 0) return default_out;
    if (input <= c.x[0]) return c.y[0];
    for (int i = 1; i < c.n; i++) {
        if (c.x[i] > input) {
            float range = c.x[i] - c.x[i - 1];
            float diff = c.y[i] - c.y[i - 1];
            float t = (input - c.x[i - 1]) / range;
            return c.y[i - 1] + diff * t;
        }
    }
    return c.y[c.n - 1];
}

#if defined(__HIPCC__)
#define RSG_CONST_TABLE __constant__ const
#else
#define RSG_CONST_TABLE static const
#endif

// Curves are tiny; keep them as inline constexpr data usable on host and device.
RSG_FN Curve curve_steer_angle() { return Curve{6, {0, 500, 1000, 1500, 1750, 3000}, {0.53356f, 0.31930f, 0.18203f, 0.10570f, 0.08507f, 0.03454f}}; }
// Three-wheel cars (CarConfig::threeWheels, src/RLConst.h:429-434, 470-475)
RSG_FN Curve curve_steer_angle_three_wheels() { return Curve{2, {0, 2300}, {0.342473f, 0.034837f}}; }
RSG_FN Curve curve_lat_friction_three_wheels() { return Curve{2, {0, 1}, {0.30f, 0.25f}}; }
RSG_FN Curve curve_powerslide_steer() { return Curve{2, {0, 2500}, {0.39235f, 0.12610f}}; }
RSG_FN Curve curve_drive_torque() { return Curve{3, {0, 1400, 1410}, {1.0f, 0.1f, 0.0f}}; }
RSG_FN Curve curve_non_sticky() { return Curve{3, {0, 0.7075f, 1}, {0.1f, 0.5f, 1.0f}}; }
RSG_FN Curve curve_lat_friction() { return Curve{2, {0, 1}, {1.0f, 0.2f}}; }
RSG_FN Curve curve_long_friction() { return Curve{0, {0}, {0}}; }
RSG_FN Curve curve_handbrake_lat() { return Curve{1, {0}, {0.1f}}; }
RSG_FN Curve curve_handbrake_long() { return Curve{2, {0, 1}, {0.5f, 0.9f}}; }
RSG_FN Curve curve_ball_extra_impulse() { return Curve{4, {0, 500.f, 2300.f, 4600.f}, {0.65f, 0.65f, 0.55f, 0.30f}}; }
RSG_FN Curve curve_bump_ground() { return Curve{3, {0.f, 1400.f, 2200.f}, {(5.f / 6.f), 1100.f, 1530.f}}; }
RSG_FN Curve curve_bump_air() { return Curve{3, {0.f, 1400.f, 2200.f}, {(5.f / 6.f), 1390.f, 1945.f}}; }
RSG_FN Curve curve_bump_upward() { return Curve{3, {0.f, 1400.f, 2200.f}, {(2.f / 6.f), 278.f, 417.f}}; }

// Car config presets (src/Sim/Car/CarConfig/CarConfig.cpp:20-100, the
// reference's tables in its order; the Psyclops hitbox adds 0.134 to the
// Octane's in float as the reference does).
RSG_FN rsg_wheel_pair_config wheel_pair(float radius, float rest, float x, float y, float z) {
    rsg_wheel_pair_config w;
    w.wheel_radius = radius;
    w.suspension_rest_length = rest;
    w.connection_point_offset[0] = x;
    w.connection_point_offset[1] = y;
    w.connection_point_offset[2] = z;
    return w;
}

RSG_FN rsg_car_config car_config(float sx, float sy, float sz, float ox, float oy, float oz,
                                 rsg_wheel_pair_config front, rsg_wheel_pair_config back, bool three_wheels) {
    rsg_car_config c{};
    c.hitbox_size[0] = sx; c.hitbox_size[1] = sy; c.hitbox_size[2] = sz;
    c.hitbox_pos_offset[0] = ox; c.hitbox_pos_offset[1] = oy; c.hitbox_pos_offset[2] = oz;
    c.front_wheels = front;
    c.back_wheels = back;
    c.three_wheels = three_wheels ? 1 : 0;
    c._pad[0] = c._pad[1] = c._pad[2] = 0;
    c.dodge_deadzone = 0.5f;  // CarConfig.h:37
    return c;
}

RSG_FN rsg_car_config preset_config(int p) {
    switch (p) {
        default:
        case 0: return car_config(120.507f, 86.6994f, 38.6591f, 13.8757f, 0, 20.755f,
                                  wheel_pair(12.50f, 38.755f, 51.25f, 25.90f, 20.755f),
                                  wheel_pair(15.00f, 37.055f, -33.75f, 29.50f, 20.755f), false);
        case 1: return car_config(130.427f, 85.7799f, 33.8f, 9.f, 0, 15.75f,
                                  wheel_pair(12.00f, 33.95f, 50.30f, 31.10f, 15.75f),
                                  wheel_pair(13.50f, 33.85f, -34.75f, 33.00f, 15.75f), false);
        case 2: return car_config(131.32f, 87.1704f, 31.8944f, 9.00857f, 0, 12.0942f,
                                  wheel_pair(12.50f, 31.9242f, 49.97f, 27.80f, 12.0942f),
                                  wheel_pair(17.00f, 27.9242f, -35.43f, 20.28f, 12.0942f), false);
        case 3: return car_config(133.992f, 83.021f, 32.8f, 12.5f, 0, 11.75f,
                                  wheel_pair(13.50f, 29.7f, 51.50f, 26.67f, 11.75f),
                                  wheel_pair(15.00f, 29.666f, -35.75f, 35.00f, 11.75f), false);
        case 4: return car_config(129.519f, 84.6879f, 36.6591f, 13.8757f, 0, 20.755f,
                                  wheel_pair(12.50f, 38.755f, 51.25f, 25.90f, 20.755f),
                                  wheel_pair(15.00f, 37.055f, -34.00f, 29.50f, 20.755f), false);
        case 5: return car_config(123.22f, 79.2103f, 44.1591f, 11.3757f, 0, 21.505f,
                                  wheel_pair(15.00f, 39.505f, 51.25f, 25.90f, 21.505f),
                                  wheel_pair(15.00f, 39.105f, -33.75f, 29.50f, 21.505f), false);
        case 6: return car_config(120.507f + 0.134f, 86.6994f + 0.134f, 38.6591f + 0.134f, 13.8757f, 0, 15.0f,
                                  wheel_pair(12.50f, 33.0f, 51.25f, 5.000f, 15.000f),
                                  wheel_pair(15.00f, 31.3f, -33.75f, 29.50f, 15.000f), true);
    }
}

// Derived per-preset physics values in Bullet units, computed the way the
// reference's setup does (src/Sim/Car/Car.cpp:195-292,
// BulletCollision/CollisionShapes/btBoxShape.cpp:17-45,
// btConvexInternalShape.h:63-78, btCompoundShape.cpp:48-82,159-179).
struct CarPhys {
    V3 box_implicit;      // half extents minus the default margin
    float box_margin;     // safe margin (<= 0.04)
    V3 box_offset;        // child offset (hitboxPosOffset in m)
    V3 local_aabb_half;   // compound local AABB half extents
    V3 local_aabb_center; // compound local AABB centre
    V3 inv_inertia_local;
    float inv_mass;
    // wheels: index 0,1 front (1 = left), 2,3 back (3 = left)
    V3 wheel_conn[4];
    float wheel_rest[4];  // suspension rest length (m)
    float wheel_radius[4];
    float wheel_force_scale[4];
    float max_susp_travel; // m
    float angular_motion_disc;
    bool three_wheels;     // CarConfig::threeWheels
    float dodge_deadzone;  // CarConfig::dodgeDeadzone
};

// btBoxShape::calculateLocalInertia (BulletCollision/CollisionShapes/
// btBoxShape.cpp:30-45) of a box with half extents `he` (margin included)
// for `mass`, inverted as btRigidBody::setMassProps does.
RSG_FN V3 box_inv_inertia(V3 he, float mass) {
    float lx = 2.f * he.x, ly = 2.f * he.y, lz = 2.f * he.z;
    V3 inertia = v3(mass / (12.0f) * (ly * ly + lz * lz),
                    mass / (12.0f) * (lx * lx + lz * lz),
                    mass / (12.0f) * (lx * lx + ly * ly));
    return v3(inertia.x != 0.f ? 1.f / inertia.x : 0.f, inertia.y != 0.f ? 1.f / inertia.y : 0.f,
              inertia.z != 0.f ? 1.f / inertia.z : 0.f);
}

RSG_FN CarPhys make_car_phys(const rsg_car_config& r) {
    CarPhys c;
    c.three_wheels = r.three_wheels != 0;
    c.dodge_deadzone = r.dodge_deadzone;
    V3 size = v3(r.hitbox_size[0], r.hitbox_size[1], r.hitbox_size[2]);
    V3 half = (size * UU_TO_BT) / 2.f;
    c.box_implicit = half - v3(CONVEX_DISTANCE_MARGIN, CONVEX_DISTANCE_MARGIN, CONVEX_DISTANCE_MARGIN);
    float mind = half.x;
    int ax = 0;
    // btVector3::minAxis: x<y ? (x<z ? 0 : 2) : (y<z ? 1 : 2)
    ax = half.x < half.y ? (half.x < half.z ? 0 : 2) : (half.y < half.z ? 1 : 2);
    mind = get(half, ax);
    float safe = 0.1f * mind;
    c.box_margin = safe < CONVEX_DISTANCE_MARGIN ? safe : CONVEX_DISTANCE_MARGIN;
    c.box_offset = v3(r.hitbox_pos_offset[0], r.hitbox_pos_offset[1], r.hitbox_pos_offset[2]) * UU_TO_BT;
    V3 he = c.box_implicit + v3(c.box_margin, c.box_margin, c.box_margin);
    V3 lmin = c.box_offset - he, lmax = c.box_offset + he;
    c.local_aabb_half = 0.5f * (lmax - lmin);
    c.local_aabb_center = 0.5f * (lmax + lmin);
    c.inv_inertia_local = box_inv_inertia(he, CAR_MASS_BT);
    c.inv_mass = 1.f / CAR_MASS_BT;
    for (int i = 0; i < 4; i++) {
        bool front = i < 2;
        bool left = i % 2;
        const rsg_wheel_pair_config& wp = front ? r.front_wheels : r.back_wheels;
        V3 conn = v3(wp.connection_point_offset[0], wp.connection_point_offset[1], wp.connection_point_offset[2]);
        if (left) conn.y *= -1;
        c.wheel_conn[i] = conn * UU_TO_BT;
        float rest = wp.suspension_rest_length;
        rest -= MAX_SUSPENSION_TRAVEL;
        c.wheel_rest[i] = rest * UU_TO_BT;
        c.wheel_radius[i] = wp.wheel_radius * UU_TO_BT;
        c.wheel_force_scale[i] = front ? SUSPENSION_FORCE_SCALE_FRONT : SUSPENSION_FORCE_SCALE_BACK;
    }
    c.max_susp_travel = ((MAX_SUSPENSION_TRAVEL * UU_TO_BT) * 100) / 100;
    // Bounding-sphere radius of the compound plus its centre offset
    // (btCollisionShape::getAngularMotionDisc, btCollisionShape.cpp:151-157):
    // computed from the compound's local AABB of the child box.
    {
        // btCollisionShape::getBoundingSphere on the identity-transformed
        // compound AABB (compound margin is 0).
        V3 amin = c.local_aabb_center - c.local_aabb_half;
        V3 amax = c.local_aabb_center + c.local_aabb_half;
        float radius = len(amax - amin) * 0.5f;
        V3 center = (amin + amax) * 0.5f;
        c.angular_motion_disc = radius + len(center);
    }
    return c;
}

RSG_FN CarPhys make_car_phys(int preset) { return make_car_phys(preset_config(preset)); }

// Ball (src/Sim/Ball/Ball.cpp:53-115, btSphereShape.cpp:37-59)
struct BallPhys {
    float radius;      // m
    float inv_mass;
    float inv_inertia;
    float angular_motion_disc;
};

RSG_FN BallPhys make_ball_phys(float radius_uu = BALL_COLLISION_RADIUS_SOCCAR, float mass = BALL_MASS_BT) {
    BallPhys b;
    b.radius = radius_uu * UU_TO_BT;
    b.inv_mass = 1.f / mass;
    float elem = 0.4f * mass * b.radius * b.radius;
    b.inv_inertia = 1.f / elem;
    // RocketSim adds 0.08 to the sphere's bounding radius
    // (BulletCollision/CollisionShapes/btCollisionShape.cpp:130-134).
    b.angular_motion_disc = (float)((double)b.radius + 0.08);
    return b;
}

// Car spawn tables (src/RLConst.h:355-427), blue side; orange is mirrored.
struct SpawnPos { float x, y, yaw; };
RSG_FN SpawnPos kickoff_spawn_hoops(int i) {
    switch (i) {
        default:
        case 0: return SpawnPos{-1536, -3072, (float)(0.78539816339744830962 * 2)};
        case 1: return SpawnPos{1536, -3072, (float)(0.78539816339744830962 * 2)};
        case 2: return SpawnPos{-256, -2816, (float)(0.78539816339744830962 * 2)};
        case 3: return SpawnPos{256, -2816, (float)(0.78539816339744830962 * 2)};
        case 4: return SpawnPos{0, -3200, (float)(0.78539816339744830962 * 2)};
    }
}
RSG_FN SpawnPos respawn_spawn_hoops(int i) {
    switch (i) {
        default:
        case 0: return SpawnPos{-1920, -3072, (float)(PI_D / 2)};
        case 1: return SpawnPos{-1152, -3072, (float)(PI_D / 2)};
        case 2: return SpawnPos{1920, -3072, (float)(PI_D / 2)};
        case 3: return SpawnPos{1152, -3072, (float)(PI_D / 2)};
    }
}
RSG_FN SpawnPos kickoff_spawn(int i) {
    switch (i) {
        default:
        case 0: return SpawnPos{-2048, -2560, (float)(0.78539816339744830962 * 1)};
        case 1: return SpawnPos{2048, -2560, (float)(0.78539816339744830962 * 3)};
        case 2: return SpawnPos{-256, -3840, (float)(0.78539816339744830962 * 2)};
        case 3: return SpawnPos{256, -3840, (float)(0.78539816339744830962 * 2)};
        case 4: return SpawnPos{0, -4608, (float)(0.78539816339744830962 * 2)};
    }
}
RSG_FN SpawnPos respawn_spawn(int i) {
    switch (i) {
        default:
        case 0: return SpawnPos{-2304, -4608, (float)(PI_D / 2)};
        case 1: return SpawnPos{-2688, -4608, (float)(PI_D / 2)};
        case 2: return SpawnPos{2304, -4608, (float)(PI_D / 2)};
        case 3: return SpawnPos{2688, -4608, (float)(PI_D / 2)};
    }
}
// Heatseeker kickoff spawns (src/RLConst.h:365-371): 4 slots, all facing +y.
RSG_FN SpawnPos kickoff_spawn_heatseeker(int i) {
    switch (i) {
        default:
        case 0: return SpawnPos{-1000, -4620, (float)(PI_D / 2)};
        case 1: return SpawnPos{1000, -4620, (float)(PI_D / 2)};
        case 2: return SpawnPos{-2000, -4620, (float)(PI_D / 2)};
        case 3: return SpawnPos{2000, -4620, (float)(PI_D / 2)};
    }
}

// Modes played in an arena (meshes, planes, boost pads and goals,
// Arena.cpp:478-520, 893-915): Soccar and Heatseeker in the Soccar arena,
// Hoops in its own.
RSG_FN bool mode_has_arena(int mode) {
    return mode == RSG_MODE_SOCCAR || mode == RSG_MODE_HEATSEEKER || mode == RSG_MODE_HOOPS;
}
// The arena a mode's meshes belong to (0: Soccar, 1: Hoops).
RSG_FN int arena_kind(int mode) { return mode == RSG_MODE_HOOPS ? 1 : 0; }
// Boost pads of the mode's arena (Arena.cpp:509-531).
RSG_FN int pad_count(int mode) { return mode == RSG_MODE_HOOPS ? PAD_NUM_BIG + PAD_NUM_SMALL_HOOPS : PAD_NUM; }
RSG_FN SpawnPos kickoff_spawn_of(int mode, int i) {
    if (mode == RSG_MODE_HOOPS) return kickoff_spawn_hoops(i);
    if (mode == RSG_MODE_HEATSEEKER) return kickoff_spawn_heatseeker(i);
    return kickoff_spawn(i);
}
// Car::Respawn's table (Car.cpp:49); Heatseeker respawns at the Soccar slots.
RSG_FN SpawnPos respawn_spawn_of(int mode, int i) {
    return mode == RSG_MODE_HOOPS ? respawn_spawn_hoops(i) : respawn_spawn(i);
}
// BallWithinHoopsGoalXYMarginSq (Arena.cpp:759-768): negative inside a
// hoop's footprint (squared margin, uu).
RSG_FN float hoops_goal_xy_margin_sq(float x, float y) {
    const float SCALE_Y = 0.9f, OFFSET_Y = 2770.f, RADIUS_SQ = 716 * 716;
    float dy = fabsf(y) * SCALE_Y - OFFSET_Y;
    float dist_sq = x * x + dy * dy;
    return dist_sq - RADIUS_SQ;
}

// MutatorConfig(GameMode) (src/Sim/MutatorConfig/MutatorConfig.h:16-79,
// MutatorConfig.cpp:5-40): the field defaults, then the mode's changes
// (Heatseeker spawns cars with full boost and never drains it).
RSG_FN rsg_mutator_config mutator_defaults(int mode) {
    rsg_mutator_config m;
    m.gravity[0] = 0.f; m.gravity[1] = 0.f; m.gravity[2] = GRAVITY_Z;
    m.car_mass = CAR_MASS_BT;
    m.car_world_friction = CARWORLD_COLLISION_FRICTION;
    m.car_world_restitution = CARWORLD_COLLISION_RESTITUTION;
    m.ball_mass = BALL_MASS_BT;
    m.ball_max_speed = BALL_MAX_SPEED;
    m.ball_drag = BALL_DRAG;
    m.ball_world_friction = BALL_FRICTION;
    m.ball_world_restitution = BALL_RESTITUTION;
    m.jump_accel = JUMP_ACCEL;
    m.jump_immediate_force = JUMP_IMMEDIATE_FORCE;
    m.boost_accel_ground = BOOST_ACCEL_GROUND;
    m.boost_accel_air = BOOST_ACCEL_AIR;
    m.boost_used_per_second = BOOST_USED_PER_SECOND;
    m.respawn_delay = DEMO_RESPAWN_TIME;
    m.bump_cooldown_time = BUMP_COOLDOWN_TIME;
    m.boost_pad_cooldown_big = PAD_COOLDOWN_BIG;
    m.boost_pad_cooldown_small = PAD_COOLDOWN_SMALL;
    m.car_spawn_boost_amount = BOOST_SPAWN_AMOUNT;
    m.ball_hit_extra_force_scale = 1.f;
    m.bump_force_scale = 1.f;
    m.ball_radius = BALL_COLLISION_RADIUS_SOCCAR;
    m.unlimited_flips = 0;
    m.unlimited_double_jumps = 0;
    m.recharge_boost_enabled = 0;
    m.demo_mode = RSG_DEMO_NORMAL;
    m.recharge_boost_per_second = RECHARGE_BOOST_PER_SECOND;
    m.recharge_boost_delay = RECHARGE_BOOST_DELAY;
    m.enable_team_demos = 0;
    m._pad[0] = m._pad[1] = m._pad[2] = 0;
    m.goal_base_threshold_y = SOCCAR_GOAL_SCORE_BASE_THRESHOLD_Y;
    if (mode == RSG_MODE_HOOPS) m.ball_radius = BALL_COLLISION_RADIUS_HOOPS;
    if (mode == RSG_MODE_HEATSEEKER) {
        m.car_spawn_boost_amount = BOOST_MAX;
        m.boost_used_per_second = 0.f;
    }
    return m;
}

// Heatseeker ball (src/RLConst.h:168-188).
constexpr float HS_INITIAL_TARGET_SPEED = 2900, HS_TARGET_SPEED_INCREMENT = 85, HS_MIN_SPEEDUP_INTERVAL = 1,
                HS_TARGET_Y = 5120, HS_TARGET_Z = 320, HS_HORIZONTAL_BLEND = 1.45f, HS_VERTICAL_BLEND = 0.78f,
                HS_SPEED_BLEND = 0.3f, HS_MAX_TURN_PITCH = (float)(7000 * PI_D / (1 << 15)), HS_MAX_SPEED = 4600,
                HS_WALL_BOUNCE_CHANGE_Y_THRESH = 300, HS_WALL_BOUNCE_CHANGE_Y_NORMAL = 0.5f,
                HS_WALL_BOUNCE_FORCE_SCALE = 1 / 3.f, HS_WALL_BOUNCE_UP_FRAC = 0.3f;
constexpr float HS_BALL_START_X = -1000, HS_BALL_START_Y = -2220, HS_BALL_START_Z = 92.75f, HS_BALL_START_VY = -65,
                HS_BALL_START_VZ = 650;

// Hoops boost pad locations (src/RLConst.h:314-340): 6 big then 14 small.
RSG_FN V3 pad_location_hoops(int i) {
    switch (i) {
        case 0: return v3(-2176.f, 2944.f, 72.f);
        case 1: return v3(2176.f, -2944.f, 72.f);
        case 2: return v3(-2176.f, -2944.f, 72.f);
        case 3: return v3(-2432.f, 0.f, 72.f);
        case 4: return v3(2432.f, 0.f, 72.f);
        case 5: return v3(2175.99f, 2944.f, 72.f);
        case 6: return v3(1536.f, -1024.f, 64.f);
        case 7: return v3(-1280.f, -2304.f, 64.f);
        case 8: return v3(0.f, -2816.f, 64.f);
        case 9: return v3(-1536.f, -1024.f, 64.f);
        case 10: return v3(1280.f, -2304.f, 64.f);
        case 11: return v3(-512.f, 512.f, 64.f);
        case 12: return v3(-1536.f, 1024.f, 64.f);
        case 13: return v3(1536.f, 1024.f, 64.f);
        case 14: return v3(1280.f, 2304.f, 64.f);
        case 15: return v3(0.f, 2816.f, 64.f);
        case 16: return v3(512.f, 512.f, 64.f);
        case 17: return v3(512.f, -512.f, 64.f);
        case 18: return v3(-512.f, -512.f, 64.f);
        default: return v3(-1280.f, 2304.f, 64.f);
    }
}

// Boost pad locations (src/RLConst.h:264-309): 6 big then 28 small.
RSG_FN V3 pad_location_soccar(int i) {
    switch (i) {
        case 0: return v3(-3584.f, 0.f, 73.f);
        case 1: return v3(3584.f, 0.f, 73.f);
        case 2: return v3(-3072.f, 4096.f, 73.f);
        case 3: return v3(3072.f, 4096.f, 73.f);
        case 4: return v3(-3072.f, -4096.f, 73.f);
        case 5: return v3(3072.f, -4096.f, 73.f);
        case 6: return v3(0.f, -4240.f, 70.f);
        case 7: return v3(-1792.f, -4184.f, 70.f);
        case 8: return v3(1792.f, -4184.f, 70.f);
        case 9: return v3(-940.f, -3308.f, 70.f);
        case 10: return v3(940.f, -3308.f, 70.f);
        case 11: return v3(0.f, -2816.f, 70.f);
        case 12: return v3(-3584.f, -2484.f, 70.f);
        case 13: return v3(3584.f, -2484.f, 70.f);
        case 14: return v3(-1788.f, -2300.f, 70.f);
        case 15: return v3(1788.f, -2300.f, 70.f);
        case 16: return v3(-2048.f, -1036.f, 70.f);
        case 17: return v3(0.f, -1024.f, 70.f);
        case 18: return v3(2048.f, -1036.f, 70.f);
        case 19: return v3(-1024.f, 0.f, 70.f);
        case 20: return v3(1024.f, 0.f, 70.f);
        case 21: return v3(-2048.f, 1036.f, 70.f);
        case 22: return v3(0.f, 1024.f, 70.f);
        case 23: return v3(2048.f, 1036.f, 70.f);
        case 24: return v3(-1788.f, 2300.f, 70.f);
        case 25: return v3(1788.f, 2300.f, 70.f);
        case 26: return v3(-3584.f, 2484.f, 70.f);
        case 27: return v3(3584.f, 2484.f, 70.f);
        case 28: return v3(0.f, 2816.f, 70.f);
        case 29: return v3(-940.f, 3308.f, 70.f);
        case 30: return v3(940.f, 3308.f, 70.f);
        case 31: return v3(-1792.f, 4184.f, 70.f);
        case 32: return v3(1792.f, 4184.f, 70.f);
        default: return v3(0.f, 4240.f, 70.f);
    }
}

// Arena::IsBallScored (Arena.cpp:893-918) for a ball at `pos_bt` (Bullet
// units): past the goal line (Soccar arena) or low inside a hoop (Hoops).
RSG_FN bool ball_scored(int mode, V3 pos_bt, const rsg_mutator_config& m) {
    if (mode == RSG_MODE_HOOPS) {
        if (!(pos_bt.z < HOOPS_GOAL_SCORE_THRESHOLD_Z * UU_TO_BT)) return false;
        V3 p = pos_bt * BT_TO_UU;
        return hoops_goal_xy_margin_sq(p.x, p.y) < 0;
    }
    if (mode != RSG_MODE_SOCCAR && mode != RSG_MODE_HEATSEEKER) return false;
    return fabsf(pos_bt.y * BT_TO_UU) > (m.goal_base_threshold_y + m.ball_radius);
}

// Boost pad i of the mode's arena.
RSG_FN V3 pad_location(int mode, int i) {
    return mode == RSG_MODE_HOOPS ? pad_location_hoops(i) : pad_location_soccar(i);
}

}  // namespace rsg
