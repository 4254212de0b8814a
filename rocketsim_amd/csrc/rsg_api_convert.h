// rsg_api_convert.h -- conversion between the C ABI state records
// (include/rsg_state.h, public units: uu, uu/s, column-major RotMat) and the
// word-plane records in memory (Bullet units, basis rows), as the reference's
// Car/Ball Get/SetState convert (src/Sim/Car/Car.cpp:10-36,
// src/Sim/Ball/Ball.cpp:27-51).  Host code only.
#pragma once

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/rsg_state.h"
#include "rsg_tick.h"

namespace rsg {

inline void rot_api_to_rows(const float* api, M3& m) {
    // api: forward, right, up columns (RotMat, src/Math/MathTypes/MathTypes.h:169-170)
    for (int r = 0; r < 3; r++) m.r[r] = v3(api[0 + r], api[3 + r], api[6 + r]);
}
inline void rot_rows_to_api(const M3& m, float* api) {
    for (int r = 0; r < 3; r++) {
        api[0 + r] = m.r[r].x;
        api[3 + r] = m.r[r].y;
        api[6 + r] = m.r[r].z;
    }
}

inline Ctrl ctrl_from_api(const rsg_controls& c) {
    Ctrl o;
    o.throttle = c.throttle; o.steer = c.steer; o.pitch = c.pitch; o.yaw = c.yaw; o.roll = c.roll;
    o.jump = c.jump != 0; o.boost = c.boost != 0; o.handbrake = c.handbrake != 0;
    return o;
}
inline rsg_controls ctrl_to_api(const Ctrl& c) {
    rsg_controls o;
    o.throttle = c.throttle; o.steer = c.steer; o.pitch = c.pitch; o.yaw = c.yaw; o.roll = c.roll;
    o.jump = c.jump; o.boost = c.boost; o.handbrake = c.handbrake; o._pad = 0;
    return o;
}

inline void car_to_api(const CarRef& r, rsg_car_state& s) {
    Car c;
    load_car(r, c);
    std::memset(&s, 0, sizeof(s));
    s.tick_count_since_update = c.ticks_since_set;
    s.hit_tick = c.hit_tick;
    s.hit_extra_tick = c.hit_extra_tick;
    V3 p = c.b.pos * BT_TO_UU, v = c.b.vel * BT_TO_UU;
    s.pos[0] = p.x; s.pos[1] = p.y; s.pos[2] = p.z;
    rot_rows_to_api(c.api_rot, s.rot);
    s.vel[0] = v.x; s.vel[1] = v.y; s.vel[2] = v.z;
    s.ang_vel[0] = c.b.ang.x; s.ang_vel[1] = c.b.ang.y; s.ang_vel[2] = c.b.ang.z;
    s.flags = c.flags;
    s.flip_rel_torque[0] = c.flip_torque.x; s.flip_rel_torque[1] = c.flip_torque.y; s.flip_rel_torque[2] = c.flip_torque.z;
    s.jump_time = c.jump_time; s.flip_time = c.flip_time;
    s.air_time = c.air_time; s.air_time_since_jump = c.air_time_since_jump;
    s.boost = c.boost; s.time_since_boosted = c.time_since_boosted; s.boosting_time = c.boosting_time;
    s.supersonic_time = c.supersonic_time; s.handbrake_val = c.handbrake;
    s.auto_flip_timer = c.af_timer; s.auto_flip_torque_scale = c.af_scale;
    s.world_contact_normal[0] = c.world_normal.x; s.world_contact_normal[1] = c.world_normal.y; s.world_contact_normal[2] = c.world_normal.z;
    s.car_contact_other_id = c.other_id;
    s.car_contact_cooldown = c.car_cooldown;
    s.demo_respawn_timer = c.demo_timer;
    s.hit_rel_pos[0] = c.hit_rel_pos.x; s.hit_rel_pos[1] = c.hit_rel_pos.y; s.hit_rel_pos[2] = c.hit_rel_pos.z;
    s.hit_ball_pos[0] = c.hit_ball_pos.x; s.hit_ball_pos[1] = c.hit_ball_pos.y; s.hit_ball_pos[2] = c.hit_ball_pos.z;
    s.hit_extra_vel[0] = c.hit_extra_vel.x; s.hit_extra_vel[1] = c.hit_extra_vel.y; s.hit_extra_vel[2] = c.hit_extra_vel.z;
    s.last_controls = ctrl_to_api(c.last);
    s.wheel_engine_force = c.engine; s.wheel_brake = c.brake; s.wheel_steer = c.steer;
    for (int w = 0; w < 4; w++) {
        s.wheel_lat_friction[w] = c.lat[w];
        s.wheel_long_friction[w] = c.lon[w];
        s.wheel_extra_pushback[w] = c.push[w];
    }
    s.bp_cell = (int32_t)r.u(CW_BP_CELL);
    s.bp_seq = r.u(CW_BP_SEQ);
}

inline void car_from_api(const CarRef& r, const rsg_car_state& s) {
    Car c;
    load_car(r, c);
    c.b.pos = v3(s.pos[0], s.pos[1], s.pos[2]) * UU_TO_BT;
    rot_api_to_rows(s.rot, c.b.rot);
    c.api_rot = c.b.rot;
    c.b.vel = v3(s.vel[0], s.vel[1], s.vel[2]) * UU_TO_BT;
    c.b.ang = v3(s.ang_vel[0], s.ang_vel[1], s.ang_vel[2]);
    c.flags = s.flags;
    c.flip_torque = v3(s.flip_rel_torque[0], s.flip_rel_torque[1], s.flip_rel_torque[2]);
    c.jump_time = s.jump_time; c.flip_time = s.flip_time;
    c.air_time = s.air_time; c.air_time_since_jump = s.air_time_since_jump;
    c.boost = s.boost; c.time_since_boosted = s.time_since_boosted; c.boosting_time = s.boosting_time;
    c.supersonic_time = s.supersonic_time; c.handbrake = s.handbrake_val;
    c.af_timer = s.auto_flip_timer; c.af_scale = s.auto_flip_torque_scale;
    c.world_normal = v3(s.world_contact_normal[0], s.world_contact_normal[1], s.world_contact_normal[2]);
    c.other_id = s.car_contact_other_id;
    c.car_cooldown = s.car_contact_cooldown;
    c.demo_timer = s.demo_respawn_timer;
    c.hit_rel_pos = v3(s.hit_rel_pos[0], s.hit_rel_pos[1], s.hit_rel_pos[2]);
    c.hit_ball_pos = v3(s.hit_ball_pos[0], s.hit_ball_pos[1], s.hit_ball_pos[2]);
    c.hit_extra_vel = v3(s.hit_extra_vel[0], s.hit_extra_vel[1], s.hit_extra_vel[2]);
    c.hit_tick = s.hit_tick;
    c.hit_extra_tick = s.hit_extra_tick;
    c.ticks_since_set = s.tick_count_since_update;
    c.last = ctrl_from_api(s.last_controls);
    c.engine = s.wheel_engine_force; c.brake = s.wheel_brake; c.steer = s.wheel_steer;
    for (int w = 0; w < 4; w++) {
        c.lat[w] = s.wheel_lat_friction[w];
        c.lon[w] = s.wheel_long_friction[w];
        c.push[w] = s.wheel_extra_pushback[w];
    }
    store_car(r, c);
    r.u(CW_BP_CELL) = (uint32_t)(s.bp_cell >= 0 && s.bp_cell < BP_NUM_CELLS ? s.bp_cell : 0);
    r.u(CW_BP_SEQ) = s.bp_seq;
}

inline void ball_to_api(const CarRef& r, rsg_ball_state& s) {
    std::memset(&s, 0, sizeof(s));
    V3 p = r.v(BW_POS) * BT_TO_UU, v = r.v(BW_VEL) * BT_TO_UU, w = r.v(BW_ANGVEL);
    s.pos[0] = p.x; s.pos[1] = p.y; s.pos[2] = p.z;
    rot_rows_to_api(load_rot(r, BW_ROT), s.rot);
    s.vel[0] = v.x; s.vel[1] = v.y; s.vel[2] = v.z;
    s.ang_vel[0] = w.x; s.ang_vel[1] = w.y; s.ang_vel[2] = w.z;
    s.tick_count_since_update = r.u64(BW_TICKS_SINCE_SET);
    s.bp_cell = (int32_t)r.u(BW_BP_CELL);
    s.bp_seq = r.u(BW_BP_SEQ);
    s.hs_y_target_dir = r.f(BW_HS);
    s.hs_cur_target_speed = r.f(BW_HS + 1);
    s.hs_time_since_hit = r.f(BW_HS + 2);
}

inline void ball_from_api(const CarRef& r, const rsg_ball_state& s) {
    r.sv(BW_POS, v3(s.pos[0], s.pos[1], s.pos[2]) * UU_TO_BT);
    M3 m;
    rot_api_to_rows(s.rot, m);
    store_rot(r, BW_ROT, m);
    r.sv(BW_VEL, v3(s.vel[0], s.vel[1], s.vel[2]) * UU_TO_BT);
    r.sv(BW_ANGVEL, v3(s.ang_vel[0], s.ang_vel[1], s.ang_vel[2]));
    r.su64(BW_TICKS_SINCE_SET, s.tick_count_since_update);
    r.u(BW_BP_CELL) = (uint32_t)(s.bp_cell >= 0 && s.bp_cell < BP_NUM_CELLS ? s.bp_cell : 0);
    r.u(BW_BP_SEQ) = s.bp_seq;
    r.sf(BW_HS, s.hs_y_target_dir);
    r.sf(BW_HS + 1, s.hs_cur_target_speed);
    r.sf(BW_HS + 2, s.hs_time_since_hit);
    // Ball::SetState wakes a moving ball (src/Sim/Ball/Ball.cpp:46-47)
    if (s.vel[0] != 0.f || s.vel[1] != 0.f || s.vel[2] != 0.f || s.ang_vel[0] != 0.f || s.ang_vel[1] != 0.f ||
        s.ang_vel[2] != 0.f)
        r.u(BW_AWAKE) = 1;
}

// The arena's MutatorConfig (the mode's defaults where the batch keeps none).
inline rsg_mutator_config arena_mutators(const BatchDev& h, int a) {
    return h.mut ? h.mut[a] : mutator_defaults(h.mode);
}

// btRigidBody::setGravity (BulletDynamics/Dynamics/btRigidBody.cpp:160-167):
// the acceleration kept, the force it gives the body's current mass.
inline void set_body_gravity(const CarRef& r, int acc_word, int force_word, const float* gravity_uu, float inv_mass) {
    V3 acc = v3(gravity_uu[0], gravity_uu[1], gravity_uu[2]) * UU_TO_BT;
    r.sv(acc_word, acc);
    r.sv(force_word, acc * (1.f / inv_mass));
}

// A body as the reference makes it, added to a world with gravity
// `gravity_uu` (btDiscreteDynamicsWorld::addRigidBody): the car's box and mass
// (Car::_BulletSetup, src/Sim/Car/Car.cpp:195-215; mass words 0 = the car
// default), the ball's sphere and mass (Ball::_BulletSetup,
// src/Sim/Ball/Ball.cpp:53-115), awake.
inline void init_car_body_words(const CarRef& cr, const float* gravity_uu) {
    cr.sf(CW_BODY_INV_MASS, 0.f);
    cr.sf(CW_BODY_INERTIA_MASS, 0.f);
    set_body_gravity(cr, CW_BODY_GRAVITY_ACC, CW_BODY_GRAVITY, gravity_uu, 1.f / CAR_MASS_BT);
}
inline void init_ball_body_words(const CarRef& br, float radius_uu, float mass, const float* gravity_uu) {
    br.sf(BW_RADIUS, radius_uu);
    br.sf(BW_MASS, mass);
    br.u(BW_AWAKE) = 1;
    set_body_gravity(br, BW_GRAVITY_ACC, BW_GRAVITY, gravity_uu, make_ball_phys(radius_uu, mass).inv_mass);
}

// A car's CarConfig: its entry in the batch's table (the presets without one).
inline rsg_car_config config_of(const BatchDev& h, int index) {
    return h.cfgs ? h.cfgs[index] : preset_config(index);
}

// Two CarConfigs with equal fields (the padding bytes are not compared).
inline bool same_car_config(const rsg_car_config& a, const rsg_car_config& b) {
    rsg_car_config x = a, y = b;
    std::memset(x._pad, 0, sizeof(x._pad));
    std::memset(y._pad, 0, sizeof(y._pad));
    x.three_wheels = x.three_wheels != 0;
    y.three_wheels = y.three_wheels != 0;
    return std::memcmp(&x, &y, sizeof(x)) == 0;
}

// The preset a config equals, or RSG_PRESET_CUSTOM.
inline int preset_of(const rsg_car_config& c) {
    for (int p = 0; p < RSG_PRESET_COUNT; p++)
        if (same_car_config(preset_config(p), c)) return p;
    return RSG_PRESET_CUSTOM;
}

// Why a CarConfig cannot be simulated ("" if it can): non-finite fields,
// a hitbox with a non-positive size, a negative wheel radius or rest length,
// or a hitbox whose AABB diagonal exceeds the broadphase cell
// (btRSBroadphase::createProxy refuses it, btRSBroadphase.cpp:227-229).
inline std::string car_config_problem(const rsg_car_config& c) {
    const float f[] = {c.hitbox_size[0], c.hitbox_size[1], c.hitbox_size[2], c.hitbox_pos_offset[0],
                       c.hitbox_pos_offset[1], c.hitbox_pos_offset[2], c.front_wheels.wheel_radius,
                       c.front_wheels.suspension_rest_length, c.front_wheels.connection_point_offset[0],
                       c.front_wheels.connection_point_offset[1], c.front_wheels.connection_point_offset[2],
                       c.back_wheels.wheel_radius, c.back_wheels.suspension_rest_length,
                       c.back_wheels.connection_point_offset[0], c.back_wheels.connection_point_offset[1],
                       c.back_wheels.connection_point_offset[2], c.dodge_deadzone};
    for (float x : f)
        if (!std::isfinite(x)) return "non-finite car config field";
    if (!(c.hitbox_size[0] > 0.f) || !(c.hitbox_size[1] > 0.f) || !(c.hitbox_size[2] > 0.f))
        return "car hitbox sizes must be positive";
    if (c.front_wheels.wheel_radius < 0.f || c.back_wheels.wheel_radius < 0.f ||
        c.front_wheels.suspension_rest_length < 0.f || c.back_wheels.suspension_rest_length < 0.f)
        return "negative wheel radius or suspension rest length";
    CarPhys cp = make_car_phys(c);
    V3 d = (cp.local_aabb_center + cp.local_aabb_half) - (cp.local_aabb_center - cp.local_aabb_half);
    float cell = bp_cell_size();
    if (dot(d, d) > cell * cell) return "car hitbox AABB exceeds the broadphase cell size";
    return "";
}

// btRSBroadphase::createProxy refuses a dynamic proxy whose AABB diagonal
// exceeds the cell size (btRSBroadphase.cpp:227-229); a ball of radius
// `radius_uu` made at (0, 0, radius) passes if this holds.
inline bool ball_fits_broadphase(float radius_uu) {
    BallPhys bp = make_ball_phys(radius_uu, BALL_MASS_BT);
    float ext = bp.radius + 0.08f;
    V3 c = v3(0, 0, bp.radius);
    V3 d = (c + v3(ext, ext, ext)) - (c - v3(ext, ext, ext));
    float cell = bp_cell_size();
    return !(dot(d, d) > cell * cell);
}

// Arena::SetMutatorConfig (src/Sim/Arena/Arena.cpp:17-48) on arena `a`'s
// records, from config `old` to `m`:
//   * the world gravity goes to the bodies that are awake
//     (btDiscreteDynamicsWorld::setGravity: a demolished car's disabled body
//     and a sleeping ball keep theirs);
//   * a new ball radius or mass remakes the ball body (Ball::_BulletSetup):
//     at rest at (0, 0, radius), awake, with the world gravity, registered
//     again in the broadphase after every other proxy;
//   * a new car mass gives every car body mass ballMass and the box inertia
//     of carMass (the reference passes mutatorConfig.ballMass to
//     setMassProps, Arena.cpp:38-39), with m_gravity = mass * acceleration.
// Ball friction, restitution and drag, and everything else, are read from
// the config each tick.
inline void apply_mutator_config(const BatchDev& h, int a, const rsg_mutator_config& old, const rsg_mutator_config& m) {
    const bool ball_changed = m.ball_radius != old.ball_radius || m.ball_mass != old.ball_mass;
    const bool car_mass_changed = m.car_mass != old.car_mass;
    const int n_cars = (int)arena_ref(h, a).u(AW_NUM_CARS);
    CarRef br = ball_ref(h, a);
    for (int i = 0; i < n_cars; i++) {
        CarRef cr = car_ref(h, a, i);
        if (cr.u(CW_FLAGS) & RSG_CF_BODY_DISABLED) continue;
        float im = cr.f(CW_BODY_INV_MASS);
        set_body_gravity(cr, CW_BODY_GRAVITY_ACC, CW_BODY_GRAVITY, m.gravity, im != 0.f ? im : 1.f / CAR_MASS_BT);
    }
    if (br.u(BW_AWAKE))
        set_body_gravity(br, BW_GRAVITY_ACC, BW_GRAVITY, m.gravity, make_ball_phys(br.f(BW_RADIUS), br.f(BW_MASS)).inv_mass);
    if (ball_changed) {
        init_ball_body_words(br, m.ball_radius, m.ball_mass, m.gravity);
        BallPhys bp = make_ball_phys(m.ball_radius, m.ball_mass);
        br.sv(BW_POS, v3(0, 0, bp.radius));
        store_rot(br, BW_ROT, m3_identity());
        br.sv(BW_VEL, v3zero());
        br.sv(BW_ANGVEL, v3zero());
        float ext = bp.radius + 0.08f;
        br.u(BW_BP_CELL) = (uint32_t)bp_cell_of(v3(0, 0, bp.radius) - v3(ext, ext, ext));
        uint32_t last = br.u(BW_BP_SEQ);
        for (int i = 0; i < n_cars; i++) {
            uint32_t q = car_ref(h, a, i).u(CW_BP_SEQ);
            last = q > last ? q : last;
        }
        br.u(BW_BP_SEQ) = last + 1;
    }
    if (car_mass_changed) {
        for (int i = 0; i < n_cars; i++) {
            CarRef cr = car_ref(h, a, i);
            cr.sf(CW_BODY_INV_MASS, 1.f / m.ball_mass);
            cr.sf(CW_BODY_INERTIA_MASS, m.car_mass);
            cr.sv(CW_BODY_GRAVITY, m.ball_mass * cr.v(CW_BODY_GRAVITY_ACC));
        }
    }
}

// A new arena's records (Arena::Create + AddCar, src/Sim/Arena/Arena.cpp:50-62,
// 416-540): cars with their config and spawn boost, the ball at rest, pads
// active; the proxies registered where the reference creates them.
// `preset` holds each car's index in the batch's config table.
inline void init_arena_record(const BatchDev& h, int a, int n_cars, const uint8_t* team, const uint8_t* preset) {
    const int mode = h.mode;
    const rsg_mutator_config mut = arena_mutators(h, a);
    CarRef ar = arena_ref(h, a);
    ar.u(AW_NUM_CARS) = (uint32_t)n_cars;
    uint32_t ow = 0;  // default car update order: reverse creation
    for (int k = 0; k < n_cars; k++) ow |= (uint32_t)(n_cars - 1 - k) << (4 * k);
    ar.u(AW_CAR_ORDER) = ow;
    for (int i = 0; i < n_cars; i++) {
        CarRef cr = car_ref(h, a, i);
        cr.u(CW_CONFIG) = (uint32_t)preset[i] | ((uint32_t)team[i] << 8);
        cr.u(CW_ID) = (uint32_t)(i + 1);
        cr.su64(CW_HIT_TICK, ~0ull);
        cr.su64(CW_HIT_EXTRA_TICK, ~0ull);
        cr.sf(CW_BOOST, mut.car_spawn_boost_amount);
        store_rot(cr, CW_ROT, m3_identity());
        init_car_body_words(cr, mut.gravity);
        // proxy created at the identity transform (Car::_BulletSetup,
        // src/Sim/Car/Car.cpp:212-233), AABB without contact threshold
        CarPhys cp = make_car_phys(config_of(h, preset[i]));
        cr.u(CW_BP_CELL) = (uint32_t)bp_cell_of(cp.local_aabb_center - cp.local_aabb_half);
        cr.u(CW_BP_SEQ) = (uint32_t)(1 + i);
    }
    CarRef br = ball_ref(h, a);
    br.sv(BW_POS, v3(0, 0, BALL_REST_Z) * UU_TO_BT);
    store_rot(br, BW_ROT, m3_identity());
    br.sf(BW_HS + 1, HS_INITIAL_TARGET_SPEED);
    init_ball_body_words(br, mut.ball_radius, mut.ball_mass, mut.gravity);
    {
        // proxy created at (0, 0, radius) (Ball::_BulletSetup, src/Sim/Ball/Ball.cpp:85-114)
        BallPhys bp = make_ball_phys(mut.ball_radius, mut.ball_mass);
        float ext = bp.radius + 0.08f;
        V3 c = v3(0, 0, bp.radius);
        br.u(BW_BP_CELL) = (uint32_t)bp_cell_of(c - v3(ext, ext, ext));
        br.u(BW_BP_SEQ) = 0;
    }
    if (mode_has_arena(mode))
        for (int p = 0; p < PAD_NUM; p++) pad_ref(h, a, p).u(PW_ACTIVE) = 1;
}

inline void pad_to_api(const CarRef& r, rsg_pad_state& s) {
    s.is_active = r.u(PW_ACTIVE);
    s.cooldown = r.f(PW_COOLDOWN);
    s.prev_locked_car_id = r.u(PW_PREV_LOCKED);
}

inline void pad_from_api(const CarRef& r, const rsg_pad_state& s) {
    r.u(PW_ACTIVE) = s.is_active ? 1u : 0u;
    r.sf(PW_COOLDOWN, s.cooldown);
    r.u(PW_PREV_LOCKED) = s.prev_locked_car_id;
}

}  // namespace rsg
