// rsg_rocketsim.hpp -- conversions between the reference's own state types
// and this library's records, for C++ code that already holds RocketSim
// types (src/Sim/Car/Car.h:17-120 CarState, src/Sim/Ball/Ball.h:18-57
// BallState, src/Sim/BoostPad/BoostPad.h:23-31 BoostPadState,
// src/Sim/CarControls.h:7-33 CarControls).  Include it after the
// reference's "RocketSim.h" (its src/ directory on the include path); with it
// a program moves its reference states into an rsg_amd::ArenaBatch and back:
//
//     rsg_car_state s = rsg_amd::ToRsg(*car);          // a RocketSim::Car
//     batch.SetCarStates({...});  batch.Step(1);
//     RocketSim::CarState back = rsg_amd::FromRsg(batch.GetCarStates()[0]);
//
// Positions and velocities stay in uu and the rotation in the reference's
// forward / right / up columns, as the reference's GetState/SetState give
// them.  The public CarState does not hold the per-wheel values btVehicleRL
// keeps between ticks nor the rigid body's contact-response flag; ToRsg(Car)
// and ApplyRsg read and write them from a Car object's members, so a state
// moved this way steps on the GPU exactly as it would in the reference.
// The broadphase registration fields (bp_cell, bp_seq) are left to the
// caller: the reference keeps them inside btRSBroadphase.
#ifndef RSG_ROCKETSIM_HPP
#define RSG_ROCKETSIM_HPP

#include <cstring>
#include <stdexcept>

#include "rsg.hpp"

namespace rsg_amd {

inline void ToRsg(const RocketSim::Vec& v, float* f) { f[0] = v.x; f[1] = v.y; f[2] = v.z; }
inline RocketSim::Vec VecFromRsg(const float* f) { return RocketSim::Vec(f[0], f[1], f[2]); }
inline void ToRsg(const RocketSim::RotMat& m, float* f) {
    ToRsg(m.forward, f);
    ToRsg(m.right, f + 3);
    ToRsg(m.up, f + 6);
}
inline RocketSim::RotMat RotFromRsg(const float* f) {
    RocketSim::RotMat m;
    m.forward = VecFromRsg(f);
    m.right = VecFromRsg(f + 3);
    m.up = VecFromRsg(f + 6);
    return m;
}

inline rsg_controls ToRsg(const RocketSim::CarControls& c) {
    rsg_controls s;
    s.throttle = c.throttle; s.steer = c.steer; s.pitch = c.pitch; s.yaw = c.yaw; s.roll = c.roll;
    s.jump = c.jump; s.boost = c.boost; s.handbrake = c.handbrake; s._pad = 0;
    return s;
}
inline RocketSim::CarControls FromRsg(const rsg_controls& s) {
    RocketSim::CarControls c;
    c.throttle = s.throttle; c.steer = s.steer; c.pitch = s.pitch; c.yaw = s.yaw; c.roll = s.roll;
    c.jump = s.jump; c.boost = s.boost; c.handbrake = s.handbrake;
    return c;
}

// The public CarState fields into `out`; the fields CarState does not hold
// (per-wheel values, body flag, broadphase record) keep out's values.
inline void ToRsg(const RocketSim::CarState& cs, rsg_car_state& out) {
    out.tick_count_since_update = cs.tickCountSinceUpdate;
    ToRsg(cs.pos, out.pos); ToRsg(cs.rotMat, out.rot); ToRsg(cs.vel, out.vel); ToRsg(cs.angVel, out.ang_vel);
    uint32_t f = out.flags & RSG_CF_BODY_DISABLED;
    if (cs.isOnGround) f |= RSG_CF_ON_GROUND;
    for (int w = 0; w < 4; w++) if (cs.wheelsWithContact[w]) f |= (RSG_CF_WHEEL0 << w);
    if (cs.hasJumped) f |= RSG_CF_HAS_JUMPED;
    if (cs.hasDoubleJumped) f |= RSG_CF_HAS_DOUBLE_JUMPED;
    if (cs.hasFlipped) f |= RSG_CF_HAS_FLIPPED;
    if (cs.isFlipping) f |= RSG_CF_IS_FLIPPING;
    if (cs.isJumping) f |= RSG_CF_IS_JUMPING;
    if (cs.isBoosting) f |= RSG_CF_IS_BOOSTING;
    if (cs.isSupersonic) f |= RSG_CF_SUPERSONIC;
    if (cs.isAutoFlipping) f |= RSG_CF_AUTO_FLIPPING;
    if (cs.worldContact.hasContact) f |= RSG_CF_WORLD_CONTACT;
    if (cs.isDemoed) f |= RSG_CF_DEMOED;
    if (cs.ballHitInfo.isValid) f |= RSG_CF_HIT_VALID;
    out.flags = f;
    ToRsg(cs.flipRelTorque, out.flip_rel_torque);
    out.jump_time = cs.jumpTime; out.flip_time = cs.flipTime;
    out.air_time = cs.airTime; out.air_time_since_jump = cs.airTimeSinceJump;
    out.boost = cs.boost; out.time_since_boosted = cs.timeSinceBoosted; out.boosting_time = cs.boostingTime;
    out.supersonic_time = cs.supersonicTime; out.handbrake_val = cs.handbrakeVal;
    out.auto_flip_timer = cs.autoFlipTimer; out.auto_flip_torque_scale = cs.autoFlipTorqueScale;
    ToRsg(cs.worldContact.contactNormal, out.world_contact_normal);
    out.car_contact_other_id = cs.carContact.otherCarID;
    out.car_contact_cooldown = cs.carContact.cooldownTimer;
    out.demo_respawn_timer = cs.demoRespawnTimer;
    ToRsg(cs.ballHitInfo.relativePosOnBall, out.hit_rel_pos);
    ToRsg(cs.ballHitInfo.ballPos, out.hit_ball_pos);
    ToRsg(cs.ballHitInfo.extraHitVel, out.hit_extra_vel);
    out.hit_tick = cs.ballHitInfo.tickCountWhenHit;
    out.hit_extra_tick = cs.ballHitInfo.tickCountWhenExtraImpulseApplied;
    out.last_controls = ToRsg(cs.lastControls);
}

inline RocketSim::CarState FromRsg(const rsg_car_state& s) {
    RocketSim::CarState cs;
    cs.pos = VecFromRsg(s.pos); cs.rotMat = RotFromRsg(s.rot); cs.vel = VecFromRsg(s.vel); cs.angVel = VecFromRsg(s.ang_vel);
    const uint32_t f = s.flags;
    cs.isOnGround = f & RSG_CF_ON_GROUND;
    for (int w = 0; w < 4; w++) cs.wheelsWithContact[w] = f & (RSG_CF_WHEEL0 << w);
    cs.hasJumped = f & RSG_CF_HAS_JUMPED;
    cs.hasDoubleJumped = f & RSG_CF_HAS_DOUBLE_JUMPED;
    cs.hasFlipped = f & RSG_CF_HAS_FLIPPED;
    cs.isFlipping = f & RSG_CF_IS_FLIPPING;
    cs.isJumping = f & RSG_CF_IS_JUMPING;
    cs.isBoosting = f & RSG_CF_IS_BOOSTING;
    cs.isSupersonic = f & RSG_CF_SUPERSONIC;
    cs.isAutoFlipping = f & RSG_CF_AUTO_FLIPPING;
    cs.worldContact.hasContact = f & RSG_CF_WORLD_CONTACT;
    cs.isDemoed = f & RSG_CF_DEMOED;
    cs.ballHitInfo.isValid = f & RSG_CF_HIT_VALID;
    cs.flipRelTorque = VecFromRsg(s.flip_rel_torque);
    cs.jumpTime = s.jump_time; cs.flipTime = s.flip_time;
    cs.airTime = s.air_time; cs.airTimeSinceJump = s.air_time_since_jump;
    cs.boost = s.boost; cs.timeSinceBoosted = s.time_since_boosted; cs.boostingTime = s.boosting_time;
    cs.supersonicTime = s.supersonic_time; cs.handbrakeVal = s.handbrake_val;
    cs.autoFlipTimer = s.auto_flip_timer; cs.autoFlipTorqueScale = s.auto_flip_torque_scale;
    cs.worldContact.contactNormal = VecFromRsg(s.world_contact_normal);
    cs.carContact.otherCarID = s.car_contact_other_id;
    cs.carContact.cooldownTimer = s.car_contact_cooldown;
    cs.demoRespawnTimer = s.demo_respawn_timer;
    cs.ballHitInfo.relativePosOnBall = VecFromRsg(s.hit_rel_pos);
    cs.ballHitInfo.ballPos = VecFromRsg(s.hit_ball_pos);
    cs.ballHitInfo.extraHitVel = VecFromRsg(s.hit_extra_vel);
    cs.ballHitInfo.tickCountWhenHit = s.hit_tick;
    cs.ballHitInfo.tickCountWhenExtraImpulseApplied = s.hit_extra_tick;
    cs.lastControls = FromRsg(s.last_controls);
    cs.tickCountSinceUpdate = s.tick_count_since_update;
    return cs;
}

// A reference car's complete record: its CarState plus the per-wheel values
// of its btVehicleRL and its body's contact-response flag (broadphase fields
// zero).
inline rsg_car_state ToRsg(RocketSim::Car& car) {
    rsg_car_state s;
    std::memset(&s, 0, sizeof(s));
    ToRsg(car.GetState(), s);
    if (car._rigidBody.m_collisionFlags & btCollisionObject::CF_NO_CONTACT_RESPONSE) s.flags |= RSG_CF_BODY_DISABLED;
    const auto& wi = car._bulletVehicle.m_wheelInfo;
    s.wheel_engine_force = wi[0].m_engineForce;
    s.wheel_brake = wi[0].m_brake;
    s.wheel_steer = wi[0].m_steerAngle;
    for (int w = 0; w < 4; w++) {
        s.wheel_lat_friction[w] = wi[w].m_latFriction;
        s.wheel_long_friction[w] = wi[w].m_longFriction;
        s.wheel_extra_pushback[w] = wi[w].m_extraPushback;
    }
    return s;
}

// The whole record into a reference car: SetState with the CarState fields,
// then the tick counter, the body flag and the per-wheel values (front wheels
// steer, rear ones do not).
inline void ApplyRsg(const rsg_car_state& s, RocketSim::Car& car) {
    car.SetState(FromRsg(s));
    car._internalState.tickCountSinceUpdate = s.tick_count_since_update;
    if (s.flags & RSG_CF_BODY_DISABLED) {
        car._rigidBody.m_activationState1 = DISABLE_SIMULATION;
        car._rigidBody.m_collisionFlags |= btCollisionObject::CF_NO_CONTACT_RESPONSE;
    } else {
        car._rigidBody.m_activationState1 = ACTIVE_TAG;
        car._rigidBody.m_collisionFlags &= ~btCollisionObject::CF_NO_CONTACT_RESPONSE;
    }
    auto& wi = car._bulletVehicle.m_wheelInfo;
    for (int w = 0; w < 4; w++) {
        wi[w].m_engineForce = s.wheel_engine_force;
        wi[w].m_brake = s.wheel_brake;
        wi[w].m_steerAngle = (w < 2) ? s.wheel_steer : 0.f;
        wi[w].m_latFriction = s.wheel_lat_friction[w];
        wi[w].m_longFriction = s.wheel_long_friction[w];
        wi[w].m_extraPushback = s.wheel_extra_pushback[w];
    }
}

// BallState (broadphase fields zero).
inline rsg_ball_state ToRsg(const RocketSim::BallState& bs) {
    rsg_ball_state s;
    std::memset(&s, 0, sizeof(s));
    s.tick_count_since_update = bs.tickCountSinceUpdate;
    ToRsg(bs.pos, s.pos); ToRsg(bs.rotMat, s.rot); ToRsg(bs.vel, s.vel); ToRsg(bs.angVel, s.ang_vel);
    s.hs_y_target_dir = bs.hsInfo.yTargetDir;
    s.hs_cur_target_speed = bs.hsInfo.curTargetSpeed;
    s.hs_time_since_hit = bs.hsInfo.timeSinceHit;
    return s;
}

inline RocketSim::BallState FromRsg(const rsg_ball_state& s) {
    RocketSim::BallState bs;
    bs.pos = VecFromRsg(s.pos); bs.rotMat = RotFromRsg(s.rot); bs.vel = VecFromRsg(s.vel); bs.angVel = VecFromRsg(s.ang_vel);
    bs.hsInfo.yTargetDir = s.hs_y_target_dir;
    bs.hsInfo.curTargetSpeed = s.hs_cur_target_speed;
    bs.hsInfo.timeSinceHit = s.hs_time_since_hit;
    bs.tickCountSinceUpdate = s.tick_count_since_update;
    return bs;
}

inline rsg_pad_state ToRsg(const RocketSim::BoostPadState& p) {
    rsg_pad_state s;
    s.is_active = p.isActive ? 1u : 0u;
    s.cooldown = p.cooldown;
    s.prev_locked_car_id = p.prevLockedCarID;
    return s;
}

inline RocketSim::BoostPadState FromRsg(const rsg_pad_state& s) {
    RocketSim::BoostPadState p;
    p.isActive = s.is_active != 0;
    p.cooldown = s.cooldown;
    p.prevLockedCarID = s.prev_locked_car_id;
    return p;
}

// CarConfig, field for field (src/Sim/Car/CarConfig/CarConfig.h:6-43).
inline rsg_car_config ToRsg(const RocketSim::CarConfig& c) {
    rsg_car_config s;
    std::memset(&s, 0, sizeof(s));
    ToRsg(c.hitboxSize, s.hitbox_size);
    ToRsg(c.hitboxPosOffset, s.hitbox_pos_offset);
    const RocketSim::WheelPairConfig* src[2] = {&c.frontWheels, &c.backWheels};
    rsg_wheel_pair_config* dst[2] = {&s.front_wheels, &s.back_wheels};
    for (int k = 0; k < 2; k++) {
        dst[k]->wheel_radius = src[k]->wheelRadius;
        dst[k]->suspension_rest_length = src[k]->suspensionRestLength;
        ToRsg(src[k]->connectionPointOffset, dst[k]->connection_point_offset);
    }
    s.three_wheels = c.threeWheels ? 1 : 0;
    s.dodge_deadzone = c.dodgeDeadzone;
    return s;
}

inline RocketSim::CarConfig FromRsg(const rsg_car_config& s) {
    RocketSim::CarConfig c;
    c.hitboxSize = VecFromRsg(s.hitbox_size);
    c.hitboxPosOffset = VecFromRsg(s.hitbox_pos_offset);
    RocketSim::WheelPairConfig* dst[2] = {&c.frontWheels, &c.backWheels};
    const rsg_wheel_pair_config* src[2] = {&s.front_wheels, &s.back_wheels};
    for (int k = 0; k < 2; k++) {
        dst[k]->wheelRadius = src[k]->wheel_radius;
        dst[k]->suspensionRestLength = src[k]->suspension_rest_length;
        dst[k]->connectionPointOffset = VecFromRsg(src[k]->connection_point_offset);
    }
    c.threeWheels = s.three_wheels != 0;
    c.dodgeDeadzone = s.dodge_deadzone;
    return c;
}

// MutatorConfig, field for field (src/Sim/MutatorConfig/MutatorConfig.h:16-79).
inline rsg_mutator_config ToRsg(const RocketSim::MutatorConfig& m) {
    rsg_mutator_config s;
    std::memset(&s, 0, sizeof(s));
    ToRsg(m.gravity, s.gravity);
    s.car_mass = m.carMass;
    s.car_world_friction = m.carWorldFriction;
    s.car_world_restitution = m.carWorldRestitution;
    s.ball_mass = m.ballMass;
    s.ball_max_speed = m.ballMaxSpeed;
    s.ball_drag = m.ballDrag;
    s.ball_world_friction = m.ballWorldFriction;
    s.ball_world_restitution = m.ballWorldRestitution;
    s.jump_accel = m.jumpAccel;
    s.jump_immediate_force = m.jumpImmediateForce;
    s.boost_accel_ground = m.boostAccelGround;
    s.boost_accel_air = m.boostAccelAir;
    s.boost_used_per_second = m.boostUsedPerSecond;
    s.respawn_delay = m.respawnDelay;
    s.bump_cooldown_time = m.bumpCooldownTime;
    s.boost_pad_cooldown_big = m.boostPadCooldown_Big;
    s.boost_pad_cooldown_small = m.boostPadCooldown_Small;
    s.car_spawn_boost_amount = m.carSpawnBoostAmount;
    s.ball_hit_extra_force_scale = m.ballHitExtraForceScale;
    s.bump_force_scale = m.bumpForceScale;
    s.ball_radius = m.ballRadius;
    s.unlimited_flips = m.unlimitedFlips;
    s.unlimited_double_jumps = m.unlimitedDoubleJumps;
    s.recharge_boost_enabled = m.rechargeBoostEnabled;
    s.demo_mode = (uint8_t)m.demoMode;
    s.recharge_boost_per_second = m.rechargeBoostPerSecond;
    s.recharge_boost_delay = m.rechargeBoostDelay;
    s.enable_team_demos = m.enableTeamDemos;
    s.goal_base_threshold_y = m.goalBaseThresholdY;
    return s;
}

inline RocketSim::MutatorConfig FromRsg(const rsg_mutator_config& s, RocketSim::GameMode mode) {
    RocketSim::MutatorConfig m(mode);
    m.gravity = VecFromRsg(s.gravity);
    m.carMass = s.car_mass;
    m.carWorldFriction = s.car_world_friction;
    m.carWorldRestitution = s.car_world_restitution;
    m.ballMass = s.ball_mass;
    m.ballMaxSpeed = s.ball_max_speed;
    m.ballDrag = s.ball_drag;
    m.ballWorldFriction = s.ball_world_friction;
    m.ballWorldRestitution = s.ball_world_restitution;
    m.jumpAccel = s.jump_accel;
    m.jumpImmediateForce = s.jump_immediate_force;
    m.boostAccelGround = s.boost_accel_ground;
    m.boostAccelAir = s.boost_accel_air;
    m.boostUsedPerSecond = s.boost_used_per_second;
    m.respawnDelay = s.respawn_delay;
    m.bumpCooldownTime = s.bump_cooldown_time;
    m.boostPadCooldown_Big = s.boost_pad_cooldown_big;
    m.boostPadCooldown_Small = s.boost_pad_cooldown_small;
    m.carSpawnBoostAmount = s.car_spawn_boost_amount;
    m.ballHitExtraForceScale = s.ball_hit_extra_force_scale;
    m.bumpForceScale = s.bump_force_scale;
    m.ballRadius = s.ball_radius;
    m.unlimitedFlips = s.unlimited_flips != 0;
    m.unlimitedDoubleJumps = s.unlimited_double_jumps != 0;
    m.rechargeBoostEnabled = s.recharge_boost_enabled != 0;
    m.demoMode = (RocketSim::DemoMode)s.demo_mode;
    m.rechargeBoostPerSecond = s.recharge_boost_per_second;
    m.rechargeBoostDelay = s.recharge_boost_delay;
    m.enableTeamDemos = s.enable_team_demos != 0;
    m.goalBaseThresholdY = s.goal_base_threshold_y;
    return m;
}

// ArenaConfig (src/Sim/Arena/ArenaConfig/ArenaConfig.h) onto a BatchConfig:
// noBallRot and the custom boost pads carry over; a broadphase other than the
// default grid cannot be simulated and throws, as RS_ERR_CLOSE would.
inline void ApplyArenaConfig(const RocketSim::ArenaConfig& a, BatchConfig& cfg) {
    const RocketSim::ArenaConfig def;
    if (!a.useCustomBroadphase || a.maxAABBLen != def.maxAABBLen || a.minPos.x != def.minPos.x ||
        a.minPos.y != def.minPos.y || a.minPos.z != def.minPos.z || a.maxPos.x != def.maxPos.x ||
        a.maxPos.y != def.maxPos.y || a.maxPos.z != def.maxPos.z)
        throw std::runtime_error("rsg: only the default broadphase grid of ArenaConfig is supported");
    cfg.ball_rotation = !a.noBallRot;
    cfg.custom_pads.clear();
    if (a.useCustomBoostPads)
        for (const auto& p : a.customBoostPads) {
            rsg_boost_pad_config c{};
            c.pos[0] = p.pos.x; c.pos[1] = p.pos.y; c.pos[2] = p.pos.z;
            c.is_big = p.isBig ? 1 : 0;
            cfg.custom_pads.push_back(c);
        }
}

}  // namespace rsg_amd

#endif  // RSG_ROCKETSIM_HPP
