// rsg_sim.h -- per-arena tick of the batched Soccar step: car game logic,
// raycast vehicle, rigid-body integration and the solver's body write-back.
//
// Everything here is __host__ __device__ so the kernels in rsg_batch.hip can
// run it per arena (one lane per arena) and per phase (one lane per car /
// wheel / pair).  The collision phases (world queries, narrowphase, contact
// solver) live in rsg_collide.h and rsg_solver.h.
//
// Semantics follow the reference tick, Arena::Step
// (src/Sim/Arena/Arena.cpp:689-756), in the same order:
//   1 ball sleep flag            Arena.cpp:694-700
//   2 car pre-tick               Car.cpp:58-131 (+ btVehicleRL.cpp:120-414)
//   3 pads pre-tick              BoostPad.cpp:51-60
//   4 world step                 btDiscreteDynamicsWorld.cpp:325-437
//   5 car post-tick / finish     Car.cpp:133-193, BoostPadGrid.cpp:5-25
//   6 pads post-tick             BoostPad.cpp:88-107
//   7 ball finish                Ball.cpp:117-143
//   8 goal test                  Arena.cpp:748-752, 893-901
#pragma once

#include "rsg_const.h"
#include "rsg_layout.h"
#include "../../include/rsg_state.h"
#include "../../include/rsg_workload.h"

namespace rsg {

struct Ctrl {
    float throttle, steer, pitch, yaw, roll;
    bool jump, boost, handbrake;
};

// Rigid body working copy (Bullet units).
struct Body {
    V3 pos;
    M3 rot;
    V3 vel, ang;
    V3 force, torque;
    M3 inv_iw;         // world inverse inertia, R diag(invI) R^T
    V3 inv_i_local;
    float inv_mass;
    V3 grav;           // gravity force (btRigidBody::m_gravity: acceleration times mass)
    bool active;       // takes part in gravity / solver / integration
    bool no_rot;       // ball: translation-only integration
};

RSG_FN void body_update_inertia(Body& b) {
    b.inv_iw = scaled(b.rot, b.inv_i_local) * transpose(b.rot);
}

RSG_FN void body_apply_central_force(Body& b, V3 f) { b.force += f; }
RSG_FN void body_apply_torque(Body& b, V3 t) { b.torque += t; }
RSG_FN void body_apply_central_impulse(Body& b, V3 imp) { b.vel += imp * b.inv_mass; }
RSG_FN void body_apply_torque_impulse(Body& b, V3 t) { b.ang += b.inv_iw * t; }
RSG_FN void body_apply_impulse(Body& b, V3 imp, V3 rel_pos) {
    if (b.inv_mass != 0.f) {
        body_apply_central_impulse(b, imp);
        body_apply_torque_impulse(b, cross(rel_pos, imp));
    }
}
RSG_FN V3 body_vel_at(const Body& b, V3 rel) { return b.vel + cross(b.ang, rel); }

// Exponential-map pose step (btTransformUtil::integrateTransform,
// LinearMath/btTransformUtil.h:37-87; Grassia 1998).
RSG_FN void integrate_transform(V3 pos, const M3& rot, V3 linvel, V3 angvel, float dt, bool no_rot,
                                V3& out_pos, M3& out_rot) {
    out_pos = pos + linvel * dt;
    if (no_rot) { out_rot = rot; return; }
    float a2 = len2(angvel);
    float a = 0.f;
    if (a2 > FLT_EPS) a = sqrtf(a2);
    if (a * dt > ANGULAR_MOTION_THRESHOLD) a = ANGULAR_MOTION_THRESHOLD / dt;
    V3 axis;
    if (a < 0.001f)
        axis = angvel * (0.5f * dt - (dt * dt * dt) * (0.020833333333f) * a * a);
    else
        axis = angvel * (sinf(0.5f * a * dt) / a);
    Q4 dorn = q4(axis.x, axis.y, axis.z, cosf(a * dt * 0.5f));
    Q4 orn0 = m3_to_quat(rot);
    Q4 pred = qmul(dorn, orn0);
    if (qlen2(pred) > FLT_EPS) pred = qnormalized(pred);
    if (qlen2(pred) > FLT_EPS)
        out_rot = quat_to_m3(pred);
    else
        out_rot = rot;
}

// Per-wheel values that only live within one tick.
struct WheelT {
    V3 hard_point, dir_ws, axle_ws;
    M3 basis;           // wheel world basis (includes steering)
    bool in_contact, with_world;
    int ground;         // -1 none, -2 static world, >=0 dynamic body index
    V3 contact_point, contact_normal;
    float susp_len, susp_rel_vel, clipped_inv;
    V3 impulse;
    float susp_force;
};

struct Car {
    Body b;
    M3 api_rot;        // CarState::rotMat: the body basis as of the last post-tick
                       // update of a car that was not demolished
    int preset, team;  // preset: the car's entry in the batch's config table
    CarPhys phys;      // its CarConfig's physics values
    uint32_t id;
    uint32_t flags;
    V3 flip_torque;
    float jump_time, flip_time, air_time, air_time_since_jump;
    float boost, time_since_boosted, boosting_time, supersonic_time, handbrake;
    float af_timer, af_scale;
    V3 world_normal;
    uint32_t other_id;
    float car_cooldown, demo_timer;
    V3 hit_rel_pos, hit_ball_pos, hit_extra_vel;
    uint64_t hit_tick, hit_extra_tick, ticks_since_set;
    Ctrl last, ctrl;
    float engine, brake, steer;
    float lat[4], lon[4], push[4];
    uint32_t gen_redraw, gen_left;
    V3 vel_cache;
};

struct Ball {
    Body b;
    float radius_uu, mass;  // the body's sphere and mass as last made
    uint64_t ticks_since_set;
    V3 vel_cache;
    // BallState::HeatseekerInfo (src/Sim/Ball/Ball.h:24-31)
    float hs_dir, hs_speed, hs_time;
};

struct Pad {
    bool active;
    float cooldown;
    uint32_t prev_locked;
    int cur_locked;     // car index locked this tick, -1 none
};

RSG_FN bool flag(uint32_t f, uint32_t bit) { return (f & bit) != 0; }
RSG_FN void setflag(uint32_t& f, uint32_t bit, bool v) { f = v ? (f | bit) : (f & ~bit); }

RSG_FN V3 car_forward(const Car& c) { return col(c.b.rot, 0); }
RSG_FN V3 car_right(const Car& c) { return col(c.b.rot, 1); }
RSG_FN V3 car_up(const Car& c) { return col(c.b.rot, 2); }

// ---------------------------------------------------------------------------
// World query interface used by the vehicle.  Implemented in rsg_collide.h.
struct RayHit {
    bool hit;
    bool is_static;
    int body;          // dynamic body index (0 = ball, 1 + i = car i), or -1
    V3 point, normal;  // normal already normalised
    float fraction;
};

// ---------------------------------------------------------------------------
// Vehicle (src/Sim/btVehicleRL/btVehicleRL.cpp)

// updateWheelTransformsWS (btVehicleRL.cpp:106-117)
RSG_FN void wheel_transforms_ws(const Car& c, const CarPhys& cp, int i, WheelT& w) {
    w.in_contact = false;
    w.with_world = false;
    w.hard_point = c.b.rot * cp.wheel_conn[i] + c.b.pos;
    w.dir_ws = c.b.rot * v3(0.f, 0.f, -1.f);
    w.axle_ws = c.b.rot * v3(0.f, -1.f, 0.f);
}

// updateWheelTransform (btVehicleRL.cpp:67-96), steering from the previous
// tick's m_steerAngle.  Coordinate system: right = 1, up = 2, forward = 0.
RSG_FN void wheel_update_transform(const Car& c, const CarPhys& cp, int i, WheelT& w) {
    wheel_transforms_ws(c, cp, i, w);
    V3 up = -w.dir_ws;
    V3 right = w.axle_ws;
    V3 fwd = normalized(cross(up, right));
    float steer = (i < 2) ? c.steer : 0.f;
    M3 steer_m = quat_to_m3(q_axis_angle(up, steer));
    M3 basis2;
    basis2.r[0] = v3(fwd.x, -right.x, up.x);
    basis2.r[1] = v3(fwd.y, -right.y, up.y);
    basis2.r[2] = v3(fwd.z, -right.z, up.z);
    w.basis = steer_m * basis2;
}

struct SolverInfo {
    float erp;        // 0.2 (btContactSolverInfo default)
    float time_step;  // timestep of the last world step
};

// resolveSingleCollision against a static object with applyImpulses = false
// (BulletDynamics/ConstraintSolver/btContactConstraint.cpp:60-103).
RSG_FN float resolve_single_collision_static(const Body& b1, V3 contact, V3 normal, float distance,
                                             const SolverInfo& si) {
    V3 rel1 = contact - b1.pos;
    V3 vel = body_vel_at(b1, rel1) - v3zero();
    float rel_vel = dot(normal, vel);
    float restitution = 0.f * -rel_vel;
    float positional_error = si.erp * -distance / si.time_step;
    float velocity_error = -(1.0f + restitution) * rel_vel;
    V3 r0 = contact - b1.pos;
    V3 c0 = cross(r0, normal);
    V3 vec = cross(mul_vm(c0, b1.inv_iw), r0);
    float denom0 = b1.inv_mass + dot(normal, vec);
    float denom1 = 0.f;
    float jac = 1.f / (denom0 + denom1);
    float imp = positional_error * jac + velocity_error * jac;
    return 0.f > imp ? 0.f : imp;
}

// resolveSingleBilateral (btContactConstraint.cpp:106-150): returns the
// impulse; `other` may be null for a static ground.
RSG_FN float resolve_single_bilateral(const Body& b1, V3 pos1, const Body* other, V3 pos2, V3 normal) {
    float nl2 = len2(normal);
    if (nl2 > 1.1f) return 0.f;
    V3 rel1 = pos1 - b1.pos;
    V3 rel2 = other ? pos2 - other->pos : pos2;
    V3 vel1 = body_vel_at(b1, rel1);
    V3 vel2 = other ? body_vel_at(*other, rel2) : v3zero();
    V3 vel = vel1 - vel2;
    V3 aJ = transpose(b1.rot) * cross(rel1, normal);
    V3 m0 = b1.inv_i_local * aJ;
    float diag = b1.inv_mass + dot(m0, aJ);
    if (other) {
        V3 bJ = transpose(other->rot) * cross(rel2, -normal);
        V3 m1 = other->inv_i_local * bJ;
        diag = (diag + other->inv_mass) + dot(m1, bJ);
    } else {
        diag = (diag + 0.f) + 0.f;
    }
    float jac_inv = 1.f / diag;
    float rel_vel = dot(normal, vel);
    return -0.2f * rel_vel * jac_inv;
}

// ---------------------------------------------------------------------------
// Car game logic (src/Sim/Car/Car.cpp)

// Euler yaw/pitch/roll of a basis (btMatrix3x3::getEulerYPR,
// LinearMath/btMatrix3x3.h:527-547) with RocketSim's sign flips
// (Angle::FromRotMat, src/Math/MathTypes/MathTypes.cpp:62-70).  Only roll is
// used by the auto-flip.
RSG_FN float basis_roll(const M3& m) {
    float pitch = asinf(clampf(-m.r[2].x, -1.f, 1.f));
    float roll = atan2f(m.r[2].y, m.r[2].z);
    if (fabsf(pitch) == 1.57079632679489661923f) {
        if (roll > 0) roll -= 3.14159265358979323846f; else roll += 3.14159265358979323846f;
    }
    return -roll;
}

// Angle(yaw, 0, 0).ToRotMat(): setEulerZYX(0, 0, yaw) (MathTypes.cpp:72-78,
// btMatrix3x3.h:304-321).
RSG_FN M3 yaw_basis(float yaw) {
    float ci = cosf(-0.f), cj = cosf(-0.f), ch = cosf(yaw);
    float si = sinf(-0.f), sj = sinf(-0.f), sh = sinf(yaw);
    float cc = ci * ch, cs = ci * sh, sc = si * ch, ss = si * sh;
    return m3(v3(cj * ch, sj * sc - cs, sj * cc + ss),
              v3(cj * sh, sj * ss + cc, sj * cs - sc),
              v3(-sj, cj * si, cj * ci));
}

RSG_FN V3 wheel_upwards_dir(const Car& c, const WheelT* wh) {
    V3 sum = v3zero();
    for (int i = 0; i < 4; i++)
        if (wh[i].in_contact) sum += wh[i].contact_normal;
    if (sum == v3zero()) return car_up(c);
    return safe_normalized(sum);
}

// _UpdateWheels (Car.cpp:339-484)
RSG_FN void car_update_wheels(Car& c, WheelT* wh, const CarPhys& cp, float dt, int n_contact, float fwd_speed_uu) {
    float abs_fwd = fabsf(fwd_speed_uu);
    bool wheels_world = false;
    for (int i = 0; i < 4; i++) wheels_world |= wh[i].with_world;

    if (c.ctrl.handbrake) c.handbrake += POWERSLIDE_RISE_RATE * dt;
    else c.handbrake -= POWERSLIDE_FALL_RATE * dt;
    c.handbrake = rs_min(rs_max(c.handbrake, 0.f), 1.f);

    float real_throttle = c.ctrl.throttle;
    float real_brake = 0;
    if (c.ctrl.boost && c.boost > 0) real_throttle = 1;

    {
        float drive_scale = curve_eval(curve_drive_torque(), abs_fwd);
        float engine_throttle = real_throttle;
        if (c.ctrl.handbrake) {
        } else {
            float abs_thr = fabsf(real_throttle);
            if (abs_thr >= THROTTLE_DEADZONE) {
                if (abs_fwd > STOPPING_FORWARD_VEL && rs_sgn(real_throttle) != rs_sgn(fwd_speed_uu)) {
                    real_brake = 1;
                    if (abs_fwd > BRAKING_NO_THROTTLE_SPEED_THRESH) engine_throttle = 0;
                }
            } else {
                engine_throttle = 0;
                bool full_stop = (abs_fwd < STOPPING_FORWARD_VEL);
                real_brake = full_stop ? 1 : COASTING_BRAKE_FACTOR;
            }
        }
        if (n_contact < 3) drive_scale /= 4;
        c.engine = engine_throttle * (THROTTLE_TORQUE_AMOUNT * UU_TO_BT) * drive_scale;
        c.brake = real_brake * (BRAKE_TORQUE_AMOUNT * UU_TO_BT);
    }

    {
        float steer_angle = curve_eval(cp.three_wheels ? curve_steer_angle_three_wheels() : curve_steer_angle(), abs_fwd);
        if (c.handbrake != 0.f)
            steer_angle += (curve_eval(curve_powerslide_steer(), abs_fwd) - steer_angle) * c.handbrake;
        steer_angle *= c.ctrl.steer;
        c.steer = steer_angle;
    }

    for (int i = 0; i < 4; i++) {
        WheelT& w = wh[i];
        if (w.ground != -1) {
            V3 lat_dir = col(w.basis, 1);
            V3 long_dir = cross(lat_dir, w.contact_normal);
            float curve_in = 0;
            V3 delta = w.hard_point - c.b.pos;
            V3 cv = (cross(c.b.ang, delta) + c.b.vel) * BT_TO_UU;
            float base_fr = fabsf(dot(cv, lat_dir));
            if (base_fr > 5) curve_in = base_fr / (fabsf(dot(cv, long_dir)) + base_fr);
            float lat_f = curve_eval(cp.three_wheels ? curve_lat_friction_three_wheels() : curve_lat_friction(), curve_in);
            float long_f = curve_eval(curve_long_friction(), curve_in);
            if (c.handbrake != 0.f) {
                float hb = c.handbrake;
                lat_f *= (curve_eval(curve_handbrake_lat(), curve_in) - 1) * hb + 1;
                long_f *= (curve_eval(curve_handbrake_long(), curve_in) - 1) * hb + 1;
            } else {
                long_f = 1;
            }
            bool sticky = real_throttle != 0;
            if (!sticky) {
                float ns = curve_eval(curve_non_sticky(), w.contact_normal.z);
                lat_f *= ns;
                long_f *= ns;
            }
            c.lat[i] = lat_f;
            c.lon[i] = long_f;
        }
    }

    if (wheels_world) {
        V3 up_dir = wheel_upwards_dir(c, wh);
        bool full_stick = (real_throttle != 0) || (abs_fwd > STOPPING_FORWARD_VEL);
        float sticky_scale = cp.three_wheels ? 0.f : 0.5f;
        if (full_stick) sticky_scale += 1 - fabsf(up_dir.z);
        body_apply_central_force(c.b, up_dir * sticky_scale * (GRAVITY_Z * UU_TO_BT) * CAR_MASS_BT);
    }
}

// _UpdateBoost (Car.cpp:486-533), Soccar mutators (no recharge).
RSG_FN void car_update_boost(Car& c, float dt, const rsg_mutator_config& M) {
    bool has_boost = c.boost > 0;
    bool boosting = flag(c.flags, RSG_CF_IS_BOOSTING);
    if (has_boost) {
        if (boosting) {
            boosting = c.ctrl.boost || c.boosting_time < BOOST_MIN_TIME;
        } else {
            if (c.ctrl.boost) boosting = true;
        }
    } else {
        boosting = false;
    }
    setflag(c.flags, RSG_CF_IS_BOOSTING, boosting);
    if (boosting) c.boosting_time += dt; else c.boosting_time = 0;
    if (boosting) {
        c.boost = rs_max(c.boost - M.boost_used_per_second * dt, 0.f);
        float accel = flag(c.flags, RSG_CF_ON_GROUND) ? M.boost_accel_ground : M.boost_accel_air;
        body_apply_central_force(c.b, accel * UU_TO_BT * car_forward(c) * CAR_MASS_BT);
        c.time_since_boosted = 0;
    } else {
        c.time_since_boosted += dt;
        // recharge (MutatorConfig::rechargeBoostEnabled)
        if (M.recharge_boost_enabled && c.time_since_boosted >= M.recharge_boost_delay)
            c.boost += M.recharge_boost_per_second * dt;
    }
    c.boost = rs_min(c.boost, BOOST_MAX);
}

// _UpdateJump (Car.cpp:535-582)
RSG_FN void car_update_jump(Car& c, float dt, bool jump_pressed, const rsg_mutator_config& M) {
    bool on_ground = flag(c.flags, RSG_CF_ON_GROUND);
    bool is_jumping = flag(c.flags, RSG_CF_IS_JUMPING);
    bool has_jumped = flag(c.flags, RSG_CF_HAS_JUMPED);
    if (on_ground && !is_jumping) {
        if (has_jumped && c.jump_time < JUMP_MIN_TIME + JUMP_RESET_TIME_PAD) {
        } else {
            has_jumped = false;
            c.jump_time = 0;
        }
    }
    if (is_jumping) {
        is_jumping = c.jump_time < JUMP_MIN_TIME || (c.ctrl.jump && c.jump_time < JUMP_MAX_TIME);
    } else if (on_ground && jump_pressed) {
        is_jumping = true;
        c.jump_time = 0;
        V3 f = car_up(c) * M.jump_immediate_force * UU_TO_BT * CAR_MASS_BT;
        body_apply_central_impulse(c.b, f);
    }
    if (is_jumping) {
        has_jumped = true;
        V3 total = car_up(c) * M.jump_accel;
        if (c.jump_time < JUMP_MIN_TIME) total = total * JUMP_PRE_MIN_ACCEL_SCALE;
        body_apply_central_force(c.b, total * UU_TO_BT * CAR_MASS_BT);
    }
    if (is_jumping || has_jumped) c.jump_time += dt;
    setflag(c.flags, RSG_CF_IS_JUMPING, is_jumping);
    setflag(c.flags, RSG_CF_HAS_JUMPED, has_jumped);
}

// _UpdateAirTorque (Car.cpp:584-669)
RSG_FN void car_update_air_torque(Car& c, float dt, bool update_air_control) {
    V3 dir_pitch = -car_right(c), dir_yaw = car_up(c), dir_roll = -car_forward(c);
    bool do_air = false;
    bool flipping = flag(c.flags, RSG_CF_IS_FLIPPING);
    if (flipping) flipping = flag(c.flags, RSG_CF_HAS_FLIPPED) && c.flip_time < FLIP_TORQUE_TIME;
    setflag(c.flags, RSG_CF_IS_FLIPPING, flipping);
    if (flipping) {
        V3 rel = c.flip_torque;
        if (!is_zero(c.flip_torque)) {
            float pitch_scale = 1;
            if (rel.y != 0 && c.ctrl.pitch != 0) {
                if (rs_sgn(rel.y) == rs_sgn(c.ctrl.pitch)) {
                    pitch_scale = 1 - rs_min(fabsf(c.ctrl.pitch), 1.f);
                    do_air = true;
                }
            }
            rel.y *= pitch_scale;
            V3 dodge = rel * v3(FLIP_TORQUE_X, FLIP_TORQUE_Y, 0);
            body_apply_torque(c.b, inverse(c.b.inv_iw) * (c.b.rot * dodge));
        } else {
            do_air = true;
        }
    } else {
        do_air = true;
    }
    do_air &= !flag(c.flags, RSG_CF_AUTO_FLIPPING);
    do_air &= update_air_control;
    if (do_air) {
        float pitch_scale = 1;
        V3 torque;
        if (c.ctrl.pitch != 0.f || c.ctrl.yaw != 0.f || c.ctrl.roll != 0.f) {
            if (flipping) pitch_scale = 0;
            else if (flag(c.flags, RSG_CF_HAS_FLIPPED)) {
                if (c.flip_time < FLIP_TORQUE_TIME + FLIP_PITCHLOCK_EXTRA_TIME) pitch_scale = 0;
            }
            torque = (c.ctrl.pitch * dir_pitch * pitch_scale * AIR_TORQUE_P) +
                     (c.ctrl.yaw * dir_yaw * AIR_TORQUE_Y) +
                     (c.ctrl.roll * dir_roll * AIR_TORQUE_R);
        } else {
            torque = v3zero();
        }
        V3 av = c.b.ang;
        float damp_p = dot(dir_pitch, av) * AIR_DAMP_P * (1 - fabsf(c.ctrl.pitch * pitch_scale));
        float damp_y = dot(dir_yaw, av) * AIR_DAMP_Y * (1 - fabsf(c.ctrl.yaw));
        float damp_r = dot(dir_roll, av) * AIR_DAMP_R;
        V3 damping = (dir_yaw * damp_y) + (dir_pitch * damp_p) + (dir_roll * damp_r);
        body_apply_torque(c.b, inverse(c.b.inv_iw) * (torque - damping) * CAR_TORQUE_SCALE);
    }
    if (c.ctrl.throttle != 0)
        body_apply_central_force(c.b, car_forward(c) * c.ctrl.throttle * THROTTLE_AIR_ACCEL * UU_TO_BT * CAR_MASS_BT);
}

// _UpdateDoubleJumpOrFlip (Car.cpp:671-781)
RSG_FN void car_update_double_jump_or_flip(Car& c, float dt, bool jump_pressed, float fwd_speed_uu,
                                            const rsg_mutator_config& M) {
    if (flag(c.flags, RSG_CF_ON_GROUND)) {
        setflag(c.flags, RSG_CF_HAS_DOUBLE_JUMPED, false);
        setflag(c.flags, RSG_CF_HAS_FLIPPED, false);
        c.air_time = 0;
        c.air_time_since_jump = 0;
        c.flip_time = 0;
    } else {
        c.air_time += dt;
        if (flag(c.flags, RSG_CF_HAS_JUMPED) && !flag(c.flags, RSG_CF_IS_JUMPING))
            c.air_time_since_jump += dt;
        else
            c.air_time_since_jump = 0;

        if (jump_pressed && c.air_time_since_jump < DOUBLEJUMP_MAX_DELAY) {
            float in_mag = fabsf(c.ctrl.yaw) + fabsf(c.ctrl.pitch) + fabsf(c.ctrl.roll);
            bool is_flip = in_mag >= c.phys.dodge_deadzone;
            bool can_use = (!flag(c.flags, RSG_CF_HAS_DOUBLE_JUMPED) && !flag(c.flags, RSG_CF_HAS_FLIPPED)) ||
                           (is_flip ? M.unlimited_flips : M.unlimited_double_jumps);
            if (flag(c.flags, RSG_CF_AUTO_FLIPPING)) can_use = false;
            if (can_use) {
                if (is_flip) {
                    c.flip_time = 0;
                    setflag(c.flags, RSG_CF_HAS_FLIPPED, true);
                    setflag(c.flags, RSG_CF_IS_FLIPPING, true);
                    float fwd_ratio = fabsf(fwd_speed_uu) / CAR_MAX_SPEED;
                    V3 dodge = v3(-c.ctrl.pitch, c.ctrl.yaw + c.ctrl.roll, 0);
                    if (fabsf(c.ctrl.yaw + c.ctrl.roll) < 0.1f && fabsf(c.ctrl.pitch) < 0.1f)
                        dodge = v3zero();
                    else
                        dodge = safe_normalized(dodge);
                    c.flip_torque = v3(-dodge.y, dodge.x, 0);
                    if (fabsf(dodge.x) < 0.1f) dodge.x = 0;
                    if (fabsf(dodge.y) < 0.1f) dodge.y = 0;
                    if (!fuzzy_zero(dodge)) {
                        bool backwards;
                        if (fabsf(fwd_speed_uu) < 100.0f) backwards = dodge.x < 0.0f;
                        else backwards = (dodge.x >= 0.0f) != (fwd_speed_uu >= 0.0f);
                        V3 init = dodge * FLIP_INITIAL_VEL_SCALE;
                        float max_x = backwards ? FLIP_BACKWARD_IMPULSE_MAX_SPEED_SCALE : FLIP_FORWARD_IMPULSE_MAX_SPEED_SCALE;
                        init.x *= ((max_x - 1) * fwd_ratio) + 1.f;
                        init.y *= ((FLIP_SIDE_IMPULSE_MAX_SPEED_SCALE - 1) * fwd_ratio) + 1.f;
                        if (backwards) init.x *= FLIP_BACKWARD_IMPULSE_SCALE_X;
                        V3 f = car_forward(c);
                        V3 f2 = rs_normalized(v3(f.x, f.y, 0));
                        V3 r2 = v3(-f2.y, f2.x, 0.f);
                        V3 dv = init.x * f2 + init.y * r2;
                        body_apply_central_impulse(c.b, dv * UU_TO_BT * CAR_MASS_BT);
                    }
                } else {
                    V3 f = car_up(c) * JUMP_IMMEDIATE_FORCE * UU_TO_BT * CAR_MASS_BT;
                    body_apply_central_impulse(c.b, f);
                    setflag(c.flags, RSG_CF_HAS_DOUBLE_JUMPED, true);
                }
            }
        }
    }
    if (flag(c.flags, RSG_CF_IS_FLIPPING)) {
        c.flip_time += dt;
        if (c.flip_time <= FLIP_TORQUE_TIME) {
            if (c.flip_time >= FLIP_Z_DAMP_START && (c.b.vel.z < 0 || c.flip_time < FLIP_Z_DAMP_END)) {
                c.b.vel.z *= powf(1 - FLIP_Z_DAMP_120, dt / (1 / 120.f));
            }
        }
    } else if (flag(c.flags, RSG_CF_HAS_FLIPPED)) {
        c.flip_time += dt;
    }
}

// _UpdateAutoFlip (Car.cpp:783-817)
RSG_FN void car_update_auto_flip(Car& c, float dt, bool jump_pressed) {
    if (jump_pressed && flag(c.flags, RSG_CF_WORLD_CONTACT) && c.world_normal.z > CAR_AUTOFLIP_NORMZ_THRESH) {
        float roll = basis_roll(c.b.rot);
        float abs_roll = fabsf(roll);
        if (abs_roll > CAR_AUTOFLIP_ROLL_THRESH) {
            c.af_timer = CAR_AUTOFLIP_TIME * (abs_roll / (float)PI_D);
            c.af_scale = (roll > 0) ? 1 : -1;
            setflag(c.flags, RSG_CF_AUTO_FLIPPING, true);
            body_apply_central_impulse(c.b, -car_up(c) * CAR_AUTOFLIP_IMPULSE * UU_TO_BT * CAR_MASS_BT);
        }
    }
    if (flag(c.flags, RSG_CF_AUTO_FLIPPING)) {
        if (c.af_timer <= 0) {
            setflag(c.flags, RSG_CF_AUTO_FLIPPING, false);
            c.af_timer = 0;
        } else {
            c.b.ang += car_forward(c) * CAR_AUTOFLIP_TORQUE * c.af_scale * dt;
            c.af_timer -= dt;
        }
    }
}

// _UpdateAutoRoll (Car.cpp:819-853)
RSG_FN void car_update_auto_roll(Car& c, WheelT* wh, int n_contact) {
    V3 ground_up = n_contact > 0 ? wheel_upwards_dir(c, wh) : c.world_normal;
    V3 ground_down = -ground_up;
    V3 fwd = car_forward(c), right = car_right(c);
    V3 cross_right = cross(ground_up, fwd);
    V3 cross_fwd = cross(ground_down, cross_right);
    float right_tf = 1 - clampf(dot(right, cross_right), 0, 1);
    float fwd_tf = 1 - clampf(dot(fwd, cross_fwd), 0, 1);
    V3 tdir_right = fwd * (dot(right, ground_up) >= 0 ? -1.f : 1.f);
    V3 tdir_fwd = right * (dot(fwd, ground_up) >= 0 ? 1.f : -1.f);
    V3 t_right = tdir_right * right_tf;
    V3 t_fwd = tdir_fwd * fwd_tf;
    body_apply_central_force(c.b, ground_down * CAR_AUTOROLL_FORCE * UU_TO_BT * CAR_MASS_BT);
    body_apply_torque(c.b, inverse(c.b.inv_iw) * (t_fwd + t_right) * CAR_AUTOROLL_TORQUE);
}

// btVehicleRL::calcFrictionImpulses (btVehicleRL.cpp:306-380)
template <class World>
RSG_FN void vehicle_calc_friction(Car& c, WheelT* wh, World& world, float dt) {
    float friction_scale = CAR_MASS_BT / 3;  // getMass() = 1 / inverseMass
    friction_scale = (1.f / c.b.inv_mass) / 3;
    for (int i = 0; i < 4; i++) {
        WheelT& w = wh[i];
        if (w.ground != -1) {
            const Body* ground = w.ground >= 0 ? world.body_ptr(w.ground) : nullptr;
            V3 axle = col(w.basis, 1);
            V3 n = w.contact_normal;
            float proj = dot(axle, n);
            axle -= n * proj;
            axle = safe_normalized(axle);
            V3 fwd = safe_normalized(cross(n, axle));
            float side = resolve_single_bilateral(c.b, w.contact_point, ground, w.contact_point, axle);
            float rolling;
            if (c.engine == 0) {
                if (c.brake != 0.f) {
                    V3 rel_car = w.contact_point - c.b.pos;
                    V3 v1 = body_vel_at(c.b, rel_car);
                    V3 v2 = ground ? body_vel_at(*ground, rel_car) : v3zero();
                    float rel_vel = dot(v1 - v2, fwd);
                    if (dt > (1 / 80.f)) {
                        float thr = -(1 / (dt * 150.f)) + 0.8f;
                        if (fabsf(rel_vel) < thr) rel_vel = 0;
                    }
                    rolling = clampf(-rel_vel * ROLLING_FRICTION_SCALE_MAGIC, -c.brake, c.brake);
                } else {
                    rolling = 0;
                }
            } else {
                rolling = -c.engine / friction_scale;
            }
            V3 total = (fwd * rolling * c.lon[i]) + (axle * side * c.lat[i]);
            w.impulse = total * friction_scale;
        } else {
            w.impulse = v3zero();
        }
    }
}

// The suspension ray of wheel i (btVehicleRL.cpp:123-131): from the wheel's
// hard point along the suspension direction, rest + travel + radius - 0.05.
RSG_FN void wheel_ray(const Car& c, const CarPhys& cp, int i, V3& source, V3& target) {
    WheelT w;
    wheel_transforms_ws(c, cp, i, w);
    float ray_len = cp.wheel_rest[i] + cp.max_susp_travel + cp.wheel_radius[i] - SUSPENSION_SUBTRACTION;
    source = w.hard_point;
    target = source + (w.dir_ws * ray_len);
}

// btVehicleRL::rayCast (btVehicleRL.cpp:120-209).  `pre`, when given, is this
// wheel's ray result cast earlier in the same pre-tick (nothing it depends on
// has changed since).
template <class World>
RSG_FN void vehicle_ray_cast(Car& c, WheelT* wh, const CarPhys& cp, int i, World& world, int self_body, const SolverInfo& si,
                             const RayHit* pre) {
    WheelT& w = wh[i];
    wheel_transforms_ws(c, cp, i, w);
    float travel = cp.max_susp_travel;
    float rest = cp.wheel_rest[i];
    float radius = cp.wheel_radius[i];
    float ray_len = rest + travel + radius - SUSPENSION_SUBTRACTION;
    V3 source = w.hard_point;
    V3 target = source + (w.dir_ws * ray_len);
    w.contact_point = target;
    w.ground = -1;
    RayHit h = pre ? *pre : world.raycast(source, target, self_body);
    V3 up = car_up(c);
    if (h.hit) {
        w.contact_point = h.point;
        w.contact_normal = h.normal;
        w.in_contact = true;
        w.with_world = h.is_static;
        w.ground = h.is_static ? -2 : h.body;
        float trace = dot(w.hard_point - w.contact_point, up);
        w.susp_len = clampf(trace - radius, rest - travel, rest + travel);
        float denom = dot(w.contact_normal, up);
        V3 relpos = w.contact_point - c.b.pos;
        V3 vel_at = body_vel_at(c.b, relpos);
        float proj_vel = dot(w.contact_normal, vel_at);
        if (denom > 0.1f) {
            float inv = 1 / denom;
            w.susp_rel_vel = proj_vel * inv;
            w.clipped_inv = inv;
        } else {
            w.susp_rel_vel = 0;
            w.clipped_inv = 10;
        }
        if (h.is_static) {
            float thresh = (rest + radius) - SUSPENSION_SUBTRACTION;
            if (trace < thresh) {
                float delta = trace - thresh;
                float r = resolve_single_collision_static(c.b, h.point, h.normal, delta, si);
                c.push[i] = r / 4;
            }
        }
    } else {
        w.susp_len = rest + travel;
        w.susp_rel_vel = 0;
        w.contact_normal = -w.dir_ws;
        w.clipped_inv = 1;
        c.push[i] = 0;
    }
}

// btVehicleRL::updateSuspension + applyFrictionImpulses (btVehicleRL.cpp:270-395)
RSG_FN void vehicle_second(Car& c, WheelT* wh, const CarPhys& cp, float dt) {
    for (int i = 0; i < 4; i++) {
        WheelT& w = wh[i];
        if (w.in_contact) {
            float force = (cp.wheel_rest[i] - w.susp_len) * SUSPENSION_STIFFNESS * w.clipped_inv;
            float damp = (w.susp_rel_vel < 0) ? WHEELS_DAMPING_COMPRESSION : WHEELS_DAMPING_RELAXATION;
            w.susp_force = force - (damp * w.susp_rel_vel);
            w.susp_force *= cp.wheel_force_scale[i];
            if (w.susp_force < 0) w.susp_force = 0;
        } else {
            w.susp_force = 0;
        }
    }
    for (int i = 0; i < 4; i++) {
        WheelT& w = wh[i];
        if (w.susp_force != 0) {
            V3 off = w.contact_point - c.b.pos;
            float base = (w.susp_force * dt) + c.push[i];
            V3 f = w.contact_normal * base;
            body_apply_impulse(c.b, f, off);
        }
    }
    V3 up = car_up(c);
    for (int i = 0; i < 4; i++) {
        WheelT& w = wh[i];
        if (!is_zero(w.impulse)) {
            V3 off = w.contact_point - c.b.pos;
            float d = dot(up, off);
            V3 rel = off - up * d;
            body_apply_impulse(c.b, w.impulse * dt, rel);
        }
    }
}

// Car::Respawn (Car.cpp:43-56) at a given respawn slot of the mode's table.
RSG_FN void car_respawn_state(Car& c, int mode, int slot, float boost_amount) {
    SpawnPos sp = respawn_spawn_of(mode, slot);
    V3 pos_uu = v3(sp.x, sp.y * (c.team == RSG_TEAM_BLUE ? 1.f : -1.f), CAR_RESPAWN_Z);
    float yaw = (float)((double)sp.yaw + (c.team == RSG_TEAM_BLUE ? 0.0 : PI_D));
    // CarState() defaults then SetState (Car.cpp:23-36)
    Car keep = c;
    c = Car();
    c.b = keep.b;
    c.preset = keep.preset; c.team = keep.team; c.id = keep.id;
    c.phys = keep.phys;
    c.ctrl = keep.ctrl;
    c.gen_redraw = keep.gen_redraw; c.gen_left = keep.gen_left;
    // Per-wheel values persist across SetState in the reference.
    c.engine = keep.engine; c.brake = keep.brake; c.steer = keep.steer;
    for (int i = 0; i < 4; i++) { c.lat[i] = keep.lat[i]; c.lon[i] = keep.lon[i]; c.push[i] = keep.push[i]; }
    c.vel_cache = v3zero();
    c.b.pos = pos_uu * UU_TO_BT;
    c.b.rot = yaw_basis(yaw);
    c.api_rot = c.b.rot;
    c.b.vel = v3zero();
    c.b.ang = v3zero();
    body_update_inertia(c.b);
    // CarState default isOnGround = true; the body flag is not part of the
    // state SetState writes
    c.flags = RSG_CF_ON_GROUND | (keep.flags & RSG_CF_BODY_DISABLED);
    c.boost = boost_amount;
    c.hit_tick = ~0ull;
    c.hit_extra_tick = ~0ull;
    c.ticks_since_set = 0;
}

}  // namespace rsg
