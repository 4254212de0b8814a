// rsg_adjoint.h -- differentiable airborne car step and its hand-written
// adjoint (reverse mode), for batched air-control optimisation.
//
// The step is the reference's tick for a car with no wheel, hitbox or ball
// contact (Car::_PreTickUpdate in the air, src/Sim/Car/Car.cpp:58-131, then
// the world step and Car::_FinishPhysicsTick):
//   * air control (Car::_UpdateAirTorque, Car.cpp:584-667): torque
//       T = -P*Tp*r + Y*Ty*u - R*Tr*f,
//       D = Dy*(1-|Y|)*(u.w)*u + Dp*(1-|P|)*(r.w)*r + Dr*(f.w)*f,
//     applied as I_world*(T - D)*S, so the solver's angular step is
//       w' = w + (T - D)*S*dt;
//   * boost (Car::_UpdateBoost, Car.cpp:486-530) relaxed to b in [0, 1]:
//       v' = v + (g + b*A*f)*dt   (b = 1 is the reference's boosting car);
//   * integration (btTransformUtil::integrateTransform, exponential map):
//       p' = p + v'*dt,  q' = normalize(dq(w', dt) (x) q);
//   * speed caps after the step (Car::_FinishPhysicsTick): |v| <= 2300 uu/s,
//     |w| <= 5.5 rad/s.
// f, r, u are the columns (forward, right, up) of R(q).  The orientation is
// carried as a unit quaternion between ticks (the reference re-derives it
// from its basis each tick; the two agree to rounding).  Bullet's clamp of
// the per-step rotation angle (ANGULAR_MOTION_THRESHOLD, pi/4 per step) never
// binds below the 5.5 rad/s cap and is not modelled.
//
// The adjoint of one step is air_step_adjoint: given the step's input state
// and the adjoint of its output, it returns the adjoint of the input and the
// gradient of the controls.  Rollouts checkpoint the state every K ticks
// and recompute the ticks in between when going backwards.
#pragma once

#include "rsg_math.h"
#include "rsg_const.h"

namespace rsg {

// State in Bullet units: position, unit quaternion (x, y, z, w), linear and
// angular velocity.
struct AirState {
    V3 p;
    Q4 q;
    V3 v, w;
};

constexpr int AIR_STATE_WORDS = 13;  // p3 q4 v3 w3

struct AirCtrl {
    float pitch, yaw, roll, boost;
};

// Columns of R(q) for a unit quaternion.
RSG_FN void air_basis(Q4 q, V3& f, V3& r, V3& u) {
    float x = q.x, y = q.y, z = q.z, w = q.w;
    f = v3(1.f - 2.f * (y * y + z * z), 2.f * (x * y + z * w), 2.f * (z * x - w * y));
    r = v3(2.f * (x * y - w * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + w * x));
    u = v3(2.f * (z * x + y * w), 2.f * (y * z - w * x), 1.f - 2.f * (x * x + y * y));
}

RSG_FN float air_sgn(float a) { return a > 0.f ? 1.f : (a < 0.f ? -1.f : 0.f); }

// Rotation quaternion of angular velocity w over dt (integrateTransform's
// exponential map, with its Taylor branch below 0.001 rad/s).
RSG_FN Q4 air_dq(V3 w, float dt) {
    float th = len(w);
    V3 a;
    if (th < 0.001f)
        a = w * (0.5f * dt - (dt * dt * dt) * (0.020833333333f) * th * th);
    else
        a = w * (sinf(0.5f * th * dt) / th);
    return q4(a.x, a.y, a.z, cosf(th * dt * 0.5f));
}

RSG_FN V3 air_cap(V3 v, float vmax) {
    float l2 = len2(v);
    if (l2 > vmax * vmax) return v * (vmax / sqrtf(l2));
    return v;
}

RSG_FN void air_step(AirState& s, const AirCtrl& c, float dt) {
    V3 f, r, u;
    air_basis(s.q, f, r, u);
    V3 torque = r * (-c.pitch * AIR_TORQUE_P) + u * (c.yaw * AIR_TORQUE_Y) + f * (-c.roll * AIR_TORQUE_R);
    float ky = AIR_DAMP_Y * (1.f - fabsf(c.yaw)), kp = AIR_DAMP_P * (1.f - fabsf(c.pitch)), kr = AIR_DAMP_R;
    V3 damp = u * (ky * dot(u, s.w)) + r * (kp * dot(r, s.w)) + f * (kr * dot(f, s.w));
    V3 w1 = s.w + (torque - damp) * (CAR_TORQUE_SCALE * dt);
    V3 v1 = s.v + (v3(0.f, 0.f, GRAVITY_Z * UU_TO_BT) + f * (c.boost * BOOST_ACCEL_AIR * UU_TO_BT)) * dt;
    V3 p1 = s.p + v1 * dt;
    Q4 h = qmul(air_dq(w1, dt), s.q);
    float hl = sqrtf(qlen2(h));
    s.p = p1;
    s.q = qscale(h, 1.f / hl);
    s.v = air_cap(v1, CAR_MAX_SPEED * UU_TO_BT);
    s.w = air_cap(w1, CAR_MAX_ANG_SPEED);
}

// Adjoint of air_cap at input v: g is the output's adjoint.
RSG_FN V3 air_cap_adjoint(V3 v, float vmax, V3 g) {
    float l2 = len2(v);
    if (l2 > vmax * vmax) {
        float l = sqrtf(l2);
        V3 n = v * (1.f / l);
        return (g - n * dot(n, g)) * (vmax / l);
    }
    return g;
}

// q_bar contribution of the basis columns' adjoints (fb, rb, ub) at q.
RSG_FN Q4 air_basis_adjoint(Q4 q, V3 fb, V3 rb, V3 ub) {
    float x = q.x, y = q.y, z = q.z, w = q.w;
    Q4 g;
    g.x = dot(fb, v3(0.f, 2.f * y, 2.f * z)) + dot(rb, v3(2.f * y, -4.f * x, 2.f * w)) +
          dot(ub, v3(2.f * z, -2.f * w, -4.f * x));
    g.y = dot(fb, v3(-4.f * y, 2.f * x, -2.f * w)) + dot(rb, v3(2.f * x, 0.f, 2.f * z)) +
          dot(ub, v3(2.f * w, 2.f * z, -4.f * y));
    g.z = dot(fb, v3(-4.f * z, 2.f * w, 2.f * x)) + dot(rb, v3(-2.f * w, -4.f * z, 2.f * y)) +
          dot(ub, v3(2.f * x, 2.f * y, 0.f));
    g.w = dot(fb, v3(0.f, 2.f * z, -2.f * y)) + dot(rb, v3(-2.f * z, 0.f, 2.f * x)) +
          dot(ub, v3(2.f * y, -2.f * x, 0.f));
    return g;
}

// Adjoint of one air_step.  `s` is the step's input; `g` holds the adjoint of
// its output on entry and the adjoint of its input on return; the control
// gradient is added to `gc`.
RSG_FN void air_step_adjoint(const AirState& s, const AirCtrl& c, float dt, AirState& g, AirCtrl& gc) {
    // --- recompute the forward intermediates
    V3 f, r, u;
    air_basis(s.q, f, r, u);
    V3 torque = r * (-c.pitch * AIR_TORQUE_P) + u * (c.yaw * AIR_TORQUE_Y) + f * (-c.roll * AIR_TORQUE_R);
    float ky = AIR_DAMP_Y * (1.f - fabsf(c.yaw)), kp = AIR_DAMP_P * (1.f - fabsf(c.pitch)), kr = AIR_DAMP_R;
    float uw = dot(u, s.w), rw = dot(r, s.w), fw = dot(f, s.w);
    V3 damp = u * (ky * uw) + r * (kp * rw) + f * (kr * fw);
    const float S = CAR_TORQUE_SCALE * dt;
    V3 w1 = s.w + (torque - damp) * S;
    const float A = BOOST_ACCEL_AIR * UU_TO_BT;
    V3 v1 = s.v + (v3(0.f, 0.f, GRAVITY_Z * UU_TO_BT) + f * (c.boost * A)) * dt;
    Q4 a = air_dq(w1, dt);
    Q4 h = qmul(a, s.q);
    float hl = sqrtf(qlen2(h));
    Q4 q1 = qscale(h, 1.f / hl);

    // --- speed caps
    V3 v1b = air_cap_adjoint(v1, CAR_MAX_SPEED * UU_TO_BT, g.v);
    V3 w1b = air_cap_adjoint(w1, CAR_MAX_ANG_SPEED, g.w);

    // --- p' = p + v1 dt
    V3 pb = g.p;
    v1b = v1b + g.p * dt;

    // --- q' = normalize(h), h = a (x) q
    Q4 qb1 = g.q;
    float qd = q1.x * qb1.x + q1.y * qb1.y + q1.z * qb1.z + q1.w * qb1.w;
    Q4 hb = q4((qb1.x - q1.x * qd) / hl, (qb1.y - q1.y * qd) / hl, (qb1.z - q1.z * qd) / hl,
               (qb1.w - q1.w * qd) / hl);
    Q4 q = s.q;
    Q4 ab;
    ab.x = q.w * hb.x - q.z * hb.y + q.y * hb.z - q.x * hb.w;
    ab.y = q.z * hb.x + q.w * hb.y - q.x * hb.z - q.y * hb.w;
    ab.z = -q.y * hb.x + q.x * hb.y + q.w * hb.z - q.z * hb.w;
    ab.w = q.x * hb.x + q.y * hb.y + q.z * hb.z + q.w * hb.w;
    Q4 qb;
    qb.x = a.w * hb.x + a.z * hb.y - a.y * hb.z - a.x * hb.w;
    qb.y = -a.z * hb.x + a.w * hb.y + a.x * hb.z - a.y * hb.w;
    qb.z = a.y * hb.x - a.x * hb.y + a.w * hb.z - a.z * hb.w;
    qb.w = a.x * hb.x + a.y * hb.y + a.z * hb.z + a.w * hb.w;

    // --- a = dq(w1): vector part k(th)*w1, scalar cos(th dt / 2)
    {
        float th = len(w1);
        V3 av = v3(ab.x, ab.y, ab.z);
        if (th < 0.001f) {
            float k = 0.5f * dt - (dt * dt * dt) * (0.020833333333f) * th * th;
            // d(k w1)/dw1 = k I - (dt^3 / 24) w1 w1^T ; dc/dw1 = -sin(th dt/2) dt/2 w1/th ~ -(dt^2/4) w1
            w1b = w1b + av * k - w1 * ((dt * dt * dt) * (2.f * 0.020833333333f) * dot(w1, av));
            w1b = w1b - w1 * (ab.w * (0.25f * dt * dt));
        } else {
            float sn = sinf(0.5f * th * dt), cs = cosf(0.5f * th * dt);
            float k = sn / th;
            float dk = (0.5f * dt * cs * th - sn) / (th * th);  // dk/dth
            V3 n = w1 * (1.f / th);
            w1b = w1b + av * k + n * (dk * dot(w1, av));
            w1b = w1b - n * (ab.w * sn * 0.5f * dt);
        }
    }

    // --- v1 = v + (g + b A f) dt
    V3 vb = v1b;
    V3 fb = v1b * (c.boost * A * dt);
    gc.boost += dot(f, v1b) * A * dt;

    // --- w1 = w + (torque - damp) S
    V3 wb = w1b;
    V3 tb = w1b * S, db = w1b * (-S);
    // torque = -P Tp r + Y Ty u - R Tr f
    gc.pitch += -AIR_TORQUE_P * dot(r, tb);
    gc.yaw += AIR_TORQUE_Y * dot(u, tb);
    gc.roll += -AIR_TORQUE_R * dot(f, tb);
    V3 rb = tb * (-c.pitch * AIR_TORQUE_P);
    V3 ub = tb * (c.yaw * AIR_TORQUE_Y);
    fb = fb + tb * (-c.roll * AIR_TORQUE_R);
    // damp = ky (u.w) u + kp (r.w) r + kr (f.w) f
    float ud = dot(u, db), rd = dot(r, db), fd = dot(f, db);
    wb = wb + u * (ky * ud) + r * (kp * rd) + f * (kr * fd);
    ub = ub + db * (ky * uw) + s.w * (ky * ud);
    rb = rb + db * (kp * rw) + s.w * (kp * rd);
    fb = fb + db * (kr * fw) + s.w * (kr * fd);
    gc.yaw += -AIR_DAMP_Y * air_sgn(c.yaw) * uw * ud;
    gc.pitch += -AIR_DAMP_P * air_sgn(c.pitch) * rw * rd;

    // --- basis columns -> q
    Q4 qbb = air_basis_adjoint(q, fb, rb, ub);
    g.p = pb;
    g.q = q4(qb.x + qbb.x, qb.y + qbb.y, qb.z + qbb.z, qb.w + qbb.w);
    g.v = vb;
    g.w = wb;
}

// Air-control loss at the final state: geodesic angle to a target
// orientation plus 1e-4 * (position error in uu)^2.
RSG_FN float air_loss(const AirState& s, Q4 tq, V3 tp_bt) {
    float d = fabsf(s.q.x * tq.x + s.q.y * tq.y + s.q.z * tq.z + s.q.w * tq.w);
    d = d < 1.f - 1e-6f ? d : 1.f - 1e-6f;
    V3 e = (s.p - tp_bt) * BT_TO_UU;
    return 2.f * acosf(d) + 1e-4f * len2(e);
}

RSG_FN void air_loss_grad(const AirState& s, Q4 tq, V3 tp_bt, float scale, AirState& g) {
    float dd = s.q.x * tq.x + s.q.y * tq.y + s.q.z * tq.z + s.q.w * tq.w;
    float d = fabsf(dd);
    float k = 0.f;
    if (d < 1.f - 1e-6f) k = -2.f / sqrtf(1.f - d * d) * air_sgn(dd);
    g.q = q4(tq.x * k * scale, tq.y * k * scale, tq.z * k * scale, tq.w * k * scale);
    V3 e = (s.p - tp_bt) * BT_TO_UU;
    g.p = e * (2e-4f * BT_TO_UU * scale);
    g.v = v3zero();
    g.w = v3zero();
}

}  // namespace rsg
