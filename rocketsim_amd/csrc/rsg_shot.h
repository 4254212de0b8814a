// rsg_shot.h -- differentiable aerial shot: one airborne car and the ball,
// through the car-ball hit, for batched shot optimisation.
//
// One tick of the reference for an airborne car and a flying ball with no
// world contact (Arena::Step, src/Sim/Arena/Arena.cpp:689-756):
//   * the car's air control and boost as in rsg_adjoint.h (Car.cpp:486-530,
//     584-667), as external impulses of the solver step;
//   * the ball's damping (btRigidBody::applyDamping, drag 0.03) and gravity;
//   * the car hitbox vs ball contact (btConvexConvexAlgorithm on the box and
//     the sphere): closest point of the box's core to the ball centre, normal
//     from the ball to the car, depth = distance - box margin - radius, a
//     contact when the depth is below the pair's breaking threshold;
//   * Ball::_OnHit's extra hit impulse (Ball.cpp:236-290), once per touch, added
//     to the ball's velocity at the end of the tick;
//   * the sequential-impulse solve of that contact as rsg_solver.h sets it up
//     (btSequentialImpulseConstraintSolver: normal row with friction 2.0 and
//     restitution 0, one friction row along the relative tangential velocity,
//     10 split-impulse passes, 10 velocity passes) and its write-back;
//   * integration (exponential map for the car, position only for the ball),
//     then the speed caps (Car::_FinishPhysicsTick, Ball::_FinishPhysicsTick).
// The car's orientation is carried as a unit quaternion, the inverse inertia
// in world space applied as R diag(I^-1) R^T to vectors.
//
// The step is written once over a scalar type: float for the forward rollout
// and Dn (value + one tangent) for its derivative.  The backward pass of a
// rollout recomputes each checkpointed segment and, per tick, takes the
// tangent of the step along each of its 26 inputs (22 state, 4 controls):
// the vector-Jacobian product g^T J is exact for the branch taken in the
// forward pass (contact tick, extra impulse, clamps of the solver rows),
// which the recomputation takes again.
#pragma once

#include "rsg_adjoint.h"
#include "rsg_const.h"

namespace rsg {

// ---------------------------------------------------------------------------
// Scalars: float, or a value with one tangent.

struct Dn {
    float v, d;
};
RSG_FN Dn dn(float v, float d = 0.f) { Dn r; r.v = v; r.d = d; return r; }
RSG_FN Dn operator+(Dn a, Dn b) { return dn(a.v + b.v, a.d + b.d); }
RSG_FN Dn operator-(Dn a, Dn b) { return dn(a.v - b.v, a.d - b.d); }
RSG_FN Dn operator-(Dn a) { return dn(-a.v, -a.d); }
RSG_FN Dn operator*(Dn a, Dn b) { return dn(a.v * b.v, a.d * b.v + a.v * b.d); }
RSG_FN Dn operator/(Dn a, Dn b) { float q = a.v / b.v; return dn(q, (a.d - q * b.d) / b.v); }
RSG_FN Dn operator+(Dn a, float b) { return dn(a.v + b, a.d); }
RSG_FN Dn operator+(float a, Dn b) { return dn(a + b.v, b.d); }
RSG_FN Dn operator-(Dn a, float b) { return dn(a.v - b, a.d); }
RSG_FN Dn operator-(float a, Dn b) { return dn(a - b.v, -b.d); }
RSG_FN Dn operator*(Dn a, float b) { return dn(a.v * b, a.d * b); }
RSG_FN Dn operator*(float a, Dn b) { return dn(a * b.v, a * b.d); }
RSG_FN Dn operator/(Dn a, float b) { return dn(a.v / b, a.d / b); }
RSG_FN Dn operator/(float a, Dn b) { float q = a / b.v; return dn(q, -q * b.d / b.v); }

RSG_FN float sval(float a) { return a; }
RSG_FN float sval(Dn a) { return a.v; }
RSG_FN float s_sqrt(float a) { return sqrtf(a); }
RSG_FN Dn s_sqrt(Dn a) { float s = sqrtf(a.v); return dn(s, s > 0.f ? a.d * 0.5f / s : 0.f); }
RSG_FN float s_sin(float a) { return sinf(a); }
RSG_FN Dn s_sin(Dn a) { return dn(sinf(a.v), cosf(a.v) * a.d); }
RSG_FN float s_cos(float a) { return cosf(a); }
RSG_FN Dn s_cos(Dn a) { return dn(cosf(a.v), -sinf(a.v) * a.d); }
RSG_FN float s_abs(float a) { return fabsf(a); }
RSG_FN Dn s_abs(Dn a) { return a.v < 0.f ? -a : a; }
template <class S> RSG_FN S s_const(float a) { return S(a); }
template <> RSG_FN Dn s_const<Dn>(float a) { return dn(a, 0.f); }
template <class S> RSG_FN S s_min(S a, S b) { return sval(a) < sval(b) ? a : b; }
template <class S> RSG_FN S s_clamp(S a, float lo, float hi) {
    return sval(a) < lo ? s_const<S>(lo) : (sval(a) > hi ? s_const<S>(hi) : a);
}

template <class S>
struct Vt {
    S x, y, z;
};
template <class S> RSG_FN Vt<S> vt(S x, S y, S z) { Vt<S> r; r.x = x; r.y = y; r.z = z; return r; }
template <class S> RSG_FN Vt<S> vt_zero() { return vt(s_const<S>(0.f), s_const<S>(0.f), s_const<S>(0.f)); }
template <class S> RSG_FN Vt<S> vt_const(V3 a) { return vt(s_const<S>(a.x), s_const<S>(a.y), s_const<S>(a.z)); }
template <class S> RSG_FN Vt<S> operator+(Vt<S> a, Vt<S> b) { return vt<S>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class S> RSG_FN Vt<S> operator-(Vt<S> a, Vt<S> b) { return vt<S>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class S> RSG_FN Vt<S> operator-(Vt<S> a) { return vt<S>(-a.x, -a.y, -a.z); }
template <class S, class K> RSG_FN Vt<S> operator*(Vt<S> a, K k) { return vt<S>(a.x * k, a.y * k, a.z * k); }
template <class S> RSG_FN Vt<S> vmul(Vt<S> a, V3 k) { return vt(a.x * k.x, a.y * k.y, a.z * k.z); }
template <class S> RSG_FN S vdot(Vt<S> a, Vt<S> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// btSimdDot3 association, x + (y + z)
template <class S> RSG_FN S vsdot(Vt<S> a, Vt<S> b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
template <class S> RSG_FN Vt<S> vcross(Vt<S> a, Vt<S> b) {
    return vt(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class S> RSG_FN S vlen2(Vt<S> a) { return vdot(a, a); }
template <class S> RSG_FN Vt<S> vcap(Vt<S> v, float vmax) {
    S l2 = vlen2(v);
    if (sval(l2) > vmax * vmax) return v * (vmax / s_sqrt(l2));
    return v;
}
// RocketSim's Vec::Normalized (1 / length, zero for a zero vector)
template <class S> RSG_FN Vt<S> vrs_normalized(Vt<S> a) {
    S l2 = (a.x * a.x + a.y * a.y) + a.z * a.z;
    if (!(sval(l2) > 0.f)) return vt(s_const<S>(0.f), s_const<S>(0.f), s_const<S>(0.f));
    S l = s_sqrt(l2);
    if (sval(l) > FLT_EPS * FLT_EPS) return vt(a.x / l, a.y / l, a.z / l);
    return vt(s_const<S>(0.f), s_const<S>(0.f), s_const<S>(0.f));
}

template <class S>
struct Qt {
    S x, y, z, w;
};
template <class S> RSG_FN Qt<S> qt(S x, S y, S z, S w) { Qt<S> q; q.x = x; q.y = y; q.z = z; q.w = w; return q; }
template <class S> RSG_FN Qt<S> qt_mul(Qt<S> a, Qt<S> b) {
    return qt((a.w * b.x - a.z * b.y) + (a.x * b.w + a.y * b.z), (a.w * b.y - a.x * b.z) + (a.y * b.w + a.z * b.x),
              (a.w * b.z - a.y * b.x) + (a.z * b.w + a.x * b.y), (a.w * b.w - a.z * b.z) + -(a.x * b.x + a.y * b.y));
}
template <class S> RSG_FN Qt<S> qt_normalized(Qt<S> q) {
    S inv = 1.f / s_sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
    return qt(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
}
// Columns of R(q) (rsg_adjoint.h air_basis).
template <class S> RSG_FN void qt_basis(Qt<S> q, Vt<S>& f, Vt<S>& r, Vt<S>& u) {
    S x = q.x, y = q.y, z = q.z, w = q.w;
    f = vt(1.f - 2.f * (y * y + z * z), 2.f * (x * y + z * w), 2.f * (z * x - w * y));
    r = vt(2.f * (x * y - w * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + w * x));
    u = vt(2.f * (z * x + y * w), 2.f * (y * z - w * x), 1.f - 2.f * (x * x + y * y));
}
// Exponential-map rotation of angular velocity w over dt (air_dq).
template <class S> RSG_FN Qt<S> qt_dq(Vt<S> w, float dt) {
    S th = s_sqrt(vlen2(w));
    Vt<S> a;
    if (sval(th) < 0.001f)
        a = w * (0.5f * dt - (dt * dt * dt) * (0.020833333333f) * th * th);
    else
        a = w * (s_sin(0.5f * th * dt) / th);
    return qt(a.x, a.y, a.z, s_cos(th * dt * 0.5f));
}

// ---------------------------------------------------------------------------
// The shot state and step

// Interface layout (uu, uu/s): car position 3, quaternion 4, velocity 3,
// angular velocity 3; ball position 3, velocity 3, angular velocity 3.
constexpr int SHOT_STATE_WORDS = 22;
constexpr int SHOT_INPUTS = SHOT_STATE_WORDS + 4;

enum { SHOT_CONTACT = 1u, SHOT_DEEP = 2u, SHOT_WORLD = 4u };

template <class S>
struct ShotState {
    Vt<S> cp;
    Qt<S> cq;
    Vt<S> cv, cw;
    Vt<S> bp, bv, bw;
};

// Flat access (Bullet units inside): word k of the interface order.
template <class S> RSG_FN S& shot_word(ShotState<S>& s, int k) {
    S* cs[SHOT_STATE_WORDS] = {&s.cp.x, &s.cp.y, &s.cp.z, &s.cq.x, &s.cq.y, &s.cq.z, &s.cq.w,
                               &s.cv.x, &s.cv.y, &s.cv.z, &s.cw.x, &s.cw.y, &s.cw.z,
                               &s.bp.x, &s.bp.y, &s.bp.z, &s.bv.x, &s.bv.y, &s.bv.z, &s.bw.x, &s.bw.y, &s.bw.z};
    return *cs[k];
}
// Scale of word k between the interface (uu) and Bullet units.
RSG_FN float shot_unit(int k) {
    return (k < 3 || (k >= 7 && k < 10) || (k >= 13 && k < 19)) ? UU_TO_BT : 1.f;
}

template <class S>
struct ShotCtrl {
    S pitch, yaw, roll, boost;
};

// Constants of the Octane car and the ball, computed once.
struct ShotConsts {
    V3 box_offset, box_he;   // hitbox child offset, core half extents
    float box_margin, radius;
    V3 car_inv_i;            // local inverse inertia
    float car_inv_m, ball_inv_m, ball_inv_i;
    float threshold;         // contact breaking threshold of the pair
    float drag;              // (1 - drag)^dt
    float dt;
};

RSG_FN ShotConsts shot_consts(float dt) {
    CarPhys cp = make_car_phys(0);  // Octane
    BallPhys bp = make_ball_phys();
    ShotConsts k;
    k.box_offset = cp.box_offset;
    k.box_he = cp.box_implicit;
    k.box_margin = cp.box_margin;
    k.radius = bp.radius;
    k.car_inv_i = cp.inv_inertia_local;
    k.car_inv_m = cp.inv_mass;
    k.ball_inv_m = bp.inv_mass;
    k.ball_inv_i = bp.inv_inertia;
    float ta = cp.angular_motion_disc * CONTACT_BREAKING_THRESHOLD;
    float tb = bp.angular_motion_disc * CONTACT_BREAKING_THRESHOLD;
    k.threshold = ta < tb ? ta : tb;
    k.drag = powf(1.f - BALL_DRAG, dt);
    k.dt = dt;
    return k;
}

// Solver body and rows of the one contact (rsg_solver.h).
template <class S>
struct ShotBody {
    Vt<S> dv, dw, push, turn;
};
template <class S>
struct ShotRow {
    Vt<S> n1, rc1, n2, rc2, ang_a, ang_b;
    S jac, rhs, rhs_pen, applied, applied_push;
};

template <class S>
RSG_FN S shot_row_dot(const ShotRow<S>& c, const ShotBody<S>& A, const ShotBody<S>& B) {
    S d1 = vsdot(c.n1, A.dv) + vsdot(c.rc1, A.dw);
    S d2 = vsdot(c.n2, B.dv) + vsdot(c.rc2, B.dw);
    return d1 * c.jac + d2 * c.jac;
}

template <class S>
RSG_FN void shot_apply(const ShotRow<S>& c, ShotBody<S>& A, ShotBody<S>& B, S delta, const ShotConsts& k) {
    A.dv = A.dv + (c.n1 * k.car_inv_m) * delta;
    A.dw = A.dw + c.ang_a * delta;
    B.dv = B.dv + (c.n2 * k.ball_inv_m) * delta;
    B.dw = B.dw + c.ang_b * delta;
}

// One tick.  `tick` counts from 0; last_extra is the tick of the last extra
// hit impulse (start below -1); flags collect SHOT_*.
template <class S>
RSG_FN void shot_step(ShotState<S>& s, const ShotCtrl<S>& c, int tick, int& last_extra, uint32_t& flags,
                      const ShotConsts& k) {
    const float dt = k.dt;
    Vt<S> f, r, u;
    qt_basis(s.cq, f, r, u);
    // car: air control torque and boost, as solver-step impulses
    Vt<S> torque = r * (-c.pitch * AIR_TORQUE_P) + u * (c.yaw * AIR_TORQUE_Y) + f * (-c.roll * AIR_TORQUE_R);
    S ky = AIR_DAMP_Y * (1.f - s_abs(c.yaw)), kp = AIR_DAMP_P * (1.f - s_abs(c.pitch));
    Vt<S> damp = u * (ky * vdot(u, s.cw)) + r * (kp * vdot(r, s.cw)) + f * (AIR_DAMP_R * vdot(f, s.cw));
    Vt<S> ext_t = (torque - damp) * (CAR_TORQUE_SCALE * dt);
    Vt<S> grav = vt_const<S>(v3(0.f, 0.f, GRAVITY_Z * UU_TO_BT));
    Vt<S> ext_f = (grav + f * (c.boost * BOOST_ACCEL_AIR * UU_TO_BT)) * dt;
    // ball: damping, gravity impulse
    s.bv = s.bv * k.drag;
    Vt<S> ext_fb = grav * dt;

    // car hitbox vs ball: closest point of the box core to the ball centre
    Vt<S> box_c = s.cp + f * k.box_offset.x + r * k.box_offset.y + u * k.box_offset.z;
    Vt<S> d = s.bp - box_c;
    S lx = vdot(f, d), ly = vdot(r, d), lz = vdot(u, d);
    const bool inside = fabsf(sval(lx)) < k.box_he.x && fabsf(sval(ly)) < k.box_he.y && fabsf(sval(lz)) < k.box_he.z;
    S qx = s_clamp(lx, -k.box_he.x, k.box_he.x), qy = s_clamp(ly, -k.box_he.y, k.box_he.y),
      qz = s_clamp(lz, -k.box_he.z, k.box_he.z);
    Vt<S> to_box = (box_c + f * qx + r * qy + u * qz) - s.bp;
    S dist_core = s_sqrt(vlen2(to_box));
    S depth = dist_core - (k.box_margin + k.radius);
    if (inside) flags |= SHOT_DEEP;
    const bool contact = !inside && sval(depth) < k.threshold && sval(dist_core) > 0.f;

    Vt<S> cache = vt(s_const<S>(0.f), s_const<S>(0.f), s_const<S>(0.f));
    Vt<S> car_v = s.cv, car_w = s.cw, ball_v = s.bv, ball_w = s.bw;
    if (contact) {
        flags |= SHOT_CONTACT;
        Vt<S> n = to_box * (1.f / dist_core);          // from the ball towards the car
        Vt<S> pt_b = s.bp + n * k.radius;
        Vt<S> pt_a = pt_b + n * depth;
        Vt<S> rel1 = pt_a - s.cp, rel2 = pt_b - s.bp;
        // Ball::_OnHit's extra impulse, once per touch
        if (tick > last_extra + 1 || last_extra > tick) {
            last_extra = tick;
            Vt<S> rel_pos = s.bp * BT_TO_UU - s.cp * BT_TO_UU;
            Vt<S> rel_vel = s.bv * BT_TO_UU - s.cv * BT_TO_UU;
            S rel_speed = s_min(s_sqrt(vlen2(rel_vel)), s_const<S>(BALL_CAR_EXTRA_IMPULSE_MAXDELTAVEL_UU));
            if (sval(rel_speed) > 0.f) {
                Vt<S> hit_dir = vrs_normalized(vmul(rel_pos, v3(1.f, 1.f, BALL_CAR_EXTRA_IMPULSE_Z_SCALE)));
                Vt<S> fwd_adj = (f * vdot(hit_dir, f)) * (1.f - BALL_CAR_EXTRA_IMPULSE_FORWARD_SCALE);
                hit_dir = vrs_normalized(hit_dir - fwd_adj);
                // LinearPieceCurve of the relative speed (curve_ball_extra_impulse)
                Curve cv = curve_ball_extra_impulse();
                S scale = s_const<S>(cv.y[cv.n - 1]);
                if (sval(rel_speed) <= cv.x[0]) {
                    scale = s_const<S>(cv.y[0]);
                } else {
                    for (int i = 1; i < cv.n; i++) {
                        if (cv.x[i] > sval(rel_speed)) {
                            S t = (rel_speed - cv.x[i - 1]) / (cv.x[i] - cv.x[i - 1]);
                            scale = cv.y[i - 1] + (cv.y[i] - cv.y[i - 1]) * t;
                            break;
                        }
                    }
                }
                Vt<S> added = ((hit_dir * rel_speed) * scale) * 1.f;
                cache = added * UU_TO_BT;
            }
        }
        // the contact row (setup_contact_row) and its friction row
        // (add_friction_for / setup_friction_row)
        auto inv_iw_car = [&](Vt<S> x) {
            return f * (k.car_inv_i.x * vdot(f, x)) + r * (k.car_inv_i.y * vdot(r, x)) + u * (k.car_inv_i.z * vdot(u, x));
        };
        ShotRow<S> cr;
        Vt<S> ta0 = vcross(rel1, n), ta1 = vcross(rel2, n);
        cr.ang_a = inv_iw_car(ta0);
        cr.ang_b = -ta1 * k.ball_inv_i;
        {
            S d0 = k.car_inv_m + vdot(n, vcross(cr.ang_a, rel1));
            S d1 = k.ball_inv_m + vdot(n, vcross(-cr.ang_b, rel2));
            cr.jac = 1.f / (d0 + d1);
        }
        cr.n1 = n; cr.rc1 = ta0; cr.n2 = -n; cr.rc2 = -ta1;
        {
            S v1n = vdot(cr.n1, s.cv + ext_f) + vdot(cr.rc1, s.cw + ext_t);
            S v2n = vdot(cr.n2, s.bv + ext_fb) + vdot(cr.rc2, s.bw);
            S vel_err = 0.f - (v1n + v2n);
            S pos_err = sval(depth) > 0.f ? s_const<S>(0.f) : (-depth * SOLVER_ERP2) * (1.f / dt);
            cr.rhs = vel_err * cr.jac;
            cr.rhs_pen = pos_err * cr.jac;
        }
        cr.applied = s_const<S>(0.f);
        cr.applied_push = s_const<S>(0.f);
        ShotRow<S> fr;
        {
            Vt<S> vel1 = (s.cv + ext_f) + vcross(s.cw + ext_t, rel1);
            Vt<S> vel2 = (s.bv + ext_fb) + vcross(s.bw, rel2);
            Vt<S> vel = vel1 - vel2;
            Vt<S> dir = vel - n * vdot(n, vel);
            S lat = vlen2(dir);
            if (sval(lat) > FLT_EPS) {
                dir = dir * (1.f / s_sqrt(lat));
            } else {
                // btPlaneSpace1 of the normal
                if (fabsf(sval(n.z)) > 0.7071067811865475244f) {
                    S a = n.y * n.y + n.z * n.z, kk = 1.f / s_sqrt(a);
                    dir = vt(s_const<S>(0.f), -n.z * kk, n.y * kk);
                } else {
                    S a = n.x * n.x + n.y * n.y, kk = 1.f / s_sqrt(a);
                    dir = vt(-n.y * kk, n.x * kk, s_const<S>(0.f));
                }
            }
            fr.n1 = dir;
            fr.rc1 = vcross(rel1, dir);
            fr.ang_a = inv_iw_car(fr.rc1);
            fr.n2 = -dir;
            fr.rc2 = vcross(rel2, fr.n2);
            fr.ang_b = fr.rc2 * k.ball_inv_i;
            S d0 = k.car_inv_m + vdot(dir, vcross(fr.ang_a, rel1));
            S d1 = k.ball_inv_m + vdot(dir, vcross(-fr.ang_b, rel2));
            fr.jac = 1.f / (d0 + d1);
            S v1n = vdot(fr.n1, s.cv + ext_f) + vdot(fr.rc1, s.cw);
            S v2n = vdot(fr.n2, s.bv + ext_fb) + vdot(fr.rc2, s.bw);
            fr.rhs = (0.f - (v1n + v2n)) * fr.jac;
            fr.applied = s_const<S>(0.f);
        }
        ShotBody<S> A, B;
        A.dv = A.dw = A.push = A.turn = vt(s_const<S>(0.f), s_const<S>(0.f), s_const<S>(0.f));
        B = A;
        // split-impulse passes (resolve_split), early out once settled
        for (int it = 0; it < SOLVER_ITERATIONS; it++) {
            float lsq = 0.f;
            if (sval(cr.rhs_pen) != 0.f) {
                S d1 = vsdot(cr.n1, A.push) + vsdot(cr.rc1, A.turn);
                S d2 = vsdot(cr.n2, B.push) + vsdot(cr.rc2, B.turn);
                S delta = cr.rhs_pen - cr.applied_push * 0.f;
                delta = delta - d1 * cr.jac;
                delta = delta - d2 * cr.jac;
                S sum = cr.applied_push + delta;
                if (sval(sum) < 0.f) {
                    delta = 0.f - cr.applied_push;
                    cr.applied_push = s_const<S>(0.f);
                } else {
                    cr.applied_push = sum;
                }
                A.push = A.push + (cr.n1 * k.car_inv_m) * delta;
                A.turn = A.turn + cr.ang_a * delta;
                B.push = B.push + (cr.n2 * k.ball_inv_m) * delta;
                B.turn = B.turn + cr.ang_b * delta;
                float rr = sval(delta) * (1.f / sval(cr.jac));
                lsq = rr * rr;
            }
            if (lsq <= 0.f || it >= SOLVER_ITERATIONS - 1) break;
        }
        // velocity passes: the contact row (resolve_lower), then its friction
        // row within +- friction * applied (resolve_generic)
        for (int it = 0; it < SOLVER_ITERATIONS; it++) {
            {
                S delta = cr.rhs - cr.applied * 0.f;
                S d1 = vsdot(cr.n1, A.dv) + vsdot(cr.rc1, A.dw);
                S d2 = vsdot(cr.n2, B.dv) + vsdot(cr.rc2, B.dw);
                delta = delta - d1 * cr.jac;
                delta = delta - d2 * cr.jac;
                S sum = cr.applied + delta;
                if (sval(sum) < 0.f) {
                    delta = 0.f - cr.applied;
                    cr.applied = s_const<S>(0.f);
                } else {
                    cr.applied = sum;
                }
                shot_apply(cr, A, B, delta, k);
            }
            S total = cr.applied;
            if (sval(total) > 0.f) {
                S lo = -(CARBALL_COLLISION_FRICTION * total), hi = CARBALL_COLLISION_FRICTION * total;
                S delta = fr.rhs - fr.applied * 0.f;
                S d1 = vsdot(fr.n1, A.dv) + vsdot(fr.rc1, A.dw);
                S d2 = vsdot(fr.n2, B.dv) + vsdot(fr.rc2, B.dw);
                delta = delta - d1 * fr.jac;
                delta = delta - d2 * fr.jac;
                S sum = fr.applied + delta;
                const bool lower_less = sval(sum) < sval(lo);
                const bool upper_less = sval(sum) < sval(hi);
                delta = lower_less ? lo - fr.applied : delta;
                S app = lower_less ? lo : sum;
                delta = upper_less ? delta : hi - fr.applied;
                fr.applied = upper_less ? app : hi;
                shot_apply(fr, A, B, delta, k);
            }
        }
        // write-back (btSolverBody::writebackVelocityAndTransform)
        car_v = (s.cv + A.dv) + ext_f;
        car_w = (s.cw + A.dw) + ext_t;
        ball_v = (s.bv + B.dv) + ext_fb;
        ball_w = (s.bw + B.dw);
        const bool car_push = sval(A.push.x) != 0.f || sval(A.push.y) != 0.f || sval(A.push.z) != 0.f ||
                              sval(A.turn.x) != 0.f || sval(A.turn.y) != 0.f || sval(A.turn.z) != 0.f;
        if (car_push) {
            s.cp = s.cp + A.push * dt;
            s.cq = qt_normalized(qt_mul(qt_dq(A.turn * SOLVER_SPLIT_TURN_ERP, dt), s.cq));
        }
        const bool ball_push = sval(B.push.x) != 0.f || sval(B.push.y) != 0.f || sval(B.push.z) != 0.f ||
                               sval(B.turn.x) != 0.f || sval(B.turn.y) != 0.f || sval(B.turn.z) != 0.f;
        if (ball_push) s.bp = s.bp + B.push * dt;
    } else {
        car_v = s.cv + ext_f;
        car_w = s.cw + ext_t;
        ball_v = s.bv + ext_fb;
    }
    // integrateTransforms
    s.cp = s.cp + car_v * dt;
    s.cq = qt_normalized(qt_mul(qt_dq(car_w, dt), s.cq));
    s.bp = s.bp + ball_v * dt;
    // post-tick: caps; the ball's extra impulse joins its velocity first
    s.cv = vcap(car_v, CAR_MAX_SPEED * UU_TO_BT);
    s.cw = vcap(car_w, CAR_MAX_ANG_SPEED);
    s.bv = vcap(ball_v + cache, BALL_MAX_SPEED * UU_TO_BT);
    s.bw = vcap(ball_w, BALL_MAX_ANG_SPEED);
    // out of the modelled domain: near the floor, ceiling or walls
    const float floor_z = 0.f, ceil_z = 2044.f * UU_TO_BT, wall_x = (4096.f - 256.f) * UU_TO_BT,
                wall_y = (5120.f - 256.f) * UU_TO_BT, car_r = 2.5f;
    if (sval(s.bp.z) < floor_z + k.radius + 0.5f || sval(s.bp.z) > ceil_z - k.radius - 0.5f ||
        fabsf(sval(s.bp.x)) > wall_x - k.radius || fabsf(sval(s.bp.y)) > wall_y - k.radius ||
        sval(s.cp.z) < floor_z + car_r || sval(s.cp.z) > ceil_z - car_r || fabsf(sval(s.cp.x)) > wall_x - car_r ||
        fabsf(sval(s.cp.y)) > wall_y - car_r)
        flags |= SHOT_WORLD;
}

// Loss: distance of the ball to the target (uu).
template <class S> RSG_FN S shot_loss(const ShotState<S>& s, V3 target_uu) {
    Vt<S> d = s.bp * BT_TO_UU - vt_const<S>(target_uu);
    return s_sqrt(vlen2(d));
}

// ---------------------------------------------------------------------------
// Rollouts (one lane per car-ball pair; structure-of-arrays by word)

constexpr int SHOT_CKPT_WORDS = SHOT_STATE_WORDS + 1;  // + the last extra-impulse tick
constexpr int SHOT_MAX_K = 32;

RSG_FN float shot_to_bt(int k, float x) { return shot_unit(k) == 1.f ? x : x * UU_TO_BT; }
RSG_FN float shot_to_uu(int k, float x) { return shot_unit(k) == 1.f ? x : x * BT_TO_UU; }

RSG_FN void shot_load(const float* src, int n, int i, ShotState<float>& s) {
    for (int k = 0; k < SHOT_STATE_WORDS; k++) shot_word(s, k) = shot_to_bt(k, src[(size_t)k * n + i]);
}
RSG_FN void shot_store(float* dst, int n, int i, ShotState<float>& s) {
    for (int k = 0; k < SHOT_STATE_WORDS; k++) dst[(size_t)k * n + i] = shot_to_uu(k, shot_word(s, k));
}
RSG_FN void shot_ckpt_store(float* ck, int n, int i, ShotState<float>& s, int last_extra) {
    for (int k = 0; k < SHOT_STATE_WORDS; k++) ck[(size_t)k * n + i] = shot_word(s, k);
    ck[(size_t)SHOT_STATE_WORDS * n + i] = (float)last_extra;
}
RSG_FN void shot_ckpt_load(const float* ck, int n, int i, ShotState<float>& s, int& last_extra) {
    for (int k = 0; k < SHOT_STATE_WORDS; k++) shot_word(s, k) = ck[(size_t)k * n + i];
    last_extra = (int)ck[(size_t)SHOT_STATE_WORDS * n + i];
}
RSG_FN ShotCtrl<float> shot_ctrl(const float* ctrl, int n, int i, int t) {
    const float* c = ctrl + (size_t)t * 4 * n;
    ShotCtrl<float> r;
    r.pitch = c[i]; r.yaw = c[n + i]; r.roll = c[2 * n + i]; r.boost = c[3 * n + i];
    return r;
}

// Forward: `ticks` steps from state0 with a checkpoint every K ticks; returns
// the loss and leaves the final state in s.
RSG_FN float shot_forward(ShotState<float>& s, const float* ctrl, int n, int i, int ticks, int K, float* ckpt,
                          V3 target, const ShotConsts& k, uint32_t& flags) {
    int last_extra = -1000;
    for (int t = 0; t < ticks; t++) {
        if (ckpt && t % K == 0) shot_ckpt_store(ckpt + (size_t)(t / K) * SHOT_CKPT_WORDS * n, n, i, s, last_extra);
        shot_step(s, shot_ctrl(ctrl, n, i, t), t, last_extra, flags, k);
    }
    return shot_loss(s, target);
}

// One tangent pass of tick t at (s, c), along input j (0-21 the state words,
// 22-25 pitch, yaw, roll, boost; any other j seeds nothing): returns
// sum_w g[w] d out_w / d in_j, entry j of the vector-Jacobian product g^T J.
RSG_FN float shot_vjp_entry(const ShotState<float>& s, const ShotCtrl<float>& c, int t, int last_extra,
                            const ShotConsts& k, const float g[SHOT_STATE_WORDS], int j) {
    ShotState<Dn> sd;
    ShotState<float> sf = s;
    for (int w = 0; w < SHOT_STATE_WORDS; w++) shot_word(sd, w) = dn(shot_word(sf, w), w == j ? 1.f : 0.f);
    ShotCtrl<Dn> cd;
    cd.pitch = dn(c.pitch, j == SHOT_STATE_WORDS ? 1.f : 0.f);
    cd.yaw = dn(c.yaw, j == SHOT_STATE_WORDS + 1 ? 1.f : 0.f);
    cd.roll = dn(c.roll, j == SHOT_STATE_WORDS + 2 ? 1.f : 0.f);
    cd.boost = dn(c.boost, j == SHOT_STATE_WORDS + 3 ? 1.f : 0.f);
    int le = last_extra;
    uint32_t fl = 0;
    shot_step(sd, cd, t, le, fl, k);
    float acc = 0.f;
    for (int w = 0; w < SHOT_STATE_WORDS; w++) acc += g[w] * shot_word(sd, w).d;
    return acc;
}

// Vector-Jacobian product of tick t at (s, c): g (adjoint of the output
// state, Bullet units) becomes the adjoint of the input state; gc receives
// the controls' gradient.  One tangent pass per input.
RSG_FN void shot_step_vjp(const ShotState<float>& s, const ShotCtrl<float>& c, int t, int last_extra,
                          const ShotConsts& k, float g[SHOT_STATE_WORDS], float gc[4]) {
    float gin[SHOT_INPUTS];
    for (int j = 0; j < SHOT_INPUTS; j++) gin[j] = shot_vjp_entry(s, c, t, last_extra, k, g, j);
    for (int w = 0; w < SHOT_STATE_WORDS; w++) g[w] = gin[w];
    for (int j = 0; j < 4; j++) gc[j] = gin[SHOT_STATE_WORDS + j];
}

// Adjoint of the loss |50 bp - target| (uu) with respect to the final state
// (Bullet units): 50 (50 bp - target) / loss on bp, scaled by dloss.
RSG_FN void shot_loss_grad(const ShotState<float>& s, V3 target, float dloss, float g[SHOT_STATE_WORDS]) {
    for (int w = 0; w < SHOT_STATE_WORDS; w++) g[w] = 0.f;
    V3 d = v3(s.bp.x, s.bp.y, s.bp.z) * BT_TO_UU - target;
    float l = len(d);
    if (l > 0.f) {
        float sc = dloss * BT_TO_UU / l;
        g[13] = d.x * sc; g[14] = d.y * sc; g[15] = d.z * sc;
    }
}

// Backward: recomputes each segment from its checkpoint and applies the
// per-tick VJPs in reverse; writes the controls' gradient [ticks][4][n] and,
// when ds0 is not null, the initial state's gradient in interface units.
RSG_FN void shot_backward(const float* ctrl, int n, int i, int ticks, int K, const float* ckpt, V3 target,
                          float dloss, const ShotConsts& k, float* dctrl, float* ds0) {
    const int nseg = (ticks + K - 1) / K;
    ShotState<float> st[SHOT_MAX_K];
    int le[SHOT_MAX_K];
    // final state from the last checkpoint
    ShotState<float> s;
    int last_extra;
    shot_ckpt_load(ckpt + (size_t)(nseg - 1) * SHOT_CKPT_WORDS * n, n, i, s, last_extra);
    uint32_t fl = 0;
    for (int t = (nseg - 1) * K; t < ticks; t++) shot_step(s, shot_ctrl(ctrl, n, i, t), t, last_extra, fl, k);
    float g[SHOT_STATE_WORDS];
    shot_loss_grad(s, target, dloss, g);
    for (int seg = nseg - 1; seg >= 0; seg--) {
        int t0 = seg * K, t1 = t0 + K < ticks ? t0 + K : ticks;
        ShotState<float> x;
        int lx;
        shot_ckpt_load(ckpt + (size_t)seg * SHOT_CKPT_WORDS * n, n, i, x, lx);
        for (int t = t0; t < t1; t++) {
            st[t - t0] = x;
            le[t - t0] = lx;
            shot_step(x, shot_ctrl(ctrl, n, i, t), t, lx, fl, k);
        }
        for (int t = t1 - 1; t >= t0; t--) {
            float gc[4];
            shot_step_vjp(st[t - t0], shot_ctrl(ctrl, n, i, t), t, le[t - t0], k, g, gc);
            float* b = dctrl + (size_t)t * 4 * n;
            b[i] = gc[0]; b[n + i] = gc[1]; b[2 * n + i] = gc[2]; b[3 * n + i] = gc[3];
        }
    }
    if (ds0)
        for (int w = 0; w < SHOT_STATE_WORDS; w++) ds0[(size_t)w * n + i] = g[w] * shot_unit(w);
}

}  // namespace rsg
