// rsg_solver.h -- sequential-impulse (projected Gauss-Seidel) contact solver
// of one arena, restating the reference's btSequentialImpulseConstraintSolver
// as configured by RocketSim (Catto, "Iterative Dynamics with Temporal
// Coherence", GDC 2005; Bullet 3.24 BulletDynamics/ConstraintSolver):
//   * 10 split-impulse (position) passes, then 10 velocity passes: normal
//     rows first, then one velocity-dependent friction row per contact
//     (btSequentialImpulseConstraintSolver.cpp:1601-1829);
//   * contact rows: erp2 = 0.8, no position bias in the velocity rows
//     (RocketSim change, :949-955), split impulse always on (threshold 1e30);
//   * ball-vs-world contacts are "special": skipped by the velocity passes and
//     replaced by one averaged row per body (:1048-1061, :1164-1211);
//   * write-back v += dv + F dt / m, w += dw + (T * invI) dt, pose nudged by
//     push/turn velocities (btSolverBody.h:265-287).
// No contact persists between ticks in the reference (pairs are rebuilt every
// tick), so there is no warm starting here either.
#pragma once

#include "rsg_contact.h"

namespace rsg {

constexpr int MAX_SOLVER_BODIES = 1 + RSG_MAX_CARS + 1;
// Contact rows and friction rows per arena and tick (each); a host test
// build may lower the capacity (RSG_TEST_MAX_SOLVER_ROWS).
#ifdef RSG_TEST_MAX_SOLVER_ROWS
constexpr int MAX_ROWS = RSG_TEST_MAX_SOLVER_ROWS;
#else
constexpr int MAX_ROWS = RSG_MAX_SOLVER_ROWS;
#endif

struct SBody {
    int body;        // dynamic body index, -1 for the fixed body
    Xf xf;
    V3 inv_mass;     // per-axis inverse mass (linear factor 1)
    V3 lin, ang;     // velocities at setup
    V3 ext_f, ext_t; // external force / torque impulses
    V3 dv, dw, push, turn;
};

struct SRow {
    int a, b;
    V3 n1, n2, rc1, rc2, ang_a, ang_b;
    float rhs, rhs_pen, cfm, lo, hi, jac_inv;
    float applied, applied_push;
    float friction;
    int friction_index;  // contact rows: first friction row; friction rows: their contact row
    bool special;
};

// btSimdDot3 association: x + (y + z)
RSG_FN float sdot(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }

struct SpecialInfo {
    int num;
    float friction, restitution;
    V3 total_normal;
    float total_dist;
};

struct Solver {
    SBody bodies[MAX_SOLVER_BODIES];
    int n_bodies;
    int fixed_id;
    int companion[MAX_SOLVER_BODIES];  // dynamic body -> solver body, -1 none
    SRow* rows;      // MAX_ROWS contact rows (storage placed by the caller)
    int n_rows;
    SRow* fric;      // MAX_ROWS friction rows
    int n_fric;
    SpecialInfo special[MAX_SOLVER_BODIES];
    float dt;
    bool overflow;
};

RSG_FN void solver_reset(Solver& S, float dt) {
    S.n_bodies = 0;
    S.fixed_id = -1;
    for (int i = 0; i < MAX_SOLVER_BODIES; i++) {
        S.companion[i] = -1;
        S.special[i].num = 0;
        S.special[i].friction = 0;
        S.special[i].restitution = 0;
        S.special[i].total_normal = v3zero();
        S.special[i].total_dist = 0;
    }
    S.n_rows = 0;
    S.n_fric = 0;
    S.dt = dt;
    S.overflow = false;
}

RSG_FN void init_solver_body(SBody& sb, const Body* rb, int body, float dt) {
    sb.body = body;
    sb.dv = v3zero(); sb.dw = v3zero(); sb.push = v3zero(); sb.turn = v3zero();
    if (rb) {
        sb.xf = xf_from(rb->pos, rb->rot);
        sb.inv_mass = v3(rb->inv_mass, rb->inv_mass, rb->inv_mass);
        sb.lin = rb->vel;
        sb.ang = rb->ang;
        sb.ext_f = rb->force * rb->inv_mass * dt;
        sb.ext_t = mul_vm(rb->torque, rb->inv_iw) * dt;
    } else {
        sb.xf = xf_ident();
        sb.inv_mass = v3zero();
        sb.lin = v3zero(); sb.ang = v3zero();
        sb.ext_f = v3zero(); sb.ext_t = v3zero();
    }
}

template <class BodyAt>
RSG_FN int solver_body_id(Solver& S, int body, BodyAt& body_at) {
    if (body >= 0) {
        if (S.companion[body] >= 0) return S.companion[body];
        const Body* rb = body_at(body);
        if (rb->inv_mass != 0.f) {
            int id = S.n_bodies++;
            init_solver_body(S.bodies[id], rb, body, S.dt);
            S.companion[body] = id;
            return id;
        }
    }
    if (S.fixed_id < 0) {
        S.fixed_id = S.n_bodies++;
        init_solver_body(S.bodies[S.fixed_id], nullptr, -1, S.dt);
    }
    return S.fixed_id;
}

// setupContactConstraint (btSequentialImpulseConstraintSolver.cpp:795-976)
RSG_FN void setup_contact_row(Solver& S, SRow& c, int ia, int ib, const Body* rb0, const Body* rb1,
                              V3 normal, float dist, float friction, float restitution_coef,
                              V3 rel1, V3 rel2) {
    SBody& A = S.bodies[ia];
    SBody& B = S.bodies[ib];
    float inv_dt = 1.f / S.dt;
    float cfm = 0.f;  // m_globalCfm
    float erp = SOLVER_ERP2;
    cfm *= inv_dt;
    V3 ta0 = cross(rel1, normal);
    c.ang_a = rb0 ? (rb0->inv_iw * ta0) * v3(1, 1, 1) : v3zero();
    V3 ta1 = cross(rel2, normal);
    c.ang_b = rb1 ? (rb1->inv_iw * -ta1) * v3(1, 1, 1) : v3zero();
    float d0 = 0.f, d1 = 0.f;
    if (rb0) {
        V3 vec = cross(c.ang_a, rel1);
        d0 = rb0->inv_mass + dot(normal, vec);
    }
    if (rb1) {
        V3 vec = cross(-c.ang_b, rel2);
        d1 = rb1->inv_mass + dot(normal, vec);
    }
    float relax = 1.f;  // m_sor
    c.jac_inv = relax / (d0 + d1 + cfm);
    if (rb0) { c.n1 = normal; c.rc1 = ta0; } else { c.n1 = v3zero(); c.rc1 = v3zero(); }
    if (rb1) { c.n2 = -normal; c.rc2 = -ta1; } else { c.n2 = v3zero(); c.rc2 = v3zero(); }
    float penetration = dist + 0.f;
    float restitution = 0.f;
    {
        V3 vel1 = rb0 ? body_vel_at(*rb0, rel1) : v3zero();
        V3 vel2 = rb1 ? body_vel_at(*rb1, rel2) : v3zero();
        V3 vel = vel1 - vel2;
        float rel_vel = dot(normal, vel);
        c.friction = friction;
        // restitutionCurve
        if (fabsf(rel_vel) < SOLVER_RESTITUTION_VEL_THRESHOLD) restitution = 0.f;
        else restitution = restitution_coef * -rel_vel;
        if (restitution <= 0.f) restitution = 0.f;
    }
    c.applied = 0.f;
    c.applied_push = 0.f;
    {
        V3 efa = A.body >= 0 ? A.ext_f : v3zero();
        V3 eta = A.body >= 0 ? A.ext_t : v3zero();
        V3 efb = B.body >= 0 ? B.ext_f : v3zero();
        V3 etb = B.body >= 0 ? B.ext_t : v3zero();
        float v1n = dot(c.n1, A.lin + efa) + dot(c.rc1, A.ang + eta);
        float v2n = dot(c.n2, B.lin + efb) + dot(c.rc2, B.ang + etb);
        float rel_vel = v1n + v2n;
        float pos_err = 0.f;
        float vel_err = restitution - rel_vel;
        if (penetration > 0) pos_err = 0;
        else pos_err = -penetration * erp * inv_dt;
        float pen_imp = pos_err * c.jac_inv;
        float vel_imp = vel_err * c.jac_inv;
        // split impulse with threshold 1e30: always split
        c.rhs = vel_imp;
        c.rhs_pen = pen_imp;
        c.cfm = cfm * c.jac_inv;
        c.lo = 0;
        c.hi = 1e10f;
    }
}

// setupFrictionConstraint (btSequentialImpulseConstraintSolver.cpp:513-602)
RSG_FN void setup_friction_row(Solver& S, SRow& c, int ia, int ib, const Body* rb0, const Body* rb1,
                               V3 axis, float friction, V3 rel1, V3 rel2) {
    SBody& A = S.bodies[ia];
    SBody& B = S.bodies[ib];
    c.a = ia; c.b = ib;
    c.friction = friction;
    c.applied = 0.f;
    c.applied_push = 0.f;
    if (rb0) {
        c.n1 = axis;
        c.rc1 = cross(rel1, c.n1);
        c.ang_a = rb0->inv_iw * c.rc1 * v3(1, 1, 1);
    } else {
        c.n1 = v3zero(); c.rc1 = v3zero(); c.ang_a = v3zero();
    }
    if (rb1) {
        c.n2 = -axis;
        c.rc2 = cross(rel2, c.n2);
        c.ang_b = rb1->inv_iw * c.rc2 * v3(1, 1, 1);
    } else {
        c.n2 = v3zero(); c.rc2 = v3zero(); c.ang_b = v3zero();
    }
    float d0 = 0.f, d1 = 0.f;
    if (rb0) {
        V3 vec = cross(c.ang_a, rel1);
        d0 = rb0->inv_mass + dot(axis, vec);
    }
    if (rb1) {
        V3 vec = cross(-c.ang_b, rel2);
        d1 = rb1->inv_mass + dot(axis, vec);
    }
    c.jac_inv = 1.f / (d0 + d1);
    float v1n = dot(c.n1, rb0 ? A.lin + A.ext_f : v3zero()) + dot(c.rc1, rb0 ? A.ang : v3zero());
    float v2n = dot(c.n2, rb1 ? B.lin + B.ext_f : v3zero()) + dot(c.rc2, rb1 ? B.ang : v3zero());
    float rel_vel = v1n + v2n;
    float vel_err = 0.f - rel_vel;
    float vel_imp = vel_err * c.jac_inv;
    c.rhs = 0.f + vel_imp;
    c.rhs_pen = 0.f;
    c.cfm = 0.f;
    c.lo = -c.friction;
    c.hi = c.friction;
    c.special = false;
}

// convertContactInner's friction part with SOLVER_USE_WARMSTARTING | SOLVER_SIMD
// (btSequentialImpulseConstraintSolver.cpp:1071-1161): one friction row along
// the relative tangential velocity, or btPlaneSpace1 of the normal when slow.
RSG_FN void add_friction_for(Solver& S, int ia, int ib, const Body* rb0, const Body* rb1,
                             V3 normal, float friction, V3 rel1, V3 rel2, int contact_index) {
    SBody& A = S.bodies[ia];
    SBody& B = S.bodies[ib];
    V3 vel1 = A.body >= 0 ? (A.lin + A.ext_f) + cross(A.ang + A.ext_t, rel1) : v3zero();
    V3 vel2 = B.body >= 0 ? (B.lin + B.ext_f) + cross(B.ang + B.ext_t, rel2) : v3zero();
    V3 vel = vel1 - vel2;
    float rel_vel = dot(normal, vel);
    V3 dir = vel - normal * rel_vel;
    float lat = len2(dir);
    if (lat > FLT_EPS) {
        dir = dir * (1.f / sqrtf(lat));
    } else {
        V3 d2;
        plane_space(normal, dir, d2);
    }
    if (S.n_fric >= MAX_ROWS) { S.overflow = true; return; }
    SRow& f = S.fric[S.n_fric++];
    setup_friction_row(S, f, ia, ib, rb0, rb1, dir, friction, rel1, rel2);
    f.friction_index = contact_index;
}

RSG_FN float resolve_lower(Solver& S, SRow& c) {
    SBody& A = S.bodies[c.a];
    SBody& B = S.bodies[c.b];
    float delta = c.rhs - c.applied * c.cfm;
    float d1 = sdot(c.n1, A.dv) + sdot(c.rc1, A.dw);
    float d2 = sdot(c.n2, B.dv) + sdot(c.rc2, B.dw);
    delta = delta - d1 * c.jac_inv;
    delta = delta - d2 * c.jac_inv;
    float sum = c.applied + delta;
    if (sum < c.lo) { delta = c.lo - c.applied; c.applied = c.lo; }
    else c.applied = sum;
    if (A.body >= 0) { A.dv += (c.n1 * A.inv_mass) * delta; A.dw += c.ang_a * delta; }
    if (B.body >= 0) { B.dv += (c.n2 * B.inv_mass) * delta; B.dw += c.ang_b * delta; }
    return delta / c.jac_inv;
}

RSG_FN float resolve_generic(Solver& S, SRow& c) {
    SBody& A = S.bodies[c.a];
    SBody& B = S.bodies[c.b];
    float delta = c.rhs - c.applied * c.cfm;
    float d1 = sdot(c.n1, A.dv) + sdot(c.rc1, A.dw);
    float d2 = sdot(c.n2, B.dv) + sdot(c.rc2, B.dw);
    delta = delta - d1 * c.jac_inv;
    delta = delta - d2 * c.jac_inv;
    float sum = c.applied + delta;
    bool lower_less = sum < c.lo;
    bool upper_less = sum < c.hi;
    float low_min_applied = c.lo - c.applied;
    delta = lower_less ? low_min_applied : delta;
    float app = lower_less ? c.lo : sum;
    float upper_min_applied = c.hi - c.applied;
    delta = upper_less ? delta : upper_min_applied;
    c.applied = upper_less ? app : c.hi;
    if (A.body >= 0) { A.dv += (c.n1 * A.inv_mass) * delta; A.dw += c.ang_a * delta; }
    if (B.body >= 0) { B.dv += (c.n2 * B.inv_mass) * delta; B.dw += c.ang_b * delta; }
    return delta / c.jac_inv;
}

RSG_FN float resolve_split(Solver& S, SRow& c) {
    if (!(c.rhs_pen != 0.f)) return 0.f;
    SBody& A = S.bodies[c.a];
    SBody& B = S.bodies[c.b];
    float delta = c.rhs_pen - c.applied_push * c.cfm;
    float d1 = sdot(c.n1, A.push) + sdot(c.rc1, A.turn);
    float d2 = sdot(c.n2, B.push) + sdot(c.rc2, B.turn);
    delta = delta - d1 * c.jac_inv;
    delta = delta - d2 * c.jac_inv;
    float sum = c.applied_push + delta;
    if (sum < c.lo) { delta = c.lo - c.applied_push; c.applied_push = c.lo; }
    else c.applied_push = sum;
    if (A.body >= 0) { A.push += (c.n1 * A.inv_mass) * delta; A.turn += c.ang_a * delta; }
    if (B.body >= 0) { B.push += (c.n2 * B.inv_mass) * delta; B.turn += c.ang_b * delta; }
    return delta * (1.f / c.jac_inv);
}

RSG_FN void solver_iterate(Solver& S) {
    // split impulse passes (:1778-1810), early out when every row is settled
    for (int it = 0; it < SOLVER_ITERATIONS; it++) {
        float lsq = 0.f;
        for (int j = 0; j < S.n_rows; j++) {
            float r = resolve_split(S, S.rows[j]);
            lsq = rs_max(lsq, r * r);
        }
        if (lsq <= 0.f || it >= SOLVER_ITERATIONS - 1) break;
    }
    for (int it = 0; it < SOLVER_ITERATIONS; it++) {
        for (int j = 0; j < S.n_rows; j++) {
            SRow& c = S.rows[j];
            if (c.special) continue;
            resolve_lower(S, c);
        }
        for (int j = 0; j < S.n_fric; j++) {
            SRow& f = S.fric[j];
            float total = S.rows[f.friction_index].applied;
            if (total > 0.f) {
                f.lo = -(f.friction * total);
                f.hi = f.friction * total;
                resolve_generic(S, f);
            }
        }
    }
}

}  // namespace rsg
