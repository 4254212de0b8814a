// rsg_collide.h -- the world step of one arena: gravity, damping, collision
// detection, contact solve and integration
// (btDiscreteDynamicsWorld::stepSimulation / internalSingleStepSimulation,
// BulletDynamics/Dynamics/btDiscreteDynamicsWorld.cpp:325-437).
#pragma once

#include "rsg_arena.h"

namespace rsg {

// Placeholder mesh type for arenas without a static world (THE_VOID).
struct NoMesh {};

template <class MeshT>
struct World {
    ArenaLocal& A;
    const MeshT& mesh;
    bool soccar;

    RSG_FN World(ArenaLocal& a, const MeshT& m, bool s) : A(a), mesh(m), soccar(s) {}

    // Dynamic body index: 0 = ball, 1 + i = car i.
    RSG_FN Body* body_ptr(int i) { return i == 0 ? &A.ball.b : &A.cars[i - 1].b; }
    RSG_FN int num_bodies() const { return 1 + A.n_cars; }

    RSG_FN RayHit raycast(V3 from, V3 to, int self_body) {
        RayHit h;
        h.hit = false;
        h.is_static = false;
        h.body = -1;
        h.fraction = 1.f;
        h.point = to;
        h.normal = v3zero();
        (void)from; (void)self_body;
        return h;
    }

    // btDiscreteDynamicsWorld::stepSimulation(dt, 0, dt)
    RSG_FN void step(float dt, const TickParams& P) {
        (void)P;
        const int nb = num_bodies();
        // applyGravity: active bodies only (btDiscreteDynamicsWorld.cpp:265-276)
        const V3 g = v3(0.f, 0.f, GRAVITY_Z) * UU_TO_BT;
        for (int i = 0; i < nb; i++) {
            Body& b = *body_ptr(i);
            if (b.active) b.force += g * (1.f / b.inv_mass);
        }
        // predictUnconstraintMotion: damping on every dynamic body
        // (btRigidBody::applyDamping, btRigidBody.cpp:153-207); the ball has
        // linear damping BALL_DRAG, cars none.
        {
            Body& bb = A.ball.b;
            bb.vel *= powf(1.f - BALL_DRAG, dt);
            bb.ang *= powf(1.f - 0.f, dt);
        }
        // Solver body write-back for every body in an active island
        // (btSequentialImpulseConstraintSolver.cpp:458-491, 1866-1890,
        // btSolverBody.h:265-287): v += F/m dt, w += (T * invI) dt.
        for (int i = 0; i < nb; i++) {
            Body& b = *body_ptr(i);
            if (!b.active) continue;
            V3 ext_f = b.force * b.inv_mass * dt;
            V3 ext_t = mul_vm(b.torque, b.inv_iw) * dt;
            b.vel = (b.vel + v3zero()) + ext_f;
            b.ang = (b.ang + v3zero()) + ext_t;
        }
        // integrateTransforms: active bodies (btDiscreteDynamicsWorld.cpp:889-983)
        for (int i = 0; i < nb; i++) {
            Body& b = *body_ptr(i);
            if (!b.active) continue;
            V3 np; M3 nr;
            integrate_transform(b.pos, b.rot, b.vel, b.ang, dt, b.no_rot, np, nr);
            b.pos = np;
            b.rot = nr;
            body_update_inertia(b);
        }
        // clearForces
        for (int i = 0; i < nb; i++) {
            Body& b = *body_ptr(i);
            b.force = v3zero();
            b.torque = v3zero();
        }
    }
};

}  // namespace rsg
