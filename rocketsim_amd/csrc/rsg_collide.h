// rsg_collide.h -- the world step of one arena
// (btDiscreteDynamicsWorld::stepSimulation / internalSingleStepSimulation,
// BulletDynamics/Dynamics/btDiscreteDynamicsWorld.cpp:325-437):
//   gravity -> damping + predicted poses -> AABBs -> broadphase pairs ->
//   narrowphase + contact callbacks -> islands -> contact solve -> integrate.
//
// Pair order, manifold order and the island sort follow the reference
// exactly, since the Gauss-Seidel solve depends on the order of its rows:
//   * pairs: for each dynamic proxy in handle order (ball, cars), the static
//     objects registered in its broadphase cell, then the other dynamic
//     proxies of the cell (btRSBroadphase.cpp:392-468);
//   * manifolds: created in dispatch order (btCollisionDispatcher.cpp:68-115),
//     compound children release theirs when their AABB stops overlapping
//     (btCompoundCollisionAlgorithm.cpp:300-338);
//   * islands: union-find with path compression over the pairs, manifolds
//     sorted by island id with Bullet's (unstable) quicksort
//     (btSimulationIslandManager.cpp:196-445, LinearMath/btAlignedObjectArray.h).
#pragma once

#include "rsg_solver.h"
#include "rsg_gjk.h"
#include "rsg_boxbox.h"

namespace rsg {

constexpr int MAX_BODIES = 1 + RSG_MAX_CARS;
// Overlapping pairs of a tick: every dynamic pair, and up to 16 statics per
// body on average (a body's box meets a few of an arena's statics; more pairs
// set RSG_STATUS_CONTACT_OVERFLOW and are dropped).
constexpr int MAX_PAIRS = MAX_BODIES * 16 + MAX_BODIES * (MAX_BODIES - 1) / 2;

struct Pair {
    int16_t a, b;  // body codes, a has the lower unique id (statics < ball < cars)
};

// Bullet's btAlignedObjectArray::quickSortInternal (Hoare partition on the
// middle element) over manifold indices keyed by island id.
RSG_FN void quicksort_by_key(int* idx, const int* key, int lo, int hi) {
    // explicit stack instead of recursion (same visiting order)
    int stack_lo[32], stack_hi[32];
    int sp = 0;
    stack_lo[sp] = lo; stack_hi[sp] = hi; sp++;
    while (sp > 0) {
        sp--;
        int l = stack_lo[sp], h = stack_hi[sp];
        int i = l, j = h;
        int x = key[idx[(l + h) / 2]];
        do {
            while (key[idx[i]] < x) i++;
            while (x < key[idx[j]]) j--;
            if (i <= j) {
                int t = idx[i]; idx[i] = idx[j]; idx[j] = t;
                i++; j--;
            }
        } while (i <= j);
        // recursion order: (l, j) first, then (i, h) -- push in reverse
        if (i < h) { stack_lo[sp] = i; stack_hi[sp] = h; sp++; }
        if (l < j) { stack_lo[sp] = l; stack_hi[sp] = j; sp++; }
    }
}

// The world step's per-arena tables: proxy bounds, cell-list order, the
// overlapping pairs of the tick and the pairs that get a contact manifold
// (plan[i] = pair index; manifold i belongs to it).
struct WorldTables {
    V3 aabb_min[MAX_BODIES], aabb_max[MAX_BODIES];
    int order[MAX_BODIES];  // dynamic proxies in cell-list order
    Pair pairs[MAX_PAIRS];
    int n_pairs;
    int plan[MAX_MANIFOLDS];
    int n_mf;
};

// The island pass's tables (btSimulationIslandManager): union-find over the
// dynamic bodies, the bodies sorted by island, the island manifolds.
struct IslandTables {
    int uf_id[MAX_BODIES], tag[MAX_BODIES];
    int elem[MAX_BODIES], ekey[MAX_BODIES];
    int bodies[MAX_BODIES];
    int im[MAX_MANIFOLDS], ikey[MAX_MANIFOLDS];
};

// What the contact-added callbacks do to the arena (Arena.cpp:244-414),
// recorded while the pairs' narrowphase runs and applied afterwards in pair
// order, which is the order the reference's dispatcher meets the contacts
// in.  The narrowphase reads only poses, shapes and the pre-tick contact
// response flags, none of which these effects change, so recording them
// changes no result.
enum EffectKind { EFF_CAR_WORLD = 1, EFF_BALL_WORLD = 2, EFF_CAR_BALL = 3, EFF_CAR_CAR = 4 };
struct Effect {
    int code;   // kind | car a << 8 | car b << 16
    V3 v0, v1;  // normal (world contacts) or local contact points
};
// Effects kept per pair (a host test build may lower it: RSG_TEST_PAIR_EFFECTS)
#ifdef RSG_TEST_PAIR_EFFECTS
constexpr int PAIR_EFFECTS = RSG_TEST_PAIR_EFFECTS;
#else
constexpr int PAIR_EFFECTS = 4;
#endif
struct PairEffects {
    int n;
    int redo;  // more effects than fit, or the pair needs the penetration pass: run it again in order
    Effect e[PAIR_EFFECTS];
};
// One arena's record of a shared narrowphase, one entry per manifold.
struct PairLog {
    PairEffects eff[MAX_MANIFOLDS];
};

// Per-arena work space of the world step: contact manifolds, solver rows,
// the penetration query's pools and the tables above (about 60 KB).
// Callers place it in private memory (one lane per arena) or in a per-arena
// block of global memory, where a lane's consecutive words share cache
// lines (private memory interleaves the lanes of a wave word by word).
struct StepSpace {
    Manifold mf[MAX_MANIFOLDS];
    SRow rows[MAX_ROWS];
    SRow fric[MAX_ROWS];
    EpaSpace epa;
    WorldTables wt;
    IslandTables isl;
    Solver solver;
};
// A kernel that keeps one StepSpace per lane in private memory needs its
// size times the lanes the device holds in flight (64 per wave, 1024 waves
// at one wave per SIMD) within the 4 GiB a queue's scratch can address.
static_assert(sizeof(StepSpace) <= 60 * 1024, "StepSpace must fit a lane's private memory");
// The plain kernel's lane holds a StepSpace, an ArenaLocal and about 2 KB of
// other private data within the device's 64 KiB per lane.
static_assert(sizeof(StepSpace) + sizeof(ArenaLocal) <= 62 * 1024, "the plain kernel's lane must fit 64 KiB");

// An arena's block of a batch's global work space: the world step's work
// space and the record of a shared narrowphase.
struct ArenaSpace {
    StepSpace step;
    PairLog log;
};

struct World {
    ArenaLocal& A;
    const ArenaGeom* geo;  // the arena's statics (null without an arena)
    bool soccar;
    const MeshView* mesh;  // the mesh of the static pair being run
    StepSpace* ws;  // null for worlds that only cast rays

    WorldTables* T;  // ws->wt
    int ray_order[MAX_BODIES];  // dynamic proxies in cell-list order as of the start of the tick (wheel rays)
    bool ray_order_ready;
    Manifold* mf;   // ws->mf
    // Narrowphase of one planned pair: where its callbacks' effects go (null:
    // applied at once) and the penetration work space (null: a query that
    // needs it marks the pair for a second, in-order run).
    PairEffects* rec;
    EpaSpace* epa;
    bool aborted;

    RSG_FN World(ArenaLocal& a, const ArenaGeom* g, bool s, StepSpace* w = nullptr)
        : A(a), geo(g), soccar(s), mesh(nullptr), ws(w), T(w ? &w->wt : nullptr), ray_order_ready(false), mf(w ? w->mf : nullptr),
          rec(nullptr), epa(w ? &w->epa : nullptr), aborted(false) {}

    // Whether the subsimplex ray cast against a convex of bounding radius
    // `radius` about `center` can report a hit with fraction below 1: only
    // if the segment passes within the cast's convergence distance of the
    // shape.  Accepted hits of the reference's cast lie within 0.01 of the
    // bounding sphere (largest seen 0.0081 over 3e5 near-miss rays), so
    // segments farther than 0.1 outside it are skipped without changing any
    // result.
    RSG_FN static bool ray_may_hit(V3 from, V3 to, V3 center, float radius) {
        V3 d = to - from;
        float dd = dot(d, d);
        float t = dd > 0.f ? dot(center - from, d) / dd : 0.f;
        t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
        V3 p = from + d * t;
        float lim = radius + 0.1f;
        return len2(center - p) <= lim * lim;
    }

    // Dynamic body index: 0 = ball, 1 + i = car i.
    RSG_FN Body* body_ptr(int i) { return i == 0 ? &A.ball.b : &A.cars[i - 1].b; }
    RSG_FN int num_bodies() const { return 1 + A.n_cars; }
    RSG_FN bool is_car(int k) const { return k >= 1; }
    // Contact response: a car demolished before this tick (or respawning in it)
    // has CF_NO_CONTACT_RESPONSE for the whole tick (Car::_PreTickUpdate,
    // src/Sim/Car/Car.cpp:67-83); a demolition during the tick takes effect on
    // the next one, so this reads the pre-tick state kept in Body::active.
    RSG_FN bool has_response(int code) {
        if (code <= 0) return true;
        return A.cars[code - 1].b.active;
    }
    // Contact response of a ray's closest hit.  Serially this is the body's
    // current flag.  When all wheel rays are cast at the start of the
    // pre-tick (ray_caster_pos >= 0: the casting car's position in the car
    // update order), a car updated before the caster already carries this
    // tick's flag (no response iff demolished), one updated after it still
    // carries the previous tick's.
    int ray_caster_pos = -1;
    RSG_FN bool ray_target_responds(int code) {
        if (code <= 0) return true;
        const Car& t = A.cars[code - 1];
        if (ray_caster_pos < 0) return t.b.active;
        int pos = 0;
        while (pos < A.n_cars && A.order[pos] != code - 1) pos++;
        return pos < ray_caster_pos ? !flag(t.flags, RSG_CF_DEMOED) : t.b.active;
    }
    RSG_FN bool active(int code) {
        if (code < 0) return false;
        return body_ptr(code)->active;
    }

    RSG_FN Xf xf_of(int code) {
        if (code >= 0) { Body* b = body_ptr(code); return xf_from(b->pos, b->rot); }
        if (static_is_mesh(*geo, -1 - code)) return xf_ident();
        return xf_from(static_plane(*geo, -1 - code).origin, m3_identity());
    }

    RSG_FN float threshold_of(int code) {
        if (code == 0) return make_ball_phys(A.ball.radius_uu, A.ball.mass).angular_motion_disc * CONTACT_BREAKING_THRESHOLD;
        if (code > 0) return A.cars[code - 1].phys.angular_motion_disc * CONTACT_BREAKING_THRESHOLD;
        return 1e30f;  // planes and the arena meshes have huge bounding spheres
    }

    RSG_FN void open_manifold(int mi, int b0, int b1) {
        Manifold& m = mf[mi];
        m.b0 = b0; m.b1 = b1;
        m.threshold = rs_min(threshold_of(b0), threshold_of(b1));
        m.n = 0;
    }

    // -----------------------------------------------------------------------
    // AABBs (btCollisionWorld::updateSingleAabb): shape AABB at the current
    // and predicted pose, merged, +- the global contact threshold.
    RSG_FN void shape_aabb(int code, V3 pos, const M3& rot, V3& mn, V3& mx) {
        if (code == 0) {
            float r = make_ball_phys(A.ball.radius_uu, A.ball.mass).radius + 0.08f;  // btSphereShape::getAabb (RocketSim +0.08)
            mn = pos - v3(r, r, r);
            mx = pos + v3(r, r, r);
        } else {
            const CarPhys& cp = A.cars[code - 1].phys;
            M3 ab = abs_m3(rot);
            V3 c = rot * cp.local_aabb_center + pos;
            V3 e = v3(dot(ab.r[0], cp.local_aabb_half), dot(ab.r[1], cp.local_aabb_half), dot(ab.r[2], cp.local_aabb_half));
            mn = c - e;
            mx = c + e;
        }
    }

    // Child box AABB of a car (btBoxShape::getAabb at hull transform, no threshold).
    RSG_FN void car_child_aabb(int code, V3& mn, V3& mx) {
        Car& c = A.cars[code - 1];
        const CarPhys& cp = c.phys;
        Xf x = xf_mul(xf_from(c.b.pos, c.b.rot), xf_from(cp.box_offset, m3_identity()));
        V3 he = cp.box_implicit + v3(cp.box_margin, cp.box_margin, cp.box_margin);
        M3 ab = abs_m3(x.R);
        V3 e = v3(dot(ab.r[0], he), dot(ab.r[1], he), dot(ab.r[2], he));
        mn = x.t - e;
        mx = x.t + e;
    }

    RSG_FN bool static_in_cell(int s, int c) {
        if (!soccar) return false;
        return (geo->cell_statics[c] >> s) & 1u;
    }

    // -----------------------------------------------------------------------
    // Ray cast for the wheels (btDefaultVehicleRaycaster::castRay,
    // BulletDynamics/Vehicle/btDefaultVehicleRaycaster.cpp:32-53): closest hit
    // over the statics registered in the ray start's cell, then the dynamic
    // bodies of that cell.
    RSG_FN RayHit raycast(V3 from, V3 to, int self_body) {
        RayHit h;
        h.hit = false;
        h.is_static = false;
        h.body = -1;
        h.fraction = 1.f;
        h.point = to;
        h.normal = v3zero();
        if (!soccar && num_bodies() < 2) return h;
        int c = bp_cell_of(from);
        float frac = 1.f;
        V3 n = v3zero();
        int hit_obj = -100;
        // static objects registered in the cell, in registration order
        const int ns = soccar ? num_statics(*geo) : 0;
        for (int s = 0; s < ns; s++) {
            if (!static_in_cell(s, c) || ((geo->car_skip >> s) & 1u)) continue;
            if (frac == 0.f) break;
            if (static_is_mesh(*geo, s)) {
                const MeshView& m = geo->meshes[s];
                struct RayCB {
                    const MeshView* m; V3 from, to; float* frac; V3* n; bool hit;
                    RSG_FN void operator()(int t) {
                        V3 v[3];
                        mesh_triangle(*m, t, v);
                        if (ray_triangle(from, to, v[0], v[1], v[2], *frac, *n)) hit = true;
                    }
                } cb{&m, from, to, &frac, &n, false};
                mesh_query_ray(m, from, to, cb);
                if (cb.hit) hit_obj = -1 - s;
            } else {
                if (ray_plane(static_plane(*geo, s), from, to, frac, n)) hit_obj = -1 - s;
            }
        }
        // then the dynamic proxies registered in the cell (cell-list order),
        // except the casting car (btCollisionWorld::RayResultCallback::needsCollision)
        const int nb = num_bodies();
        if (!ray_order_ready) {
            // registration order does not change before the world step
            for (int i = 0; i < nb; i++) ray_order[i] = i;
            for (int i = 1; i < nb; i++) {
                int v = ray_order[i], j = i - 1;
                while (j >= 0 && A.bp_seq[ray_order[j]] > A.bp_seq[v]) { ray_order[j + 1] = ray_order[j]; j--; }
                ray_order[j + 1] = v;
            }
            ray_order_ready = true;
        }
        for (int r = 0; r < nb; r++) {
            int q = ray_order[r];
            if (!bp_cells_adjacent(A.bp_cell[q], c)) continue;
            if (frac == 0.f) break;
            if (q == self_body) continue;
            float f;
            V3 nn;
            if (q == 0) {
                CShape sph = shape_sphere(make_ball_phys(A.ball.radius_uu, A.ball.mass).radius);
                if (!ray_may_hit(from, to, A.ball.b.pos, sph.margin)) continue;
                if (ray_convex(from, to, sph, xf_from(A.ball.b.pos, A.ball.b.rot), f, nn)) {
                    if (len2(nn) > 0.0001f && f < frac) { frac = f; n = normalized(nn); hit_obj = q; }
                }
            } else {
                CShape box = car_box_shape(q);
                Xf cx = car_child_xf(q);
                if (!ray_may_hit(from, to, cx.t, len(box.half + v3(box.margin, box.margin, box.margin)))) continue;
                if (ray_convex(from, to, box, cx, f, nn)) {
                    if (len2(nn) > 0.0001f && f < frac) { frac = f; n = normalized(nn); hit_obj = q; }
                }
            }
        }
        if (hit_obj != -100) {
            // btDefaultVehicleRaycaster::castRay: the closest hit must respond
            if (hit_obj >= 0 && !ray_target_responds(hit_obj)) return h;
            h.hit = true;
            h.is_static = hit_obj < 0;
            h.body = hit_obj >= 0 ? hit_obj : -1;
            h.fraction = frac;
            float sm = 1.f - frac;
            h.point = from * sm + to * frac;
            h.normal = normalized(n);
        }
        return h;
    }

    // -----------------------------------------------------------------------
    // Contact point creation (btManifoldResult::addContactPoint) for a
    // manifold whose body0 is dynamic `b0`; `normal` points from B to A,
    // `point_b` lies on B, depth < 0 is penetration.
    RSG_FN int add_contact(int mi, V3 normal, V3 point_b, float depth, int tri) {
        Manifold& m = mf[mi];
        if (depth > m.threshold) return -1;
        V3 point_a = point_b + normal * depth;
        CPoint p;
        p.local_a = xf_inv_apply(xf_of(m.b0), point_a);
        p.local_b = xf_inv_apply(xf_of(m.b1), point_b);
        p.pos_a = point_a;
        p.pos_b = point_b;
        p.normal = normal;
        p.dist = depth;
        p.special = false;
        p.tri = tri;
        // combined friction / restitution (btManifoldResult.cpp:56-77)
        float fa, ra, fb, rb;
        body_material(m.b0, fa, ra);
        body_material(m.b1, fb, rb);
        bool stat = m.b0 < 0 || m.b1 < 0;
        p.friction = stat ? (fa < fb ? fa : fb) : fa * fb;
        p.restitution = stat ? (ra > rb ? ra : rb) : ra * rb;
        int idx = manifold_add_point(m, p);
        contact_added(m.p[idx], m.b0, m.b1);
        return idx;
    }

    RSG_FN void body_material(int code, float& f, float& r) {
        if (code < 0) { f = ARENA_COLLISION_BASE_FRICTION; r = ARENA_COLLISION_BASE_RESTITUTION; }
        else if (code == 0) { f = A.mut.ball_world_friction; r = A.mut.ball_world_restitution; }
        else { f = CAR_COLLISION_FRICTION; r = CAR_COLLISION_RESTITUTION; }
    }

    // Arena::_BulletContactAddedCallback (src/Sim/Arena/Arena.cpp:244-314):
    // the point's combined material and flags here, the arena's side through
    // effect().
    RSG_FN void contact_added(CPoint& cp, int b0, int b1) {
        if (!has_response(b0) || !has_response(b1)) return;
        if (b0 > 0 && b1 < 0) {
            // car + world (Arena.cpp:407-414)
            cp.friction = A.mut.car_world_friction;
            cp.restitution = A.mut.car_world_restitution;
            effect(EFF_CAR_WORLD, b0 - 1, 0, cp.normal, v3zero());
        } else if (b0 == 0 && b1 < 0) {
            // ball + world: Ball::_OnWorldCollision, then the averaged
            // "special" row (Arena.cpp:296-304)
            if (A.mode == RSG_MODE_HEATSEEKER) effect(EFF_BALL_WORLD, 0, 0, cp.normal, v3zero());
            cp.special = true;
        } else if (b0 > 0 && b1 == 0) {
            cp.friction = CARBALL_COLLISION_FRICTION;
            cp.restitution = CARBALL_COLLISION_RESTITUTION;
            effect(EFF_CAR_BALL, b0 - 1, 0, cp.local_b, v3zero());
        } else if (b0 > 0 && b1 > 0) {
            cp.friction = CARCAR_COLLISION_FRICTION;
            cp.restitution = CARCAR_COLLISION_RESTITUTION;
            effect(EFF_CAR_CAR, b0 - 1, b1 - 1, cp.local_a, cp.local_b);
        }
        if (b1 < 0 && cp.tri >= 0 && static_is_mesh(*geo, -1 - b1)) adjust_internal_edge(cp, geo->meshes[-1 - b1], cp.tri);
    }

    RSG_FN void effect(int kind, int a, int b, V3 v0, V3 v1) {
        if (!rec) { apply_effect(kind, a, b, v0, v1); return; }
        if (kind == EFF_CAR_WORLD && rec->n > 0 && (rec->e[rec->n - 1].code & 0xff) == EFF_CAR_WORLD) {
            rec->e[rec->n - 1].v0 = v0;  // the flag is set either way; the last normal stays
            return;
        }
        if (rec->n >= PAIR_EFFECTS) { rec->redo = 1; return; }
        Effect& e = rec->e[rec->n++];
        e.code = kind | (a << 8) | (b << 16);
        e.v0 = v0;
        e.v1 = v1;
    }

    RSG_FN void apply_effect(int kind, int a, int b, V3 v0, V3 v1) {
        if (kind == EFF_CAR_WORLD) {
            Car& c = A.cars[a];
            setflag(c.flags, RSG_CF_WORLD_CONTACT, true);
            c.world_normal = v0;
        } else if (kind == EFF_BALL_WORLD) {
            ball_world_hit_heatseeker(v0);
        } else if (kind == EFF_CAR_BALL) {
            car_ball_hit(a, v0);
        } else if (kind == EFF_CAR_CAR) {
            car_car_hit(a, b, v0, v1);
        }
    }

    // Ball::_OnWorldCollision for Heatseeker (src/Sim/Ball/Ball.cpp:326-358):
    // a hit on the back wall of the net the ball seeks turns it round, with
    // an extra impulse towards the other net.  `normal` points from the world
    // to the ball.
    RSG_FN void ball_world_hit_heatseeker(V3 normal) {
        Ball& ball = A.ball;
        if (ball.hs_dir == 0) return;
        V3 pos = ball.b.pos * BT_TO_UU;
        float rel_ny = normal.y * ball.hs_dir, rel_y = pos.y * ball.hs_dir;
        if (rel_ny <= -HS_WALL_BOUNCE_CHANGE_Y_NORMAL && rel_y >= ARENA_EXTENT_Y - HS_WALL_BOUNCE_CHANGE_Y_THRESH) {
            ball.hs_dir *= -1;
            V3 vel = ball.b.vel * BT_TO_UU;
            V3 to_goal = rs_normalized(v3(0, HS_TARGET_Y * ball.hs_dir, HS_TARGET_Z) - pos);
            V3 dir = to_goal * (1 - HS_WALL_BOUNCE_UP_FRAC) + v3(0, 0, 1) * HS_WALL_BOUNCE_UP_FRAC;
            V3 impulse = dir * len(vel) * HS_WALL_BOUNCE_FORCE_SCALE;
            ball.vel_cache += impulse * UU_TO_BT;
        }
    }

    // Arena::_BtCallback_OnCarBallCollision + Ball::_OnHit
    // (src/Sim/Arena/Arena.cpp:316-322, src/Sim/Ball/Ball.cpp:236-290).  The
    // manifold's body0 is the car, so the point on the ball is local point B.
    RSG_FN void car_ball_hit(int ci, V3 local_b) {
        Car& c = A.cars[ci];
        Body& bb = A.ball.b;
        setflag(c.flags, RSG_CF_HIT_VALID, true);
        c.hit_rel_pos = local_b * BT_TO_UU;
        c.hit_tick = A.tick;
        V3 ball_pos = bb.pos * BT_TO_UU;
        c.hit_ball_pos = ball_pos;
        c.hit_extra_vel = v3zero();
        if ((A.tick > c.hit_extra_tick + 1) || (c.hit_extra_tick > A.tick)) {
            c.hit_extra_tick = A.tick;
            V3 fwd = car_forward(c);
            V3 rel_pos = ball_pos - c.b.pos * BT_TO_UU;
            V3 rel_vel = bb.vel * BT_TO_UU - c.b.vel * BT_TO_UU;
            float rel_speed = rs_min(len(rel_vel), BALL_CAR_EXTRA_IMPULSE_MAXDELTAVEL_UU);
            if (rel_speed > 0) {
                // Hoops lifts harder for a car on the ground (Ball.cpp:272-276)
                const bool hoops_z = A.mode == RSG_MODE_HOOPS && flag(c.flags, RSG_CF_ON_GROUND) &&
                                     car_up(c).z > BALL_CAR_EXTRA_IMPULSE_Z_SCALE_HOOPS_NORMAL_Z_THRESH;
                const float z_scale = hoops_z ? BALL_CAR_EXTRA_IMPULSE_Z_SCALE_HOOPS_GROUND : BALL_CAR_EXTRA_IMPULSE_Z_SCALE;
                V3 hit_dir = rs_normalized(rel_pos * v3(1, 1, z_scale));
                V3 fwd_adj = fwd * dot(hit_dir, fwd) * (1 - BALL_CAR_EXTRA_IMPULSE_FORWARD_SCALE);
                hit_dir = rs_normalized(hit_dir - fwd_adj);
                V3 added = (hit_dir * rel_speed) * curve_eval(curve_ball_extra_impulse(), rel_speed) *
                           A.mut.ball_hit_extra_force_scale;
                c.hit_extra_vel = added;
                A.ball.vel_cache += added * UU_TO_BT;
            }
            if (A.mode == RSG_MODE_HEATSEEKER) {
                // Ball.cpp:291-299: a touch by the other team speeds the
                // ball up (at most once a second) and retargets it
                Ball& ball = A.ball;
                bool can_increase = (ball.hs_time > HS_MIN_SPEEDUP_INTERVAL) || (ball.hs_dir == 0);
                float new_dir = c.team == RSG_TEAM_BLUE ? 1.f : -1.f;
                if (can_increase && new_dir != ball.hs_dir) {
                    ball.hs_time = 0;
                    ball.hs_speed = rs_min(ball.hs_speed + HS_TARGET_SPEED_INCREMENT, HS_MAX_SPEED);
                }
                ball.hs_dir = new_dir;
            }
        }
    }

    // Arena::_BtCallback_OnCarCarCollision (src/Sim/Arena/Arena.cpp:324-395):
    // bumps and demolitions, tested from both cars.  car1 is the manifold's
    // body0 (point A), car2 its body1 (point B).
    RSG_FN void car_car_hit(int c1, int c2, V3 local_a, V3 local_b) {
        for (int it = 0; it < 2; it++) {
            bool swapped = it == 1;
            Car& me = A.cars[swapped ? c2 : c1];
            Car& other = A.cars[swapped ? c1 : c2];
            if (flag(me.flags, RSG_CF_DEMOED) || flag(other.flags, RSG_CF_DEMOED)) return;
            if (me.other_id == other.id && me.car_cooldown > 0) continue;
            V3 vel = me.b.vel * BT_TO_UU;
            V3 other_vel = other.b.vel * BT_TO_UU;
            V3 delta_pos = other.b.pos * BT_TO_UU - me.b.pos * BT_TO_UU;
            if (dot(vel, delta_pos) > 0) {
                V3 vel_dir = rs_normalized(vel);
                V3 dir_to_other = rs_normalized(delta_pos);
                float speed_towards = dot(vel, dir_to_other);
                float other_away = dot(other_vel, vel_dir);
                if (speed_towards > other_away) {
                    V3 local = swapped ? local_b : local_a;
                    if ((local.x * BT_TO_UU) > BUMP_MIN_FORWARD_DIST) {
                        bool is_demo = A.mut.demo_mode == RSG_DEMO_ON_CONTACT    ? true
                                       : A.mut.demo_mode == RSG_DEMO_DISABLED ? false
                                                                              : flag(me.flags, RSG_CF_SUPERSONIC);
                        if (is_demo && !A.mut.enable_team_demos) is_demo = me.team != other.team;
                        emit_event(A, is_demo ? RSG_EVENT_DEMO : RSG_EVENT_BUMP, me.id, other.id);
                        if (is_demo) {
                            setflag(other.flags, RSG_CF_DEMOED, true);
                            other.demo_timer = A.mut.respawn_delay;
                        } else {
                            bool ground = flag(other.flags, RSG_CF_ON_GROUND);
                            float base = curve_eval(ground ? curve_bump_ground() : curve_bump_air(), speed_towards);
                            V3 up = ground ? car_up(other) : v3(0, 0, 1);
                            V3 imp = vel_dir * base + up * curve_eval(curve_bump_upward(), speed_towards) *
                                                      A.mut.bump_force_scale;
                            other.vel_cache += imp * UU_TO_BT;
                        }
                        me.other_id = other.id;
                        me.car_cooldown = A.mut.bump_cooldown_time;
                    }
                }
            }
        }
    }

    RSG_FN void refresh(int mi) {
        Manifold& m = mf[mi];
        if (m.n == 0) return;
        manifold_refresh(m, xf_of(m.b0), xf_of(m.b1));
    }

    // -----------------------------------------------------------------------
    // Narrowphase

    // btConvexPlaneCollisionAlgorithm::processCollision for the ball
    RSG_FN void ball_vs_plane(int mi, int s) {
        const PlaneDef& pd = static_plane(*geo, s);
        Xf px = xf_from(pd.origin, m3_identity());
        Body& b = A.ball.b;
        Xf bx = xf_from(b.pos, b.rot);
        Xf plane_in_convex = xf_mul(xf_inverse(bx), px);
        Xf convex_in_plane = xf_mul(xf_inverse(px), bx);
        V3 dir = plane_in_convex.R * -pd.n;
        V3 vtx;
        float r = make_ball_phys(A.ball.radius_uu, A.ball.mass).radius;
        if (len2(dir) < FLT_EPS * FLT_EPS) vtx = normalized(v3(-1, -1, -1)) * r;
        else vtx = r * normalized(dir);
        V3 vin = xf_apply(convex_in_plane, vtx);
        float distance = dot(pd.n, vin) - 0.f;
        if (distance < mf[mi].threshold + 0.f) {
            V3 proj = vin - distance * pd.n;
            V3 w = xf_apply(px, proj);
            V3 nB = px.R * pd.n;
            add_contact(mi, nB, w, distance, -1);
        }
        refresh(mi);
    }

    // Child box of a car vs a plane (same algorithm, box support with margin)
    RSG_FN void car_vs_plane(int mi, int code, int s) {
        const PlaneDef& pd = static_plane(*geo, s);
        Xf px = xf_from(pd.origin, m3_identity());
        Car& c = A.cars[code - 1];
        const CarPhys& cp = c.phys;
        Xf cx = xf_mul(xf_from(c.b.pos, c.b.rot), xf_from(cp.box_offset, m3_identity()));
        Xf plane_in_convex = xf_mul(xf_inverse(cx), px);
        Xf convex_in_plane = xf_mul(xf_inverse(px), cx);
        V3 dir = plane_in_convex.R * -pd.n;
        V3 he = cp.box_implicit + v3(cp.box_margin, cp.box_margin, cp.box_margin);
        V3 vtx = v3(dir.x >= 0 ? he.x : -he.x, dir.y >= 0 ? he.y : -he.y, dir.z >= 0 ? he.z : -he.z);
        V3 vin = xf_apply(convex_in_plane, vtx);
        float distance = dot(pd.n, vin) - 0.f;
        if (distance < mf[mi].threshold + 0.f) {
            V3 proj = vin - distance * pd.n;
            V3 w = xf_apply(px, proj);
            V3 nB = px.R * pd.n;
            add_contact(mi, nB, w, distance, -1);
        }
        refresh(mi);
    }

    // Sphere vs one mesh triangle (SphereTriangleDetector::collide, with the
    // closest-point-on-triangle capsule test RocketSim uses).
    RSG_FN void sphere_triangle(int mi, const V3 v[3], int tri) {
        Body& b = A.ball.b;
        V3 center = b.pos;  // triangle space = world (mesh at identity)
        float radius = make_ball_phys(A.ball.radius_uu, A.ball.mass).radius;
        float thr = mf[mi].threshold + 0.f;
        float rwt = radius + thr;
        V3 normal = cross(v[1] - v[0], v[2] - v[0]);
        float l2 = len2(normal);
        bool has = false;
        V3 cpt;
        if (l2 >= FLT_EPS * FLT_EPS) {
            normal = normal * (1.f / sqrtf(l2));
            V3 p1c = center - v[0];
            float dfp = dot(p1c, normal);
            if (dfp < 0.f) { dfp *= -1.f; normal = normal * -1.f; }
            if (dfp < rwt) {
                if (point_in_triangle(v, normal, center)) {
                    has = true;
                    cpt = center - normal * dfp;
                } else {
                    float min_d2 = rwt * rwt;
                    V3 ne = closest_point_triangle(center, v[0], v[1], v[2]);
                    V3 dd = ne - center;
                    float d2 = dot(dd, dd);
                    if (d2 < min_d2) { has = true; cpt = ne; }
                }
            }
        }
        if (!has) return;
        V3 c2c = center - cpt;
        float d2 = len2(c2c);
        if (!(d2 < rwt * rwt)) return;
        V3 rn, pt;
        float depth;
        if (d2 > FLT_EPS) {
            float dist = sqrtf(d2);
            rn = normalized(c2c);
            pt = cpt;
            depth = -(radius - dist);
        } else {
            rn = normal;
            pt = cpt;
            depth = -radius;
        }
        add_contact(mi, rn, pt, depth, tri);
    }

    // SphereTriangleDetector::pointInTriangle (RocketSim variant)
    RSG_FN static bool point_in_triangle(const V3 v[3], V3 normal, V3 p) {
        V3 e0 = v[1] - v[0], e1 = v[2] - v[1], e2 = v[0] - v[2];
        V3 n0 = cross(e0, normal), n1 = cross(e1, normal), n2 = cross(e2, normal);
        float r0 = dot(p - v[0], n0), r1 = dot(p - v[1], n1), r2 = dot(p - v[2], n2);
        return (r0 <= 0 && r1 <= 0 && r2 <= 0) || (r0 > 0 && r1 > 0 && r2 > 0);
    }

    // closestPointTriangle (Ericson, Real-Time Collision Detection 5.1.5)
    RSG_FN static V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c) {
        V3 ab = b - a, ac = c - a, ap = p - a;
        float d1 = dot(ab, ap), d2 = dot(ac, ap);
        if (d1 <= 0.f && d2 <= 0.f) return a;
        V3 bp = p - b;
        float d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= 0.f && d4 <= d3) return b;
        V3 cp = p - c;
        float d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d6 >= 0.f && d5 <= d6) return c;
        float vc = d1 * d4 - d3 * d2;
        if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) { float v = d1 / (d1 - d3); return a + v * ab; }
        float vb = d5 * d2 - d1 * d6;
        if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) { float v = d2 / (d2 - d6); return a + v * ac; }
        float va = d3 * d6 - d5 * d4;
        if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
            float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            return b + v * (c - b);
        }
        float denom = 1.f / (va + vb + vc);
        float v = vb * denom, w = vc * denom;
        return a + v * ab + w * ac;
    }

    RSG_FN Xf car_child_xf(int code) {
        Car& c = A.cars[code - 1];
        const CarPhys& cp = c.phys;
        return xf_mul(xf_from(c.b.pos, c.b.rot), xf_from(cp.box_offset, m3_identity()));
    }
    RSG_FN CShape car_box_shape(int code) {
        const CarPhys& cp = A.cars[code - 1].phys;
        return shape_box(cp.box_implicit, cp.box_margin);
    }

    // Car hitbox vs an arena mesh (World::mesh): btConvexConcaveCollisionAlgorithm with the
    // per-triangle early out, then btConvexConvexAlgorithm (GJK) per triangle.
    RSG_FN void car_vs_mesh(int mi, int code) {
        Xf cx = car_child_xf(code);
        const CarPhys& cp = A.cars[code - 1].phys;
        V3 he = cp.box_implicit + v3(cp.box_margin, cp.box_margin, cp.box_margin);
        M3 ab = abs_m3(cx.R);
        V3 e = v3(dot(ab.r[0], he), dot(ab.r[1], he), dot(ab.r[2], he));
        V3 amin = cx.t - e, amax = cx.t + e;
        struct CB {
            World* w; int mi; int code; V3 amin, amax; Xf cx; V3 he;
            RSG_FN void operator()(int t) { w->car_triangle(mi, code, t, amin, amax, cx, he); }
        } cb{this, mi, code, amin, amax, cx, he};
        mesh_query_aabb(*mesh, amin, amax, cb);
        if (!aborted) refresh(mi);
    }

    RSG_FN void car_triangle(int mi, int code, int t, V3 amin, V3 amax, const Xf& cx, V3 he) {
        if (aborted) return;
        V3 v[3];
        mesh_triangle(*mesh, t, v);
        if (!triangle_aabb_overlap(v, amin, amax)) return;
        float thr = mf[mi].threshold + 0.f;
        {
            V3 tn = normalized(cross(v[1] - v[0], v[2] - v[0]));
            M3 inv = inverse(cx.R);
            for (int side = 0; side < 2; side++) {
                V3 d = inv * tn;
                V3 lp = v3(fsel(d.x, he.x, -he.x), fsel(d.y, he.y, -he.y), fsel(d.z, he.z, -he.z));
                V3 wp = xf_apply(cx, lp);
                float dist = dot(tn, v[0]) - dot(tn, wp);
                if (dist > thr) return;
                tn = tn * -1.f;
            }
        }
        CShape box = car_box_shape(code);
        CShape tri = shape_triangle(v);
        float md = box.margin + tri.margin + mf[mi].threshold;
        md *= md;
        V3 n, pb;
        float dist;
        int r = gjk_closest(box, tri, cx, xf_ident(), md, n, pb, dist, epa);
        if (r < 0) aborted = true;
        else if (r) add_contact(mi, n, pb, dist, t);
    }

    // Car hitbox vs the ball: btConvexConvexAlgorithm (GJK) on the box and the
    // sphere; the manifold (car, ball) exists once the child AABBs overlap.
    RSG_FN void car_vs_ball(int mi, int code) {
        Xf cx = car_child_xf(code);
        CShape box = car_box_shape(code);
        CShape sph = shape_sphere(make_ball_phys(A.ball.radius_uu, A.ball.mass).radius);
        float md = box.margin + sph.margin + mf[mi].threshold;
        md *= md;
        V3 n, pb;
        float dist;
        int r = gjk_closest(box, sph, cx, xf_from(A.ball.b.pos, A.ball.b.rot), md, n, pb, dist, epa);
        if (r < 0) { aborted = true; return; }
        if (r) add_contact(mi, n, pb, dist, -1);
        refresh(mi);
    }

    // Car hitbox vs car hitbox: btBoxBoxCollisionAlgorithm (dBoxBox2) with
    // box0 = the car of manifold body0.
    RSG_FN void car_vs_car(int mi, int c0, int c1) {
        Xf x0 = car_child_xf(c0), x1 = car_child_xf(c1);
        const CarPhys& p0 = A.cars[c0 - 1].phys;
        const CarPhys& p1 = A.cars[c1 - 1].phys;
        V3 h0 = p0.box_implicit + v3(p0.box_margin, p0.box_margin, p0.box_margin);
        V3 h1 = p1.box_implicit + v3(p1.box_margin, p1.box_margin, p1.box_margin);
        BoxBoxContacts bc;
        box_box(x0, h0, x1, h1, bc);
        for (int i = 0; i < bc.n; i++) add_contact(mi, bc.normal, bc.point[i], bc.depth[i], -1);
        refresh(mi);
    }

    // btConvexConcaveCollisionAlgorithm::processCollision for the ball vs
    // an arena mesh (World::mesh)
    RSG_FN void ball_vs_mesh(int mi) {
        Body& b = A.ball.b;
        float r = make_ball_phys(A.ball.radius_uu, A.ball.mass).radius + 0.08f;
        V3 amin = b.pos - v3(r, r, r), amax = b.pos + v3(r, r, r);  // convex AABB in mesh space, + margin 0
        struct CB {
            World* w; int mi; V3 amin, amax;
            RSG_FN void operator()(int t) { w->ball_triangle(mi, t, amin, amax); }
        } cb{this, mi, amin, amax};
        mesh_query_aabb(*mesh, amin, amax, cb);
        refresh(mi);
    }

    RSG_FN void ball_triangle(int mi, int t, V3 amin, V3 amax) {
        V3 v[3];
        mesh_triangle(*mesh, t, v);
        if (!triangle_aabb_overlap(v, amin, amax)) return;
        // early out on the triangle plane (btConvexConcaveCollisionAlgorithm.cpp:86-125)
        {
            V3 tn = normalized(cross(v[1] - v[0], v[2] - v[0]));
            Body& b = A.ball.b;
            float r = make_ball_phys(A.ball.radius_uu, A.ball.mass).radius;
            V3 d = inverse(b.rot) * tn;  // getBasis().inverse() * n
            V3 lp = (len2(d) < FLT_EPS * FLT_EPS) ? normalized(v3(-1, -1, -1)) * r : r * normalized(d);
            V3 wp = b.rot * lp + b.pos;
            float thr = mf[mi].threshold + 0.f;
            float dist = dot(tn, v[0]) - dot(tn, wp);
            if (dist > thr) return;
            tn = tn * -1.f;
            d = inverse(b.rot) * tn;
            lp = (len2(d) < FLT_EPS * FLT_EPS) ? normalized(v3(-1, -1, -1)) * r : r * normalized(d);
            wp = b.rot * lp + b.pos;
            dist = dot(tn, v[0]) - dot(tn, wp);
            if (dist > thr) return;
        }
        sphere_triangle(mi, v, t);
    }

    // -----------------------------------------------------------------------
    // The step

    // btDiscreteDynamicsWorld::internalSingleStepSimulation
    // (btDiscreteDynamicsWorld.cpp:395-440) on one lane.
    RSG_FN void step(float dt) {
        step_begin(dt);
        run_pairs();
        step_end(dt, nullptr);
    }

    // Gravity, damping, bounds, broadphase and the manifold plan.
    RSG_FN void step_begin(float dt) {
        step_prologue(dt);
        for (int i = 0; i < num_bodies(); i++) body_aabb(i, dt);
        step_broadphase();
    }

    // The world step's start in parts: gravity and damping; each body's
    // proxy AABB (independent per body, so the arena's lanes may share them);
    // broadphase cells, pairs and the manifold plan.
    RSG_FN void step_prologue(float dt) {
        const int nb = num_bodies();
        // btRigidBody::applyGravity: each body's own gravity
        for (int i = 0; i < nb; i++) {
            Body& b = *body_ptr(i);
            if (b.active) b.force += b.grav;
        }
        // predictUnconstraintMotion (damping on every dynamic body; the
        // ball's linear damping is the arena's ballDrag, clamped to [0, 1]
        // by btRigidBody::setDamping)
        {
            Body& bb = A.ball.b;
            const float drag = fminf(fmaxf(A.mut.ball_drag, 0.f), 1.f);
            bb.vel *= powf(1.f - drag, dt);
            bb.ang *= powf(1.f - 0.f, dt);
        }
    }

    // Proxy AABB of body i: its shape at the current and the predicted pose,
    // merged, +- the contact threshold (btCollisionWorld::updateSingleAabb
    // over predictIntegratedTransform).
    RSG_FN void body_aabb(int i, float dt) {
        Body& b = *body_ptr(i);
        V3 pp; M3 pr;
        integrate_transform(b.pos, b.rot, b.vel, b.ang, dt, b.no_rot, pp, pr);
        V3 mn, mx, mn2, mx2;
        shape_aabb(i, b.pos, b.rot, mn, mx);
        V3 t = v3(CONTACT_BREAKING_THRESHOLD, CONTACT_BREAKING_THRESHOLD, CONTACT_BREAKING_THRESHOLD);
        mn = mn - t; mx = mx + t;
        shape_aabb(i, pp, pr, mn2, mx2);
        mn2 = mn2 - t; mx2 = mx2 + t;
        T->aabb_min[i] = v3(fminf(mn.x, mn2.x), fminf(mn.y, mn2.y), fminf(mn.z, mn2.z));
        T->aabb_max[i] = v3(fmaxf(mx.x, mx2.x), fmaxf(mx.y, mx2.y), fmaxf(mx.z, mx2.z));
    }

    RSG_FN void step_broadphase() {
        const int nb = num_bodies();
        for (int i = 0; i < nb; i++) {
            // btRSBroadphase::setAabb: a proxy whose cell changed is removed
            // from its old cells and appended to the new ones, i.e. it moves
            // to the end of every cell list it is in.
            int c = bp_cell_of(T->aabb_min[i]);
            if (c != A.bp_cell[i]) {
                A.bp_cell[i] = c;
                if (nb > 1) {
                    uint32_t mxs = 0;
                    for (int q = 0; q < nb; q++) mxs = A.bp_seq[q] > mxs ? A.bp_seq[q] : mxs;
                    A.bp_seq[i] = mxs + 1;
                }
            }
        }
        // cell-list order of the dynamic proxies (registration order)
        for (int i = 0; i < nb; i++) T->order[i] = i;
        for (int i = 1; i < nb; i++) {
            int v = T->order[i], j = i - 1;
            while (j >= 0 && A.bp_seq[T->order[j]] > A.bp_seq[v]) { T->order[j + 1] = T->order[j]; j--; }
            T->order[j + 1] = v;
        }
        // keep the order numbers small: rank them 0..nb-1
        for (int r = 0; r < nb; r++) A.bp_seq[T->order[r]] = (uint32_t)r;
        find_pairs();
        plan_pairs();
    }

    // The narrowphase's recorded effects (when it was shared: `log`),
    // islands and solver, integration.
    RSG_FN void step_end(float dt, const PairLog* log) {
        step_solve(dt, log);
        for (int i = 0; i < num_bodies(); i++) body_integrate(i, dt);
    }
    RSG_FN void step_solve(float dt, const PairLog* log) {
        if (log) apply_recorded(log);
        solve_islands(dt);
    }

    // integrateTransforms for body i (active bodies only), then
    // clearForces; independent per body.
    RSG_FN void body_integrate(int i, float dt) {
        Body& b = *body_ptr(i);
        if (b.active) {
            V3 np; M3 nr;
            integrate_transform(b.pos, b.rot, b.vel, b.ang, dt, b.no_rot, np, nr);
            b.pos = np;
            b.rot = nr;
            body_update_inertia(b);
        }
        b.force = v3zero();
        b.torque = v3zero();
    }

    // btRSBroadphase::calculateOverlappingPairs
    RSG_FN void find_pairs() {
        T->n_pairs = 0;
        const int nb = num_bodies();
        for (int p = 0; p < nb; p++) {
            int c = A.bp_cell[p];
            const int ns = soccar ? num_statics(*geo) : 0;
            const uint32_t skip = (soccar && p > 0) ? geo->car_skip : 0u;  // no statics, no geometry
            for (int s = 0; s < ns; s++) {
                if (!static_in_cell(s, c) || ((skip >> s) & 1u)) continue;
                V3 smin, smax;
                static_proxy_aabb(geo, s, smin, smax);
                if (aabb_overlap(T->aabb_min[p], T->aabb_max[p], smin, smax)) {
                    if (T->n_pairs >= MAX_PAIRS) { A.status |= RSG_STATUS_CONTACT_OVERFLOW; continue; }
                    T->pairs[T->n_pairs].a = (int16_t)(-1 - s);
                    T->pairs[T->n_pairs].b = (int16_t)p;
                    T->n_pairs++;
                }
            }
            if (nb > 1) {
                // the cell's dynamic list: every proxy registered within one
                // cell of `c`, in registration order
                int cnt = 0;
                for (int q = 0; q < nb; q++)
                    if (bp_cells_adjacent(A.bp_cell[q], c)) cnt++;
                if (cnt > 1) {
                    for (int r = 0; r < nb; r++) {
                        int q = T->order[r];
                        if (q == p || !bp_cells_adjacent(A.bp_cell[q], c)) continue;
                        if (!aabb_overlap(T->aabb_min[p], T->aabb_max[p], T->aabb_min[q], T->aabb_max[q])) continue;
                        int lo = p < q ? p : q, hi = p < q ? q : p;
                        bool found = false;
                        for (int k = 0; k < T->n_pairs; k++)
                            if (T->pairs[k].a == lo && T->pairs[k].b == hi) found = true;
                        if (found) continue;
                        if (T->n_pairs >= MAX_PAIRS) { A.status |= RSG_STATUS_CONTACT_OVERFLOW; continue; }
                        T->pairs[T->n_pairs].a = (int16_t)lo;
                        T->pairs[T->n_pairs].b = (int16_t)hi;
                        T->n_pairs++;
                    }
                }
            }
        }
    }

    RSG_FN bool needs_collision(int a, int b) { return active(a) || active(b); }

    // The dispatcher's manifolds (btCollisionDispatcher::dispatchAllCollisionPairs,
    // btCollisionDispatcher.cpp:221-271), decided before any narrowphase runs:
    // every algorithm's creation and release of its manifold depends only on
    // bounding boxes.  Ball + static and car + static pairs and the car + car
    // box-box child create their manifold with the pair, and the compound
    // ones release it unless the child AABB tests pass; the swapped compound
    // of ball + car creates it once the child AABB meets the ball's.  Each
    // kept manifold is (body b, body a) of its pair.  A manifold past the
    // capacity is dropped where the pair would have created it.
    RSG_FN void plan_pairs() {
        int n = 0;
        for (int k = 0; k < T->n_pairs; k++) {
            int a = T->pairs[k].a, b = T->pairs[k].b;
            if (!needs_collision(a, b)) continue;
            bool creates = true, keeps = true;
            if (a < 0) {
                if (b != 0) {
                    V3 cmin, cmax, smin, smax;
                    car_child_aabb(b, cmin, cmax);
                    static_shape_aabb(geo, -1 - a, smin, smax);
                    keeps = aabb_overlap(cmin, cmax, smin, smax);
                }
            } else if (a == 0) {
                V3 cmin, cmax;
                car_child_aabb(b, cmin, cmax);
                float r = make_ball_phys(A.ball.radius_uu, A.ball.mass).radius + 0.08f;
                V3 bmin = A.ball.b.pos - v3(r, r, r), bmax = A.ball.b.pos + v3(r, r, r);
                creates = keeps = aabb_overlap(cmin, cmax, bmin, bmax);
            } else {
                V3 amin, amax, bcmin, bcmax, bmin, bmax;
                car_child_aabb(a, amin, amax);
                Body& bb = A.cars[b - 1].b;
                shape_aabb(b, bb.pos, bb.rot, bcmin, bcmax);
                keeps = false;
                if (aabb_overlap(amin, amax, bcmin, bcmax)) {
                    car_child_aabb(b, bmin, bmax);
                    keeps = aabb_overlap(bmin, bmax, amin, amax);
                }
            }
            if (!creates) continue;
            if (n >= MAX_MANIFOLDS) { A.status |= RSG_STATUS_CONTACT_OVERFLOW; continue; }
            if (!keeps) continue;
            T->plan[n++] = k;
        }
        T->n_mf = n;
    }

    // The narrowphase of planned pair `mi` into manifold mi.
    RSG_FN void run_pair(int mi) {
        const Pair pr = T->pairs[T->plan[mi]];
        const int a = pr.a, b = pr.b;
        aborted = false;
        open_manifold(mi, b, a);
        if (a < 0) {
            const int s = -1 - a;
            const bool is_mesh = static_is_mesh(*geo, s);
            mesh = is_mesh ? &geo->meshes[s] : nullptr;
            if (b == 0) {
                if (is_mesh) ball_vs_mesh(mi);
                else ball_vs_plane(mi, s);
            } else {
                if (is_mesh) car_vs_mesh(mi, b);
                else car_vs_plane(mi, b, s);
            }
        } else if (a == 0) {
            car_vs_ball(mi, b);
        } else {
            car_vs_car(mi, b, a);
        }
    }

    // The narrowphase on one lane, pair by pair, effects applied as they come.
    RSG_FN void run_pairs() {
        rec = nullptr;
        epa = &ws->epa;
        for (int mi = 0; mi < T->n_mf; mi++) run_pair(mi);
    }

    // One lane's share of the narrowphase when the lanes of an arena run the
    // planned pairs side by side: effects recorded in `log`, no penetration
    // pass (`space` null) or the lane's own.
    RSG_FN void run_pairs_recorded(int first, int stride, EpaSpace* space, PairLog* log) {
        epa = space;
        for (int mi = first; mi < T->n_mf; mi += stride) {
            PairEffects& pe = log->eff[mi];
            pe.n = 0;
            pe.redo = 0;
            rec = &pe;
            run_pair(mi);
            if (aborted) pe.redo = 1;
        }
        rec = nullptr;
        epa = &ws->epa;
    }

    // The recorded effects in pair order; a pair marked for it runs again,
    // with its effects applied as they come.
    RSG_FN void apply_recorded(const PairLog* log) {
        rec = nullptr;
        epa = &ws->epa;
        for (int mi = 0; mi < T->n_mf; mi++) {
            const PairEffects& pe = log->eff[mi];
            if (pe.redo) {
                run_pair(mi);
                continue;
            }
            for (int j = 0; j < pe.n; j++) {
                const Effect& e = pe.e[j];
                apply_effect(e.code & 0xff, (e.code >> 8) & 0xff, e.code >> 16, e.v0, e.v1);
            }
        }
    }

    // -----------------------------------------------------------------------
    // Islands + solve

    // Union-find over the dynamic bodies joined by overlapping pairs with
    // contact response (btSimulationIslandManager::findUnions); the
    // reference's set sizes never decide a union, so none are kept.
    RSG_FN void island_unions(int* uf_id) {
        const int nb = num_bodies();
        for (int i = 0; i < nb; i++) uf_id[i] = i;
        for (int k = 0; k < T->n_pairs; k++) {
            int a = T->pairs[k].a, b = T->pairs[k].b;
            if (a < 0 || b < 0) continue;
            if (!has_response(a) || !has_response(b)) continue;
            int i = uf_find(uf_id, a), j = uf_find(uf_id, b);
            if (i == j) continue;
            uf_id[i] = j;
        }
    }

    // The islands and solver of a tick in which no manifold holds a point:
    // no rows, so the solver's write-back leaves each awake body's velocity
    // plus its external force and torque impulses (the full path's
    // expressions, with zero solver deltas) and its pose as it was.  Islands
    // only decide whether a sleeping ball wakes: it does when its island
    // holds an awake body.
    RSG_FN void solve_without_contacts(float dt) {
        const int nb = num_bodies();
        if (!A.ball.b.active) {
            int* uf_id = ws->isl.uf_id;
            island_unions(uf_id);
            const int ball_root = uf_find(uf_id, 0);
            for (int i = 1; i < nb; i++)
                if (A.cars[i - 1].b.active && uf_find(uf_id, i) == ball_root) { A.ball.b.active = true; break; }
        }
        for (int i = 0; i < nb; i++) {
            Body& b = *body_ptr(i);
            if (!b.active || b.inv_mass == 0.f) continue;
            const V3 ext_f = b.force * b.inv_mass * dt;
            const V3 ext_t = mul_vm(b.torque, b.inv_iw) * dt;
            b.vel = (b.vel + v3zero()) + ext_f;
            b.ang = (b.ang + v3zero()) + ext_t;
        }
    }

    RSG_FN void solve_islands(float dt) {
        bool any_point = false;
        for (int m = 0; m < T->n_mf; m++) any_point = any_point || mf[m].n > 0;
        if (!any_point) {
            solve_without_contacts(dt);
            return;
        }
        const int nb = num_bodies();
        IslandTables& I = ws->isl;
        int* uf_id = I.uf_id;
        island_unions(uf_id);
        int* tag = I.tag;
        for (int i = 0; i < nb; i++) tag[i] = uf_find(uf_id, i);
        // sortIslands: elements sorted by island id (quicksort)
        int* elem = I.elem;
        int* ekey = I.ekey;
        for (int i = 0; i < nb; i++) { elem[i] = i; ekey[i] = uf_find(uf_id, i); }
        if (nb > 1) quicksort_by_key(elem, ekey, 0, nb - 1);
        // activation: wake islands that contain an active (non sleeping) body
        // (btSimulationIslandManager::buildIslands)
        bool sleeping[MAX_BODIES];
        for (int i = 0; i < nb; i++) sleeping[i] = (i == 0) ? !A.ball.b.active : false;
        // cars: ACTIVE_TAG (awake) or DISABLE_SIMULATION (demoed, never sleeps)
        for (int s0 = 0; s0 < nb;) {
            int id = ekey[elem[s0]];
            int e0 = s0;
            while (e0 < nb && ekey[elem[e0]] == id) e0++;
            bool all_sleeping = true;
            for (int q = s0; q < e0; q++) {
                int bi = elem[q];
                bool active_tag = (bi == 0) ? A.ball.b.active : A.cars[bi - 1].b.active;
                if (active_tag) { all_sleeping = false; break; }
            }
            if (!all_sleeping) {
                for (int q = s0; q < e0; q++) {
                    int bi = elem[q];
                    if (bi == 0 && sleeping[0]) {
                        sleeping[0] = false;
                        A.ball.b.active = true;  // WANTS_DEACTIVATION: integrated, but got no gravity
                    }
                }
            }
            s0 = e0;
        }
        // island manifolds in dispatcher order
        int* im = I.im;
        int* ikey = I.ikey;
        int n_im = 0;
        for (int m = 0; m < T->n_mf; m++) {
            int b0 = mf[m].b0, b1 = mf[m].b1;
            bool b0_sleep = b0 < 0 ? true : (b0 == 0 ? sleeping[0] : false);
            bool b1_sleep = b1 < 0 ? true : (b1 == 0 ? sleeping[0] : false);
            if (b0_sleep && b1_sleep) continue;
            if (!(has_response(b0) && has_response(b1))) continue;
            if (b0 < 0 && b1 < 0) continue;
            im[n_im] = m;
            ikey[m] = b0 >= 0 ? tag[b0] : tag[b1];
            n_im++;
        }
        if (n_im > 1) quicksort_by_key(im, ikey, 0, n_im - 1);
        // collect bodies of non-sleeping islands, in island order
        int* bodies = I.bodies;
        int n_bodies = 0;
        for (int s0 = 0; s0 < nb;) {
            int id = ekey[elem[s0]];
            int e0 = s0;
            bool island_sleeping = true;
            while (e0 < nb && ekey[elem[e0]] == id) {
                int bi = elem[e0];
                if (body_ptr(bi)->active) island_sleeping = false;
                e0++;
            }
            if (!island_sleeping)
                for (int q = s0; q < e0; q++) bodies[n_bodies++] = elem[q];
            s0 = e0;
        }
        // manifolds whose island is processed (all islands are batched)
        Solver& S = ws->solver;
        S.rows = ws->rows;
        S.fric = ws->fric;
        solver_reset(S, dt);
        struct BodyAt {
            World* w;
            RSG_FN const Body* operator()(int k) { return w->body_ptr(k); }
        } body_at{this};
        for (int i = 0; i < n_bodies; i++) solver_body_id(S, bodies[i], body_at);
        for (int q = 0; q < n_im; q++) {
            int m = im[q];
            // manifolds of sleeping islands are not solved
            int b0 = mf[m].b0, b1 = mf[m].b1;
            int owner = b0 >= 0 ? b0 : b1;
            bool in = false;
            for (int i = 0; i < n_bodies; i++) if (bodies[i] == owner) in = true;
            if (!in) continue;
            convert_manifold(S, mf[m], body_at);
        }
        for (int i = 0; i < n_bodies; i++) {
            int k = bodies[i];
            if (S.special[k].num > 0) convert_special(S, k, body_at);
        }
        solver_iterate(S);
        // write back (btSequentialImpulseConstraintSolver::writeBackBodies)
        for (int i = 0; i < S.n_bodies; i++) {
            SBody& sb = S.bodies[i];
            if (sb.body < 0) continue;
            Body& b = *body_ptr(sb.body);
            sb.lin = sb.lin + sb.dv;
            sb.ang = sb.ang + sb.dw;
            if (sb.push.x != 0.f || sb.push.y != 0 || sb.push.z != 0 || sb.turn.x != 0.f || sb.turn.y != 0 || sb.turn.z != 0) {
                V3 np; M3 nr;
                integrate_transform(sb.xf.t, sb.xf.R, sb.push, sb.turn * SOLVER_SPLIT_TURN_ERP, dt, b.no_rot, np, nr);
                sb.xf.t = np;
                sb.xf.R = nr;
            }
            b.vel = sb.lin + sb.ext_f;
            b.ang = sb.ang + sb.ext_t;
            b.pos = sb.xf.t;
            b.rot = sb.xf.R;
        }
        if (S.overflow) A.status |= RSG_STATUS_SOLVER_OVERFLOW;
    }

    RSG_FN static int uf_find(int* id, int x) {
        while (x != id[x]) {
            id[x] = id[id[x]];
            x = id[x];
        }
        return x;
    }

    template <class BodyAt>
    RSG_FN void convert_manifold(Solver& S, const Manifold& m, BodyAt& body_at) {
        int ia = solver_body_id(S, m.b0, body_at);
        int ib = solver_body_id(S, m.b1, body_at);
        SBody& A_ = S.bodies[ia];
        SBody& B_ = S.bodies[ib];
        bool a_zero = fuzzy_zero(A_.inv_mass), b_zero = fuzzy_zero(B_.inv_mass);
        if (a_zero && b_zero) return;
        const Body* rb0 = m.b0 >= 0 ? body_ptr(m.b0) : nullptr;
        const Body* rb1 = m.b1 >= 0 ? body_ptr(m.b1) : nullptr;
        Xf x0 = xf_of(m.b0), x1 = xf_of(m.b1);
        for (int j = 0; j < m.n; j++) {
            const CPoint& cp = m.p[j];
            if (S.n_rows >= MAX_ROWS) { S.overflow = true; return; }
            int ci = S.n_rows++;
            SRow& c = S.rows[ci];
            c.a = ia; c.b = ib;
            V3 rel1 = cp.pos_a - x0.t;
            V3 rel2 = cp.pos_b - x1.t;
            setup_contact_row(S, c, ia, ib, rb0, rb1, cp.normal, cp.dist, cp.friction, cp.restitution, rel1, rel2);
            c.special = cp.special;
            c.friction_index = S.n_fric;
            if (cp.special) {
                for (int i = 0; i < 2; i++) {
                    int code = i ? m.b1 : m.b0;
                    if (code >= 0) {
                        SpecialInfo& sp = S.special[code];
                        sp.num++;
                        sp.friction = cp.friction;
                        sp.restitution = cp.restitution;
                        sp.total_normal += cp.normal;
                        sp.total_dist += len(i ? rel2 : rel1);
                    }
                }
            }
            add_friction_for(S, ia, ib, rb0, rb1, cp.normal, cp.friction, rel1, rel2, ci);
        }
    }

    template <class BodyAt>
    RSG_FN void convert_special(Solver& S, int k, BodyAt& body_at) {
        SpecialInfo& sp = S.special[k];
        float distance = sp.total_dist / (float)sp.num;
        V3 normal = sp.total_normal * (1.f / (float)sp.num);
        int ia = solver_body_id(S, k, body_at);
        int ib = solver_body_id(S, -1000, body_at);  // fixed body
        if (S.n_rows >= MAX_ROWS) { S.overflow = true; return; }
        int ci = S.n_rows++;
        SRow& c = S.rows[ci];
        c.a = ia; c.b = ib;
        V3 rel1 = normal * -distance;
        V3 rel2 = v3zero();
        const Body* rb0 = body_ptr(k);
        setup_contact_row(S, c, ia, ib, rb0, nullptr, normal, distance, sp.friction, sp.restitution, rel1, rel2);
        c.special = false;
        c.friction_index = S.n_fric;
        add_friction_for(S, ia, ib, rb0, nullptr, normal, sp.friction, rel1, rel2, ci);
        sp.num = 0;
    }
};

}  // namespace rsg
