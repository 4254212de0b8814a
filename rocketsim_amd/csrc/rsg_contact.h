// rsg_contact.h -- contact generation against the static arena (the shared
// triangle meshes + 4 planes), contact records, the reference's contact-added
// rules and the internal-edge normal fix-up.
//
// Reference behaviour restated here:
//   static planes and their AABBs     src/Sim/Arena/Arena.cpp:1036-1068,
//                                     BulletCollision/CollisionShapes/btStaticPlaneShape.cpp:40-82
//   convex vs plane                   CollisionDispatch/btConvexPlaneCollisionAlgorithm.cpp:92-161
//   sphere vs triangle                CollisionDispatch/SphereTriangleDetector.cpp:88-243
//   convex vs triangle mesh           CollisionDispatch/btConvexConcaveCollisionAlgorithm.cpp:71-250
//   contact point creation / refresh  CollisionDispatch/btManifoldResult.cpp:56-207,
//                                     NarrowPhaseCollision/btPersistentManifold.cpp:110-304
//   contact callback                  src/Sim/Arena/Arena.cpp:244-314, 407-414
//   internal-edge fix-up              CollisionDispatch/btInternalEdgeUtility.cpp:370-798
//   ray vs triangle                   NarrowPhaseCollision/btRaycastCallback.cpp:35-118
#pragma once

#include "rsg_arena.h"
#include "rsg_mesh.h"

namespace rsg {

// Static collision objects in the reference's creation order
// (Arena::_SetupArenaCollisionShapes, src/Sim/Arena/Arena.cpp:990-1068): the
// arena's collision meshes in load order, then its planes (the floor, the
// ceiling, the two side walls and, in Hoops, the two end walls).  In an arena
// of M meshes static object s is mesh s for s < M and plane s - M otherwise;
// its body code is -1 - s.
enum PlaneId { PL_FLOOR = 0, PL_CEIL = 1, PL_WALL_NX = 2, PL_WALL_PX = 3, PL_WALL_NY = 4, PL_WALL_PY = 5,
               MAX_PLANES = 6 };
constexpr int MAX_MESHES = 26;
constexpr int MAX_STATICS = MAX_MESHES + MAX_PLANES;  // one bit each in a cell's mask

struct PlaneDef {
    V3 n;        // plane normal
    V3 origin;   // rigid body origin (plane constant is 0)
    int axis;    // single axis of the normal
    bool backwards;
};

// An arena's static geometry as the kernels see it, shared by every arena
// of the process (the reference keeps its meshes in process-wide storage,
// src/RocketSim.cpp:50-53).
struct ArenaGeom {
    const MeshView* meshes;        // n_meshes views, in load order
    int n_meshes;
    int n_planes;
    const uint32_t* cell_statics;  // BP_NUM_CELLS masks: bit s = static object s registered in the cell
    // Statics that cars and their wheel rays pass through: the Hoops net
    // meshes, whose collision group only the ball's filter mask holds
    // (Arena.cpp:1005-1024, Ball.cpp:111-114, Car.cpp:227-233).
    uint32_t car_skip;
    PlaneDef planes[MAX_PLANES];
};
RSG_FN int num_statics(const ArenaGeom& g) { return g.n_meshes + g.n_planes; }
RSG_FN bool static_is_mesh(const ArenaGeom& g, int s) { return s < g.n_meshes; }
RSG_FN const PlaneDef& static_plane(const ArenaGeom& g, int s) { return g.planes[s - g.n_meshes]; }

// The planes of an arena kind (0: Soccar, 1: Hoops) in creation order
// (Arena.cpp:1033-1068); returns their count.  `unit` is the stored length
// of the axis-aligned normals: the reference keeps planeNormal.normalized()
// (btStaticPlaneShape.cpp:21), whose SSE reciprocal-square-root normalisation
// of a unit axis is the host CPU's rsqrtss(1) after one Newton step
// (0.99999994 on common x86 cores), measured once on the host.
RSG_FN int arena_planes(int kind, float unit, PlaneDef out[MAX_PLANES]) {
    const bool hoops = kind == 1;
    const float ex = hoops ? ARENA_EXTENT_X_HOOPS : ARENA_EXTENT_X;
    const float ey = ARENA_EXTENT_Y_HOOPS;
    const float h = hoops ? ARENA_HEIGHT_HOOPS : ARENA_HEIGHT;
    out[PL_FLOOR] = PlaneDef{v3(0, 0, unit), v3(0, 0, 0) * UU_TO_BT, 2, false};
    out[PL_CEIL] = PlaneDef{v3(0, 0, -unit), v3(0, 0, h) * UU_TO_BT, 2, true};
    out[PL_WALL_NX] = PlaneDef{v3(unit, 0, 0), v3(-ex, 0, h / 2) * UU_TO_BT, 0, false};
    out[PL_WALL_PX] = PlaneDef{v3(-unit, 0, 0), v3(ex, 0, h / 2) * UU_TO_BT, 0, true};
    if (!hoops) return 4;
    out[PL_WALL_NY] = PlaneDef{v3(0, unit, 0), v3(0, -ey, h / 2) * UU_TO_BT, 1, false};
    out[PL_WALL_PY] = PlaneDef{v3(0, -unit, 0), v3(0, ey, h / 2) * UU_TO_BT, 1, true};
    return 6;
}

// btStaticPlaneShape::getAabb with RocketSim's single-axis slab (+-0.2 m,
// unbounded behind the plane).
RSG_FN void plane_shape_aabb(const PlaneDef& p, V3& mn, V3& mx) {
    const float L = 1e18f;
    mn = v3(-L, -L, -L);
    mx = v3(L, L, L);
    float o = get(p.origin, p.axis);
    set(mn, p.axis, o + (0.f - 0.2f));
    set(mx, p.axis, o + (0.f + 0.2f));
    if (p.backwards) set(mx, p.axis, L); else set(mn, p.axis, -L);
}

// btTriangleMeshShape::getAabb at the identity transform (margin 0).
RSG_FN void mesh_shape_aabb(const MeshView& m, V3& mn, V3& mx) {
    V3 he = 0.5f * (m.aabb_max - m.aabb_min);
    V3 c = 0.5f * (m.aabb_max + m.aabb_min);
    mn = c - he;
    mx = c + he;
}

RSG_FN void static_shape_aabb(const ArenaGeom* g, int s, V3& mn, V3& mx) {
    if (static_is_mesh(*g, s)) mesh_shape_aabb(g->meshes[s], mn, mx);
    else plane_shape_aabb(static_plane(*g, s), mn, mx);
}

// Broadphase proxy AABB of a static object: shape AABB +- the global contact
// breaking threshold (btCollisionWorld::updateSingleAabb).
RSG_FN void static_proxy_aabb(const ArenaGeom* g, int s, V3& mn, V3& mx) {
    static_shape_aabb(g, s, mn, mx);
    V3 t = v3(CONTACT_BREAKING_THRESHOLD, CONTACT_BREAKING_THRESHOLD, CONTACT_BREAKING_THRESHOLD);
    mn = mn - t;
    mx = mx + t;
}

// Broadphase grid (btRSBroadphase with the reference ArenaConfig: min
// (-5600,-6000,0) uu, max (5600,6000,2200) uu, cell 370 uu;
// src/Sim/Arena/ArenaConfig/ArenaConfig.h:20-41, Arena.cpp:438-447).
RSG_FN V3 bp_min_pos() { return v3(-5600, -6000, 0) * UU_TO_BT; }
RSG_FN V3 bp_max_pos() { return v3(5600, 6000, 2200) * UU_TO_BT; }
RSG_FN float bp_cell_size() { return 370 * UU_TO_BT; }

// btRSBroadphase::GetCellIndices (btRSBroadphase.h:90-98).  Coordinates far
// outside the grid (the planes' unbounded AABBs) clamp to the border cells.
RSG_FN void bp_cell_indices(V3 p, int& i, int& j, int& k) {
    V3 f = (p - bp_min_pos()) * (1.f / bp_cell_size());
    f.x = fminf(fmaxf(f.x, -1e9f), 1e9f);
    f.y = fminf(fmaxf(f.y, -1e9f), 1e9f);
    f.z = fminf(fmaxf(f.z, -1e9f), 1e9f);
    i = (int)f.x; j = (int)f.y; k = (int)f.z;
    i = i < 0 ? 0 : (i > BP_CELLS_X - 1 ? BP_CELLS_X - 1 : i);
    j = j < 0 ? 0 : (j > BP_CELLS_Y - 1 ? BP_CELLS_Y - 1 : j);
    k = k < 0 ? 0 : (k > BP_CELLS_Z - 1 ? BP_CELLS_Z - 1 : k);
}
RSG_FN int bp_cell_index(int i, int j, int k) { return i * BP_CELLS_Y * BP_CELLS_Z + j * BP_CELLS_Z + k; }
RSG_FN int bp_cell_of(V3 p) {
    int i, j, k;
    bp_cell_indices(p, i, j, k);
    return bp_cell_index(i, j, k);
}
// Two cells within one cell of each other on every axis: a dynamic proxy is
// registered in the 27 cells around its own (_UpdateCellsDynamic).
RSG_FN bool bp_cells_adjacent(int a, int b) {
    int ai = a / (BP_CELLS_Y * BP_CELLS_Z), aj = (a / BP_CELLS_Z) % BP_CELLS_Y, ak = a % BP_CELLS_Z;
    int bi = b / (BP_CELLS_Y * BP_CELLS_Z), bj = (b / BP_CELLS_Z) % BP_CELLS_Y, bk = b % BP_CELLS_Z;
    int di = ai - bi, dj = aj - bj, dk = ak - bk;
    return di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1;
}

RSG_FN bool aabb_overlap(V3 amin, V3 amax, V3 bmin, V3 bmax) {
    bool o = true;
    o = (amin.x > bmax.x || amax.x < bmin.x) ? false : o;
    o = (amin.z > bmax.z || amax.z < bmin.z) ? false : o;
    o = (amin.y > bmax.y || amax.y < bmin.y) ? false : o;
    return o;
}

// TestTriangleAgainstAabb2 (LinearMath/btAabbUtil2.h:49-63): x, z, then y.
RSG_FN bool triangle_aabb_overlap(const V3 v[3], V3 amin, V3 amax) {
    if (fminf(fminf(v[0].x, v[1].x), v[2].x) > amax.x) return false;
    if (fmaxf(fmaxf(v[0].x, v[1].x), v[2].x) < amin.x) return false;
    if (fminf(fminf(v[0].z, v[1].z), v[2].z) > amax.z) return false;
    if (fmaxf(fmaxf(v[0].z, v[1].z), v[2].z) < amin.z) return false;
    if (fminf(fminf(v[0].y, v[1].y), v[2].y) > amax.y) return false;
    if (fmaxf(fmaxf(v[0].y, v[1].y), v[2].y) < amin.y) return false;
    return true;
}

RSG_FN M3 abs_m3(const M3& m) {
    M3 a;
    for (int r = 0; r < 3; r++) a.r[r] = v3(fabsf(m.r[r].x), fabsf(m.r[r].y), fabsf(m.r[r].z));
    return a;
}

// Transform helpers (btTransform: x -> R x + t).
struct Xf {
    M3 R;
    V3 t;
};
RSG_FN V3 xf_apply(const Xf& x, V3 p) { return x.R * p + x.t; }
// invXform: R^T (p - t)
RSG_FN V3 xf_inv_apply(const Xf& x, V3 p) { V3 v = p - x.t; return mul_vm(v, x.R); }
RSG_FN Xf xf_inverse(const Xf& x) { Xf r; r.R = transpose(x.R); r.t = r.R * (-x.t); return r; }
RSG_FN Xf xf_mul(const Xf& a, const Xf& b) { Xf r; r.R = a.R * b.R; r.t = xf_apply(a, b.t); return r; }
RSG_FN Xf xf_ident() { Xf r; r.R = m3_identity(); r.t = v3zero(); return r; }
RSG_FN Xf xf_from(V3 t, const M3& R) { Xf r; r.R = R; r.t = t; return r; }

// ---------------------------------------------------------------------------
// Contact records

// Manifolds per arena and tick.  A host test build may lower the capacity
// (RSG_TEST_MAX_MANIFOLDS) to reach the overflow path with ordinary play.
#ifdef RSG_TEST_MAX_MANIFOLDS
constexpr int MAX_MANIFOLDS = RSG_TEST_MAX_MANIFOLDS;
#else
constexpr int MAX_MANIFOLDS = RSG_MAX_MANIFOLDS;
#endif
constexpr int MAX_POINTS_PER_MANIFOLD = 4;

struct CPoint {
    V3 local_a, local_b;
    V3 pos_a, pos_b;
    V3 normal;       // normal on B, world
    float dist;
    float friction, restitution;
    bool special;
    int tri;         // mesh triangle (or -1)
};

struct Manifold {
    int b0, b1;      // body codes: >= 0 dynamic (0 ball, 1+i car i), < 0 static (-1 - StaticId)
    float threshold; // contact breaking threshold
    int n;
    CPoint p[MAX_POINTS_PER_MANIFOLD];
};

// Index of the point to replace when the 4-point cache is full
// (btPersistentManifold::sortCachedPoints with gContactCalcArea3Points = true,
// btPersistentManifold.cpp:30-32, 110-191): keep the deepest point, replace
// the one whose removal leaves the largest (3-point approximated) area.
RSG_FN int sort_cached_points(const Manifold& m, const CPoint& pt) {
    int max_pen = -1;
    float max_pen_d = pt.dist;
    for (int i = 0; i < 4; i++)
        if (m.p[i].dist < max_pen_d) { max_pen = i; max_pen_d = m.p[i].dist; }
    const V3 n = pt.local_a;
    const V3 p0 = m.p[0].local_a, p1 = m.p[1].local_a, p2 = m.p[2].local_a, p3 = m.p[3].local_a;
    float res[4] = {0, 0, 0, 0};
    if (max_pen != 0) res[0] = len2(cross(n - p1, p3 - p2));
    if (max_pen != 1) res[1] = len2(cross(n - p0, p3 - p2));
    if (max_pen != 2) res[2] = len2(cross(n - p0, p3 - p1));
    if (max_pen != 3) res[3] = len2(cross(n - p0, p2 - p1));
    // btVector4::closestAxis4: index of the largest absolute value, first wins
    int best = 0;
    float bv = fabsf(res[0]);
    for (int i = 1; i < 4; i++)
        if (fabsf(res[i]) > bv) { bv = fabsf(res[i]); best = i; }
    return best;
}

RSG_FN int manifold_add_point(Manifold& m, const CPoint& pt) {
    int idx = m.n;
    if (idx == MAX_POINTS_PER_MANIFOLD) {
        idx = sort_cached_points(m, pt);
    } else {
        m.n++;
    }
    if (idx < 0) idx = 0;
    m.p[idx] = pt;
    return idx;
}

RSG_FN void manifold_remove_point(Manifold& m, int i) {
    int last = m.n - 1;
    if (i != last) m.p[i] = m.p[last];
    m.n--;
}

// btPersistentManifold::refreshContactPoints (btPersistentManifold.cpp:247-304)
RSG_FN void manifold_refresh(Manifold& m, const Xf& ta, const Xf& tb) {
    for (int i = m.n - 1; i >= 0; i--) {
        CPoint& p = m.p[i];
        p.pos_a = xf_apply(ta, p.local_a);
        p.pos_b = xf_apply(tb, p.local_b);
        p.dist = dot(p.pos_a - p.pos_b, p.normal);
    }
    for (int i = m.n - 1; i >= 0; i--) {
        CPoint& p = m.p[i];
        if (!(p.dist <= m.threshold)) {
            manifold_remove_point(m, i);
        } else {
            V3 proj = p.pos_a - p.normal * p.dist;
            V3 diff = p.pos_b - proj;
            float d2 = dot(diff, diff);
            if (d2 > m.threshold * m.threshold) manifold_remove_point(m, i);
        }
    }
}

// ---------------------------------------------------------------------------
// Internal-edge fix-up (btAdjustInternalEdgeContacts with the mesh at the
// identity transform and normalAdjustFlags = 0).

constexpr uint32_t TRI_V0V1_CONVEX = 1, TRI_V1V2_CONVEX = 2, TRI_V2V0_CONVEX = 4;
constexpr uint32_t TRI_V0V1_SWAP = 8, TRI_V1V2_SWAP = 16, TRI_V2V0_SWAP = 32;
constexpr uint32_t TRI_HAS_INFO = 1u << 31;
constexpr float TWO_PI_F = 6.283185307179586232f;

RSG_FN void nearest_on_segment(V3 p, V3 a, V3 b, V3& out) {
    V3 d = b - a;
    if (fuzzy_zero(d)) { out = a; return; }
    float t = dot(p - a, d) / dot(d, d);
    if (t < 0) t = 0; else if (t > 1) t = 1;
    out = a + d * t;
}

// quatRotate(q, v) with the reference's SSE association
// (LinearMath/btQuaternion.h:707-733, 912-923).
RSG_FN V3 quat_rotate(Q4 q, V3 w) {
    Q4 a = q4((q.w * w.x + q.y * w.z) - q.z * w.y,
              (q.w * w.y + q.z * w.x) - q.x * w.z,
              (q.w * w.z + q.x * w.y) - q.y * w.x,
              -(q.x * w.x + q.y * w.y) - q.z * w.z);
    Q4 inv = q4(-q.x, -q.y, -q.z, q.w);
    Q4 r = qmul(a, inv);
    return v3(r.x, r.y, r.z);
}

RSG_FN float edge_angle_of(V3 a, V3 b, V3 c) { return atan2f(dot(c, a), dot(c, b)); }

// btClampNormal (btInternalEdgeUtility.cpp:385-412)
RSG_FN bool clamp_normal(V3 edge, V3 tri_normal, V3 local_n, float corrected, V3& out) {
    V3 edge_cross = normalized(cross(edge, tri_normal));
    float cur = edge_angle_of(edge_cross, tri_normal, local_n);
    if (corrected < 0) {
        if (cur < corrected) {
            float diff = corrected - cur;
            out = quat_to_m3(q_axis_angle(edge, diff)) * local_n;
            return true;
        }
    }
    if (corrected >= 0) {
        if (cur > corrected) {
            float diff = corrected - cur;
            out = quat_to_m3(q_axis_angle(edge, diff)) * local_n;
            return true;
        }
    }
    return false;
}

RSG_FN void adjust_internal_edge(CPoint& cp, const MeshView& mesh, int tri) {
    if (!mesh.tri_info) return;
    const TriInfo& info = mesh.tri_info[tri];
    if (!(info.flags & TRI_HAS_INFO)) return;
    V3 v[3];
    mesh_triangle(mesh, tri, v);
    V3 tri_normal = normalized(cross(v[1] - v[0], v[2] - v[0]));
    V3 nearest;
    nearest_on_segment(cp.local_b, v[0], v[1], nearest);
    V3 contact = cp.local_b;
    bool near_edge = false;
    int concave_hits = 0;
    V3 local_n = normalized(cp.normal);
    int best = -1;
    float best_d = 1e18f;
    const float max_angle = TWO_PI_F;
    const int ea[3] = {0, 1, 2}, eb[3] = {1, 2, 0};
    for (int e = 0; e < 3; e++) {
        if (fabsf(info.edge_angle[e]) < max_angle) {
            V3 nn;
            nearest_on_segment(cp.local_b, v[ea[e]], v[eb[e]], nn);
            float l = len(contact - nn);
            if (l < best_d) { best = e; best_d = l; }
        }
    }
    const uint32_t convex_bit[3] = {TRI_V0V1_CONVEX, TRI_V1V2_CONVEX, TRI_V2V0_CONVEX};
    const uint32_t swap_bit[3] = {TRI_V0V1_SWAP, TRI_V1V2_SWAP, TRI_V2V0_SWAP};
    for (int e = 0; e < 3; e++) {
        if (e > 0) nearest_on_segment(contact, v[ea[e]], v[eb[e]], nearest);
        if (!(fabsf(info.edge_angle[e]) < max_angle)) continue;
        float l = len(contact - nearest);
        if (!(l < 0.1f)) continue;  // m_edgeDistanceThreshold
        if (best != e) continue;
        V3 edge = v[ea[e]] - v[eb[e]];
        near_edge = true;
        float ang = info.edge_angle[e];
        if (ang == 0.f) {
            concave_hits++;
        } else {
            bool convex = (info.flags & convex_bit[e]) != 0;
            float swap = convex ? 1.f : -1.f;
            V3 nA = swap * tri_normal;
            V3 comp_b = quat_rotate(q_axis_angle(edge, ang), tri_normal);
            if (info.flags & swap_bit[e]) comp_b = comp_b * -1.f;
            V3 nB = swap * comp_b;
            float nda = dot(local_n, nA), ndb = dot(local_n, nB);
            bool back = (nda < 0.f) && (ndb < 0.f);  // m_convexEpsilon = 0
            if (back) {
                concave_hits++;
            } else {
                V3 ln = (e == 0) ? local_n : cp.normal;
                V3 clamped;
                if (clamp_normal(edge, swap * tri_normal, ln, ang, clamped)) {
                    if (dot(clamped, 1.f * tri_normal) > 0) {
                        cp.normal = clamped;
                        cp.pos_b = cp.pos_a - cp.normal * cp.dist;
                        cp.local_b = cp.pos_b;
                    }
                }
            }
        }
    }
    if (near_edge && concave_hits > 0) {
        V3 nn = tri_normal * 1.f;
        float d = dot(nn, local_n);
        if (d < 0) return;
        cp.normal = nn;
        cp.pos_b = cp.pos_a - cp.normal * cp.dist;
        cp.local_b = cp.pos_b;
    }
}

// ---------------------------------------------------------------------------
// Ray casts (wheel suspension rays) against the static world

// btTriangleRaycastCallback::processTriangle (btRaycastCallback.cpp:35-118),
// flags 0.  Returns true and updates hit_fraction / normal when closer.
RSG_FN bool ray_triangle(V3 from, V3 to, V3 v0, V3 v1, V3 v2, float& hit_fraction, V3& hit_normal) {
    V3 v10 = v1 - v0, v20 = v2 - v0;
    V3 tn = cross(v10, v20);
    float dist = dot(v0, tn);
    float da = dot(tn, from) - dist;
    float db = dot(tn, to) - dist;
    if (da * db >= 0.f) return false;
    float proj = da - db;
    float distance = da / proj;
    if (!(distance < hit_fraction)) return false;
    float tol = len2(tn) * -0.0001f;
    float s = 1.f - distance;
    V3 point = from * s + to * distance;
    V3 v0p = v0 - point, v1p = v1 - point;
    V3 cp0 = cross(v0p, v1p);
    if (!(dot(cp0, tn) >= tol)) return false;
    V3 v2p = v2 - point;
    V3 cp1 = cross(v1p, v2p);
    if (!(dot(cp1, tn) >= tol)) return false;
    V3 cp2 = cross(v2p, v0p);
    if (!(dot(cp2, tn) >= tol)) return false;
    tn = normalized(tn);
    hit_normal = (da <= 0.f) ? -tn : tn;
    hit_fraction = distance;
    return true;
}

// btPlaneSpace1 (LinearMath/btVector3.h:1266-1294)
RSG_FN void plane_space(V3 n, V3& p, V3& q) {
    if (fabsf(n.z) > 0.7071067811865475244f) {
        float a = n.y * n.y + n.z * n.z;
        float k = 1.f / sqrtf(a);
        p = v3(0, -n.z * k, n.y * k);
        q = v3(a * k, -n.x * p.z, n.x * p.y);
    } else {
        float a = n.x * n.x + n.y * n.y;
        float k = 1.f / sqrtf(a);
        p = v3(-n.y * k, n.x * k, 0);
        q = v3(-n.z * p.y, n.z * p.x, a * k);
    }
}

// Ray vs static plane: the plane shape emits two triangles around the ray's
// AABB (btStaticPlaneShape::processAllTriangles) which are ray-tested.
RSG_FN bool ray_plane(const PlaneDef& pd, V3 from_w, V3 to_w, float& hit_fraction, V3& hit_normal) {
    // world -> plane local (identity basis)
    V3 from = from_w - pd.origin, to = to_w - pd.origin;
    V3 amin = v3(fminf(from.x, to.x), fminf(from.y, to.y), fminf(from.z, to.z));
    V3 amax = v3(fmaxf(from.x, to.x), fmaxf(from.y, to.y), fmaxf(from.z, to.z));
    V3 he = (amax - amin) * 0.5f;
    float radius = len(he);
    V3 center = (amax + amin) * 0.5f;
    V3 t0, t1;
    plane_space(pd.n, t0, t1);
    V3 pc = center - (dot(pd.n, center) - 0.f) * pd.n;
    V3 a = pc + t0 * radius + t1 * radius;
    V3 b = pc + t0 * radius - t1 * radius;
    V3 c = pc - t0 * radius - t1 * radius;
    bool hit = false;
    V3 n;
    if (ray_triangle(from, to, a, b, c, hit_fraction, n)) { hit = true; hit_normal = n; }
    V3 a2 = pc - t0 * radius - t1 * radius;
    V3 b2 = pc - t0 * radius + t1 * radius;
    V3 c2 = pc + t0 * radius + t1 * radius;
    if (ray_triangle(from, to, a2, b2, c2, hit_fraction, n)) { hit = true; hit_normal = n; }
    return hit;
}

}  // namespace rsg
