// rsg_gjk.h -- convex-convex narrowphase of the car hitbox against the ball
// and against mesh triangles: the reference's GJK pair detector with its
// Voronoi simplex solver, and its GJK/EPA penetration solver for overlapping
// cores.  The steps (termination tests, simplex reduction order, barycentric
// bookkeeping, EPA face list order) follow the reference so that the contact
// point and normal it reports -- which are not unique when faces are parallel
// -- come out the same:
//
//   btGjkPairDetector::getClosestPointsNonVirtual
//       BulletCollision/NarrowPhaseCollision/btGjkPairDetector.cpp:686-959
//   btVoronoiSimplexSolver (closest point on point/segment/triangle/
//   tetrahedron, Ericson "Real-Time Collision Detection" 5.1)
//       BulletCollision/NarrowPhaseCollision/btVoronoiSimplexSolver.cpp:29-577
//   btGjkEpaPenetrationDepthSolver::calcPenDepth
//       BulletCollision/NarrowPhaseCollision/btGjkEpaPenetrationDepthSolver.cpp:22-79
//   btGjkEpaSolver2 (Presson's GJK + expanding polytope)
//       BulletCollision/NarrowPhaseCollision/btGjkEpa2.cpp:35-1000
//   support mappings: btConvexShape::localGetSupportVertex(WithoutMargin)NonVirtual
//       BulletCollision/CollisionShapes/btConvexShape.cpp:126-250
//
// Everything lives in registers / private memory of one lane; the EPA pools
// use 16-bit indices instead of pointers.
#pragma once

#include "rsg_contact.h"

namespace rsg {

enum ShapeKind { SHAPE_SPHERE = 0, SHAPE_BOX = 1, SHAPE_TRIANGLE = 2 };

struct CShape {
    int kind;
    V3 half;      // box: half extents without margin
    float margin; // box margin, sphere radius, triangle 0
    V3 v[3];      // triangle vertices in the shape frame
};

RSG_FN CShape shape_sphere(float r) { CShape s; s.kind = SHAPE_SPHERE; s.half = v3zero(); s.margin = r; return s; }
RSG_FN CShape shape_box(V3 half_no_margin, float margin) { CShape s; s.kind = SHAPE_BOX; s.half = half_no_margin; s.margin = margin; return s; }
RSG_FN CShape shape_triangle(const V3 v[3]) {
    CShape s; s.kind = SHAPE_TRIANGLE; s.half = v3zero(); s.margin = 0.f;
    s.v[0] = v[0]; s.v[1] = v[1]; s.v[2] = v[2];
    return s;
}

RSG_FN float fsel(float a, float b, float c) { return a >= 0.f ? b : c; }
RSG_FN int max_axis3(V3 v) { return v.x < v.y ? (v.y < v.z ? 2 : 1) : (v.x < v.z ? 2 : 0); }

// localGetSupportVertexWithoutMarginNonVirtual
RSG_FN V3 support_core(const CShape& s, V3 d) {
    if (s.kind == SHAPE_BOX) return v3(fsel(d.x, s.half.x, -s.half.x), fsel(d.y, s.half.y, -s.half.y), fsel(d.z, s.half.z, -s.half.z));
    if (s.kind == SHAPE_TRIANGLE) {
        V3 dots = v3(dot(d, s.v[0]), dot(d, s.v[1]), dot(d, s.v[2]));
        return s.v[max_axis3(dots)];
    }
    return v3zero();
}

// localGetSupportVertexNonVirtual: core support + margin along the normalised direction
RSG_FN V3 support_full(const CShape& s, V3 d) {
    V3 n = d;
    if (len2(n) < FLT_EPS * FLT_EPS) n = v3(-1.f, -1.f, -1.f);
    n = normalized(n);
    return support_core(s, n) + s.margin * n;
}

// Centre of the shape's AABB in a frame (getAabb), used for the reference's
// contact normal direction fix.
RSG_FN V3 shape_aabb_center(const CShape& s, const Xf& t) {
    V3 mn, mx;
    if (s.kind == SHAPE_SPHERE) {
        float m = s.margin + 0.08f;  // btSphereShape::getAabb (RocketSim +0.08)
        mn = t.t - v3(m, m, m);
        mx = t.t + v3(m, m, m);
    } else if (s.kind == SHAPE_BOX) {
        V3 he = s.half + v3(s.margin, s.margin, s.margin);
        M3 ab = abs_m3(t.R);
        V3 e = v3(dot(ab.r[0], he), dot(ab.r[1], he), dot(ab.r[2], he));
        mn = t.t - e;
        mx = t.t + e;
    } else {
        // btConvexInternalShape::getAabbSlow via support vertices along +-axes
        for (int i = 0; i < 3; i++) {
            V3 vec = v3zero();
            set(vec, i, 1.f);
            V3 sv = support_core(s, mul_vm(vec, t.R));
            set(mx, i, get(xf_apply(t, sv), i) + s.margin);
            set(vec, i, -1.f);
            sv = support_core(s, mul_vm(vec, t.R));
            set(mn, i, get(xf_apply(t, sv), i) - s.margin);
        }
    }
    return (mx + mn) * 0.5f;
}

// ---------------------------------------------------------------------------
// Voronoi simplex solver

struct VoronoiSimplex {
    V3 W[4], P[4], Q[4];
    int n;
    V3 last_w;
    bool needs_update, cached_valid;
    V3 cp1, cp2, cv;
    float bc[4];
    uint32_t used;   // bit i: vertex i used by the closest feature
    bool degenerate;

    RSG_FN void reset() {
        cached_valid = false;
        n = 0;
        needs_update = true;
        last_w = v3(1e18f, 1e18f, 1e18f);
        bc_reset();
    }
    RSG_FN void bc_reset() { degenerate = false; bc[0] = bc[1] = bc[2] = bc[3] = 0.f; used = 0; }
    RSG_FN bool bc_valid() const { return bc[0] >= 0.f && bc[1] >= 0.f && bc[2] >= 0.f && bc[3] >= 0.f; }
    RSG_FN void remove_vertex(int i) {
        n--;
        W[i] = W[n]; P[i] = P[n]; Q[i] = Q[n];
    }
    RSG_FN void reduce_vertices(uint32_t u) {
        if (n >= 4 && !(u & 8)) remove_vertex(3);
        if (n >= 3 && !(u & 4)) remove_vertex(2);
        if (n >= 2 && !(u & 2)) remove_vertex(1);
        if (n >= 1 && !(u & 1)) remove_vertex(0);
    }
    RSG_FN void add_vertex(V3 w, V3 p, V3 q) {
        last_w = w;
        needs_update = true;
        W[n] = w; P[n] = p; Q[n] = q;
        n++;
    }
    RSG_FN bool in_simplex(V3 w) const {
        bool found = false;
        for (int i = 0; i < n; i++) {
            V3 d = w - W[i];
            if (len2(d) <= 0.0001f) { found = true; break; }
        }
        if (w.x == last_w.x && w.y == last_w.y && w.z == last_w.z) return true;
        return found;
    }
    RSG_FN bool full() const { return n == 4; }

    // closestPtPointTriangle (origin against triangle abc)
    RSG_FN static void closest_tri(V3 p, V3 a, V3 b, V3 c, V3& q, float w[3], uint32_t& u) {
        u = 0;
        V3 ab = b - a, ac = c - a, ap = p - a;
        float d1 = dot(ab, ap), d2 = dot(ac, ap);
        if (d1 <= 0.f && d2 <= 0.f) { q = a; u = 1; w[0] = 1; w[1] = 0; w[2] = 0; return; }
        V3 bp = p - b;
        float d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= 0.f && d4 <= d3) { q = b; u = 2; w[0] = 0; w[1] = 1; w[2] = 0; return; }
        float vc = d1 * d4 - d3 * d2;
        if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
            float v = d1 / (d1 - d3);
            q = a + v * ab; u = 3; w[0] = 1 - v; w[1] = v; w[2] = 0; return;
        }
        V3 cp = p - c;
        float d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d6 >= 0.f && d5 <= d6) { q = c; u = 4; w[0] = 0; w[1] = 0; w[2] = 1; return; }
        float vb = d5 * d2 - d1 * d6;
        if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
            float ww = d2 / (d2 - d6);
            q = a + ww * ac; u = 5; w[0] = 1 - ww; w[1] = 0; w[2] = ww; return;
        }
        float va = d3 * d6 - d5 * d4;
        if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
            float ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            q = b + ww * (c - b); u = 6; w[0] = 0; w[1] = 1 - ww; w[2] = ww; return;
        }
        float denom = 1.f / (va + vb + vc);
        float v = vb * denom, ww = vc * denom;
        q = a + ab * v + ac * ww;
        u = 7;
        w[0] = 1 - v - ww; w[1] = v; w[2] = ww;
    }

    // pointOutsideOfPlane: 1 outside, 0 inside, -1 degenerate
    RSG_FN static int outside_plane(V3 p, V3 a, V3 b, V3 c, V3 d) {
        V3 nrm = cross(b - a, c - a);
        float signp = dot(p - a, nrm);
        float signd = dot(d - a, nrm);
        if (signd * signd < (1e-4f * 1e-4f)) return -1;
        return signp * signd < 0.f ? 1 : 0;
    }

    RSG_FN bool closest_tetra(V3 p, V3 a, V3 b, V3 c, V3 d) {
        // finalResult starts as "the point itself, all four vertices"
        used = 15;
        int oabc = outside_plane(p, a, b, c, d);
        int oacd = outside_plane(p, a, c, d, b);
        int oadb = outside_plane(p, a, d, b, c);
        int obdc = outside_plane(p, b, d, c, a);
        if (oabc < 0 || oacd < 0 || oadb < 0 || obdc < 0) { degenerate = true; return false; }
        if (!oabc && !oacd && !oadb && !obdc) return false;
        float best = 3.402823466e+38f;
        V3 q; float w[3]; uint32_t u;
        if (oabc) {
            closest_tri(p, a, b, c, q, w, u);
            float sq = dot(q - p, q - p);
            if (sq < best) {
                best = sq;
                used = (u & 1) | (u & 2) | (u & 4);
                bc[0] = w[0]; bc[1] = w[1]; bc[2] = w[2]; bc[3] = 0;
            }
        }
        if (oacd) {
            closest_tri(p, a, c, d, q, w, u);
            float sq = dot(q - p, q - p);
            if (sq < best) {
                best = sq;
                used = (u & 1) | ((u & 2) ? 4u : 0u) | ((u & 4) ? 8u : 0u);
                bc[0] = w[0]; bc[1] = 0; bc[2] = w[1]; bc[3] = w[2];
            }
        }
        if (oadb) {
            closest_tri(p, a, d, b, q, w, u);
            float sq = dot(q - p, q - p);
            if (sq < best) {
                best = sq;
                used = (u & 1) | ((u & 4) ? 2u : 0u) | ((u & 2) ? 8u : 0u);
                bc[0] = w[0]; bc[1] = w[2]; bc[2] = 0; bc[3] = w[1];
            }
        }
        if (obdc) {
            closest_tri(p, b, d, c, q, w, u);
            float sq = dot(q - p, q - p);
            if (sq < best) {
                best = sq;
                used = ((u & 1) ? 2u : 0u) | ((u & 4) ? 4u : 0u) | ((u & 2) ? 8u : 0u);
                bc[0] = 0; bc[1] = w[0]; bc[2] = w[2]; bc[3] = w[1];
            }
        }
        return true;
    }

    RSG_FN bool update() {
        if (!needs_update) return cached_valid;
        bc_reset();
        needs_update = false;
        if (n == 0) {
            cached_valid = false;
        } else if (n == 1) {
            cp1 = P[0];
            cp2 = Q[0];
            cv = cp1 - cp2;
            bc_reset();
            bc[0] = 1; bc[1] = 0; bc[2] = 0; bc[3] = 0;
            cached_valid = bc_valid();
        } else if (n == 2) {
            V3 from = W[0], to = W[1];
            V3 diff = v3zero() - from;
            V3 v = to - from;
            float t = dot(v, diff);
            if (t > 0) {
                float vv = dot(v, v);
                if (t < vv) {
                    t /= vv;
                    diff -= t * v;
                    used = 3;
                } else {
                    t = 1;
                    diff -= v;
                    used = 2;
                }
            } else {
                t = 0;
                used = 1;
            }
            bc[0] = 1 - t; bc[1] = t; bc[2] = 0; bc[3] = 0;
            cp1 = P[0] + t * (P[1] - P[0]);
            cp2 = Q[0] + t * (Q[1] - Q[0]);
            cv = cp1 - cp2;
            reduce_vertices(used);
            cached_valid = bc_valid();
        } else if (n == 3) {
            V3 q; float w[3]; uint32_t u;
            closest_tri(v3zero(), W[0], W[1], W[2], q, w, u);
            used = u;
            bc[0] = w[0]; bc[1] = w[1]; bc[2] = w[2]; bc[3] = 0;
            cp1 = P[0] * bc[0] + P[1] * bc[1] + P[2] * bc[2];
            cp2 = Q[0] * bc[0] + Q[1] * bc[1] + Q[2] * bc[2];
            cv = cp1 - cp2;
            reduce_vertices(used);
            cached_valid = bc_valid();
        } else if (n == 4) {
            bool sep = closest_tetra(v3zero(), W[0], W[1], W[2], W[3]);
            if (sep) {
                cp1 = P[0] * bc[0] + P[1] * bc[1] + P[2] * bc[2] + P[3] * bc[3];
                cp2 = Q[0] * bc[0] + Q[1] * bc[1] + Q[2] * bc[2] + Q[3] * bc[3];
                cv = cp1 - cp2;
                reduce_vertices(used);
            } else {
                if (degenerate) {
                    cached_valid = false;
                } else {
                    cached_valid = true;
                    cv = v3zero();
                }
                return cached_valid;
            }
            cached_valid = bc_valid();
        } else {
            cached_valid = false;
        }
        return cached_valid;
    }

    RSG_FN bool closest(V3& v) {
        bool ok = update();
        v = cv;
        return ok;
    }
    RSG_FN void compute_points(V3& p1, V3& p2) {
        update();
        p1 = cp1;
        p2 = cp2;
    }
};

// ---------------------------------------------------------------------------
// Ray cast against a convex shape: btSubsimplexConvexCast of a point
// (BulletCollision/NarrowPhaseCollision/btSubSimplexConvexCast.cpp:27-140) as
// run by btCollisionWorld::rayTestSingleInternal for convex shapes
// (BulletCollision/CollisionDispatch/btCollisionWorld.cpp:267-312).  The shape
// does not move; supports include its margin.  Returns true with the hit
// fraction and the (unit) hit normal.

RSG_FN V3 support_with_margin(const CShape& s, V3 d) {
    if (s.kind == SHAPE_BOX) {
        V3 h = s.half + v3(s.margin, s.margin, s.margin);
        return v3(fsel(d.x, h.x, -h.x), fsel(d.y, h.y, -h.y), fsel(d.z, h.z, -h.z));
    }
    // sphere: margin (= radius) along the normalised direction
    V3 n = d;
    if (len2(n) < FLT_EPS * FLT_EPS) n = v3(-1.f, -1.f, -1.f);
    n = normalized(n);
    return v3zero() + s.margin * n;
}

RSG_FN bool ray_convex(V3 from, V3 to, const CShape& B, const Xf& tb, float& fraction, V3& normal) {
    VoronoiSimplex sx;
    sx.reset();
    V3 r = to - from;  // linVelA - linVelB (the shape is static)
    float lambda = 0.f;
    V3 ia = from;      // interpolated origin of the ray point
    Xf ib = tb;
    V3 sup_a = from;
    V3 sup_b = xf_apply(tb, support_with_margin(B, mul_vm(r, tb.R)));
    V3 v = sup_a - sup_b;
    int max_iter = 32;          // float build: 32 iterations, epsilon 1e-4
    V3 n = v3zero();
    float dist2 = len2(v);
    while (dist2 > 0.0001f && max_iter--) {
        sup_a = ia;
        sup_b = xf_apply(ib, support_with_margin(B, mul_vm(v, ib.R)));
        V3 w = sup_a - sup_b;
        float vdw = dot(v, w);
        if (lambda > 1.f) return false;
        if (vdw > 0.f) {
            float vdr = dot(v, r);
            if (vdr >= -(FLT_EPS * FLT_EPS)) return false;
            lambda = lambda - vdw / vdr;
            float sm = 1.f - lambda;
            ia = from * sm + to * lambda;
            ib.t = tb.t * sm + tb.t * lambda;
            w = sup_a - sup_b;
            n = v;
        }
        if (!sx.in_simplex(w)) sx.add_vertex(w, sup_a, sup_b);
        if (sx.closest(v)) dist2 = len2(v);
        else dist2 = 0.f;
    }
    V3 nn = len2(n) >= FLT_EPS * FLT_EPS ? normalized(n) : v3zero();
    if (dot(nn, r) >= -0.f) return false;
    fraction = lambda;
    normal = nn;
    return true;
}

// ---------------------------------------------------------------------------
// Penetration depth of two overlapping convex shapes, as the reference's
// btGjkEpaPenetrationDepthSolver computes it (btGjkEpa2.cpp, Bullet's
// btGjkEpaSolver2 by N. Presson): a GJK distance query on the Minkowski
// difference with margins (Gilbert, Johnson & Keerthi 1988, closest points
// of the sub-simplices by signed volumes), and, when the origin is inside,
// the expanding polytope algorithm (van den Bergen 2001): the polytope grows
// towards the support point along its face nearest the origin until that
// face's distance converges.  Tolerances, iteration limits, pool sizes and
// the order in which faces are created, visited and compared follow the
// reference, since the depth, normal and witness points must agree with it
// bit for bit.

constexpr float GJK_EPS = 0.0001f;          // accuracy, minimum distance and duplicate-vertex tolerance
constexpr int GJK_MAX_STEPS = 128;
constexpr float POLY_EPS = 0.0001f;         // polytope convergence, minimum face area
constexpr float POLY_PLANE_EPS = 0.00001f;  // a vertex this far below a face still sees it
constexpr int POLY_MAX_VERTS = 128, POLY_MAX_FACES = 256, POLY_MAX_STEPS = 255;
constexpr int POLY_WALK_DEPTH = 128;

// Support mapping of the Minkowski difference A - B in A's frame.
struct MinkDiff {
    const CShape* s0;
    const CShape* s1;
    M3 to1;     // shape0 frame -> shape1 frame rotation
    Xf to0;     // shape1 frame -> shape0 frame
    bool margins;
    RSG_FN V3 sup(const CShape& s, V3 d) const { return margins ? support_full(s, d) : support_core(s, d); }
    RSG_FN V3 support0(V3 d) const { return sup(*s0, d); }
    RSG_FN V3 support1(V3 d) const { return xf_apply(to0, sup(*s1, to1 * d)); }
    RSG_FN V3 support(V3 d) const { return support0(d) - support1(-d); }
    RSG_FN V3 support(V3 d, int idx) const { return idx ? support1(d) : support0(d); }
};

// A support point: the unit direction it was found along and the point.
struct SupportPt { V3 dir, w; };

// A simplex of up to four pool entries with barycentric weights.
struct Simplex {
    int16_t idx[4];
    float bary[4];
    int n;
};

// Pool entries 0-3 serve the GJK simplices; the polytope's vertices follow.
constexpr int SUPPORT_POOL = 4 + POLY_MAX_VERTS;

// Signed volume (triple product) a . (b x c), expanded in the reference's
// term order.
RSG_FN float triple(V3 a, V3 b, V3 c) {
    return (a.y * b.z * c.x + a.z * b.x * c.y - a.x * b.z * c.y - a.y * b.x * c.z + a.x * b.y * c.z - a.z * b.y * c.x);
}

// Closest point to the origin on segment ab: squared distance (or -1 for a
// degenerate segment), weights of a and b, and the mask of the vertices
// that support it (bit 0 a, bit 1 b).
RSG_FN float nearest_on_segment(V3 a, V3 b, float* wt, uint32_t& used) {
    const V3 ab = b - a;
    const float l2 = len2(ab);
    if (!(l2 > 0.f)) return -1;
    const float t = l2 > 0 ? -dot(a, ab) / l2 : 0;
    if (t >= 1) { wt[0] = 0; wt[1] = 1; used = 2; return len2(b); }
    if (t <= 0) { wt[0] = 1; wt[1] = 0; used = 1; return len2(a); }
    wt[1] = t;
    wt[0] = 1 - t;
    used = 3;
    return len2(a + ab * t);
}

// Closest point to the origin on triangle abc (bits 0-2 of `used`): on an
// edge whose outer side holds the origin, else inside the face.
RSG_FN float nearest_on_triangle(V3 a, V3 b, V3 c, float* wt, uint32_t& used) {
    const V3 p[3] = {a, b, c};
    const V3 e[3] = {a - b, b - c, c - a};  // edge i runs from p[i] to p[i+1], reversed
    const V3 n = cross(e[0], e[1]);
    const float n2 = len2(n);
    if (!(n2 > 0.f)) return -1;
    float best = -1;
    float ew[2] = {0.f, 0.f};
    uint32_t eu = 0;
    for (int i = 0; i < 3; i++) {
        if (!(dot(p[i], cross(e[i], n)) > 0)) continue;  // origin not outside edge i
        const int j = i == 2 ? 0 : i + 1;
        const float d2 = nearest_on_segment(p[i], p[j], ew, eu);
        if (best < 0 || d2 < best) {
            best = d2;
            used = ((eu & 1) ? 1u << i : 0u) + ((eu & 2) ? 1u << j : 0u);
            wt[i] = ew[0];
            wt[j] = ew[1];
            wt[j == 2 ? 0 : j + 1] = 0;
        }
    }
    if (best < 0) {
        const float k = dot(a, n);
        const float nl = sqrtf(n2);
        const V3 q = n * (k / n2);
        best = len2(q);
        used = 7;
        wt[0] = len(cross(e[1], b - q)) / nl;
        wt[1] = len(cross(e[2], c - q)) / nl;
        wt[2] = 1 - (wt[0] + wt[1]);
    }
    return best;
}

// Closest point to the origin in tetrahedron abcd (bits 0-3): on a face
// through d that the origin is outside of, else the origin itself.
RSG_FN float nearest_in_tetra(V3 a, V3 b, V3 c, V3 d, float* wt, uint32_t& used) {
    const V3 p[4] = {a, b, c, d};
    const V3 e[3] = {a - d, b - d, c - d};
    const float vol = triple(e[0], e[1], e[2]);
    const bool facing = (vol * dot(a, cross(b - c, a - b))) <= 0;
    if (!(facing && fabsf(vol) > 0.f)) return -1;
    float best = -1;
    float fw[3] = {0.f, 0.f, 0.f};
    uint32_t fu = 0;
    for (int i = 0; i < 3; i++) {
        const int j = i == 2 ? 0 : i + 1;
        if (!(vol * dot(d, cross(e[i], e[j])) > 0)) continue;
        const float d2 = nearest_on_triangle(p[i], p[j], d, fw, fu);
        if (best < 0 || d2 < best) {
            best = d2;
            used = ((fu & 1) ? 1u << i : 0u) + ((fu & 2) ? 1u << j : 0u) + ((fu & 4) ? 8u : 0u);
            wt[i] = fw[0];
            wt[j] = fw[1];
            wt[j == 2 ? 0 : j + 1] = 0;
            wt[3] = fw[2];
        }
    }
    if (best < 0) {
        best = 0;
        used = 15;
        wt[0] = triple(c, b, d) / vol;
        wt[1] = triple(a, c, d) / vol;
        wt[2] = triple(b, a, d) / vol;
        wt[3] = 1 - (wt[0] + wt[1] + wt[2]);
    }
    return best;
}

// GJK on the Minkowski difference: the simplex closest to the origin.
// Outcome: 0 separated (dist is the distance), 1 the origin is inside (or
// touching), 2 out of steps.
struct GjkQuery {
    MinkDiff md;
    SupportPt* pool;
    V3 v;                 // current closest point to the origin
    float dist;
    Simplex sets[2];      // the live simplex and the one being built
    int live;
    int16_t spare[4];     // pool slots 0-3 not in the live simplex
    int n_spare;
    Simplex* result;
    int outcome;

    RSG_FN void support(V3 d, SupportPt& sp) const {
        sp.dir = d * (1.f / len(d));
        sp.w = md.support(sp.dir);
    }
    RSG_FN void push(Simplex& sx, V3 d) {
        sx.bary[sx.n] = 0;
        sx.idx[sx.n] = spare[--n_spare];
        support(d, pool[sx.idx[sx.n++]]);
    }
    RSG_FN void pop(Simplex& sx) { spare[n_spare++] = sx.idx[--sx.n]; }
    RSG_FN V3 wof(const Simplex& sx, int k) const { return pool[sx.idx[k]].w; }

    RSG_FN int run(V3 guess) {
        for (int k = 0; k < 4; k++) spare[k] = (int16_t)k;
        n_spare = 4;
        live = 0;
        outcome = 0;
        dist = 0;
        sets[0].n = 0;
        v = guess;
        const float g2 = len2(v);
        push(sets[0], g2 > 0 ? -v : v3(1, 0, 0));
        sets[0].bary[0] = 1;
        v = pool[sets[0].idx[0]].w;
        float d2 = g2;
        float lower = 0;  // best lower bound on the distance so far
        V3 recent[4] = {v, v, v, v};
        int slot = 0;
        int steps = 0;
        do {
            Simplex& cur = sets[live];
            Simplex& nxt = sets[1 - live];
            const float vl = len(v);
            if (vl < GJK_EPS) { outcome = 1; break; }
            push(cur, -v);
            const V3 w = wof(cur, cur.n - 1);
            bool repeat = false;
            for (int k = 0; k < 4 && !repeat; k++) repeat = len2(w - recent[k]) < GJK_EPS;
            if (repeat) { pop(sets[live]); break; }
            slot = (slot + 1) & 3;
            recent[slot] = w;
            lower = fmaxf(dot(v, w) / vl, lower);
            if (((vl - lower) - (GJK_EPS * vl)) <= 0) { pop(sets[live]); break; }
            float wt[4];
            uint32_t used = 0;
            if (cur.n == 2) d2 = nearest_on_segment(wof(cur, 0), wof(cur, 1), wt, used);
            else if (cur.n == 3) d2 = nearest_on_triangle(wof(cur, 0), wof(cur, 1), wof(cur, 2), wt, used);
            else d2 = nearest_in_tetra(wof(cur, 0), wof(cur, 1), wof(cur, 2), wof(cur, 3), wt, used);
            if (!(d2 >= 0)) { pop(sets[live]); break; }
            // keep the supporting vertices, rebuild v from their weights
            nxt.n = 0;
            v = v3zero();
            live = 1 - live;
            for (int k = 0; k < cur.n; k++) {
                if (used & (1u << k)) {
                    nxt.idx[nxt.n] = cur.idx[k];
                    nxt.bary[nxt.n++] = wt[k];
                    v += pool[cur.idx[k]].w * wt[k];
                } else {
                    spare[n_spare++] = cur.idx[k];
                }
            }
            if (used == 15) outcome = 1;
            outcome = (++steps < GJK_MAX_STEPS) ? outcome : 2;
        } while (outcome == 0);
        result = &sets[live];
        if (outcome == 0) dist = len(v);
        else if (outcome == 1) dist = 0;
        return outcome;
    }

    RSG_FN bool tetra_has_volume() const {
        const Simplex& r = *result;
        return fabsf(triple(wof(r, 0) - wof(r, 3), wof(r, 1) - wof(r, 3), wof(r, 2) - wof(r, 3))) > 0;
    }

    // Trial direction k at simplex size n (n = 1: +-x, +-y, +-z; n = 2: +-
    // (edge x axis) per axis; n = 3: +- face normal), or false when k is past
    // the last one.  Degenerate cross products offer no trial.
    RSG_FN bool trial_dir(int n, int& k, V3& d) const {
        const Simplex& r = *result;
        while (true) {
            if (n == 1) {
                if (k >= 6) return false;
                V3 axis = v3zero();
                set(axis, k >> 1, 1.f);
                d = (k & 1) ? -axis : axis;
                return true;
            }
            if (n == 2) {
                if (k >= 6) return false;
                V3 axis = v3zero();
                set(axis, k >> 1, 1.f);
                const V3 c = cross(wof(r, 1) - wof(r, 0), axis);
                if (!(len2(c) > 0)) { k = (k | 1) + 1; continue; }
                d = (k & 1) ? -c : c;
                return true;
            }
            if (k >= 2) return false;
            const V3 c = cross(wof(r, 1) - wof(r, 0), wof(r, 2) - wof(r, 0));
            if (!(len2(c) > 0)) return false;
            d = (k & 1) ? -c : c;
            return true;
        }
    }

    // Grows the final simplex to a tetrahedron of non-zero volume around the
    // origin, trying the trial directions depth first (the simplex of a
    // touching contact is flat).
    RSG_FN bool grow_to_tetra() {
        Simplex& r = *result;
        const int base = r.n;
        if (base == 4) return tetra_has_volume();
        int next_trial[5] = {0, 0, 0, 0, 0};
        int n = base;
        while (true) {
            if (n == 4) {
                if (tetra_has_volume()) return true;
                pop(r);
                n--;
                continue;
            }
            V3 d;
            int k = next_trial[n];
            const bool more = trial_dir(n, k, d);
            next_trial[n] = k + 1;
            if (!more) {
                if (n == base) return false;
                pop(r);
                n--;
                continue;
            }
            push(r, d);
            n++;
            next_trial[n] = 0;
        }
    }
};

// A polytope face: outward unit normal, distance of its plane from the
// origin, vertices (pool indices), neighbours across its edges (edge k runs
// from v[k] to v[k+1]) with the neighbour's matching edge, the walk mark, and
// the creation stamp that orders the live faces (newest first).
struct PolyFace {
    V3 n;
    float d;
    int16_t v[3];
    int16_t nb[3];
    uint8_t nb_edge[3];
    uint8_t mark;
    uint16_t stamp;
    bool live;
};

// Expanding polytope over a vertex pool (entries 4 on) and a face pool.
// Outcome: 0 valid, 1 touching, 2 degenerate face, 3 non-convex, 4 bad
// horizon, 5 out of faces, 6 out of vertices, 7 converged, 8 fallback
// (no enclosing tetrahedron), 9 failed.
struct Polytope {
    int outcome;
    Simplex result;   // the converged face's vertices and barycentric weights
    V3 normal;
    float depth;
    SupportPt* pool;
    PolyFace* faces;
    int n_used;       // faces[0 .. n_used) have been handed out at least once
    int16_t free_list[POLY_MAX_FACES];
    int n_free;
    int n_live;
    uint16_t clock;
    int next_vert;

    RSG_FN void reset(SupportPt* sp, PolyFace* fp) {
        pool = sp;
        faces = fp;
        outcome = 9;
        normal = v3zero();
        depth = 0;
        next_vert = 0;
        n_used = 0;
        n_free = 0;
        n_live = 0;
        clock = 0;
    }
    RSG_FN void link(int fa, int ea, int fb, int eb) {
        faces[fa].nb_edge[ea] = (uint8_t)eb; faces[fa].nb[ea] = (int16_t)fb;
        faces[fb].nb_edge[eb] = (uint8_t)ea; faces[fb].nb[eb] = (int16_t)fa;
    }
    RSG_FN void retire(int f) {
        faces[f].live = false;
        n_live--;
        free_list[n_free++] = (int16_t)f;
    }
    // Distance of the face plane's nearest point to the origin when that
    // point lies outside edge ab (the origin projects past the edge).
    RSG_FN bool edge_dist(int f, int a, int b, float& out) const {
        const V3 wa = pool[a].w, wb = pool[b].w;
        const V3 ab = wb - wa;
        const V3 edge_out = cross(ab, faces[f].n);
        if (!(dot(wa, edge_out) < 0)) return false;
        const float ab2 = len2(ab);
        const float a_ab = dot(wa, ab);
        const float b_ab = dot(wb, ab);
        if (a_ab > 0) out = len(wa);
        else if (b_ab < 0) out = len(wb);
        else {
            const float ab_dot = dot(wa, wb);
            out = sqrtf(fmaxf((len2(wa) * len2(wb) - ab_dot * ab_dot) / ab2, 0.f));
        }
        return true;
    }
    // A new face abc, or -1 (outcome set) when it is degenerate, faces away
    // from the origin (unless forced) or no face is free.
    RSG_FN int add_face(int a, int b, int c, bool forced) {
        int f;
        if (n_free > 0) f = free_list[--n_free];
        else if (n_used < POLY_MAX_FACES) f = n_used++;
        else {
            outcome = 5;
            return -1;
        }
        PolyFace& F = faces[f];
        F.live = true;
        F.stamp = ++clock;
        n_live++;
        F.mark = 0;
        F.v[0] = (int16_t)a; F.v[1] = (int16_t)b; F.v[2] = (int16_t)c;
        F.n = cross(pool[b].w - pool[a].w, pool[c].w - pool[a].w);
        const float nl = len(F.n);
        if (nl > POLY_EPS) {
            if (!(edge_dist(f, a, b, F.d) || edge_dist(f, b, c, F.d) || edge_dist(f, c, a, F.d)))
                F.d = dot(pool[a].w, F.n) / nl;
            F.n = F.n * (1.f / nl);
            if (forced || F.d >= -POLY_PLANE_EPS) return f;
            outcome = 3;
        } else {
            outcome = 2;
        }
        retire(f);
        return -1;
    }
    // The live face nearest the origin (the newest one on equal distance).
    RSG_FN int nearest_face() const {
        int best = -1;
        for (int f = 0; f < n_used; f++)
            if (faces[f].live && (best < 0 || faces[f].stamp > faces[best].stamp)) best = f;
        float bd = faces[best].d * faces[best].d;
        for (int f = 0; f < n_used; f++) {
            if (!faces[f].live || f == best) continue;
            const float d2 = faces[f].d * faces[f].d;
            if (d2 < bd || (d2 == bd && faces[f].stamp > faces[best].stamp)) {
                best = f;
                bd = d2;
            }
        }
        return best;
    }

    struct Rim { int first, last, count; };  // the new faces along the horizon

    // Walks the faces visible from vertex w, starting across edge `e0` of the
    // face being replaced (into face f0), depth first: a face that w sees is
    // marked, its two other neighbours walked, and it is retired once both
    // walks succeed; a face w does not see gets a new face on the shared
    // edge, linked to the previous new face along the horizon.
    RSG_FN bool walk_horizon(int mark, int w, int f0, int e0, Rim& rim) {
        struct Step { int16_t f; uint8_t e, stage; };
        Step st[POLY_WALK_DEPTH];
        int top = 0;
        st[0].f = (int16_t)f0; st[0].e = (uint8_t)e0; st[0].stage = 0;
        bool ok = false;
        while (top >= 0) {
            Step& s = st[top];
            const int f = s.f;
            const int ea = s.e == 2 ? 0 : s.e + 1;   // the edge after e
            const int eb = ea == 2 ? 0 : ea + 1;      // and the one after that
            if (s.stage == 0) {
                if (faces[f].mark == mark) { ok = false; top--; continue; }
                if ((dot(faces[f].n, pool[w].w) - faces[f].d) < -POLY_PLANE_EPS) {
                    // w is below this face: the shared edge is on the horizon
                    const int nf = add_face(faces[f].v[ea], faces[f].v[s.e], w, false);
                    ok = nf >= 0;
                    if (ok) {
                        link(nf, 0, f, s.e);
                        if (rim.last >= 0) link(rim.last, 1, nf, 2);
                        else rim.first = nf;
                        rim.last = nf;
                        rim.count++;
                    }
                    top--;
                    continue;
                }
                faces[f].mark = (uint8_t)mark;
                s.stage = 1;
                if (top + 1 >= POLY_WALK_DEPTH) { outcome = 4; return false; }
                Step& c = st[++top];
                c.f = faces[f].nb[ea]; c.e = faces[f].nb_edge[ea]; c.stage = 0;
            } else if (s.stage == 1) {
                if (!ok) { top--; continue; }
                s.stage = 2;
                if (top + 1 >= POLY_WALK_DEPTH) { outcome = 4; return false; }
                Step& c = st[++top];
                c.f = faces[f].nb[eb]; c.e = faces[f].nb_edge[eb]; c.stage = 0;
            } else {
                if (ok) retire(f);
                top--;
            }
        }
        return ok;
    }

    // Expands from the GJK query's final simplex (which encloses the origin).
    RSG_FN int expand(GjkQuery& gjk, V3 guess) {
        Simplex& sx = *gjk.result;
        if (sx.n > 1 && gjk.grow_to_tetra()) {
            outcome = 0;
            next_vert = 0;
            if (triple(gjk.wof(sx, 0) - gjk.wof(sx, 3), gjk.wof(sx, 1) - gjk.wof(sx, 3),
                       gjk.wof(sx, 2) - gjk.wof(sx, 3)) < 0) {
                // orient the tetrahedron so that its faces point outwards
                const int16_t ti = sx.idx[0]; sx.idx[0] = sx.idx[1]; sx.idx[1] = ti;
                const float tb = sx.bary[0]; sx.bary[0] = sx.bary[1]; sx.bary[1] = tb;
            }
            const int t0 = add_face(sx.idx[0], sx.idx[1], sx.idx[2], true);
            const int t1 = add_face(sx.idx[1], sx.idx[0], sx.idx[3], true);
            const int t2 = add_face(sx.idx[2], sx.idx[1], sx.idx[3], true);
            const int t3 = add_face(sx.idx[0], sx.idx[2], sx.idx[3], true);
            if (n_live == 4) {
                int best = nearest_face();
                PolyFace final_face = faces[best];
                int mark = 0;
                link(t0, 0, t1, 0);
                link(t0, 1, t2, 0);
                link(t0, 2, t3, 0);
                link(t1, 1, t3, 2);
                link(t1, 2, t2, 1);
                link(t2, 2, t3, 1);
                outcome = 0;
                for (int step = 0; step < POLY_MAX_STEPS; step++) {
                    if (next_vert >= POLY_MAX_VERTS) { outcome = 6; break; }
                    Rim rim{-1, -1, 0};
                    const int w = 4 + next_vert++;
                    faces[best].mark = (uint8_t)(++mark);
                    gjk.support(faces[best].n, pool[w]);
                    const float gain = dot(faces[best].n, pool[w].w) - faces[best].d;
                    if (!(gain > POLY_EPS)) { outcome = 7; break; }
                    bool ok = true;
                    for (int k = 0; k < 3 && ok; k++) ok &= walk_horizon(mark, w, faces[best].nb[k], faces[best].nb_edge[k], rim);
                    if (!(ok && rim.count >= 3)) { outcome = 4; break; }
                    link(rim.last, 1, rim.first, 2);
                    retire(best);
                    best = nearest_face();
                    final_face = faces[best];
                }
                // barycentric weights of the origin's projection on the face
                const V3 proj = final_face.n * final_face.d;
                normal = final_face.n;
                depth = final_face.d;
                result.n = 3;
                for (int k = 0; k < 3; k++) result.idx[k] = final_face.v[k];
                for (int k = 0; k < 3; k++) {
                    const int a = final_face.v[k == 2 ? 0 : k + 1], b = final_face.v[k == 0 ? 2 : k - 1];
                    result.bary[k] = len(cross(pool[a].w - proj, pool[b].w - proj));
                }
                const float total = result.bary[0] + result.bary[1] + result.bary[2];
                for (int k = 0; k < 3; k++) result.bary[k] /= total;
                return outcome;
            }
        }
        // no tetrahedron: the guess direction at zero depth
        outcome = 8;
        normal = -guess;
        const float nl = len(normal);
        normal = nl > 0 ? normal * (1.f / nl) : v3(1, 0, 0);
        depth = 0;
        result.n = 1;
        result.idx[0] = sx.idx[0];
        result.bary[0] = 1;
        return outcome;
    }
};

struct PenResult {
    V3 witness[2];
    V3 normal;
    float distance;
};

RSG_FN void mink_init(MinkDiff& md, const CShape& s0, const Xf& t0, const CShape& s1, const Xf& t1, bool margins) {
    md.s0 = &s0;
    md.s1 = &s1;
    md.to1 = transpose(t1.R) * t0.R;
    // shape 1's frame in shape 0's: t0^-1 * t1
    md.to0.R = transpose(t0.R) * t1.R;
    md.to0.t = mul_vm(t1.t - t0.t, t0.R);
    md.margins = margins;
}

// Penetration of the shapes with their margins (btGjkEpaSolver2::Penetration):
// GJK, then the polytope when the origin is inside the difference.
RSG_FN bool penetration(const CShape& s0, const Xf& t0, const CShape& s1, const Xf& t1, V3 guess, PenResult& res,
                        SupportPt* pool, PolyFace* faces, Polytope& poly) {
    GjkQuery gjk;
    mink_init(gjk.md, s0, t0, s1, t1, true);
    gjk.pool = pool;
    res.witness[0] = res.witness[1] = v3zero();
    if (gjk.run(-guess) != 1) return false;
    poly.reset(pool, faces);
    if (poly.expand(gjk, -guess) == 9) return false;
    V3 w0 = v3zero();
    for (int k = 0; k < poly.result.n; k++) w0 += gjk.md.support(pool[poly.result.idx[k]].dir, 0) * poly.result.bary[k];
    res.witness[0] = xf_apply(t0, w0);
    res.witness[1] = xf_apply(t0, w0 - poly.normal * poly.depth);
    res.normal = -poly.normal;
    res.distance = -poly.depth;
    return true;
}

// Distance of the shapes without margins (btGjkEpaSolver2::Distance).
RSG_FN bool separation(const CShape& s0, const Xf& t0, const CShape& s1, const Xf& t1, V3 guess, PenResult& res,
                       SupportPt* pool) {
    GjkQuery gjk;
    mink_init(gjk.md, s0, t0, s1, t1, false);
    gjk.pool = pool;
    res.witness[0] = res.witness[1] = v3zero();
    if (gjk.run(guess) != 0) return false;
    V3 w0 = v3zero(), w1 = v3zero();
    for (int k = 0; k < gjk.result->n; k++) {
        const float b = gjk.result->bary[k];
        w0 += gjk.md.support(pool[gjk.result->idx[k]].dir, 0) * b;
        w1 += gjk.md.support(-pool[gjk.result->idx[k]].dir, 1) * b;
    }
    res.witness[0] = xf_apply(t0, w0);
    res.witness[1] = xf_apply(t0, w1);
    res.normal = w0 - w1;
    res.distance = len(res.normal);
    res.normal = res.normal * (1.f / (res.distance > GJK_EPS ? res.distance : 1.f));
    return true;
}

// Work space of a penetration query: the support-point pool and the
// polytope's face pool.  The caller places it (private memory, or a
// per-arena block in global memory that a lane reads without the per-lane
// interleaving of private memory).
struct EpaSpace {
    SupportPt pool[SUPPORT_POOL];
    PolyFace faces[POLY_MAX_FACES];
};

// Penetration depth with the reference's nine guess directions
// (btGjkEpaPenetrationDepthSolver::calcPenDepth): the first guess for which
// the penetration query succeeds gives the contact; a guess whose distance
// query succeeds instead reports the shapes apart.
RSG_FN bool calc_pen_depth(const CShape& a, const CShape& b, const Xf& ta, const Xf& tb, V3& v, V3& wa, V3& wb,
                           EpaSpace& space) {
    Polytope poly;
    const V3 guesses[9] = {safe_normalized(tb.t - ta.t), safe_normalized(ta.t - tb.t), v3(0, 0, 1), v3(0, 1, 0),
                           v3(1, 0, 0), v3(1, 1, 0), v3(1, 1, 1), v3(0, 1, 1), v3(1, 0, 1)};
    for (int i = 0; i < 9; i++) {
        PenResult r;
        if (penetration(a, ta, b, tb, guesses[i], r, space.pool, space.faces, poly)) {
            wa = r.witness[0]; wb = r.witness[1]; v = r.normal;
            return true;
        }
        if (separation(a, ta, b, tb, guesses[i], r, space.pool)) {
            wa = r.witness[0]; wb = r.witness[1]; v = r.normal;
            return false;
        }
    }
    wa = v3zero(); wb = v3zero(); v = v3zero();
    return false;
}

// ---------------------------------------------------------------------------
// btGjkPairDetector::getClosestPointsNonVirtual.  Returns 1 with the
// reported contact (normal on B pointing to A, point on B in world space,
// signed distance) when the shapes are closer than sqrt(max_dist2), else 0.
// Without a penetration work space (`space` null) a query that needs the
// penetration pass returns -1 and reports nothing.

RSG_FN int gjk_closest(const CShape& A, const CShape& B, const Xf& ta, const Xf& tb, float max_dist2,
                       V3& normal_out, V3& point_b_out, float& dist_out, EpaSpace* space) {
    const float REL_ERROR2 = 1.0e-6f;
    float distance = 0.f;
    V3 normalInB = v3zero();
    V3 pointOnA = v3zero(), pointOnB = v3zero();
    Xf lta = ta, ltb = tb;
    V3 offset = (ta.t + tb.t) * 0.5f;
    lta.t -= offset;
    ltb.t -= offset;
    float marginA = A.margin, marginB = B.margin;
    int cur_iter = 0;
    V3 sep = v3(0, 1, 0);
    bool isValid = false, checkSimplex = false;
    int degenerate = 0;
    VoronoiSimplex sx;
    float squaredDistance = 1e18f;
    float delta = 0.f;
    float margin = marginA + marginB;
    sx.reset();
    while (true) {
        V3 sepA = mul_vm(-sep, ta.R);
        V3 sepB = mul_vm(sep, tb.R);
        V3 pInA = support_core(A, sepA);
        V3 qInB = support_core(B, sepB);
        V3 pW = xf_apply(lta, pInA);
        V3 qW = xf_apply(ltb, qInB);
        V3 w = pW - qW;
        delta = dot(sep, w);
        if (delta > 0.f && delta * delta > squaredDistance * max_dist2) { degenerate = 10; checkSimplex = true; break; }
        if (sx.in_simplex(w)) { degenerate = 1; checkSimplex = true; break; }
        float f0 = squaredDistance - delta;
        float f1 = squaredDistance * REL_ERROR2;
        if (f0 <= f1) { degenerate = f0 <= 0.f ? 2 : 11; checkSimplex = true; break; }
        sx.add_vertex(w, pW, qW);
        V3 nsep;
        if (!sx.closest(nsep)) { degenerate = 3; checkSimplex = true; break; }
        if (len2(nsep) < REL_ERROR2) { sep = nsep; degenerate = 6; checkSimplex = true; break; }
        float prev = squaredDistance;
        squaredDistance = len2(nsep);
        if (prev - squaredDistance <= FLT_EPS * prev) { checkSimplex = true; degenerate = 12; break; }
        sep = nsep;
        if (cur_iter++ > 1000) break;
        if (sx.full()) { degenerate = 13; break; }
    }
    if (checkSimplex) {
        sx.compute_points(pointOnA, pointOnB);
        normalInB = sep;
        float lenSqr = len2(sep);
        if (lenSqr < 0.0001f) degenerate = 5;
        if (lenSqr > FLT_EPS * FLT_EPS) {
            float rlen = 1.f / sqrtf(lenSqr);
            normalInB *= rlen;
            float s = sqrtf(squaredDistance);
            pointOnA -= sep * (marginA / s);
            pointOnB += sep * (marginB / s);
            distance = (1.f / rlen) - margin;
            isValid = true;
        }
    }
    bool catchDegenerate = degenerate != 0 && (distance + margin) < 0.01f;
    if (!isValid || catchDegenerate) {
        if (!space) return -1;
        V3 tA, tB;
        sep = v3zero();
        bool ok2 = calc_pen_depth(A, B, lta, ltb, sep, tA, tB, *space);
        if (ok2) {
            V3 tn = tB - tA;
            float ls = len2(tn);
            if (ls <= FLT_EPS * FLT_EPS) { tn = sep; ls = len2(sep); }
            if (ls > FLT_EPS * FLT_EPS) {
                tn = tn * (1.f / sqrtf(ls));
                float d2 = -len(tA - tB);
                if (!isValid || d2 < distance) {
                    distance = d2; pointOnA = tA; pointOnB = tB; normalInB = tn; isValid = true;
                }
            }
        } else {
            if (len2(sep) > 0.f) {
                float d2 = len(tA - tB) - margin;
                if (!isValid || d2 < distance) {
                    distance = d2;
                    pointOnA = tA - sep * marginA;
                    pointOnB = tB + sep * marginB;
                    normalInB = normalized(sep);
                    isValid = true;
                }
            }
        }
    }
    if (isValid && (distance < 0.f || distance * distance < max_dist2)) {
        V3 posA = shape_aabb_center(A, lta);
        V3 posB = shape_aabb_center(B, ltb);
        if (dot(posA - posB, normalInB) < 0.f) normalInB = normalInB * -1.f;
        normal_out = normalInB;
        point_b_out = pointOnB + offset;
        dist_out = distance;
        return 1;
    }
    return 0;
}

}  // namespace rsg
