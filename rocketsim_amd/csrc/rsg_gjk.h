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
// btGjkEpaSolver2: GJK (Presson) + EPA, used when the cores overlap

constexpr float G2_ACCURACY = 0.0001f, G2_MIN_DISTANCE = 0.0001f, G2_DUP_EPS = 0.0001f;
constexpr int G2_MAX_ITER = 128;
constexpr float EPA_ACCURACY = 0.0001f, EPA_PLANE_EPS = 0.00001f;
constexpr int EPA_MAX_VERTICES = 128, EPA_MAX_FACES = 256, EPA_MAX_ITER = 255;
constexpr int EPA_STACK = 128;

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

struct SV { V3 d, w; };

struct G2Simplex {
    int16_t c[4];
    float p[4];
    int rank;
};

// vertex pool shared by the GJK stage (entries 0..3) and EPA (4..)
constexpr int SV_POOL = 4 + EPA_MAX_VERTICES;

struct Gjk2 {
    MinkDiff shape;
    SV* sv;           // pool
    V3 ray;
    float distance;
    G2Simplex simplices[2];
    int16_t freel[4];
    int nfree;
    int current;
    G2Simplex* simplex;
    int status;       // 0 valid, 1 inside, 2 failed

    RSG_FN void getsupport(V3 d, SV& s) const {
        s.d = d * (1.f / len(d));
        s.w = shape.support(s.d);
    }
    RSG_FN void removevertice(G2Simplex& sx) { freel[nfree++] = sx.c[--sx.rank]; }
    RSG_FN void appendvertice(G2Simplex& sx, V3 v) {
        sx.p[sx.rank] = 0;
        sx.c[sx.rank] = freel[--nfree];
        getsupport(v, sv[sx.c[sx.rank++]]);
    }
    RSG_FN static float det3(V3 a, V3 b, V3 c) {
        return (a.y * b.z * c.x + a.z * b.x * c.y - a.x * b.z * c.y - a.y * b.x * c.z + a.x * b.y * c.z - a.z * b.y * c.x);
    }
    RSG_FN static float project2(V3 a, V3 b, float* w, uint32_t& m) {
        V3 d = b - a;
        float l = len2(d);
        if (l > 0.f) {
            float t = l > 0 ? -dot(a, d) / l : 0;
            if (t >= 1) { w[0] = 0; w[1] = 1; m = 2; return len2(b); }
            if (t <= 0) { w[0] = 1; w[1] = 0; m = 1; return len2(a); }
            w[1] = t; w[0] = 1 - t; m = 3;
            return len2(a + d * t);
        }
        return -1;
    }
    RSG_FN static float project3(V3 a, V3 b, V3 c, float* w, uint32_t& m) {
        const int imd3[3] = {1, 2, 0};
        V3 vt[3] = {a, b, c};
        V3 dl[3] = {a - b, b - c, c - a};
        V3 n = cross(dl[0], dl[1]);
        float l = len2(n);
        if (l > 0.f) {
            float mindist = -1;
            float subw[2] = {0.f, 0.f};
            uint32_t subm = 0;
            for (int i = 0; i < 3; i++) {
                if (dot(vt[i], cross(dl[i], n)) > 0) {
                    int j = imd3[i];
                    float subd = project2(vt[i], vt[j], subw, subm);
                    if (mindist < 0 || subd < mindist) {
                        mindist = subd;
                        m = ((subm & 1) ? 1u << i : 0u) + ((subm & 2) ? 1u << j : 0u);
                        w[i] = subw[0];
                        w[j] = subw[1];
                        w[imd3[j]] = 0;
                    }
                }
            }
            if (mindist < 0) {
                float d = dot(a, n);
                float s = sqrtf(l);
                V3 p = n * (d / l);
                mindist = len2(p);
                m = 7;
                w[0] = len(cross(dl[1], b - p)) / s;
                w[1] = len(cross(dl[2], c - p)) / s;
                w[2] = 1 - (w[0] + w[1]);
            }
            return mindist;
        }
        return -1;
    }
    RSG_FN static float project4(V3 a, V3 b, V3 c, V3 d, float* w, uint32_t& m) {
        const int imd3[3] = {1, 2, 0};
        V3 vt[4] = {a, b, c, d};
        V3 dl[3] = {a - d, b - d, c - d};
        float vl = det3(dl[0], dl[1], dl[2]);
        bool ng = (vl * dot(a, cross(b - c, a - b))) <= 0;
        if (ng && fabsf(vl) > 0.f) {
            float mindist = -1;
            float subw[3] = {0.f, 0.f, 0.f};
            uint32_t subm = 0;
            for (int i = 0; i < 3; i++) {
                int j = imd3[i];
                float s = vl * dot(d, cross(dl[i], dl[j]));
                if (s > 0) {
                    float subd = project3(vt[i], vt[j], d, subw, subm);
                    if (mindist < 0 || subd < mindist) {
                        mindist = subd;
                        m = ((subm & 1) ? 1u << i : 0u) + ((subm & 2) ? 1u << j : 0u) + ((subm & 4) ? 8u : 0u);
                        w[i] = subw[0];
                        w[j] = subw[1];
                        w[imd3[j]] = 0;
                        w[3] = subw[2];
                    }
                }
            }
            if (mindist < 0) {
                mindist = 0;
                m = 15;
                w[0] = det3(c, b, d) / vl;
                w[1] = det3(a, c, d) / vl;
                w[2] = det3(b, a, d) / vl;
                w[3] = 1 - (w[0] + w[1] + w[2]);
            }
            return mindist;
        }
        return -1;
    }

    RSG_FN int evaluate(V3 guess) {
        int iterations = 0;
        float sqdist = 0, alpha = 0;
        V3 lastw[4];
        int clastw = 0;
        freel[0] = 0; freel[1] = 1; freel[2] = 2; freel[3] = 3;
        nfree = 4;
        current = 0;
        status = 0;
        distance = 0;
        simplices[0].rank = 0;
        ray = guess;
        float sqrl = len2(ray);
        appendvertice(simplices[0], sqrl > 0 ? -ray : v3(1, 0, 0));
        simplices[0].p[0] = 1;
        ray = sv[simplices[0].c[0]].w;
        sqdist = sqrl;
        lastw[0] = lastw[1] = lastw[2] = lastw[3] = ray;
        do {
            int next = 1 - current;
            G2Simplex& cs = simplices[current];
            G2Simplex& ns = simplices[next];
            float rl = len(ray);
            if (rl < G2_MIN_DISTANCE) { status = 1; break; }
            appendvertice(cs, -ray);
            V3 w = sv[cs.c[cs.rank - 1]].w;
            bool found = false;
            for (int i = 0; i < 4; i++)
                if (len2(w - lastw[i]) < G2_DUP_EPS) { found = true; break; }
            if (found) { removevertice(simplices[current]); break; }
            clastw = (clastw + 1) & 3;
            lastw[clastw] = w;
            float omega = dot(ray, w) / rl;
            alpha = fmaxf(omega, alpha);
            if (((rl - alpha) - (G2_ACCURACY * rl)) <= 0) { removevertice(simplices[current]); break; }
            float weights[4];
            uint32_t mask = 0;
            if (cs.rank == 2) sqdist = project2(sv[cs.c[0]].w, sv[cs.c[1]].w, weights, mask);
            else if (cs.rank == 3) sqdist = project3(sv[cs.c[0]].w, sv[cs.c[1]].w, sv[cs.c[2]].w, weights, mask);
            else sqdist = project4(sv[cs.c[0]].w, sv[cs.c[1]].w, sv[cs.c[2]].w, sv[cs.c[3]].w, weights, mask);
            if (sqdist >= 0) {
                ns.rank = 0;
                ray = v3zero();
                current = next;
                for (int i = 0, ni = cs.rank; i < ni; i++) {
                    if (mask & (1u << i)) {
                        ns.c[ns.rank] = cs.c[i];
                        ns.p[ns.rank++] = weights[i];
                        ray += sv[cs.c[i]].w * weights[i];
                    } else {
                        freel[nfree++] = cs.c[i];
                    }
                }
                if (mask == 15) status = 1;
            } else {
                removevertice(simplices[current]);
                break;
            }
            status = (++iterations < G2_MAX_ITER) ? status : 2;
        } while (status == 0);
        simplex = &simplices[current];
        if (status == 0) distance = len(ray);
        else if (status == 1) distance = 0;
        return status;
    }

    // EncloseOrigin, unrolled over the simplex rank (1 -> 4)
    RSG_FN bool enclose4() {
        return fabsf(det3(sv[simplex->c[0]].w - sv[simplex->c[3]].w, sv[simplex->c[1]].w - sv[simplex->c[3]].w,
                          sv[simplex->c[2]].w - sv[simplex->c[3]].w)) > 0;
    }
    RSG_FN bool enclose3() {
        V3 nn = cross(sv[simplex->c[1]].w - sv[simplex->c[0]].w, sv[simplex->c[2]].w - sv[simplex->c[0]].w);
        if (len2(nn) > 0) {
            appendvertice(*simplex, nn);
            if (enclose4()) return true;
            removevertice(*simplex);
            appendvertice(*simplex, -nn);
            if (enclose4()) return true;
            removevertice(*simplex);
        }
        return false;
    }
    RSG_FN bool enclose2() {
        V3 d = sv[simplex->c[1]].w - sv[simplex->c[0]].w;
        for (int i = 0; i < 3; i++) {
            V3 axis = v3zero();
            set(axis, i, 1.f);
            V3 p = cross(d, axis);
            if (len2(p) > 0) {
                appendvertice(*simplex, p);
                if (enclose3()) return true;
                removevertice(*simplex);
                appendvertice(*simplex, -p);
                if (enclose3()) return true;
                removevertice(*simplex);
            }
        }
        return false;
    }
    RSG_FN bool enclose1() {
        for (int i = 0; i < 3; i++) {
            V3 axis = v3zero();
            set(axis, i, 1.f);
            appendvertice(*simplex, axis);
            if (enclose2()) return true;
            removevertice(*simplex);
            appendvertice(*simplex, -axis);
            if (enclose2()) return true;
            removevertice(*simplex);
        }
        return false;
    }
    RSG_FN bool enclose_origin() {
        switch (simplex->rank) {
            case 1: return enclose1();
            case 2: return enclose2();
            case 3: return enclose3();
            case 4: return enclose4();
        }
        return false;
    }
};

struct EpaFace {
    V3 n;
    float d;
    int16_t c[3];
    int16_t f[3];
    int16_t l[2];
    uint8_t e[3];
    uint8_t pass;
};

struct Epa {
    int status;  // 0 valid, 1 touching, 2 degenerated, 3 nonconvex, 4 invalid hull, 5 out of faces,
                 // 6 out of vertices, 7 accuracy reached, 8 fallback, 9 failed
    G2Simplex result;
    V3 normal;
    float depth;
    SV* sv;               // shared pool (EPA vertices from index 4)
    EpaFace* fc;          // face pool, EPA_MAX_FACES entries
    int nextsv;
    int16_t hull_root, stock_root;
    int hull_count, stock_count;

    RSG_FN void bind(int fa, int ea, int fb, int eb) {
        fc[fa].e[ea] = (uint8_t)eb; fc[fa].f[ea] = (int16_t)fb;
        fc[fb].e[eb] = (uint8_t)ea; fc[fb].f[eb] = (int16_t)fa;
    }
    RSG_FN void list_append(int16_t& root, int& count, int f) {
        fc[f].l[0] = -1;
        fc[f].l[1] = root;
        if (root >= 0) fc[root].l[0] = (int16_t)f;
        root = (int16_t)f;
        count++;
    }
    RSG_FN void list_remove(int16_t& root, int& count, int f) {
        if (fc[f].l[1] >= 0) fc[fc[f].l[1]].l[0] = fc[f].l[0];
        if (fc[f].l[0] >= 0) fc[fc[f].l[0]].l[1] = fc[f].l[1];
        if (f == root) root = fc[f].l[1];
        count--;
    }
    RSG_FN void init(SV* pool, EpaFace* faces) {
        sv = pool;
        fc = faces;
        status = 9;
        normal = v3zero();
        depth = 0;
        nextsv = 0;
        hull_root = -1; stock_root = -1; hull_count = 0; stock_count = 0;
        for (int i = 0; i < EPA_MAX_FACES; i++) list_append(stock_root, stock_count, EPA_MAX_FACES - i - 1);
    }
    RSG_FN bool edgedist(int face, int a, int b, float& dist) const {
        V3 ba = sv[b].w - sv[a].w;
        V3 n_ab = cross(ba, fc[face].n);
        float a_dot_nab = dot(sv[a].w, n_ab);
        if (a_dot_nab < 0) {
            float ba_l2 = len2(ba);
            float a_dot_ba = dot(sv[a].w, ba);
            float b_dot_ba = dot(sv[b].w, ba);
            if (a_dot_ba > 0) dist = len(sv[a].w);
            else if (b_dot_ba < 0) dist = len(sv[b].w);
            else {
                float a_dot_b = dot(sv[a].w, sv[b].w);
                dist = sqrtf(fmaxf((len2(sv[a].w) * len2(sv[b].w) - a_dot_b * a_dot_b) / ba_l2, 0.f));
            }
            return true;
        }
        return false;
    }
    RSG_FN int newface(int a, int b, int c, bool forced) {
        if (stock_root >= 0) {
            int face = stock_root;
            list_remove(stock_root, stock_count, face);
            list_append(hull_root, hull_count, face);
            EpaFace& F = fc[face];
            F.pass = 0;
            F.c[0] = (int16_t)a; F.c[1] = (int16_t)b; F.c[2] = (int16_t)c;
            F.n = cross(sv[b].w - sv[a].w, sv[c].w - sv[a].w);
            float l = len(F.n);
            bool v = l > EPA_ACCURACY;
            if (v) {
                if (!(edgedist(face, a, b, F.d) || edgedist(face, b, c, F.d) || edgedist(face, c, a, F.d)))
                    F.d = dot(sv[a].w, F.n) / l;
                F.n = F.n * (1.f / l);
                if (forced || F.d >= -EPA_PLANE_EPS) return face;
                status = 3;
            } else {
                status = 2;
            }
            list_remove(hull_root, hull_count, face);
            list_append(stock_root, stock_count, face);
            return -1;
        }
        status = stock_root >= 0 ? 6 : 5;
        return -1;
    }
    RSG_FN int findbest() const {
        int minf = hull_root;
        float mind = fc[minf].d * fc[minf].d;
        for (int f = fc[minf].l[1]; f >= 0; f = fc[f].l[1]) {
            float sqd = fc[f].d * fc[f].d;
            if (sqd < mind) { minf = f; mind = sqd; }
        }
        return minf;
    }

    struct Horizon { int cf, ff, nf; };

    // expand(), iterative form of the reference's recursion
    RSG_FN bool expand(int pass, int w, int f0, int e0, Horizon& hz) {
        const int i1m3[3] = {1, 2, 0};
        const int i2m3[3] = {2, 0, 1};
        struct Frame { int16_t f; uint8_t e, phase; };
        Frame stk[EPA_STACK];
        int sp = 0;
        stk[0].f = (int16_t)f0; stk[0].e = (uint8_t)e0; stk[0].phase = 0;
        bool ret = false;
        while (sp >= 0) {
            Frame& fr = stk[sp];
            int f = fr.f;
            if (fr.phase == 0) {
                if (fc[f].pass == pass) { ret = false; sp--; continue; }
                int e1 = i1m3[fr.e];
                if ((dot(fc[f].n, sv[w].w) - fc[f].d) < -EPA_PLANE_EPS) {
                    int nf = newface(fc[f].c[e1], fc[f].c[fr.e], w, false);
                    if (nf >= 0) {
                        bind(nf, 0, f, fr.e);
                        if (hz.cf >= 0) bind(hz.cf, 1, nf, 2);
                        else hz.ff = nf;
                        hz.cf = nf;
                        hz.nf++;
                        ret = true;
                    } else {
                        ret = false;
                    }
                    sp--;
                    continue;
                }
                fc[f].pass = (uint8_t)pass;
                fr.phase = 1;
                if (sp + 1 >= EPA_STACK) { status = 4; return false; }
                Frame& ch = stk[++sp];
                ch.f = fc[f].f[e1]; ch.e = fc[f].e[e1]; ch.phase = 0;
            } else if (fr.phase == 1) {
                if (!ret) { sp--; continue; }
                int e2 = i2m3[fr.e];
                fr.phase = 2;
                if (sp + 1 >= EPA_STACK) { status = 4; return false; }
                Frame& ch = stk[++sp];
                ch.f = fc[f].f[e2]; ch.e = fc[f].e[e2]; ch.phase = 0;
            } else {
                if (ret) {
                    list_remove(hull_root, hull_count, f);
                    list_append(stock_root, stock_count, f);
                }
                sp--;
            }
        }
        return ret;
    }

    RSG_FN int evaluate(Gjk2& gjk, V3 guess) {
        G2Simplex& simplex = *gjk.simplex;
        if (simplex.rank > 1 && gjk.enclose_origin()) {
            while (hull_root >= 0) {
                int f = hull_root;
                list_remove(hull_root, hull_count, f);
                list_append(stock_root, stock_count, f);
            }
            status = 0;
            nextsv = 0;
            if (Gjk2::det3(sv[simplex.c[0]].w - sv[simplex.c[3]].w, sv[simplex.c[1]].w - sv[simplex.c[3]].w,
                           sv[simplex.c[2]].w - sv[simplex.c[3]].w) < 0) {
                int16_t tc = simplex.c[0]; simplex.c[0] = simplex.c[1]; simplex.c[1] = tc;
                float tp = simplex.p[0]; simplex.p[0] = simplex.p[1]; simplex.p[1] = tp;
            }
            int tetra[4] = {newface(simplex.c[0], simplex.c[1], simplex.c[2], true),
                            newface(simplex.c[1], simplex.c[0], simplex.c[3], true),
                            newface(simplex.c[2], simplex.c[1], simplex.c[3], true),
                            newface(simplex.c[0], simplex.c[2], simplex.c[3], true)};
            if (hull_count == 4) {
                int best = findbest();
                EpaFace outer = fc[best];
                int pass = 0;
                int iterations = 0;
                bind(tetra[0], 0, tetra[1], 0);
                bind(tetra[0], 1, tetra[2], 0);
                bind(tetra[0], 2, tetra[3], 0);
                bind(tetra[1], 1, tetra[3], 2);
                bind(tetra[1], 2, tetra[2], 1);
                bind(tetra[2], 2, tetra[3], 1);
                status = 0;
                for (; iterations < EPA_MAX_ITER; iterations++) {
                    if (nextsv < EPA_MAX_VERTICES) {
                        Horizon hz{-1, -1, 0};
                        int w = 4 + nextsv++;
                        bool valid = true;
                        fc[best].pass = (uint8_t)(++pass);
                        gjk.getsupport(fc[best].n, sv[w]);
                        float wdist = dot(fc[best].n, sv[w].w) - fc[best].d;
                        if (wdist > EPA_ACCURACY) {
                            for (int j = 0; j < 3 && valid; j++)
                                valid &= expand(pass, w, fc[best].f[j], fc[best].e[j], hz);
                            if (valid && hz.nf >= 3) {
                                bind(hz.cf, 1, hz.ff, 2);
                                list_remove(hull_root, hull_count, best);
                                list_append(stock_root, stock_count, best);
                                best = findbest();
                                outer = fc[best];
                            } else {
                                status = 4;
                                break;
                            }
                        } else {
                            status = 7;
                            break;
                        }
                    } else {
                        status = 6;
                        break;
                    }
                }
                V3 projection = outer.n * outer.d;
                normal = outer.n;
                depth = outer.d;
                result.rank = 3;
                result.c[0] = outer.c[0]; result.c[1] = outer.c[1]; result.c[2] = outer.c[2];
                result.p[0] = len(cross(sv[outer.c[1]].w - projection, sv[outer.c[2]].w - projection));
                result.p[1] = len(cross(sv[outer.c[2]].w - projection, sv[outer.c[0]].w - projection));
                result.p[2] = len(cross(sv[outer.c[0]].w - projection, sv[outer.c[1]].w - projection));
                float sum = result.p[0] + result.p[1] + result.p[2];
                result.p[0] /= sum;
                result.p[1] /= sum;
                result.p[2] /= sum;
                return status;
            }
        }
        status = 8;
        normal = -guess;
        float nl = len(normal);
        if (nl > 0) normal = normal * (1.f / nl);
        else normal = v3(1, 0, 0);
        depth = 0;
        result.rank = 1;
        result.c[0] = simplex.c[0];
        result.p[0] = 1;
        return status;
    }
};

struct G2Results {
    V3 witness[2];
    V3 normal;
    float distance;
};

RSG_FN void g2_init_shape(MinkDiff& md, const CShape& s0, const Xf& t0, const CShape& s1, const Xf& t1, bool margins) {
    md.s0 = &s0;
    md.s1 = &s1;
    md.to1 = transpose(t1.R) * t0.R;
    // wtrs0.inverseTimes(wtrs1)
    md.to0.R = transpose(t0.R) * t1.R;
    md.to0.t = mul_vm(t1.t - t0.t, t0.R);
    md.margins = margins;
}

// btGjkEpaSolver2::Penetration (with margins)
RSG_FN bool g2_penetration(const CShape& s0, const Xf& t0, const CShape& s1, const Xf& t1, V3 guess,
                           G2Results& res, SV* pool, EpaFace* faces, Epa& epa) {
    Gjk2 gjk;
    g2_init_shape(gjk.shape, s0, t0, s1, t1, true);
    gjk.sv = pool;
    res.witness[0] = res.witness[1] = v3zero();
    int gs = gjk.evaluate(-guess);
    if (gs == 1) {
        epa.init(pool, faces);  // EPA vertices follow the four GJK entries of the pool
        int es = epa.evaluate(gjk, -guess);
        if (es != 9) {
            V3 w0 = v3zero();
            for (int i = 0; i < epa.result.rank; i++)
                w0 += gjk.shape.support(pool[epa.result.c[i]].d, 0) * epa.result.p[i];
            res.witness[0] = xf_apply(t0, w0);
            res.witness[1] = xf_apply(t0, w0 - epa.normal * epa.depth);
            res.normal = -epa.normal;
            res.distance = -epa.depth;
            return true;
        }
    }
    return false;
}

// btGjkEpaSolver2::Distance (without margins)
RSG_FN bool g2_distance(const CShape& s0, const Xf& t0, const CShape& s1, const Xf& t1, V3 guess,
                        G2Results& res, SV* pool) {
    Gjk2 gjk;
    g2_init_shape(gjk.shape, s0, t0, s1, t1, false);
    gjk.sv = pool;
    res.witness[0] = res.witness[1] = v3zero();
    int gs = gjk.evaluate(guess);
    if (gs == 0) {
        V3 w0 = v3zero(), w1 = v3zero();
        for (int i = 0; i < gjk.simplex->rank; i++) {
            float p = gjk.simplex->p[i];
            w0 += gjk.shape.support(pool[gjk.simplex->c[i]].d, 0) * p;
            w1 += gjk.shape.support(-pool[gjk.simplex->c[i]].d, 1) * p;
        }
        res.witness[0] = xf_apply(t0, w0);
        res.witness[1] = xf_apply(t0, w1);
        res.normal = w0 - w1;
        res.distance = len(res.normal);
        res.normal = res.normal * (1.f / (res.distance > G2_MIN_DISTANCE ? res.distance : 1.f));
        return true;
    }
    return false;
}

// Work space of a penetration query: the GJK/EPA vertex pool and EPA's face
// pool.  The caller places it (private memory, or a per-arena block in global
// memory that a lane reads without the per-lane interleaving of private
// memory).
struct EpaSpace {
    SV pool[SV_POOL];
    EpaFace fc[EPA_MAX_FACES];
};

// btGjkEpaPenetrationDepthSolver::calcPenDepth
RSG_FN bool calc_pen_depth(const CShape& a, const CShape& b, const Xf& ta, const Xf& tb, V3& v, V3& wa, V3& wb,
                           EpaSpace& space) {
    SV* pool = space.pool;
    Epa epa;
    V3 guesses[9] = {safe_normalized(tb.t - ta.t), safe_normalized(ta.t - tb.t), v3(0, 0, 1), v3(0, 1, 0),
                     v3(1, 0, 0), v3(1, 1, 0), v3(1, 1, 1), v3(0, 1, 1), v3(1, 0, 1)};
    for (int i = 0; i < 9; i++) {
        G2Results r;
        if (g2_penetration(a, ta, b, tb, guesses[i], r, pool, space.fc, epa)) {
            wa = r.witness[0]; wb = r.witness[1]; v = r.normal;
            return true;
        }
        if (g2_distance(a, ta, b, tb, guesses[i], r, pool)) {
            wa = r.witness[0]; wb = r.witness[1]; v = r.normal;
            return false;
        }
    }
    wa = v3zero(); wb = v3zero(); v = v3zero();
    return false;
}

// ---------------------------------------------------------------------------
// btGjkPairDetector::getClosestPointsNonVirtual.  Returns true with the
// reported contact (normal on B pointing to A, point on B in world space,
// signed distance) when the shapes are closer than sqrt(max_dist2).

RSG_FN bool gjk_closest(const CShape& A, const CShape& B, const Xf& ta, const Xf& tb, float max_dist2,
                        V3& normal_out, V3& point_b_out, float& dist_out, EpaSpace& space) {
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
        V3 tA, tB;
        sep = v3zero();
        bool ok2 = calc_pen_depth(A, B, lta, ltb, sep, tA, tB, space);
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
        return true;
    }
    return false;
}

}  // namespace rsg
