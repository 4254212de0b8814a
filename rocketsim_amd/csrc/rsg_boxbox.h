// rsg_boxbox.h -- contacts between two oriented boxes (car hitboxes), giving
// the same points, normal and depths as the reference's car-car narrowphase
// (btBoxBoxDetector, BulletCollision/CollisionDispatch/btBoxBoxDetector.cpp:
// 272-767, itself ODE's dBoxBox).  Method (Gottschalk's separating-axis test
// for OBBs, then Sutherland-Hodgman clipping):
//   1. Fifteen candidate axes: the three face normals of each box and the nine
//      edge-edge cross products.  The axis of least penetration wins; an
//      edge axis must beat a face axis by 5 % to be chosen.  Any separating
//      axis ends the test.
//   2. An edge-edge axis gives one point: the closest approach of the two
//      edges.  A face axis makes that face the reference face: the most
//      anti-parallel face of the other box (the incident face) is clipped
//      against the reference face's rectangle, and the clipped points below
//      the reference face are kept; more than four are reduced to four
//      spread by angle around the polygon's centroid, starting with the
//      deepest.
// Every sum, product and comparison is evaluated in the reference's order so
// that the points come out bit for bit (tests/test_components.py).
#pragma once

#include "rsg_contact.h"

namespace rsg {

struct BoxBoxContacts {
    V3 normal;        // normal on box 1 (pointing towards box 0)
    int n;
    V3 point[4];      // points on box 1
    float depth[4];   // signed distance (negative = penetration)
};

// One box: centre, the three axes (columns of its basis) and half extents.
struct BoxFrame {
    V3 c;
    V3 ax[3];
    float h[3];
};

RSG_FN BoxFrame box_frame(const Xf& x, V3 half) {
    BoxFrame b;
    b.c = x.t;
    for (int k = 0; k < 3; k++) b.ax[k] = v3(get(x.R.r[0], k), get(x.R.r[1], k), get(x.R.r[2], k));
    b.h[0] = half.x; b.h[1] = half.y; b.h[2] = half.z;
    return b;
}

// Parameters of the points of closest approach of the lines pa + s ua and
// pb + t ub (unit directions); both 0 for near-parallel lines.
RSG_FN void closest_line_params(V3 pa, V3 ua, V3 pb, V3 ub, float& s, float& t) {
    V3 d = pb - pa;
    float cos_ab = dot(ua, ub);
    float da = dot(ua, d);
    float db = -dot(ub, d);
    float den = 1 - cos_ab * cos_ab;
    if (den <= 0.0001f) {
        s = 0;
        t = 0;
        return;
    }
    den = 1.f / den;
    s = (da + cos_ab * db) * den;
    t = (cos_ab * da + db) * den;
}

// Sutherland-Hodgman: clips the polygon `poly` (n 2D points, x y pairs)
// against the rectangle |x| < half[0], |y| < half[1], one half-plane at a
// time (x >= -hx, x <= hx, y >= -hy, y <= hy).  At most 8 points are kept: a
// clip step that would emit a ninth stops the whole clip with its first 8.
// Returns the point count; the polygon is left in `poly`.
RSG_FN int clip_to_rect(const float half[2], float poly[16], int n) {
    float next[16];
    for (int axis = 0; axis < 2; axis++) {
        for (int side = -1; side <= 1; side += 2) {
            const float bound = half[axis];
            int m = 0;
            bool full = false;
            for (int i = 0; i < n && !full; i++) {
                const float* a = poly + 2 * i;
                const float* b = poly + 2 * (i + 1 < n ? i + 1 : 0);
                const bool a_in = side * a[axis] < bound;
                const bool b_in = side * b[axis] < bound;
                if (a_in) {
                    next[2 * m] = a[0];
                    next[2 * m + 1] = a[1];
                    full = ++m == 8;
                }
                if (!full && a_in != b_in) {
                    const int o = 1 - axis;
                    next[2 * m + o] = a[o] + (b[o] - a[o]) / (b[axis] - a[axis]) * (side * bound - a[axis]);
                    next[2 * m + axis] = side * bound;
                    full = ++m == 8;
                }
            }
            for (int k = 0; k < 2 * m; k++) poly[k] = next[k];
            n = m;
            if (full) return n;
        }
    }
    return n;
}

// Chooses `keep` of the n polygon points: `first`, then for each of the keep-1
// directions spaced evenly in angle from it around the polygon's centroid,
// the unused point whose direction is nearest (first one on ties).
RSG_FN void spread_by_angle(int n, const float pts[16], int keep, int first, int out[4]) {
    const float pi = 3.14159265f;
    float cx, cy;
    if (n == 1) {
        cx = pts[0];
        cy = pts[1];
    } else if (n == 2) {
        cx = 0.5f * (pts[0] + pts[2]);
        cy = 0.5f * (pts[1] + pts[3]);
    } else {
        // area-weighted centroid (shoelace), closing edge last
        float area2 = 0, sx = 0, sy = 0, w = 0;
        for (int i = 0; i < n; i++) {
            const int j = i + 1 < n ? i + 1 : 0;
            w = pts[2 * i] * pts[2 * j + 1] - pts[2 * j] * pts[2 * i + 1];
            if (j == 0) break;
            area2 += w;
            sx += w * (pts[2 * i] + pts[2 * j]);
            sy += w * (pts[2 * i + 1] + pts[2 * j + 1]);
        }
        float inv = fabsf(area2 + w) > FLT_EPS ? 1.f / (3.0f * (area2 + w)) : 1e18f;
        cx = inv * (sx + w * (pts[2 * n - 2] + pts[0]));
        cy = inv * (sy + w * (pts[2 * n - 1] + pts[1]));
    }
    float ang[8];
    bool used[8];
    for (int i = 0; i < n; i++) {
        ang[i] = atan2f(pts[2 * i + 1] - cy, pts[2 * i] - cx);
        used[i] = false;
    }
    used[first] = true;
    out[0] = first;
    for (int k = 1; k < keep; k++) {
        float target = (float)k * (2 * pi / keep) + ang[first];
        if (target > pi) target -= 2 * pi;
        float best = 1e9f;
        int pick = first;
        for (int i = 0; i < n; i++) {
            if (used[i]) continue;
            float d = fabsf(ang[i] - target);
            if (d > pi) d = 2 * pi - d;
            if (d < best) {
                best = d;
                pick = i;
            }
        }
        used[pick] = true;
        out[k] = pick;
    }
}

// Box 0 (transform x0, half extents h0) against box 1.  Contacts are
// reported as the manifold receives them: normal on box 1, point on box 1,
// signed distance.
RSG_FN void box_box(const Xf& x0, V3 h0, const Xf& x1, V3 h1, BoxBoxContacts& out) {
    out.n = 0;
    const BoxFrame b0 = box_frame(x0, h0), b1 = box_frame(x1, h1);
    const V3 d = b1.c - b0.c;
    // centre offset in box 0's frame; axis cosines and their magnitudes
    float off[3], cs[3][3], ac[3][3];
    for (int i = 0; i < 3; i++) off[i] = dot(b0.ax[i], d);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cs[i][j] = dot(b0.ax[i], b1.ax[j]);
            ac[i][j] = fabsf(cs[i][j]);
        }
    const float* A = b0.h;
    const float* B = b1.h;

    // 1. separating-axis search.  `best` is the largest (least negative)
    // separation, `axis_id` 1-3 box 0 faces, 4-6 box 1 faces, 7-15 edge pairs.
    float best = -3.402823466e+38f;
    int axis_id = 0;
    bool flip = false;
    V3 edge_n = v3zero();  // chosen edge axis in box 0's frame (unit)
    for (int k = 0; k < 3; k++) {
        const float dist = off[k];
        const float reach = A[k] + B[0] * ac[k][0] + B[1] * ac[k][1] + B[2] * ac[k][2];
        const float sep = fabsf(dist) - reach;
        if (sep > 0) return;
        if (sep > best) { best = sep; axis_id = 1 + k; flip = dist < 0; }
    }
    for (int k = 0; k < 3; k++) {
        const float dist = dot(b1.ax[k], d);
        const float reach = A[0] * ac[0][k] + A[1] * ac[1][k] + A[2] * ac[2][k] + B[k];
        const float sep = fabsf(dist) - reach;
        if (sep > 0) return;
        if (sep > best) { best = sep; axis_id = 4 + k; flip = dist < 0; }
    }
    // edge axes: box 0 axis i x box 1 axis j, measured in box 0's frame; the
    // cosine magnitudes get a small floor against near-parallel edges
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ac[i][j] += 1.0e-5f;
    for (int i = 0; i < 3; i++) {
        const int u = (i + 1) % 3, w = (i + 2) % 3;         // cyclic successors
        const int lo = u < w ? u : w, hi = u < w ? w : u;   // the other two, ascending
        for (int j = 0; j < 3; j++) {
            const int jlo = j == 0 ? 1 : 0, jhi = j == 2 ? 1 : 2;
            const float dist = off[w] * cs[u][j] - off[u] * cs[w][j];
            const float reach = A[lo] * ac[hi][j] + A[hi] * ac[lo][j] + B[jlo] * ac[i][jhi] + B[jhi] * ac[i][jlo];
            float n3[3];
            n3[i] = 0;
            n3[u] = -cs[w][j];
            n3[w] = cs[u][j];
            float sep = fabsf(dist) - reach;
            if (sep > FLT_EPS) return;
            const float len = sqrtf(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
            if (len > FLT_EPS) {
                sep /= len;
                if (sep * 1.05f > best) {
                    best = sep;
                    axis_id = 7 + 3 * i + j;
                    edge_n = v3(n3[0] / len, n3[1] / len, n3[2] / len);
                    flip = dist < 0;
                }
            }
        }
    }
    if (!axis_id) return;
    V3 normal;  // world normal pointing from box 0 to box 1
    if (axis_id <= 3) normal = b0.ax[axis_id - 1];
    else if (axis_id <= 6) normal = b1.ax[axis_id - 4];
    else normal = x0.R * edge_n;  // back from box 0's frame
    if (flip) normal = -normal;
    const float pen = -best;
    out.normal = -normal;

    // 2a. edge-edge: closest approach of the two edges nearest each other
    if (axis_id > 6) {
        V3 ea = b0.c, eb = b1.c;
        for (int k = 0; k < 3; k++) ea += (dot(normal, b0.ax[k]) > 0 ? 1.0f : -1.0f) * A[k] * b0.ax[k];
        for (int k = 0; k < 3; k++) eb += (dot(normal, b1.ax[k]) > 0 ? -1.0f : 1.0f) * B[k] * b1.ax[k];
        const V3 ua = b0.ax[(axis_id - 7) / 3], ub = b1.ax[(axis_id - 7) % 3];
        float s, t;
        closest_line_params(ea, ua, eb, ub, s, t);
        out.n = 1;
        out.point[0] = eb + ub * t;
        out.depth[0] = -pen;
        return;
    }

    // 2b. face contact: `ref` holds the reference face, `inc` the incident one
    const bool ref_is_0 = axis_id <= 3;
    const BoxFrame& ref = ref_is_0 ? b0 : b1;
    const BoxFrame& inc = ref_is_0 ? b1 : b0;
    const V3 ref_n = ref_is_0 ? normal : -normal;  // reference face normal, towards the incident box
    const int face = ref_is_0 ? axis_id - 1 : axis_id - 4;
    const int fu = face == 0 ? 1 : 0, fv = face == 2 ? 1 : 2;  // the face's in-plane axes
    // the incident face: the incident box's axis most parallel to ref_n
    float cn[3], acn[3];
    for (int k = 0; k < 3; k++) {
        cn[k] = dot(inc.ax[k], ref_n);
        acn[k] = fabsf(cn[k]);
    }
    int up, ia, ib;  // incident face axis, then its in-plane axes
    if (acn[1] > acn[0]) {
        if (acn[1] > acn[2]) { up = 1; ia = 0; ib = 2; }
        else { up = 2; ia = 0; ib = 1; }
    } else {
        if (acn[0] > acn[2]) { up = 0; ia = 1; ib = 2; }
        else { up = 2; ia = 0; ib = 1; }
    }
    const V3 rel = inc.c - ref.c;
    const V3 face_c = cn[up] < 0 ? rel + inc.h[up] * inc.ax[up] : rel - inc.h[up] * inc.ax[up];
    // incident face corners in the reference face's 2D coordinates
    const float fc_u = dot(face_c, ref.ax[fu]), fc_v = dot(face_c, ref.ax[fv]);
    float m_ua = dot(ref.ax[fu], inc.ax[ia]), m_ub = dot(ref.ax[fu], inc.ax[ib]);
    float m_va = dot(ref.ax[fv], inc.ax[ia]), m_vb = dot(ref.ax[fv], inc.ax[ib]);
    float poly[16];
    {
        const float ua = m_ua * inc.h[ia], va = m_va * inc.h[ia], ub = m_ub * inc.h[ib], vb = m_vb * inc.h[ib];
        poly[0] = fc_u - ua - ub; poly[1] = fc_v - va - vb;
        poly[2] = fc_u - ua + ub; poly[3] = fc_v - va + vb;
        poly[4] = fc_u + ua + ub; poly[5] = fc_v + va + vb;
        poly[6] = fc_u + ua - ub; poly[7] = fc_v + va - vb;
    }
    const float rect[2] = {ref.h[fu], ref.h[fv]};
    int n = clip_to_rect(rect, poly, 4);
    if (n < 1) return;
    // back to 3D (relative to the reference box's centre): invert the 2x2 map
    const float inv_det = 1.f / (m_ua * m_vb - m_ub * m_va);
    m_ua *= inv_det; m_ub *= inv_det; m_va *= inv_det; m_vb *= inv_det;
    V3 pts[8];
    float depth[8];
    int kept = 0;
    for (int k = 0; k < n; k++) {
        const float du = poly[2 * k] - fc_u, dv = poly[2 * k + 1] - fc_v;
        const float sa = m_vb * du - m_ub * dv;
        const float sb = -m_va * du + m_ua * dv;
        pts[kept] = face_c + sa * inc.ax[ia] + sb * inc.ax[ib];
        depth[kept] = ref.h[face] - dot(ref_n, pts[kept]);
        if (depth[kept] >= 0) {
            poly[2 * kept] = poly[2 * k];
            poly[2 * kept + 1] = poly[2 * k + 1];
            kept++;
        }
    }
    if (kept < 1) return;
    int pick[4];
    int count = kept;
    if (kept > 4) {
        int deepest = 0;
        for (int k = 1; k < kept; k++)
            if (depth[k] > depth[deepest]) deepest = k;
        spread_by_angle(kept, poly, 4, deepest, pick);
        count = 4;
    } else {
        for (int k = 0; k < kept; k++) pick[k] = k;
    }
    // points on box 1: as found when box 1 is the incident box, moved down to
    // the reference face when box 1 holds it
    for (int k = 0; k < count; k++) {
        V3 p = pts[pick[k]] + ref.c;
        if (!ref_is_0) p = p - normal * depth[pick[k]];
        out.point[k] = p;
        out.depth[k] = -depth[pick[k]];
    }
    out.n = count;
}

}  // namespace rsg
