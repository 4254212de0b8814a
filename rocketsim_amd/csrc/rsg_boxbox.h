// rsg_boxbox.h -- box vs box contacts for car hitboxes: the separating-axis
// box-box test with face clipping and contact culling that the reference uses
// for car-car pairs (btBoxBoxDetector / dBoxBox2, originally from the Open
// Dynamics Engine; BulletCollision/CollisionDispatch/btBoxBoxDetector.cpp:
// 27-767).  Fifteen axes (3 + 3 face normals, 9 edge crosses with a 5%
// preference for face axes), then either one edge-edge point or the incident
// face clipped against the reference face rectangle, at most four points
// chosen by angle around their centroid.
#pragma once

#include "rsg_contact.h"

namespace rsg {

struct BoxBoxContacts {
    V3 normal;        // normal on box 1 (pointing towards box 0)
    int n;
    V3 point[4];      // points on box 1
    float depth[4];   // signed distance (negative = penetration)
};

// Column k of a basis (the reference reads its 4-stride matrices by column).
RSG_FN V3 bcol(const M3& m, int k) { return v3(get(m.r[0], k), get(m.r[1], k), get(m.r[2], k)); }

// dLineClosestApproach
RSG_FN void line_closest_approach(V3 pa, V3 ua, V3 pb, V3 ub, float& alpha, float& beta) {
    V3 p = pb - pa;
    float uaub = dot(ua, ub);
    float q1 = dot(ua, p);
    float q2 = -dot(ub, p);
    float d = 1 - uaub * uaub;
    if (d <= 0.0001f) {
        alpha = 0;
        beta = 0;
    } else {
        d = 1.f / d;
        alpha = (q1 + uaub * q2) * d;
        beta = (uaub * q1 + q2) * d;
    }
}

// intersectRectQuad2: clip quad q (4 2D points) against the rectangle
// [-h0,h0] x [-h1,h1]; returns the polygon in ret (up to 8 points).
RSG_FN int intersect_rect_quad(const float h[2], const float p[8], float ret[16]) {
    int nq = 4, nr = 0;
    float buffer[16];
    float qbuf[16];
    for (int i = 0; i < 8; i++) qbuf[i] = p[i];
    // q / r alternate between (qbuf|ret|buffer); track them as pointers
    float* q = qbuf;
    float* r = ret;
    for (int dir = 0; dir <= 1; dir++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            float* pq = q;
            float* pr = r;
            nr = 0;
            for (int i = nq; i > 0; i--) {
                if (sign * pq[dir] < h[dir]) {
                    pr[0] = pq[0];
                    pr[1] = pq[1];
                    pr += 2;
                    nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                float* nextq = (i > 1) ? pq + 2 : q;
                if ((sign * pq[dir] < h[dir]) ^ (sign * nextq[dir] < h[dir])) {
                    pr[1 - dir] = pq[1 - dir] + (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) * (sign * h[dir] - pq[dir]);
                    pr[dir] = sign * h[dir];
                    pr += 2;
                    nr++;
                    if (nr & 8) { q = r; goto done; }
                }
                pq += 2;
            }
            q = r;
            r = (q == ret) ? buffer : ret;
            nq = nr;
        }
    }
done:
    if (q != ret)
        for (int i = 0; i < nr * 2; i++) ret[i] = q[i];
    return nr;
}

// cullPoints2: choose m of the n polygon points, starting with i0, spread by
// angle around the polygon's centroid.
RSG_FN void cull_points(int n, const float p[], int m, int i0, int iret[]) {
    const float PI_ODE = 3.14159265f;
    float a, cx, cy, q;
    if (n == 1) {
        cx = p[0];
        cy = p[1];
    } else if (n == 2) {
        cx = 0.5f * (p[0] + p[2]);
        cy = 0.5f * (p[1] + p[3]);
    } else {
        a = 0; cx = 0; cy = 0;
        for (int i = 0; i < n - 1; i++) {
            q = p[i * 2] * p[i * 2 + 3] - p[i * 2 + 2] * p[i * 2 + 1];
            a += q;
            cx += q * (p[i * 2] + p[i * 2 + 2]);
            cy += q * (p[i * 2 + 1] + p[i * 2 + 3]);
        }
        q = p[n * 2 - 2] * p[1] - p[0] * p[n * 2 - 1];
        if (fabsf(a + q) > FLT_EPS) a = 1.f / (3.0f * (a + q));
        else a = 1e18f;
        cx = a * (cx + q * (p[n * 2 - 2] + p[0]));
        cy = a * (cy + q * (p[n * 2 - 1] + p[1]));
    }
    float ang[8];
    for (int i = 0; i < n; i++) ang[i] = atan2f(p[i * 2 + 1] - cy, p[i * 2] - cx);
    int avail[8];
    for (int i = 0; i < n; i++) avail[i] = 1;
    avail[i0] = 0;
    iret[0] = i0;
    int k = 1;
    for (int j = 1; j < m; j++) {
        a = (float)j * (2 * PI_ODE / m) + ang[i0];
        if (a > PI_ODE) a -= 2 * PI_ODE;
        float maxdiff = 1e9f, diff;
        iret[k] = i0;
        for (int i = 0; i < n; i++) {
            if (avail[i]) {
                diff = fabsf(ang[i] - a);
                if (diff > PI_ODE) diff = 2 * PI_ODE - diff;
                if (diff < maxdiff) { maxdiff = diff; iret[k] = i; }
            }
        }
        avail[iret[k]] = 0;
        k++;
    }
}

// Box 0 (centre x0.t, basis x0.R, half extents h0) against box 1.  Contacts
// are reported as the reference reports them to the manifold: normal on
// box 1, point on box 1, negative depth.
RSG_FN void box_box(const Xf& x0, V3 h0, const Xf& x1, V3 h1, BoxBoxContacts& out) {
    out.n = 0;
    const float fudge_factor = 1.05f;
    const M3& R1 = x0.R;
    const M3& R2 = x1.R;
    V3 p1 = x0.t, p2 = x1.t;
    V3 p = p2 - p1;
    V3 pp = v3(dot(bcol(R1, 0), p), dot(bcol(R1, 1), p), dot(bcol(R1, 2), p));
    float A[3] = {h0.x, h0.y, h0.z};
    float B[3] = {h1.x, h1.y, h1.z};
    float R[3][3], Q[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            R[i][j] = dot(bcol(R1, i), bcol(R2, j));
            Q[i][j] = fabsf(R[i][j]);
        }
    float s = -3.402823466e+38f, s2, l;
    int norm_col = -1;   // >= 0: face normal = column of R1 (0..2) or R2 (3..5)
    V3 normalC = v3zero();
    bool invert_normal = false;
    int code = 0;

    // face axes
    float e1v[6], e2v[6];
    e1v[0] = pp.x; e2v[0] = A[0] + B[0] * Q[0][0] + B[1] * Q[0][1] + B[2] * Q[0][2];
    e1v[1] = pp.y; e2v[1] = A[1] + B[0] * Q[1][0] + B[1] * Q[1][1] + B[2] * Q[1][2];
    e1v[2] = pp.z; e2v[2] = A[2] + B[0] * Q[2][0] + B[1] * Q[2][1] + B[2] * Q[2][2];
    e1v[3] = dot(bcol(R2, 0), p); e2v[3] = A[0] * Q[0][0] + A[1] * Q[1][0] + A[2] * Q[2][0] + B[0];
    e1v[4] = dot(bcol(R2, 1), p); e2v[4] = A[0] * Q[0][1] + A[1] * Q[1][1] + A[2] * Q[2][1] + B[1];
    e1v[5] = dot(bcol(R2, 2), p); e2v[5] = A[0] * Q[0][2] + A[1] * Q[1][2] + A[2] * Q[2][2] + B[2];
    for (int k = 0; k < 6; k++) {
        s2 = fabsf(e1v[k]) - e2v[k];
        if (s2 > 0) return;
        if (s2 > s) {
            s = s2;
            norm_col = k;
            invert_normal = e1v[k] < 0;
            code = k + 1;
        }
    }
    // edge x edge axes
    const float fudge2 = 1.0e-5f;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Q[i][j] += fudge2;
    float R11 = R[0][0], R12 = R[0][1], R13 = R[0][2], R21 = R[1][0], R22 = R[1][1], R23 = R[1][2],
          R31 = R[2][0], R32 = R[2][1], R33 = R[2][2];
    float Q11 = Q[0][0], Q12 = Q[0][1], Q13 = Q[0][2], Q21 = Q[1][0], Q22 = Q[1][1], Q23 = Q[1][2],
          Q31 = Q[2][0], Q32 = Q[2][1], Q33 = Q[2][2];
    struct EdgeAxis { float e1, e2, n1, n2, n3; };
    EdgeAxis ax[9] = {
        {pp.z * R21 - pp.y * R31, (A[1] * Q31 + A[2] * Q21 + B[1] * Q13 + B[2] * Q12), 0, -R31, R21},
        {pp.z * R22 - pp.y * R32, (A[1] * Q32 + A[2] * Q22 + B[0] * Q13 + B[2] * Q11), 0, -R32, R22},
        {pp.z * R23 - pp.y * R33, (A[1] * Q33 + A[2] * Q23 + B[0] * Q12 + B[1] * Q11), 0, -R33, R23},
        {pp.x * R31 - pp.z * R11, (A[0] * Q31 + A[2] * Q11 + B[1] * Q23 + B[2] * Q22), R31, 0, -R11},
        {pp.x * R32 - pp.z * R12, (A[0] * Q32 + A[2] * Q12 + B[0] * Q23 + B[2] * Q21), R32, 0, -R12},
        {pp.x * R33 - pp.z * R13, (A[0] * Q33 + A[2] * Q13 + B[0] * Q22 + B[1] * Q21), R33, 0, -R13},
        {pp.y * R11 - pp.x * R21, (A[0] * Q21 + A[1] * Q11 + B[1] * Q33 + B[2] * Q32), -R21, R11, 0},
        {pp.y * R12 - pp.x * R22, (A[0] * Q22 + A[1] * Q12 + B[0] * Q33 + B[2] * Q31), -R22, R12, 0},
        {pp.y * R13 - pp.x * R23, (A[0] * Q23 + A[1] * Q13 + B[0] * Q32 + B[1] * Q31), -R23, R13, 0},
    };
    for (int k = 0; k < 9; k++) {
        const EdgeAxis& e = ax[k];
        s2 = fabsf(e.e1) - e.e2;
        if (s2 > FLT_EPS) return;
        l = sqrtf(e.n1 * e.n1 + e.n2 * e.n2 + e.n3 * e.n3);
        if (l > FLT_EPS) {
            s2 /= l;
            if (s2 * fudge_factor > s) {
                s = s2;
                norm_col = -1;
                normalC = v3(e.n1 / l, e.n2 / l, e.n3 / l);
                invert_normal = e.e1 < 0;
                code = 7 + k;
            }
        }
    }
    if (!code) return;
    V3 normal;
    if (norm_col >= 0) normal = norm_col < 3 ? bcol(R1, norm_col) : bcol(R2, norm_col - 3);
    else normal = R1 * normalC;
    if (invert_normal) normal = -normal;
    float depth = -s;

    if (code > 6) {
        // edge-edge
        V3 pa = p1;
        for (int j = 0; j < 3; j++) {
            float sign = dot(normal, bcol(R1, j)) > 0 ? 1.0f : -1.0f;
            pa += sign * A[j] * bcol(R1, j);
        }
        V3 pb = p2;
        for (int j = 0; j < 3; j++) {
            float sign = dot(normal, bcol(R2, j)) > 0 ? -1.0f : 1.0f;
            pb += sign * B[j] * bcol(R2, j);
        }
        V3 ua = bcol(R1, (code - 7) / 3);
        V3 ub = bcol(R2, (code - 7) % 3);
        float alpha, beta;
        line_closest_approach(pa, ua, pb, ub, alpha, beta);
        pb += ub * beta;
        out.normal = -normal;
        out.n = 1;
        out.point[0] = pb;
        out.depth[0] = -depth;
        return;
    }

    // face-something: box a holds the reference face
    const M3& Ra = code <= 3 ? R1 : R2;
    const M3& Rb = code <= 3 ? R2 : R1;
    V3 pa = code <= 3 ? p1 : p2;
    V3 pb = code <= 3 ? p2 : p1;
    const float* Sa = code <= 3 ? A : B;
    const float* Sb = code <= 3 ? B : A;
    V3 normal2 = code <= 3 ? normal : -normal;
    V3 nr = v3(dot(bcol(Rb, 0), normal2), dot(bcol(Rb, 1), normal2), dot(bcol(Rb, 2), normal2));
    V3 anr = v3(fabsf(nr.x), fabsf(nr.y), fabsf(nr.z));
    int lanr, a1, a2;
    if (anr.y > anr.x) {
        if (anr.y > anr.z) { a1 = 0; lanr = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
        if (anr.x > anr.z) { lanr = 0; a1 = 1; a2 = 2; }
        else { a1 = 0; a2 = 1; lanr = 2; }
    }
    V3 center;
    if (get(nr, lanr) < 0) center = (pb - pa) + Sb[lanr] * bcol(Rb, lanr);
    else center = (pb - pa) - Sb[lanr] * bcol(Rb, lanr);
    int codeN = code <= 3 ? code - 1 : code - 4;
    int code1, code2;
    if (codeN == 0) { code1 = 1; code2 = 2; }
    else if (codeN == 1) { code1 = 0; code2 = 2; }
    else { code1 = 0; code2 = 1; }
    float c1 = dot(center, bcol(Ra, code1));
    float c2 = dot(center, bcol(Ra, code2));
    float m11 = dot(bcol(Ra, code1), bcol(Rb, a1));
    float m12 = dot(bcol(Ra, code1), bcol(Rb, a2));
    float m21 = dot(bcol(Ra, code2), bcol(Rb, a1));
    float m22 = dot(bcol(Ra, code2), bcol(Rb, a2));
    float quad[8];
    {
        float k1 = m11 * Sb[a1], k2 = m21 * Sb[a1], k3 = m12 * Sb[a2], k4 = m22 * Sb[a2];
        quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
        quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
        quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
        quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    float rect[2] = {Sa[code1], Sa[code2]};
    float ret[16];
    int n = intersect_rect_quad(rect, quad, ret);
    if (n < 1) return;
    V3 point[8];
    float dep[8];
    float det1 = 1.f / (m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    int cnum = 0;
    for (int j = 0; j < n; j++) {
        float k1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
        float k2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
        point[cnum] = center + k1 * bcol(Rb, a1) + k2 * bcol(Rb, a2);
        dep[cnum] = Sa[codeN] - dot(normal2, point[cnum]);
        if (dep[cnum] >= 0) {
            ret[cnum * 2] = ret[j * 2];
            ret[cnum * 2 + 1] = ret[j * 2 + 1];
            cnum++;
        }
    }
    if (cnum < 1) return;
    int maxc = 4;
    if (maxc > cnum) maxc = cnum;
    if (maxc < 1) maxc = 1;
    out.normal = -normal;
    if (cnum <= maxc) {
        for (int j = 0; j < cnum; j++) {
            V3 w = point[j] + pa;
            if (code >= 4) w = w - normal * dep[j];
            out.point[out.n] = w;
            out.depth[out.n] = -dep[j];
            out.n++;
        }
    } else {
        int i1 = 0;
        float maxdepth = dep[0];
        for (int i = 1; i < cnum; i++)
            if (dep[i] > maxdepth) { maxdepth = dep[i]; i1 = i; }
        int iret[8];
        cull_points(cnum, ret, maxc, i1, iret);
        for (int j = 0; j < maxc; j++) {
            V3 w = point[iret[j]] + pa;
            if (code >= 4) w = w - normal * dep[iret[j]];
            out.point[out.n] = w;
            out.depth[out.n] = -dep[iret[j]];
            out.n++;
        }
    }
}

}  // namespace rsg
