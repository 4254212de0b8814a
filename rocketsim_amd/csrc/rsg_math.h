// rsg_math.h -- small fixed-size FP32 linear algebra for the tick kernels.
//
// Vectors are 3 floats, matrices are row-major 3x3 (rows r[0..2], the same
// convention as Bullet's btMatrix3x3, LinearMath/btMatrix3x3.h), quaternions
// are (x, y, z, w).  Where the reference's x86-64 build evaluates an
// expression in a particular association order (its SSE code paths,
// LinearMath/btScalar.h:211-235 turns them on), the same association is used
// here so that one tick rounds the way the reference does in the common case.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define RSG_FN __host__ __device__ __forceinline__
#else
#define RSG_FN inline
#endif

namespace rsg {

struct V3 {
    float x, y, z;
};

RSG_FN V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
RSG_FN V3 v3zero() { return v3(0.f, 0.f, 0.f); }
RSG_FN V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RSG_FN V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RSG_FN V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
RSG_FN V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RSG_FN V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
RSG_FN V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RSG_FN V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
RSG_FN V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
RSG_FN V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
RSG_FN V3& operator*=(V3& a, float s) { a = a * s; return a; }
RSG_FN bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

RSG_FN float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RSG_FN V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RSG_FN float len2(V3 a) { return dot(a, a); }
RSG_FN float len(V3 a) { return sqrtf(len2(a)); }
RSG_FN bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
RSG_FN float get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RSG_FN void set(V3& a, int i, float v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }

// v / |v| (Bullet normalize(): multiply by the reciprocal of the length).
// A host build may substitute its own rounding of this step (RSG_NORMALIZE_HOOK)
// to study how sensitive a trajectory is to the last bits of a normal.
#ifdef RSG_NORMALIZE_HOOK
RSG_FN V3 normalized(V3 a) { return RSG_NORMALIZE_HOOK(a); }
#else
RSG_FN V3 normalized(V3 a) { return a * (1.f / len(a)); }
#endif

// Bullet safeNormalize(): leave as (1,0,0) when the length is ~0
// (LinearMath/btVector3.h:287-302).
RSG_FN V3 safe_normalized(V3 a) {
    float l2 = len2(a);
    if (l2 >= 1.1920928955078125e-07f * 1.1920928955078125e-07f) return a * (1.f / sqrtf(l2));
    return v3(1.f, 0.f, 0.f);
}

// RocketSim Vec::Normalized(): zero vector when the length is tiny
// (src/Math/MathTypes/MathTypes.h:96-103).
RSG_FN V3 rs_normalized(V3 a) {
    float l2 = (a.x * a.x + a.y * a.y) + a.z * a.z;
    float l = l2 > 0.f ? sqrtf(l2) : 0.f;
    if (l > 1.1920928955078125e-07f * 1.1920928955078125e-07f) return a / l;
    return v3zero();
}

RSG_FN bool fuzzy_zero(V3 a) { return len2(a) < 1.1920928955078125e-07f; }

struct M3 {
    V3 r[3];
};

RSG_FN M3 m3(V3 r0, V3 r1, V3 r2) { M3 m; m.r[0] = r0; m.r[1] = r1; m.r[2] = r2; return m; }
RSG_FN M3 m3_identity() { return m3(v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 1)); }
RSG_FN V3 col(const M3& m, int c) { return v3(get(m.r[0], c), get(m.r[1], c), get(m.r[2], c)); }
RSG_FN V3 operator*(const M3& m, V3 v) { return v3(dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)); }
// v * M = M^T v, evaluated column-wise as the reference's SSE path does.
RSG_FN V3 mul_vm(V3 v, const M3& m) {
    return v3((v.x * m.r[0].x + v.y * m.r[1].x) + v.z * m.r[2].x,
              (v.x * m.r[0].y + v.y * m.r[1].y) + v.z * m.r[2].y,
              (v.x * m.r[0].z + v.y * m.r[1].z) + v.z * m.r[2].z);
}
RSG_FN M3 operator*(const M3& a, const M3& b) {
    M3 o;
    for (int i = 0; i < 3; i++) {
        V3 ai = a.r[i];
        o.r[i] = (b.r[0] * ai.x + b.r[1] * ai.y) + b.r[2] * ai.z;
    }
    return o;
}
RSG_FN M3 transpose(const M3& m) { return m3(col(m, 0), col(m, 1), col(m, 2)); }
RSG_FN M3 scaled(const M3& m, V3 s) { return m3(m.r[0] * s, m.r[1] * s, m.r[2] * s); }

// btMatrix3x3::inverse() by cofactors (LinearMath/btMatrix3x3.h:1093-1103).
RSG_FN float cofac(const M3& m, int r1, int c1, int r2, int c2) {
    return get(m.r[r1], c1) * get(m.r[r2], c2) - get(m.r[r1], c2) * get(m.r[r2], c1);
}
RSG_FN M3 inverse(const M3& m) {
    V3 co = v3(cofac(m, 1, 1, 2, 2), cofac(m, 1, 2, 2, 0), cofac(m, 1, 0, 2, 1));
    float det = dot(m.r[0], co);
    float s = 1.f / det;
    return m3(v3(co.x * s, cofac(m, 0, 2, 2, 1) * s, cofac(m, 0, 1, 1, 2) * s),
              v3(co.y * s, cofac(m, 0, 0, 2, 2) * s, cofac(m, 0, 2, 1, 0) * s),
              v3(co.z * s, cofac(m, 0, 1, 2, 0) * s, cofac(m, 0, 0, 1, 1) * s));
}

struct Q4 {
    float x, y, z, w;
};

RSG_FN Q4 q4(float x, float y, float z, float w) { Q4 q; q.x = x; q.y = y; q.z = z; q.w = w; return q; }
// (x*x + z*z) + (y*y + w*w): the horizontal-add order of the SSE dot.
RSG_FN float qlen2(Q4 q) { return (q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w); }
RSG_FN Q4 qscale(Q4 q, float s) { return q4(q.x * s, q.y * s, q.z * s, q.w * s); }
RSG_FN Q4 qnormalized(Q4 q) { return qscale(q, 1.f / sqrtf(qlen2(q))); }

// Hamilton product with the association of the SSE path
// (LinearMath/btQuaternion.h:619-650).
RSG_FN Q4 qmul(Q4 a, Q4 b) {
    return q4((a.w * b.x - a.z * b.y) + (a.x * b.w + a.y * b.z),
              (a.w * b.y - a.x * b.z) + (a.y * b.w + a.z * b.x),
              (a.w * b.z - a.y * b.x) + (a.z * b.w + a.x * b.y),
              (a.w * b.w - a.z * b.z) + -(a.x * b.x + a.y * b.y));
}

// Axis-angle quaternion (btQuaternion::setRotation, LinearMath/btQuaternion.h:108-115).
RSG_FN Q4 q_axis_angle(V3 axis, float angle) {
    float d = len(axis);
    float s = sinf(angle * 0.5f) / d;
    return q4(axis.x * s, axis.y * s, axis.z * s, cosf(angle * 0.5f));
}

// Quaternion -> rotation matrix (btMatrix3x3::setRotation, SSE association,
// LinearMath/btMatrix3x3.h:216-282).
RSG_FN M3 quat_to_m3(Q4 q) {
    float d = qlen2(q);
    float s = 2.f / d;
    float X = q.x, Y = q.y, Z = q.z, W = q.w;
    V3 r0 = v3((-(Y * Y) + -(Z * Z)) * s + 1.f, (X * Y + -(W * Z)) * s, (Z * X + Y * W) * s);
    V3 r1 = v3((X * Y + Z * W) * s, (-(X * X) + -(Z * Z)) * s + 1.f, (Y * Z + -(W * X)) * s);
    V3 r2 = v3((Z * X + -(W * Y)) * s, (Y * Z + W * X) * s, (-(X * X) + -(Y * Y)) * s + 1.f);
    return m3(r0, r1, r2);
}

// Rotation matrix -> quaternion (btMatrix3x3::getRotation, SSE variant,
// LinearMath/btMatrix3x3.h:421-480).
RSG_FN Q4 m3_to_quat(const M3& m) {
    float trace = (m.r[0].x + m.r[1].y) + m.r[2].z;
    float t[4];
    float x;
    if (trace > 0.f) {
        x = trace + 1.f;
        t[0] = m.r[2].y - m.r[1].z;
        t[1] = m.r[0].z - m.r[2].x;
        t[2] = m.r[1].x - m.r[0].y;
        t[3] = x;
    } else {
        int i, j, k;
        if (m.r[0].x < m.r[1].y) {
            if (m.r[1].y < m.r[2].z) { i = 2; j = 0; k = 1; } else { i = 1; j = 2; k = 0; }
        } else {
            if (m.r[0].x < m.r[2].z) { i = 2; j = 0; k = 1; } else { i = 0; j = 1; k = 2; }
        }
        x = get(m.r[i], i) - get(m.r[j], j) - get(m.r[k], k) + 1.f;
        t[3] = get(m.r[k], j) - get(m.r[j], k);
        t[j] = get(m.r[j], i) + get(m.r[i], j);
        t[k] = get(m.r[k], i) + get(m.r[i], k);
        t[i] = x;
    }
    float s = 0.5f / sqrtf(x);
    return q4(t[0] * s, t[1] * s, t[2] * s, t[3] * s);
}

RSG_FN float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
RSG_FN float rs_max(float a, float b) { return a > b ? a : b; }
RSG_FN float rs_min(float a, float b) { return a < b ? a : b; }
RSG_FN int rs_sgn(float v) { return (v > 0.f) - (v < 0.f); }

}  // namespace rsg
