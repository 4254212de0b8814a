// Host build of single narrowphase components of the tick code
// (rocketsim_amd/csrc/rsg_gjk.h, rsg_boxbox.h), with the reference's SSE
// normalisation (btVector3::normalize: rsqrtss plus one Newton step,
// LinearMath/btVector3.h), for known-answer tests against the same
// components of the reference (oracle/ref_harness.cpp rsref_gjk,
// rsref_box_box, rsref_ray_convex).  Signatures mirror those probes.
#include <immintrin.h>

template <class T>
static inline T sse_normalized(T a) {
    float l2 = (a.x * a.x + a.y * a.y) + a.z * a.z;
    float y = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(l2)));
    float h = l2 * 0.5f;
    h = h * y;
    h = h * y;
    y = y * (1.5f - h);
    T r = a;
    r.x = a.x * y;
    r.y = a.y * y;
    r.z = a.z * y;
    return r;
}
#define RSG_NORMALIZE_HOOK(a) sse_normalized(a)

#include <memory>

#include "../../rocketsim_amd/csrc/rsg_collide.h"

using namespace rsg;

namespace {
Xf xf_from_floats(const float* x) {
    Xf t;
    t.R = m3(v3(x[0], x[1], x[2]), v3(x[3], x[4], x[5]), v3(x[6], x[7], x[8]));
    t.t = v3(x[9], x[10], x[11]);
    return t;
}
void put(float* o, V3 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
}  // namespace

extern "C" {

// Box (implicit half extents + margin, as btBoxShape derives them from the
// full half extents) against a sphere (kind 0) or a triangle (kind 2).
int ch_gjk(const float* box_implicit, float box_margin, const float* xa, int kind_b, float radius, const float* tri,
           const float* xb, float max_dist2, float* out_n, float* out_p, float* out_d) {
    CShape box = shape_box(v3(box_implicit[0], box_implicit[1], box_implicit[2]), box_margin);
    V3 v[3] = {v3(tri[0], tri[1], tri[2]), v3(tri[3], tri[4], tri[5]), v3(tri[6], tri[7], tri[8])};
    CShape b = kind_b == 0 ? shape_sphere(radius) : shape_triangle(v);
    auto space = std::make_unique<EpaSpace>();
    V3 n, p;
    float d;
    if (gjk_closest(box, b, xf_from_floats(xa), xf_from_floats(xb), max_dist2, n, p, d, space.get()) != 1) return 0;
    put(out_n, n);
    put(out_p, p);
    *out_d = d;
    return 1;
}

// Two boxes given by their full half extents (margin included).
int ch_box_box(const float* h0, const float* x0, const float* h1, const float* x1, float* normals, float* points,
               float* depths, int cap) {
    BoxBoxContacts bc;
    box_box(xf_from_floats(x0), v3(h0[0], h0[1], h0[2]), xf_from_floats(x1), v3(h1[0], h1[1], h1[2]), bc);
    for (int i = 0; i < bc.n && i < cap; i++) {
        put(normals + 3 * i, bc.normal);
        put(points + 3 * i, bc.point[i]);
        depths[i] = bc.depth[i];
    }
    return bc.n;
}

// A ray (a point cast from `from` to `to`) against a sphere (kind 0) or a box
// (kind 1, implicit half extents + margin), accepted as the world ray query
// accepts it.
int ch_ray_convex(int kind, float radius, const float* box_implicit, float box_margin, const float* xb,
                  const float* from, const float* to, float* frac, float* normal) {
    CShape b = kind == 0 ? shape_sphere(radius)
                         : shape_box(v3(box_implicit[0], box_implicit[1], box_implicit[2]), box_margin);
    float f;
    V3 n;
    if (!ray_convex(v3(from[0], from[1], from[2]), v3(to[0], to[1], to[2]), b, xf_from_floats(xb), f, n)) return 0;
    // the world ray query accepts the hit and normalises its normal again
    // (World::raycast, as btCollisionWorld::rayTestSingleInternal does)
    if (!(len2(n) > 0.0001f)) return 0;
    *frac = f;
    put(normal, normalized(n));
    return 1;
}

}  // extern "C"

// The arena mesh as the library builds it (rsg_mesh_build.h): triangle
// internal-edge records and AABB queries over the quantized tree.
#include "../../rocketsim_amd/csrc/rsg_mesh_build.h"

extern "C" {

void* ch_mesh_create(const void* cmf, uint64_t size) {
    auto* m = new HostMesh();
    if (!build_host_mesh(cmf, size, *m).empty()) {
        delete m;
        return nullptr;
    }
    return m;
}

void ch_mesh_destroy(void* m) { delete (HostMesh*)m; }

int ch_mesh_num_tris(void* m) { return (int)(((HostMesh*)m)->tris.size() / 3); }

void ch_tri_info(void* m, int tri, float* angles, uint32_t* flags) {
    const TriInfo& t = ((HostMesh*)m)->tri_info[tri];
    for (int k = 0; k < 3; k++) angles[k] = t.edge_angle[k];
    *flags = t.flags;
}

// Triangles the query box [mn, mx] reaches, in tree visit order, kept when
// their own AABB overlaps the box (the narrowphase's first test).
int ch_mesh_query_aabb(void* m, const float* mn, const float* mx, int* out, int max_out) {
    MeshView mv = ((HostMesh*)m)->view();
    struct CB {
        const MeshView* mv; V3 a, b; int* out; int max_out; int n;
        void operator()(int t) {
            V3 v[3];
            mesh_triangle(*mv, t, v);
            if (!triangle_aabb_overlap(v, a, b)) return;
            if (n < max_out) out[n] = t;
            n++;
        }
    } cb{&mv, v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]), out, max_out, 0};
    mesh_query_aabb(mv, cb.a, cb.b, cb);
    return cb.n;
}

// Keeps the triangles of `in` whose AABB overlaps [mn, mx] (applied to the
// reference's raw visit list for the comparison above).
int ch_filter_tris(void* m, const float* mn, const float* mx, const int* in, int n, int* out) {
    MeshView mv = ((HostMesh*)m)->view();
    int k = 0;
    for (int i = 0; i < n; i++) {
        V3 v[3];
        mesh_triangle(mv, in[i], v);
        if (triangle_aabb_overlap(v, v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]))) out[k++] = in[i];
    }
    return k;
}

}  // extern "C"

extern "C" {

// Centroid of mesh triangle t (for placing queries on the surface).
void ch_mesh_tri_center(void* m, int t, float* out) {
    MeshView mv = ((HostMesh*)m)->view();
    V3 v[3];
    mesh_triangle(mv, t, v);
    V3 c = (v[0] + v[1] + v[2]) * (1.f / 3.f);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

}  // extern "C"
