// rsg_mesh_build.h -- host-side construction of the shared arena geometry
// (MeshView in rsg_mesh.h, ArenaGeom in rsg_contact.h): per mesh the
// quantized AABB tree and the internal-edge table, per arena the broadphase
// cell table of its static objects.
//
// The tree is built exactly as the reference builds its BVH
// (btOptimizedBvh::build, BulletCollision/CollisionShapes/btOptimizedBvh.cpp:28-188;
// btQuantizedBvh::setQuantizationValues/buildTree/sortAndCalcSplittingIndex/
// calcSplittingAxis/updateSubtreeHeaders, btQuantizedBvh.cpp:69-305), because
// the node order it produces is the order in which contacts are generated.
#pragma once

#include <algorithm>
#include <cmath>
#include <immintrin.h>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "rsg_contact.h"

namespace rsg {

struct HostMesh {
    std::vector<float> verts;     // 3 per vertex, Bullet units
    std::vector<int32_t> tris;    // 3 per triangle
    std::vector<QNode> nodes;
    std::vector<QSubtree> subtrees;
    std::vector<TriInfo> tri_info;
    V3 bvh_min, bvh_max, quant;
    V3 aabb_min, aabb_max;
    std::vector<uint32_t> occ;          // MeshView::occ bitmask
    int occ_dims[3] = {0, 0, 0};

    MeshView view() const {
        MeshView v{};
        v.verts = verts.data();
        v.tris = tris.data();
        v.nodes = nodes.data();
        v.subtrees = subtrees.data();
        v.tri_info = tri_info.empty() ? nullptr : tri_info.data();
        v.n_tris = (int)tris.size() / 3;
        v.n_nodes = (int)nodes.size();
        v.n_subtrees = (int)subtrees.size();
        v.bvh_min = bvh_min; v.bvh_max = bvh_max; v.quant = quant;
        v.aabb_min = aabb_min; v.aabb_max = aabb_max;
        v.occ = occ.empty() ? nullptr : occ.data();
        v.occ_nx = occ_dims[0]; v.occ_ny = occ_dims[1]; v.occ_nz = occ_dims[2];
        v.occ_origin = bvh_min;
        v.occ_inv_cell = 1.f / MESH_OCC_CELL;
        return v;
    }
};

// CollisionMeshFile::ReadFromStream + MakeBulletMesh
// (src/CollisionMeshFile/CollisionMeshFile.cpp:11-73): int32 triangle and
// vertex counts, int32[3] per triangle, float32[3] per vertex (Bullet units,
// no scaling on load).  Returns an empty string or an error message.
inline std::string parse_cmf(const void* data, uint64_t size, HostMesh& M) {
    const uint8_t* p = (const uint8_t*)data;
    if (!p || size < 8) return "truncated .cmf header";
    int32_t nt, nv;
    std::memcpy(&nt, p, 4);
    std::memcpy(&nv, p + 4, 4);
    if ((nt < nv ? nt : nv) <= 0 || (nt > nv ? nt : nv) > 1000 * 1000) return "bad triangle/vertex count";
    uint64_t need = 8 + (uint64_t)nt * 12 + (uint64_t)nv * 12;
    if (size < need) return "input data overflown";
    M.tris.resize((size_t)nt * 3);
    M.verts.resize((size_t)nv * 3);
    std::memcpy(M.tris.data(), p + 8, (size_t)nt * 12);
    std::memcpy(M.verts.data(), p + 8 + (size_t)nt * 12, (size_t)nv * 12);
    for (int32_t idx : M.tris)
        if (idx < 0 || idx >= nv) return "bad triangle vertex index";
    return "";
}

struct BvhBuilder {
    HostMesh& M;
    std::vector<QNode> leaves;
    int cur = 0;

    explicit BvhBuilder(HostMesh& m) : M(m) {}

    MeshView view() const {
        MeshView v{};
        v.bvh_min = M.bvh_min; v.bvh_max = M.bvh_max; v.quant = M.quant;
        return v;
    }

    void set_quantization(V3 amin, V3 amax, float margin) {
        V3 clamp = v3(margin, margin, margin);
        M.bvh_min = amin - clamp;
        M.bvh_max = amax + clamp;
        V3 size = M.bvh_max - M.bvh_min;
        M.quant = v3(65533.f / size.x, 65533.f / size.y, 65533.f / size.z);
        {
            uint16_t q[3];
            MeshView mv = view();
            quantize(mv, q, M.bvh_min, 0);
            V3 v = unquantize(mv, q);
            V3 c = v - clamp;
            M.bvh_min = v3(fminf(M.bvh_min.x, c.x), fminf(M.bvh_min.y, c.y), fminf(M.bvh_min.z, c.z));
        }
        size = M.bvh_max - M.bvh_min;
        M.quant = v3(65533.f / size.x, 65533.f / size.y, 65533.f / size.z);
        {
            uint16_t q[3];
            MeshView mv = view();
            quantize(mv, q, M.bvh_max, 1);
            V3 v = unquantize(mv, q);
            V3 c = v + clamp;
            M.bvh_max = v3(fmaxf(M.bvh_max.x, c.x), fmaxf(M.bvh_max.y, c.y), fmaxf(M.bvh_max.z, c.z));
        }
        size = M.bvh_max - M.bvh_min;
        M.quant = v3(65533.f / size.x, 65533.f / size.y, 65533.f / size.z);
    }

    V3 leaf_min(int i) const { MeshView mv = view(); return unquantize(mv, leaves[i].qmin); }
    V3 leaf_max(int i) const { MeshView mv = view(); return unquantize(mv, leaves[i].qmax); }

    static int max_axis(V3 v) { return v.x < v.y ? (v.y < v.z ? 2 : 1) : (v.x < v.z ? 2 : 0); }

    int calc_splitting_axis(int s, int e) {
        V3 means = v3zero(), var = v3zero();
        int n = e - s;
        for (int i = s; i < e; i++) means += 0.5f * (leaf_max(i) + leaf_min(i));
        means *= (1.f / (float)n);
        for (int i = s; i < e; i++) {
            V3 c = 0.5f * (leaf_max(i) + leaf_min(i));
            V3 d = c - means;
            d = d * d;
            var += d;
        }
        var *= (1.f / ((float)n - 1));
        return max_axis(var);
    }

    int sort_and_calc_split(int s, int e, int axis) {
        int split = s;
        int n = e - s;
        V3 means = v3zero();
        for (int i = s; i < e; i++) means += 0.5f * (leaf_max(i) + leaf_min(i));
        means *= (1.f / (float)n);
        float split_value = get(means, axis);
        for (int i = s; i < e; i++) {
            V3 c = 0.5f * (leaf_max(i) + leaf_min(i));
            if (get(c, axis) > split_value) {
                QNode t = leaves[i];
                leaves[i] = leaves[split];
                leaves[split] = t;
                split++;
            }
        }
        int balanced = n / 3;
        bool unbalanced = (split <= (s + balanced)) || (split >= (e - 1 - balanced));
        if (unbalanced) split = s + (n >> 1);
        return split;
    }

    void update_subtree_headers(int left, int right) {
        const QNode& l = M.nodes[left];
        int lsize = l.esc_or_tri >= 0 ? 1 : -l.esc_or_tri;
        const QNode& r = M.nodes[right];
        int rsize = r.esc_or_tri >= 0 ? 1 : -r.esc_or_tri;
        if (lsize * (int)sizeof(QNode) <= 2048) {
            QSubtree st;
            std::memcpy(st.qmin, l.qmin, sizeof(st.qmin));
            std::memcpy(st.qmax, l.qmax, sizeof(st.qmax));
            st.root = left;
            st.size = lsize;
            M.subtrees.push_back(st);
        }
        if (rsize * (int)sizeof(QNode) <= 2048) {
            QSubtree st;
            std::memcpy(st.qmin, r.qmin, sizeof(st.qmin));
            std::memcpy(st.qmax, r.qmax, sizeof(st.qmax));
            st.root = right;
            st.size = rsize;
            M.subtrees.push_back(st);
        }
    }

    void build_tree(int s, int e) {
        int n = e - s;
        int cur_index = cur;
        if (n == 1) {
            M.nodes[cur] = leaves[s];
            cur++;
            return;
        }
        int axis = calc_splitting_axis(s, e);
        int split = sort_and_calc_split(s, e, axis);
        int internal = cur;
        MeshView mv = view();
        quantize(mv, M.nodes[cur].qmin, M.bvh_max, 0);
        quantize(mv, M.nodes[cur].qmax, M.bvh_min, 1);
        for (int i = s; i < e; i++) {
            uint16_t qmin[3], qmax[3];
            quantize(mv, qmin, leaf_min(i), 0);
            quantize(mv, qmax, leaf_max(i), 1);
            for (int k = 0; k < 3; k++) {
                if (M.nodes[cur].qmin[k] > qmin[k]) M.nodes[cur].qmin[k] = qmin[k];
                if (M.nodes[cur].qmax[k] < qmax[k]) M.nodes[cur].qmax[k] = qmax[k];
            }
        }
        cur++;
        int left = cur;
        build_tree(s, split);
        int right = cur;
        build_tree(split, e);
        int escape = cur - cur_index;
        if (escape * (int)sizeof(QNode) > 2048) update_subtree_headers(left, right);
        M.nodes[internal].esc_or_tri = -escape;
    }

    void build() {
        int nt = (int)M.tris.size() / 3;
        // local AABB of the mesh: support vertices along +-axes over all
        // triangles (btTriangleMeshShape::recalcLocalAabb), margin 0.
        V3 amin = v3(1e18f, 1e18f, 1e18f), amax = v3(-1e18f, -1e18f, -1e18f);
        for (int t = 0; t < nt; t++)
            for (int j = 0; j < 3; j++) {
                int vi = M.tris[3 * t + j];
                V3 p = v3(M.verts[3 * vi], M.verts[3 * vi + 1], M.verts[3 * vi + 2]);
                amin = v3(fminf(amin.x, p.x), fminf(amin.y, p.y), fminf(amin.z, p.z));
                amax = v3(fmaxf(amax.x, p.x), fmaxf(amax.y, p.y), fmaxf(amax.z, p.z));
            }
        M.aabb_min = amin;
        M.aabb_max = amax;
        set_quantization(amin, amax, 1.0f);
        MeshView mv = view();
        leaves.resize(nt);
        for (int t = 0; t < nt; t++) {
            V3 p[3];
            for (int j = 0; j < 3; j++) {
                int vi = M.tris[3 * t + j];
                p[j] = v3(M.verts[3 * vi], M.verts[3 * vi + 1], M.verts[3 * vi + 2]);
            }
            V3 bmin = v3(1e18f, 1e18f, 1e18f), bmax = v3(-1e18f, -1e18f, -1e18f);
            for (int j = 0; j < 3; j++) {
                bmin = v3(fminf(bmin.x, p[j].x), fminf(bmin.y, p[j].y), fminf(bmin.z, p[j].z));
                bmax = v3(fmaxf(bmax.x, p[j].x), fmaxf(bmax.y, p[j].y), fmaxf(bmax.z, p[j].z));
            }
            const float MIN_DIM = 0.002f, MIN_HALF = 0.001f;
            if (bmax.x - bmin.x < MIN_DIM) { bmax.x = bmax.x + MIN_HALF; bmin.x = bmin.x - MIN_HALF; }
            if (bmax.y - bmin.y < MIN_DIM) { bmax.y = bmax.y + MIN_HALF; bmin.y = bmin.y - MIN_HALF; }
            if (bmax.z - bmin.z < MIN_DIM) { bmax.z = bmax.z + MIN_HALF; bmin.z = bmin.z - MIN_HALF; }
            quantize(mv, leaves[t].qmin, bmin, 0);
            quantize(mv, leaves[t].qmax, bmax, 1);
            leaves[t].esc_or_tri = t;
        }
        M.nodes.assign(2 * nt, QNode{});
        cur = 0;
        build_tree(0, nt);
        if (M.subtrees.empty()) {
            QSubtree st;
            std::memcpy(st.qmin, M.nodes[0].qmin, sizeof(st.qmin));
            std::memcpy(st.qmax, M.nodes[0].qmax, sizeof(st.qmax));
            st.root = 0;
            st.size = M.nodes[0].esc_or_tri >= 0 ? 1 : -M.nodes[0].esc_or_tri;
            M.subtrees.push_back(st);
        }
        M.nodes.resize(cur);
    }
};

// ---------------------------------------------------------------------------
// Internal-edge table (btGenerateInternalEdgeInfo + btConnectivityProcessor,
// BulletCollision/CollisionDispatch/btInternalEdgeUtility.cpp:38-368, with the
// btTriangleInfoMap defaults of btTriangleInfoMap.h:61-69).  For every triangle
// A, the triangles whose BVH leaves overlap A's AABB are visited in tree order;
// a neighbour sharing exactly two vertices writes the angle of the shared edge
// and ORs its convex / swap flags (a later neighbour on the same edge
// overwrites the angle, as in the reference).

// btVector3::normalize of the reference's x86-64 build (BT_USE_SSE_IN_API,
// LinearMath/btVector3.h:304-345): one reciprocal-square-root estimate and one
// Newton step.  Used here so the host-side tables come out bit-identical to
// the reference's on the same CPU; the device code divides by the length.
inline V3 host_sse_normalized(V3 a) {
    float l2 = (a.x * a.x + a.y * a.y) + a.z * a.z;
    float y0 = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(l2)));
    float vd = ((l2 * 0.5f) * y0) * y0;
    float y = y0 * (1.5f - vd);
    return v3(a.x * y, a.y * y, a.z * y);
}

inline V3 host_vertex(const HostMesh& M, int vi) {
    return v3(M.verts[3 * vi], M.verts[3 * vi + 1], M.verts[3 * vi + 2]);
}

inline void build_tri_info(HostMesh& M) {
    const float equal_vertex_threshold = 0.0001f * 0.0001f;
    const float planar_epsilon = 0.0001f;
    const float simd_pi = 3.1415926535897932384626433832795029f;
    int nt = (int)M.tris.size() / 3;
    M.tri_info.assign(nt, TriInfo{{TWO_PI_F, TWO_PI_F, TWO_PI_F}, 0u});
    MeshView mv = M.view();
    mv.tri_info = nullptr;
    for (int ta = 0; ta < nt; ta++) {
        V3 A[3];
        for (int j = 0; j < 3; j++) A[j] = host_vertex(M, M.tris[3 * ta + j]);
        V3 amin = v3(1e18f, 1e18f, 1e18f), amax = v3(-1e18f, -1e18f, -1e18f);
        for (int j = 0; j < 3; j++) {
            amin = v3(fminf(amin.x, A[j].x), fminf(amin.y, A[j].y), fminf(amin.z, A[j].z));
            amax = v3(fmaxf(amax.x, A[j].x), fmaxf(amax.y, A[j].y), fmaxf(amax.z, A[j].z));
        }
        struct Conn {
            HostMesh* M; int ta; const V3* A; float eq, planar, pi;
            void operator()(int tb) {
                if (tb == ta) return;
                V3 B[3];
                for (int j = 0; j < 3; j++) B[j] = host_vertex(*M, M->tris[3 * tb + j]);
                if (len2(cross(B[1] - B[0], B[2] - B[0])) < eq) return;
                if (len2(cross(A[1] - A[0], A[2] - A[0])) < eq) return;
                int numshared = 0;
                int sa[3] = {-1, -1, -1}, sb[3] = {-1, -1, -1};
                for (int i = 0; i < 3; i++) {
                    for (int j = 0; j < 3; j++) {
                        if (len2(A[i] - B[j]) < eq) {
                            sa[numshared] = i;
                            sb[numshared] = j;
                            numshared++;
                            if (numshared >= 3) return;
                        }
                    }
                    if (numshared >= 3) return;
                }
                if (numshared != 2) return;
                if (sa[0] == 0 && sa[1] == 2) {
                    sa[0] = 2; sa[1] = 0;
                    int t = sb[1]; sb[1] = sb[0]; sb[0] = t;
                }
                TriInfo& info = M->tri_info[ta];
                info.flags |= TRI_HAS_INFO;
                int sum_a = sa[0] + sa[1];
                int other_a = 3 - sum_a;
                V3 edge = A[sa[1]] - A[sa[0]];
                int other_b = 3 - (sb[0] + sb[1]);
                V3 tb0 = B[sb[1]], tb1 = B[sb[0]], tb2 = B[other_b];
                V3 normal_a = host_sse_normalized(cross(A[1] - A[0], A[2] - A[0]));
                V3 normal_b = host_sse_normalized(cross(tb1 - tb0, tb2 - tb0));
                edge = host_sse_normalized(edge);
                V3 edge_cross_a = host_sse_normalized(cross(edge, normal_a));
                if (dot(edge_cross_a, A[other_a] - A[sa[0]]) < 0) edge_cross_a = edge_cross_a * -1.f;
                V3 edge_cross_b = host_sse_normalized(cross(edge, normal_b));
                if (dot(edge_cross_b, B[other_b] - B[sb[0]]) < 0) edge_cross_b = edge_cross_b * -1.f;
                V3 calc_edge = cross(edge_cross_a, edge_cross_b);
                float l2 = len2(calc_edge);
                float corrected = 0;
                bool is_convex = false;
                if (!(l2 < planar)) {
                    calc_edge = host_sse_normalized(calc_edge);
                    V3 calc_normal_a = host_sse_normalized(cross(calc_edge, edge_cross_a));
                    float angle2 = atan2f(dot(edge_cross_b, calc_normal_a), dot(edge_cross_b, edge_cross_a));
                    float ang4 = pi - angle2;
                    is_convex = dot(normal_a, edge_cross_b) < 0.f;
                    corrected = is_convex ? ang4 : -ang4;
                }
                int e;
                V3 axis;
                uint32_t swap_bit, convex_bit;
                if (sum_a == 1) { e = 0; axis = A[0] - A[1]; swap_bit = TRI_V0V1_SWAP; convex_bit = TRI_V0V1_CONVEX; }
                else if (sum_a == 2) { e = 2; axis = A[2] - A[0]; swap_bit = TRI_V2V0_SWAP; convex_bit = TRI_V2V0_CONVEX; }
                else { e = 1; axis = A[1] - A[2]; swap_bit = TRI_V1V2_SWAP; convex_bit = TRI_V1V2_CONVEX; }
                V3 computed_b = quat_rotate(q_axis_angle(axis, -corrected), normal_a);
                if (dot(computed_b, normal_b) < 0) info.flags |= swap_bit;
                info.edge_angle[e] = -corrected;
                if (is_convex) info.flags |= convex_bit;
            }
        } conn{&M, ta, A, equal_vertex_threshold, planar_epsilon, simd_pi};
        mesh_query_aabb(mv, amin, amax, conn);
    }
}

// ---------------------------------------------------------------------------
// An arena's static geometry on the host: its collision meshes in load order
// and the broadphase cell masks of all its static objects (ArenaGeom).
struct HostArena {
    std::vector<HostMesh> meshes;
    std::vector<MeshView> views;        // host views of `meshes` (refresh_views)
    std::vector<uint32_t> cell_statics; // BP_NUM_CELLS masks
    int n_planes = 0;
    PlaneDef planes[MAX_PLANES];
    uint32_t car_skip = 0;              // ArenaGeom::car_skip

    void refresh_views() {
        views.clear();
        for (const HostMesh& m : meshes) views.push_back(m.view());
    }
    ArenaGeom geom() const {
        ArenaGeom g{};
        g.meshes = views.data();
        g.n_meshes = (int)views.size();
        g.cell_statics = cell_statics.empty() ? nullptr : cell_statics.data();
        g.n_planes = n_planes;
        g.car_skip = car_skip;
        for (int i = 0; i < MAX_PLANES; i++) g.planes[i] = planes[i];
        return g;
    }
};

// Static objects per broadphase cell (_UpdateCellsStatic<true>,
// BulletCollision/BroadphaseCollision/btRSBroadphase.cpp:96-185): for each
// cell (i,j,k) covered by the static proxy's AABB (max clamped to the grid
// max), the proxy is registered in the 27 cells around it -- for a mesh only
// if one of its triangles' AABB overlaps the cell's box.  The reference reaches the
// mesh through a btTriangleMeshShape pointer there, and processAllTriangles is
// not virtual in its Bullet copy (btTriangleMeshShape.h:54), so the test is
// the brute-force filtered walk over every triangle
// (btTriangleMeshShape.cpp:116-141), not the BVH query.

inline void build_cell_statics(HostArena& H) {
    H.refresh_views();
    H.cell_statics.assign(BP_NUM_CELLS, 0u);
    const ArenaGeom g = H.geom();
    V3 gmax = bp_max_pos();
    for (int s = 0; s < num_statics(g); s++) {
        V3 mn, mx;
        static_proxy_aabb(&g, s, mn, mx);
        mx = v3(fminf(mx.x, gmax.x), fminf(mx.y, gmax.y), fminf(mx.z, gmax.z));
        int i0, j0, k0, i1, j1, k1;
        bp_cell_indices(mn, i0, j0, k0);
        bp_cell_indices(mx, i1, j1, k1);
        for (int i = i0; i <= i1; i++)
            for (int j = j0; j <= j1; j++)
                for (int k = k0; k <= k1; k++) {
                    if (static_is_mesh(g, s)) {
                        const MeshView& mv = g.meshes[s];
                        V3 cmin = bp_min_pos() + v3((float)i, (float)j, (float)k) * bp_cell_size();
                        V3 cmax = cmin + v3(bp_cell_size(), bp_cell_size(), bp_cell_size());
                        bool hit = false;
                        for (int t = 0; t < mv.n_tris && !hit; t++) {
                            V3 v[3];
                            mesh_triangle(mv, t, v);
                            hit = triangle_aabb_overlap(v, cmin, cmax);
                        }
                        if (!hit) continue;
                    }
                    for (int di = -1; di <= 1; di++)
                        for (int dj = -1; dj <= 1; dj++)
                            for (int dk = -1; dk <= 1; dk++) {
                                int ci = i + di, cj = j + dj, ck = k + dk;
                                if (ci < 0 || cj < 0 || ck < 0) continue;
                                if (ci >= BP_CELLS_X || cj >= BP_CELLS_Y || ck >= BP_CELLS_Z) continue;
                                H.cell_statics[bp_cell_index(ci, cj, ck)] |= 1u << s;
                            }
                }
    }
}

// Occupancy grid (MeshView::occ): mark every cell a triangle's box, grown by
// MESH_OCC_SLACK, touches (same cell arithmetic as mesh_region_empty).
inline void build_occupancy(HostMesh& M) {
    V3 ext = M.bvh_max - M.bvh_min;
    M.occ_dims[0] = (int)std::ceil(ext.x / MESH_OCC_CELL) + 1;
    M.occ_dims[1] = (int)std::ceil(ext.y / MESH_OCC_CELL) + 1;
    M.occ_dims[2] = (int)std::ceil(ext.z / MESH_OCC_CELL) + 1;
    size_t bits = (size_t)M.occ_dims[0] * M.occ_dims[1] * M.occ_dims[2];
    M.occ.assign((bits + 31) / 32, 0u);
    const float inv = 1.f / MESH_OCC_CELL;
    const int nt = (int)M.tris.size() / 3;
    for (int t = 0; t < nt; t++) {
        V3 lo = v3(1e30f, 1e30f, 1e30f), hi = v3(-1e30f, -1e30f, -1e30f);
        for (int j = 0; j < 3; j++) {
            int vi = M.tris[3 * t + j];
            V3 p = v3(M.verts[3 * vi], M.verts[3 * vi + 1], M.verts[3 * vi + 2]);
            lo = v3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
            hi = v3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
        }
        V3 a = (lo - v3(MESH_OCC_SLACK, MESH_OCC_SLACK, MESH_OCC_SLACK) - M.bvh_min) * inv;
        V3 b = (hi + v3(MESH_OCC_SLACK, MESH_OCC_SLACK, MESH_OCC_SLACK) - M.bvh_min) * inv;
        int i0 = std::max(0, (int)std::floor(a.x)), j0 = std::max(0, (int)std::floor(a.y)),
            k0 = std::max(0, (int)std::floor(a.z));
        int i1 = std::min(M.occ_dims[0] - 1, (int)std::floor(b.x)),
            j1 = std::min(M.occ_dims[1] - 1, (int)std::floor(b.y)),
            k1 = std::min(M.occ_dims[2] - 1, (int)std::floor(b.z));
        for (int i = i0; i <= i1; i++)
            for (int j = j0; j <= j1; j++)
                for (int k = k0; k <= k1; k++) {
                    size_t bit = ((size_t)i * M.occ_dims[1] + j) * M.occ_dims[2] + k;
                    M.occ[bit >> 5] |= 1u << (bit & 31);
                }
    }
}

// Full preparation of one collision mesh from .cmf bytes (one entry of
// RocketSim::InitFromMem's list, src/RocketSim.cpp:129-155: the mesh, its
// BVH and its triangle info map).
inline std::string build_host_mesh(const void* data, uint64_t size, HostMesh& M) {
    std::string err = parse_cmf(data, size, M);
    if (!err.empty()) return err;
    BvhBuilder b(M);
    b.build();
    build_tri_info(M);
    build_occupancy(M);
    return "";
}

// Vertex count by which the reference recognises a Hoops net mesh
// (HOOPS_NET_NUM_VERTS, src/Sim/Arena/Arena.cpp:1008-1018).
constexpr int HOOPS_NET_NUM_VERTS = 505;

// An arena of kind `kind` (0: Soccar, 1: Hoops) from its n collision mesh
// files, in the order given (the reference adds one static body per mesh in
// list order, then the planes, src/Sim/Arena/Arena.cpp:990-1068).  Returns ""
// or why the files were refused.
inline std::string build_host_arena(const void* const* data, const uint64_t* sizes, int n, int kind,
                                    HostArena& H) {
    if (kind != 0 && kind != 1) return "unknown arena kind";
    if (n < 1 || n > MAX_MESHES)
        return "between 1 and " + std::to_string(MAX_MESHES) + " mesh files are supported, got " + std::to_string(n);
    H.meshes.assign((size_t)n, HostMesh());
    for (int i = 0; i < n; i++) {
        if (!data[i]) return "mesh file " + std::to_string(i) + " is null";
        std::string err = build_host_mesh(data[i], sizes[i], H.meshes[(size_t)i]);
        if (!err.empty()) return "mesh file " + std::to_string(i) + ": " + err;
        if (kind == 1 && (int)H.meshes[(size_t)i].verts.size() / 3 == HOOPS_NET_NUM_VERTS) H.car_skip |= 1u << i;
    }
    H.n_planes = arena_planes(kind, host_sse_normalized(v3(0.f, 0.f, 1.f)).z, H.planes);
    build_cell_statics(H);
    return "";
}

}  // namespace rsg
