// rsg_mesh.h -- a shared, read-only arena collision mesh: triangles and a
// 16-bit quantized AABB tree over them.
//
// One copy lives in HBM for the whole batch (every arena uses the same
// mesh), so after the first ticks it is L2-resident.  The tree follows the
// published quantized-AABB-tree scheme the reference uses
// (BulletCollision/BroadphaseCollision/btQuantizedBvh.cpp:116-305,
// BulletCollision/CollisionShapes/btOptimizedBvh.cpp:28-188): leaves in
// triangle order, mean-split along the axis of largest centroid variance,
// depth-first node array with escape indices, 16-bit quantization with a 1 m
// margin, and subtree headers for trees larger than 2 KiB.  Traversal order is
// part of the contract: it decides the order contacts reach the solver.
#pragma once

#include "rsg_math.h"

namespace rsg {

struct QNode {
    uint16_t qmin[3];
    uint16_t qmax[3];
    int32_t esc_or_tri;  // >= 0: leaf, triangle index (part 0); < 0: -escape index
};

struct QSubtree {
    uint16_t qmin[3];
    uint16_t qmax[3];
    int32_t root;
    int32_t size;
};

constexpr int MAX_TRI_INFO_EDGES = 3;

// Per-triangle internal-edge data (btTriangleInfo, BulletCollision/CollisionShapes/btTriangleInfoMap.h:37-54).
struct TriInfo {
    float edge_angle[3];
    uint32_t flags;
};

// Broadphase grid of the reference arena (btRSBroadphase with
// ArenaConfig defaults, src/Sim/Arena/ArenaConfig/ArenaConfig.h:20-41):
// cell = 370 uu, bounds (-5600,-6000,0)..(5600,6000,2200) uu.
constexpr int BP_CELLS_X = 31, BP_CELLS_Y = 33, BP_CELLS_Z = 6;
constexpr int BP_NUM_CELLS = BP_CELLS_X * BP_CELLS_Y * BP_CELLS_Z;

struct MeshView {
    const float* verts;      // 3 floats per vertex (Bullet units)
    const int32_t* tris;     // 3 indices per triangle
    const QNode* nodes;
    const QSubtree* subtrees;
    const TriInfo* tri_info; // may be null
    int n_tris, n_nodes, n_subtrees;
    V3 bvh_min, bvh_max, quant;
    V3 aabb_min, aabb_max;   // local AABB of the mesh (no margin)
    float tri_info_equal_vertex_threshold, tri_info_edge_distance_threshold,
          tri_info_convex_epsilon, tri_info_planar_epsilon, tri_info_max_edge_angle;
    // Occupancy bitmask on a uniform grid over the tree bounds: bit set iff a
    // triangle's box (grown by MESH_OCC_SLACK) touches the cell.  Lets a query
    // whose box touches no marked cell skip the tree walk (no triangle there
    // can yield a contact or a ray hit).  May be null.
    const uint32_t* occ;
    int occ_nx, occ_ny, occ_nz;
    V3 occ_origin;
    float occ_inv_cell;
};

constexpr float MESH_OCC_CELL = 2.f;     // Bullet units
constexpr float MESH_OCC_SLACK = 0.01f;

// True when no triangle of the mesh touches the box [lo, hi]: every grid
// cell the (slightly grown) box touches is unmarked.  Boxes touching many
// cells, or with non-finite bounds, report false (walk the tree).
RSG_FN bool mesh_region_empty(const MeshView& m, V3 lo, V3 hi) {
    if (!m.occ) return false;
    if (!(lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z)) return false;
    V3 a = (lo - v3(MESH_OCC_SLACK, MESH_OCC_SLACK, MESH_OCC_SLACK) - m.occ_origin) * m.occ_inv_cell;
    V3 b = (hi + v3(MESH_OCC_SLACK, MESH_OCC_SLACK, MESH_OCC_SLACK) - m.occ_origin) * m.occ_inv_cell;
    if (!(fabsf(a.x) < 1e6f && fabsf(a.y) < 1e6f && fabsf(a.z) < 1e6f && fabsf(b.x) < 1e6f &&
          fabsf(b.y) < 1e6f && fabsf(b.z) < 1e6f))
        return false;
    int i0 = (int)floorf(a.x), j0 = (int)floorf(a.y), k0 = (int)floorf(a.z);
    int i1 = (int)floorf(b.x), j1 = (int)floorf(b.y), k1 = (int)floorf(b.z);
    i0 = i0 < 0 ? 0 : i0; j0 = j0 < 0 ? 0 : j0; k0 = k0 < 0 ? 0 : k0;
    i1 = i1 >= m.occ_nx ? m.occ_nx - 1 : i1;
    j1 = j1 >= m.occ_ny ? m.occ_ny - 1 : j1;
    k1 = k1 >= m.occ_nz ? m.occ_nz - 1 : k1;
    if (i0 > i1 || j0 > j1 || k0 > k1) return true;  // entirely outside the grid, which covers the mesh
    if ((i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1) > 64) return false;
    for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++)
            for (int k = k0; k <= k1; k++) {
                int bit = (i * m.occ_ny + j) * m.occ_nz + k;
                if (m.occ[bit >> 5] >> (bit & 31) & 1u) return false;
            }
    return true;
}

RSG_FN void quantize(const MeshView& m, uint16_t out[3], V3 p, int is_max) {
    V3 v = (p - m.bvh_min) * m.quant;
    if (is_max) {
        out[0] = (uint16_t)(((uint16_t)(v.x + 1.f)) | 1);
        out[1] = (uint16_t)(((uint16_t)(v.y + 1.f)) | 1);
        out[2] = (uint16_t)(((uint16_t)(v.z + 1.f)) | 1);
    } else {
        out[0] = (uint16_t)(((uint16_t)(v.x)) & 0xfffe);
        out[1] = (uint16_t)(((uint16_t)(v.y)) & 0xfffe);
        out[2] = (uint16_t)(((uint16_t)(v.z)) & 0xfffe);
    }
}

RSG_FN void quantize_clamp(const MeshView& m, uint16_t out[3], V3 p, int is_max) {
    V3 c = p;
    c.x = c.x < m.bvh_min.x ? m.bvh_min.x : c.x;
    c.y = c.y < m.bvh_min.y ? m.bvh_min.y : c.y;
    c.z = c.z < m.bvh_min.z ? m.bvh_min.z : c.z;
    c.x = c.x > m.bvh_max.x ? m.bvh_max.x : c.x;
    c.y = c.y > m.bvh_max.y ? m.bvh_max.y : c.y;
    c.z = c.z > m.bvh_max.z ? m.bvh_max.z : c.z;
    quantize(m, out, c, is_max);
}

RSG_FN V3 unquantize(const MeshView& m, const uint16_t q[3]) {
    return v3((float)q[0] / m.quant.x, (float)q[1] / m.quant.y, (float)q[2] / m.quant.z) + m.bvh_min;
}

RSG_FN bool q_overlap(const uint16_t amin[3], const uint16_t amax[3], const uint16_t bmin[3], const uint16_t bmax[3]) {
    return !(amin[0] > bmax[0] || amax[0] < bmin[0] || amin[1] > bmax[1] || amax[1] < bmin[1] ||
             amin[2] > bmax[2] || amax[2] < bmin[2]);
}

RSG_FN void mesh_triangle(const MeshView& m, int t, V3 out[3]) {
    for (int j = 0; j < 3; j++) {
        int vi = m.tris[3 * t + j];
        out[j] = v3(m.verts[3 * vi], m.verts[3 * vi + 1], m.verts[3 * vi + 2]);
    }
}

// AABB query in the reference's order (walkStacklessQuantizedTreeCacheFriendly,
// btQuantizedBvh.cpp:605-625 -> walkStacklessQuantizedTree :627-665).
// cb(int tri) is called for every leaf whose quantized box overlaps.
template <class CB>
RSG_FN void mesh_query_aabb(const MeshView& m, V3 qmin_f, V3 qmax_f, CB& cb) {
    if (mesh_region_empty(m, qmin_f, qmax_f)) return;
    uint16_t qmin[3], qmax[3];
    quantize_clamp(m, qmin, qmin_f, 0);
    quantize_clamp(m, qmax, qmax_f, 1);
    for (int s = 0; s < m.n_subtrees; s++) {
        const QSubtree& st = m.subtrees[s];
        if (!q_overlap(qmin, qmax, st.qmin, st.qmax)) continue;
        int cur = st.root, end = st.root + st.size;
        while (cur < end) {
            const QNode& n = m.nodes[cur];
            bool ov = q_overlap(qmin, qmax, n.qmin, n.qmax);
            bool leaf = n.esc_or_tri >= 0;
            if (leaf && ov) cb(n.esc_or_tri);
            if (ov || leaf) cur++;
            else cur += -n.esc_or_tri;
        }
    }
}

// Ray query (walkStacklessQuantizedTreeAgainstRay with RAYAABB2,
// btQuantizedBvh.cpp:531-603 and LinearMath/btAabbUtil2.h:82-115).
template <class CB>
RSG_FN void mesh_query_ray(const MeshView& m, V3 from, V3 to, CB& cb) {
    V3 dir = safe_normalized(to - from);
    float lambda_max = dot(dir, to - from);
    V3 inv = v3(dir.x == 0.f ? 1e18f : 1.f / dir.x, dir.y == 0.f ? 1e18f : 1.f / dir.y,
                dir.z == 0.f ? 1e18f : 1.f / dir.z);
    int sign[3] = {inv.x < 0.f, inv.y < 0.f, inv.z < 0.f};
    V3 rmin = v3(fminf(from.x, to.x), fminf(from.y, to.y), fminf(from.z, to.z));
    V3 rmax = v3(fmaxf(from.x, to.x), fmaxf(from.y, to.y), fmaxf(from.z, to.z));
    if (mesh_region_empty(m, rmin, rmax)) return;
    uint16_t qmin[3], qmax[3];
    quantize_clamp(m, qmin, rmin, 0);
    quantize_clamp(m, qmax, rmax, 1);
    int cur = 0;
    while (cur < m.n_nodes) {
        const QNode& n = m.nodes[cur];
        bool ray_ov = false;
        if (q_overlap(qmin, qmax, n.qmin, n.qmax)) {
            V3 b[2] = {unquantize(m, n.qmin), unquantize(m, n.qmax)};
            float tmin = (get(b[sign[0]], 0) - from.x) * inv.x;
            float tmax = (get(b[1 - sign[0]], 0) - from.x) * inv.x;
            float tymin = (get(b[sign[1]], 1) - from.y) * inv.y;
            float tymax = (get(b[1 - sign[1]], 1) - from.y) * inv.y;
            if (!((tmin > tymax) || (tymin > tmax))) {
                if (tymin > tmin) tmin = tymin;
                if (tymax < tmax) tmax = tymax;
                float tzmin = (get(b[sign[2]], 2) - from.z) * inv.z;
                float tzmax = (get(b[1 - sign[2]], 2) - from.z) * inv.z;
                if (!((tmin > tzmax) || (tzmin > tmax))) {
                    if (tzmin > tmin) tmin = tzmin;
                    if (tzmax < tmax) tmax = tzmax;
                    ray_ov = (tmin < lambda_max) && (tmax > 0.f);
                }
            }
        }
        bool leaf = n.esc_or_tri >= 0;
        if (leaf && ray_ov) cb(n.esc_or_tri);
        if (ray_ov || leaf) cur++;
        else cur += -n.esc_or_tri;
    }
}

}  // namespace rsg
