"""Procedural Soccar arena collision mesh, written in the reference's
CollisionMeshFile (.cmf) format.

The reference ships no arena meshes (RocketSim.cpp loads them at run time,
src/RocketSim.cpp:92-185), so every workload here uses one synthetic arena:

  * playing volume x in [-4096, 4096], y in [-5120, 5120], z in [0, 2044] uu;
  * the floor (z = 0), ceiling (z = 2048) and side walls (x = +-4096) are the
    reference's static planes (src/Sim/Arena/Arena.cpp:1036-1068) and are not
    part of the mesh;
  * the mesh holds the back walls (y = +-5120), 45 degree corner bevels of
    1152 uu, 256 uu fillets where walls meet floor and ceiling (also along the
    side walls), and both goals: a mouth 1786 wide x 642.8 tall cut into the
    back wall and an 880 deep box behind it;
  * about 8k triangles, every triangle's normal pointing into the play space.

.cmf layout (src/CollisionMeshFile/CollisionMeshFile.cpp:11-61): int32
numTris, int32 numVerts, numTris x int32[3] vertex indices, numVerts x
float32[3] vertices, in Bullet units (uu / 50, no scaling on load,
CollisionMeshFile.cpp:63-73).
"""
import math
import struct

import numpy as np

EXTENT_X = 4096.0
EXTENT_Y = 5120.0
HEIGHT = 2044.0
FILLET_R = 256.0
BEVEL = 1152.0
GOAL_HALF_W = 1786.0 / 2
GOAL_H = 642.8
GOAL_DEPTH = 880.0
UU_TO_BT = 1.0 / 50.0

FILLET_SEGS = 8      # arc segments per fillet
WALL_SEGS = 6        # vertical wall segments between the fillets
EDGE_STEP = 164.0    # target edge length along the walls


def _outline():
    """Arena outline (counter-clockwise seen from above) as a list of corners."""
    X, Y, B = EXTENT_X, EXTENT_Y, BEVEL
    return [
        (X, -(Y - B)), (X, Y - B),          # +x side wall
        (X - B, Y),                         # bevel to the +y back wall
        (-(X - B), Y),                      # +y back wall
        (-X, Y - B),                        # bevel
        (-X, -(Y - B)),                     # -x side wall
        (-(X - B), -Y),                     # bevel
        (X - B, -Y),                        # -y back wall
    ]


def _profile():
    """(inset d, height z) samples of the wall cross-section, bottom to top."""
    pts = []
    for k in range(FILLET_SEGS + 1):
        t = (math.pi / 2) * k / FILLET_SEGS
        pts.append((FILLET_R - FILLET_R * math.sin(t), FILLET_R - FILLET_R * math.cos(t)))
    z0, z1 = FILLET_R, HEIGHT - FILLET_R
    # goal crossbar height must be a profile sample so the mouth is cut cleanly
    zs = list(np.linspace(z0, z1, WALL_SEGS + 1)[1:-1]) + [GOAL_H]
    for z in sorted(zs):
        pts.append((0.0, float(z)))
    for k in range(FILLET_SEGS + 1):
        t = (math.pi / 2) * k / FILLET_SEGS
        pts.append((FILLET_R - FILLET_R * math.cos(t), z1 + FILLET_R * math.sin(t)))
    return pts


def _offset_polygon(corners, d):
    """Inward offset of a convex CCW polygon by distance d (mitred corners)."""
    n = len(corners)
    P = np.array(corners, dtype=np.float64)
    normals = []
    for i in range(n):
        a, b = P[i], P[(i + 1) % n]
        e = b - a
        e /= np.linalg.norm(e)
        normals.append(np.array([-e[1], e[0]]))  # left normal = inward for CCW
    out = []
    for i in range(n):
        n0 = normals[(i - 1) % n]
        n1 = normals[i]
        m = n0 + n1
        m /= np.linalg.norm(m)
        scale = d / max(np.dot(m, n1), 1e-9)
        out.append(P[i] + m * scale)
    return out


def build_mesh():
    """Returns (vertices [V,3] float32 in uu, triangles [T,3] int32)."""
    corners = _outline()
    prof = _profile()
    n_edges = len(corners)
    rings = [_offset_polygon(corners, d) for d, _ in prof]
    verts = []
    tris = []
    vid = {}

    def vertex(p):
        key = (round(p[0], 4), round(p[1], 4), round(p[2], 4))
        if key not in vid:
            vid[key] = len(verts)
            verts.append(p)
        return vid[key]

    interior = np.array([0.0, 0.0, HEIGHT / 2])

    def add_quad(a, b, c, d, inward_ref):
        # a-b-c-d in order around the quad; orient both triangles towards inward_ref
        for tri in ((a, b, c), (a, c, d)):
            p = [np.asarray(verts[i]) for i in tri]
            nrm = np.cross(p[1] - p[0], p[2] - p[0])
            if np.linalg.norm(nrm) < 1e-9:
                continue
            cen = (p[0] + p[1] + p[2]) / 3
            if np.dot(nrm, inward_ref - cen) < 0:
                tri = (tri[0], tri[2], tri[1])
            tris.append(tri)

    for e in range(n_edges):
        c0, c1 = corners[e], corners[(e + 1) % n_edges]
        is_side = abs(c0[0] - c1[0]) < 1e-6 and abs(abs(c0[0]) - EXTENT_X) < 1e-6
        is_back = abs(c0[1] - c1[1]) < 1e-6 and abs(abs(c0[1]) - EXTENT_Y) < 1e-6
        length = math.dist(c0, c1)
        # split the back walls exactly at the goal posts
        if is_back:
            xa, xb = c0[0], c1[0]
            cuts = sorted({xa, xb, -GOAL_HALF_W, GOAL_HALF_W})
            cuts = [x for x in cuts if min(xa, xb) - 1e-6 <= x <= max(xa, xb) + 1e-6]
            ts = []
            for k in range(len(cuts) - 1):
                seg = cuts[k + 1] - cuts[k]
                m = max(1, int(round(seg / EDGE_STEP)))
                ts += [(cuts[k] + seg * j / m - xa) / (xb - xa) for j in range(m)]
            ts.append(1.0)
            ts = sorted(set(round(t, 9) for t in ts))
        else:
            m = max(1, int(round(length / EDGE_STEP)))
            ts = [j / m for j in range(m + 1)]
        for k in range(len(prof) - 1):
            (d0, z0), (d1, z1) = prof[k], prof[k + 1]
            if is_side and d0 == 0.0 and d1 == 0.0:
                continue  # flat side wall: the x = +-4096 planes cover it
            r0a, r0b = rings[k][e], rings[k][(e + 1) % n_edges]
            r1a, r1b = rings[k + 1][e], rings[k + 1][(e + 1) % n_edges]
            for j in range(len(ts) - 1):
                ta, tb = ts[j], ts[j + 1]
                pa0 = r0a + (r0b - r0a) * ta
                pb0 = r0a + (r0b - r0a) * tb
                pa1 = r1a + (r1b - r1a) * ta
                pb1 = r1a + (r1b - r1a) * tb
                if is_back:
                    xm = 0.5 * (pa0[0] + pb0[0])
                    zm = 0.5 * (z0 + z1)
                    if abs(xm) < GOAL_HALF_W and zm < GOAL_H:
                        continue  # goal mouth
                q = [vertex((float(pa0[0]), float(pa0[1]), float(z0))),
                     vertex((float(pb0[0]), float(pb0[1]), float(z0))),
                     vertex((float(pb1[0]), float(pb1[1]), float(z1))),
                     vertex((float(pa1[0]), float(pa1[1]), float(z1)))]
                add_quad(*q, interior)

    # goals: side walls, back wall and ceiling of an 880 deep box behind each mouth
    def grid_quads(origin, du, dv, nu, nv, inward_ref):
        origin, du, dv = (np.asarray(v, dtype=np.float64) for v in (origin, du, dv))
        for i in range(nu):
            for j in range(nv):
                p00 = origin + du * (i / nu) + dv * (j / nv)
                p10 = origin + du * ((i + 1) / nu) + dv * (j / nv)
                p11 = origin + du * ((i + 1) / nu) + dv * ((j + 1) / nv)
                p01 = origin + du * (i / nu) + dv * ((j + 1) / nv)
                q = [vertex(tuple(float(x) for x in p)) for p in (p00, p10, p11, p01)]
                add_quad(*q, inward_ref)

    for s in (1.0, -1.0):
        y0 = s * EXTENT_Y
        y1 = s * (EXTENT_Y + GOAL_DEPTH)
        inside = np.array([0.0, s * (EXTENT_Y + GOAL_DEPTH / 2), GOAL_H / 2])
        nw, nd, nh = 8, 4, 3
        grid_quads((-GOAL_HALF_W, y0, 0), (0, y1 - y0, 0), (0, 0, GOAL_H), nd, nh, inside)  # left
        grid_quads((GOAL_HALF_W, y0, 0), (0, y1 - y0, 0), (0, 0, GOAL_H), nd, nh, inside)   # right
        grid_quads((-GOAL_HALF_W, y1, 0), (2 * GOAL_HALF_W, 0, 0), (0, 0, GOAL_H), nw, nh, inside)  # back
        grid_quads((-GOAL_HALF_W, y0, GOAL_H), (2 * GOAL_HALF_W, 0, 0), (0, y1 - y0, 0), nw, nd, inside)  # roof

    V = np.array(verts, dtype=np.float32)
    T = np.array(tris, dtype=np.int32)
    return V, T


def to_cmf(verts_uu, tris):
    """Serialise to .cmf bytes (vertices in Bullet units)."""
    v_bt = (np.asarray(verts_uu, dtype=np.float64) * UU_TO_BT).astype(np.float32)
    out = struct.pack("<ii", len(tris), len(v_bt))
    out += np.ascontiguousarray(tris, dtype="<i4").tobytes()
    out += np.ascontiguousarray(v_bt, dtype="<f4").tobytes()
    return out


def from_cmf(data):
    """Parse .cmf bytes -> (vertices [V,3] float32 in Bullet units, tris [T,3] int32)."""
    nt, nv = struct.unpack_from("<ii", data, 0)
    if nt <= 0 or nv <= 0 or max(nt, nv) > 1000 * 1000:
        raise ValueError(f"bad .cmf counts ({nt}, {nv})")
    need = 8 + nt * 12 + nv * 12
    if len(data) < need:
        raise ValueError(".cmf data truncated")
    tris = np.frombuffer(data, dtype="<i4", count=nt * 3, offset=8).reshape(nt, 3)
    verts = np.frombuffer(data, dtype="<f4", count=nv * 3, offset=8 + nt * 12).reshape(nv, 3)
    if tris.min() < 0 or tris.max() >= nv:
        raise ValueError("bad .cmf triangle vertex index")
    return verts, tris


_CACHE = None
_SPLIT_CACHE = None

# The pieces of soccar_split_cmfs(), in load order.
SPLIT_PIECES = ("blue goal", "orange goal", "-x -y", "+x -y", "-x +y", "+x +y")


def split_mesh(V, T):
    """Split the arena's triangles into SPLIT_PIECES by centroid: the two goal
    boxes (behind the back walls), then the four quadrants of the rest.  Each
    piece keeps its triangles in their original order and its used vertices
    in theirs.  Returns a list of (verts, tris)."""
    V = np.asarray(V, dtype=np.float32)
    T = np.asarray(T, dtype=np.int32)
    c = V[T].mean(axis=1)
    goal_b = c[:, 1] < -EXTENT_Y + 1.0
    goal_o = c[:, 1] > EXTENT_Y - 1.0
    rest = ~(goal_b | goal_o)
    nx, ny = c[:, 0] < 0, c[:, 1] < 0
    masks = [goal_b, goal_o, rest & nx & ny, rest & ~nx & ny, rest & nx & ~ny, rest & ~nx & ~ny]
    out = []
    for m in masks:
        t = T[m]
        used = np.unique(t)
        remap = np.full(len(V), -1, dtype=np.int32)
        remap[used] = np.arange(len(used), dtype=np.int32)
        out.append((V[used], remap[t]))
    return out


def soccar_split_cmfs():
    """The synthetic Soccar arena as six .cmf files (SPLIT_PIECES), the way
    the reference's arena comes as a list of mesh files (src/RocketSim.cpp:
    20-28,92-185); each becomes its own static body."""
    global _SPLIT_CACHE
    if _SPLIT_CACHE is None:
        V, T = build_mesh()
        _SPLIT_CACHE = [to_cmf(v, t) for v, t in split_mesh(V, T)]
    return list(_SPLIT_CACHE)


# -- the synthetic Hoops arena -------------------------------------------------
# The Hoops arena's planes (floor, ceiling, side and end walls at
# x = +-8900/3, y = +-3581, z = 1820 uu; src/RLConst.h:18-20) are added by the
# arena itself.  Its meshes here: 45 degree bevels in the four vertical
# corners, and per side a hoop -- a rim band and, below it, a funnel net of
# exactly 505 vertices (the count by which the reference tells a net mesh,
# src/Sim/Arena/Arena.cpp:1008-1018).  A ball dropped through a rim comes to
# rest in the net low inside the hoop's scoring footprint
# (Arena.cpp:759-768,902-915).

HOOPS_EXTENT_X = 8900.0 / 3
HOOPS_EXTENT_Y = 3581.0
HOOPS_HEIGHT = 1820.0
HOOPS_BEVEL = 600.0
HOOP_Y = 2900.0          # hoop centre |y|
HOOP_R = 600.0           # rim radius
RIM_Z0, RIM_Z1 = 330.0, 380.0
NET_BOTTOM_Z = 30.0
NET_BOTTOM_SCALE = 0.4
HOOP_SEGS = 24
NET_RINGS = 21           # 21 rings of 24 + the bottom centre = 505 vertices

_HOOPS_CACHE = None


def _facing(V, tris, inside):
    """Orient every triangle so its normal points towards inside(centroid)."""
    out = []
    for t in tris:
        a, b, c = (np.asarray(V[i], dtype=np.float64) for i in t)
        n = np.cross(b - a, c - a)
        if np.dot(n, inside((a + b + c) / 3) - (a + b + c) / 3) < 0:
            t = (t[0], t[2], t[1])
        out.append(t)
    return out


def _hoop(sign):
    """(rim (verts, tris), net (verts, tris)) of the hoop at y = sign * HOOP_Y."""
    cy = sign * HOOP_Y
    ang = [2 * math.pi * k / HOOP_SEGS for k in range(HOOP_SEGS)]
    rim_v = [(HOOP_R * math.cos(a), cy + HOOP_R * math.sin(a), z) for z in (RIM_Z0, RIM_Z1) for a in ang]
    rim_t = []
    for k in range(HOOP_SEGS):
        k2 = (k + 1) % HOOP_SEGS
        rim_t += [(k, k2, HOOP_SEGS + k2), (k, HOOP_SEGS + k2, HOOP_SEGS + k)]
    axis = lambda p: np.array([0.0, cy, p[2]])
    rim_t = _facing(rim_v, rim_t, axis)
    net_v = []
    for i in range(NET_RINGS):
        f = i / (NET_RINGS - 1)
        z = RIM_Z0 + (NET_BOTTOM_Z - RIM_Z0) * f
        r = HOOP_R * (1 + (NET_BOTTOM_SCALE - 1) * f)
        net_v += [(r * math.cos(a), cy + r * math.sin(a), z) for a in ang]
    net_v.append((0.0, cy, NET_BOTTOM_Z))
    net_t = []
    for i in range(NET_RINGS - 1):
        for k in range(HOOP_SEGS):
            k2 = (k + 1) % HOOP_SEGS
            a, b = i * HOOP_SEGS + k, i * HOOP_SEGS + k2
            c, d = a + HOOP_SEGS, b + HOOP_SEGS
            net_t += [(a, b, d), (a, d, c)]
    last, centre = (NET_RINGS - 1) * HOOP_SEGS, len(net_v) - 1
    net_t += [(last + k, last + (k + 1) % HOOP_SEGS, centre) for k in range(HOOP_SEGS)]
    net_t = _facing(net_v, net_t, lambda p: np.array([0.0, cy, p[2] + 200.0]))
    assert len(net_v) == 505
    return (rim_v, rim_t), (net_v, net_t)


def _hoops_shell():
    """The four corner bevels, facing the arena's centre."""
    V, T = [], []
    for sx in (-1, 1):
        for sy in (-1, 1):
            p = (sx * HOOPS_EXTENT_X, sy * (HOOPS_EXTENT_Y - HOOPS_BEVEL))
            q = (sx * (HOOPS_EXTENT_X - HOOPS_BEVEL), sy * HOOPS_EXTENT_Y)
            base = len(V)
            V += [(p[0], p[1], 0.0), (q[0], q[1], 0.0), (q[0], q[1], HOOPS_HEIGHT), (p[0], p[1], HOOPS_HEIGHT)]
            T += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
    return V, _facing(V, T, lambda p: np.array([0.0, 0.0, p[2]]))


def hoops_cmfs():
    """The synthetic Hoops arena as five .cmf files: the corner bevels, the
    two rims, then the two nets (deterministic, cached)."""
    global _HOOPS_CACHE
    if _HOOPS_CACHE is None:
        (rim_b, net_b), (rim_o, net_o) = _hoop(-1), _hoop(1)
        parts = [_hoops_shell(), rim_b, rim_o, net_b, net_o]
        _HOOPS_CACHE = [to_cmf(np.array(v, dtype=np.float32), np.array(t, dtype=np.int32)) for v, t in parts]
    return list(_HOOPS_CACHE)


def soccar_cmf():
    """The synthetic Soccar arena as .cmf bytes (deterministic, cached)."""
    global _CACHE
    if _CACHE is None:
        V, T = build_mesh()
        _CACHE = to_cmf(V, T)
    return _CACHE


if __name__ == "__main__":
    V, T = build_mesh()
    print("verts", len(V), "tris", len(T))
    print("bounds", V.min(0), V.max(0))
