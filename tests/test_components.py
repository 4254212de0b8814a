"""Known-answer tests of single narrowphase and mesh components against the
same components of the oracle (the reference built from source), on the CPU.

The library's component code (rocketsim_amd/csrc/rsg_gjk.h, rsg_boxbox.h,
rsg_mesh.h, rsg_mesh_build.h) is built for the host with the reference's SSE
normalisation (tests/cpp/components_host.cpp) and compared bit for bit with
the reference's on seeded random inputs:
  * GJK with EPA fallback (btGjkPairDetector with btGjkEpaPenetrationDepthSolver,
    btGjkPairDetector.cpp:686-959, btGjkEpa2.cpp): box against triangle and
    box against sphere, separated, touching and penetrating; contact flag,
    normal, point and distance;
  * box-box contacts (btBoxBoxDetector / dBoxBox2, btBoxBoxDetector.cpp:272-767):
    every reported point, normal and depth, in report order;
  * the ray cast against a sphere or a box (btSubsimplexConvexCast,
    btSubSimplexConvexCast.cpp:31-142, accepted as
    btCollisionWorld::rayTestSingleInternal accepts it): fraction and normal;
  * the triangle internal-edge records of the arena mesh
    (btGenerateInternalEdgeInfo, btInternalEdgeUtility.cpp:297-353): three
    edge angles and flags of every triangle;
  * AABB queries over the quantized tree (btQuantizedBvh.cpp:307-325): the
    triangles reached whose AABB overlaps the query, in visit order.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
F = C.POINTER(C.c_float)
I = C.POINTER(C.c_int)
N = 3000


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    so = str(tmp_path_factory.mktemp("components") / "components_host.so")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "components_host.cpp"),
                        "-o", so], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    H = C.CDLL(so)
    R = ref.lib()
    R.rsref_gjk.argtypes = [F, F, C.c_int, C.c_float, F, F, C.c_float, F, F, F, C.POINTER(C.c_int)]
    R.rsref_box_dims.argtypes = [F, F, F]
    R.rsref_box_box.argtypes = [F, F, F, F, F, F, F, C.c_int]
    R.rsref_ray_convex.argtypes = [C.c_int, C.c_float, F, F, F, F, F, F]
    R.rsref_tri_info.argtypes = [C.c_int, F, C.POINTER(C.c_uint32)]
    R.rsref_mesh_query_aabb.argtypes = [F, F, I, C.c_int]
    H.ch_gjk.argtypes = [F, C.c_float, F, C.c_int, C.c_float, F, F, C.c_float, F, F, F]
    H.ch_box_box.argtypes = [F, F, F, F, F, F, F, C.c_int]
    H.ch_ray_convex.argtypes = [C.c_int, C.c_float, F, C.c_float, F, F, F, F, F]
    H.ch_mesh_create.restype = C.c_void_p
    H.ch_mesh_create.argtypes = [C.c_char_p, C.c_uint64]
    H.ch_mesh_destroy.argtypes = [C.c_void_p]
    H.ch_mesh_num_tris.argtypes = [C.c_void_p]
    H.ch_tri_info.argtypes = [C.c_void_p, C.c_int, F, C.POINTER(C.c_uint32)]
    H.ch_mesh_query_aabb.argtypes = [C.c_void_p, F, F, I, C.c_int]
    H.ch_filter_tris.argtypes = [C.c_void_p, F, F, I, C.c_int, I]
    H.ch_mesh_tri_center.argtypes = [C.c_void_p, C.c_int, F]
    return R, H


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(F)


def _rot(rng):
    q = rng.normal(size=4)
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]], dtype=np.float32)


def _xf(r, t):
    return _f(np.concatenate([r.reshape(-1), t]))


def _box_dims(R, half):
    """btBoxShape's implicit half extents and margin for full half extents
    `half` (its constructor lowers the margin for thin boxes)."""
    imp, mg = _f(np.zeros(3)), _f(np.zeros(1))
    R.rsref_box_dims(_p(_f(half)), _p(imp), _p(mg))
    return imp, float(mg[0])


def test_gjk_epa_box_vs_triangle_and_sphere(libs):
    R, H = libs
    rng = np.random.default_rng(1)
    hits = penetrating = 0
    for it in range(N):
        half = _f(rng.uniform(0.3, 1.5, 3))
        imp, mg = _box_dims(R, half)
        xa = _xf(_rot(rng), rng.uniform(-2, 2, 3))
        kind = 0 if it % 4 == 0 else 2
        if kind == 2:
            c = xa[9:] + rng.normal(size=3) * 1.2
            tri = _f((c + rng.normal(size=(3, 3)) * 1.5).reshape(-1))
            xb, rad = _xf(np.eye(3), np.zeros(3)), 0.0
        else:
            tri, rad = _f(np.zeros(9)), 1.8263
            xb = _xf(_rot(rng), xa[9:] + rng.normal(size=3) * 1.5)
        md = (mg + rad + 0.04) ** 2
        out_r = [_f(np.zeros(3)), _f(np.zeros(3)), _f(np.zeros(1))]
        out_h = [_f(np.zeros(3)), _f(np.zeros(3)), _f(np.zeros(1))]
        method = C.c_int()
        a = R.rsref_gjk(_p(half), _p(xa), kind, rad, _p(tri), _p(xb), md, *map(_p, out_r), C.byref(method))
        b = H.ch_gjk(_p(imp), mg, _p(xa), kind, rad, _p(tri), _p(xb), md, *map(_p, out_h))
        assert a == b, f"case {it}: contact {a} vs {b}"
        if a:
            hits += 1
            penetrating += out_r[2][0] < 0
            assert all(x.tobytes() == y.tobytes() for x, y in zip(out_r, out_h)), \
                f"case {it}: ref {out_r} got {out_h}"
    assert hits > N // 3 and penetrating > N // 10


def test_box_box_contacts(libs):
    R, H = libs
    rng = np.random.default_rng(2)
    hits = four = 0
    for it in range(N):
        h0, h1 = _f(rng.uniform(0.3, 1.5, 3)), _f(rng.uniform(0.3, 1.5, 3))
        g0 = _f(sum(_box_dims(R, h0)))
        g1 = _f(sum(_box_dims(R, h1)))
        x0 = _xf(_rot(rng), rng.uniform(-1, 1, 3))
        x1 = _xf(_rot(rng) if it % 3 else np.eye(3), x0[9:] + rng.normal(size=3) * 1.3)
        ref_out = [_f(np.zeros(24)), _f(np.zeros(24)), _f(np.zeros(8))]
        got_out = [_f(np.zeros(24)), _f(np.zeros(24)), _f(np.zeros(8))]
        a = R.rsref_box_box(_p(h0), _p(x0), _p(h1), _p(x1), *map(_p, ref_out), 8)
        b = H.ch_box_box(_p(g0), _p(x0), _p(g1), _p(x1), *map(_p, got_out), 8)
        assert a == b, f"case {it}: {a} vs {b} contacts"
        hits += a > 0
        four += a == 4
        assert all(x.tobytes() == y.tobytes() for x, y in zip(ref_out, got_out)), f"case {it}"
    assert hits > N // 2 and four > N // 20


def test_ray_vs_sphere_and_box(libs):
    R, H = libs
    rng = np.random.default_rng(3)
    hits = 0
    for it in range(N):
        kind = it % 2
        half = _f(rng.uniform(0.3, 1.5, 3))
        imp, mg = _box_dims(R, half)
        xb = _xf(_rot(rng), rng.uniform(-1, 1, 3))
        fr = _f(xb[9:] + rng.normal(size=3) * 3)
        to = _f(xb[9:] + rng.normal(size=3) * 1.0)
        f1, n1, f2, n2 = _f(np.zeros(1)), _f(np.zeros(3)), _f(np.zeros(1)), _f(np.zeros(3))
        a = R.rsref_ray_convex(kind, 1.8263, _p(half), _p(xb), _p(fr), _p(to), _p(f1), _p(n1))
        b = H.ch_ray_convex(kind, 1.8263, _p(imp), mg, _p(xb), _p(fr), _p(to), _p(f2), _p(n2))
        assert a == b, f"case {it}: hit {a} vs {b}"
        if a:
            hits += 1
            assert f1.tobytes() == f2.tobytes() and n1.tobytes() == n2.tobytes(), f"case {it}"
    assert hits > N // 3


@pytest.fixture(scope="module")
def mesh(libs):
    from rocketsim_amd import arena_mesh
    R, H = libs
    cmf = arena_mesh.soccar_cmf()
    m = H.ch_mesh_create(cmf, len(cmf))
    assert m
    yield m
    H.ch_mesh_destroy(m)


def test_triangle_internal_edge_records(libs, mesh):
    R, H = libs
    n = H.ch_mesh_num_tris(mesh)
    assert n > 5000
    with_info = 0
    for t in range(n):
        ra, ga = _f(np.zeros(3)), _f(np.zeros(3))
        rf, gf = C.c_uint32(), C.c_uint32()
        rc = R.rsref_tri_info(t, _p(ra), C.byref(rf))
        H.ch_tri_info(mesh, t, _p(ga), C.byref(gf))
        # the library marks triangles that have an entry in the reference's
        # map with bit 31 (TRI_HAS_INFO); the rest of the word is the map's flags
        has = bool(gf.value >> 31)
        assert has == (rc == 0), f"triangle {t}: entry {rc == 0} vs {has}"
        if rc == 0:
            with_info += 1
            assert ra.tobytes() == ga.tobytes() and rf.value == gf.value & 0x7FFFFFFF, \
                f"triangle {t}: {ra} {rf.value} vs {ga} {gf.value}"
    assert with_info > n // 2


def test_mesh_aabb_queries(libs, mesh):
    R, H = libs
    rng = np.random.default_rng(4)
    nonempty = 0
    buf_r, buf_g, buf_f = (C.c_int * 4096)(), (C.c_int * 4096)(), (C.c_int * 4096)()
    for it in range(N):
        # boxes of car / ball size (Bullet units) anywhere in the arena, or
        # about a random triangle of the mesh
        c = np.array([rng.uniform(-85, 85), rng.uniform(-105, 105), rng.uniform(-1, 42)])
        if it % 2:
            tc = _f(np.zeros(3))
            H.ch_mesh_tri_center(mesh, int(rng.integers(0, H.ch_mesh_num_tris(mesh))), _p(tc))
            c = tc + rng.normal(size=3) * 1.5
        e = rng.uniform(0.5, 4.0, 3)
        mn, mx = _f(c - e), _f(c + e)
        nr = R.rsref_mesh_query_aabb(_p(mn), _p(mx), buf_r, 4096)
        ng = H.ch_mesh_query_aabb(mesh, _p(mn), _p(mx), buf_g, 4096)
        assert 0 <= nr <= 4096 and 0 <= ng <= 4096
        nf = H.ch_filter_tris(mesh, _p(mn), _p(mx), buf_r, nr, buf_f)
        assert list(buf_f[:nf]) == list(buf_g[:ng]), f"case {it}"
        nonempty += ng > 0
    assert nonempty > N // 4
