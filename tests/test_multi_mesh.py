"""Arenas made of several collision meshes (RocketSim::InitFromMem with a list
of mesh files, src/RocketSim.cpp:92-185; one static body per mesh in list
order, then the planes, src/Sim/Arena/Arena.cpp:990-1068).

The synthetic arena is split into six mesh files
(arena_mesh.soccar_split_cmfs).  Contacts then come from several static
bodies, in the broadphase's static order, and the internal-edge fix-up sees
each piece's own triangle info map, so the split arena steps differently
from the one-mesh arena; the oracle loads the same six files in the same
order.  The CPU test compares the host build of the tick with the oracle,
the GPU tests compare the step kernels with the committed fixture and with
fresh oracle samples."""
import ctypes as C
import os

import numpy as np
import pytest

import parity
import tick_host
from rocketsim_amd import arena_mesh, native
from rocketsim_amd.state import KERNEL_COOP, MODE_SOCCAR

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "soccar_split_tf.npz")


def _oracle_built():
    from oracle import ref
    return ref.available()


def test_split_pieces_cover_the_arena():
    V, T = arena_mesh.build_mesh()
    pieces = arena_mesh.split_mesh(V, T)
    assert len(pieces) == len(arena_mesh.SPLIT_PIECES)
    assert sum(len(t) for _, t in pieces) == len(T)
    # every piece's triangles are original triangles with the same corners
    whole = {tuple(np.round(V[t].ravel(), 3)) for t in T}
    for v, t in pieces:
        assert len(t) > 0
        assert {tuple(np.round(v[x].ravel(), 3)) for x in t} <= whole
    files = arena_mesh.soccar_split_cmfs()
    for f, (v, t) in zip(files, pieces):
        fv, ft = arena_mesh.from_cmf(f)
        assert len(ft) == len(t) and len(fv) == len(v)


def test_init_meshes_refuses_bad_lists():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("librsg.so not built")
    L = native.lib()
    assert L.rsg_init_meshes(None, None, 0) != 0
    assert b"between 1 and 26" in L.rsg_last_error()
    one = arena_mesh.soccar_cmf()
    bufs = [C.create_string_buffer(one, len(one))] * 27
    ptrs = (C.c_void_p * 27)(*[C.cast(b, C.c_void_p) for b in bufs])
    sizes = (C.c_uint64 * 27)(*([len(one)] * 27))
    assert L.rsg_init_meshes(ptrs, sizes, 27) != 0
    bad = (C.c_void_p * 1)(C.cast(C.create_string_buffer(b"\x01\x00", 2), C.c_void_p))
    assert L.rsg_init_meshes(bad, (C.c_uint64 * 1)(2), 1) != 0
    assert b"mesh file 0" in L.rsg_last_error()


def test_split_arena_host_build_matches_oracle(tmp_path):
    if not _oracle_built():
        pytest.skip("oracle library not built on this machine")
    import multi_mesh_case
    samples = multi_mesh_case.generate(str(tmp_path / "split.npz"), seed=8101, n_traj=4)
    L = tick_host.get("sse")
    results, _ = tick_host.run_samples(L, samples, MODE_SOCCAR, cmfs=arena_mesh.soccar_split_cmfs())
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True, cmfs=arena_mesh.soccar_split_cmfs())


def test_split_fixture_host_build():
    samples = parity.load_samples(FIXTURE)
    L = tick_host.get("sse")
    results, _ = tick_host.run_samples(L, samples, MODE_SOCCAR, cmfs=arena_mesh.soccar_split_cmfs())
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True, cmfs=arena_mesh.soccar_split_cmfs())


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,per_block", [(0, 0), (KERNEL_COOP, 4)])
def test_split_arena_fixture_parity(kernel, per_block):
    samples = parity.load_samples(FIXTURE)
    try:
        results = parity.run_on_gpu(samples, MODE_SOCCAR, kernel=kernel, per_block=per_block,
                                    cmfs=arena_mesh.soccar_split_cmfs())
    finally:
        from rocketsim_amd.batch import ensure_meshes
        ensure_meshes(None)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True, cmfs=arena_mesh.soccar_split_cmfs())


@pytest.mark.gpu
def test_split_arena_live_oracle_parity(tmp_path):
    if not _oracle_built():
        pytest.skip("oracle library not built on this machine")
    import multi_mesh_case
    samples = multi_mesh_case.generate(str(tmp_path / "split.npz"), seed=8201, n_traj=8)
    try:
        results = parity.run_on_gpu(samples, MODE_SOCCAR, cmfs=arena_mesh.soccar_split_cmfs())
    finally:
        from rocketsim_amd.batch import ensure_meshes
        ensure_meshes(None)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True, cmfs=arena_mesh.soccar_split_cmfs())
