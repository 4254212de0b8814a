"""Snapshots in the reference's byte format (Arena::Serialize /
Arena::DeserializeNew, src/Sim/Arena/Arena.cpp:546-644) through
rsg_snapshot_* and rsg_batch_{get,set}_snapshot (include/rsg.h).

CPU, against the reference's own serializer in the oracle:
  * the bytes written for an arena equal the reference's, byte for byte
    (Soccar and THE_VOID, 2v2 with random presets, during play with hits and
    demolitions), except the 8 bytes of the ball's DropshotInfo::lastDamageTick,
    which the reference never initializes (Ball.h:45);
  * reading the reference's bytes and writing them again gives them back;
  * the extension block restores every state record exactly;
  * the reference reads the bytes written here back to the same arena;
  * what is not supported (other modes, custom car configs or mutators,
    truncated data) is refused with an error.
GPU:
  * a batch arena's snapshot taken and loaded into another batch resumes
    bit for bit;
  * a reference arena's snapshot loaded into a batch reads back the same
    (positions and velocities to the ulp of a pass through Bullet units).
"""
import ctypes as C
import struct

import pytest

import snapshot_cases
from rocketsim_amd.batch import deserialize_snapshot, serialize_snapshot
from rocketsim_amd.state import ArenaSnapshot, Controls, MODE_HEATSEEKER, MODE_HOOPS, MODE_SOCCAR, MODE_THE_VOID, PRESET_CUSTOM


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


@pytest.fixture(scope="module")
def ref():
    r = _oracle()
    if r is None:
        pytest.skip("oracle library not built")
    return r


def _by_id(snap):
    """The snapshot with its cars in ascending id order."""
    s = ArenaSnapshot.from_buffer_copy(bytes(snap))
    order = sorted(range(snap.n_cars), key=lambda k: snap.car_id[k])
    for dst, k in enumerate(order):
        s.car_id[dst] = snap.car_id[k]
        s.car_team[dst] = snap.car_team[k]
        s.car_preset[dst] = snap.car_preset[k]
        s.car_config[dst] = snap.car_config[k]
        s.controls[dst] = snap.controls[k]
        s.cars[dst] = snap.cars[k]
    return s


def _assert_records_close(x_snap, y_snap):
    """Car and ball records equal, positions and velocities within the ulp
    that a pass through Bullet units and back can move them."""
    pairs = [(x_snap.cars[k], y_snap.cars[k]) for k in range(y_snap.n_cars)] + [(x_snap.ball, y_snap.ball)]
    for x, y in pairs:
        x = type(x).from_buffer_copy(bytes(x))
        for f in ("pos", "vel"):
            assert max(abs(a - b) for a, b in zip(getattr(x, f), getattr(y, f))) <= 1e-3
            setattr(x, f, getattr(y, f))
        assert bytes(x) == bytes(y)


@pytest.mark.parametrize("mode", [MODE_SOCCAR, MODE_HOOPS, MODE_HEATSEEKER, MODE_THE_VOID])
def test_serialize_matches_reference_bytes(ref, mode):
    n = 0
    for A, presets in snapshot_cases.played_arenas(ref, mode, n_arenas=2, ticks=240, seed=11 + mode, snap_every=40):
        rb = A.serialize()
        snap = snapshot_cases.snapshot_from_ref(A, presets, mode)
        ours = serialize_snapshot(snap)
        assert len(ours) > len(rb)
        assert snapshot_cases.mask_uninitialized(ours[:len(rb)], len(rb)) == \
            snapshot_cases.mask_uninitialized(rb, len(rb))
        # reading the reference's bytes and writing them again
        d = deserialize_snapshot(rb)
        assert d.raw_words == 0 and d.n_cars == 4
        again = serialize_snapshot(d)
        assert again[:len(rb)] == snapshot_cases.mask_uninitialized(rb, len(rb))
        # the extension block carries every record: hidden values exactly,
        # positions and velocities in Bullet units (an ulp from uu and back)
        e = deserialize_snapshot(ours)
        assert e.raw_words > 0
        _assert_records_close(e, snap)
        assert serialize_snapshot(e)[len(rb):] == ours[len(rb):]   # the raw records as read
        n += 1
    assert n == 12


def test_snapshots_cover_hits_and_demolitions(ref):
    from rocketsim_amd.state import CF_DEMOED, CF_HIT_VALID
    flags = 0
    for A, presets in snapshot_cases.played_arenas(ref, MODE_SOCCAR, n_arenas=2, ticks=240, seed=11, snap_every=5):
        for i in range(4):
            flags |= A.get_car(i).flags
    assert flags & CF_HIT_VALID and flags & CF_DEMOED


def test_reference_reads_written_bytes(ref):
    for A, presets in snapshot_cases.played_arenas(ref, MODE_SOCCAR, n_arenas=1, ticks=200, seed=5, snap_every=50):
        # The reference's reader keeps every car only when ids ascend in the
        # file (Arena.cpp:596-610 with _AddCarFromPtr, :64-78), so write the
        # cars in id order.
        snap = _by_id(snapshot_cases.snapshot_from_ref(A, presets, MODE_SOCCAR))
        ours = serialize_snapshot(snap)
        B = ref.RefArena.deserialize(ours)
        try:
            assert B.num_cars == 4
            back = _by_id(deserialize_snapshot(B.serialize()))
            want = _by_id(deserialize_snapshot(ours[:len(B.serialize())]))
            # Car/Ball::SetState and GetState scale positions and velocities
            # to Bullet units and back, which can move them by an ulp; the
            # reference's reader resets every car's boost to the spawn amount
            # (Car::_BulletSetup, Car.cpp:291, runs after the state is read)
            for k in range(4):
                x, y = back.cars[k], want.cars[k]
                for f in ("pos", "vel"):
                    assert max(abs(a - b) for a, b in zip(getattr(x, f), getattr(y, f))) <= 1e-3, f
                    setattr(x, f, getattr(y, f))
                assert x.boost == C.c_float(100 / 3).value
                x.boost = y.boost
                # broadphase records are not in the format (a new arena's,
                # from the file's car order)
                x.bp_cell, x.bp_seq = y.bp_cell, y.bp_seq
                assert bytes(x) == bytes(y)
            for f in ("pos", "vel"):
                assert max(abs(a - b) for a, b in zip(getattr(back.ball, f), getattr(want.ball, f))) <= 1e-3
                setattr(back.ball, f, getattr(want.ball, f))
            assert serialize_snapshot(back) == serialize_snapshot(want)
        finally:
            B.close()


def test_deserialize_refuses_unsupported(ref):
    A = ref.RefArena(MODE_SOCCAR)
    A.add_car(0, 0)
    A.step(3)
    rb = A.serialize()
    A.close()
    deserialize_snapshot(rb)   # accepted as is
    bad_mode = rb[:4] + bytes([3]) + rb[5:]                       # GameMode SNOWDAY
    with pytest.raises(RuntimeError, match="mode"):
        deserialize_snapshot(bad_mode)
    as_hoops = rb[:4] + bytes([1]) + rb[5:]                       # HOOPS: 20 pads, not 34
    with pytest.raises(RuntimeError, match="pad count"):
        deserialize_snapshot(as_hoops)
    with pytest.raises(RuntimeError, match="truncated|count"):
        deserialize_snapshot(rb[:len(rb) // 2])
    # car config: any config is read (hitboxPosOffset.x of the first car
    # changed: a custom car); one the kernels cannot simulate is refused
    cfg_at = 4 + 1 + 4 + 8 + 4 + 42 + 1 + 4 + 1 + 4 + (4 + 5 * 4 + 3) + 4 + 4
    custom = deserialize_snapshot(rb[:cfg_at] + struct.pack("<f", 14.0) + rb[cfg_at + 4:])
    assert custom.car_preset[0] == PRESET_CUSTOM and custom.car_config[0].hitbox_pos_offset[0] == 14.0
    size_at = cfg_at + 16
    with pytest.raises(RuntimeError, match="positive"):
        deserialize_snapshot(rb[:size_at] + struct.pack("<f", -1.0) + rb[size_at + 4:])
    # mutator config: any values are read (gravity z here); a config with
    # another field count is refused, as MutatorConfig::Deserialize does
    m_at = len(rb) - snapshot_cases.MUTATOR_BYTES
    g_at = m_at + 2 + 4 + 8
    low_g = rb[:g_at] + struct.pack("<f", -500.0) + rb[g_at + 4:]
    assert deserialize_snapshot(low_g).mutators.gravity[2] == -500.0
    bad_mut = rb[:m_at] + struct.pack("<H", 29) + rb[m_at + 2:]
    with pytest.raises(RuntimeError, match="different version"):
        deserialize_snapshot(bad_mut)


# ---------------------------------------------------------------------------
# GPU

def _controls(rng, n):
    import parity
    return [parity.rand_controls(rng) for _ in range(n)]


@pytest.mark.gpu
def test_batch_snapshot_resumes_bit_for_bit():
    import random
    from rocketsim_amd.batch import ArenaBatch
    rng = random.Random(4)
    n = 16
    a = ArenaBatch.uniform(n, 2, 2, seed=3, auto_reset=True)
    b = ArenaBatch.uniform(n, 2, 2, seed=3, auto_reset=True)
    try:
        for _ in range(6):
            a.set_controls(_controls(rng, 4 * n))
            a.step(20)
        for i in range(n):
            b.deserialize_arena(i, a.serialize_arena(i))
        assert bytes(b.get_cars()) == bytes(a.get_cars())
        assert bytes(b.get_balls()) == bytes(a.get_balls())
        for _ in range(4):
            c = _controls(rng, 4 * n)
            a.set_controls(c)
            b.set_controls(c)
            a.step(25)
            b.step(25)
        assert bytes(b.get_cars()) == bytes(a.get_cars())
        assert bytes(b.get_balls()) == bytes(a.get_balls())
        assert bytes(b.get_pads()) == bytes(a.get_pads())
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_batch_loads_reference_snapshot():
    r = _oracle()
    if r is None:
        pytest.skip("oracle library not built on this machine")
    from rocketsim_amd.batch import ArenaBatch
    for A, presets in snapshot_cases.played_arenas(r, MODE_SOCCAR, n_arenas=1, ticks=120, seed=8, snap_every=60):
        snap = snapshot_cases.snapshot_from_ref(A, presets, MODE_SOCCAR)
        order = A.car_order()
        teams = [A.car_team(i) for i in range(4)]
        b = ArenaBatch([4], teams, presets)
        try:
            b.set_snapshot(0, snap)
            got = b.get_snapshot(0)
            assert got.raw_words > 0
            _assert_records_close(got, snap)
            assert list(got.car_id)[:4] == list(snap.car_id)[:4]
            assert bytes(got.controls) == bytes(snap.controls) and bytes(got.pads) == bytes(snap.pads)
            assert got.tick_count == snap.tick_count
            assert b.get_car_order() == order
        finally:
            b.close()
