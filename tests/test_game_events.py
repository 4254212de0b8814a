"""Game-event tracking: rsg_game_events_* (include/rsg.h), the reference's
GameEventTracker (src/Sim/GameEventTracker/GameEventTracker.cpp:1-166) for
every arena of a batch, against the reference's own tracker in the oracle.

Both trackers are driven through the same recorded states of scripted Soccar
play (tests/game_event_cases.py: shots at goal, some after a team mate's
touch, some into a defender on the goal line, goals), updated every tick and
every 8 ticks; the shot, goal (scorer, passer) and save events must be the
same, tick for tick.
CPU: the tracker update the GPU kernel runs (rsg_game_events.h), built for
the host.  GPU: the kernel through the C ABI.
"""
import collections
import ctypes as C
import os
import subprocess

import pytest

import game_event_cases
from rocketsim_amd.state import (BallState, CarState, Event, EVENT_SAVE, EVENT_SCORER, EVENT_SHOT, GameEventConfig,
                                 MODE_HOOPS, MODE_SOCCAR)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


@pytest.fixture(scope="module")
def reference_runs():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built")
    return {every: game_event_cases.run_reference(ref, n_arenas=6, ticks=1200, update_every=every, seed=40 + every)
            for every in (1, 8)}


def _kinds(events):
    c = collections.Counter()
    for per_arena in events:
        for e in per_arena:
            c[(e.type, e.b != 0)] += 1
    return c


def test_config_defaults():
    from rocketsim_amd import native
    got = GameEventConfig(shot_min_speed=0)
    native.lib().rsg_game_event_config_default(C.byref(got))
    assert bytes(got) == bytes(GameEventConfig())
    # GameEventTrackerConfig's defaults (GameEventTracker.h:10-41)
    want = (1750.0, 0.3, 0.0, 1.0, 2.0, 4.0, 2.0)
    have = (got.shot_min_speed, got.shot_touch_min_delay, got.pred_score_extra_margin, got.shot_event_cooldown,
            got.shot_min_score_time, got.goal_max_touch_time, got.pass_max_touch_time)
    assert have == tuple(C.c_float(x).value for x in want) and got.auto_state_set_detection == 1


def test_scenarios_cover_every_event(reference_runs):
    kinds = _kinds(reference_runs[1][3])
    assert kinds[(EVENT_SHOT, False)] and kinds[(EVENT_SHOT, True)]
    assert kinds[(EVENT_SCORER, False)] and kinds[(EVENT_SCORER, True)]
    assert kinds[(EVENT_SAVE, False)]


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("ge") / "gehost.so")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "game_events_host.cpp"),
                        "-o", so], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = C.CDLL(so)
    L.ge_create.restype = C.c_void_p
    L.ge_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                            C.POINTER(GameEventConfig), C.c_float, C.c_int]
    L.ge_destroy.argtypes = [C.c_void_p]
    L.ge_update.argtypes = [C.c_void_p, C.POINTER(CarState), C.POINTER(BallState), C.POINTER(C.c_uint64),
                            C.POINTER(C.c_int32)]
    L.ge_pop.argtypes = [C.c_void_p, C.POINTER(Event), C.c_int]
    return L


def _by_arena(events, n):
    out = [[] for _ in range(n)]
    for e in events:
        out[e.arena].append(game_event_cases.event_key(e))
    return out


@pytest.mark.parametrize("every", [1, 8])
def test_host_tracker_matches_reference(reference_runs, host_lib, every):
    teams, presets, updates, ref_events = reference_runs[every]
    n = len(ref_events)
    L = host_lib
    cfg = GameEventConfig()
    h = L.ge_create(n, 4, (C.c_uint8 * (4 * n))(*teams), (C.c_uint8 * (4 * n))(*presets), C.byref(cfg), 120.0, MODE_SOCCAR)
    got = []
    try:
        for cars, balls, ticks, orders in updates:
            L.ge_update(h, (CarState * (4 * n))(*[c for per in cars for c in per]), (BallState * n)(*balls),
                        (C.c_uint64 * n)(*ticks), (C.c_int32 * (4 * n))(*[o for per in orders for o in per]))
            buf = (Event * 1024)()
            got.extend(buf[:L.ge_pop(h, buf, 1024)])
    finally:
        L.ge_destroy(h)
    want = [[game_event_cases.event_key(e) for e in per] for per in ref_events]
    assert _by_arena(got, n) == want


@pytest.mark.gpu
@pytest.mark.parametrize("every", [1, 8])
def test_gpu_tracker_matches_reference(reference_runs, every):
    from rocketsim_amd.batch import ArenaBatch
    from rocketsim_amd.game_events import GameEventTracker
    teams, presets, updates, ref_events = reference_runs[every]
    n = len(ref_events)
    b = ArenaBatch([4] * n, teams, presets, mode=MODE_SOCCAR)
    tr = GameEventTracker(b)
    got = []
    try:
        for cars, balls, ticks, orders in updates:
            b.set_cars([c for per in cars for c in per])
            b.set_balls(balls)
            b.set_tick_counts(ticks)
            b.set_car_order([o for per in orders for o in per])
            tr.update(b)
            got.extend(tr.poll())
        assert tr.lost == 0
    finally:
        tr.close()
        b.close()
    want = [[game_event_cases.event_key(e) for e in per] for per in ref_events]
    assert _by_arena(got, n) == want


@pytest.mark.gpu
def test_gpu_tracker_on_a_stepping_batch():
    from rocketsim_amd.batch import ArenaBatch
    from rocketsim_amd.game_events import GameEventTracker
    n = 256
    b = ArenaBatch.uniform(n, 2, 2, seed=9, synth_controls=True, auto_reset=True)
    tr = GameEventTracker(b)
    try:
        events = []
        for _ in range(60):
            b.step(8)
            tr.update(b)
            events.extend(tr.poll())
        assert tr.lost == 0
        ids = set(range(1, 5))
        for e in events:
            assert e.type in (EVENT_SHOT, EVENT_SCORER, EVENT_SAVE)
            assert 0 <= e.arena < n and e.a in ids and (e.b == 0 or e.b in ids)
        tr.reset_persistent_info()
        tr.reset_persistent_info(3)
        with pytest.raises(RuntimeError):
            tr.reset_persistent_info(n)
    finally:
        tr.close()
        b.close()


@pytest.fixture(scope="module")
def hoops_runs():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built")
    return game_event_cases.run_reference(ref, n_arenas=6, ticks=1200, update_every=1, seed=71, mode=MODE_HOOPS)


def test_hoops_scenarios_have_shots_and_goals(hoops_runs):
    """Lobs at the hoops (Arena::IsBallProbablyGoingIn's Hoops branch,
    Arena.cpp:808-885): the reference's tracker reports shots and goals."""
    kinds = _kinds(hoops_runs[3])
    assert kinds[(EVENT_SHOT, False)] + kinds[(EVENT_SHOT, True)] > 0
    assert kinds[(EVENT_SCORER, False)] + kinds[(EVENT_SCORER, True)] > 0


def test_host_tracker_matches_reference_hoops(hoops_runs, host_lib):
    teams, presets, updates, ref_events = hoops_runs
    n = len(ref_events)
    L = host_lib
    cfg = GameEventConfig()
    h = L.ge_create(n, 4, (C.c_uint8 * (4 * n))(*teams), (C.c_uint8 * (4 * n))(*presets), C.byref(cfg), 120.0,
                    MODE_HOOPS)
    got = []
    try:
        for cars, balls, ticks, orders in updates:
            L.ge_update(h, (CarState * (4 * n))(*[c for per in cars for c in per]), (BallState * n)(*balls),
                        (C.c_uint64 * n)(*ticks), (C.c_int32 * (4 * n))(*[o for per in orders for o in per]))
            buf = (Event * 1024)()
            got.extend(buf[:L.ge_pop(h, buf, 1024)])
    finally:
        L.ge_destroy(h)
    want = [[game_event_cases.event_key(e) for e in per] for per in ref_events]
    assert _by_arena(got, n) == want


@pytest.mark.gpu
def test_gpu_tracker_matches_reference_hoops(hoops_runs):
    from rocketsim_amd.batch import ArenaBatch
    from rocketsim_amd.game_events import GameEventTracker
    teams, presets, updates, ref_events = hoops_runs
    n = len(ref_events)
    b = ArenaBatch([4] * n, teams, presets, mode=MODE_HOOPS)
    tr = GameEventTracker(b)
    got = []
    try:
        for cars, balls, ticks, orders in updates:
            b.set_cars([c for per in cars for c in per])
            b.set_balls(balls)
            b.set_tick_counts(ticks)
            b.set_car_order([o for per in orders for o in per])
            tr.update(b)
            got.extend(tr.poll())
        assert tr.lost == 0
    finally:
        tr.close()
        b.close()
    want = [[game_event_cases.event_key(e) for e in per] for per in ref_events]
    assert _by_arena(got, n) == want
