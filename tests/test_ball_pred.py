"""Ball prediction: the batched BallPredTracker (rsg_ball_pred_*, include/rsg.h)
against the reference's BallPredTracker (src/Sim/BallPredTracker/
BallPredTracker.cpp) run in the oracle.

CPU:
  * the tracker update the GPU kernel runs (rocketsim_amd/csrc/rsg_ball_pred.h),
    built for the host with the reference's SSE normalisation, reproduces the
    reference tracker bit for bit: the full prediction from a ball thrown at
    the floor, walls, corners and goals, and a sequence of UpdatePredManual
    calls that takes every branch (reuse, no change, re-prediction outside the
    Matches margins, beyond the prediction, negative ticks);
  * the index rule of GetBallStateForTime, checked against the reference;
  * argument errors of rsg_ball_pred_create (no GPU work).
GPU:
  * the committed reference runs (tests/golden/soccar_ball_pred.npz): the same
    branches and ring order (the tick counts since the last SetState, exact)
    and states within the one-tick parity tolerances, plus one-tick
    teacher-forced parity along the GPU's own predictions when the oracle is
    present;
  * UpdatePredFromArena on a stepping batch equals UpdatePredManual with the
    batch's balls and tick counts, bit for bit;
  * GetBallStateForTime and the empty-prediction error.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import ball_pred_cases
import parity
from rocketsim_amd.state import BallState, MODE_HEATSEEKER, MODE_SOCCAR, MODE_THE_VOID

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIXTURE = os.path.join(HERE, "golden", "soccar_ball_pred.npz")


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def expected_index(pred_time, size, tick_time):
    """RS_CLAMP(predTime / tickTime, 0, size - 1) evaluated in float32 and
    truncated (BallPredTracker.cpp:77, Framework.h:47-50)."""
    x = np.float32(pred_time) / np.float32(tick_time)
    x = x if x > 0 else np.float32(0)
    hi = np.float32(size - 1)
    return int(x if x < hi else hi)


# ---------------------------------------------------------------------------
# CPU

@pytest.fixture(scope="module")
def host_pred(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("bpred") / "bphost.so")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "ball_pred_host.cpp"),
                        "-o", so], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = C.CDLL(so)
    L.bp_create.restype = C.c_void_p
    L.bp_create.argtypes = [C.c_char_p, C.c_uint64, C.c_int, C.c_int, C.c_float]
    L.bp_destroy.argtypes = [C.c_void_p]
    L.bp_update.argtypes = [C.c_void_p, C.c_int, C.POINTER(BallState), C.POINTER(C.c_int32),
                            C.POINTER(C.c_uint64)]
    L.bp_get.argtypes = [C.c_void_p, C.c_int, C.POINTER(BallState)]
    return L


def _host_preds(L, h, n, T):
    out = []
    for i in range(n):
        buf = (BallState * T)()
        size = L.bp_get(h, i, buf)
        out.append([bytes(buf[t]) for t in range(size)])
    return out


@pytest.mark.parametrize("mode", [MODE_SOCCAR, MODE_THE_VOID])
def test_host_tracker_matches_reference_bit_for_bit(host_pred, mode):
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built")
    from rocketsim_amd import arena_mesh
    L = host_pred
    n, T = (10, 120) if mode == MODE_SOCCAR else (3, 60)
    states = ball_pred_cases.initial_states(ref, n, seed=7 + mode, mode=mode)
    run = ball_pred_cases.reference_run(ref, states, T, mode=mode)
    cmf = arena_mesh.soccar_cmf() if mode == MODE_SOCCAR else None
    h = L.bp_create(cmf, len(cmf) if cmf else 0, n, T, 120.0)
    assert h
    try:
        L.bp_update(h, 0, (BallState * n)(*run["init"]), None, None)
        want = [[bytes(s) for s in p] for p in run["pred0"]]
        assert _host_preds(L, h, n, T) == want
        for name, curs, ks, preds in run["steps"]:
            L.bp_update(h, 1, (BallState * n)(*curs), (C.c_int32 * n)(*ks), None)
            want = [[bytes(s) for s in p] for p in preds]
            assert _host_preds(L, h, n, T) == want, name
    finally:
        L.bp_destroy(h)


def test_state_for_time_index_rule_matches_reference():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built")
    T = 50
    A = ref.RefArena(MODE_THE_VOID)
    b = A.get_ball()
    b.pos[:] = [0.0, 0.0, 500.0]
    b.tick_count_since_update = 0
    A.set_ball(b)
    tr = ref.RefBallPredTracker(A, T)
    try:
        for t in (-1.0, 0.0, 0.004, 1 / 120, 0.0084, 0.1, 0.25, 0.4, 0.41, 100.0, float("inf"), float("-inf")):
            got = tr.get_ball_state_for_time(t).tick_count_since_update   # predData[i] has i ticks since the set
            assert got == expected_index(t, T, 1 / 120), t
    finally:
        tr.close()
        A.close()


def test_create_rejects_bad_arguments():
    from rocketsim_amd import native
    L = native.lib()
    for args in ((MODE_THE_VOID, 0, 10, 120.0), (MODE_THE_VOID, 4, 0, 120.0), (MODE_THE_VOID, 4, 10, 500.0),
                 (7, 4, 10, 120.0)):
        assert not L.rsg_ball_pred_create(*args)
        assert L.rsg_last_error()


# ---------------------------------------------------------------------------
# GPU

def _check_states(ref_states, got_states, tag):
    """Tick counts since the last SetState exactly (they tell which branch
    ran and where the ring starts); states within the one-tick parity
    tolerances, which a prediction of this length keeps."""
    assert len(ref_states) == len(got_states), tag
    for t, (a, b) in enumerate(zip(ref_states, got_states)):
        assert a.tick_count_since_update == b.tick_count_since_update, (tag, t)
    worst = 0.0
    for a, b in zip(ref_states, got_states):
        for k in ("pos", "vel", "ang_vel"):
            x = np.asarray(getattr(a, k)[:], np.float64)
            y = np.asarray(getattr(b, k)[:], np.float64)
            worst = max(worst, float(np.max(np.abs(x - y) / np.maximum(np.abs(x), parity.REL_FLOOR[k]))))
    return worst


@pytest.mark.gpu
def test_gpu_tracker_matches_reference_runs():
    from rocketsim_amd.ball_pred import BallPredTracker
    run, T = ball_pred_cases.load(FIXTURE)
    n = len(run["init"])
    tr = BallPredTracker(mode=MODE_SOCCAR, n=n, num_pred_ticks=T)
    try:
        tr.force_update_all_pred(run["init"])
        worst = {}
        w = 0.0
        for i in range(n):
            w = max(w, _check_states(run["pred0"][i], tr.pred_data(i), ("force", i)))
        worst["force"] = w
        for name, curs, ks, preds in run["steps"]:
            tr.update_pred_manual(curs, ks)
            w = 0.0
            for i in range(n):
                w = max(w, _check_states(preds[i], tr.pred_data(i), (name, i)))
            worst[name] = w
        print("worst relative error per update:", worst)
        assert max(worst.values()) <= parity.REL_TOL, worst
    finally:
        tr.close()


@pytest.mark.gpu
def test_gpu_predictions_one_tick_parity_with_oracle():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    from rocketsim_amd.ball_pred import BallPredTracker
    n, T = 40, 240
    states = ball_pred_cases.initial_states(ref, n, seed=99)
    tr = BallPredTracker(mode=MODE_SOCCAR, n=n, num_pred_ticks=T)
    A = ref.RefArena(MODE_SOCCAR)
    try:
        tr.force_update_all_pred(states)
        bad = total = 0
        for i in range(n):
            pd = tr.pred_data(i)
            for t in range(T - 1):
                s = BallState.from_buffer_copy(bytes(pd[t]))
                A.set_ball(s)
                A.step(1)
                r = A.get_ball()
                r.bp_cell, r.bp_seq = 0, 0
                total += 1
                if parity.compare_ball(r, pd[t + 1]):
                    bad += 1
        print(f"one-tick parity along GPU predictions: {bad}/{total} outside tolerance")
        assert bad <= parity.CHAOTIC_FRACTION * total
    finally:
        A.close()
        tr.close()


@pytest.mark.gpu
def test_gpu_update_from_batch_equals_manual():
    from rocketsim_amd.ball_pred import BallPredTracker
    from rocketsim_amd.batch import ArenaBatch
    n, T = 64, 90
    b = ArenaBatch.uniform(n, 2, 2, seed=5, synth_controls=True, auto_reset=True)
    p1 = BallPredTracker(b, T)                 # constructor: UpdatePredFromArena
    p2 = BallPredTracker(mode=MODE_SOCCAR, n=n, num_pred_ticks=T)
    last = [0] * n
    try:
        def manual():
            ticks = b.get_tick_counts()
            p2.update_pred_manual(list(b.get_balls()), [int(t) - l for t, l in zip(ticks, last)])
            for a in range(n):
                last[a] = int(ticks[a])
        manual()
        assert bytes(p1.pred_data_all()) == bytes(p2.pred_data_all())
        for k in (1, 8, 0, 30, 200):
            b.step(k)
            p1.update_pred_from_arena(b)
            manual()
            assert bytes(p1.pred_data_all()) == bytes(p2.pred_data_all()), k
    finally:
        p1.close()
        p2.close()
        b.close()


@pytest.mark.gpu
def test_gpu_state_for_time_and_empty_error():
    from rocketsim_amd.ball_pred import BallPredTracker
    T = 50
    tr = BallPredTracker(mode=MODE_THE_VOID, n=3, num_pred_ticks=T)
    try:
        with pytest.raises(RuntimeError, match="empty"):
            tr.get_ball_state_for_time(0, 0.5)
        s = BallState()
        s.pos[:] = [0.0, 0.0, 500.0]
        s.rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        tr.force_update_all_pred([s, s, s])
        for t in (-1.0, 0.0, 0.004, 1 / 120, 0.0084, 0.1, 0.25, 0.4, 0.41, 100.0, float("inf")):
            got = tr.get_ball_state_for_time(1, t)
            assert got.tick_count_since_update == expected_index(t, T, 1 / 120), t
        v = tr.view()
        assert v.n == 3 and v.num_pred_ticks == T and v.traj and v.head and v.size
        assert not math.isnan(tr.pred_data(2)[T - 1].pos[2])
    finally:
        tr.close()


@pytest.mark.gpu
def test_gpu_heatseeker_predictions_one_tick_parity_with_oracle():
    """A Heatseeker tracker predicts the seeking ball: the reference's
    tracker makes its arena in the source arena's mode (BallPredTracker.cpp:7)
    and carries the seeking state in each BallState.  Each GPU prediction,
    set on the oracle and stepped one tick, gives the next one, seeking words
    included.  A seeking turn can land one rotator step away from the
    reference's (the last-place difference of atan2f, as in the Heatseeker
    parity test), so up to 1 % of ticks may differ."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    import random
    from rocketsim_amd.ball_pred import BallPredTracker
    n, T = 24, 240
    rng = random.Random(5)
    A = ref.RefArena(MODE_HEATSEEKER)
    states = []
    for _ in range(n):
        s = A.get_ball()
        d = rng.choice([-1.0, 1.0])
        s.pos[:] = [rng.uniform(-3000, 3000), d * rng.uniform(0, 4000), rng.uniform(200, 1500)]
        s.vel[:] = [rng.gauss(0, 800), d * rng.uniform(1000, 3000), rng.gauss(0, 400)]
        s.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
        s.hs_y_target_dir, s.hs_cur_target_speed = d, rng.choice([2900.0, 3500.0, 4600.0])
        s.hs_time_since_hit = rng.uniform(0, 2)
        states.append(s)
    tr = BallPredTracker(mode=MODE_HEATSEEKER, n=n, num_pred_ticks=T)
    try:
        tr.force_update_all_pred(states)
        bad = total = turned = rounding = 0
        for i in range(n):
            pd = tr.pred_data(i)
            for t in range(T - 1):
                s = BallState.from_buffer_copy(bytes(pd[t]))
                A.set_ball(s)
                A.step(1)
                r = A.get_ball()
                r.bp_cell, r.bp_seq = 0, 0
                total += 1
                if parity.compare_ball(r, pd[t + 1]):
                    # only the seeking direction's rotator rounding may differ
                    if pd[t].hs_y_target_dir != 0 and parity.seeking_direction_rounding(r, pd[t + 1]):
                        rounding += 1
                    else:
                        bad += 1
                turned += 1 if pd[t].hs_y_target_dir * pd[t + 1].hs_y_target_dir < 0 else 0
        print(f"Heatseeker one-tick parity along GPU predictions: {bad}/{total} outside tolerance, "
              f"{rounding} seeking directions one rotator step apart, {turned} turns")
        assert bad == 0 and rounding <= 0.01 * total and turned >= 1
    finally:
        A.close()
        tr.close()


@pytest.mark.gpu
def test_gpu_hoops_predictions_one_tick_parity_with_oracle():
    """A Hoops tracker (BallPredTracker.cpp:7: the source arena's mode) predicts
    in the Hoops arena: its planes, rims and nets.  Each GPU prediction of a
    ball lobbed at a hoop, set on the oracle and stepped one tick, gives the
    next one within the tolerances.  The balls move, so the kickoff launch
    rule (which counts ticks since the ball was set) never applies."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    import random
    from rocketsim_amd.ball_pred import BallPredTracker
    from rocketsim_amd.state import MODE_HOOPS
    n, T = 24, 240
    rng = random.Random(7)
    A = ref.RefArena(MODE_HOOPS)
    states = []
    for _ in range(n):
        s = A.get_ball()
        d = rng.choice([-1.0, 1.0])
        s.pos[:] = [rng.uniform(-600, 600), d * rng.uniform(1500, 2600), rng.uniform(200, 900)]
        s.vel[:] = [rng.gauss(0, 200), d * rng.uniform(300, 1200), rng.uniform(200, 900)]
        s.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
        states.append(s)
    tr = BallPredTracker(mode=MODE_HOOPS, n=n, num_pred_ticks=T)
    try:
        tr.force_update_all_pred(states)
        bad = total = 0
        low = 0
        for i in range(n):
            pd = tr.pred_data(i)
            for t in range(T - 1):
                s = BallState.from_buffer_copy(bytes(pd[t]))
                A.set_ball(s)
                A.step(1)
                r = A.get_ball()
                r.bp_cell, r.bp_seq = 0, 0
                total += 1
                bad += 1 if parity.compare_ball(r, pd[t + 1]) else 0
                low += 1 if pd[t + 1].pos[2] < 270 and abs(abs(pd[t + 1].pos[1]) - 2900) < 600 else 0
        print(f"Hoops one-tick parity along GPU predictions: {bad}/{total} outside tolerance, "
              f"{low} ticks low inside a hoop")
        assert bad <= 0.005 * total and low > 0
    finally:
        A.close()
        tr.close()
