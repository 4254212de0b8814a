"""Cases for the ball-prediction tracker tests: initial ball states and a
sequence of tracker updates, each run through the reference's
BallPredTracker (src/Sim/BallPredTracker/BallPredTracker.cpp) in the oracle.

Test infrastructure only.
"""
import ctypes as C
import random

import numpy as np

from rocketsim_amd.state import BallState, MODE_SOCCAR

KINDS = ("drop", "wall", "corner", "goal", "roll")


def initial_states(ref, n, seed, mode=MODE_SOCCAR):
    """n ball states thrown at the floor, a side wall, a corner, into a goal
    or rolling on the floor, as Ball::GetState returns them after SetState
    (so a tracker built on the arena starts from the same state)."""
    rng = random.Random(seed)
    A = ref.RefArena(mode)
    out = []
    for k in range(n):
        kind = KINDS[k % len(KINDS)]
        b = A.get_ball()
        sx, sy = rng.choice([-1, 1]), rng.choice([-1, 1])
        if kind == "drop":
            b.pos[:] = [rng.uniform(-3500, 3500), rng.uniform(-4500, 4500), rng.uniform(150, 1800)]
            b.vel[:] = [rng.gauss(0, 800), rng.gauss(0, 800), rng.gauss(0, 600)]
        elif kind == "wall":
            b.pos[:] = [sx * rng.uniform(3000, 3900), rng.uniform(-3000, 3000), rng.uniform(150, 1500)]
            b.vel[:] = [sx * rng.uniform(500, 3000), rng.gauss(0, 500), rng.gauss(0, 500)]
        elif kind == "corner":
            b.pos[:] = [sx * rng.uniform(2800, 3700), sy * rng.uniform(3800, 4800), rng.uniform(150, 1500)]
            b.vel[:] = [sx * rng.uniform(500, 3000), sy * rng.uniform(500, 3000), rng.gauss(0, 500)]
        elif kind == "goal":
            b.pos[:] = [rng.uniform(-1200, 1200), sy * rng.uniform(4200, 5000), rng.uniform(150, 900)]
            b.vel[:] = [rng.gauss(0, 800), sy * rng.uniform(500, 3000), rng.gauss(0, 500)]
        else:
            b.pos[:] = [rng.uniform(-3000, 3000), rng.uniform(-4000, 4000), 93.15]
            b.vel[:] = [rng.gauss(0, 1000), rng.gauss(0, 1000), 0.0]
        b.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
        b.tick_count_since_update = 0
        A.set_ball(b)
        s = A.get_ball()
        s.bp_cell, s.bp_seq = 0, 0
        out.append(s)
    A.close()
    return out


def _copy(s):
    return BallState.from_buffer_copy(bytes(s))


def _perturbed(pd, k, dpos=(0.0, 0.0, 0.0), dvel=(0.0, 0.0, 0.0)):
    c = _copy(pd[k])
    for j in range(3):
        c.pos[j] += dpos[j]
        c.vel[j] += dvel[j]
    return c


def update_sequence(T):
    """(name, make(pd) -> (current state, ticks since the last update)) for
    UpdatePredManual: reuse paths (the state at that tick, or within the
    Matches margins), full re-predictions (outside the margins, too many or
    negative ticks) and no-change."""
    return [
        ("reuse_5", lambda pd: (_copy(pd[5]), 5)),
        ("same_0", lambda pd: (_copy(pd[0]), 0)),
        ("outside_margin", lambda pd: (_perturbed(pd, 3, dpos=(1.0, 0.0, 0.0)), 3)),
        ("beyond_size", lambda pd: (_copy(pd[1]), T + 2)),
        ("negative", lambda pd: (_copy(pd[1]), -1)),
        ("within_margin_7", lambda pd: (_perturbed(pd, 7, dpos=(0.0, 0.0, 0.5), dvel=(0.0, 0.3, 0.0)), 7)),
        ("reuse_last", lambda pd: (_copy(pd[T - 1]), T - 1)),
        ("at_size", lambda pd: (_copy(pd[T - 1]), T)),
    ]


def reference_run(ref, states, T, with_sequence=True, mode=MODE_SOCCAR):
    """Runs the reference tracker for each initial state (built on an arena
    holding it, which predicts at once), then the update sequence.  Returns
    {"init": [n] the trackers' first states (predData[0], what
    ForceUpdateAllPred started from), "pred0": [n][T] BallState,
    "steps": [(name, curs, ks, preds)]}."""
    arenas, trs = [], []
    for s in states:
        A = ref.RefArena(mode)
        A.set_ball(s)
        arenas.append(A)
        trs.append(ref.RefBallPredTracker(A, T))
    pred0 = [tr.pred_data() for tr in trs]
    steps = []
    if with_sequence:
        for name, make in update_sequence(T):
            curs, ks = [], []
            for tr in trs:
                c, k = make(tr.pred_data())
                curs.append(c)
                ks.append(k)
                tr.update_pred_manual(c, k)
            steps.append((name, curs, ks, [tr.pred_data() for tr in trs]))
    for tr in trs:
        tr.close()
    for A in arenas:
        A.close()
    return {"init": [_copy(p[0]) for p in pred0], "pred0": pred0, "steps": steps}


def as_array(states):
    """BallState records -> raw bytes array (uint8 [..., sizeof])."""
    return np.frombuffer(b"".join(bytes(s) for s in states), dtype=np.uint8).reshape(len(states), -1)


def from_array(arr):
    # records written before the Heatseeker fields were appended are zero-padded
    return [BallState.from_buffer_copy(arr[i].tobytes().ljust(C.sizeof(BallState), b"\0")) for i in range(arr.shape[0])]


def save(path, run, T):
    d = {"T": np.array(T), "init": as_array(run["init"]),
         "pred0": np.stack([as_array(p) for p in run["pred0"]])}
    for j, (name, curs, ks, preds) in enumerate(run["steps"]):
        d[f"s{j}_cur"] = as_array(curs)
        d[f"s{j}_k"] = np.array(ks, dtype=np.int32)
        d[f"s{j}_pred"] = np.stack([as_array(p) for p in preds])
    np.savez_compressed(path, **d)


def load(path):
    z = np.load(path)
    T = int(z["T"])
    init = from_array(z["init"])
    pred0 = [from_array(p) for p in z["pred0"]]
    steps = []
    for j, (name, _) in enumerate(update_sequence(T)):
        if f"s{j}_cur" not in z:
            break
        steps.append((name, from_array(z[f"s{j}_cur"]), [int(k) for k in z[f"s{j}_k"]],
                      [from_array(p) for p in z[f"s{j}_pred"]]))
    return {"init": init, "pred0": pred0, "steps": steps}, T
