"""Divergence horizon of free runs: Soccar 2v2 arenas started from the same
state on the GPU and in the oracle (the reference's Arena::Step), then run
side by side with the same controls and no teacher forcing.  For each arena
the first tick at which any car or ball value leaves the one-tick parity
tolerances (tests/parity.py), and the first at which a position is more than
1 uu or a discrete flag differs, are reported.

Free runs of contact-rich play are chaotic: an ulp of difference (the
reference's SSE reciprocal square root against the GPU's) grows through
contacts, so the horizon is a measured quantity.  The test asserts only that
every arena stays within the tolerances for its first 20 ticks (the measured
minimum is about 110).
"""
import random

import numpy as np
import pytest

import parity
from rocketsim_amd.state import Controls, MODE_SOCCAR


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def _settle(A, n):
    """Set the oracle to its own state read through the API until it is a
    fixed point, so that both sides step from the same Bullet-unit values."""
    for _ in range(8):
        before = [bytes(A.get_car(i)) for i in range(n)] + [bytes(A.get_ball())]
        for i in range(n):
            A.set_car(i, A.get_car(i))
        A.set_ball(A.get_ball())
        if [bytes(A.get_car(i)) for i in range(n)] + [bytes(A.get_ball())] == before:
            return


def _pos_err(ref_state, got_state):
    return float(np.max(np.abs(np.asarray(ref_state.pos[:]) - np.asarray(got_state.pos[:]))))


@pytest.mark.gpu
def test_free_run_divergence_horizon():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    from rocketsim_amd.batch import ArenaBatch
    n_arenas, ticks = 24, 1200
    rng = random.Random(77)
    teams = [0, 1, 0, 1]
    arenas = []
    for a in range(n_arenas):
        A = ref.RefArena(MODE_SOCCAR)
        for t in teams:
            A.add_car(t, 0)
        A.reset_kickoff(rng.randint(0, 10000))
        _settle(A, 4)
        arenas.append(A)
    b = ArenaBatch.uniform(n_arenas, 2, 2)
    try:
        b.set_cars([A.get_car(i) for A in arenas for i in range(4)])
        b.set_balls([A.get_ball() for A in arenas])
        b.set_pads([p for A in arenas for p in A.get_pads()])
        b.set_tick_counts([A.tick_count for A in arenas])
        b.set_car_order([o for A in arenas for o in A.car_order()])
        ctrls = [Controls() for _ in range(4 * n_arenas)]
        left = [0] * (4 * n_arenas)
        tol_h = [None] * n_arenas     # first tick outside the parity tolerances
        macro_h = [None] * n_arenas   # first tick with > 1 uu or a flag apart
        for t in range(1, ticks + 1):
            for k in range(4 * n_arenas):
                if left[k] == 0:
                    ctrls[k] = parity.rand_controls(rng)
                    left[k] = rng.randint(2, 60)
                left[k] -= 1
            for a, A in enumerate(arenas):
                for i in range(4):
                    A.set_controls(i, ctrls[4 * a + i])
                A.step(1)
            b.set_controls(ctrls)
            b.step(1)
            cars, balls = b.get_cars(), b.get_balls()
            for a, A in enumerate(arenas):
                if macro_h[a] is not None:
                    continue
                rc = [A.get_car(i) for i in range(4)]
                rb = A.get_ball()
                gc = cars[4 * a:4 * a + 4]
                bad = [m for i in range(4) for m in parity.compare_car(rc[i], gc[i])] + parity.compare_ball(rb, balls[a])
                if bad and tol_h[a] is None:
                    tol_h[a] = t
                far = max([_pos_err(rc[i], gc[i]) for i in range(4)] + [_pos_err(rb, balls[a])]) > 1.0
                flags = any(rc[i].flags != gc[i].flags for i in range(4))
                if far or flags:
                    macro_h[a] = t
            if all(h is not None for h in macro_h):
                break
        def summary(h):
            v = np.array([x if x is not None else ticks + 1 for x in h])
            return f"min {v.min()}, median {int(np.median(v))}, max {v.max()} ticks ({int((v > ticks).sum())} of {len(v)} arenas still together at {ticks})"
        print(f"\nfree-run divergence horizon, {n_arenas} Soccar 2v2 arenas, random controls:")
        print("  first tick outside the one-tick tolerances:", summary(tol_h))
        print("  first tick > 1 uu apart or with a flag apart:", summary(macro_h))
        assert all(h is None or h > 20 for h in tol_h)
    finally:
        b.close()
        for A in arenas:
            A.close()
