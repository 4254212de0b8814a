"""Heatseeker (the Soccar arena with the seeking ball) against the oracle:
teacher-forced one-tick parity of the HIP kernels with the reference's
Arena::Step in GameMode::HEATSEEKER, and the mode's kickoff.

Scenarios (every `stride`-th tick of each trajectory is a sample):
  * kickoff: cars driven at the kickoff ball, whose first touch starts the
    seeking (Ball::_OnHit, src/Sim/Ball/Ball.cpp:291-299);
  * seeking: a ball already seeking a net at 2900-4600 uu/s, turning towards
    it each tick (Ball::_PreTickUpdate, Ball.cpp:161-206);
  * back wall: a seeking ball about to hit the back wall of its target net,
    which turns it round with an extra impulse (Ball::_OnWorldCollision,
    Ball.cpp:326-358).
Tolerances are those of tests/parity.py; the Heatseeker fields of the ball
(target net, target speed, time since the last speed-up) are compared too.
"""
import math
import os

import pytest

import parity
from rocketsim_amd.state import MODE_HEATSEEKER

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def test_heatseeker_scenarios_exercise_the_mode():
    """The oracle's trajectories behind the GPU test reach every Heatseeker
    branch: seeking starts, speeds up and turns round at a back wall."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_heatseeker(ref, n_traj=9, ticks=90, seed=3, stride=1)
    started = sum(1 for s in samples if s.ball_in.hs_y_target_dir == 0 and s.ball_out.hs_y_target_dir != 0)
    turned = sum(1 for s in samples if s.ball_in.hs_y_target_dir * s.ball_out.hs_y_target_dir < 0)
    seeking = sum(1 for s in samples if s.ball_in.hs_y_target_dir != 0)
    assert seeking > 50 and turned >= 1 and started + turned >= 1


def _one_rotator_step_apart(s, res):
    """Whether the only difference of a sample is the seeking ball's new
    direction, one or two rotator steps apart at equal speed
    (parity.seeking_direction_rounding), every car matching."""
    cars, ball = res[0], res[1]
    if any(parity.compare_car(s.cars_out[j], cars[j]) for j in range(len(s.teams))):
        return False
    return s.ball_in.hs_y_target_dir != 0 and parity.seeking_direction_rounding(s.ball_out, ball)


def _check(samples, results):
    off = [i for i, (s, r) in enumerate(zip(samples, results)) if parity.check_samples([s], [r])[0]]
    rounding = [i for i in off if _one_rotator_step_apart(samples[i], results[i])]
    rest = [i for i in off if i not in rounding]
    print(f"\n{len(samples)} samples: {len(rounding)} seeking directions one rotator step apart, {len(rest)} other")
    _, msgs = parity.check_samples([samples[i] for i in rest], [results[i] for i in rest])
    assert not rest, f"{len(rest)}/{len(samples)} samples off\n" + "\n".join(msgs)
    assert len(rounding) <= 0.01 * len(samples)


@pytest.mark.gpu
def test_heatseeker_fixture_parity():
    samples = parity.load_samples(os.path.join(HERE, "golden", "heatseeker_tf.npz"))
    _check(samples, parity.run_on_gpu(samples, MODE_HEATSEEKER))


@pytest.mark.gpu
def test_heatseeker_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_heatseeker(ref, n_traj=36, ticks=240, seed=21, stride=2)
    _check(samples, parity.run_on_gpu(samples, MODE_HEATSEEKER))


@pytest.mark.gpu
def test_heatseeker_kickoff():
    """Arena::ResetToRandomKickoff in Heatseeker (Arena.cpp:113-223): four
    spawn slots at y = -4620 facing +y (mirrored for orange), full boost that
    never drains, the ball launched from (-1000, +-2220, 92.75) at
    (0, -+65, 650) uu/s, towards a random side."""
    from rocketsim_amd.batch import ArenaBatch
    n = 64
    b = ArenaBatch.uniform(n, 2, 2, mode=MODE_HEATSEEKER)
    try:
        b.reset_kickoff(-1, 5)
        cars, balls = b.get_cars(), b.get_balls()
        slots = {(-1000.0, -4620.0), (1000.0, -4620.0), (-2000.0, -4620.0), (2000.0, -4620.0)}
        sides = set()
        for a in range(n):
            ball = balls[a]
            sy = 1 if ball.pos[1] > 0 else -1
            sides.add(sy)
            assert ball.pos[:] == pytest.approx([-1000, 2220 * sy, 92.75], abs=1e-3)
            assert ball.vel[:] == pytest.approx([0, 65 * sy, 650], abs=1e-3)
            assert (ball.hs_y_target_dir, ball.hs_cur_target_speed, ball.hs_time_since_hit) == (0, 2900, 0)
            for i in range(4):
                c = cars[4 * a + i]
                x, y = c.pos[0], c.pos[1]
                if i >= 2:  # orange: mirrored
                    x, y = -x, -y
                assert (round(x), round(y)) in slots
                assert c.boost == 100
        assert sides == {-1, 1}
        b.step(60)
        assert all(c.boost == 100 for c in b.get_cars())
    finally:
        b.close()
