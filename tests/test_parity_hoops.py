"""Hoops (GameMode::HOOPS) in the synthetic Hoops arena
(rocketsim_amd.arena_mesh.hoops_cmfs): the arena's six planes
(src/Sim/Arena/Arena.cpp:1036-1068), the net meshes that only the ball
collides with (Arena.cpp:1005-1024), the ball's kickoff launch
(Ball.cpp:209-233), the stronger upward extra hit of cars on the ground
(Ball.cpp:272-276), the Hoops boost pads and spawn tables
(RLConst.h:314-427) and the hoop goal test (Arena.cpp:759-768,902-915).

The host build and the GPU kernels are compared with the oracle's
teacher-forced samples (committed fixture and fresh ones); the same
tolerances as Soccar apply (tests/parity.py)."""
import math
import os

import pytest

import parity
import tick_host
from rocketsim_amd import arena_mesh
from rocketsim_amd.state import KERNEL_COOP, MODE_HOOPS, NUM_PADS

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "hoops_tf.npz")


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def test_hoops_arena_meshes():
    """Five files: the corner bevels, two rims, then two nets of exactly 505
    vertices (the reference's net test)."""
    files = arena_mesh.hoops_cmfs()
    counts = [len(arena_mesh.from_cmf(f)[0]) for f in files]
    assert counts[3] == counts[4] == 505
    assert all(c != 505 for c in counts[:3])


def test_hoops_mutator_defaults():
    """MutatorConfig(HOOPS) (MutatorConfig.cpp:8-10): the Hoops ball radius."""
    from rocketsim_amd.state import MutatorConfig
    from rocketsim_amd import native
    import ctypes as C
    m = MutatorConfig()
    native.check(native.lib().rsg_mutator_config_default(MODE_HOOPS, C.byref(m)))
    assert m.ball_radius == pytest.approx(96.3831)


def test_hoops_fixture_host_build():
    samples = parity.load_samples(FIXTURE)
    results, _ = tick_host.run_samples(tick_host.get("sse"), samples, MODE_HOOPS)
    parity.assert_parity(samples, results, MODE_HOOPS, chaotic=True, cmfs=arena_mesh.hoops_cmfs())


def test_hoops_live_oracle_host_build():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_hoops(ref, n_traj=8, ticks=90, seed=61, stride=2)
    results, _ = tick_host.run_samples(tick_host.get("sse"), samples, MODE_HOOPS)
    parity.assert_parity(samples, results, MODE_HOOPS, chaotic=True, cmfs=arena_mesh.hoops_cmfs())


def test_hoops_goal_in_oracle():
    """A ball dropped through a rim settles in the net low enough to score."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    A = ref.RefArena(MODE_HOOPS)
    A.add_car(0, 0)
    b = A.get_ball()
    b.pos[:] = [0.0, 2900.0, 700.0]
    b.vel[:] = [0.0, 0.0, -200.0]
    A.set_ball(b)
    scored = False
    for _ in range(200):
        A.step(1)
        scored = scored or A.is_ball_scored()
    A.close()
    assert scored


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,per_block", [(0, 0), (KERNEL_COOP, 4)])
def test_hoops_fixture_parity(kernel, per_block):
    samples = parity.load_samples(FIXTURE)
    results = parity.run_on_gpu(samples, MODE_HOOPS, kernel=kernel, per_block=per_block)
    parity.assert_parity(samples, results, MODE_HOOPS, chaotic=True, cmfs=arena_mesh.hoops_cmfs())


@pytest.mark.gpu
def test_hoops_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_hoops(ref, n_traj=16, ticks=90, seed=67, stride=1)
    results = parity.run_on_gpu(samples, MODE_HOOPS)
    parity.assert_parity(samples, results, MODE_HOOPS, chaotic=True, cmfs=arena_mesh.hoops_cmfs())


@pytest.mark.gpu
def test_hoops_kickoff_launch_and_goal_on_gpu():
    """Kickoff (Arena.cpp:113-223 with the Hoops tables): cars on Hoops spawn
    slots, 20 active pads, the ball frozen at the centre until the launch
    tick, which then moves it up; a ball dropped into a hoop scores, and with
    auto-reset the batch restarts the kickoff."""
    from rocketsim_amd.batch import ArenaBatch
    b = ArenaBatch([4], [0, 0, 1, 1], [0, 1, 2, 3], mode=MODE_HOOPS, auto_reset=True)
    try:
        b.reset_kickoff(0, 11)
        spawns = {(-1536.0, -3072.0), (1536.0, -3072.0), (-256.0, -2816.0), (256.0, -2816.0), (0.0, -3200.0)}
        for c, team in zip(b.get_cars(), [0, 0, 1, 1]):
            x, y = c.pos[0], c.pos[1]
            key = (x, y) if team == 0 else (-x, -y)
            assert (round(key[0], 3), round(key[1], 3)) in spawns
            assert c.pos[2] == pytest.approx(17.0)
        pads = b.get_pads()
        assert all(p.is_active for p in pads[:20])
        ball = b.get_balls()[0]
        assert list(ball.pos) == pytest.approx([0.0, 0.0, 93.15])
        b.step(32)  # launch on the tick whose pre-tick sees 32 ticks (0.2667 s) since the reset
        assert b.get_balls()[0].vel[2] == 0.0
        b.step(1)
        assert b.get_balls()[0].vel[2] > 900.0
        ball = b.get_balls()[0]
        ball.pos[0], ball.pos[1], ball.pos[2] = 0.0, 2900.0, 700.0
        ball.vel[0], ball.vel[1], ball.vel[2] = 0.0, 0.0, -200.0
        b.set_balls([ball])
        goals = 0
        for _ in range(200):
            b.step(1)
            if b.get_balls()[0].pos[2] == pytest.approx(93.15) and b.get_balls()[0].pos[1] == 0.0:
                goals += 1
                break
        assert goals == 1, "the hoop goal did not reset the kickoff"
    finally:
        b.close()
