"""ArenaConfig other than the defaults.

noBallRot = false (src/Sim/Arena/ArenaConfig/ArenaConfig.h,
Ball.cpp:109): the ball's pose rotates with its angular velocity instead of
keeping its rotation (the batch's `ball_rotation`).  Spinning balls against
the arena and a car, from an oracle arena created with that config: the host
build and the GPU match the oracle, ball rotation included, and the same
samples stepped without rotation do not."""
import math
import random

import pytest

import parity
import tick_host
from rocketsim_amd.state import Controls, MODE_SOCCAR


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def _samples(ref, n_traj=6, ticks=60, seed=91):
    rng = random.Random(seed)
    out = []
    for k in range(n_traj):
        A = ref.RefArena(MODE_SOCCAR, no_ball_rot=False)
        teams, presets = [0, 1], [rng.randint(0, 5), rng.randint(0, 5)]
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 1000))
        b = A.get_ball()
        sx = rng.choice([-1, 1])
        b.pos[:] = [sx * rng.uniform(2500, 3800), rng.uniform(-3000, 3000), rng.uniform(150, 900)]
        b.vel[:] = [sx * rng.uniform(300, 2000), rng.gauss(0, 600), rng.gauss(-200, 400)]
        b.ang_vel[:] = [rng.gauss(0, 4) for _ in range(3)]
        A.set_ball(b)
        if k % 2:
            s = A.get_car(0)
            yaw = math.atan2(b.pos[1] - 0.0, b.pos[0] - 0.0)
            s.pos[:] = [b.pos[0] - 400 * math.cos(yaw), b.pos[1] - 400 * math.sin(yaw), 17.0]
            s.rot[:] = [math.cos(yaw), math.sin(yaw), 0, -math.sin(yaw), math.cos(yaw), 0, 0, 0, 1]
            s.vel[:] = [1500 * math.cos(yaw), 1500 * math.sin(yaw), 0]
            A.set_car(0, s)
        ctrls = [Controls(), Controls()]
        ctrls[0].throttle = 1.0
        for t in range(ticks):
            if t % 2 == 0:
                out.append(parity._one_tick(A, MODE_SOCCAR, teams, presets, ctrls))
            else:
                A.step(1)
        A.close()
    return out


def test_rotating_ball_host_build_matches_oracle():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = _samples(ref)
    L = tick_host.get("sse")
    results, _ = tick_host.run_samples(L, samples, MODE_SOCCAR, ball_rotation=True)
    parity.assert_parity(samples, results, MODE_SOCCAR)
    fixed, _ = tick_host.run_samples(L, samples, MODE_SOCCAR, ball_rotation=False)
    assert parity.check_samples(samples, fixed)[0] > len(samples) // 2, "rotation made no difference"


@pytest.mark.gpu
def test_rotating_ball_gpu_matches_oracle():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = _samples(ref, seed=93)
    results = parity.run_on_gpu(samples, MODE_SOCCAR, ball_rotation=True)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)


# Custom pads along the cars' paths from the kickoff slots (big and small)
PADS = [(-1800.0, -2300.0, 73.0, True), (1800.0, -2300.0, 73.0, False), (-200.0, -3300.0, 70.0, False),
        (200.0, -3300.0, 70.0, True), (0.0, -3900.0, 70.0, False), (-1800.0, 2300.0, 73.0, True),
        (1800.0, 2300.0, 73.0, False), (0.0, 3900.0, 70.0, True)]


def _pad_samples(ref, n_traj=6, ticks=90, seed=95):
    """ArenaConfig::customBoostPads: cars driving from the kickoff over the
    custom pads, some with full boost (they still take a custom pad: the
    reference tests every pad without the grid's early outs)."""
    rng = random.Random(seed)
    out = []
    for k in range(n_traj):
        A = ref.RefArena(MODE_SOCCAR, custom_pads=PADS)
        teams, presets = [0, 1, 0, 1], [rng.randint(0, 5) for _ in range(4)]
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 1000))
        for i in range(4):
            s = A.get_car(i)
            s.boost = 100.0 if (i + k) % 3 == 0 else rng.uniform(0, 60)
            A.set_car(i, s)
        ctrls = []
        for _ in range(4):
            c = Controls()
            c.throttle = 1.0
            c.steer = rng.uniform(-0.3, 0.3)
            ctrls.append(c)
        for t in range(ticks):
            if t % 2 == 0:
                out.append(parity._one_tick(A, MODE_SOCCAR, teams, presets, ctrls))
            else:
                A.step(1)
        A.close()
    return out


def test_custom_pads_host_build_matches_oracle():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = _pad_samples(ref)
    assert all(len(s.pads_in) >= len(PADS) for s in samples)
    L = tick_host.get("sse")
    results, _ = tick_host.run_samples(L, samples, MODE_SOCCAR, custom_pads=PADS)
    parity.assert_parity(samples, results, MODE_SOCCAR)
    used = sum(1 for s in samples for p in s.pads_out[:len(PADS)] if not p.is_active)
    assert used > 0, "no custom pad was taken"
    plain, _ = tick_host.run_samples(L, samples, MODE_SOCCAR)
    assert parity.check_samples(samples, plain)[0] > 0, "custom pads made no difference"


@pytest.mark.gpu
def test_custom_pads_gpu_matches_oracle():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = _pad_samples(ref, seed=97)
    results = parity.run_on_gpu(samples, MODE_SOCCAR, custom_pads=PADS)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)
