"""The launch path bench.py times, checked against single-tick stepping and
against the oracle (the reference's Arena::Step, src/Sim/Arena/Arena.cpp:689-756).

bench.py steps a batch in multi-tick launches (8 ticks per launch, the arena
kept on chip across the launch), with the kernel and arenas per workgroup the
autotuner picks.  Here:
  * step(k) equals k calls of step(1), bit for bit, for every kernel and
    launch shape the bench can run, uniform and uneven batches;
  * the teacher-forced parity sets pass with the bench's launch shapes;
  * one parity launch mixes 1v1, 2v2 and 3v3 Soccar samples (ball, drive,
    melee, walls) in one batch;
  * the 4096-arena 2v2 workload runs on its own for 960 ticks with the seeded
    controls and auto-reset; at three ticks 256 sampled arenas are set into
    the oracle and stepped one tick (teacher-forced), and the GPU steps the
    same states with the bench's launch shape.
"""
import os
import random

import pytest

import parity
from rocketsim_amd.state import KERNEL_AUTO, KERNEL_COOP, KERNEL_LDS, KERNEL_SCRATCH, MODE_SOCCAR

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def _state(b):
    return (b"".join(bytes(c) for c in b.get_cars()), b"".join(bytes(x) for x in b.get_balls()),
            b"".join(bytes(p) for p in b.get_pads()), b.get_tick_counts(), b.get_goals())


def _batch(shape, n, seed):
    from rocketsim_amd.batch import ArenaBatch
    if shape == "mixed":
        rng = random.Random(seed)
        counts, teams, presets = [], [], []
        for _ in range(n):
            u = rng.random()
            per_team = 1 if u < 0.40 else (2 if u < 0.75 else 3)
            counts.append(2 * per_team)
            teams += [0] * per_team + [1] * per_team
            presets += [rng.randrange(6) for _ in range(2 * per_team)]
        return ArenaBatch(counts, teams, presets, mode=MODE_SOCCAR, seed=seed, synth_controls=True, auto_reset=True)
    nb, no = shape
    return ArenaBatch.uniform(n, nb, no, 0, mode=MODE_SOCCAR, seed=seed, synth_controls=True, auto_reset=True)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,per_block,shape", [
    (KERNEL_COOP, 4, (2, 2)),    # the bench's 2v2 shape
    (KERNEL_COOP, 1, (2, 2)),
    (KERNEL_COOP, 2, (3, 3)),
    (KERNEL_COOP, 0, "mixed"),
    (KERNEL_SCRATCH, 64, (2, 2)),  # the plain kernel as bench.py times it
    (KERNEL_SCRATCH, 16, "mixed"),
    (KERNEL_LDS, 0, (1, 1)),
])
def test_multi_tick_launch_equals_single_ticks(kernel, per_block, shape):
    """Three 8-tick launches against 24 one-tick launches, after a 240-tick
    pre-roll into contact-rich play (both batches pre-rolled alike)."""
    n = 512
    a, b = _batch(shape, n, 4242), _batch(shape, n, 4242)
    try:
        for x in (a, b):
            x.set_kernel(kernel)
            x.set_launch(per_block)
            for _ in range(30):
                x.step(8)
        assert _state(a) == _state(b)
        for _ in range(3):
            a.step(8)
            for _ in range(8):
                b.step(1)
            assert _state(a) == _state(b)
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,per_block", [(KERNEL_COOP, 4), (KERNEL_SCRATCH, 64)])
@pytest.mark.parametrize("name", ["soccar_ball_tf.npz", "soccar_drive_tf.npz", "soccar_melee_tf.npz",
                                  "soccar_walls_tf.npz"])
def test_fixture_parity_with_bench_launch(name, kernel, per_block):
    samples = parity.load_samples(os.path.join(HERE, "golden", name))
    results = parity.run_on_gpu(samples, MODE_SOCCAR, kernel, per_block)
    nbad = parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)
    print(f"\n{name} kernel {kernel} x{per_block}: {nbad}/{len(samples)} samples out of tolerance")


@pytest.mark.gpu
def test_mixed_1v1_2v2_3v3_parity_one_launch():
    """Ball, drive, melee and wall samples with 2, 4 and 6 cars, shuffled into
    one batch and stepped in one launch (automatic and bench launch shapes)."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = []
    samples += parity.sample_soccar_ball(ref, n_traj=8, ticks_per_traj=60, seed=101, stride=3)
    samples += parity.sample_soccar_drive(ref, n_traj=4, ticks_per_traj=120, seed=102, n_cars=2, stride=3)
    samples += parity.sample_soccar_melee(ref, n_traj=8, ticks_per_traj=60, seed=103, n_cars=4, stride=2)
    samples += parity.sample_soccar_melee(ref, n_traj=6, ticks_per_traj=60, seed=104, n_cars=6, stride=2,
                                          near_walls=True)
    samples += parity.sample_soccar_drive(ref, n_traj=4, ticks_per_traj=120, seed=105, n_cars=6, stride=3)
    random.Random(7).shuffle(samples)
    assert {len(s.teams) for s in samples} == {2, 4, 6}
    for kernel, per_block in ((KERNEL_AUTO, 0), (KERNEL_COOP, 4), (KERNEL_SCRATCH, 64)):
        results = parity.run_on_gpu(samples, MODE_SOCCAR, kernel, per_block)
        nbad = parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)
        print(f"\nmixed launch kernel {kernel} x{per_block}: {nbad}/{len(samples)} samples out of tolerance")


@pytest.mark.gpu
def test_c1_sampled_arenas_teacher_forced():
    """4096 Soccar 2v2 Octane arenas run the seeded control process with
    auto-reset in the bench's launches (cooperative kernel, 4 arenas per
    workgroup, 8-tick launches); at ticks 240, 600 and 960, 256 sampled
    arenas' full states are set into the oracle with the controls the GPU
    applied on the next tick, and both step one tick."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    n = 4096
    b = _batch((2, 2), n, 20261202)
    rng = random.Random(99)
    samples = []
    try:
        b.set_kernel(KERNEL_COOP)
        b.set_launch(4)
        t = 0
        for mark in (240, 600, 960):
            while t < mark:
                b.step(8)
                t += 8
            cars, balls, pads, ticks = b.get_cars(), b.get_balls(), b.get_pads(), b.get_tick_counts()
            b.step(1)
            t += 1
            ctrls = b.get_controls()  # the controls the GPU drew and applied in that tick
            print(f"\n  tick {t}: setting 256 sampled arenas into the oracle", flush=True)
            for a in rng.sample(range(n), 256):
                samples.append(parity.sample_from_state(
                    ref, MODE_SOCCAR, [0, 0, 1, 1], [0, 0, 0, 0], list(cars[4 * a:4 * a + 4]), balls[a],
                    list(pads[a * 34:(a + 1) * 34]), ticks[a], list(ctrls[4 * a:4 * a + 4])))
        goals = b.get_goals()
    finally:
        b.close()
    results = parity.run_on_gpu(samples, MODE_SOCCAR, KERNEL_COOP, 4)
    nbad = parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)
    print(f"\nC1 sampled arenas: {nbad}/{len(samples)} samples out of tolerance; "
          f"goals in the run: {sum(goals[0]) + sum(goals[1])}")
