"""Kickoff reset, auto-reset on goal and demolition respawn against the
oracle (the reference's Arena::ResetToRandomKickoff, src/Sim/Arena/Arena.cpp:113-223,
and Car::Respawn, src/Sim/Car/Car.cpp:43-56).

The reference draws the kickoff slot order and the respawn slot from
std::default_random_engine; the batch draws them from its own counter-based
generator (include/rsg_workload.h).  The tests therefore take the GPU's draw
as given, find the reference's seed (or run) that draws the same slots, and
compare everything else exactly or within the one-tick tolerances:
  * a ball sent across the goal line in a batch with auto-reset on: the
    goal is counted and reported as the reference's goal callback reports
    it, and the tick ends in the reference's kickoff state for the same slot
    order (cars on their spawn slots, mirrored for orange, boost, flags,
    ball at rest at the centre, pads reset); the next tick then matches the
    reference stepped from that state;
  * rsg_batch_reset_kickoff places every arena as the reference's reset with
    the same slot order does;
  * a demolished car whose timer runs out respawns on one of the reference's
    respawn slots with the reference's state, on the same tick as the
    reference's car, and the tick after matches.
"""
import random

import pytest

import parity
from rocketsim_amd.state import CF_DEMOED, Controls, EVENT_GOAL, MODE_SOCCAR


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


def _ranked(cars, teams, order):
    """Cars keyed by (team, rank in the car update order): a kickoff gives the
    r-th car of each team, in that order, the r-th slot of its draw
    (Arena.cpp:153-195)."""
    out, seen = {}, [0, 0]
    for i in order:
        out[(teams[i], seen[teams[i]])] = cars[i]
        seen[teams[i]] += 1
    return out


def _slot_key(cars, teams, order):
    """The spawn positions (uu, rounded) of each team's cars by update rank,
    i.e. the slot order a kickoff drew."""
    r = _ranked(cars, teams, order)
    return tuple(sorted((k, tuple(round(v, 1) for v in c.pos[:2])) for k, c in r.items()))


def _seed_for(ref, teams, presets, key, cache):
    """A reference seed whose ResetToRandomKickoff draws the slot order `key`
    for cars of `teams`."""
    tk = (tuple(teams), tuple(presets))
    if tk not in cache:
        cache[tk] = {}
        A = parity.pooled_arena(ref, MODE_SOCCAR, teams, presets)
        order = A.car_order()
        # seeds spread over the int range: std::default_random_engine seeded
        # with small integers never draws the first two slots first
        for k in range(4000):
            seed = (k * 2654435761 + 12345) % 2147483647
            A.reset_kickoff(seed)
            cache[tk].setdefault(_slot_key([A.get_car(i) for i in range(len(teams))], teams, order), seed)
    return cache[tk].get(key)


@pytest.mark.gpu
@pytest.mark.parametrize("per_team", [1, 2, 3])
def test_goal_auto_reset_matches_reference_kickoff(per_team):
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    rng = random.Random(500 + per_team)
    teams = [i % 2 for i in range(2 * per_team)]
    samples, arenas = [], []
    for k in range(24):
        presets = [rng.randrange(6) for _ in teams]
        A = ref.RefArena(MODE_SOCCAR)
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 100000))
        side = rng.choice([-1.0, 1.0])
        b = A.get_ball()
        b.pos[:] = [rng.uniform(-600, 600), side * (5124.25 + 91.25 - 8.0), rng.uniform(120, 500)]
        b.vel[:] = [rng.gauss(0, 200), side * rng.uniform(1600, 2400), rng.gauss(0, 200)]
        A.set_ball(b)
        ctrls = [parity.rand_controls(rng) for _ in teams]
        s = parity._one_tick(A, MODE_SOCCAR, teams, presets, ctrls)
        assert any(e[0] == EVENT_GOAL for e in s.events), "scenario did not score"
        samples.append(s)
        arenas.append(A)
    results = parity.run_on_gpu(samples, MODE_SOCCAR, auto_reset=True, seed=77)
    cache = {}
    try:
        for i, (s, r, A) in enumerate(zip(samples, results, arenas)):
            cars, ball, pads, tick, events = r
            # the goal, as the reference's goal callback reported it
            assert [e for e in events if e[0] == EVENT_GOAL] == [e for e in s.events if e[0] == EVENT_GOAL]
            assert tick == s.tick_out
            # the reference's kickoff with the GPU's slot order, from the
            # reference's post-tick state
            key = _slot_key(cars, s.teams, s.car_order)
            seed = _seed_for(ref, s.teams, s.presets, key, cache)
            assert seed is not None, f"arena {i}: GPU slot order {key} is no kickoff"
            A.reset_kickoff(seed)
            kick = parity.Sample()
            kick.mode, kick.teams, kick.presets = MODE_SOCCAR, s.teams, s.presets
            kick.cars_out = [A.get_car(j) for j in range(len(s.teams))]
            kick.ball_out, kick.pads_out, kick.tick_out = A.get_ball(), A.get_pads(), A.tick_count
            nbad, msgs = parity.check_samples([kick], [(cars, ball, pads, tick)])
            assert nbad == 0, f"arena {i}: " + "\n".join(msgs)
            for j, c in enumerate(cars):
                ref_c = kick.cars_out[j]
                assert (c.boost, c.flags) == (ref_c.boost, ref_c.flags)
                assert list(c.vel[:]) == [0.0, 0.0, 0.0] and list(c.ang_vel[:]) == [0.0, 0.0, 0.0]
            assert list(ball.vel[:]) == [0.0, 0.0, 0.0] and list(ball.ang_vel[:]) == [0.0, 0.0, 0.0]
            assert all(p.is_active and p.cooldown == 0.0 for p in pads)
        # the tick after the reset, teacher-forced from the reset state
        nxt = [parity._one_tick(A, MODE_SOCCAR, s.teams, s.presets, [Controls() for _ in s.teams])
               for s, A in zip(samples, arenas)]
        parity.assert_parity(nxt, parity.run_on_gpu(nxt, MODE_SOCCAR), MODE_SOCCAR)
    finally:
        for A in arenas:
            A.close()


@pytest.mark.gpu
def test_reset_kickoff_matches_reference():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    from rocketsim_amd.batch import ArenaBatch
    rng = random.Random(3)
    counts, teams, presets = [], [], []
    for _ in range(64):
        per_team = rng.choice([1, 2, 3])
        counts.append(2 * per_team)
        teams += [0, 1] * per_team
        presets += [rng.randrange(6) for _ in range(2 * per_team)]
    b = ArenaBatch(counts, teams, presets, mode=MODE_SOCCAR, seed=5)
    cache = {}
    try:
        b.step(30)
        b.reset_kickoff(-1, 12345)
        cars, balls, pads = b.get_cars(), b.get_balls(), b.get_pads()
        orders = b.get_car_order()
        k = 0
        for a, n in enumerate(counts):
            t, p = teams[k:k + n], presets[k:k + n]
            got = list(cars[k:k + n])
            seed = _seed_for(ref, t, p, _slot_key(got, t, orders[k:k + n]), cache)
            assert seed is not None
            A = parity.pooled_arena(ref, MODE_SOCCAR, t, p)
            A.reset_kickoff(seed)
            ref_ranked = _ranked([A.get_car(j) for j in range(n)], t, A.car_order())
            for key, g in _ranked(got, t, orders[k:k + n]).items():
                r = ref_ranked[key]
                assert max(abs(x - y) for x, y in zip(g.pos[:], r.pos[:])) <= 1e-3
                assert max(abs(x - y) for x, y in zip(g.rot[:], r.rot[:])) <= 1e-6
                assert (g.boost, g.flags, list(g.vel[:])) == (r.boost, r.flags, list(r.vel[:]))
            rb = A.get_ball()
            assert max(abs(x - y) for x, y in zip(balls[a].pos[:], rb.pos[:])) <= 1e-3
            assert list(balls[a].vel[:]) == list(rb.vel[:])
            assert all(x.is_active and x.cooldown == 0.0 for x in pads[a * 34:(a + 1) * 34])
            k += n
    finally:
        b.close()


def _respawn_positions(ref, teams, presets, before, car, runs=48):
    """The reference's respawn of `car` (timer running out in the next tick)
    from the state `before`: its positions over `runs` runs (the reference
    draws the slot from its own random engine) with one full after-state per
    position."""
    out = {}
    A = parity.pooled_arena(ref, MODE_SOCCAR, teams, presets)
    for _ in range(runs):
        for j, c in enumerate(before):
            A.set_car(j, c)
        A.step(1)
        c = A.get_car(car)
        out.setdefault(tuple(round(v, 1) for v in c.pos[:2]), c)
    return out


@pytest.mark.gpu
def test_demolished_car_respawns_like_reference():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    rng = random.Random(9)
    teams, samples, victims = [0, 1, 0, 1], [], []
    for k in range(16):
        presets = [rng.randrange(6) for _ in teams]
        A = ref.RefArena(MODE_SOCCAR)
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 10000))
        A.step(rng.randint(10, 60))
        v = k % 4
        c = A.get_car(v)
        c.flags |= CF_DEMOED
        c.demo_respawn_timer = 1.0 / 120.0  # runs out in the next tick
        A.set_car(v, c)
        samples.append(parity._one_tick(A, MODE_SOCCAR, teams, presets, [Controls() for _ in teams]))
        victims.append(v)
        A.close()
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    seen = set()
    for i, (s, r, v) in enumerate(zip(samples, results, victims)):
        got = r[0][v]
        assert not got.flags & CF_DEMOED
        options = _respawn_positions(ref, s.teams, s.presets, s.cars_in, v)
        key = tuple(round(x, 1) for x in got.pos[:2])
        assert key in options, f"sample {i}: GPU respawn at {key}, reference's at {sorted(options)}"
        seen.add((s.teams[v], key))
        # everything but the drawn slot as the reference's respawn on that
        # slot; the respawned car's broadphase cell is the one of its slot.
        # The registration order of the bodies is left out here: the
        # reference run that drew the GPU's slot has another history (every
        # other parity test compares it).
        opt = type(options[key]).from_buffer_copy(bytes(options[key]))
        assert opt.bp_cell == got.bp_cell, f"sample {i}: cell {got.bp_cell} vs {opt.bp_cell}"
        s.cars_out = list(s.cars_out)
        s.cars_out[v] = opt
        for j, c in enumerate(s.cars_out):
            c.bp_seq = r[0][j].bp_seq
        s.ball_out.bp_seq = r[1].bp_seq
        nbad, msgs = parity.check_samples([s], [r])
        assert nbad == 0, f"sample {i}: " + "\n".join(msgs)
    assert len(seen) >= 4  # several slots were drawn


@pytest.mark.gpu
def test_full_demo_timer_respawns_on_reference_tick():
    """A car demolished with the full 3 s timer respawns on the same tick as
    the reference's (the timer counts down by the tick time with the
    reference's float rounding)."""
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    from rocketsim_amd.batch import ArenaBatch
    teams, presets = [0, 1], [0, 0]
    A = ref.RefArena(MODE_SOCCAR)
    for t, p in zip(teams, presets):
        A.add_car(t, p)
    A.reset_kickoff(11)
    c = A.get_car(0)
    c.flags |= CF_DEMOED
    c.demo_respawn_timer = 3.0
    A.set_car(0, c)
    b = ArenaBatch([2], teams, presets, mode=MODE_SOCCAR)
    try:
        b.set_cars([A.get_car(0), A.get_car(1)])
        b.set_balls([A.get_ball()])
        b.set_pads(A.get_pads())
        b.set_car_order(A.car_order())
        ref_tick = gpu_tick = None
        for t in range(1, 400):
            A.step(1)
            b.step(1)
            if ref_tick is None and not A.get_car(0).flags & CF_DEMOED:
                ref_tick = t
            if gpu_tick is None and not b.get_cars()[0].flags & CF_DEMOED:
                gpu_tick = t
            if ref_tick and gpu_tick:
                break
        assert ref_tick is not None and ref_tick == gpu_tick
    finally:
        b.close()
        A.close()
