"""MutatorConfig (src/Sim/MutatorConfig/MutatorConfig.h:16-79) and
Arena::SetMutatorConfig (src/Sim/Arena/Arena.cpp:17-48) against the oracle.

  * The defaults of each mode are the reference's, field for field.
  * SetMutatorConfig's side effects match the reference's on the same
    arena history: the ball remade at rest at (0, 0, radius) when its radius
    or mass changes, every car given the ball's mass when the car mass
    changes, and the new gravity given only to bodies that are awake (a ball
    asleep at kickoff keeps the old gravity after it is woken; so does a
    demolished car after it respawns).
  * Teacher-forced one-tick parity in contact-rich melees under two
    non-default configs: the host build of the tick code with the reference's
    SSE normalisation matches the oracle exactly on the CPU, and the GPU
    matches it within the parity tolerances (tests/parity.py).
"""
import copy
import ctypes as C
import math

import pytest

import parity
import tick_host
from rocketsim_amd.state import (DEMO_ON_CONTACT, MODE_HEATSEEKER, MODE_SOCCAR, MODE_THE_VOID, MutatorConfig,
                                 BallState, CarState, CF_DEMOED)


@pytest.fixture(scope="module")
def ref():
    from oracle import ref as r
    if not r.available():
        pytest.skip("oracle library not built on this machine")
    r.init()
    return r


def _defaults(mode):
    from rocketsim_amd import native
    m = MutatorConfig()
    native.check(native.lib().rsg_mutator_config_default(mode, C.byref(m)))
    return m


def config_big_ball_low_gravity():
    """A remade ball (radius and mass), lower gravity, more drag and bounce, a
    lower speed cap, a nearer goal line."""
    m = _defaults(MODE_SOCCAR)
    m.gravity[2] = -420.0
    m.ball_radius = 100.0
    m.ball_mass = 42.0
    m.ball_drag = 0.06
    m.ball_max_speed = 4200.0
    m.ball_world_restitution = 0.85
    m.ball_world_friction = 0.25
    m.ball_hit_extra_force_scale = 1.6
    m.goal_base_threshold_y = 4900.0
    return m


def config_heavy_cars_demos():
    """The car-mass path (cars get the ball's mass), demolitions on any bumper
    contact including team mates, stronger bumps and jumps, boost recharge,
    unlimited flips and double jumps, quick respawns and pads."""
    m = _defaults(MODE_SOCCAR)
    m.car_mass = 240.0
    m.car_world_friction = 0.5
    m.car_world_restitution = 0.1
    m.jump_accel = 2000.0
    m.jump_immediate_force = 400.0
    m.boost_accel_ground = 1200.0
    m.boost_accel_air = 1300.0
    m.boost_used_per_second = 20.0
    m.respawn_delay = 0.5
    m.bump_cooldown_time = 0.1
    m.boost_pad_cooldown_big = 3.0
    m.boost_pad_cooldown_small = 1.0
    m.car_spawn_boost_amount = 80.0
    m.bump_force_scale = 2.5
    m.unlimited_flips = 1
    m.unlimited_double_jumps = 1
    m.recharge_boost_enabled = 1
    m.recharge_boost_per_second = 15.0
    m.recharge_boost_delay = 0.1
    m.demo_mode = DEMO_ON_CONTACT
    m.enable_team_demos = 1
    return m


CONFIGS = {"big_ball_low_gravity": config_big_ball_low_gravity, "heavy_cars_demos": config_heavy_cars_demos}


@pytest.mark.parametrize("mode", [MODE_SOCCAR, MODE_HEATSEEKER, MODE_THE_VOID])
def test_defaults_are_the_references(ref, mode):
    A = ref.RefArena(mode)
    try:
        assert bytes(_defaults(mode)) == bytes(A.get_mutators())
    finally:
        A.close()


def _host_batch(L, ref_arena, mode, teams, presets):
    """A one-arena host batch in the oracle arena's current state."""
    from rocketsim_amd import arena_mesh
    n = len(teams)
    cmfs = [arena_mesh.soccar_cmf()] if mode != MODE_THE_VOID else []
    h = L.th_create(mode, 1, (C.c_int * 1)(n), (C.c_uint8 * n)(*teams), (C.c_uint8 * n)(*presets),
                    *tick_host.mesh_args(cmfs), 120.0, None, None)
    assert h
    L.th_set_cars(h, (CarState * n)(*[ref_arena.get_car(i) for i in range(n)]))
    L.th_set_balls(h, (BallState * 1)(ref_arena.get_ball()))
    L.th_set_tick_counts(h, (C.c_uint64 * 1)(ref_arena.tick_count))
    L.th_set_car_order(h, (C.c_int32 * n)(*ref_arena.car_order()))
    return h


def _host_state(L, h, n):
    cars = (CarState * n)()
    ball = (BallState * 1)()
    L.th_get_cars(h, cars)
    L.th_get_balls(h, ball)
    return list(cars), ball[0]


def test_remade_ball_and_heavy_cars_match_reference(ref):
    """SetMutatorConfig mid-game on the oracle and on the host build from the
    same state: the remade ball (position, velocity, broadphase record) is
    the reference's, and 90 free ticks later, with the cars driven, the
    arenas still agree (mass and inertia of the car-mass path, the new
    gravity and ball)."""
    L = tick_host.get("sse")
    teams, presets = [0, 1, 0, 1], [0, 1, 2, 3]
    A = ref.RefArena(MODE_SOCCAR)
    for t, p in zip(teams, presets):
        A.add_car(t, p)
    A.reset_kickoff(5)
    A.step(30)
    h = _host_batch(L, A, MODE_SOCCAR, teams, presets)
    try:
        m = config_big_ball_low_gravity()
        m.car_mass = 210.0
        A.set_mutators(m)
        L.th_set_mutators(h, C.byref(m))
        cars, ball = _host_state(L, h, 4)
        rb = A.get_ball()
        assert list(ball.pos) == list(rb.pos) and list(ball.vel) == list(rb.vel)
        assert (ball.bp_cell, ball.bp_seq) == (rb.bp_cell, rb.bp_seq) or \
            parity.bp_order(ball, cars) == parity.bp_order(rb, [A.get_car(i) for i in range(4)])
        ctrl = parity.rand_controls(__import__("random").Random(3))
        ctrl.throttle, ctrl.boost = 1.0, 1
        from rocketsim_amd.state import Controls
        ctrls = [ctrl, Controls(), ctrl, Controls()]
        for i in range(4):
            A.set_controls(i, ctrls[i])
        L.th_set_controls(h, (Controls * 4)(*ctrls))
        for _ in range(90):
            A.step(1)
            L.th_step(h, 1)
        cars, ball = _host_state(L, h, 4)
        for i in range(4):
            msgs = parity.compare_car(A.get_car(i), cars[i])
            assert not msgs, msgs
        assert not parity.compare_ball(A.get_ball(), ball)
    finally:
        L.th_destroy(h)
        A.close()


def test_sleeping_ball_and_demolished_car_keep_old_gravity(ref):
    """The new gravity goes to awake bodies only: the kickoff ball, asleep at
    rest, and a demolished car keep the old gravity once woken or
    respawned, in the reference and in the host build alike."""
    L = tick_host.get("sse")
    teams, presets = [0, 1], [0, 0]
    A = ref.RefArena(MODE_SOCCAR)
    for t, p in zip(teams, presets):
        A.add_car(t, p)
    A.reset_kickoff(2)
    demo = A.get_car(1)
    demo.flags |= CF_DEMOED
    demo.demo_respawn_timer = 0.25
    A.set_car(1, demo)
    A.step(5)  # ball asleep at rest; car 1's body disabled
    h = _host_batch(L, A, MODE_SOCCAR, teams, presets)
    try:
        # the host batch needs the same history for the ball's sleep flag
        L.th_destroy(h)
        A2 = ref.RefArena(MODE_SOCCAR)
        for t, p in zip(teams, presets):
            A2.add_car(t, p)
        A2.reset_kickoff(2)
        A2.set_car(1, demo)
        h = _host_batch(L, A2, MODE_SOCCAR, teams, presets)
        A2.close()
        L.th_step(h, 5)
        m = _defaults(MODE_SOCCAR)
        m.gravity[2] = -1500.0
        A.set_mutators(m)
        L.th_set_mutators(h, C.byref(m))
        b = A.get_ball()
        b.vel[:] = [0.0, 0.0, 400.0]
        A.set_ball(b)
        L.th_set_balls(h, (BallState * 1)(b))
        zs_ref, zs_host = [], []
        for _ in range(60):
            A.step(1)
            L.th_step(h, 1)
            cars, ball = _host_state(L, h, 2)
            zs_ref.append((A.get_ball().pos[2], A.get_car(1).pos[2]))
            zs_host.append((ball.pos[2], cars[1].pos[2]))
        for (rb, rc), (hb, hc) in zip(zs_ref, zs_host):
            assert abs(rb - hb) <= 1e-3 * max(1.0, abs(rb)), (rb, hb)
            assert abs(rc - hc) <= 1e-3 * max(1.0, abs(rc)), (rc, hc)
        # the ball rose and fell at the old gravity: apex about v^2 / (2 * 650)
        apex = max(z for z, _ in zs_ref) - 93.15
        assert abs(apex - 400.0 ** 2 / (2 * 650.0)) < 8.0, apex
    finally:
        L.th_destroy(h)
        A.close()


# The big-ball config is set before the cars are added (Arena::Create,
# SetMutatorConfig, AddCar: the usual order), the heavy-car config after them
# (so the car-mass path applies).  A ball remade after its arena's cars were
# added moves to the end of the reference's collision-object order, which
# reorders islands and solver bodies; this library keeps the ball first.
FIRST = {"big_ball_low_gravity": True, "heavy_cars_demos": False}


def _samples(ref, name):
    m = CONFIGS[name]()
    seed = 101 if name == "big_ball_low_gravity" else 202
    return m, parity.sample_soccar_melee(ref, n_traj=10, ticks_per_traj=60, seed=seed, stride=2, mutators=m,
                                         mutators_first=FIRST[name])


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_mutator_melee_host_build_matches_oracle(ref, name):
    """The host build with the reference's normalisation steps every sample
    exactly as the oracle does (live samples: same CPU)."""
    m, samples = _samples(ref, name)
    results, status = tick_host.run_samples(tick_host.get("sse"), samples, MODE_SOCCAR, mutators=m,
                                            mutators_first=FIRST[name])
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, "\n".join(msgs)
    assert not any(status)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_mutator_melee_gpu_parity(ref, name):
    m, samples = _samples(ref, name)
    results = parity.run_on_gpu(samples, MODE_SOCCAR, mutators=m, mutators_first=FIRST[name])
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)


@pytest.mark.gpu
def test_gpu_set_mutator_config_matches_host_build(ref):
    """rsg_batch_set_mutator_config on the GPU batch's records equals the host
    build's on the same state, arena by arena (-1 = all, or one arena), and
    the config reads back; bad configs are refused."""
    from rocketsim_amd.batch import ArenaBatch
    b = ArenaBatch.uniform(6, 2, 2, 0, mode=MODE_SOCCAR, seed=9)
    try:
        b.step(40)
        m = config_big_ball_low_gravity()
        b.set_mutator_config(m, arena=2)
        assert bytes(b.get_mutator_config(2)) == bytes(m)
        assert bytes(b.get_mutator_config(1)) == bytes(_defaults(MODE_SOCCAR))
        balls = b.get_balls()
        r = m.ball_radius
        assert list(balls[2].pos) == pytest.approx([0.0, 0.0, r], abs=1e-3) and list(balls[2].vel) == [0, 0, 0]
        assert balls[1].pos[2] != pytest.approx(r, abs=1e-3) or balls[1].vel[2] != 0.0
        bad = copy.copy(m)
        bad.ball_radius = -1.0
        with pytest.raises(RuntimeError, match="positive"):
            b.set_mutator_config(bad)
        bad = copy.copy(m)
        bad.gravity[0] = math.nan
        with pytest.raises(RuntimeError, match="non-finite"):
            b.set_mutator_config(bad)
        # a ball the reference's broadphase refuses (AABB diagonal above 370 uu)
        bad = copy.copy(m)
        bad.ball_radius = 110.0
        with pytest.raises(RuntimeError, match="cell size"):
            b.set_mutator_config(bad)
    finally:
        b.close()
