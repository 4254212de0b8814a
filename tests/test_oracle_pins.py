"""Pins the oracle (the reference RocketSim compiled from its own sources,
oracle/_ref) against what the mathematics and the reference's constants fix,
independently of the GPU code: closed-form free flight with Bullet's damping
order, the speed caps, boost use and air acceleration, big-pad pickup, the
kickoff spawn table, and resting equilibria that the contact solver and the
suspension must reach.  The reference ships no golden vectors or known-answer
tests of its own, so these are the checks that would catch a mis-built or
mis-driven oracle (a wrong unit, gravity sign, drag exponent, tick order).

Constants are from src/RLConst.h (cited per test)."""
import math

import pytest

from rocketsim_amd.state import Controls, MODE_SOCCAR, MODE_THE_VOID, CF_ON_GROUND, CF_WHEEL

DT = 1.0 / 120.0
GRAVITY_Z = -650.0          # RLConst.h:12
BALL_DRAG = 0.03            # RLConst.h:45 (Bullet linear damping, per second)
BALL_MAX_SPEED = 6000.0     # RLConst.h:54
CAR_MAX_SPEED = 2300.0      # RLConst.h:53
CAR_MAX_ANG_SPEED = 5.5     # RLConst.h:66
BOOST_USED_PER_SECOND = 100.0 / 3.0   # RLConst.h:57
BOOST_ACCEL_AIR = 3175.0 / 3.0        # RLConst.h:60
BALL_REST_Z = 93.15         # RLConst.h:43
CAR_SPAWN_REST_Z = 17.0     # RLConst.h:141


@pytest.fixture(scope="module")
def ref():
    from oracle import ref as r
    if not r.available():
        pytest.skip("oracle library not built (make -C oracle)")
    r.init()
    return r


def _void_arena(ref, n_cars=0):
    A = ref.RefArena(MODE_THE_VOID)
    for i in range(n_cars):
        A.add_car(i % 2, 0)
    return A


def test_ball_free_flight_closed_form(ref):
    """Ball in THE_VOID: per tick Bullet damps first (v *= (1-d)^dt,
    btRigidBody::applyDamping in predictUnconstraintMotion), the solver adds
    g*dt, then the pose advances by v*dt."""
    A = _void_arena(ref)
    b = A.get_ball()
    b.pos[:] = [100.0, -200.0, 1500.0]
    b.vel[:] = [400.0, -250.0, 300.0]
    b.ang_vel[:] = [0.0, 0.0, 0.0]
    A.set_ball(b)
    x = [100.0, -200.0, 1500.0]
    v = [400.0, -250.0, 300.0]
    damp = (1.0 - BALL_DRAG) ** DT
    for t in range(90):
        A.step(1)
        v = [v[0] * damp, v[1] * damp, v[2] * damp + GRAVITY_Z * DT]
        x = [x[k] + v[k] * DT for k in range(3)]
    got = A.get_ball()
    for k in range(3):
        assert abs(got.vel[k] - v[k]) <= 1e-4 * max(1.0, abs(v[k])), (k, got.vel[k], v[k])
        assert abs(got.pos[k] - x[k]) <= 1e-4 * max(1.0, abs(x[k])), (k, got.pos[k], x[k])
    A.close()


def test_car_free_flight_linear(ref):
    """A car in THE_VOID with no input: no linear damping, gravity only."""
    A = _void_arena(ref, 1)
    s = A.get_car(0)
    s.pos[:] = [0.0, 0.0, 1000.0]
    s.vel[:] = [300.0, -120.0, 50.0]
    s.ang_vel[:] = [0.0, 0.0, 0.0]
    s.flags = 0
    A.set_car(0, s)
    A.set_controls(0, Controls())
    n = 60
    A.step(n)
    got = A.get_car(0)
    assert abs(got.vel[0] - 300.0) < 1e-3
    assert abs(got.vel[1] + 120.0) < 1e-3
    vz = 50.0 + GRAVITY_Z * DT * n
    assert abs(got.vel[2] - vz) < 1e-3 * abs(vz)
    A.close()


def test_speed_caps(ref):
    """Ball and car velocity clamps at the end of the tick
    (Ball::_FinishPhysicsTick, Car::_FinishPhysicsTick)."""
    A = _void_arena(ref, 1)
    b = A.get_ball()
    b.pos[:] = [0.0, 0.0, 1500.0]
    b.vel[:] = [9000.0, 0.0, 0.0]
    A.set_ball(b)
    s = A.get_car(0)
    s.pos[:] = [2000.0, 0.0, 1000.0]
    s.vel[:] = [0.0, 3000.0, 0.0]
    s.ang_vel[:] = [10.0, 0.0, 0.0]
    s.flags = 0
    A.set_car(0, s)
    A.set_controls(0, Controls())
    A.step(1)
    bv = A.get_ball().vel
    cv = A.get_car(0).vel
    cw = A.get_car(0).ang_vel
    assert math.sqrt(sum(x * x for x in bv)) <= BALL_MAX_SPEED * (1 + 1e-6)
    assert math.sqrt(sum(x * x for x in cv)) <= CAR_MAX_SPEED * (1 + 1e-6)
    assert math.sqrt(sum(x * x for x in cw)) <= CAR_MAX_ANG_SPEED * (1 + 1e-6)
    A.close()


def test_boost_use_and_air_acceleration(ref):
    """Holding boost in the air for 0.5 s uses 100/3 per second and adds
    3175/3 uu/s^2 along the car's forward axis (no rotation, no input)."""
    A = _void_arena(ref, 1)
    s = A.get_car(0)
    s.pos[:] = [0.0, 0.0, 1000.0]
    s.vel[:] = [0.0, 0.0, 0.0]
    s.ang_vel[:] = [0.0, 0.0, 0.0]
    s.boost = 50.0
    s.flags = 0
    A.set_car(0, s)
    c = Controls()
    c.boost = 1
    A.set_controls(0, c)
    n = 60
    A.step(n)
    got = A.get_car(0)
    assert abs(got.boost - (50.0 - BOOST_USED_PER_SECOND * DT * n)) < 1e-3
    fwd = got.rot[0:3]
    along = sum(got.vel[k] * fwd[k] for k in range(3))
    # the first boosting tick applies the force too: n ticks of acceleration
    assert abs(along - BOOST_ACCEL_AIR * DT * n) < 1e-3 * BOOST_ACCEL_AIR * DT * n
    A.close()


def test_big_pad_pickup(ref):
    """A car resting on a big pad with no boost is filled to 100 and the pad
    goes on its 10 s cooldown (BoostPad.cpp:62-107, RLConst.h:256-266)."""
    import ctypes as C
    A = ref.RefArena(MODE_SOCCAR)
    A.add_car(0, 0)
    L = ref.lib()
    big = None
    for i in range(34):
        pos = (C.c_float * 3)()
        is_big = C.c_int()
        L.rsref_pad_config(A.h, i, pos, C.byref(is_big))
        if is_big.value and abs(pos[1]) < 1 and pos[0] > 0:
            big = (i, list(pos))
    assert big is not None
    i, pos = big
    s = A.get_car(0)
    s.pos[:] = [pos[0], pos[1], CAR_SPAWN_REST_Z]
    s.rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    s.vel[:] = [0, 0, 0]
    s.ang_vel[:] = [0, 0, 0]
    s.boost = 0.0
    s.flags = CF_ON_GROUND | sum(CF_WHEEL)
    A.set_car(0, s)
    A.set_controls(0, Controls())
    A.step(2)
    assert A.get_car(0).boost == pytest.approx(100.0)
    pad = A.get_pads()[i]
    assert pad.is_active == 0
    assert 10.0 - 2 * DT - 1e-4 <= pad.cooldown <= 10.0
    A.close()


def test_kickoff_spawns_from_table(ref):
    """ResetToRandomKickoff places cars on the kickoff spawn table
    (RLConst.h CAR_SPAWN_LOCATIONS_SOCCAR, mirrored for orange) at rest
    height, the ball at rest at BALL_REST_Z."""
    table = {(-2048.0, -2560.0), (2048.0, -2560.0), (-256.0, -3840.0), (256.0, -3840.0), (0.0, -4608.0)}
    for seed in range(6):
        A = ref.RefArena(MODE_SOCCAR)
        for i in range(4):
            A.add_car(i % 2, 0)
        A.reset_kickoff(seed)
        seen = set()
        for i in range(4):
            c = A.get_car(i)
            x, y = c.pos[0], c.pos[1]
            if i % 2 == 1:
                x, y = -x, -y
            key = (round(x), round(y))
            assert (float(key[0]), float(key[1])) in table, key
            assert abs(c.pos[2] - CAR_SPAWN_REST_Z) < 1e-3
            seen.add((i % 2, key))
        assert len(seen) == 4
        b = A.get_ball()
        assert list(b.pos) == pytest.approx([0.0, 0.0, BALL_REST_Z])
        A.close()


def test_resting_equilibria(ref):
    """A dropped ball settles on the floor at BALL_REST_Z (sphere radius plus
    the floor contact); a car at a kickoff spawn settles on its four wheels
    near its spawn height and stays on the ground.  (A ball with exactly zero
    velocity is put to sleep and does not fall, Arena.cpp:694-700, so it is
    released with a small downward speed.)"""
    A = ref.RefArena(MODE_SOCCAR)
    A.add_car(0, 0)
    A.reset_kickoff(3)
    b = A.get_ball()
    b.pos[:] = [0.0, 0.0, 250.0]
    b.vel[:] = [0.0, 0.0, -1.0]
    A.set_ball(b)
    A.set_controls(0, Controls())
    A.step(1200)
    b = A.get_ball()
    assert abs(b.pos[2] - BALL_REST_Z) < 0.5, b.pos[2]
    assert math.sqrt(sum(v * v for v in b.vel)) < 2.0
    c = A.get_car(0)
    assert c.flags & CF_ON_GROUND
    assert all(c.flags & w for w in CF_WHEEL)
    assert abs(c.pos[2] - CAR_SPAWN_REST_Z) < 1.0, c.pos[2]
    assert math.sqrt(sum(v * v for v in c.vel)) < 1.0
    A.close()
