"""Scripted Soccar play for the game-event tracker tests (test infrastructure
only): reference arenas where cars shoot at goal, some with a pass before and
some with a defender on the line, with the reference's GameEventTracker
updated as it goes.  Every update's state is recorded so another tracker can
be driven through the same states."""
import math
import random

import parity
from rocketsim_amd.state import Controls, GameEventConfig, MODE_HOOPS, MODE_SOCCAR

TEAMS = [0, 1, 0, 1]


def _yaw_rot(yaw):
    c, s = math.cos(yaw), math.sin(yaw)
    return [c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0]


def _attempt(A, rng, k):
    """Attempt k: blue (even k) or orange shoots from 1200-3000 uu out.  Half
    the attempts are a car driven into the ball at rest; the others roll the
    ball at the net as if the shooter had touched it 40 ticks before, with a
    defender parked in its path most of the time (saves)."""
    blue = k % 2 == 0
    sgn = 1.0 if blue else -1.0
    shooter, defender, mate = (0, 1, 2) if blue else (1, 0, 3)
    driven = rng.random() < 0.5
    x = rng.uniform(-600, 600)
    y0 = sgn * rng.uniform(1200, 3000)
    tick = A.tick_count
    b = A.get_ball()
    b.pos[:] = [x, y0, 93.15]
    b.vel[:] = [0.0, 0.0, 0.0] if driven else [rng.uniform(-150, 150), sgn * rng.uniform(2000, 2800), 0.0]
    b.ang_vel[:] = [0.0, 0.0, 0.0]
    A.set_ball(b)
    c = A.get_car(shooter)
    c.pos[:] = [x + rng.uniform(-40, 40), y0 - sgn * (260.0 if driven else 600.0), 17.0]
    c.rot[:] = _yaw_rot(sgn * math.pi / 2)
    c.vel[:] = [0.0, sgn * rng.uniform(1400, 2250), 0.0] if driven else [0.0, 0.0, 0.0]
    c.ang_vel[:] = [0.0, 0.0, 0.0]
    if not driven:
        c.flags |= 1 << 15
        c.hit_tick = max(0, tick - 40)
    A.set_car(shooter, c)
    ctl = Controls()
    ctl.throttle = 1.0 if driven else 0.0
    ctl.boost = 1 if driven else 0
    A.set_controls(shooter, ctl)
    if rng.random() < 0.5:
        # a team mate touched the ball shortly before: a pass
        m = A.get_car(mate)
        m.flags |= 1 << 15
        m.hit_tick = max(0, tick - rng.randint(5, 200))
        A.set_car(mate, m)
    d = A.get_car(defender)
    if not driven and rng.random() < 0.7:
        # a defender parked side-on in the ball's path
        d.pos[:] = [x + rng.uniform(-40, 40), y0 + sgn * rng.uniform(1000, 1400), 17.0]
        d.rot[:] = _yaw_rot(0.0)
    else:
        d.pos[:] = [rng.uniform(-3000, 3000), -sgn * 3000.0, 17.0]
        d.rot[:] = _yaw_rot(-sgn * math.pi / 2)
    d.vel[:] = [0.0, 0.0, 0.0]
    d.ang_vel[:] = [0.0, 0.0, 0.0]
    A.set_car(defender, d)
    A.set_controls(defender, Controls())


def _attempt_hoops(A, rng, k):
    """Hoops attempt k: blue (even k) or orange lobs the ball at the far hoop
    from 900-1500 uu out at about 2000 uu/s, touched by the shooter 40 ticks before, aimed
    within 250 uu of the hoop's centre (some hit the rim or miss); half the
    time a team mate touched it earlier (a pass)."""
    blue = k % 2 == 0
    sgn = 1.0 if blue else -1.0
    shooter, mate = (0, 2) if blue else (1, 3)
    tick = A.tick_count
    x = rng.uniform(-300, 300)
    y0 = sgn * rng.uniform(900, 1500)
    ty = sgn * (2900.0 + rng.uniform(-250, 250))
    t = rng.uniform(0.8, 1.0)  # fast enough for a shot (GameEventTrackerConfig::shotMinSpeed)
    z0, zt, g = 150.0, 560.0, -650.0  # over the rim height the tracker predicts from (365 uu + 1.2 radii)
    b = A.get_ball()
    b.pos[:] = [x, y0, z0]
    b.vel[:] = [rng.uniform(-60, 60), (ty - y0) / t, (zt - z0 - 0.5 * g * t * t) / t]
    b.ang_vel[:] = [0.0, 0.0, 0.0]
    A.set_ball(b)
    c = A.get_car(shooter)
    c.pos[:] = [x, y0 - sgn * 500.0, 17.0]
    c.rot[:] = _yaw_rot(sgn * math.pi / 2)
    c.vel[:] = [0.0, 0.0, 0.0]
    c.ang_vel[:] = [0.0, 0.0, 0.0]
    c.flags |= 1 << 15
    c.hit_tick = max(0, tick - 40)
    A.set_car(shooter, c)
    A.set_controls(shooter, Controls())
    if rng.random() < 0.5:
        m = A.get_car(mate)
        m.flags |= 1 << 15
        m.hit_tick = max(0, tick - rng.randint(5, 200))
        A.set_car(mate, m)


def run_reference(ref, n_arenas, ticks, update_every, seed, config=None, mode=MODE_SOCCAR):
    """Returns (teams, presets, updates, events): updates[u] = (cars, balls,
    ticks, orders) with cars[a] the arena's CarState records in creation
    order; events[a] = the reference tracker's events for arena a.  In
    MODE_HOOPS the attempts are lobs at the hoops."""
    rng = random.Random(seed)
    config = config or GameEventConfig()
    presets = [rng.randint(0, 5) for _ in range(4 * n_arenas)]
    arenas, trackers = [], []
    for a in range(n_arenas):
        A = ref.RefArena(mode)
        for i in range(4):
            A.add_car(TEAMS[i], presets[4 * a + i])
        A.reset_kickoff(rng.randint(0, 1000))
        arenas.append(A)
        trackers.append(ref.RefGameEventTracker(config))
    updates, events = [], [[] for _ in range(n_arenas)]
    attempt_len = rng.choice([150, 170, 190]) * (2 if mode == MODE_HOOPS else 1)
    for t in range(ticks):
        for a, A in enumerate(arenas):
            if t % attempt_len == 0:
                (_attempt_hoops if mode == MODE_HOOPS else _attempt)(A, rng, t // attempt_len + a)
            elif t % 40 == 20:
                for i in (2, 3):
                    A.set_controls(i, parity.rand_controls(rng))
            A.step(1)
        if t % update_every == update_every - 1:
            rec = ([[A.get_car(i) for i in range(4)] for A in arenas], [A.get_ball() for A in arenas],
                   [A.tick_count for A in arenas], [A.car_order() for A in arenas])
            updates.append(rec)
            for a, (A, tr) in enumerate(zip(arenas, trackers)):
                tr.update(A)
                events[a].extend(tr.pop())
    for tr in trackers:
        tr.close()
    for A in arenas:
        A.close()
    teams = TEAMS * n_arenas
    return teams, presets, updates, events


def event_key(e):
    return (e.type, e.tick, e.a, e.b)
