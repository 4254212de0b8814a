"""Teacher-forced one-tick parity: sampling from the oracle and comparison.

A sample is one tick of the reference Arena::Step (src/Sim/Arena/Arena.cpp:689-756)
taken from an oracle trajectory: the full state before the tick (cars incl.
the per-wheel values btVehicleRL keeps between ticks, ball, pads, tick count),
the controls held during it, and the full state after it.  The GPU batch is
set to the "before" state of many samples at once (one sample per arena),
stepped once, and compared with the "after" state.

Tolerances (north star: 1e-3 relative on position, velocity and angular
velocity, 1e-4 on rotation; discrete state exact):
  pos      |d| <= 1e-3 * max(|ref|, 1 uu)
  vel      |d| <= 1e-3 * max(|ref|, 1 uu/s)
  ang_vel  |d| <= 1e-3 * max(|ref|, 1e-2 rad/s)
  rot      |d| <= 1e-4 (absolute, unit vectors)
  other continuous fields (timers, boost, wheel values, hit info)
           |d| <= 1e-3 * max(|ref|, 1)
  flags, ids, tick counters, boolean controls: exact
"""
import ctypes as C
import io
import math
import random

import numpy as np

from rocketsim_amd.state import (BallState, CarState, Controls, PadState, CF_HAS_JUMPED,
                                 CF_ON_GROUND, CF_SUPERSONIC, ARENA_MODES, MODE_HEATSEEKER, MODE_HOOPS, MODE_SOCCAR,
                                 MODE_THE_VOID, NUM_PADS)

EXACT_CAR = {"flags", "car_contact_other_id", "hit_tick", "hit_extra_tick", "tick_count_since_update",
             "last_controls.jump", "last_controls.boost", "last_controls.handbrake", "last_controls._pad",
             "bp_cell"}
# Registration order numbers: only their relative order across the arena's
# bodies is state (compared by bp_order below).
SKIP_CAR = {"bp_seq"}
REL_FLOOR = {"pos": 1.0, "vel": 1.0, "ang_vel": 1e-2,
             # wheel pushback is an impulse (kg m/s) taken from a ray depth that
             # is a difference of nearly equal coordinates; 10 units move the
             # car (180 kg) by 0.05 m/s, so 1e-3 of that is far below the
             # velocity tolerance
             "wheel_extra_pushback": 10.0}
ROT_TOL = 1e-4
REL_TOL = 1e-3


def flat_fields(s, prefix=""):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if hasattr(v, "_fields_"):
            out.update(flat_fields(v, prefix + name + "."))
        elif hasattr(v, "__len__"):
            out[prefix + name] = [float(x) for x in v]
        else:
            out[prefix + name] = v
    return out


def compare_car(ref, got):
    """Return a list of (field, ref, got, err) mismatches."""
    bad = []
    fr, fg = flat_fields(ref), flat_fields(got)
    for k, rv in fr.items():
        gv = fg[k]
        if k in SKIP_CAR:
            continue
        if k in EXACT_CAR:
            if rv != gv:
                bad.append((k, rv, gv, None))
            continue
        a = np.atleast_1d(np.asarray(rv, dtype=np.float64))
        b = np.atleast_1d(np.asarray(gv, dtype=np.float64))
        if k == "rot":
            err = np.max(np.abs(a - b))
            if not err <= ROT_TOL:
                bad.append((k, rv, gv, err))
            continue
        floor = REL_FLOOR.get(k, 1.0)
        err = np.max(np.abs(a - b) / np.maximum(np.abs(a), floor))
        if not err <= REL_TOL:
            bad.append((k, rv, gv, err))
    return bad


def compare_ball(ref, got):
    bad = []
    for k in ("pos", "vel", "ang_vel"):
        a = np.asarray(getattr(ref, k)[:], dtype=np.float64)
        b = np.asarray(getattr(got, k)[:], dtype=np.float64)
        err = np.max(np.abs(a - b) / np.maximum(np.abs(a), REL_FLOOR[k]))
        if not err <= REL_TOL:
            bad.append(("ball." + k, list(a), list(b), err))
    a = np.asarray(ref.rot[:]); b = np.asarray(got.rot[:])
    if np.max(np.abs(a - b)) > ROT_TOL:
        bad.append(("ball.rot", list(a), list(b), np.max(np.abs(a - b))))
    if ref.tick_count_since_update != got.tick_count_since_update:
        bad.append(("ball.tick_count_since_update", ref.tick_count_since_update, got.tick_count_since_update, None))
    if ref.bp_cell != got.bp_cell:
        bad.append(("ball.bp_cell", ref.bp_cell, got.bp_cell, None))
    # Heatseeker state: the target net and speed are discrete, the time a sum of tick times
    if (ref.hs_y_target_dir, ref.hs_cur_target_speed) != (got.hs_y_target_dir, got.hs_cur_target_speed):
        bad.append(("ball.hs", (ref.hs_y_target_dir, ref.hs_cur_target_speed),
                    (got.hs_y_target_dir, got.hs_cur_target_speed), None))
    if abs(ref.hs_time_since_hit - got.hs_time_since_hit) > 1e-5 * max(1.0, abs(ref.hs_time_since_hit)):
        bad.append(("ball.hs_time_since_hit", ref.hs_time_since_hit, got.hs_time_since_hit, None))
    return bad


def bp_order(ball, cars):
    """Broadphase registration order of (ball, cars...) as a permutation."""
    seq = [ball.bp_seq] + [c.bp_seq for c in cars]
    return sorted(range(len(seq)), key=lambda i: (seq[i], i))


def compare_pads(ref, got):
    bad = []
    for i, (r, g) in enumerate(zip(ref, got)):
        if r.is_active != g.is_active or r.prev_locked_car_id != g.prev_locked_car_id:
            bad.append((f"pad{i}", (r.is_active, r.prev_locked_car_id), (g.is_active, g.prev_locked_car_id), None))
        if abs(r.cooldown - g.cooldown) > 1e-4 * max(1.0, abs(r.cooldown)):
            bad.append((f"pad{i}.cooldown", r.cooldown, g.cooldown, abs(r.cooldown - g.cooldown)))
    return bad


class Sample:
    __slots__ = ("mode", "teams", "presets", "car_order", "cars_in", "ball_in", "pads_in", "ctrls", "tick_in",
                 "cars_out", "ball_out", "pads_out", "tick_out", "events", "live", "configs")

    def __init__(self):
        self.events = None  # goal/bump/demo events of the tick (live oracle samples)
        self.live = False   # stepped by the oracle in this process (not loaded from a fixture)
        self.configs = None  # the cars' CarConfigs when not presets


def _rand_quat(rng):
    u1, u2, u3 = rng.random(), rng.random(), rng.random()
    return (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
            math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))


def quat_to_rotmat(q):
    """Unit quaternion -> RotMat words (forward, right, up columns)."""
    x, y, z, w = q
    m = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    return [m[0][0], m[1][0], m[2][0], m[0][1], m[1][1], m[2][1], m[0][2], m[1][2], m[2][2]]


def rand_controls(rng):
    c = Controls()
    for f in ("throttle", "steer", "pitch", "yaw", "roll"):
        setattr(c, f, rng.uniform(-1, 1))
    c.jump = rng.random() < 0.5
    c.boost = rng.random() < 0.5
    c.handbrake = rng.random() < 0.5
    return c


def sample_void(ref, n_traj, ticks_per_traj, seed):
    """Airborne cars and ball flight in THE_VOID (no contacts): samples from
    oracle trajectories with random controls redrawn every tick."""
    rng = random.Random(seed)
    out = []
    for _ in range(n_traj):
        A = ref.RefArena(MODE_THE_VOID)
        ncar = rng.randint(1, 4)
        teams = [rng.randint(0, 1) for _ in range(ncar)]
        presets = [rng.randint(0, 5) for _ in range(ncar)]
        for i in range(ncar):
            A.add_car(teams[i], presets[i])
        for i in range(ncar):
            s = A.get_car(i)
            s.pos[:] = [i * 1000.0 + rng.uniform(-100, 100), rng.uniform(-100, 100), rng.uniform(800, 1800)]
            s.rot[:] = quat_to_rotmat(_rand_quat(rng))
            s.vel[:] = [rng.gauss(0, 500) for _ in range(3)]
            s.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
            s.boost = rng.uniform(0, 100)
            s.flags = CF_HAS_JUMPED if rng.random() < 0.5 else 0
            s.air_time_since_jump = rng.uniform(0, 2)
            A.set_car(i, s)
        b = A.get_ball()
        b.pos[:] = [-3000, 0, 1000]
        b.vel[:] = [rng.gauss(0, 300) for _ in range(3)]
        b.ang_vel[:] = [rng.gauss(0, 1) for _ in range(3)]
        A.set_ball(b)
        for _ in range(ticks_per_traj):
            out.append(_one_tick(A, MODE_THE_VOID, teams, presets, [rand_controls(rng) for _ in range(ncar)]))
        A.close()
    return out


def _soccar_scenario(ref, rng, kind, teams, presets):
    """A Soccar arena at a random kickoff with the ball thrown at the floor,
    a side wall, a corner or into a goal (the arena mesh and the four planes)."""
    A = ref.RefArena(MODE_SOCCAR)
    for t, p in zip(teams, presets):
        A.add_car(t, p)
    A.reset_kickoff(rng.randint(0, 1000))
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
    else:  # goal
        b.pos[:] = [rng.uniform(-1200, 1200), sy * rng.uniform(4200, 5000), rng.uniform(150, 900)]
        b.vel[:] = [rng.gauss(0, 800), sy * rng.uniform(500, 3000), rng.gauss(0, 500)]
    b.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
    A.set_ball(b)
    return A


def sample_soccar_ball(ref, n_traj, ticks_per_traj, seed, stride=1):
    """Ball against the arena (floor, walls, corners, goals) with two idle
    cars resting on their wheels at kickoff: every `stride`-th tick of each
    trajectory becomes a sample."""
    rng = random.Random(seed)
    kinds = ("drop", "wall", "corner", "goal")
    out = []
    for k in range(n_traj):
        teams, presets = [0, 1], [rng.randint(0, 5), rng.randint(0, 5)]
        A = _soccar_scenario(ref, rng, kinds[k % len(kinds)], teams, presets)
        zero = [Controls() for _ in teams]
        for t in range(ticks_per_traj):
            if t % stride == 0:
                out.append(_one_tick(A, MODE_SOCCAR, teams, presets, zero))
            else:
                A.step(1)
        A.close()
    return out


def sample_soccar_drive(ref, n_traj, ticks_per_traj, seed, n_cars=4, stride=1):
    """Cars driving from a random kickoff with random controls held for
    10-60 ticks (throttle mostly forward, no jumps until the first redraw):
    wheels on the floor and ramps, boost pads, car-car bumps."""
    rng = random.Random(seed)
    out = []
    for _ in range(n_traj):
        A = ref.RefArena(MODE_SOCCAR)
        teams = [i % 2 for i in range(n_cars)]
        presets = [rng.randint(0, 5) for _ in range(n_cars)]
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 1000))
        ctrls = [rand_controls(rng) for _ in range(n_cars)]
        for c in ctrls:
            c.throttle = 1.0
            c.jump = 0
        hold = [rng.randint(10, 60) for _ in range(n_cars)]
        for t in range(ticks_per_traj):
            for i in range(n_cars):
                hold[i] -= 1
                if hold[i] <= 0:
                    ctrls[i] = rand_controls(rng)
                    ctrls[i].throttle = rng.choice([1.0, 1.0, 0.5, -1.0])
                    hold[i] = rng.randint(10, 60)
            if t % stride == 0:
                out.append(_one_tick(A, MODE_SOCCAR, teams, presets, ctrls))
            else:
                for i in range(n_cars):
                    A.set_controls(i, ctrls[i])
                A.step(1)
        A.close()
    return out


def sample_soccar_melee(ref, n_traj, ticks_per_traj, seed, n_cars=4, stride=1, near_walls=False, mutators=None,
                        mutators_first=False, configs=None):
    """Cars placed 250-700 uu around the ball and driven at it (some
    supersonic, some jumping): car-ball hits with the extra impulse,
    car-car box-box contacts, bumps and demolitions, wheels resting on
    other cars.  With near_walls the ball starts by a side wall, a back wall
    or a corner, so cars drive onto the mesh's fillets and walls.  With
    `mutators` the arena's MutatorConfig is set before its cars are added
    (`mutators_first`) or once they are.  With `configs` (CarConfigs) each car
    takes one of them at random instead of a preset."""
    rng = random.Random(seed)
    out = []
    for _ in range(n_traj):
        A = ref.RefArena(MODE_SOCCAR)
        teams = [i % 2 for i in range(n_cars)]
        presets = [rng.randint(0, 5) for _ in range(n_cars)]
        cfgs = [rng.choice(configs) for _ in range(n_cars)] if configs else None
        if mutators is not None and mutators_first:
            A.set_mutators(mutators)
        for i, (t, p) in enumerate(zip(teams, presets)):
            if cfgs:
                A.add_car_config(t, cfgs[i])
            else:
                A.add_car(t, p)
        if mutators is not None and not mutators_first:
            A.set_mutators(mutators)
        bx, by = rng.uniform(-3000, 3000), rng.uniform(-4000, 4000)
        if near_walls:
            bx, by = rng.choice([-1, 1]) * rng.uniform(3300, 3800), rng.uniform(-4500, 4500)
            if rng.random() < 0.4:
                bx, by = rng.uniform(-3600, 3600), rng.choice([-1, 1]) * rng.uniform(4300, 4800)
        b = A.get_ball()
        b.pos[:] = [bx, by, rng.choice([93.15, rng.uniform(100, 400)])]
        b.vel[:] = [rng.gauss(0, 300), rng.gauss(0, 300), 0]
        A.set_ball(b)
        for i in range(n_cars):
            s = A.get_car(i)
            ang, dist = rng.uniform(0, 2 * math.pi), rng.uniform(250, 700)
            x, y = bx + dist * math.cos(ang), by + dist * math.sin(ang)
            yaw = math.atan2(by - y, bx - x) + rng.gauss(0, 0.3)
            cy, sy = math.cos(yaw), math.sin(yaw)
            spd = rng.choice([0, 800, 1500, 2300])
            s.pos[:] = [x, y, 17.0]
            s.rot[:] = [cy, sy, 0, -sy, cy, 0, 0, 0, 1]
            s.vel[:] = [spd * cy, spd * sy, 0]
            s.ang_vel[:] = [0, 0, 0]
            s.flags = CF_ON_GROUND | 0x1E | (CF_SUPERSONIC if spd >= 2200 else 0)
            s.supersonic_time = 0.5 if spd >= 2200 else 0
            s.boost = 50
            A.set_car(i, s)
        A.step(2)  # let the wheels settle
        ctrls = [rand_controls(rng) for _ in range(n_cars)]
        for c in ctrls:
            c.throttle = 1.0
            c.boost = rng.random() < 0.5
            c.jump = rng.random() < 0.2
        hold = [rng.randint(10, 60) for _ in range(n_cars)]
        for t in range(ticks_per_traj):
            for i in range(n_cars):
                hold[i] -= 1
                if hold[i] <= 0:
                    ctrls[i] = rand_controls(rng)
                    hold[i] = rng.randint(10, 60)
            if t % stride == 0:
                out.append(_one_tick(A, MODE_SOCCAR, teams, presets, ctrls))
                out[-1].configs = cfgs
            else:
                for i in range(n_cars):
                    A.set_controls(i, ctrls[i])
                A.step(1)
        A.close()
    return out


# The reference normalises vectors with SSE rsqrtss plus one Newton step
# (btVector3::normalize, LinearMath/btVector3.h:304-340), whose last bits
# depend on the x86 CPU; the GPU divides by the length.  GJK/EPA (which stop
# at a 1e-4 squared distance) and the subsimplex ray cast can turn such ulp
# differences into visibly different contacts in rare configurations.
#
# Contact-rich sets therefore allow samples out of tolerance for this one
# mechanism only, and check each such sample against it with the host build
# of the same tick code (tests/cpp/tick_host.cpp):
#   * built with the GPU's normalisation, it must give the GPU's result
#     within the one-tick tolerances (so the GPU's departure is the
#     normalisation's), and
#   * built with the reference's SSE normalisation, it must match the oracle
#     on that sample (so everything but the normalisation agrees).  This leg
#     needs the rsqrtss of the CPU the oracle ran on: it is checked for live
#     samples, which the oracle stepped on this machine; for recorded
#     fixtures it is checked on the recording CPU by
#     test_tick_host.test_sse_host_build_reproduces_oracle_fixtures.
# At most CHAOTIC_FRACTION of a set may be out of tolerance even so.
CHAOTIC_FRACTION = 0.005


def explain_outliers(samples, results, mode, bad_idx, cmfs=None):
    """Check the out-of-tolerance samples `bad_idx` against the normalisation
    mechanism above (host builds on the arena's mesh files `cmfs`, None: the
    one-mesh synthetic arena).  Returns messages for the samples it does not
    explain."""
    import tick_host
    sub = [samples[i] for i in bad_idx]
    sse_res, _ = tick_host.run_samples(tick_host.get("sse"), sub, mode, cmfs=cmfs)
    plain_res, _ = tick_host.run_samples(tick_host.get("plain"), sub, mode, cmfs=cmfs)
    out = []
    for k, i in enumerate(bad_idx):
        s = samples[i]
        n_sse, m_sse = check_samples([s], [sse_res[k]])
        # the host build with the GPU's normalisation in the reference's place
        as_ref = Sample()
        as_ref.mode, as_ref.teams, as_ref.presets = s.mode, s.teams, s.presets
        as_ref.cars_out, as_ref.ball_out, as_ref.pads_out, as_ref.tick_out = plain_res[k]
        n_plain, m_plain = check_samples([as_ref], [results[i][:4]])
        if n_sse and s.live:
            out.append(f"sample {i}: the SSE host build departs from the oracle too: {m_sse[0]}")
        elif n_plain:
            out.append(f"sample {i}: the GPU departs from the host build with its own normalisation: {m_plain[0]}")
    return out


# One step of the rotator rounding (Math::RoundAngleUE3, src/Math/Math.cpp:
# 75-91): 4 of the 65536 units of a turn.
UE3_STEP = 4 * 2 * math.pi / 65536


def seeking_direction_rounding(ref_ball, got_ball):
    """Whether the only difference between two ball records is a Heatseeker
    seeking ball's direction, at most two rotator steps apart at equal speed
    (position following from it): the seeking turn rounds the direction to
    rotator units after two atan2f calls, and where the exact angle lies
    within an ulp of a rounding boundary the reference's glibc atan2f and the
    GPU's land one step apart.  The seeking words themselves must match."""
    bad = [b[0] for b in compare_ball(ref_ball, got_ball)]
    if not bad or set(bad) - {"ball.vel", "ball.pos"} or ref_ball.hs_y_target_dir == 0:
        return False
    a = [float(x) for x in ref_ball.vel]
    b = [float(x) for x in got_ball.vel]
    na, nb = math.sqrt(sum(x * x for x in a)), math.sqrt(sum(x * x for x in b))
    cosang = max(-1.0, min(1.0, sum(x * y for x, y in zip(a, b)) / (na * nb)))
    return abs(na - nb) <= 1e-4 * na and math.acos(cosang) <= 2.05 * UE3_STEP


def assert_parity(samples, results, mode, chaotic=False, cmfs=None):
    """Every sample within the tolerances, or (chaotic=True) at most
    CHAOTIC_FRACTION of them out, each explained by the normalisation
    mechanism (explain_outliers, on the arena's mesh files `cmfs`).  Prints
    and returns the out-of-tolerance count."""
    nbad, msgs, bad_idx = check_samples(samples, results, return_indices=True)
    print(f"\n  {nbad}/{len(samples)} samples out of tolerance", flush=True)
    if nbad == 0:
        return 0
    assert chaotic, f"{nbad}/{len(samples)} samples off\n" + "\n".join(msgs)
    allowed = int(CHAOTIC_FRACTION * len(samples))
    assert nbad <= allowed, f"{nbad}/{len(samples)} samples off (allowed {allowed})\n" + "\n".join(msgs)
    unexplained = explain_outliers(samples, results, mode, bad_idx, cmfs=cmfs)
    assert not unexplained, "\n".join(unexplained + msgs)
    print(f"  all {nbad} explained by the normalisation rounding (host builds)", flush=True)
    return nbad


def sample_hoops(ref, n_traj, ticks, seed, stride):
    """Hoops trajectories (GameMode::HOOPS in the synthetic Hoops arena,
    rocketsim_amd.arena_mesh.hoops_cmfs), cycling through four scenarios: the
    kickoff, whose frozen ball is launched 0.265 s after the reset (every tick
    around the launch is a sample); a ball dropped through a rim into its net
    (a goal); cars driven under a hoop through its net, which only the ball
    collides with; cars driven at a ground ball (the Hoops extra hit lift).
    Every `stride`-th tick is a sample."""
    rng = random.Random(seed)
    out = []
    for k in range(n_traj):
        kind = ("kickoff", "hoop", "net", "melee")[k % 4]
        A = ref.RefArena(MODE_HOOPS)
        n = rng.choice([2, 4])
        teams = [i % 2 for i in range(n)]
        presets = [rng.randint(0, 5) for _ in range(n)]
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 1000))
        sy = rng.choice([-1, 1])
        b = A.get_ball()
        if kind in ("hoop", "net"):
            b.pos[:] = [rng.uniform(-300, 300), sy * (2900 + rng.uniform(-300, 300)), rng.uniform(500, 900)]
            b.vel[:] = [rng.gauss(0, 200), rng.gauss(0, 200), rng.uniform(-600, 0)]
        elif kind == "melee":
            b.pos[:] = [rng.uniform(-2000, 2000), rng.uniform(-2500, 2500), 96.4]
            b.vel[:] = [rng.gauss(0, 300), rng.gauss(0, 300), 0]
        if kind != "kickoff":
            b.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
            A.set_ball(b)
            for i in range(n):
                s = A.get_car(i)
                if kind == "melee":
                    ang, dist = rng.uniform(0, 2 * math.pi), rng.uniform(250, 700)
                    x, y = b.pos[0] + dist * math.cos(ang), b.pos[1] + dist * math.sin(ang)
                    yaw = math.atan2(b.pos[1] - y, b.pos[0] - x) + rng.gauss(0, 0.3)
                else:
                    x, y = rng.uniform(-500, 500), sy * rng.uniform(1900, 2300)
                    yaw = math.atan2(sy, 0.0) + rng.gauss(0, 0.3)
                cy, syw = math.cos(yaw), math.sin(yaw)
                spd = rng.choice([0, 800, 1500])
                s.pos[:] = [x, y, 17.0]
                s.rot[:] = [cy, syw, 0, -syw, cy, 0, 0, 0, 1]
                s.vel[:] = [spd * cy, spd * syw, 0]
                s.ang_vel[:] = [0, 0, 0]
                s.flags = CF_ON_GROUND | 0x1E
                s.boost = 50
                A.set_car(i, s)
        ctrls = [rand_controls(rng) for _ in range(n)]
        for c in ctrls:
            c.throttle, c.jump = 1.0, 0
        hold = [rng.randint(10, 60) for _ in range(n)]
        for t in range(ticks):
            for i in range(n):
                hold[i] -= 1
                if hold[i] <= 0:
                    ctrls[i] = rand_controls(rng)
                    ctrls[i].throttle = 1.0
                    hold[i] = rng.randint(10, 60)
            if t % stride == 0 or (kind == "kickoff" and 28 <= t <= 36):
                out.append(_one_tick(A, MODE_HOOPS, teams, presets, ctrls))
            else:
                A.step(1)
        A.close()
    return out


def sample_heatseeker(ref, n_traj, ticks, seed, stride):
    """Heatseeker trajectories (GameMode::HEATSEEKER in the Soccar arena),
    cycling through three scenarios: cars driven at the kickoff ball, whose
    first touch starts the seeking; a ball already seeking a net at
    2900-4600 uu/s; a seeking ball about to hit the back wall of its target
    net, which turns it round.  Every `stride`-th tick is a sample."""
    rng = random.Random(seed)
    out = []
    for k in range(n_traj):
        kind = ("kickoff", "seeking", "back_wall")[k % 3]
        A = ref.RefArena(MODE_HEATSEEKER)
        n = rng.choice([2, 4])
        teams = [i % 2 for i in range(n)]
        presets = [rng.randint(0, 5) for _ in range(n)]
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 1000))
        b = A.get_ball()
        if kind == "seeking":
            d = rng.choice([-1.0, 1.0])
            b.pos[:] = [rng.uniform(-3000, 3000), rng.uniform(-3500, 3500), rng.uniform(200, 1500)]
            ang = rng.uniform(0, 2 * math.pi)
            spd = rng.uniform(1500, 3500)
            b.vel[:] = [spd * math.cos(ang), spd * math.sin(ang), rng.gauss(0, 300)]
            b.hs_y_target_dir, b.hs_cur_target_speed = d, rng.choice([2900.0, 2985.0, 3500.0, 4600.0])
            b.hs_time_since_hit = rng.uniform(0, 3)
        elif kind == "back_wall":
            d = rng.choice([-1.0, 1.0])
            b.pos[:] = [rng.choice([-1, 1]) * rng.uniform(1100, 3000), d * rng.uniform(4500, 4800), rng.uniform(300, 1500)]
            b.vel[:] = [rng.gauss(0, 300), d * rng.uniform(2000, 4000), rng.gauss(0, 300)]
            b.hs_y_target_dir, b.hs_cur_target_speed, b.hs_time_since_hit = d, 3500.0, 0.5
        if kind != "kickoff":
            b.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
            A.set_ball(b)
        ctrls = [rand_controls(rng) for _ in range(n)]
        for c in ctrls:
            c.throttle, c.jump = 1.0, 0
        hold = [rng.randint(10, 60) for _ in range(n)]
        for t in range(ticks):
            for i in range(n):
                hold[i] -= 1
                if hold[i] <= 0:
                    ctrls[i] = rand_controls(rng)
                    ctrls[i].throttle = rng.choice([1.0, 1.0, 0.5])
                    hold[i] = rng.randint(10, 60)
            if t % stride == 0:
                out.append(_one_tick(A, MODE_HEATSEEKER, teams, presets, ctrls))
            else:
                for i in range(n):
                    A.set_controls(i, ctrls[i])
                A.step(1)
        A.close()
    return out


_pool = {}


def pooled_arena(ref, mode, teams, presets):
    """An oracle arena with cars of `teams` / `presets`, reused across calls
    (creating a reference Soccar arena takes about 0.35 s); the caller sets
    every state it relies on."""
    key = (mode, tuple(teams), tuple(presets))
    if key not in _pool:
        A = ref.RefArena(mode)
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        _pool[key] = A
    return _pool[key]


def sample_from_state(ref, mode, teams, presets, cars, ball, pads, tick, ctrls):
    """A teacher-forced sample whose "before" state is given in API units (for
    instance a state the GPU reached in a long batched run): an oracle arena
    is set to it and stepped once with `ctrls`."""
    A = pooled_arena(ref, mode, teams, presets)
    for i, c in enumerate(cars):
        A.set_car(i, c)
    A.set_ball(ball)
    if mode in ARENA_MODES:
        A.set_pads(pads[:len(A.get_pads())])
    A.tick_count = tick
    return _one_tick(A, mode, teams, presets, ctrls)


def _arena_pads(A, mode):
    """The oracle arena's boost pads in the batch's NUM_PADS slots: a Hoops
    arena has 20 and leaves the rest, like an arena without pads, at their
    initial state."""
    pads = A.get_pads() if mode in ARENA_MODES else []
    return pads + [PadState(1, 0.0, 0) for _ in range(NUM_PADS - len(pads))]


def _one_tick(A, mode, teams, presets, ctrls):
    """One teacher-forced sample.  The oracle is first set to its own state as
    read through the API (positions and velocities in uu), repeatedly until
    reading it back gives the same record, so that it steps from exactly the
    Bullet-unit values the GPU derives from that record (one uu -> bt -> uu
    round trip can move a value by one ulp)."""
    n = len(teams)
    for _ in range(8):
        before = [bytes(A.get_car(i)) for i in range(n)] + [bytes(A.get_ball())]
        for i in range(n):
            A.set_car(i, A.get_car(i))
        A.set_ball(A.get_ball())
        after = [bytes(A.get_car(i)) for i in range(n)] + [bytes(A.get_ball())]
        if after == before:
            break
    s = Sample()
    s.live = True
    s.mode = mode
    s.teams = list(teams)
    s.presets = list(presets)
    s.car_order = A.car_order()
    s.cars_in = [A.get_car(i) for i in range(n)]
    s.ball_in = A.get_ball()
    s.pads_in = _arena_pads(A, mode)
    s.tick_in = A.tick_count
    s.ctrls = [Controls.from_buffer_copy(c) for c in ctrls]
    for i in range(n):
        A.set_controls(i, ctrls[i])
    A.pop_events()
    A.step(1)
    s.events = [(e.type, e.tick, e.a, e.b) for e in A.pop_events()]
    s.cars_out = [A.get_car(i) for i in range(n)]
    s.ball_out = A.get_ball()
    s.pads_out = _arena_pads(A, mode)
    s.tick_out = A.tick_count
    return s


# -- fixture (de)serialisation: raw struct bytes in an .npz ------------------

def _pack(structs, cls):
    return np.frombuffer(b"".join(bytes(x) for x in structs), dtype=np.uint8) if structs else np.zeros(0, np.uint8)


def _unpack(arr, cls, count=None):
    """Records of `cls` from packed bytes.  With `count`, the stored record
    size is len / count, and records written before fields were appended to
    the struct are zero-padded (the ball's Heatseeker fields)."""
    raw = arr.tobytes()
    sz = C.sizeof(cls)
    rec = len(raw) // count if count else sz
    return [cls.from_buffer_copy(raw[i:i + rec].ljust(sz, b"\0")) for i in range(0, len(raw), rec)]


def save_samples(path, samples):
    ncars = np.array([len(s.teams) for s in samples], dtype=np.int32)
    np.savez_compressed(
        path,
        mode=np.array([s.mode for s in samples], dtype=np.int32),
        ncars=ncars,
        teams=np.array(sum((s.teams for s in samples), []), dtype=np.uint8),
        presets=np.array(sum((s.presets for s in samples), []), dtype=np.uint8),
        car_order=np.array(sum((s.car_order for s in samples), []), dtype=np.uint8),
        cars_in=_pack(sum((s.cars_in for s in samples), []), CarState),
        cars_out=_pack(sum((s.cars_out for s in samples), []), CarState),
        ctrls=_pack(sum((s.ctrls for s in samples), []), Controls),
        ball_in=_pack([s.ball_in for s in samples], BallState),
        ball_out=_pack([s.ball_out for s in samples], BallState),
        pads_in=_pack(sum((list(s.pads_in) for s in samples), []), PadState),
        pads_out=_pack(sum((list(s.pads_out) for s in samples), []), PadState),
        tick_in=np.array([s.tick_in for s in samples], dtype=np.uint64),
        tick_out=np.array([s.tick_out for s in samples], dtype=np.uint64),
    )


def load_samples(path):
    z = np.load(path, allow_pickle=False)
    ncars = z["ncars"]
    teams, presets = list(z["teams"]), list(z["presets"])
    order = list(z["car_order"])
    cars_in, cars_out = _unpack(z["cars_in"], CarState), _unpack(z["cars_out"], CarState)
    ctrls = _unpack(z["ctrls"], Controls)
    ball_in, ball_out = _unpack(z["ball_in"], BallState, len(ncars)), _unpack(z["ball_out"], BallState, len(ncars))
    pads_in, pads_out = _unpack(z["pads_in"], PadState), _unpack(z["pads_out"], PadState)
    out = []
    k = 0
    for i, n in enumerate(ncars):
        s = Sample()
        s.mode = int(z["mode"][i])
        s.teams = [int(t) for t in teams[k:k + n]]
        s.presets = [int(p) for p in presets[k:k + n]]
        s.car_order = [int(o) for o in order[k:k + n]]
        s.cars_in, s.cars_out, s.ctrls = cars_in[k:k + n], cars_out[k:k + n], ctrls[k:k + n]
        s.ball_in, s.ball_out = ball_in[i], ball_out[i]
        s.pads_in = pads_in[i * NUM_PADS:(i + 1) * NUM_PADS]
        s.pads_out = pads_out[i * NUM_PADS:(i + 1) * NUM_PADS]
        s.tick_in, s.tick_out = int(z["tick_in"][i]), int(z["tick_out"][i])
        out.append(s)
        k += n
    return out


def run_on_gpu(samples, mode, kernel=0, per_block=0, auto_reset=False, seed=0, mutators=None, mutators_first=False,
               cmfs=None, ball_rotation=False, custom_pads=None):
    """Step every sample once on the GPU, one sample per arena.  `kernel` and
    `per_block` pick the step kernel and arenas per workgroup (0: the
    library's automatic choice); `auto_reset` / `seed` set the batch's kickoff
    reset on goal; `mutators` is every arena's MutatorConfig, given at
    creation (`mutators_first`: before the cars are added) or set after;
    `cmfs` the arena's mesh files (None: the one-mesh synthetic arena)."""
    from rocketsim_amd.batch import ArenaBatch, ensure_meshes
    if mode in ARENA_MODES:
        ensure_meshes(cmfs, mode)
    cpa = [len(s.teams) for s in samples]
    car_configs = None
    if any(getattr(s, "configs", None) for s in samples):
        from tick_host import preset_config
        car_configs = sum((list(s.configs) if getattr(s, "configs", None) else [preset_config(p) for p in s.presets]
                           for s in samples), [])
    batch = ArenaBatch(cpa, sum((s.teams for s in samples), []), sum((s.presets for s in samples), []),
                       mode=mode, auto_reset=auto_reset, seed=seed,
                       mutators=mutators if mutators_first else None, car_configs=car_configs,
                       ball_rotation=ball_rotation, custom_pads=custom_pads)
    try:
        batch.set_kernel(kernel)
        batch.set_launch(per_block)
        if mutators is not None and not mutators_first:
            batch.set_mutator_config(mutators)
        batch.set_cars(sum((list(s.cars_in) for s in samples), []))
        batch.set_balls([s.ball_in for s in samples])
        batch.set_pads(sum((list(s.pads_in) for s in samples), []))
        batch.set_controls(sum((list(s.ctrls) for s in samples), []))
        batch.set_tick_counts([s.tick_in for s in samples])
        batch.set_car_order(sum((list(s.car_order) for s in samples), []))
        batch.enable_events(1 << 16)
        batch.step(1)
        events, lost = batch.poll_events(1 << 16)
        assert lost == 0
        cars = batch.get_cars()
        balls = batch.get_balls()
        pads = batch.get_pads()
        ticks = batch.get_tick_counts()
    finally:
        batch.close()
    per_arena = {}
    for e in events:
        per_arena.setdefault(e.arena, []).append((e.type, e.tick, e.a, e.b))
    results = []
    k = 0
    for i, s in enumerate(samples):
        n = len(s.teams)
        results.append((cars[k:k + n], balls[i], pads[i * NUM_PADS:(i + 1) * NUM_PADS], ticks[i],
                        per_arena.get(i, [])))
        k += n
    return results


def check_samples(samples, results, max_report=10, return_indices=False):
    """Return (n_bad_samples, messages[, indices of the bad samples])."""
    msgs = []
    nbad = 0
    bad_idx = []
    for i, (s, res) in enumerate(zip(samples, results)):
        cars, ball, pads, tick = res[:4]
        bad = []
        if s.events is not None and len(res) > 4 and s.events != res[4]:
            bad.append(("events", s.events, res[4], None))
        for j in range(len(s.teams)):
            bad += [(f"car{j}." + b[0],) + b[1:] for b in compare_car(s.cars_out[j], cars[j])]
        bad += compare_ball(s.ball_out, ball)
        if bp_order(s.ball_out, s.cars_out) != bp_order(ball, cars):
            bad.append(("bp_order", bp_order(s.ball_out, s.cars_out), bp_order(ball, cars), None))
        if s.mode in ARENA_MODES:
            bad += compare_pads(s.pads_out, pads)
        if tick != s.tick_out:
            bad.append(("tick", s.tick_out, tick, None))
        if bad:
            nbad += 1
            bad_idx.append(i)
            if len(msgs) < max_report:
                msgs.append(f"sample {i}: " + "; ".join(f"{b[0]} ref={b[1]} got={b[2]} err={b[3]}" for b in bad[:6]))
    if return_indices:
        return nbad, msgs, bad_idx
    return nbad, msgs
