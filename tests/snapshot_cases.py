"""Reference arenas for the snapshot-format tests (test infrastructure only):
arenas of the oracle stepped with random controls, and their state as an
ArenaSnapshot read through the oracle's getters."""
import ctypes as C
import random

import parity
from rocketsim_amd.state import ARENA_MODES, ArenaSnapshot, MODE_HEATSEEKER, MODE_SOCCAR

# Ball.h:45: DropshotInfo::lastDamageTick has no initializer, so the
# reference serializes whatever its memory holds.  It is the last field of the
# ball record, just before the mutator config (123 bytes for these modes).
MUTATOR_BYTES = 123


def mask_uninitialized(data, ref_len):
    """Zero the reference's uninitialized lastDamageTick in `data` (whose
    first ref_len bytes are in the reference's format)."""
    k = ref_len - MUTATOR_BYTES - 8
    return data[:k] + bytes(8) + data[k + 8:]


def snapshot_from_ref(A, presets, mode):
    s = ArenaSnapshot()
    s.mode = mode
    s.tick_time = C.c_float(1 / 120).value
    s.tick_count = A.tick_count
    s.last_car_id = A.last_car_id()
    order = A.car_order()
    s.n_cars = len(order)
    for k, i in enumerate(order):
        s.car_id[k] = A.L.rsref_car_id(A.h, i)
        s.car_team[k] = A.car_team(i)
        s.car_preset[k] = presets[i]
        s.controls[k] = A.get_controls(i)
        s.cars[k] = A.get_car(i)
    s.ball = A.get_ball()
    if mode in ARENA_MODES:
        for p, pd in enumerate(A.get_pads()):
            s.pads[p] = pd
    s.mutators = A.get_mutators()
    return s


def played_arenas(ref, mode, n_arenas, ticks, seed, snap_every):
    """Yields (arena, presets) after every `snap_every` ticks of play with
    random controls redrawn every 20 ticks (2v2, random presets; every 60
    ticks a car is driven into the ball and a blue car at supersonic speed
    into an orange one, so hits and demolitions happen)."""
    rng = random.Random(seed)
    for _ in range(n_arenas):
        A = ref.RefArena(mode)
        presets = [rng.randint(0, 5) for _ in range(4)]
        for i, p in enumerate(presets):
            A.add_car(i % 2, p)
        if mode in ARENA_MODES:
            A.reset_kickoff(rng.randint(0, 1000))
        for t in range(ticks):
            if t % 20 == 0:
                for i in range(4):
                    A.set_controls(i, parity.rand_controls(rng))
            if t % 60 == 30:
                # car 2 driven into the ball; car 0 (blue) at supersonic
                # speed into car 1 (orange), facing it along +x
                b = A.get_ball()
                c = A.get_car(2)
                c.pos[:] = [b.pos[0] - 250.0, b.pos[1], b.pos[2]]
                c.rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
                c.vel[:] = [2000.0, 0.0, 0.0]
                A.set_car(2, c)
                x0, y0 = rng.uniform(-2000, 1500), rng.uniform(-3000, 3000)
                for i, dx, vx in ((0, 0.0, 2290.0), (1, 260.0, 0.0)):
                    c = A.get_car(i)
                    c.pos[:] = [x0 + dx, y0, 17.0]
                    c.rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
                    c.vel[:] = [vx, 0.0, 0.0]
                    c.ang_vel[:] = [0.0, 0.0, 0.0]
                    A.set_car(i, c)
            A.step(1)
            if t % snap_every == snap_every - 1:
                yield A, presets
        A.close()
