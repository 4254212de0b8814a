"""Python front end of the batched tick step (wraps include/rsg.h).

`ArenaBatch` is the batched counterpart of the reference's Arena
(src/Sim/Arena/Arena.h): `step(ticks)` replaces `Arena::Step(ticks)`
(src/Sim/Arena/Arena.cpp:689-756) for every arena of the batch at once, and
the get/set methods replace Car/Ball/BoostPad Get/SetState.
"""
import ctypes as C

from . import native
from .state import (ARENA_MODES, ArenaSnapshot, BallState, BatchDesc, BoostPadConfig, CarState, Controls, Event, MODE_HEATSEEKER,
                    MODE_HOOPS, MODE_SOCCAR, MODE_THE_VOID,
                    CarConfig, MutatorConfig, NUM_PADS, PadState)


_mesh_files = {}  # arena kind (0: Soccar, 1: Hoops) -> the .cmf files loaded last, in order


def _kind(mode):
    return 1 if mode == MODE_HOOPS else 0


def _default_cmfs(mode):
    from . import arena_mesh
    return arena_mesh.hoops_cmfs() if mode == MODE_HOOPS else [arena_mesh.soccar_cmf()]


def init_meshes(cmfs, mode=MODE_SOCCAR):
    """Load an arena's collision meshes (RocketSim::InitFromMem,
    src/RocketSim.cpp:92-185): `cmfs` is the list of .cmf file bytes, in the
    order the reference's list would hold them, for the arena of `mode` (the
    Soccar arena for SOCCAR and HEATSEEKER, the Hoops arena for HOOPS); each
    becomes one static body of every arena stepped afterwards."""
    files = [bytes(c) for c in cmfs]
    bufs = [C.create_string_buffer(c, len(c)) for c in files]
    n = len(files)
    ptrs = (C.c_void_p * max(n, 1))(*[C.cast(b, C.c_void_p) for b in bufs])
    sizes = (C.c_uint64 * max(n, 1))(*[len(c) for c in files])
    native.check(native.lib().rsg_init_arena_meshes(mode, ptrs, sizes, n))
    _mesh_files[_kind(mode)] = files


def init_mesh(cmf=None):
    """Load a one-mesh Soccar arena: `cmf` is the bytes of one .cmf file; by
    default the procedural arena of rocketsim_amd.arena_mesh."""
    init_meshes([cmf] if cmf is not None else _default_cmfs(MODE_SOCCAR))


def ensure_meshes(cmfs=None, mode=MODE_SOCCAR):
    """init_meshes(cmfs, mode) unless exactly these files are loaded already
    (None: the mode's procedural arena)."""
    files = [bytes(c) for c in (cmfs if cmfs is not None else _default_cmfs(mode))]
    if _mesh_files.get(_kind(mode)) != files:
        init_meshes(files, mode)


def ensure_arena(mode):
    """The mode's procedural arena unless meshes for its arena are loaded."""
    if mode in ARENA_MODES and _kind(mode) not in _mesh_files:
        init_meshes(_default_cmfs(mode), mode)


def serialize_snapshot(snap):
    """Arena::Serialize (src/Sim/Arena/Arena.cpp:546-573) bytes of an
    ArenaSnapshot, followed by this library's extension block."""
    L = native.lib()
    n = native.check(L.rsg_snapshot_serialize(C.byref(snap), None, 0))
    buf = (C.c_uint8 * n)()
    native.check(L.rsg_snapshot_serialize(C.byref(snap), buf, n))
    return bytes(buf)


def deserialize_snapshot(data):
    """Arena::DeserializeNew (Arena.cpp:575-644) of snapshot bytes."""
    out = ArenaSnapshot()
    native.check(native.lib().rsg_snapshot_deserialize(bytes(data), len(data), C.byref(out)))
    return out


class ArenaBatch:
    def __init__(self, cars_per_arena, teams, presets, mode=MODE_SOCCAR, tick_rate=120.0,
                 seed=0, auto_reset=False, synth_controls=False, mutators=None, car_configs=None,
                 ball_rotation=False, custom_pads=None):
        """`presets`: RSG_PRESET_* per car, or None with `car_configs` (one
        CarConfig per car); `mutators`: every arena's MutatorConfig, set
        before its cars are added."""
        L = native.lib()
        ensure_arena(mode)
        n = len(cars_per_arena)
        total = sum(cars_per_arena)
        if car_configs is not None:
            presets = [0] * total
            if len(car_configs) != total:
                raise ValueError("car_configs must have one entry per car")
        if len(teams) != total or len(presets) != total:
            raise ValueError("teams/presets must have one entry per car")
        self._cfgs = (CarConfig * max(total, 1))(*car_configs) if car_configs is not None else None
        self._cpa = (C.c_int * n)(*cars_per_arena)
        self._team = (C.c_uint8 * max(total, 1))(*teams)
        self._preset = (C.c_uint8 * max(total, 1))(*presets)
        self._mut = mutators
        d = BatchDesc(n_arenas=n, mode=mode, tick_rate=tick_rate, seed=seed,
                      auto_reset=int(auto_reset), synth_controls=int(synth_controls),
                      cars_per_arena=self._cpa, car_team=self._team, car_preset=self._preset,
                      mutators=C.pointer(mutators) if mutators is not None else None,
                      car_configs=self._cfgs, ball_rotation=int(ball_rotation))
        if custom_pads:
            # ArenaConfig::customBoostPads: [(x, y, z, is_big), ...]
            self._custom = (BoostPadConfig * len(custom_pads))(
                *[BoostPadConfig((C.c_float * 3)(x, y, z), int(bool(big))) for x, y, z, big in custom_pads])
            d.custom_pads = self._custom
            d.n_custom_pads = len(custom_pads)
        h = L.rsg_batch_create(C.byref(d))
        if not h:
            raise RuntimeError(L.rsg_last_error().decode())
        self.h = h
        self.L = L
        self.n_arenas = n
        self.mode = mode
        self.tick_rate = tick_rate
        self.n_cars = total
        self.cars_per_arena = list(cars_per_arena)
        self.offsets = [L.rsg_batch_car_offset(h, a) for a in range(n)]

    @classmethod
    def uniform(cls, n_arenas, n_blue, n_orange, preset=0, **kw):
        per = n_blue + n_orange
        teams = ([0] * n_blue + [1] * n_orange) * n_arenas
        return cls([per] * n_arenas, teams, [preset] * (per * n_arenas), **kw)

    # -- snapshots (Arena::Serialize / DeserializeNew) ---------------------
    def get_snapshot(self, arena):
        out = ArenaSnapshot()
        native.check(self.L.rsg_batch_get_snapshot(self.h, arena, C.byref(out)))
        return out

    def set_snapshot(self, arena, snap):
        native.check(self.L.rsg_batch_set_snapshot(self.h, arena, C.byref(snap)))

    def serialize_arena(self, arena):
        """The arena in the reference's snapshot byte format."""
        return serialize_snapshot(self.get_snapshot(arena))

    def deserialize_arena(self, arena, data):
        self.set_snapshot(arena, deserialize_snapshot(data))

    # -- stepping ---------------------------------------------------------
    def step(self, ticks=1, stream=None):
        native.check(self.L.rsg_batch_step(self.h, ticks, stream))

    def sync(self):
        native.check(self.L.rsg_batch_sync(self.h))

    def reset_kickoff(self, arena=-1, seed=0):
        native.check(self.L.rsg_batch_reset_kickoff(self.h, arena, seed))

    # -- state ------------------------------------------------------------
    def get_cars(self):
        out = (CarState * max(self.n_cars, 1))()
        native.check(self.L.rsg_batch_get_cars(self.h, out))
        return out

    def set_cars(self, states):
        arr = (CarState * max(self.n_cars, 1))(*states)
        native.check(self.L.rsg_batch_set_cars(self.h, arr))

    def get_balls(self):
        out = (BallState * self.n_arenas)()
        native.check(self.L.rsg_batch_get_balls(self.h, out))
        return out

    def set_balls(self, states):
        arr = (BallState * self.n_arenas)(*states)
        native.check(self.L.rsg_batch_set_balls(self.h, arr))

    def get_pads(self):
        out = (PadState * (NUM_PADS * self.n_arenas))()
        native.check(self.L.rsg_batch_get_pads(self.h, out))
        return out

    def set_pads(self, states):
        arr = (PadState * (NUM_PADS * self.n_arenas))(*states)
        native.check(self.L.rsg_batch_set_pads(self.h, arr))

    def set_controls(self, controls):
        arr = (Controls * max(self.n_cars, 1))(*controls)
        native.check(self.L.rsg_batch_set_controls(self.h, arr))

    def get_controls(self):
        out = (Controls * max(self.n_cars, 1))()
        native.check(self.L.rsg_batch_get_controls(self.h, out))
        return out

    def set_launch(self, arenas_per_block):
        """Arenas per workgroup of the step kernel (0 = device default)."""
        native.check(self.L.rsg_batch_set_launch(self.h, int(arenas_per_block)))

    def enable_events(self, capacity=1 << 16):
        """Goal / bump / demolition events into a device ring of `capacity`."""
        native.check(self.L.rsg_batch_enable_events(self.h, capacity))

    def poll_events(self, max_out=1 << 16):
        """Events since the last poll, ordered by (arena, sequence); returns
        (list of Event, number lost)."""
        out = (Event * max_out)()
        lost = C.c_int()
        n = native.check(self.L.rsg_batch_poll_events(self.h, out, max_out, C.byref(lost)))
        return list(out[:n]), lost.value

    def graph_create(self, ticks_per_launch, launches):
        """Capture `launches` launches of `ticks_per_launch` ticks in a hipGraph."""
        native.check(self.L.rsg_batch_graph_create(self.h, ticks_per_launch, launches))

    def graph_launch(self, stream=None):
        native.check(self.L.rsg_batch_graph_launch(self.h, stream))

    def autotune(self, ticks=8, stream=None):
        """Time every kernel / launch shape on a copy of the state, keep the
        fastest; returns (kernel, arenas_per_block, ms_per_tick)."""
        k, pb, ms = C.c_int(), C.c_int(), C.c_float()
        native.check(self.L.rsg_batch_autotune(self.h, ticks, stream, C.byref(k), C.byref(pb), C.byref(ms)))
        return k.value, pb.value, ms.value

    def set_kernel(self, kernel):
        """KERNEL_AUTO, KERNEL_SCRATCH or KERNEL_LDS (rocketsim_amd.state)."""
        native.check(self.L.rsg_batch_set_kernel(self.h, int(kernel)))

    def get_kernel(self):
        return native.check(self.L.rsg_batch_get_kernel(self.h))

    def get_launch(self):
        return native.check(self.L.rsg_batch_get_launch(self.h))

    def get_car_order(self):
        """Per-arena car update order (flattened like the car arrays)."""
        out = (C.c_int32 * max(1, self.n_cars))()
        native.check(self.L.rsg_batch_get_car_order(self.h, out))
        return list(out)[:self.n_cars]

    def set_car_order(self, order):
        arr = (C.c_int32 * max(1, self.n_cars))(*order)
        native.check(self.L.rsg_batch_set_car_order(self.h, arr))

    def get_tick_counts(self):
        out = (C.c_uint64 * self.n_arenas)()
        native.check(self.L.rsg_batch_get_tick_counts(self.h, out))
        return list(out)

    def set_tick_counts(self, ticks):
        arr = (C.c_uint64 * self.n_arenas)(*ticks)
        native.check(self.L.rsg_batch_set_tick_counts(self.h, arr))

    def get_goals(self):
        b = (C.c_uint32 * self.n_arenas)()
        o = (C.c_uint32 * self.n_arenas)()
        t = (C.c_uint64 * self.n_arenas)()
        native.check(self.L.rsg_batch_get_goals(self.h, b, o, t))
        return list(b), list(o), list(t)

    def get_status(self):
        """Per-arena RSG_STATUS_* bits (non-finite state, contact or solver
        capacity exceeded) since the last clear_status()."""
        out = (C.c_uint32 * self.n_arenas)()
        native.check(self.L.rsg_batch_get_status(self.h, out, None))
        return list(out)

    def status_summary(self, out, stream=None):
        """Enqueues the device-side summary of the status words into `out`, a
        torch uint32/int32 tensor of 2 elements on the GPU: [OR of the bits,
        number of flagged arenas]."""
        native.check(self.L.rsg_batch_status_summary(self.h, C.c_void_p(out.data_ptr()), stream))

    def clear_status(self):
        native.check(self.L.rsg_batch_clear_status(self.h))

    def set_strict(self, on=True):
        """step() and graph_launch() raise when an arena carries a status bit,
        as the reference's Arena::Step throws on its own errors."""
        native.check(self.L.rsg_batch_set_strict(self.h, int(bool(on))))

    def set_mutator_config(self, cfg, arena=-1):
        """Arena::SetMutatorConfig on one arena, or on all (-1)."""
        native.check(self.L.rsg_batch_set_mutator_config(self.h, arena, C.byref(cfg)))

    def get_mutator_config(self, arena=0):
        out = MutatorConfig()
        native.check(self.L.rsg_batch_get_mutator_config(self.h, arena, C.byref(out)))
        return out

    def is_ball_scored(self):
        out = (C.c_uint8 * self.n_arenas)()
        native.check(self.L.rsg_batch_is_ball_scored(self.h, out))
        return [bool(x) for x in out]

    def close(self):
        if getattr(self, "h", None):
            self.L.rsg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["ArenaBatch", "MODE_HEATSEEKER", "MODE_SOCCAR", "MODE_THE_VOID"]
