"""ctypes loader for the oracle: the reference RocketSim compiled from its own
sources into oracle/_ref/librsref.so (recipe: oracle/Makefile, C entry points:
oracle/ref_harness.cpp).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may use this module, and only as the checker or
the timed CPU baseline -- never as the product path.
"""
import ctypes as C
import os

from rocketsim_amd.state import BallState, CarConfig, CarState, Controls, MutatorConfig, PadState

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "librsref.so")

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.rsref_last_error.restype = C.c_char_p
        L.rsref_init.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_int]
        L.rsref_init_modes.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_int,
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_int]
        L.rsref_arena_create.restype = C.c_void_p
        L.rsref_arena_create.argtypes = [C.c_int, C.c_float]
        L.rsref_arena_create_cfg.restype = C.c_void_p
        L.rsref_arena_create_cfg.argtypes = [C.c_int, C.c_float, C.c_int, C.POINTER(C.c_float), C.c_int]
        L.rsref_arena_destroy.argtypes = [C.c_void_p]
        L.rsref_add_car.restype = C.c_uint32
        L.rsref_add_car.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.rsref_num_cars.argtypes = [C.c_void_p]
        L.rsref_car_id.restype = C.c_uint32
        L.rsref_car_id.argtypes = [C.c_void_p, C.c_int]
        L.rsref_set_controls.argtypes = [C.c_void_p, C.c_int, C.POINTER(Controls)]
        L.rsref_get_car.argtypes = [C.c_void_p, C.c_int, C.POINTER(CarState)]
        L.rsref_set_car.argtypes = [C.c_void_p, C.c_int, C.POINTER(CarState)]
        L.rsref_get_ball.argtypes = [C.c_void_p, C.POINTER(BallState)]
        L.rsref_set_ball.argtypes = [C.c_void_p, C.POINTER(BallState)]
        L.rsref_num_pads.argtypes = [C.c_void_p]
        L.rsref_get_pad.argtypes = [C.c_void_p, C.c_int, C.POINTER(PadState)]
        L.rsref_set_pad.argtypes = [C.c_void_p, C.c_int, C.POINTER(PadState)]
        L.rsref_get_tick_count.restype = C.c_uint64
        L.rsref_get_tick_count.argtypes = [C.c_void_p]
        L.rsref_set_tick_count.argtypes = [C.c_void_p, C.c_uint64]
        L.rsref_step.argtypes = [C.c_void_p, C.c_int]
        L.rsref_car_order.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        from rocketsim_amd.state import Event
        L.rsref_pop_events.argtypes = [C.c_void_p, C.POINTER(Event), C.c_int]
        L.rsref_is_ball_scored.argtypes = [C.c_void_p]
        L.rsref_reset_kickoff.argtypes = [C.c_void_p, C.c_int]
        L.rsref_pad_config.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.rsref_ball_pred_create.restype = C.c_void_p
        L.rsref_ball_pred_create.argtypes = [C.c_void_p, C.c_int]
        L.rsref_ball_pred_destroy.argtypes = [C.c_void_p]
        L.rsref_ball_pred_update_from_arena.argtypes = [C.c_void_p, C.c_void_p]
        L.rsref_ball_pred_update_manual.argtypes = [C.c_void_p, C.POINTER(BallState), C.c_int]
        L.rsref_ball_pred_force.argtypes = [C.c_void_p, C.POINTER(BallState)]
        L.rsref_ball_pred_get.argtypes = [C.c_void_p, C.POINTER(BallState), C.c_int]
        L.rsref_ball_pred_for_time.argtypes = [C.c_void_p, C.c_float, C.POINTER(BallState)]
        L.rsref_serialize.restype = C.c_int64
        L.rsref_serialize.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.rsref_deserialize.restype = C.c_void_p
        L.rsref_deserialize.argtypes = [C.c_char_p, C.c_uint64]
        L.rsref_last_car_id.restype = C.c_uint32
        L.rsref_last_car_id.argtypes = [C.c_void_p]
        L.rsref_get_mutators.argtypes = [C.c_void_p, C.POINTER(MutatorConfig)]
        L.rsref_add_car_config.restype = C.c_uint32
        L.rsref_add_car_config.argtypes = [C.c_void_p, C.c_int, C.POINTER(CarConfig)]
        L.rsref_car_config_preset.argtypes = [C.c_int, C.POINTER(CarConfig)]
        L.rsref_set_mutators.argtypes = [C.c_void_p, C.POINTER(MutatorConfig)]
        L.rsref_car_team.argtypes = [C.c_void_p, C.c_int]
        L.rsref_get_controls.argtypes = [C.c_void_p, C.c_int, C.POINTER(Controls)]
        from rocketsim_amd.state import GameEventConfig
        L.rsref_tracker_create.restype = C.c_void_p
        L.rsref_tracker_create.argtypes = [C.POINTER(GameEventConfig)]
        L.rsref_tracker_destroy.argtypes = [C.c_void_p]
        L.rsref_tracker_update.argtypes = [C.c_void_p, C.c_void_p]
        L.rsref_tracker_reset.argtypes = [C.c_void_p]
        L.rsref_tracker_pop.argtypes = [C.c_void_p, C.POINTER(Event), C.c_int]
        L.rsref_run_workload.restype = C.c_double
        L.rsref_run_workload.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        _lib = L
    return _lib


_initialised = False


def car_config_preset(preset):
    """The reference's CAR_CONFIG_* preset as an rsg_car_config."""
    c = CarConfig()
    if lib().rsref_car_config_preset(preset, C.byref(c)) != 0:
        raise ValueError(preset)
    return c


def _blob_args(meshes):
    n = len(meshes)
    bufs = [C.create_string_buffer(m, len(m)) for m in meshes]
    ptrs = (C.c_void_p * max(n, 1))(*[C.cast(b, C.c_void_p) for b in bufs])
    ptrs._keep = bufs
    return ptrs, (C.c_uint64 * max(n, 1))(*[len(m) for m in meshes]), n


def init(meshes=None, hoops_meshes=None):
    """RocketSim::InitFromMem with the given .cmf blobs (bytes) for SOCCAR
    and for HOOPS, by default the procedural arenas of
    rocketsim_amd.arena_mesh.  The reference initialises once per process;
    later calls are no-ops."""
    global _initialised
    if _initialised:
        return
    from rocketsim_amd import arena_mesh
    if meshes is None:
        meshes = [arena_mesh.soccar_cmf()]
    if hoops_meshes is None:
        hoops_meshes = arena_mesh.hoops_cmfs()
    L = lib()
    if L.rsref_init_modes(*_blob_args(meshes), *_blob_args(hoops_meshes)) != 0:
        raise RuntimeError(L.rsref_last_error().decode())
    _initialised = True


class RefArena:
    """One reference Arena with cars added in a fixed order."""

    def __init__(self, mode, tick_rate=120.0, no_ball_rot=True, custom_pads=None):
        """custom_pads: ArenaConfig::customBoostPads as [(x, y, z, is_big), ...]."""
        L = lib()
        flat = [float(v) for p in (custom_pads or []) for v in (p[0], p[1], p[2], 1.0 if p[3] else 0.0)]
        buf = (C.c_float * max(len(flat), 1))(*flat)
        self.h = L.rsref_arena_create_cfg(mode, tick_rate, 1 if no_ball_rot else 0, buf, len(flat) // 4)
        if not self.h:
            raise RuntimeError(L.rsref_last_error().decode())
        self.L = L

    def add_car(self, team, preset):
        cid = self.L.rsref_add_car(self.h, team, preset)
        if cid == 0:
            raise RuntimeError(self.L.rsref_last_error().decode())
        return cid

    @property
    def num_cars(self):
        return self.L.rsref_num_cars(self.h)

    def get_car(self, i):
        s = CarState()
        self.L.rsref_get_car(self.h, i, C.byref(s))
        return s

    def set_car(self, i, s):
        self.L.rsref_set_car(self.h, i, C.byref(s))

    def get_ball(self):
        s = BallState()
        self.L.rsref_get_ball(self.h, C.byref(s))
        return s

    def set_ball(self, s):
        self.L.rsref_set_ball(self.h, C.byref(s))

    def get_pads(self):
        out = []
        for i in range(self.L.rsref_num_pads(self.h)):
            p = PadState()
            self.L.rsref_get_pad(self.h, i, C.byref(p))
            out.append(p)
        return out

    def set_pads(self, pads):
        for i, p in enumerate(pads):
            self.L.rsref_set_pad(self.h, i, C.byref(p))

    def set_controls(self, i, c):
        self.L.rsref_set_controls(self.h, i, C.byref(c))

    def pop_events(self):
        """Goal / bump / demo callbacks fired since the last call."""
        from rocketsim_amd.state import Event
        out = (Event * 256)()
        n = self.L.rsref_pop_events(self.h, out, 256)
        return list(out[:n])

    def serialize(self):
        """Arena::Serialize bytes (src/Sim/Arena/Arena.cpp:546-573)."""
        n = self.L.rsref_serialize(self.h, None, 0)
        buf = (C.c_uint8 * n)()
        self.L.rsref_serialize(self.h, buf, n)
        return bytes(buf)

    @classmethod
    def deserialize(cls, data):
        """Arena::DeserializeNew (Arena.cpp:575-644); cars in id order."""
        self = cls.__new__(cls)
        self.L = lib()
        self.h = self.L.rsref_deserialize(bytes(data), len(data))
        if not self.h:
            raise RuntimeError(self.L.rsref_last_error().decode())
        return self

    def add_car_config(self, team, cfg):
        cid = self.L.rsref_add_car_config(self.h, team, C.byref(cfg))
        if cid == 0:
            raise RuntimeError("rsref_add_car_config failed")
        return cid

    def get_mutators(self):
        m = MutatorConfig()
        self.L.rsref_get_mutators(self.h, C.byref(m))
        return m

    def set_mutators(self, m):
        self.L.rsref_set_mutators(self.h, C.byref(m))

    def last_car_id(self):
        return self.L.rsref_last_car_id(self.h)

    def car_team(self, i):
        return self.L.rsref_car_team(self.h, i)

    def get_controls(self, i):
        c = Controls()
        self.L.rsref_get_controls(self.h, i, C.byref(c))
        return c

    def car_order(self):
        """Iteration order of the reference's car set (creation indices)."""
        out = (C.c_int32 * 8)()
        n = self.L.rsref_car_order(self.h, out)
        return list(out)[:n]

    @property
    def tick_count(self):
        return self.L.rsref_get_tick_count(self.h)

    @tick_count.setter
    def tick_count(self, t):
        self.L.rsref_set_tick_count(self.h, t)

    def step(self, ticks=1):
        if self.L.rsref_step(self.h, ticks) != 0:
            raise RuntimeError(self.L.rsref_last_error().decode())

    def is_ball_scored(self):
        return bool(self.L.rsref_is_ball_scored(self.h))

    def reset_kickoff(self, seed=-1):
        self.L.rsref_reset_kickoff(self.h, seed)

    def close(self):
        if self.h:
            self.L.rsref_arena_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefBallPredTracker:
    """The reference's BallPredTracker (src/Sim/BallPredTracker/BallPredTracker.h)
    over a RefArena; constructing it predicts from the arena's ball."""

    def __init__(self, arena, num_pred_ticks):
        self.L = lib()
        self.h = self.L.rsref_ball_pred_create(arena.h, num_pred_ticks)
        self.num_pred_ticks = num_pred_ticks

    def update_pred_from_arena(self, arena):
        self.L.rsref_ball_pred_update_from_arena(self.h, arena.h)

    def update_pred_manual(self, state, ticks_since_last_update):
        self.L.rsref_ball_pred_update_manual(self.h, C.byref(state), ticks_since_last_update)

    def force_update_all_pred(self, state):
        self.L.rsref_ball_pred_force(self.h, C.byref(state))

    def pred_data(self):
        out = (BallState * self.num_pred_ticks)()
        n = self.L.rsref_ball_pred_get(self.h, out, self.num_pred_ticks)
        return list(out[:n])

    def get_ball_state_for_time(self, t):
        out = BallState()
        self.L.rsref_ball_pred_for_time(self.h, t, C.byref(out))
        return out

    def close(self):
        if self.h:
            self.L.rsref_ball_pred_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefGameEventTracker:
    """The reference's GameEventTracker (src/Sim/GameEventTracker/) with its
    shot, goal and save callbacks recorded as Event records."""

    def __init__(self, config):
        self.L = lib()
        self.h = self.L.rsref_tracker_create(C.byref(config))

    def update(self, arena):
        self.L.rsref_tracker_update(self.h, arena.h)

    def reset_persistent_info(self):
        self.L.rsref_tracker_reset(self.h)

    def pop(self):
        from rocketsim_amd.state import Event
        out = (Event * 256)()
        n = self.L.rsref_tracker_pop(self.h, out, 256)
        return list(out[:n])

    def close(self):
        if self.h:
            self.L.rsref_tracker_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_workload(threads, arenas_per_thread, n_blue, n_orange, preset, ticks, seed, cpus=None):
    """Timed CPU baseline: returns (arena_ticks, wall_seconds).  `cpus` pins
    thread t to CPU cpus[t]."""
    L = lib()
    n = C.c_uint64(0)
    pin = (C.c_int * threads)(*cpus) if cpus else None
    secs = L.rsref_run_workload(threads, arenas_per_thread, n_blue, n_orange, preset, ticks, seed, C.byref(n), pin)
    return n.value, secs
