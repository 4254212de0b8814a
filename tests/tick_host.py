"""ctypes front end of the host build of the batched tick (tests/cpp/tick_host.cpp):
the code the step kernels run, compiled for the CPU with g++, in the GPU's
struct-of-arrays layout.  Used by the CPU tests as a second implementation
to explain where the GPU and the reference part."""
import ctypes as C
import os
import subprocess

from rocketsim_amd.state import (ARENA_MODES, BallState, CarConfig, CarState, Controls, MODE_HEATSEEKER, MODE_HOOPS,
                                 MODE_SOCCAR, MutatorConfig,
                                 NUM_PADS, PadState)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "tick_host.cpp")


def build(out_path, defines=(), extra=(), load_it=True):
    """Compile the host tick library with the given -D switches (and load it,
    unless load_it is False: a sanitizer build loads only under its runtime)."""
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", *["-D" + d for d in defines],
           *extra, "-I" + os.path.join(ROOT, "include"), SRC, "-o", out_path]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-4000:])
    return load(out_path) if load_it else out_path


_cache = {}
VARIANTS = {"sse": ["RSG_SSE_NORMALIZE"], "plain": [],
            # the world step as the cooperative kernel runs it (narrowphase
            # over 16 lanes, effects applied afterwards in pair order)
            "sse_split": ["RSG_SSE_NORMALIZE", "RSG_TEST_SPLIT_WORLD=16"],
            # ... keeping one effect per pair, so pairs with more run again in order
            "sse_split_1eff": ["RSG_SSE_NORMALIZE", "RSG_TEST_SPLIT_WORLD=16", "RSG_TEST_PAIR_EFFECTS=1"],
            # contact capacities low enough for ordinary play to reach them
            "small_manifolds": ["RSG_TEST_MAX_MANIFOLDS=1"], "small_rows": ["RSG_TEST_MAX_SOLVER_ROWS=4"]}


def get(variant):
    """The host build `variant` (VARIANTS), compiled once per process."""
    if variant not in _cache:
        import tempfile
        d = tempfile.mkdtemp(prefix="tick_host_")
        _cache[variant] = build(os.path.join(d, f"tick_host_{variant}.so"), VARIANTS[variant])
    return _cache[variant]


def load(path):
    L = C.CDLL(path)
    L.th_create.restype = C.c_void_p
    L.th_create.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                            C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_int, C.c_float,
                            C.POINTER(MutatorConfig), C.POINTER(CarConfig)]
    L.th_destroy.argtypes = [C.c_void_p]
    for name, t in (("cars", CarState), ("balls", BallState), ("pads", PadState)):
        getattr(L, "th_set_" + name).argtypes = [C.c_void_p, C.POINTER(t)]
        getattr(L, "th_get_" + name).argtypes = [C.c_void_p, C.POINTER(t)]
    L.th_set_controls.argtypes = [C.c_void_p, C.POINTER(Controls)]
    L.th_set_tick_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    L.th_get_tick_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    L.th_set_car_order.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.th_set_ball_rotation.argtypes = [C.c_void_p, C.c_int]
    L.th_set_custom_pads.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.th_get_status.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
    L.th_step.argtypes = [C.c_void_p, C.c_int]
    L.th_rsqrt_signature.argtypes = [C.POINTER(C.c_float)]
    L.th_set_mutators.argtypes = [C.c_void_p, C.POINTER(MutatorConfig)]
    return L


# rsqrtss of th_rsqrt_signature's inputs on the CPU that recorded the
# committed fixtures (Intel Xeon @ 2.70GHz), as float bit patterns.
RECORDING_CPU_RSQRT = [0x3f7ff000, 0x3f34f800, 0x3f13c800, 0x3fb4f800, 0x3f666000, 0x3eb91000, 0x41fd0000,
                       0x3db85000]


def same_rsqrt_as_recording_cpu():
    """Whether this CPU's rsqrtss agrees with the fixtures' recording CPU's
    on the signature inputs."""
    import struct
    out = (C.c_float * 8)()
    get("sse").th_rsqrt_signature(out)
    return [struct.unpack("<I", struct.pack("<f", x))[0] for x in out] == RECORDING_CPU_RSQRT


def _car_configs(samples):
    """The samples' cars' configs (a sample may carry `configs`, one per car),
    or None when every car uses a preset."""
    if not any(getattr(s, "configs", None) for s in samples):
        return None
    out = []
    for s in samples:
        cfgs = getattr(s, "configs", None)
        out += list(cfgs) if cfgs else [preset_config(p) for p in s.presets]
    return (CarConfig * len(out))(*out)


def preset_config(p):
    from rocketsim_amd import native
    c = CarConfig()
    native.check(native.lib().rsg_car_config_preset(p, C.byref(c)))
    return c


def mesh_args(cmfs):
    """th_create's (files, sizes, count) arguments for a list of .cmf file
    bytes (the buffers stay alive with the returned tuple's first item)."""
    cmfs = [bytes(c) for c in (cmfs or [])]
    bufs = [C.create_string_buffer(c, len(c)) for c in cmfs]
    n = len(cmfs)
    ptrs = (C.c_void_p * max(n, 1))(*[C.cast(b, C.c_void_p) for b in bufs])
    ptrs._keep = bufs
    return ptrs, (C.c_uint64 * max(n, 1))(*[len(c) for c in cmfs]), n


def run_samples(L, samples, mode, cmf=None, inject=None, mutators=None, mutators_first=False, cmfs=None,
                ball_rotation=False, custom_pads=None):
    """Step every teacher-forced sample once on the host build, one sample per
    arena (the host counterpart of parity.run_on_gpu).  `inject(cars, balls)`
    may edit the "before" records; `mutators` is every arena's MutatorConfig,
    set before its cars are added (`mutators_first`) or after, as
    Arena::SetMutatorConfig.  Returns (results, status words)."""
    if cmfs is None and cmf is not None:
        cmfs = [cmf]
    if mode in ARENA_MODES and cmfs is None:
        from rocketsim_amd import arena_mesh
        cmfs = arena_mesh.hoops_cmfs() if mode == MODE_HOOPS else [arena_mesh.soccar_cmf()]
    if mode not in ARENA_MODES:
        cmfs = []
    n = len(samples)
    cpa = [len(s.teams) for s in samples]
    teams = sum((list(s.teams) for s in samples), [])
    presets = sum((list(s.presets) for s in samples), [])
    total = sum(cpa)
    h = L.th_create(mode, n, (C.c_int * n)(*cpa), (C.c_uint8 * max(total, 1))(*teams),
                    (C.c_uint8 * max(total, 1))(*presets), *mesh_args(cmfs), 120.0,
                    C.byref(mutators) if mutators is not None and mutators_first else None,
                    _car_configs(samples))
    if not h:
        raise RuntimeError("th_create failed")
    try:
        if mutators is not None and not mutators_first:
            L.th_set_mutators(h, C.byref(mutators))
        L.th_set_ball_rotation(h, 1 if ball_rotation else 0)
        if custom_pads:
            from rocketsim_amd.state import BoostPadConfig
            arr = (BoostPadConfig * len(custom_pads))(
                *[BoostPadConfig((C.c_float * 3)(x, y, z), int(bool(big))) for x, y, z, big in custom_pads])
            L.th_set_custom_pads(h, arr, len(custom_pads))
        cars = (CarState * max(total, 1))(*sum((list(s.cars_in) for s in samples), []))
        balls = (BallState * n)(*[s.ball_in for s in samples])
        if inject:
            inject(cars, balls)
        L.th_set_cars(h, cars)
        L.th_set_balls(h, balls)
        L.th_set_pads(h, (PadState * (n * NUM_PADS))(*sum((list(s.pads_in) for s in samples), [])))
        L.th_set_controls(h, (Controls * max(total, 1))(*sum((list(s.ctrls) for s in samples), [])))
        L.th_set_tick_counts(h, (C.c_uint64 * n)(*[s.tick_in for s in samples]))
        L.th_set_car_order(h, (C.c_int32 * max(total, 1))(*sum((list(s.car_order) for s in samples), [])))
        L.th_step(h, 1)
        out_cars = (CarState * max(total, 1))()
        out_balls = (BallState * n)()
        out_pads = (PadState * (n * NUM_PADS))()
        ticks = (C.c_uint64 * n)()
        status = (C.c_uint32 * n)()
        L.th_get_cars(h, out_cars)
        L.th_get_balls(h, out_balls)
        L.th_get_pads(h, out_pads)
        L.th_get_tick_counts(h, ticks)
        L.th_get_status(h, status)
    finally:
        L.th_destroy(h)
    results, k = [], 0
    for i, s in enumerate(samples):
        c = len(s.teams)
        results.append((out_cars[k:k + c], out_balls[i], out_pads[i * NUM_PADS:(i + 1) * NUM_PADS], ticks[i]))
        k += c
    return results, list(status)
