"""The host build of the batched tick (tests/cpp/tick_host.cpp: the code the
step kernels run, compiled with g++) on the CPU.

  * Built with the reference's SSE normalisation it reproduces the oracle's
    recorded ticks on every committed fixture, with no exception: the tick
    code is the reference's arithmetic, and the one thing the GPU does
    differently is the last bits of a normalised vector.  (The parity tests
    use this build to explain each GPU sample out of tolerance.)  The same
    holds with the world step split as the cooperative kernel splits it
    (narrowphase shared by 16 lanes, contact effects applied afterwards in
    pair order, penetration queries run again in order), and with room for
    one effect per pair, so that every pair with more runs again in order.
  * Capacity overflow is signalled, not silent: with the contact capacities
    lowered, arenas that exceed them carry RSG_STATUS_CONTACT_OVERFLOW or
    RSG_STATUS_SOLVER_OVERFLOW, and every arena whose result differs from the
    full-capacity build carries one of them.
  * A non-finite state sets RSG_STATUS_NONFINITE on its arena only, and the
    tick code stays inside its arrays while stepping it (an AddressSanitizer
    build over 60 ticks).
"""
import os
import subprocess
import sys

import pytest

import parity
import tick_host
from rocketsim_amd.state import (MODE_HEATSEEKER, MODE_SOCCAR, MODE_THE_VOID, STATUS_CONTACT_OVERFLOW,
                                 STATUS_NONFINITE, STATUS_SOLVER_OVERFLOW)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

FIXTURES = [("void_tf.npz", MODE_THE_VOID), ("soccar_ball_tf.npz", MODE_SOCCAR),
            ("soccar_drive_tf.npz", MODE_SOCCAR), ("soccar_melee_tf.npz", MODE_SOCCAR),
            ("soccar_walls_tf.npz", MODE_SOCCAR), ("heatseeker_tf.npz", MODE_HEATSEEKER)]


@pytest.mark.parametrize("variant", ["sse", "sse_split", "sse_split_1eff"])
@pytest.mark.parametrize("name,mode", FIXTURES)
def test_sse_host_build_reproduces_oracle_fixtures(name, mode, variant):
    if not tick_host.same_rsqrt_as_recording_cpu():
        pytest.skip("this CPU's rsqrtss differs from the fixtures' recording CPU's")
    samples = parity.load_samples(os.path.join(HERE, "golden", name))
    results, status = tick_host.run_samples(tick_host.get(variant), samples, mode)
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, "\n".join(msgs)
    assert not any(status)


@pytest.mark.parametrize("variant,bit", [("small_manifolds", STATUS_CONTACT_OVERFLOW),
                                         ("small_rows", STATUS_SOLVER_OVERFLOW)])
def test_capacity_overflow_is_signalled(variant, bit):
    samples = parity.load_samples(os.path.join(HERE, "golden", "soccar_melee_tf.npz"))
    full, full_status = tick_host.run_samples(tick_host.get("plain"), samples, MODE_SOCCAR)
    small, small_status = tick_host.run_samples(tick_host.get(variant), samples, MODE_SOCCAR)
    assert not any(full_status)
    flagged = [i for i, s in enumerate(small_status) if s & bit]
    assert all(s in (0, bit) for s in small_status)
    assert 0 < len(flagged) < len(samples)
    differs = []
    for i, (a, b) in enumerate(zip(full, small)):
        same = (b"".join(bytes(c) for c in a[0]) == b"".join(bytes(c) for c in b[0]) and
                bytes(a[1]) == bytes(b[1]))
        if not same:
            differs.append(i)
    assert differs, "the lowered capacity changed no result"
    assert set(differs) <= set(flagged), f"unflagged arenas changed: {sorted(set(differs) - set(flagged))[:10]}"


_NAN_SCRIPT = r"""
import math, os, sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import parity, tick_host
from rocketsim_amd.state import MODE_SOCCAR
L = tick_host.load({so!r})
samples = parity.load_samples(os.path.join({tests!r}, "golden", "soccar_melee_tf.npz"))[:64]
def inject(cars, balls):
    cars[4 * 5 + 1].vel[0] = math.nan      # arena 5, car 1
    balls[9].pos[2] = math.inf             # arena 9's ball
for t in range(60):
    res, status = tick_host.run_samples(L, samples, MODE_SOCCAR, inject=inject if t == 0 else None)
    for s, r in zip(samples, res):
        s.cars_in, s.ball_in, s.pads_in, s.tick_in = list(r[0]), r[1], list(r[2]), r[3]
    if t == 0:
        first = status
print("STATUS", first[5], first[9], sum(1 for i, x in enumerate(first) if x and i not in (5, 9)))
"""


def test_nonfinite_state_is_flagged_and_stays_in_bounds(tmp_path):
    """NaN / inf injected into two arenas: only they are flagged, and an
    AddressSanitizer build of the tick code steps them for 60 ticks without
    an out-of-bounds access (the GPU test injects the same state)."""
    so = str(tmp_path / "tick_host_asan.so")
    tick_host.build(so, [], ["-fsanitize=address", "-fno-omit-frame-pointer"], load_it=False)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    code = _NAN_SCRIPT.format(root=ROOT, tests=HERE, so=so)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    line = [x for x in r.stdout.splitlines() if x.startswith("STATUS")][0].split()
    assert int(line[1]) & STATUS_NONFINITE and int(line[2]) & STATUS_NONFINITE
    assert int(line[3]) == 0


_UB_SCRIPT = r"""
import os, sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import parity, tick_host
from rocketsim_amd import arena_mesh
from rocketsim_amd.state import MODE_HEATSEEKER, MODE_HOOPS, MODE_SOCCAR, MODE_THE_VOID
L = tick_host.load({so!r})
for name, mode, cmfs in (("void_tf.npz", MODE_THE_VOID, None), ("soccar_split_tf.npz", MODE_SOCCAR, arena_mesh.soccar_split_cmfs()),
                         ("hoops_tf.npz", MODE_HOOPS, None), ("heatseeker_tf.npz", MODE_HEATSEEKER, None),
                         ("soccar_walls_tf.npz", MODE_SOCCAR, None)):
    samples = parity.load_samples(os.path.join({tests!r}, "golden", name))[:256]
    tick_host.run_samples(L, samples, mode, cmfs=cmfs)
print("OK")
"""


@pytest.mark.parametrize("variant", ["sse", "sse_split"])
def test_undefined_behaviour_sanitizer_on_every_arena_kind(tmp_path, variant):
    """An UndefinedBehaviorSanitizer + AddressSanitizer build of the tick code
    (aborting on the first report) steps samples of every arena kind: no
    statics (THE_VOID), the split Soccar arena, Hoops, Heatseeker.  It caught a
    read of the arena geometry through a null pointer in an arena without
    statics, which faulted the GPU."""
    so = str(tmp_path / f"tick_host_ub_{variant}.so")
    tick_host.build(so, tick_host.VARIANTS[variant],
                    ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
                    load_it=False)
    libs = [subprocess.run(["g++", f"-print-file-name={n}"], capture_output=True, text=True).stdout.strip()
            for n in ("libasan.so", "libubsan.so")]
    env = dict(os.environ, LD_PRELOAD=" ".join(libs), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    code = _UB_SCRIPT.format(root=ROOT, tests=HERE, so=so)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
