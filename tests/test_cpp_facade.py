"""The C++ facade (include/rsg.hpp) compiles and links against librsg.so with
the host compiler alone, and on the GPU a C++ program written against it
steps an arena to the same state as the Python host layer."""
import os
import subprocess

import pytest

from rocketsim_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "arena_step_demo.cpp")


def _build(tmpdir):
    exe = os.path.join(tmpdir, "arena_step_demo")
    libdir = os.path.dirname(native.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L" + libdir, "-lrsg", "-Wl,-rpath," + libdir, "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_facade_compiles_and_links(tmp_path):
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("librsg.so not built")
    _build(str(tmp_path))


@pytest.mark.gpu
def test_facade_matches_python_layer(tmp_path):
    from rocketsim_amd import arena_mesh
    from rocketsim_amd.batch import ArenaBatch
    from rocketsim_amd.state import Controls, MODE_SOCCAR
    exe = _build(str(tmp_path))
    cmf = os.path.join(str(tmp_path), "soccar.cmf")
    with open(cmf, "wb") as f:
        f.write(arena_mesh.soccar_cmf())
    out = subprocess.run([exe, cmf, "90"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    vals = {line.split()[0]: [float(x) for x in line.split()[1:]] for line in out.stdout.strip().splitlines()}

    b = ArenaBatch([2], [0, 1], [0, 1], mode=MODE_SOCCAR)
    try:
        b.reset_kickoff(0, 7)
        ball = b.get_balls()[0]
        ball.vel[0], ball.vel[1], ball.vel[2] = 300.0, -200.0, 650.0
        b.set_balls([ball])
        c = Controls()
        c.throttle = 1.0
        c.boost = 1
        b.set_controls([c, Controls()])
        b.step(90)
        ball = b.get_balls()[0]
        car = b.get_cars()[0]
        assert vals["ball"] == pytest.approx(list(ball.pos), abs=1e-3)
        assert vals["car0"] == pytest.approx(list(car.pos), abs=1e-3)
        assert int(vals["tick"][0]) == b.get_tick_counts()[0]
        assert vals["resume"][0] == 1.0 and vals["resume"][1] > 0
        # the object API (AddCar, car->controls, ball->SetState) steps the same arena
        assert vals["objball"] == vals["ball"] and vals["objcar0"] == vals["car0"]
        assert vals["objapi"] == [1.0, 1.0, 1.0]
    finally:
        b.close()


@pytest.mark.gpu
def test_reference_typed_program_through_facade(tmp_path):
    """oracle/adapter_demo.cpp: a program holding RocketSim::Arena / Car /
    CarState moves its state each tick into an rsg_amd::Arena through
    include/rsg_rocketsim.hpp, steps both, and compares the GPU's records,
    converted back to RocketSim types, with the reference's (built where the
    reference sources are, by oracle/Makefile)."""
    from rocketsim_amd import arena_mesh
    exe = os.path.join(ROOT, "oracle", "_ref", "adapter_demo")
    if not os.path.exists(exe):
        pytest.skip("adapter demo not built on this machine (needs the reference sources)")
    cmf = tmp_path / "soccar.cmf"
    cmf.write_bytes(arena_mesh.soccar_cmf())
    r = subprocess.run([exe, str(cmf), "120"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "120 ticks match" in r.stdout
