"""CarConfig (src/Sim/Car/CarConfig/CarConfig.h:6-43) beyond the six presets:
the Psyclops preset (three-wheel steering and friction curves, no sticky
force) and arbitrary configs given per car at batch creation
(Arena::AddCar(team, config), src/Sim/Arena/Arena.h:71).

  * The seven presets are the reference's CAR_CONFIG_* field for field.
  * Teacher-forced melees with Psyclops and custom cars: the host build with
    the reference's SSE normalisation matches the oracle exactly on the CPU;
    the GPU matches within the parity tolerances (tests/parity.py).
  * Snapshots write any config in the reference's byte format and read it
    back (threeWheels, which the reference does not serialize, reads as
    false); configs the kernels cannot simulate are refused.
"""
import copy
import ctypes as C
import math

import pytest

import parity
import snapshot_cases
import tick_host
from rocketsim_amd.state import MODE_SOCCAR, PRESET_COUNT, PRESET_CUSTOM, PRESET_PSYCLOPS, CarConfig


@pytest.fixture(scope="module")
def ref():
    from oracle import ref as r
    if not r.available():
        pytest.skip("oracle library not built on this machine")
    r.init()
    return r


def custom_configs():
    """A long low car with big wheels and a short tall one with a narrow
    track and a wide flip deadzone, beside the Psyclops and the Octane."""
    long_low = tick_host.preset_config(0)
    long_low.hitbox_size[:] = [150.0, 80.0, 30.0]
    long_low.hitbox_pos_offset[:] = [10.0, 0.0, 14.0]
    long_low.front_wheels.wheel_radius = 16.0
    long_low.front_wheels.connection_point_offset[:] = [60.0, 30.0, 14.0]
    long_low.back_wheels.wheel_radius = 18.0
    long_low.back_wheels.suspension_rest_length = 30.0
    tall = tick_host.preset_config(5)
    tall.hitbox_size[:] = [105.0, 70.0, 55.0]
    tall.front_wheels.connection_point_offset[:] = [45.0, 20.0, 22.0]
    tall.back_wheels.connection_point_offset[:] = [-30.0, 24.0, 22.0]
    tall.dodge_deadzone = 0.8
    return [long_low, tall, tick_host.preset_config(PRESET_PSYCLOPS), tick_host.preset_config(0)]


@pytest.mark.parametrize("preset", list(range(PRESET_COUNT)))
def test_presets_are_the_references(ref, preset):
    assert bytes(tick_host.preset_config(preset)) == bytes(ref.car_config_preset(preset))


def _samples(ref):
    return parity.sample_soccar_melee(ref, n_traj=10, ticks_per_traj=60, seed=303, stride=2, configs=custom_configs())


def test_custom_configs_host_build_matches_oracle(ref):
    samples = _samples(ref)
    results, status = tick_host.run_samples(tick_host.get("sse"), samples, MODE_SOCCAR)
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, "\n".join(msgs)
    assert not any(status)


@pytest.mark.gpu
def test_custom_configs_gpu_parity(ref):
    samples = _samples(ref)
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)


def test_snapshot_bytes_with_custom_configs_match_reference(ref):
    """Arena::Serialize of an arena with custom and Psyclops cars: our bytes
    are the reference's; DeserializeNew of them gives the configs back, the
    Psyclops without threeWheels as the reference's does."""
    from rocketsim_amd.batch import deserialize_snapshot, serialize_snapshot
    cfgs = custom_configs()
    A = ref.RefArena(MODE_SOCCAR)
    try:
        for i, c in enumerate(cfgs):
            A.add_car_config(i % 2, c)
        A.reset_kickoff(3)
        A.step(30)
        rb = A.serialize()
        snap = snapshot_cases.snapshot_from_ref(A, [0, 0, 0, 0], MODE_SOCCAR)
        order = A.car_order()
        for k, i in enumerate(order):
            snap.car_config[k] = cfgs[i]
            snap.car_preset[k] = PRESET_CUSTOM
        ours = serialize_snapshot(snap)
        assert snapshot_cases.mask_uninitialized(ours[:len(rb)], len(rb)) == \
            snapshot_cases.mask_uninitialized(rb, len(rb))
        d = deserialize_snapshot(rb)
        for k, i in enumerate(order):
            want = copy.copy(cfgs[i])
            want.three_wheels = 0
            assert bytes(d.car_config[k]) == bytes(want)
            assert d.car_preset[k] == (0 if i == 3 else PRESET_CUSTOM)
    finally:
        A.close()


@pytest.mark.gpu
def test_batch_refuses_unsimulatable_configs():
    from rocketsim_amd.batch import ArenaBatch
    bad = tick_host.preset_config(0)
    bad.hitbox_size[2] = -1.0
    with pytest.raises(RuntimeError, match="positive"):
        ArenaBatch([1], [0], None, car_configs=[bad])
    huge = tick_host.preset_config(0)
    huge.hitbox_size[:] = [400.0, 200.0, 100.0]
    with pytest.raises(RuntimeError, match="cell size"):
        ArenaBatch([1], [0], None, car_configs=[huge])
    nan = tick_host.preset_config(0)
    nan.dodge_deadzone = math.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        ArenaBatch([1], [0], None, car_configs=[nan])
