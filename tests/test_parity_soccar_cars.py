"""Teacher-forced one-tick parity in SOCCAR with cars in contact: driving
from kickoff (wheel rays on the floor, ramps and walls, tire friction,
boost pads) and melees around the ball (car-ball hits with the extra
impulse, box-box car contacts, bumps, demolitions, wheels resting on other
cars).  Tolerances, and the bounded share of samples that ulp-level
normalisation differences may push out of them, are stated in
tests/parity.py."""
import os

import pytest

import parity
from rocketsim_amd.state import MODE_SOCCAR

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["soccar_drive_tf.npz", "soccar_melee_tf.npz", "soccar_walls_tf.npz"])
def test_soccar_cars_fixture_parity(name):
    samples = parity.load_samples(os.path.join(HERE, "golden", name))
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)


@pytest.mark.gpu
def test_soccar_drive_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_soccar_drive(ref, n_traj=12, ticks_per_traj=200, seed=31, n_cars=6, stride=4)
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)


@pytest.mark.gpu
def test_soccar_melee_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_soccar_melee(ref, n_traj=40, ticks_per_traj=90, seed=37, n_cars=4, stride=1)
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)


@pytest.mark.gpu
def test_soccar_walls_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_soccar_melee(ref, n_traj=30, ticks_per_traj=90, seed=41, n_cars=6, stride=1,
                                         near_walls=True)
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    parity.assert_parity(samples, results, MODE_SOCCAR, chaotic=True)
