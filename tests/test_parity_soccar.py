"""Teacher-forced one-tick parity in SOCCAR between the HIP kernels and the
oracle (the reference RocketSim built from source): the ball against the
arena's planes and collision mesh (floor, side walls, corners, goals) with
two cars resting on their wheels.  Tolerances are stated in tests/parity.py."""
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
def test_soccar_ball_fixture_parity():
    samples = parity.load_samples(os.path.join(HERE, "golden", "soccar_ball_tf.npz"))
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, f"{nbad}/{len(samples)} samples off\n" + "\n".join(msgs)


@pytest.mark.gpu
def test_soccar_ball_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_soccar_ball(ref, n_traj=48, ticks_per_traj=120, seed=11, stride=3)
    results = parity.run_on_gpu(samples, MODE_SOCCAR)
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, f"{nbad}/{len(samples)} samples off\n" + "\n".join(msgs)
