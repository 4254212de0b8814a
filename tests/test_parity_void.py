"""Teacher-forced one-tick parity in THE_VOID (airborne cars and ball flight,
no contacts) between the HIP kernels and the oracle (the reference RocketSim
built from source).  Tolerances are stated in tests/parity.py."""
import os

import pytest

import parity
from rocketsim_amd.state import MODE_THE_VOID

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle():
    from oracle import ref
    if not ref.available():
        return None
    ref.init()
    return ref


@pytest.mark.gpu
def test_void_fixture_parity():
    samples = parity.load_samples(os.path.join(HERE, "golden", "void_tf.npz"))
    results = parity.run_on_gpu(samples, MODE_THE_VOID)
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, "\n".join(msgs)


@pytest.mark.gpu
def test_void_live_oracle_parity():
    ref = _oracle()
    if ref is None:
        pytest.skip("oracle library not built on this machine")
    samples = parity.sample_void(ref, n_traj=64, ticks_per_traj=16, seed=7)
    results = parity.run_on_gpu(samples, MODE_THE_VOID)
    nbad, msgs = parity.check_samples(samples, results)
    assert nbad == 0, f"{nbad}/{len(samples)} samples off\n" + "\n".join(msgs)
