"""Per-arena status bits on the GPU (the reference throws where the batch
sets a bit, RS_ERR_CLOSE in src/Framework.h:67-72): a non-finite state
injected into three arenas of a batch flags exactly those arenas, the
device-side summary agrees with the per-arena words, strict mode turns the
flag into a failed step naming the arena, and clearing resets the words.
The same injection stays within every array of the tick code in an
AddressSanitizer build on the CPU (test_tick_host.py)."""
import math

import pytest

from rocketsim_amd.state import MODE_SOCCAR, STATUS_NONFINITE

BAD = (5, 77, 200)


@pytest.mark.gpu
def test_nonfinite_arenas_are_flagged_and_summarised():
    import torch
    from rocketsim_amd.batch import ArenaBatch
    b = ArenaBatch.uniform(256, 2, 2, 0, mode=MODE_SOCCAR, seed=3)
    try:
        b.step(30)
        cars = list(b.get_cars())
        for a in BAD:
            cars[4 * a + 1].vel[0] = math.nan
        b.set_cars(cars)
        b.step(4)
        st = b.get_status()
        assert [a for a, s in enumerate(st) if s] == list(BAD)
        assert all(st[a] & STATUS_NONFINITE for a in BAD)
        out = torch.zeros(2, dtype=torch.int32, device="cuda")
        stream = torch.cuda.current_stream()
        b.status_summary(out, stream.cuda_stream)
        torch.cuda.synchronize()
        assert out.tolist() == [STATUS_NONFINITE, len(BAD)]
        b.set_strict(True)
        with pytest.raises(RuntimeError, match="arena 5: non-finite"):
            b.step(1)
        b.set_strict(False)
        b.clear_status()
        assert not any(b.get_status())
        b.status_summary(out, stream.cuda_stream)
        torch.cuda.synchronize()
        assert out.tolist() == [0, 0]
    finally:
        b.close()


@pytest.mark.gpu
def test_clean_play_sets_no_status():
    """The 2v2 workload with seeded controls and auto-reset keeps every
    arena clear of status bits (no capacity is exceeded in play)."""
    from rocketsim_amd.batch import ArenaBatch
    b = ArenaBatch.uniform(2048, 2, 2, 0, mode=MODE_SOCCAR, seed=11, synth_controls=True, auto_reset=True)
    try:
        b.set_strict(True)
        for _ in range(60):
            b.step(8)
        assert not any(b.get_status())
    finally:
        b.close()
