"""rsg_batch_status_summary_noalloc (the allocation-free variant, partial
results in a file-scope device array) against rsg_batch_status_summary
(partial results in a buffer the library allocates), through the C ABI: two
calls in a row, the second on a smaller batch, must give the same words; a
batch above the array's capacity is refused."""
import ctypes as C
import math

import pytest

from rocketsim_amd.state import MODE_SOCCAR, STATUS_NONFINITE


def _flagged_batch(n, bad, seed):
    from rocketsim_amd.batch import ArenaBatch
    b = ArenaBatch.uniform(n, 1, 1, 0, mode=MODE_SOCCAR, seed=seed)
    b.step(8)
    cars = list(b.get_cars())
    for a in bad:
        cars[2 * a].vel[2] = math.nan
    b.set_cars(cars)
    b.step(2)
    return b


@pytest.mark.gpu
def test_noalloc_summary_matches_allocating_summary():
    import torch
    from rocketsim_amd import native
    L = native.lib()
    stream = torch.cuda.current_stream().cuda_stream
    for n, bad in ((5000, (0, 1234, 4999, 2500)), (300, (17,))):
        b = _flagged_batch(n, bad, seed=n)
        try:
            ref = torch.full((2,), -1, dtype=torch.int32, device="cuda")
            got = torch.full((2,), -1, dtype=torch.int32, device="cuda")
            native.check(L.rsg_batch_status_summary(b.h, C.c_void_p(ref.data_ptr()), C.c_void_p(stream)))
            native.check(L.rsg_batch_status_summary_noalloc(b.h, C.c_void_p(got.data_ptr()), C.c_void_p(stream)))
            torch.cuda.synchronize()
            assert got.tolist() == ref.tolist() == [STATUS_NONFINITE, len(bad)]
            assert sum(1 for s in b.get_status() if s) == len(bad)
        finally:
            b.close()


@pytest.mark.gpu
def test_noalloc_summary_refuses_oversized_batches():
    """A batch of RSG_STATUS_NOALLOC_MAX_ARENAS + 1 ball-only arenas is
    refused with an error naming the limit; the allocating entry point takes it."""
    import torch
    from rocketsim_amd import native
    from rocketsim_amd.batch import ArenaBatch
    L = native.lib()
    n = (1 << 17) + 1
    stream = torch.cuda.current_stream().cuda_stream
    b = ArenaBatch([0] * n, [], [], mode=MODE_SOCCAR)
    try:
        out = torch.full((2,), -1, dtype=torch.int32, device="cuda")
        assert L.rsg_batch_status_summary_noalloc(b.h, C.c_void_p(out.data_ptr()), C.c_void_p(stream)) == -1
        assert "RSG_STATUS_NOALLOC_MAX_ARENAS" in L.rsg_last_error().decode()
        native.check(L.rsg_batch_status_summary(b.h, C.c_void_p(out.data_ptr()), C.c_void_p(stream)))
        torch.cuda.synchronize()
        assert out.tolist() == [0, 0]
    finally:
        b.close()
