"""Neither the kernel variant (arena working copy in private scratch or in
LDS) nor the launch geometry (arenas per workgroup) changes results: the same
batch stepped each way ends bit-identical."""
import pytest

from rocketsim_amd.state import KERNEL_AUTO, KERNEL_COOP, KERNEL_LDS, KERNEL_SCRATCH, MODE_SOCCAR


@pytest.mark.gpu
@pytest.mark.parametrize("cars", [(2, 2), (3, 3), (1, 1)])
def test_kernel_and_launch_bit_identical(cars):
    from rocketsim_amd.batch import ArenaBatch
    outs = []
    for kernel, per_block in ((KERNEL_SCRATCH, 64), (KERNEL_SCRATCH, 1), (KERNEL_SCRATCH, 7),
                              (KERNEL_LDS, 1), (KERNEL_LDS, 3), (KERNEL_COOP, 1), (KERNEL_COOP, 2),
                              (KERNEL_AUTO, 0)):
        b = ArenaBatch.uniform(300, cars[0], cars[1], 0, mode=MODE_SOCCAR, seed=99, synth_controls=True,
                               auto_reset=True)
        try:
            b.set_kernel(kernel)
            b.set_launch(per_block)
            b.step(90)
            outs.append((b"".join(bytes(c) for c in b.get_cars()), b"".join(bytes(x) for x in b.get_balls()),
                         b"".join(bytes(p) for p in b.get_pads()), b.get_tick_counts()))
        finally:
            b.close()
    for o in outs[1:]:
        assert o == outs[0]


@pytest.mark.gpu
def test_autotune_leaves_state_and_results_unchanged():
    """The autotuner times candidates on a copy of the state: the batch is not
    advanced, and stepping afterwards gives the same result as without it."""
    from rocketsim_amd.batch import ArenaBatch
    outs = []
    for tune in (False, True):
        b = ArenaBatch.uniform(500, 2, 2, 0, mode=MODE_SOCCAR, seed=5, synth_controls=True, auto_reset=True)
        try:
            b.step(30)
            before = b"".join(bytes(c) for c in b.get_cars())
            if tune:
                kernel, per_block, ms = b.autotune(4)
                assert kernel in (KERNEL_SCRATCH, KERNEL_LDS, KERNEL_COOP)
                assert 1 <= per_block <= 64 and ms > 0
                assert b.get_kernel() == kernel and b.get_launch() == per_block
                assert b"".join(bytes(c) for c in b.get_cars()) == before
                assert b.get_tick_counts() == [30] * 500
            b.step(60)
            outs.append((b"".join(bytes(c) for c in b.get_cars()), b"".join(bytes(x) for x in b.get_balls())))
        finally:
            b.close()
    assert outs[0] == outs[1]
