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
