"""The launch geometry (arenas per workgroup) does not change results: the
same batch stepped with different geometries ends bit-identical."""
import pytest

from rocketsim_amd.state import MODE_SOCCAR


@pytest.mark.gpu
def test_launch_geometry_bit_identical():
    from rocketsim_amd.batch import ArenaBatch
    outs = []
    for per_block in (1, 7, 64, 0):
        b = ArenaBatch.uniform(300, 2, 2, 0, mode=MODE_SOCCAR, seed=99, synth_controls=True, auto_reset=True)
        try:
            b.set_launch(per_block)
            b.step(90)
            outs.append((b"".join(bytes(c) for c in b.get_cars()), b"".join(bytes(x) for x in b.get_balls()),
                         b"".join(bytes(p) for p in b.get_pads()), b.get_tick_counts()))
        finally:
            b.close()
    for o in outs[1:]:
        assert o == outs[0]
