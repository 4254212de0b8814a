"""hipGraph-captured multi-tick chunks step exactly like eager launches, on
every replay, with eager work and host reads between the replays."""
import random

import pytest

from rocketsim_amd.state import MODE_SOCCAR


def _mixed_batch(n, seed):
    from rocketsim_amd.batch import ArenaBatch
    rng = random.Random(seed)
    counts, teams, presets = [], [], []
    for _ in range(n):
        per_team = rng.choice([1, 1, 2, 2, 3])
        counts.append(2 * per_team)
        teams += [0] * per_team + [1] * per_team
        presets += [rng.randint(0, 5) for _ in range(2 * per_team)]
    return ArenaBatch(counts, teams, presets, mode=MODE_SOCCAR, seed=seed, synth_controls=True, auto_reset=True)


def _state(b):
    return (b"".join(bytes(c) for c in b.get_cars()), b"".join(bytes(x) for x in b.get_balls()),
            b"".join(bytes(p) for p in b.get_pads()), b.get_tick_counts())


@pytest.mark.gpu
def test_graph_replays_match_eager_steps():
    eager = _mixed_batch(300, 11)
    graph = _mixed_batch(300, 11)
    try:
        graph.graph_create(4, 3)  # 12 ticks per replay
        for _ in range(4):
            eager.step(4)
            eager.step(4)
            eager.step(4)
            graph.graph_launch()
            assert _state(graph) == _state(eager)
    finally:
        eager.close()
        graph.close()
