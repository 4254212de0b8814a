"""rocketsim_amd -- MI355X-native batched counterpart of RocketSim's Arena::Step.

The tick step runs as hand-written HIP kernels for gfx950 (rocketsim_amd/csrc,
built into rocketsim_amd/lib/librsg.so) behind the C ABI of include/rsg.h.
"""
from .state import (BallState, CarState, Controls, PadState, MODE_SOCCAR, MODE_THE_VOID,  # noqa: F401
                    TEAM_BLUE, TEAM_ORANGE, NUM_PADS, MAX_CARS)

__version__ = "0.1.0"
