"""The C-ABI library loads and exports every entry point include/rsg.h declares,
and the Python ctypes records match the C layout (no compute calls: this runs
without a GPU)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from rocketsim_amd import native
from rocketsim_amd import state
from rocketsim_amd.state import (ArenaSnapshot, BallPredView, BallState, BatchDesc, CarState, Controls, Event,
                                 CarConfig, GameEventConfig, MutatorConfig, PadState, WheelPairConfig)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rsg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rsg_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_listed():
    assert sorted(native.EXPORTED) == declared_symbols()


def test_library_exports_all_symbols():
    lib = C.CDLL(native.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


RECORDS = {"rsg_controls": Controls, "rsg_car_state": CarState, "rsg_ball_state": BallState,
           "rsg_pad_state": PadState, "rsg_batch_desc": BatchDesc, "rsg_event": Event,
           "rsg_ball_pred_view_t": BallPredView, "rsg_arena_snapshot": ArenaSnapshot,
           "rsg_game_event_config": GameEventConfig, "rsg_mutator_config": MutatorConfig,
           "rsg_car_config": CarConfig, "rsg_wheel_pair_config": WheelPairConfig}
# enum constants of include/rsg_state.h mirrored in rocketsim_amd.state
CONSTANTS = {"RSG_PRED_POS": "PRED_POS", "RSG_PRED_ROT": "PRED_ROT", "RSG_PRED_VEL": "PRED_VEL",
             "RSG_PRED_ANG_VEL": "PRED_ANG_VEL", "RSG_PRED_TICKS_SINCE_UPDATE": "PRED_TICKS_SINCE_UPDATE",
             "RSG_PRED_WORDS": "PRED_WORDS", "RSG_EVENT_SHOT": "EVENT_SHOT", "RSG_EVENT_SCORER": "EVENT_SCORER",
             "RSG_EVENT_SAVE": "EVENT_SAVE"}


def test_ctypes_layout_matches_c_header():
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rsg.h"', "int main(void){"]
    for cname, py in RECORDS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    for cname, pyname in CONSTANTS.items():
        lines.append(f'printf("const {cname} %d\\n", (int){cname});')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        r = subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), src, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.fail(r.stderr)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    for line in out.strip().splitlines():
        cname, field, value = line.split()
        if cname == "const":
            assert getattr(state, CONSTANTS[field]) == int(value), (field, value)
            continue
        py = RECORDS[cname]
        got = C.sizeof(py) if field == "sizeof" else getattr(py, field).offset
        assert got == int(value), (cname, field, got, value)
