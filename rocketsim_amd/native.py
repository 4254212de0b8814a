"""Loader for librsg.so (the HIP tick kernels behind include/rsg.h).

There is no CPU fallback: if the library is missing or cannot be loaded, every
entry point raises.
"""
import ctypes as C
import os

from .state import (ArenaSnapshot, BallPredView, BallState, BatchDesc, CarState, Controls, Event, GameEventConfig,
                    CarConfig, MutatorConfig, PadState)

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "librsg.so")

_lib = None


class NativeMissing(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeMissing(f"{LIB_PATH} is not built; run `python -m rocketsim_amd.build`")
        # One HIP runtime per process: PyTorch ships its own libamdhip64, and
        # a process that loads librsg.so first (resolving it to the system
        # runtime) leaves PyTorch unable to find a GPU afterwards.  Loading
        # PyTorch first makes librsg.so bind to the runtime already loaded.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.rsg_last_error.restype = C.c_char_p
        L.rsg_init_mesh.argtypes = [C.c_void_p, C.c_uint64]
        L.rsg_init_meshes.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_int]
        L.rsg_init_arena_meshes.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.c_int]
        L.rsg_batch_create.restype = C.c_void_p
        L.rsg_batch_create.argtypes = [C.POINTER(BatchDesc)]
        L.rsg_batch_destroy.argtypes = [C.c_void_p]
        for name in ("rsg_batch_num_arenas", "rsg_batch_num_cars"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.rsg_batch_car_offset.argtypes = [C.c_void_p, C.c_int]
        L.rsg_batch_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.rsg_batch_sync.argtypes = [C.c_void_p]
        L.rsg_batch_get_cars.argtypes = [C.c_void_p, C.POINTER(CarState)]
        L.rsg_batch_set_cars.argtypes = [C.c_void_p, C.POINTER(CarState)]
        L.rsg_batch_get_balls.argtypes = [C.c_void_p, C.POINTER(BallState)]
        L.rsg_batch_set_balls.argtypes = [C.c_void_p, C.POINTER(BallState)]
        L.rsg_batch_get_pads.argtypes = [C.c_void_p, C.POINTER(PadState)]
        L.rsg_batch_set_pads.argtypes = [C.c_void_p, C.POINTER(PadState)]
        L.rsg_batch_set_controls.argtypes = [C.c_void_p, C.POINTER(Controls)]
        L.rsg_batch_get_controls.argtypes = [C.c_void_p, C.POINTER(Controls)]
        L.rsg_batch_set_launch.argtypes = [C.c_void_p, C.c_int]
        L.rsg_batch_get_launch.argtypes = [C.c_void_p]
        L.rsg_batch_set_kernel.argtypes = [C.c_void_p, C.c_int]
        L.rsg_batch_get_kernel.argtypes = [C.c_void_p]
        P = C.c_void_p
        L.rsg_air_rollout.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, P, P, P, P, P, P, P]
        L.rsg_air_rollout_backward.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, P, P, P, P, P, P, P]
        L.rsg_shot_rollout.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, P, P, P, P, P, P, P, P]
        L.rsg_shot_rollout_backward.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, P, P, P, P, P, P, P]
        L.rsg_batch_enable_events.argtypes = [C.c_void_p, C.c_int]
        L.rsg_batch_poll_events.argtypes = [C.c_void_p, C.POINTER(Event), C.c_int, C.POINTER(C.c_int)]
        L.rsg_batch_graph_create.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.rsg_batch_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
        L.rsg_batch_graph_ticks.argtypes = [C.c_void_p]
        L.rsg_batch_autotune.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.rsg_batch_set_car_order.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.rsg_batch_get_car_order.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.rsg_batch_get_tick_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.rsg_batch_set_tick_counts.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.rsg_batch_get_goals.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint64)]
        L.rsg_batch_reset_kickoff.argtypes = [C.c_void_p, C.c_int, C.c_uint64]
        L.rsg_batch_is_ball_scored.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        L.rsg_mutator_config_default.argtypes = [C.c_int, C.POINTER(MutatorConfig)]
        L.rsg_car_config_preset.argtypes = [C.c_int, C.POINTER(CarConfig)]
        L.rsg_batch_set_mutator_config.argtypes = [C.c_void_p, C.c_int, C.POINTER(MutatorConfig)]
        L.rsg_batch_get_mutator_config.argtypes = [C.c_void_p, C.c_int, C.POINTER(MutatorConfig)]
        L.rsg_batch_get_status.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.rsg_batch_clear_status.argtypes = [C.c_void_p]
        L.rsg_batch_set_strict.argtypes = [C.c_void_p, C.c_int]
        L.rsg_batch_status_summary.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rsg_batch_status_summary_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rsg_ball_pred_create.restype = C.c_void_p
        L.rsg_ball_pred_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float]
        L.rsg_ball_pred_destroy.argtypes = [C.c_void_p]
        L.rsg_ball_pred_force_update.argtypes = [C.c_void_p, C.POINTER(BallState), C.c_void_p]
        L.rsg_ball_pred_update_manual.argtypes = [C.c_void_p, C.POINTER(BallState), C.POINTER(C.c_int32),
                                                  C.c_void_p]
        L.rsg_ball_pred_update_from_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rsg_ball_pred_get.argtypes = [C.c_void_p, C.c_int, C.POINTER(BallState), C.c_int]
        L.rsg_ball_pred_get_all.argtypes = [C.c_void_p, C.POINTER(BallState)]
        L.rsg_ball_pred_state_for_time.argtypes = [C.c_void_p, C.c_int, C.c_float, C.POINTER(BallState)]
        L.rsg_ball_pred_view.argtypes = [C.c_void_p, C.POINTER(BallPredView)]
        L.rsg_ball_pred_set_launch.argtypes = [C.c_void_p, C.c_int]
        L.rsg_game_event_config_default.argtypes = [C.POINTER(GameEventConfig)]
        L.rsg_game_events_create.restype = C.c_void_p
        L.rsg_game_events_create.argtypes = [C.c_void_p, C.POINTER(GameEventConfig), C.c_int]
        L.rsg_game_events_destroy.argtypes = [C.c_void_p]
        L.rsg_game_events_update.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rsg_game_events_poll.argtypes = [C.c_void_p, C.POINTER(Event), C.c_int, C.POINTER(C.c_int)]
        L.rsg_game_events_reset.argtypes = [C.c_void_p, C.c_int]
        L.rsg_snapshot_serialize.restype = C.c_int64
        L.rsg_snapshot_serialize.argtypes = [C.POINTER(ArenaSnapshot), C.c_void_p, C.c_uint64]
        L.rsg_snapshot_deserialize.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(ArenaSnapshot)]
        L.rsg_batch_get_snapshot.argtypes = [C.c_void_p, C.c_int, C.POINTER(ArenaSnapshot)]
        L.rsg_batch_set_snapshot.argtypes = [C.c_void_p, C.c_int, C.POINTER(ArenaSnapshot)]
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise RuntimeError(lib().rsg_last_error().decode())
    return rc


# Every symbol include/rsg.h declares, for the ABI export test.
EXPORTED = [
    "rsg_last_error", "rsg_init_mesh", "rsg_init_meshes", "rsg_init_arena_meshes", "rsg_batch_create", "rsg_batch_destroy",
    "rsg_batch_num_arenas", "rsg_batch_num_cars", "rsg_batch_car_offset",
    "rsg_batch_step", "rsg_batch_sync",
    "rsg_batch_get_cars", "rsg_batch_set_cars", "rsg_batch_get_balls", "rsg_batch_set_balls",
    "rsg_batch_get_pads", "rsg_batch_set_pads", "rsg_batch_set_controls", "rsg_batch_get_controls",
    "rsg_batch_set_car_order", "rsg_batch_get_car_order", "rsg_batch_set_launch", "rsg_batch_get_launch",
    "rsg_batch_set_kernel", "rsg_batch_get_kernel", "rsg_batch_autotune",
    "rsg_batch_graph_create", "rsg_batch_graph_launch", "rsg_batch_graph_ticks",
    "rsg_batch_enable_events", "rsg_batch_poll_events", "rsg_air_rollout", "rsg_air_rollout_backward",
    "rsg_shot_rollout", "rsg_shot_rollout_backward",
    "rsg_batch_get_tick_counts", "rsg_batch_set_tick_counts", "rsg_batch_get_goals",
    "rsg_batch_reset_kickoff", "rsg_batch_is_ball_scored", "rsg_mutator_config_default",
    "rsg_batch_set_mutator_config", "rsg_batch_get_mutator_config", "rsg_car_config_preset",
    "rsg_batch_get_status", "rsg_batch_clear_status", "rsg_batch_set_strict", "rsg_batch_status_summary",
    "rsg_batch_status_summary_noalloc",
    "rsg_ball_pred_create", "rsg_ball_pred_destroy", "rsg_ball_pred_force_update", "rsg_ball_pred_update_manual",
    "rsg_ball_pred_update_from_batch", "rsg_ball_pred_get", "rsg_ball_pred_get_all",
    "rsg_ball_pred_state_for_time", "rsg_ball_pred_view", "rsg_ball_pred_set_launch",
    "rsg_snapshot_serialize", "rsg_snapshot_deserialize", "rsg_batch_get_snapshot", "rsg_batch_set_snapshot",
    "rsg_game_event_config_default", "rsg_game_events_create", "rsg_game_events_destroy", "rsg_game_events_update",
    "rsg_game_events_poll", "rsg_game_events_reset",
]
