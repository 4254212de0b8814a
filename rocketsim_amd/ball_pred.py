"""Batched ball prediction: the GPU counterpart of the reference's
BallPredTracker (src/Sim/BallPredTracker/BallPredTracker.h:7-36) over
rsg_ball_pred_* of include/rsg.h.

One object holds n trackers; tracker i predicts `num_pred_ticks` ticks of the
ball of arena i (its own ball-only arena, as the reference's tracker keeps).

    tracker = BallPredTracker(batch, 720)      # predicts every arena's ball
    batch.step(8)
    tracker.update_pred_from_arena(batch)      # reuses what still matches
    s = tracker.get_ball_state_for_time(3, 1.5)
"""
import ctypes as C

from . import native
from .state import BallPredView, BallState, MODE_HEATSEEKER, MODE_SOCCAR


class BallPredTracker:
    def __init__(self, arena=None, num_pred_ticks=120, *, mode=None, n=None, tick_rate=120.0, stream=None):
        """BallPredTracker(Arena*, numPredTicks) (BallPredTracker.cpp:5-12):
        with an ArenaBatch `arena`, one tracker per arena, first updated from
        it; otherwise `mode`, `n` and `tick_rate` give an empty prediction."""
        if arena is not None:
            mode, n, tick_rate = arena.mode, arena.n_arenas, arena.tick_rate
        if mode is None or n is None:
            raise ValueError("BallPredTracker needs an ArenaBatch or mode and n")
        from . import batch
        batch.ensure_arena(mode)
        L = native.lib()
        h = L.rsg_ball_pred_create(mode, n, num_pred_ticks, tick_rate)
        if not h:
            raise RuntimeError(L.rsg_last_error().decode())
        self.h, self.L = h, L
        self.n, self.num_pred_ticks, self.tick_time = n, num_pred_ticks, 1.0 / tick_rate
        if arena is not None:
            self.update_pred_from_arena(arena, stream)

    def update_pred_from_arena(self, arena, stream=None):
        """UpdatePredFromArena (BallPredTracker.cpp:18-23), tracker i <- arena i."""
        native.check(self.L.rsg_ball_pred_update_from_batch(self.h, arena.h, stream))

    def update_pred_manual(self, cur_ball_states, ticks_since_last_update, stream=None):
        """UpdatePredManual (BallPredTracker.cpp:25-61): one state and one tick
        count per tracker."""
        cur = (BallState * self.n)(*cur_ball_states)
        ticks = (C.c_int32 * self.n)(*ticks_since_last_update)
        native.check(self.L.rsg_ball_pred_update_manual(self.h, cur, ticks, stream))

    def force_update_all_pred(self, initial_ball_states, stream=None):
        """ForceUpdateAllPred (BallPredTracker.cpp:63-71) for every tracker."""
        init = (BallState * self.n)(*initial_ball_states)
        native.check(self.L.rsg_ball_pred_force_update(self.h, init, stream))

    def pred_data(self, tracker):
        """predData of one tracker (a list of BallState, tick 0 first)."""
        out = (BallState * self.num_pred_ticks)()
        size = native.check(self.L.rsg_ball_pred_get(self.h, tracker, out, self.num_pred_ticks))
        return list(out[:size])

    def pred_data_all(self):
        """Every tracker's predData, tracker-major ((n * num_pred_ticks) BallState)."""
        out = (BallState * (self.n * self.num_pred_ticks))()
        native.check(self.L.rsg_ball_pred_get_all(self.h, out))
        return out

    def get_ball_state_for_time(self, tracker, pred_time):
        """GetBallStateForTime (BallPredTracker.cpp:73-79)."""
        out = BallState()
        native.check(self.L.rsg_ball_pred_state_for_time(self.h, tracker, pred_time, C.byref(out)))
        return out

    def set_launch(self, trackers_per_wave):
        """Trackers per 64-lane wave (0 = one wave per SIMD)."""
        native.check(self.L.rsg_ball_pred_set_launch(self.h, trackers_per_wave))

    def view(self):
        """Device pointers of the predictions (rsg_ball_pred_view_t)."""
        v = BallPredView()
        native.check(self.L.rsg_ball_pred_view(self.h, C.byref(v)))
        return v

    def close(self):
        if getattr(self, "h", None):
            self.L.rsg_ball_pred_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
