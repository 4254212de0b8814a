"""Game-event tracking on the GPU: the reference's GameEventTracker
(src/Sim/GameEventTracker/GameEventTracker.h:44-106) for every arena of a
Soccar ArenaBatch, over rsg_game_events_* of include/rsg.h.

    tracker = GameEventTracker(batch)            # GameEventTrackerConfig defaults
    batch.step(8)
    tracker.update(batch)                         # GameEventTracker::Update per arena
    for e in tracker.poll():                      # shots, goals (scorer, passer), saves
        ...

Where the reference calls a shot / goal / save callback, an Event is
recorded: type EVENT_SHOT / EVENT_SCORER / EVENT_SAVE, the arena, its tick
count, `a` = the shooter / scorer / saver car id and `b` = the passer's id
(0 if none).
"""
import ctypes as C

from . import native
from .state import Event, GameEventConfig


class GameEventTracker:
    def __init__(self, batch, config=None, capacity=1 << 16):
        L = native.lib()
        self.config = config if config is not None else GameEventConfig()
        h = L.rsg_game_events_create(batch.h, C.byref(self.config), capacity)
        if not h:
            raise RuntimeError(L.rsg_last_error().decode())
        self.h, self.L, self.capacity = h, L, capacity

    def update(self, batch, stream=None):
        """GameEventTracker::Update (GameEventTracker.cpp:51-158) for every arena."""
        native.check(self.L.rsg_game_events_update(self.h, batch.h, stream))

    def poll(self):
        """Events since the last poll, by arena and then in order."""
        out = (Event * self.capacity)()
        lost = C.c_int(0)
        n = native.check(self.L.rsg_game_events_poll(self.h, out, self.capacity, C.byref(lost)))
        self.lost = lost.value
        return list(out[:n])

    def reset_persistent_info(self, arena=-1):
        """ResetPersistentInfo (GameEventTracker.cpp:160-166); -1 = every arena."""
        native.check(self.L.rsg_game_events_reset(self.h, arena))

    def close(self):
        if getattr(self, "h", None):
            self.L.rsg_game_events_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
