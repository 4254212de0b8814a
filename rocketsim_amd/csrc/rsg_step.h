// rsg_step.h -- one arena stepped `ticks` ticks by one lane: gather the
// arena from the struct-of-arrays state, run Arena::Step's body
// (src/Sim/Arena/Arena.cpp:689-756) `ticks` times on the working copy, scatter
// it back.
#pragma once

#include "rsg_tick.h"

namespace rsg {

// Waves per SIMD the step kernels are compiled for (register budget).  The
// plain and LDS kernels run fastest with the whole register file (1); the
// cooperative kernel, whose lanes wait on each other at every barrier, gains
// from a second wave per SIMD to switch to (measured at 2v2 x 16384 arenas:
// 0.98 ms per tick at 2 against 1.16 ms at 1, 1.03 at 3 and 1.08 at 4).
#ifndef RSG_STEP_WAVES_PER_EU
#define RSG_STEP_WAVES_PER_EU 1
#endif
#ifndef RSG_COOP_WAVES_PER_EU
#define RSG_COOP_WAVES_PER_EU 2
#endif

// Arena `a`'s block of the batch's global work space.
RSG_FN StepSpace* arena_space(const BatchDev& d, int a) {
    return &reinterpret_cast<ArenaSpace*>(d.space)[a].step;
}
RSG_FN PairLog* arena_log(const BatchDev& d, int a) {
    return &reinterpret_cast<ArenaSpace*>(d.space)[a].log;
}

RSG_FN void step_arena(const BatchDev& d, const ArenaGeom& geom, int a, int ticks, ArenaLocal& A, StepSpace* ws) {
    load_arena(d, a, A);
    TickParams P;
    P.mode = d.mode;
    P.dt = d.tick_time;
    P.seed = d.seed;
    P.arena_id = (uint32_t)a;
    P.synth_controls = d.synth_controls != 0;
    P.auto_reset = d.auto_reset != 0;
    const ArenaGeom* mp = mode_has_arena(d.mode) ? &geom : nullptr;
    for (int t = 0; t < ticks; t++) arena_tick(A, P, mp, ws);
    store_arena(d, a, A);
}

}  // namespace rsg
