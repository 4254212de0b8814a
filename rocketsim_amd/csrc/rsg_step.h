// rsg_step.h -- one arena stepped `ticks` ticks by one lane: gather the
// arena from the struct-of-arrays state, run Arena::Step's body
// (src/Sim/Arena/Arena.cpp:689-756) `ticks` times on the working copy, scatter
// it back.
#pragma once

#include "rsg_tick.h"

namespace rsg {

// Waves per SIMD the step kernels are compiled for (register budget).
#ifndef RSG_STEP_WAVES_PER_EU
#define RSG_STEP_WAVES_PER_EU 1
#endif

// The world-step work space of arena `a` in the batch's global block.
RSG_FN StepSpace* arena_space(const BatchDev& d, int a) {
    return reinterpret_cast<StepSpace*>(d.space + (size_t)a * sizeof(StepSpace));
}

RSG_FN void step_arena(const BatchDev& d, const MeshView& mesh, int a, int ticks, ArenaLocal& A, StepSpace* ws) {
    RSG_PROF_MARK(15);
    load_arena(d, a, A);
    RSG_PROF_MARK(15);
    TickParams P;
    P.mode = d.mode;
    P.dt = d.tick_time;
    P.seed = d.seed;
    P.arena_id = (uint32_t)a;
    P.synth_controls = d.synth_controls != 0;
    P.auto_reset = d.auto_reset != 0;
    const MeshView* mp = mode_has_arena(d.mode) ? &mesh : nullptr;
    for (int t = 0; t < ticks; t++) arena_tick(A, P, mp, ws);
    store_arena(d, a, A);
}

}  // namespace rsg
