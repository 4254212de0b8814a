// rsg_arena.h -- the working copy of one arena during a tick.
#pragma once

#include "rsg_sim.h"

namespace rsg {

constexpr int MAX_CARS = RSG_MAX_CARS;

// The car array comes last: an arena with n cars touches only the first
// offsetof(ArenaLocal, cars) + n * sizeof(Car) bytes, which lets a kernel
// pack arenas of a batch with at most n cars per arena at that stride.
struct ArenaLocal {
    int n_cars;
    int mode;           // RSG_MODE_*
    rsg_mutator_config mut;  // Arena::_mutatorConfig
    const rsg_boost_pad_config* custom_pads;  // the batch's custom pads (null: the mode's)
    int n_pads;         // boost pads in use
    Ball ball;
    Pad pads[PAD_NUM];
    uint64_t tick;
    uint32_t goals_blue, goals_orange;
    uint64_t last_goal_tick;
    uint32_t resets;
    uint32_t status;
    SolverInfo si;
    // broadphase registration of the dynamic proxies (0 = ball, 1 + i = car i)
    int bp_cell[1 + MAX_CARS];
    uint32_t bp_seq[1 + MAX_CARS];
    // Order in which the per-car updates run (Arena::_cars is an
    // unordered_set<Car*>, src/Sim/Arena/Arena.h:35; it is iterated for the
    // pre-tick, post-tick/boost-pad and kickoff loops).
    uint8_t order[MAX_CARS];
    // event output (the batch's ring, or null) and this arena's sequence
    rsg_event* ev_buf;
    uint32_t* ev_count;
    uint32_t ev_cap;
    uint32_t ev_seq;
    uint32_t arena_id;
    Car cars[MAX_CARS];
};

// Appends an event to the batch's ring (dropped when the ring is full; the
// counter keeps counting, so the host sees how many were lost).
RSG_FN void emit_event(ArenaLocal& A, uint32_t type, uint32_t a, uint32_t b) {
    uint32_t seq = A.ev_seq++;
    if (!A.ev_buf) return;
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t idx = atomicAdd(A.ev_count, 1u);
#else
    uint32_t idx = (*A.ev_count)++;
#endif
    if (idx >= A.ev_cap) return;
    rsg_event e;
    e.arena = A.arena_id;
    e.type = type;
    e.tick = A.tick;
    e.a = a;
    e.b = b;
    e.seq = seq;
    e._pad = 0;
    A.ev_buf[idx] = e;
}

struct TickParams {
    int mode;
    float dt;
    uint64_t seed;
    uint32_t arena_id;
    bool synth_controls;
    bool auto_reset;
};

}  // namespace rsg
