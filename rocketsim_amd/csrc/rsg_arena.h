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
    Car cars[MAX_CARS];
};

struct TickParams {
    int mode;
    float dt;
    uint64_t seed;
    uint32_t arena_id;
    bool synth_controls;
    bool auto_reset;
};

}  // namespace rsg
