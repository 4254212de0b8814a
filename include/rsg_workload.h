/*
 * rsg_workload.h -- the seeded synthetic control process used by every
 * throughput and parity workload, in one header that compiles for the host
 * (C or C++), for HIP device code and for the reference harness alike, so the
 * GPU step and the CPU baseline drive cars with bit-identical inputs.
 *
 * Per car, controls are redrawn every U{2..60} ticks; throttle, steer, pitch,
 * yaw and roll are drawn from U[-1,1) and jump, boost and handbrake from
 * Bernoulli(0.5).  Every draw is a pure function of
 * (seed, arena, car slot, redraw number, channel), so no generator state
 * beyond (redraw number, ticks until the next redraw) is carried per car.
 *
 * The mixer is SplitMix64's finaliser (Steele, Lea & Flood, "Fast splittable
 * pseudorandom number generators", OOPSLA 2014).
 */
#ifndef RSG_WORKLOAD_H
#define RSG_WORKLOAD_H

#include <stdint.h>
#include "rsg_state.h"

#if defined(__HIPCC__)
#define RSG_HD __host__ __device__ inline
#else
#define RSG_HD static inline
#endif

RSG_HD uint64_t rsg_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Counter-based draw: channel k of redraw r of car c of arena a. */
RSG_HD uint64_t rsg_draw(uint64_t seed, uint32_t arena, uint32_t car, uint32_t redraw, uint32_t k) {
    uint64_t key = rsg_mix64(seed ^ 0x5EEDC0DEull);
    key = rsg_mix64(key ^ ((uint64_t)arena << 20) ^ (uint64_t)car);
    key = rsg_mix64(key ^ ((uint64_t)redraw << 8) ^ (uint64_t)k);
    return key;
}

/* Uniform in [-1, 1): 24 random bits, exact in float. */
RSG_HD float rsg_unit_signed(uint64_t h) {
    return (float)(uint32_t)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
}

/* Per-car generator state: number of redraws done and ticks left on the
 * current draw.  Both start at 0, which draws on the first tick. */
typedef struct rsg_ctrl_gen {
    uint32_t redraw;
    uint32_t ticks_left;
} rsg_ctrl_gen;

/* Advance one tick; returns the controls in effect for this tick. */
RSG_HD void rsg_ctrl_next(uint64_t seed, uint32_t arena, uint32_t car,
                          rsg_ctrl_gen* g, rsg_controls* c) {
    if (g->ticks_left == 0) {
        uint32_t r = g->redraw;
        c->throttle = rsg_unit_signed(rsg_draw(seed, arena, car, r, 0));
        c->steer    = rsg_unit_signed(rsg_draw(seed, arena, car, r, 1));
        c->pitch    = rsg_unit_signed(rsg_draw(seed, arena, car, r, 2));
        c->yaw      = rsg_unit_signed(rsg_draw(seed, arena, car, r, 3));
        c->roll     = rsg_unit_signed(rsg_draw(seed, arena, car, r, 4));
        uint64_t b = rsg_draw(seed, arena, car, r, 5);
        c->jump      = (uint8_t)((b >> 61) & 1u);
        c->boost     = (uint8_t)((b >> 62) & 1u);
        c->handbrake = (uint8_t)((b >> 63) & 1u);
        c->_pad = 0;
        g->ticks_left = 2u + (uint32_t)((b & 0xFFFFFFFFull) % 59u); /* U{2..60} */
        g->redraw = r + 1u;
    }
    g->ticks_left -= 1u;
}

#undef RSG_HD
#endif /* RSG_WORKLOAD_H */
