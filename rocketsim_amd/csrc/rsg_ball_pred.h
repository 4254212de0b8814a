// rsg_ball_pred.h -- one update of one ball-prediction tracker
// (BallPredTracker, src/Sim/BallPredTracker/BallPredTracker.cpp), shared by
// the GPU kernel and host builds.  Records are RSG_PRED_WORDS word planes
// (include/rsg_state.h) in public units.
#pragma once

#include "../../include/rsg_state.h"
#include "rsg_tick.h"

namespace rsg {

enum { PRED_FORCE = 0, PRED_MANUAL = 1, PRED_FROM_BATCH = 2 };

struct PredDev {
    uint32_t* traj;        // [T][RSG_PRED_WORDS][n], row of tick t = (head + t) % T
    int32_t* head;         // [n]
    int32_t* size;         // [n] predData.size()
    int32_t* last;         // [n] lastUpdateTickCount
    const uint32_t* cur;   // [RSG_PRED_WORDS][n] current states (FORCE, MANUAL)
    const int32_t* ticks;  // [n] ticks since the last update (MANUAL)
    int T;
};

// Ball::GetState (src/Sim/Ball/Ball.cpp:27-33) into a record, word stride s.
RSG_FN void pred_put(uint32_t* r, size_t s, const Ball& ball) {
    V3 p = ball.b.pos * BT_TO_UU, v = ball.b.vel * BT_TO_UU, w = ball.b.ang;
    r[0] = fw(p.x); r[s] = fw(p.y); r[2 * s] = fw(p.z);
    for (int row = 0; row < 3; row++) {
        r[(3 + row) * s] = fw(ball.b.rot.r[row].x);
        r[(6 + row) * s] = fw(ball.b.rot.r[row].y);
        r[(9 + row) * s] = fw(ball.b.rot.r[row].z);
    }
    r[12 * s] = fw(v.x); r[13 * s] = fw(v.y); r[14 * s] = fw(v.z);
    r[15 * s] = fw(w.x); r[16 * s] = fw(w.y); r[17 * s] = fw(w.z);
    r[18 * s] = (uint32_t)ball.ticks_since_set;
    r[19 * s] = (uint32_t)(ball.ticks_since_set >> 32);
    r[20 * s] = fw(ball.hs_dir); r[21 * s] = fw(ball.hs_speed); r[22 * s] = fw(ball.hs_time);
}

RSG_FN V3 pred_v3(const uint32_t* r, size_t s, int w) {
    return v3(wf(r[w * s]), wf(r[(w + 1) * s]), wf(r[(w + 2) * s]));
}

// Ball::SetState (Ball.cpp:35-51) from a record.
RSG_FN void pred_set(Ball& ball, const uint32_t* r, size_t s) {
    ball.b.pos = pred_v3(r, s, RSG_PRED_POS) * UU_TO_BT;
    for (int row = 0; row < 3; row++)
        ball.b.rot.r[row] = v3(wf(r[(3 + row) * s]), wf(r[(6 + row) * s]), wf(r[(9 + row) * s]));
    ball.b.vel = pred_v3(r, s, RSG_PRED_VEL) * UU_TO_BT;
    ball.b.ang = pred_v3(r, s, RSG_PRED_ANG_VEL);
    ball.ticks_since_set = 0;
    ball.hs_dir = wf(r[20 * s]); ball.hs_speed = wf(r[21 * s]); ball.hs_time = wf(r[22 * s]);
    ball_init_body(ball);
}

RSG_FN float pred_dist_sq(V3 a, V3 b) {
    V3 d = a - b;
    return d.x * d.x + d.y * d.y + d.z * d.z;
}

// BallState::Matches (Ball.cpp:12-17) with its default margins (Ball.h:53).
RSG_FN bool pred_matches(const uint32_t* a, size_t sa, const uint32_t* b, size_t sb) {
    const float mp = 0.8f, mv = 0.4f, mw = 0.02f;
    return pred_dist_sq(pred_v3(a, sa, RSG_PRED_POS), pred_v3(b, sb, RSG_PRED_POS)) < mp * mp &&
           pred_dist_sq(pred_v3(a, sa, RSG_PRED_VEL), pred_v3(b, sb, RSG_PRED_VEL)) < mv * mv &&
           pred_dist_sq(pred_v3(a, sa, RSG_PRED_ANG_VEL), pred_v3(b, sb, RSG_PRED_ANG_VEL)) < mw * mw;
}

// Tracker i's update.  kind: PRED_FORCE = ForceUpdateAllPred (:63-71),
// PRED_MANUAL = UpdatePredManual (:25-61), PRED_FROM_BATCH =
// UpdatePredFromArena (:18-23) on arena i of `src`.  `pa` holds the trackers'
// ball-only prediction arenas; `A` is the working copy of arena i.
RSG_FN void ball_pred_update(const BatchDev& pa, const ArenaGeom* mp, const PredDev& p, int kind, const BatchDev& src,
                             int i, ArenaLocal& A, StepSpace* ws) {
    load_arena(pa, i, A);
    TickParams P;
    P.mode = pa.mode;
    P.dt = pa.tick_time;
    P.seed = pa.seed;
    P.arena_id = (uint32_t)i;
    P.synth_controls = false;
    P.auto_reset = false;
    const int T = p.T;
    const size_t s = (size_t)pa.n;

    uint32_t cur[RSG_PRED_WORDS];
    int k = 0;
    if (kind == PRED_FROM_BATCH) {
        CarRef br = ball_ref(src, i);
        Ball cb;
        cb.b.pos = br.v(BW_POS);
        cb.b.rot = load_rot(br, BW_ROT);
        cb.b.vel = br.v(BW_VEL);
        cb.b.ang = br.v(BW_ANGVEL);
        cb.ticks_since_set = br.u64(BW_TICKS_SINCE_SET);
        cb.hs_dir = br.f(BW_HS); cb.hs_speed = br.f(BW_HS + 1); cb.hs_time = br.f(BW_HS + 2);
        pred_put(cur, 1, cb);
        // int ticksSinceLastUpdate = arena->tickCount - lastUpdateTickCount
        uint64_t tick = arena_ref(src, i).u64(AW_TICK);
        k = (int)(uint32_t)(tick - (uint64_t)(int64_t)p.last[i]);
    } else {
        for (int w = 0; w < RSG_PRED_WORDS; w++) cur[w] = p.cur[w * s + i];
        if (kind == PRED_MANUAL) k = p.ticks[i];
    }

    uint32_t* traj = p.traj + i;
    const size_t row_words = (size_t)RSG_PRED_WORDS * s;
    int head = p.head[i], size = p.size[i];
    bool full = true;
    if (kind != PRED_FORCE && k >= 0 && k < size) {
        if (pred_matches(traj + (size_t)((head + k) % T) * row_words, s, cur, 1)) {
            full = false;
            if (k > 0) {
                // drop the k oldest ticks, predict k new ones from the last
                head = (head + k) % T;
                size -= k;
                pred_set(A.ball, traj + (size_t)((head + size - 1) % T) * row_words, s);
                for (; size < T; size++) {
                    arena_tick(A, P, mp, ws);
                    pred_put(traj + (size_t)((head + size) % T) * row_words, s, A.ball);
                }
            }
        }
    }
    if (full) {
        pred_set(A.ball, cur, 1);
        uint32_t* r0 = traj + (size_t)head * row_words;
        for (int w = 0; w < RSG_PRED_WORDS; w++) r0[w * s] = cur[w];
        for (int t = 1; t < T; t++) {
            arena_tick(A, P, mp, ws);
            pred_put(traj + (size_t)((head + t) % T) * row_words, s, A.ball);
        }
        size = T;
    }
    if (kind != PRED_FORCE) p.last[i] = (int32_t)((uint32_t)p.last[i] + (uint32_t)k);
    p.head[i] = head;
    p.size[i] = size;
    store_arena(pa, i, A);
}

}  // namespace rsg
