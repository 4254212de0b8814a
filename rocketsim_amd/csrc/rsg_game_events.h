// rsg_game_events.h -- GameEventTracker (src/Sim/GameEventTracker/
// GameEventTracker.cpp:1-166) for one arena of a batch: shots, goals with
// scorer and passer, and saves, from the arena's state after a step.  Shared
// by the GPU kernel and host builds.
#pragma once

#include "../../include/rsg_state.h"
#include "rsg_tick.h"

namespace rsg {

// Tracker state between updates (GameEventTracker.h:88-97).
enum { GE_BALL_SHOT = 1u, GE_SHOT_GOAL_ORANGE = 2u, GE_SCORED_LAST = 4u };

struct GameEventDev {
    float* cooldown;        // [n] _shotCooldown
    uint32_t* flags;        // [n] GE_*
    uint64_t* last_count;   // [n] _lastBallUpdateCount
    uint32_t* seq;          // [n] events emitted per arena
    rsg_event* ev_buf;
    uint32_t* ev_count;
    uint32_t ev_cap;
};

RSG_FN void ge_emit(const GameEventDev& g, int a, uint64_t tick, uint32_t type, uint32_t x, uint32_t y) {
    uint32_t seq = g.seq[a]++;
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t idx = atomicAdd(g.ev_count, 1u);
#else
    uint32_t idx = (*g.ev_count)++;
#endif
    if (idx >= g.ev_cap) return;
    rsg_event e;
    e.arena = (uint32_t)a;
    e.type = type;
    e.tick = tick;
    e.a = x;
    e.b = y;
    e.seq = seq;
    e._pad = 0;
    g.ev_buf[idx] = e;
}

// GetShooterPasser (GameEventTracker.cpp:5-49): over the arena's cars in
// update order, the team's car with a valid hit within max_shooter ticks and
// the latest hit tick (the first such car on ties), then likewise the passer
// among the others, relative to the shooter's hit.  Returns slot indices or -1.
RSG_FN bool ge_shooter_passer(const BatchDev& d, int a, uint32_t team, uint64_t tick, bool find_passer,
                              uint64_t max_shooter, uint64_t max_passer, int& shooter, int& passer) {
    shooter = passer = -1;
    const int n = (int)arena_ref(d, a).u(AW_NUM_CARS);
    const uint32_t ow = arena_ref(d, a).u(AW_CAR_ORDER);
    uint64_t shooter_tick = 0;
    for (int k = 0; k < n; k++) {
        int i = (int)((ow >> (4 * k)) & 0xFu);
        if (i >= n) continue;
        CarRef cr = car_ref(d, a, i);
        if (((cr.u(CW_CONFIG) >> 8) & 0xFFu) != team) continue;
        if (!(cr.u(CW_FLAGS) & RSG_CF_HIT_VALID)) continue;
        uint64_t ht = cr.u64(CW_HIT_TICK);
        if (ht + max_shooter >= tick && (shooter < 0 || ht > shooter_tick)) {
            shooter = i;
            shooter_tick = ht;
        }
    }
    if (shooter >= 0 && find_passer) {
        uint64_t passer_tick = 0;
        for (int k = 0; k < n; k++) {
            int i = (int)((ow >> (4 * k)) & 0xFu);
            if (i >= n || i == shooter) continue;
            CarRef cr = car_ref(d, a, i);
            if (((cr.u(CW_CONFIG) >> 8) & 0xFFu) != team) continue;
            if (!(cr.u(CW_FLAGS) & RSG_CF_HIT_VALID)) continue;
            uint64_t ht = cr.u64(CW_HIT_TICK);
            if (ht + max_passer >= shooter_tick && (passer < 0 || ht > passer_tick)) {
                passer = i;
                passer_tick = ht;
            }
        }
    }
    return shooter >= 0;
}

// RS_TEAM_FROM_Y (src/Sim/Car/Car.h): the team whose half y is in.
RSG_FN uint32_t ge_team_from_y(float y) { return y < 0 ? (uint32_t)RSG_TEAM_BLUE : (uint32_t)RSG_TEAM_ORANGE; }

// Arena::IsBallProbablyGoingIn for Hoops (src/Sim/Arena/Arena.cpp:808-885):
// a falling ball already low inside a hoop, or a ballistic path that clears
// the rim on the way up and comes down inside a hoop's footprint (with a
// rough bounce off the end wall).
RSG_FN bool ge_probably_going_in_hoops(V3 p, V3 v, const rsg_mutator_config& M, uint32_t& goal_team) {
    const float APPROX_RIM_HEIGHT = 365;
    float min_height = APPROX_RIM_HEIGHT + M.ball_radius * 1.2f;
    if (v.z < -FLT_EPS && p.z < min_height) {
        if (hoops_goal_xy_margin_sq(p.x, p.y) < 0) {
            goal_team = ge_team_from_y(p.y);
            return true;
        }
    }
    float margin = M.ball_radius * 1.0f;
    float margin_sq = margin * margin;
    float up_t, down_t;
    {
        float g = M.gravity[2];
        if (g > -FLT_EPS) return false;
        float vz = v.z;
        float h = p.z - min_height;
        float sqrt_in = vz * vz - 2 * g * h;
        if (sqrt_in > 0) {
            float sq = sqrtf(sqrt_in);
            up_t = (-vz + sq) / g;
            down_t = (-vz - sq) / g;
        } else {
            // never reaches the rim height: in only if it starts inside a hoop
            // (the reference leaves the team unset here)
            return hoops_goal_xy_margin_sq(p.x, p.y) < -margin_sq;
        }
    }
    if (up_t >= 0) {
        V3 up = p + (v * up_t);
        float up_margin_sq = hoops_goal_xy_margin_sq(up.x, up.y);
        float clearance = 60 + M.ball_radius;
        if (up_margin_sq > -margin_sq && up_margin_sq < (clearance * clearance)) return false;  // the rim
    }
    V3 down = p + (v * down_t);
    down.y = fabsf(down.y);
    {
        float wall_y = ARENA_EXTENT_Y_HOOPS - M.ball_radius;
        if (down.y > wall_y) {
            float over = down.y - wall_y;
            down.y -= over * (1 + M.ball_world_restitution);
        }
    }
    if (hoops_goal_xy_margin_sq(down.x, down.y) < -margin_sq) {
        goal_team = ge_team_from_y(down.y);
        return true;
    }
    return false;
}

// Arena::IsBallProbablyGoingIn for Soccar (src/Sim/Arena/Arena.cpp:771-807):
// the ball's ballistic path reaches the goal plane within max_time inside the
// approximate goal mouth; Hoops as above.  goal_team: the team whose net it is.
RSG_FN bool ge_probably_going_in(int mode, V3 pos_bt, V3 vel_bt, float max_time, float extra_margin,
                                 const rsg_mutator_config& M, uint32_t& goal_team) {
    V3 p = pos_bt * BT_TO_UU, v = vel_bt * BT_TO_UU;
    if (mode == RSG_MODE_HOOPS) return ge_probably_going_in_hoops(p, v, M, goal_team);
    if (fabsf(v.y) < FLT_EPS) return false;
    float sgn = (float)((v.y > 0) - (v.y < 0));
    float goal_y = M.goal_base_threshold_y * sgn;
    float dist = fabsf(p.y - goal_y);
    float t = dist / fabsf(v.y);
    if (t > max_time) return false;
    V3 g = v3(M.gravity[0], M.gravity[1], M.gravity[2]);
    V3 e = (p + v * t) + ((g * t) * t) / 2.f;
    const float half_width = 892.755f, height = 642.775f;
    float margin = M.ball_radius * 0.1f + extra_margin;
    if (e.z > height + margin) return false;
    if (fabsf(e.x) > half_width + margin) return false;
    goal_team = sgn < 0 ? (uint32_t)RSG_TEAM_BLUE : (uint32_t)RSG_TEAM_ORANGE;
    return true;
}

// GameEventTracker::Update (GameEventTracker.cpp:51-158) for arena a.
RSG_FN void game_events_update(const BatchDev& d, const GameEventDev& g, const rsg_game_event_config& cfg, int a) {
    CarRef ar = arena_ref(d, a), br = ball_ref(d, a);
    const uint64_t tick = ar.u64(AW_TICK);
    const V3 bpos = br.v(BW_POS), bvel = br.v(BW_VEL);
    const rsg_mutator_config M = d.mut ? d.mut[a] : mutator_defaults(d.mode);
    const bool scored = ball_scored(d.mode, bpos, M);
    const float tick_time = d.tick_time;
    const float tickrate = 1.f / tick_time;
    const uint64_t count = br.u64(BW_TICKS_SINCE_SET);
    const uint64_t last = g.last_count[a];
    float cooldown = g.cooldown[a];
    uint32_t fl = g.flags[a];

    if (count > last || !cfg.auto_state_set_detection) {
        const uint64_t dticks = count - last;
        const float dtime = (float)dticks * tick_time;
        if (scored && !(fl & GE_SCORED_LAST)) {
            // goal: the scoring team is the one away from the ball's side
            uint32_t team = (-bpos.y < 0) ? (uint32_t)RSG_TEAM_BLUE : (uint32_t)RSG_TEAM_ORANGE;
            int s, p;
            if (ge_shooter_passer(d, a, team, tick, true, (uint64_t)(cfg.goal_max_touch_time * tickrate),
                                  (uint64_t)(cfg.pass_max_touch_time * tickrate), s, p))
                ge_emit(g, a, tick, RSG_EVENT_SCORER, car_ref(d, a, s).u(CW_ID), p >= 0 ? car_ref(d, a, p).u(CW_ID) : 0u);
        } else if (!(fl & GE_BALL_SHOT)) {
            if (cooldown > 0) {
                float c = cooldown - dtime;
                cooldown = c > 0 ? c : 0;
            } else {
                V3 v = bvel * BT_TO_UU;
                float speed_sq = v.x * v.x + v.y * v.y + v.z * v.z;
                uint32_t goal_team = 0;  // left BLUE where the reference leaves it unset
                if (speed_sq >= cfg.shot_min_speed * cfg.shot_min_speed &&
                    ge_probably_going_in(d.mode, bpos, bvel, cfg.shot_min_score_time, cfg.pred_score_extra_margin, M, goal_team)) {
                    uint32_t shooter_team = goal_team == RSG_TEAM_BLUE ? (uint32_t)RSG_TEAM_ORANGE : (uint32_t)RSG_TEAM_BLUE;
                    const uint64_t min_delay = (uint64_t)(cfg.shot_touch_min_delay * tickrate);
                    int s, p;
                    if (ge_shooter_passer(d, a, shooter_team, tick, true, dticks + min_delay,
                                          (uint64_t)(cfg.pass_max_touch_time * tickrate), s, p)) {
                        uint64_t since_hit = tick - car_ref(d, a, s).u64(CW_HIT_TICK);
                        if (since_hit >= min_delay) {
                            fl |= GE_BALL_SHOT;
                            fl = goal_team == RSG_TEAM_ORANGE ? (fl | GE_SHOT_GOAL_ORANGE) : (fl & ~GE_SHOT_GOAL_ORANGE);
                            cooldown = cfg.shot_event_cooldown;
                            ge_emit(g, a, tick, RSG_EVENT_SHOT, car_ref(d, a, s).u(CW_ID),
                                    p >= 0 ? car_ref(d, a, p).u(CW_ID) : 0u);
                        }
                    }
                }
            }
        } else {
            uint32_t unused_team = 0;
            if (!ge_probably_going_in(d.mode, bpos, bvel, cfg.shot_min_score_time, cfg.pred_score_extra_margin, M, unused_team)) {
                // no longer going in: a save if the defending team touched it
                uint32_t defending = (fl & GE_SHOT_GOAL_ORANGE) ? (uint32_t)RSG_TEAM_ORANGE : (uint32_t)RSG_TEAM_BLUE;
                int s, p;
                if (ge_shooter_passer(d, a, defending, tick, false, dticks, 0, s, p))
                    ge_emit(g, a, tick, RSG_EVENT_SAVE, car_ref(d, a, s).u(CW_ID), 0u);
                fl &= ~GE_BALL_SHOT;
            }
        }
    } else if (count == last) {
        return;  // no tick since the last update
    } else {
        // the ball was set since: ResetPersistentInfo (GameEventTracker.cpp:160-166)
        fl &= GE_SHOT_GOAL_ORANGE;
        cooldown = 0;
    }
    fl = scored ? (fl | GE_SCORED_LAST) : (fl & ~GE_SCORED_LAST);
    g.flags[a] = fl;
    g.cooldown[a] = cooldown;
    g.last_count[a] = count;
}

}  // namespace rsg
