// rsg_tick.h -- one arena tick (Arena::Step body, src/Sim/Arena/Arena.cpp:689-756)
// over an arena held in registers / private memory, plus the gather/scatter
// between that working copy and the struct-of-arrays state in HBM.
#pragma once

#include "rsg_sim.h"
#include "rsg_collide.h"

namespace rsg {

// ---------------------------------------------------------------------------
// gather / scatter

RSG_FN float wf(uint32_t u) { return __builtin_bit_cast(float, u); }
RSG_FN uint32_t fw(float f) { return __builtin_bit_cast(uint32_t, f); }

struct CarRef {
    uint32_t* base;
    int stride;  // words between consecutive word planes
    RSG_FN uint32_t& u(int w) const { return base[(size_t)w * stride]; }
    RSG_FN float f(int w) const { return wf(u(w)); }
    RSG_FN void sf(int w, float v) const { u(w) = fw(v); }
    RSG_FN V3 v(int w) const { return v3(f(w), f(w + 1), f(w + 2)); }
    RSG_FN void sv(int w, V3 x) const { sf(w, x.x); sf(w + 1, x.y); sf(w + 2, x.z); }
    RSG_FN uint64_t u64(int w) const { return (uint64_t)u(w) | ((uint64_t)u(w + 1) << 32); }
    RSG_FN void su64(int w, uint64_t x) const { u(w) = (uint32_t)x; u(w + 1) = (uint32_t)(x >> 32); }
};

RSG_FN CarRef car_ref(const BatchDev& d, int arena, int slot) {
    CarRef r;
    r.base = d.car + (size_t)slot * d.n + arena;
    r.stride = d.max_cars * d.n;
    return r;
}
RSG_FN CarRef ball_ref(const BatchDev& d, int arena) { CarRef r; r.base = d.ball + arena; r.stride = d.n; return r; }
RSG_FN CarRef pad_ref(const BatchDev& d, int arena, int pad) {
    CarRef r; r.base = d.pad + (size_t)pad * d.n + arena; r.stride = PAD_NUM * d.n; return r;
}
RSG_FN CarRef arena_ref(const BatchDev& d, int arena) { CarRef r; r.base = d.arena + arena; r.stride = d.n; return r; }

RSG_FN Ctrl load_ctrl(const CarRef& r, int w) {
    Ctrl c;
    c.throttle = r.f(w); c.steer = r.f(w + 1); c.pitch = r.f(w + 2); c.yaw = r.f(w + 3); c.roll = r.f(w + 4);
    uint32_t b = r.u(w + 5);
    c.jump = b & CTRL_JUMP; c.boost = b & CTRL_BOOST; c.handbrake = b & CTRL_HANDBRAKE;
    return c;
}
RSG_FN void store_ctrl(const CarRef& r, int w, const Ctrl& c) {
    r.sf(w, c.throttle); r.sf(w + 1, c.steer); r.sf(w + 2, c.pitch); r.sf(w + 3, c.yaw); r.sf(w + 4, c.roll);
    r.u(w + 5) = (c.jump ? CTRL_JUMP : 0u) | (c.boost ? CTRL_BOOST : 0u) | (c.handbrake ? CTRL_HANDBRAKE : 0u);
}

RSG_FN M3 load_rot(const CarRef& r, int w) { return m3(r.v(w), r.v(w + 3), r.v(w + 6)); }
RSG_FN void store_rot(const CarRef& r, int w, const M3& m) { r.sv(w, m.r[0]); r.sv(w + 3, m.r[1]); r.sv(w + 6, m.r[2]); }

RSG_FN void load_car(const CarRef& r, Car& c) {
    c.b.pos = r.v(CW_POS);
    c.api_rot = load_rot(r, CW_ROT);
    c.b.vel = r.v(CW_VEL);
    c.b.ang = r.v(CW_ANGVEL);
    c.flags = r.u(CW_FLAGS);
    c.b.rot = flag(c.flags, RSG_CF_DEMOED) ? load_rot(r, CW_DEMO_BODY_ROT) : c.api_rot;
    c.flip_torque = r.v(CW_FLIP_TORQUE);
    c.jump_time = r.f(CW_JUMP_TIME);
    c.flip_time = r.f(CW_FLIP_TIME);
    c.air_time = r.f(CW_AIR_TIME);
    c.air_time_since_jump = r.f(CW_AIR_TIME_SINCE_JUMP);
    c.boost = r.f(CW_BOOST);
    c.time_since_boosted = r.f(CW_TIME_SINCE_BOOSTED);
    c.boosting_time = r.f(CW_BOOSTING_TIME);
    c.supersonic_time = r.f(CW_SUPERSONIC_TIME);
    c.handbrake = r.f(CW_HANDBRAKE_VAL);
    c.af_timer = r.f(CW_AUTOFLIP_TIMER);
    c.af_scale = r.f(CW_AUTOFLIP_SCALE);
    c.world_normal = r.v(CW_WORLD_NORMAL);
    c.other_id = r.u(CW_OTHER_CAR_ID);
    c.car_cooldown = r.f(CW_CAR_COOLDOWN);
    c.demo_timer = r.f(CW_DEMO_TIMER);
    c.hit_rel_pos = r.v(CW_HIT_REL_POS);
    c.hit_ball_pos = r.v(CW_HIT_BALL_POS);
    c.hit_extra_vel = r.v(CW_HIT_EXTRA_VEL);
    c.hit_tick = r.u64(CW_HIT_TICK);
    c.hit_extra_tick = r.u64(CW_HIT_EXTRA_TICK);
    c.ticks_since_set = r.u64(CW_TICKS_SINCE_SET);
    c.last = load_ctrl(r, CW_LAST_CTRL);
    c.engine = r.f(CW_ENGINE_FORCE);
    c.brake = r.f(CW_BRAKE);
    c.steer = r.f(CW_STEER);
    for (int i = 0; i < 4; i++) {
        c.lat[i] = r.f(CW_LAT_FRICTION + i);
        c.lon[i] = r.f(CW_LONG_FRICTION + i);
        c.push[i] = r.f(CW_PUSHBACK + i);
    }
    c.ctrl = load_ctrl(r, CW_CTRL);
    c.gen_redraw = r.u(CW_GEN_REDRAW);
    c.gen_left = r.u(CW_GEN_TICKS_LEFT);
    uint32_t cfg = r.u(CW_CONFIG);
    c.preset = cfg & 0xFF;
    c.team = (cfg >> 8) & 0xFF;
    c.id = r.u(CW_ID);
    c.vel_cache = v3zero();
}

RSG_FN void load_bp(const CarRef& r, int w_cell, int w_seq, int& cell, uint32_t& seq) {
    cell = (int)r.u(w_cell);
    seq = r.u(w_seq);
    if (cell < 0 || cell >= BP_NUM_CELLS) cell = 0;
}

RSG_FN void store_car(const CarRef& r, const Car& c) {
    r.sv(CW_POS, c.b.pos);
    store_rot(r, CW_ROT, c.api_rot);
    if (flag(c.flags, RSG_CF_DEMOED)) store_rot(r, CW_DEMO_BODY_ROT, c.b.rot);
    r.sv(CW_VEL, c.b.vel);
    r.sv(CW_ANGVEL, c.b.ang);
    r.u(CW_FLAGS) = c.flags;
    r.sv(CW_FLIP_TORQUE, c.flip_torque);
    r.sf(CW_JUMP_TIME, c.jump_time);
    r.sf(CW_FLIP_TIME, c.flip_time);
    r.sf(CW_AIR_TIME, c.air_time);
    r.sf(CW_AIR_TIME_SINCE_JUMP, c.air_time_since_jump);
    r.sf(CW_BOOST, c.boost);
    r.sf(CW_TIME_SINCE_BOOSTED, c.time_since_boosted);
    r.sf(CW_BOOSTING_TIME, c.boosting_time);
    r.sf(CW_SUPERSONIC_TIME, c.supersonic_time);
    r.sf(CW_HANDBRAKE_VAL, c.handbrake);
    r.sf(CW_AUTOFLIP_TIMER, c.af_timer);
    r.sf(CW_AUTOFLIP_SCALE, c.af_scale);
    r.sv(CW_WORLD_NORMAL, c.world_normal);
    r.u(CW_OTHER_CAR_ID) = c.other_id;
    r.sf(CW_CAR_COOLDOWN, c.car_cooldown);
    r.sf(CW_DEMO_TIMER, c.demo_timer);
    r.sv(CW_HIT_REL_POS, c.hit_rel_pos);
    r.sv(CW_HIT_BALL_POS, c.hit_ball_pos);
    r.sv(CW_HIT_EXTRA_VEL, c.hit_extra_vel);
    r.su64(CW_HIT_TICK, c.hit_tick);
    r.su64(CW_HIT_EXTRA_TICK, c.hit_extra_tick);
    r.su64(CW_TICKS_SINCE_SET, c.ticks_since_set);
    store_ctrl(r, CW_LAST_CTRL, c.last);
    r.sf(CW_ENGINE_FORCE, c.engine);
    r.sf(CW_BRAKE, c.brake);
    r.sf(CW_STEER, c.steer);
    for (int i = 0; i < 4; i++) {
        r.sf(CW_LAT_FRICTION + i, c.lat[i]);
        r.sf(CW_LONG_FRICTION + i, c.lon[i]);
        r.sf(CW_PUSHBACK + i, c.push[i]);
    }
    store_ctrl(r, CW_CTRL, c.ctrl);
    r.u(CW_GEN_REDRAW) = c.gen_redraw;
    r.u(CW_GEN_TICKS_LEFT) = c.gen_left;
}

// The car body's mass properties: the preset's unless the arena's
// SetMutatorConfig replaced them (`inv_mass`, `inertia_mass` nonzero).
RSG_FN void car_init_body(Car& c, float inv_mass, float inertia_mass) {
    const CarPhys& cp = c.phys;
    c.b.force = v3zero();
    c.b.torque = v3zero();
    c.b.inv_i_local = cp.inv_inertia_local;
    if (inertia_mass != 0.f)
        c.b.inv_i_local = box_inv_inertia(cp.box_implicit + v3(cp.box_margin, cp.box_margin, cp.box_margin), inertia_mass);
    c.b.inv_mass = inv_mass != 0.f ? inv_mass : cp.inv_mass;
    c.b.no_rot = false;
    c.b.active = true;
    body_update_inertia(c.b);
}

RSG_FN void ball_init_body(Ball& b) {
    BallPhys bp = make_ball_phys(b.radius_uu, b.mass);
    b.b.force = v3zero();
    b.b.torque = v3zero();
    b.b.inv_i_local = v3(bp.inv_inertia, bp.inv_inertia, bp.inv_inertia);
    b.b.inv_mass = bp.inv_mass;
    b.b.no_rot = true;
    b.b.active = true;
    body_update_inertia(b.b);
    b.vel_cache = v3zero();
}

RSG_FN void load_arena(const BatchDev& d, int a, ArenaLocal& A) {
    CarRef ar = arena_ref(d, a);
    A.tick = ar.u64(AW_TICK);
    A.n_cars = (int)ar.u(AW_NUM_CARS);
    A.mode = d.mode;
    A.mut = d.mut ? d.mut[a] : mutator_defaults(d.mode);
    A.custom_pads = d.n_custom_pads > 0 ? d.custom_pads : nullptr;
    A.n_pads = d.n_custom_pads > 0 ? d.n_custom_pads : pad_count(d.mode);
    A.goals_blue = ar.u(AW_GOALS_BLUE);
    A.goals_orange = ar.u(AW_GOALS_ORANGE);
    A.last_goal_tick = ar.u64(AW_LAST_GOAL_TICK);
    A.resets = ar.u(AW_RESETS);
    A.status = ar.u(AW_STATUS);
    A.ev_buf = d.ev_buf;
    A.ev_count = d.ev_count;
    A.ev_cap = d.ev_cap;
    A.ev_seq = ar.u(AW_EVENT_SEQ);
    A.arena_id = (uint32_t)a;
    {
        // Anything but a permutation of 0..n_cars-1 falls back to the
        // default (reverse creation) order, so no index leaves the arena.
        uint32_t ow = ar.u(AW_CAR_ORDER), seen = 0;
        bool ok = true;
        for (int k = 0; k < MAX_CARS; k++) {
            uint32_t v = (ow >> (4 * k)) & 0xFu;
            A.order[k] = (uint8_t)v;
            if (k < A.n_cars) {
                ok = ok && v < (uint32_t)A.n_cars && !(seen >> v & 1u);
                seen |= 1u << (v & 31u);
            }
        }
        if (!ok)
            for (int k = 0; k < MAX_CARS; k++) A.order[k] = (uint8_t)(k < A.n_cars ? A.n_cars - 1 - k : k);
    }
    A.si.erp = 0.2f;
    A.si.time_step = d.tick_time;
    for (int i = 0; i < A.n_cars; i++) {
        CarRef cr = car_ref(d, a, i);
        load_car(cr, A.cars[i]);
        A.cars[i].phys = make_car_phys(d.cfgs ? d.cfgs[A.cars[i].preset] : preset_config(A.cars[i].preset));
        car_init_body(A.cars[i], cr.f(CW_BODY_INV_MASS), cr.f(CW_BODY_INERTIA_MASS));
        A.cars[i].b.grav = cr.v(CW_BODY_GRAVITY);
        A.cars[i].b.active = !flag(A.cars[i].flags, RSG_CF_BODY_DISABLED);
        load_bp(cr, CW_BP_CELL, CW_BP_SEQ, A.bp_cell[1 + i], A.bp_seq[1 + i]);
    }
    CarRef br = ball_ref(d, a);
    A.ball.b.pos = br.v(BW_POS);
    A.ball.b.rot = load_rot(br, BW_ROT);
    A.ball.b.vel = br.v(BW_VEL);
    A.ball.b.ang = br.v(BW_ANGVEL);
    A.ball.ticks_since_set = br.u64(BW_TICKS_SINCE_SET);
    A.ball.radius_uu = br.f(BW_RADIUS);
    A.ball.mass = br.f(BW_MASS);
    ball_init_body(A.ball);
    A.ball.b.no_rot = d.ball_rot == 0;  // btRigidBody::m_noRot (Ball.cpp:109)
    A.ball.b.active = br.u(BW_AWAKE) != 0;
    A.ball.b.grav = br.v(BW_GRAVITY);
    A.ball.hs_dir = br.f(BW_HS);
    A.ball.hs_speed = br.f(BW_HS + 1);
    A.ball.hs_time = br.f(BW_HS + 2);
    load_bp(br, BW_BP_CELL, BW_BP_SEQ, A.bp_cell[0], A.bp_seq[0]);
    if (mode_has_arena(d.mode)) {
        for (int p = 0; p < PAD_NUM; p++) {
            CarRef pr = pad_ref(d, a, p);
            A.pads[p].active = pr.u(PW_ACTIVE) != 0;
            A.pads[p].cooldown = pr.f(PW_COOLDOWN);
            A.pads[p].prev_locked = pr.u(PW_PREV_LOCKED);
            A.pads[p].cur_locked = -1;
        }
    }
}

RSG_FN void store_arena(const BatchDev& d, int a, const ArenaLocal& A) {
    CarRef ar = arena_ref(d, a);
    ar.su64(AW_TICK, A.tick);
    ar.u(AW_GOALS_BLUE) = A.goals_blue;
    ar.u(AW_GOALS_ORANGE) = A.goals_orange;
    ar.su64(AW_LAST_GOAL_TICK, A.last_goal_tick);
    ar.u(AW_RESETS) = A.resets;
    ar.u(AW_STATUS) = A.status;
    ar.u(AW_EVENT_SEQ) = A.ev_seq;
    {
        uint32_t ow = 0;
        for (int k = 0; k < MAX_CARS; k++) ow |= (uint32_t)(A.order[k] & 0xFu) << (4 * k);
        ar.u(AW_CAR_ORDER) = ow;
    }
    for (int i = 0; i < A.n_cars; i++) {
        CarRef cr = car_ref(d, a, i);
        store_car(cr, A.cars[i]);
        cr.u(CW_BP_CELL) = (uint32_t)A.bp_cell[1 + i];
        cr.u(CW_BP_SEQ) = A.bp_seq[1 + i];
    }
    CarRef br = ball_ref(d, a);
    br.u(BW_BP_CELL) = (uint32_t)A.bp_cell[0];
    br.u(BW_BP_SEQ) = A.bp_seq[0];
    br.sv(BW_POS, A.ball.b.pos);
    store_rot(br, BW_ROT, A.ball.b.rot);
    br.sv(BW_VEL, A.ball.b.vel);
    br.sv(BW_ANGVEL, A.ball.b.ang);
    br.su64(BW_TICKS_SINCE_SET, A.ball.ticks_since_set);
    br.sf(BW_HS, A.ball.hs_dir);
    br.sf(BW_HS + 1, A.ball.hs_speed);
    br.sf(BW_HS + 2, A.ball.hs_time);
    br.u(BW_AWAKE) = A.ball.b.active ? 1u : 0u;
    if (mode_has_arena(d.mode)) {
        for (int p = 0; p < PAD_NUM; p++) {
            CarRef pr = pad_ref(d, a, p);
            pr.u(PW_ACTIVE) = A.pads[p].active ? 1u : 0u;
            pr.sf(PW_COOLDOWN, A.pads[p].cooldown);
            pr.u(PW_PREV_LOCKED) = A.pads[p].prev_locked;
        }
    }
}

// ---------------------------------------------------------------------------
// Kickoff reset (Arena::ResetToRandomKickoff, src/Sim/Arena/Arena.cpp:113-223)
// with a counter-based shuffle of the 5 kickoff slots in place of the
// reference's std::default_random_engine.  Cars are taken in slot order per
// team.
RSG_FN void arena_kickoff_reset(ArenaLocal& A, uint64_t seed, uint32_t arena_id) {
    const bool hs = A.mode == RSG_MODE_HEATSEEKER;
    const int n_slots = hs ? 4 : 5;  // CAR_SPAWN_LOCATION_AMOUNT(_HEATSEEKER)
    int slots[5] = {0, 1, 2, 3, 4};
    uint64_t h = rsg_draw(seed ^ 0x4B1C0FFull, arena_id, 0xFFFFu, A.resets, 7);
    for (int i = n_slots - 1; i > 0; i--) {
        int j = (int)((h >> (i * 8)) % (uint64_t)(i + 1));
        int t = slots[i]; slots[i] = slots[j]; slots[j] = t;
    }
    A.resets++;
    int idx_team[2] = {0, 0};
    int num_at_respawn[4] = {0, 0, 0, 0};
    for (int oi = 0; oi < A.n_cars; oi++) {
        Car& c = A.cars[A.order[oi]];
        int k = idx_team[c.team]++;
        SpawnPos sp;
        if (k < n_slots) {
            sp = kickoff_spawn_of(A.mode, slots[k]);
        } else {
            int ri = (k - n_slots) % n_slots;
            ri = ri < 4 ? ri : 3;
            sp = respawn_spawn_of(A.mode, ri);
            sp.y += 250.f * num_at_respawn[ri];
            num_at_respawn[ri]++;
        }
        car_respawn_state(c, A.mode, 0, A.mut.car_spawn_boost_amount);
        V3 pos = v3(sp.x, sp.y, CAR_SPAWN_REST_Z);
        float yaw = sp.yaw;
        if (c.team != RSG_TEAM_BLUE) {
            pos = pos * v3(-1, -1, 1);
            yaw = (float)((double)yaw + PI_D);
        }
        c.b.pos = pos * UU_TO_BT;
        c.b.rot = yaw_basis(yaw);
        c.api_rot = c.b.rot;
        body_update_inertia(c.b);
        c.flags = RSG_CF_ON_GROUND | (c.flags & RSG_CF_BODY_DISABLED);
    }
    // BallState() then SetState (Ball.cpp:36-51); Heatseeker starts the ball
    // rolling towards a random side's corner (Arena.cpp:200-204), the side
    // from the same counter-based draw
    A.ball.b.pos = v3(0, 0, BALL_REST_Z) * UU_TO_BT;
    A.ball.b.rot = m3_identity();
    A.ball.b.vel = v3zero();
    A.ball.b.ang = v3zero();
    if (hs) {
        const float sy = ((h >> 56) & 1u) ? 1.f : -1.f;
        A.ball.b.pos = v3(HS_BALL_START_X, HS_BALL_START_Y * sy, HS_BALL_START_Z) * UU_TO_BT;
        A.ball.b.vel = v3(0.f, HS_BALL_START_VY * sy, HS_BALL_START_VZ) * UU_TO_BT;
        A.ball.b.active = true;  // Ball::SetState wakes a moving ball
    }
    A.ball.hs_dir = 0;
    A.ball.hs_speed = HS_INITIAL_TARGET_SPEED;
    A.ball.hs_time = 0;
    A.ball.ticks_since_set = 0;
    A.ball.vel_cache = v3zero();
    body_update_inertia(A.ball.b);
    for (int p = 0; p < A.n_pads; p++) {
        A.pads[p].active = true;
        A.pads[p].cooldown = 0;
        A.pads[p].prev_locked = 0;
        A.pads[p].cur_locked = -1;
    }
}

// ---------------------------------------------------------------------------
// Boost pads (src/Sim/BoostPad/BoostPad.cpp, BoostPadGrid.cpp)

RSG_FN bool pad_is_big(const ArenaLocal& A, int p) {
    return A.custom_pads ? A.custom_pads[p].is_big != 0 : p < PAD_NUM_BIG;
}
RSG_FN V3 arena_pad_location(const ArenaLocal& A, int p) {
    if (A.custom_pads) return v3(A.custom_pads[p].pos[0], A.custom_pads[p].pos[1], A.custom_pads[p].pos[2]);
    return pad_location(A.mode, p);
}

// BoostPadGrid::Add cell of each pad; the grid maps each occupied (x,y) cell
// to one pad (src/Sim/BoostPad/BoostPadGrid/BoostPadGrid.cpp:27-41).
RSG_FN int pad_cell_x(V3 p) { return (int)(p.x / PAD_GRID_CELL_SIZE_X + (PAD_GRID_CELLS_X / 2)); }
RSG_FN int pad_cell_y(V3 p) { return (int)(p.y / PAD_GRID_CELL_SIZE_Y + (PAD_GRID_CELLS_Y / 2)); }

RSG_FN void pad_check_collide(const ArenaLocal& A, Pad& pad, int p, const Car& c, int car_idx, const CarPhys& cp) {
    V3 pos_uu = arena_pad_location(A, p);
    V3 pos_bt = pos_uu * UU_TO_BT;
    V3 cpos = c.b.pos;
    bool colliding = false;
    if (pad.prev_locked == c.id) {
        float box_rad = (pad_is_big(A, p) ? PAD_BOX_RAD_BIG : PAD_BOX_RAD_SMALL) * UU_TO_BT;
        V3 bmin = pos_bt - v3(box_rad, box_rad, 0);
        V3 bmax = pos_bt + v3(box_rad, box_rad, PAD_BOX_HEIGHT * UU_TO_BT);
        // car->_rigidBody.getAabb: compound AABB without contact threshold
        V3 center = c.b.rot * cp.local_aabb_center + c.b.pos;
        M3 ab = c.b.rot;
        for (int r = 0; r < 3; r++) ab.r[r] = v3(fabsf(ab.r[r].x), fabsf(ab.r[r].y), fabsf(ab.r[r].z));
        V3 ext = ab * cp.local_aabb_half;
        V3 cmin = center - ext, cmax = center + ext;
        colliding = (bmax.x > cmin.x && bmax.y > cmin.y && bmax.z > cmin.z) &&
                    (bmin.x < cmax.x && bmin.y < cmax.y && bmin.z < cmax.z);
    } else {
        float rad = (pad_is_big(A, p) ? PAD_CYL_RAD_BIG : PAD_CYL_RAD_SMALL) * UU_TO_BT;
        float dx = cpos.x - pos_bt.x, dy = cpos.y - pos_bt.y;
        if ((dx * dx + dy * dy) < (rad * rad))
            colliding = fabsf(cpos.z - pos_bt.z) < (PAD_CYL_HEIGHT * UU_TO_BT);
    }
    if (colliding) pad.cur_locked = car_idx;
}

// ---------------------------------------------------------------------------
// The tick

// Start of Arena::Step: the synthetic controls and the ball's sleep flag.
// Returns whether every wheel ray of the tick may be cast before the per-car
// pre-tick loop: true unless a demolished car respawns this tick, which moves
// it in the middle of the loop.
RSG_FN bool tick_begin(ArenaLocal& A, const TickParams& P) {
    if (P.synth_controls) {
        for (int i = 0; i < A.n_cars; i++) {
            Car& c = A.cars[i];
            rsg_ctrl_gen g = {c.gen_redraw, c.gen_left};
            rsg_controls rc;
            rc.throttle = c.ctrl.throttle; rc.steer = c.ctrl.steer; rc.pitch = c.ctrl.pitch;
            rc.yaw = c.ctrl.yaw; rc.roll = c.ctrl.roll;
            rc.jump = c.ctrl.jump; rc.boost = c.ctrl.boost; rc.handbrake = c.ctrl.handbrake;
            rsg_ctrl_next(P.seed, P.arena_id, (uint32_t)i, &g, &rc);
            c.gen_redraw = g.redraw; c.gen_left = g.ticks_left;
            c.ctrl.throttle = rc.throttle; c.ctrl.steer = rc.steer; c.ctrl.pitch = rc.pitch;
            c.ctrl.yaw = rc.yaw; c.ctrl.roll = rc.roll;
            c.ctrl.jump = rc.jump; c.ctrl.boost = rc.boost; c.ctrl.handbrake = rc.handbrake;
        }
    }

    // 1. ball sleep flag (Arena.cpp:694-700)
    A.ball.b.active = !(len2(A.ball.b.vel) == 0.f && len2(A.ball.b.ang) == 0.f);

    for (int i = 0; i < A.n_cars; i++) {
        const Car& c = A.cars[i];
        if (flag(c.flags, RSG_CF_DEMOED) && rs_max(c.demo_timer - P.dt, 0.f) == 0.f) return false;
    }
    return true;
}

// Wheel ray `wheel` of car `ci` as the pre-tick loop will cast it, computed
// ahead of the loop (valid when tick_begin returned true): poses, broadphase
// cells and contact-response flags are those the car sees at its turn.
RSG_FN void tick_cast_ray(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, int ci, int wheel, RayHit& out) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr);
    int pos = 0;
    while (pos < A.n_cars && A.order[pos] != ci) pos++;
    world.ray_caster_pos = pos;
    const Car& c = A.cars[ci];
    const CarPhys& cp = c.phys;
    V3 source, target;
    wheel_ray(c, cp, wheel, source, target);
    out = world.raycast(source, target, 1 + ci);
}

// Car pre-tick (Car.cpp:58-131) of car `ci`.  `pre` holds every wheel's ray
// result (car-major, 4 per car) when cast ahead, else null.
RSG_FN void car_pre_tick(ArenaLocal& A, const TickParams& P, World& world, int ci, const RayHit* pre) {
    const float dt = P.dt;
    Car& c = A.cars[ci];
    const CarPhys& cp = c.phys;
    c.ctrl.throttle = clampf(c.ctrl.throttle, -1, 1);
    c.ctrl.steer = clampf(c.ctrl.steer, -1, 1);
    c.ctrl.pitch = clampf(c.ctrl.pitch, -1, 1);
    c.ctrl.yaw = clampf(c.ctrl.yaw, -1, 1);
    c.ctrl.roll = clampf(c.ctrl.roll, -1, 1);

    if (flag(c.flags, RSG_CF_DEMOED)) {
        c.demo_timer = rs_max(c.demo_timer - dt, 0.f);
        if (c.demo_timer == 0) {
            int slot = (int)(rsg_draw(P.seed ^ 0xDE40ull, P.arena_id, (uint32_t)ci, (uint32_t)A.tick, 9) % 4u);
            car_respawn_state(c, A.mode, slot, A.mut.car_spawn_boost_amount);
        }
        c.b.active = false;  // DISABLE_SIMULATION + no contact response
    } else {
        c.b.active = true;
    }
    setflag(c.flags, RSG_CF_BODY_DISABLED, !c.b.active);
    if (flag(c.flags, RSG_CF_DEMOED)) return;

    // updateVehicleFirst (btVehicleRL.cpp:215-228)
    // btVehicleRL's wheel records: rebuilt every tick, used only in this pre-tick
    WheelT wh[4];
    for (int i = 0; i < 4; i++) wheel_update_transform(c, cp, i, wh[i]);
    for (int i = 0; i < 4; i++) {
        vehicle_ray_cast(c, wh, cp, i, world, 1 + ci, A.si, pre ? &pre[ci * 4 + i] : nullptr);
    }
    vehicle_calc_friction(c, wh, world, dt);

    bool jump_pressed = c.ctrl.jump && !c.last.jump;
    int n_contact = 0;
    for (int i = 0; i < 4; i++) {
        setflag(c.flags, RSG_CF_WHEEL0 << i, wh[i].in_contact);
        n_contact += wh[i].in_contact ? 1 : 0;
    }
    setflag(c.flags, RSG_CF_ON_GROUND, n_contact >= 3);
    float fwd_speed_uu = dot(c.b.vel, car_forward(c)) * BT_TO_UU;
    car_update_wheels(c, wh, cp, dt, n_contact, fwd_speed_uu);
    if (n_contact < 3)
        car_update_air_torque(c, dt, n_contact == 0);
    else
        setflag(c.flags, RSG_CF_IS_FLIPPING, false);
    car_update_jump(c, dt, jump_pressed, A.mut);
    car_update_auto_flip(c, dt, jump_pressed);
    car_update_double_jump_or_flip(c, dt, jump_pressed, fwd_speed_uu, A.mut);
    if (c.ctrl.throttle != 0.f && ((n_contact > 0 && n_contact < 4) || flag(c.flags, RSG_CF_WORLD_CONTACT)))
        car_update_auto_roll(c, wh, n_contact);
    setflag(c.flags, RSG_CF_WORLD_CONTACT, false);
    vehicle_second(c, wh, cp, dt);
    car_update_boost(c, dt, A.mut);
}

// Whether the car pre-ticks of this tick are independent of their order, so
// that they may run concurrently: the rays were cast ahead and no wheel rests
// on a car (tire friction reads the velocity of a car ground, which that
// car's own pre-tick changes).
RSG_FN bool car_pre_ticks_independent(const ArenaLocal& A, const RayHit* pre) {
    if (!pre) return false;
    for (int k = 0; k < 4 * A.n_cars; k++)
        if (pre[k].hit && !pre[k].is_static && pre[k].body >= 1) return false;
    return true;
}

// Pads pre-tick (BoostPad.cpp:51-60).
RSG_FN void pads_pre_tick(ArenaLocal& A, const TickParams& P) {
    if (!mode_has_arena(P.mode) || A.n_cars == 0) return;
    for (int p = 0; p < A.n_pads; p++) {
        Pad& pad = A.pads[p];
        if (pad.cooldown > 0) pad.cooldown = rs_max(pad.cooldown - P.dt, 0.f);
        pad.active = (pad.cooldown == 0);
        pad.cur_locked = -1;
    }
}

// Car post-tick (Car.cpp:133-193) and _FinishPhysicsTick of car `ci`.
RSG_FN void car_post_tick(ArenaLocal& A, const TickParams& P, int ci) {
    const float dt = P.dt;
    Car& c = A.cars[ci];
    if (flag(c.flags, RSG_CF_DEMOED)) return;
    c.api_rot = c.b.rot;
    float s2 = len2(c.b.vel * BT_TO_UU);
    bool ss = flag(c.flags, RSG_CF_SUPERSONIC);
    if (ss && c.supersonic_time < SUPERSONIC_MAINTAIN_MAX_TIME)
        ss = s2 >= SUPERSONIC_MAINTAIN_MIN_SPEED * SUPERSONIC_MAINTAIN_MIN_SPEED;
    else
        ss = s2 >= SUPERSONIC_START_SPEED * SUPERSONIC_START_SPEED;
    setflag(c.flags, RSG_CF_SUPERSONIC, ss);
    if (ss) c.supersonic_time += dt; else c.supersonic_time = 0;
    if (c.car_cooldown > 0) c.car_cooldown = rs_max(c.car_cooldown - dt, 0.f);
    c.last = c.ctrl;

    // _FinishPhysicsTick
    if (!is_zero(c.vel_cache)) {
        c.b.vel += c.vel_cache;
        c.vel_cache = v3zero();
    }
    V3 v = c.b.vel, w = c.b.ang;
    if (len2(v) > (CAR_MAX_SPEED * UU_TO_BT) * (CAR_MAX_SPEED * UU_TO_BT))
        v = normalized(v) * (CAR_MAX_SPEED * UU_TO_BT);
    if (len2(w) > (CAR_MAX_ANG_SPEED * CAR_MAX_ANG_SPEED))
        w = normalized(w) * CAR_MAX_ANG_SPEED;
    c.b.vel = v;
    c.b.ang = w;
    c.ticks_since_set++;
}

// Boost pad p against every car, in the car update order (the reference
// checks each car against the pads of its 3x3 grid window right after the
// car's post-tick, Arena.cpp:720-734 and BoostPadGrid.cpp:5-25; a pad's
// test reads only that car's pose, boost and flags, which no later post-tick
// changes, and the last colliding car locks the pad).
RSG_FN void pad_check_cars(ArenaLocal& A, int p) {
    if (A.custom_pads) {
        // custom pads: every car against every pad, without the grid's
        // early outs (Arena.cpp:723-728)
        for (int oi = 0; oi < A.n_cars; oi++) {
            const int ci = A.order[oi];
            pad_check_collide(A, A.pads[p], p, A.cars[ci], ci, A.cars[ci].phys);
        }
        return;
    }
    V3 pl = pad_location(A.mode, p);
    int px = pad_cell_x(pl), py = pad_cell_y(pl);
    for (int oi = 0; oi < A.n_cars; oi++) {
        const int ci = A.order[oi];
        const Car& c = A.cars[ci];
        if (flag(c.flags, RSG_CF_DEMOED) || c.boost >= BOOST_MAX) continue;
        V3 cp_uu = c.b.pos * BT_TO_UU;
        if (cp_uu.z > PAD_GRID_EXTENT_Z) continue;
        int ix = (int)(cp_uu.x / PAD_GRID_CELL_SIZE_X + (PAD_GRID_CELLS_X / 2));
        int iy = (int)(cp_uu.y / PAD_GRID_CELL_SIZE_Y + (PAD_GRID_CELLS_Y / 2));
        int x0 = ix - 1 > 0 ? ix - 1 : 0, x1 = ix + 1 < PAD_GRID_CELLS_X - 1 ? ix + 1 : PAD_GRID_CELLS_X - 1;
        int y0 = iy - 1 > 0 ? iy - 1 : 0, y1 = iy + 1 < PAD_GRID_CELLS_Y - 1 ? iy + 1 : PAD_GRID_CELLS_Y - 1;
        if (px >= x0 && px <= x1 && py >= y0 && py <= y1) pad_check_collide(A, A.pads[p], p, c, ci, c.phys);
    }
}

// End of Arena::Step: pads post-tick, ball finish, goal test, tick count,
// kickoff reset on goal.
RSG_FN void tick_end(ArenaLocal& A, const TickParams& P) {
    const float dt = P.dt;
    const bool soccar = mode_has_arena(P.mode);
    // 6. pads post-tick (BoostPad.cpp:88-107)
    if (soccar && A.n_cars > 0) {
        for (int p = 0; p < A.n_pads; p++) {
            Pad& pad = A.pads[p];
            uint32_t locked_id = 0;
            if (pad.cur_locked >= 0) {
                Car& c = A.cars[pad.cur_locked];
                locked_id = c.id;
                if (pad.active) {
                    float add = pad_is_big(A, p) ? PAD_BOOST_AMOUNT_BIG : PAD_BOOST_AMOUNT_SMALL;
                    c.boost = rs_min(c.boost + add, BOOST_MAX);
                    pad.active = false;
                    pad.cooldown = pad_is_big(A, p) ? A.mut.boost_pad_cooldown_big : A.mut.boost_pad_cooldown_small;
                }
            }
            pad.prev_locked = locked_id;
        }
    }

    // 7. ball finish (Ball.cpp:117-143)
    {
        Body& b = A.ball.b;
        if (!is_zero(A.ball.vel_cache)) {
            b.vel += A.ball.vel_cache;
            A.ball.vel_cache = v3zero();
        }
        V3 v = b.vel, w = b.ang;
        float max_bt = A.mut.ball_max_speed * UU_TO_BT;
        if (len2(v) > max_bt * max_bt) v = normalized(v) * max_bt;
        if (len2(w) > (BALL_MAX_ANG_SPEED * BALL_MAX_ANG_SPEED)) w = normalized(w) * BALL_MAX_ANG_SPEED;
        b.vel = v;
        b.ang = w;
        A.ball.ticks_since_set++;
    }

    A.si.time_step = dt;

    // 8. goal test (Arena.cpp:893-901) and optional kickoff reset
    bool scored = false;
    if (soccar) {
        float y = A.ball.b.pos.y * BT_TO_UU;
        if (ball_scored(A.mode, A.ball.b.pos, A.mut)) {
            scored = true;
            if (y > 0) A.goals_blue++; else A.goals_orange++;
            A.last_goal_tick = A.tick;
            emit_event(A, RSG_EVENT_GOAL, y > 0 ? (uint32_t)RSG_TEAM_BLUE : (uint32_t)RSG_TEAM_ORANGE, 0u);
        }
    }
    A.tick++;
    // non-finite poses or velocities leave the reference's results for good
    {
        V3 s = A.ball.b.pos + A.ball.b.vel;
        for (int i = 0; i < A.n_cars; i++) s += A.cars[i].b.pos + A.cars[i].b.vel + A.cars[i].b.ang;
        if (!(isfinite(s.x) && isfinite(s.y) && isfinite(s.z))) A.status |= RSG_STATUS_NONFINITE;
    }
    if (scored && P.auto_reset) arena_kickoff_reset(A, P.seed, P.arena_id);
}

// Angle::FromVec (src/Math/MathTypes/MathTypes.cpp:80-100): yaw and pitch
// of a direction.
RSG_FN void hs_angle_from_vec(V3 f, float& yaw, float& pitch) {
    const float eps = 1.1920928955078125e-07f;
    if (fabsf(f.y) > eps || fabsf(f.x) > eps) {
        yaw = atan2f(f.y, f.x);
        float d2 = sqrtf(f.x * f.x + f.y * f.y);
        pitch = -atan2f(-f.z, d2);
    } else {
        yaw = 0;
        pitch = f.z > eps ? (float)(PI_D / 2) : (f.z < -eps ? (float)(-PI_D / 2) : 0.f);
    }
}

// Math::WrapNormalizeFloat (src/Math/Math.cpp:66-73).
RSG_FN float hs_wrap(float v, float mm) {
    float r = fmodf(v, mm * 2);
    if (r > mm) r -= mm * 2;
    else if (r < -mm) r += mm * 2;
    return r;
}

// Math::RoundAngleUE3 (src/Math/Math.cpp:75-91): the rotator's 16-bit
// units, rounded down to multiples of 4.
RSG_FN float hs_round_ue3(float a) {
    constexpr float to_ints = (float)((double)(1 << 15) / PI_D);
    constexpr float back = (1.f / to_ints) * 4.f;
    int r = ((int)(a * to_ints) >> 2) & (0x4000 - 1);
    return (float)r * back;
}

// Ball::_PreTickUpdate for Heatseeker (src/Sim/Ball/Ball.cpp:161-206): once
// a team has touched it, the ball turns its velocity towards the target net
// and blends its speed towards the target speed.
RSG_FN void ball_pre_tick_heatseeker(Ball& ball, float dt) {
    if (ball.hs_dir == 0) return;
    const float pi_f = (float)PI_D, half_pi_f = (float)(PI_D / 2);
    V3 pos = ball.b.pos * BT_TO_UU, vel = ball.b.vel * BT_TO_UU;
    float vyaw, vpitch, gyaw, gpitch;
    hs_angle_from_vec(vel, vyaw, vpitch);
    hs_angle_from_vec(v3(0, HS_TARGET_Y * ball.hs_dir, HS_TARGET_Z) - pos, gyaw, gpitch);
    // Angle::GetDeltaTo, normalised (MathTypes.h:305-310)
    float dyaw = hs_wrap(gyaw - vyaw, pi_f), dpitch = hs_wrap(gpitch - vpitch, half_pi_f);
    float speed = len(vel);
    float base = (speed / HS_MAX_SPEED) * dt;
    float yaw = vyaw + dyaw * base * HS_HORIZONTAL_BLEND;
    float pitch = vpitch + dpitch * base * HS_VERTICAL_BLEND;
    yaw = hs_wrap(yaw, pi_f);
    pitch = hs_wrap(pitch, half_pi_f);
    pitch = clampf(pitch, -HS_MAX_TURN_PITCH, HS_MAX_TURN_PITCH);
    yaw = hs_round_ue3(yaw);
    pitch = hs_round_ue3(pitch);
    float new_speed = speed + ((ball.hs_speed - speed) * HS_SPEED_BLEND);
    // Angle::GetForwardVec (MathTypes.cpp:102-110)
    float cp = cosf(-pitch), cy = cosf(yaw), sy = sinf(yaw), sp = sinf(-pitch);
    V3 dir = v3(cp * cy, cp * sy, -sp);
    ball.b.vel = (dir * new_speed) * UU_TO_BT;
    ball.hs_time += dt;
}

// Ball::_PreTickUpdate for Hoops (src/Sim/Ball/Ball.cpp:209-233): the tick
// at which BALL_HOOPS_LAUNCH_DELAY passes since the ball was last set, a ball
// still frozen at the kickoff spot gets an upward impulse and wakes.
RSG_FN void ball_pre_tick_hoops(Ball& ball, float dt) {
    float cur = (float)ball.ticks_since_set * dt;
    float prev = cur - dt;
    if (!(prev < BALL_HOOPS_LAUNCH_DELAY && cur >= BALL_HOOPS_LAUNCH_DELAY)) return;
    Body& b = ball.b;
    V3 vel = b.vel * BT_TO_UU, pos = b.pos * BT_TO_UU;
    if (!(is_zero(vel) && is_zero(b.ang) && pos.x == 0.f && pos.y == 0.f)) return;
    float mass = b.inv_mass == 0.f ? 0.f : 1.f / b.inv_mass;  // btRigidBody::getMass
    V3 impulse = (v3(0, 0, BALL_HOOPS_LAUNCH_Z_VEL) * mass) * UU_TO_BT;
    b.vel += impulse * b.inv_mass;  // applyCentralImpulse (linear factor 1)
    b.active = true;
}

// The ball's pre-tick of the mode.
RSG_FN void ball_pre_tick(Ball& ball, int mode, float dt) {
    if (mode == RSG_MODE_HEATSEEKER) ball_pre_tick_heatseeker(ball, dt);
    else if (mode == RSG_MODE_HOOPS) ball_pre_tick_hoops(ball, dt);
}

// The world step of the tick (btDiscreteDynamicsWorld::stepSimulation),
// preceded by the ball's pre-tick, which runs after the pads' (Arena.cpp:715).
RSG_FN void tick_world(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws) {
    ball_pre_tick(A.ball, P.mode, P.dt);
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.step(P.dt);
}

// The same world step in three parts for an arena whose lanes share the
// narrowphase: the first and last on one lane, the pairs on every lane
// (`lane` of `lanes`, effects recorded in `log`), with the arena unchanged
// in between.  A pair whose contact needs the penetration pass is run again
// by the last part, on the arena's own work space.
RSG_FN void tick_world_begin(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws) {
    ball_pre_tick(A.ball, P.mode, P.dt);
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.step_begin(P.dt);
}
RSG_FN void tick_world_pairs(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws, PairLog* log,
                             int lane, int lanes) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.run_pairs_recorded(lane, lanes, nullptr, log);
}
RSG_FN void tick_world_end(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws,
                           const PairLog* log) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.step_end(P.dt, log);
}

// The begin and end parts in finer pieces, for lanes that share an arena:
// the per-body work (proxy AABBs, integration) goes one body per lane
// (`lane` of `lanes`), the rest stays on one lane, with the arena unchanged
// by other lanes in between.  In order they equal tick_world_begin and
// tick_world_end.
RSG_FN void tick_world_prologue(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws) {
    ball_pre_tick(A.ball, P.mode, P.dt);
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.step_prologue(P.dt);
}
RSG_FN void tick_world_aabbs(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws, int lane,
                             int lanes) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    for (int i = lane; i < world.num_bodies(); i += lanes) world.body_aabb(i, P.dt);
}
RSG_FN void tick_world_broadphase(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.step_broadphase();
}
RSG_FN void tick_world_solve(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws,
                             const PairLog* log) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    world.step_solve(P.dt, log);
}
RSG_FN void tick_world_integrate(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws, int lane,
                                 int lanes) {
    World world(A, geo, mode_has_arena(P.mode) && geo != nullptr, ws);
    for (int i = lane; i < world.num_bodies(); i += lanes) world.body_integrate(i, P.dt);
}

// The rest of Arena::Step after tick_begin, serially.  `pre` holds every
// wheel's ray result (car-major, 4 per car) when cast ahead, else null.
RSG_FN void tick_rest(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, const RayHit* pre, StepSpace* ws) {
    {
        World world(A, geo, mode_has_arena(P.mode) && geo != nullptr);
        // 2. car pre-tick, in the arena's car update order
        for (int oi = 0; oi < A.n_cars; oi++) car_pre_tick(A, P, world, A.order[oi], pre);
    }
    // 3. pads pre-tick
    pads_pre_tick(A, P);
    // 4. world step
    tick_world(A, P, geo, ws);
    // 5. cars post-tick + pad collisions
    for (int oi = 0; oi < A.n_cars; oi++) car_post_tick(A, P, A.order[oi]);
    if (mode_has_arena(P.mode))
        for (int p = 0; p < A.n_pads; p++) pad_check_cars(A, p);
    tick_end(A, P);
}

// Arena::Step (src/Sim/Arena/Arena.cpp:689-756), one tick.
RSG_FN void arena_tick(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws) {
    tick_begin(A, P);
    tick_rest(A, P, geo, nullptr, ws);
}

}  // namespace rsg
