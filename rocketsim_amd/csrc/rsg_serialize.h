// rsg_serialize.h -- the reference's arena snapshot byte format
// (Arena::Serialize / Arena::DeserializeNew, src/Sim/Arena/Arena.cpp:546-644,
// over DataStreamOut / DataStreamIn, src/DataStream/DataStreamOut.h:9-59,
// DataStreamIn.h:9-86), for rsg_arena_snapshot records.  Host code only.
//
// Layout rules of the format (little-endian, no padding between fields):
//   * WriteMultiple(a, b, ...) writes a uint32 field count, then each field's
//     raw bytes (SerializeObject, src/DataStream/SerializeObject.h:6-22);
//   * Vec is 16 bytes (x, y, z and a fourth component that is always 0),
//     RotMat three Vecs (forward, right, up), bool one byte, GameMode,
//     Team and DemoMode one byte each, CarControls its 24 raw bytes.
// After the reference's bytes, the serializer appends an extension block
// that DeserializeNew never reads (it stops after the mutator config): the
// arena's raw records in this library's word layout (cars in snapshot order,
// then ball, then arena words), which hold positions and velocities in
// Bullet units and the values the reference keeps outside CarState /
// BallState (per-wheel vehicle values, wheel contact and supersonic flags,
// tick counters, broadphase records).  With it a snapshot resumes exactly.
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "../../include/rsg_state.h"
#include "rsg_api_convert.h"

namespace rsg {

constexpr uint32_t SNAP_EXT_MAGIC = 0x58475352u;  // "RSGX"
constexpr uint32_t SNAP_EXT_VERSION = 3;

static_assert(RSG_MAX_CARS * CAR_WORDS + BALL_WORDS + ARENA_WORDS <= RSG_SNAPSHOT_RAW_WORDS,
              "snapshot raw record too small");

inline uint32_t raw_layout_tag() { return ((uint32_t)CAR_WORDS << 16) | ((uint32_t)BALL_WORDS << 8) | ARENA_WORDS; }
inline uint32_t raw_words_for(int n_cars) { return (uint32_t)(n_cars * CAR_WORDS + BALL_WORDS + ARENA_WORDS); }

// One arena's raw records as a batch of one: car k of the snapshot at slot k.
inline BatchDev raw_view(uint32_t* raw, int n_cars) {
    BatchDev d{};
    d.car = raw;
    d.ball = raw + (size_t)n_cars * CAR_WORDS;
    d.arena = d.ball + BALL_WORDS;
    d.pad = nullptr;
    d.n = 1;
    d.max_cars = n_cars;
    return d;
}

struct ByteOut {
    std::vector<uint8_t> d;
    void raw(const void* p, size_t n) {
        const uint8_t* b = (const uint8_t*)p;
        d.insert(d.end(), b, b + n);
    }
    template <class T>
    void put(T v) { raw(&v, sizeof(T)); }
    void count(uint32_t n) { put<uint32_t>(n); }
    void b8(bool v) { put<uint8_t>(v ? 1 : 0); }
    void vec(float x, float y, float z) {
        float w[4] = {x, y, z, 0.f};
        raw(w, 16);
    }
    void vec(const float* v) { vec(v[0], v[1], v[2]); }
    void rot(const float* api) { vec(api); vec(api + 3); vec(api + 6); }
};

struct ByteIn {
    const uint8_t* p;
    size_t n, pos = 0;
    bool bad = false;
    ByteIn(const uint8_t* data, size_t size) : p(data), n(size) {}
    void raw(void* out, size_t k) {
        if (bad || n - pos < k || pos > n) { bad = true; std::memset(out, 0, k); return; }
        std::memcpy(out, p + pos, k);
        pos += k;
    }
    template <class T>
    T get() { T v; raw(&v, sizeof(T)); return v; }
    bool b8() { return get<uint8_t>() != 0; }
    void vec(float* v) {
        float w[4];
        raw(w, 16);
        v[0] = w[0]; v[1] = w[1]; v[2] = w[2];
    }
    void rot(float* api) { vec(api); vec(api + 3); vec(api + 6); }
    size_t left() const { return bad || pos > n ? 0 : n - pos; }
};

// A snapshot car's CarConfig: car_config[k], or car_preset[k]'s when that
// entry is all zero.
inline rsg_car_config snapshot_car_config(const rsg_arena_snapshot& s, int k) {
    static const rsg_car_config zero = {};
    if (!same_car_config(s.car_config[k], zero)) return s.car_config[k];
    return preset_config(s.car_preset[k] < RSG_PRESET_COUNT ? s.car_preset[k] : 0);
}

// A CarConfig in its serialized field order (CAR_CONFIG_SERIALIZATION_FIELDS,
// CarConfig.h:45-49): dodgeDeadzone, hitboxPosOffset, hitboxSize, front
// connection point, rest length, radius, back connection point, rest length,
// radius.  threeWheels is not serialized.
inline void config_words(const rsg_car_config& r, float out[1 + 3 + 3 + 5 + 5]) {
    float* o = out;
    *o++ = r.dodge_deadzone;
    for (int i = 0; i < 3; i++) *o++ = r.hitbox_pos_offset[i];
    for (int i = 0; i < 3; i++) *o++ = r.hitbox_size[i];
    const rsg_wheel_pair_config* w[2] = {&r.front_wheels, &r.back_wheels};
    for (int k = 0; k < 2; k++) {
        for (int i = 0; i < 3; i++) *o++ = w[k]->connection_point_offset[i];
        *o++ = w[k]->suspension_rest_length;
        *o++ = w[k]->wheel_radius;
    }
}

// MutatorConfig::Serialize (MutatorConfig.cpp:42-45): the field count, then
// MUTATOR_CONFIG_SERIALIZATION_FIELDS (MutatorConfig.h:81-88) in that order.
constexpr int MUTATOR_FIELDS = 30;
inline void write_mutators(ByteOut& o, const rsg_mutator_config& m) {
    o.put<uint16_t>(MUTATOR_FIELDS);
    o.count(MUTATOR_FIELDS);
    o.vec(m.gravity);
    const float f1[] = {m.car_mass, m.car_world_friction, m.car_world_restitution, m.ball_mass, m.ball_max_speed,
                        m.ball_drag, m.ball_world_friction, m.ball_world_restitution, m.jump_accel,
                        m.jump_immediate_force, m.boost_accel_ground, m.boost_accel_air, m.boost_used_per_second,
                        m.respawn_delay, m.car_spawn_boost_amount, m.bump_cooldown_time, m.boost_pad_cooldown_big,
                        m.boost_pad_cooldown_small, m.ball_hit_extra_force_scale, m.bump_force_scale, m.ball_radius};
    for (float f : f1) o.put<float>(f);
    o.b8(m.unlimited_flips);
    o.b8(m.unlimited_double_jumps);
    o.b8(m.recharge_boost_enabled);
    o.put<float>(m.recharge_boost_per_second);
    o.put<float>(m.recharge_boost_delay);
    o.put<uint8_t>(m.demo_mode);
    o.b8(m.enable_team_demos);
    o.put<float>(m.goal_base_threshold_y);
}

// MutatorConfig::Deserialize (MutatorConfig.cpp:47-55), which refuses a
// config written with another field count.
inline std::string read_mutators(ByteIn& in, rsg_mutator_config& m) {
    uint16_t fields = in.get<uint16_t>();
    uint32_t count = in.get<uint32_t>();
    if (in.bad) return "truncated";
    if (fields != MUTATOR_FIELDS || count != (uint32_t)MUTATOR_FIELDS)
        return "mutator config is from a different version of RocketSim, fields don't match";
    in.vec(m.gravity);
    float* f1[] = {&m.car_mass, &m.car_world_friction, &m.car_world_restitution, &m.ball_mass, &m.ball_max_speed,
                   &m.ball_drag, &m.ball_world_friction, &m.ball_world_restitution, &m.jump_accel,
                   &m.jump_immediate_force, &m.boost_accel_ground, &m.boost_accel_air, &m.boost_used_per_second,
                   &m.respawn_delay, &m.car_spawn_boost_amount, &m.bump_cooldown_time, &m.boost_pad_cooldown_big,
                   &m.boost_pad_cooldown_small, &m.ball_hit_extra_force_scale, &m.bump_force_scale, &m.ball_radius};
    for (float* f : f1) *f = in.get<float>();
    m.unlimited_flips = in.b8();
    m.unlimited_double_jumps = in.b8();
    m.recharge_boost_enabled = in.b8();
    m.recharge_boost_per_second = in.get<float>();
    m.recharge_boost_delay = in.get<float>();
    m.demo_mode = in.get<uint8_t>();
    m.enable_team_demos = in.b8();
    m._pad[0] = m._pad[1] = m._pad[2] = 0;
    m.goal_base_threshold_y = in.get<float>();
    if (in.bad) return "truncated";
    if (m.demo_mode > RSG_DEMO_DISABLED) return "unknown DemoMode";
    return "";
}

inline void write_default_arena_config(ByteOut& o) {
    // ArenaConfig defaults (src/Sim/Arena/ArenaConfig/ArenaConfig.h:19-46)
    o.count(5);
    o.vec(-5600.f, -6000.f, 0.f);
    o.vec(5600.f, 6000.f, 2200.f);
    o.put<float>(370.f);
    o.b8(true);   // noBallRot
    o.b8(true);   // useCustomBroadphase
    o.b8(false);  // useCustomBoostPads (ArenaConfig.cpp:8)
}

inline void write_controls_raw(ByteOut& o, const rsg_controls& c) {
    // CarControls as raw bytes: five floats, jump, boost, handbrake, padding
    o.put<float>(c.throttle); o.put<float>(c.steer);
    o.put<float>(c.pitch); o.put<float>(c.yaw); o.put<float>(c.roll);
    o.b8(c.jump != 0); o.b8(c.boost != 0); o.b8(c.handbrake != 0);
    o.put<uint8_t>(0);
}

// Raw records from a snapshot's state records, as set_cars / set_balls write
// them into a new arena's records.
inline void raw_from_records(const rsg_arena_snapshot& s, std::vector<uint32_t>& raw) {
    raw.assign(raw_words_for(s.n_cars), 0u);
    BatchDev v = raw_view(raw.data(), s.n_cars);
    for (int k = 0; k < s.n_cars; k++) {
        CarRef cr = car_ref(v, 0, k);
        cr.u(CW_CONFIG) = (uint32_t)s.car_preset[k] | ((uint32_t)s.car_team[k] << 8);
        cr.u(CW_ID) = s.car_id[k];
        car_from_api(cr, s.cars[k]);
        store_ctrl(cr, CW_CTRL, ctrl_from_api(s.controls[k]));
        init_car_body_words(cr, s.mutators.gravity);
    }
    {
        // the mode's default ball body with the config's gravity, as
        // Arena::DeserializeNew leaves it
        const rsg_mutator_config def = mutator_defaults(s.mode);
        init_ball_body_words(ball_ref(v, 0), def.ball_radius, def.ball_mass, s.mutators.gravity);
    }
    ball_from_api(ball_ref(v, 0), s.ball);
    CarRef ar = arena_ref(v, 0);
    ar.su64(AW_TICK, s.tick_count);
    ar.u(AW_NUM_CARS) = (uint32_t)s.n_cars;
    uint32_t ow = 0;
    for (int k = 0; k < s.n_cars; k++) ow |= (uint32_t)k << (4 * k);
    ar.u(AW_CAR_ORDER) = ow;
}

// State records from raw records (what rsg_batch_get_snapshot reports).
// False when the raw records disagree with the snapshot's cars.
inline bool records_from_raw(rsg_arena_snapshot& s) {
    BatchDev v = raw_view(s.raw, s.n_cars);
    for (int k = 0; k < s.n_cars; k++) {
        CarRef cr = car_ref(v, 0, k);
        uint32_t cfg = cr.u(CW_CONFIG);
        if (cr.u(CW_ID) != s.car_id[k] || ((cfg >> 8) & 0xFF) != s.car_team[k]) return false;
    }
    for (int k = 0; k < s.n_cars; k++) {
        CarRef cr = car_ref(v, 0, k);
        car_to_api(cr, s.cars[k]);
        s.controls[k] = ctrl_to_api(load_ctrl(cr, CW_CTRL));
    }
    ball_to_api(ball_ref(v, 0), s.ball);
    return true;
}

// Arena::Serialize (Arena.cpp:546-573).
inline std::string snapshot_serialize(const rsg_arena_snapshot& s, std::vector<uint8_t>& out) {
    if (s.mode != RSG_MODE_SOCCAR && s.mode != RSG_MODE_HOOPS && s.mode != RSG_MODE_HEATSEEKER &&
        s.mode != RSG_MODE_THE_VOID)
        return "mode must be SOCCAR, HOOPS, HEATSEEKER or THE_VOID";
    if (s.n_cars < 0 || s.n_cars > RSG_MAX_CARS) return "n_cars out of range";
    ByteOut o;
    o.count(4);
    o.put<uint8_t>((uint8_t)s.mode);
    o.put<float>(s.tick_time);
    o.put<uint64_t>(s.tick_count);
    o.put<uint32_t>(s.last_car_id);
    write_default_arena_config(o);

    o.put<uint32_t>((uint32_t)s.n_cars);
    for (int k = 0; k < s.n_cars; k++) {
        if (s.car_team[k] > 1) return "bad car team";
        const rsg_controls& c = s.controls[k];
        const rsg_car_state& cs = s.cars[k];
        o.put<uint8_t>(s.car_team[k]);
        o.put<uint32_t>(s.car_id[k]);
        // Car::Serialize (src/Sim/Car/Car.cpp:325-329): controls, config, state
        o.count(8);
        o.put<float>(c.throttle); o.put<float>(c.steer);
        o.put<float>(c.pitch); o.put<float>(c.yaw); o.put<float>(c.roll);
        o.b8(c.boost != 0); o.b8(c.jump != 0); o.b8(c.handbrake != 0);
        float cfg[17];
        config_words(snapshot_car_config(s, k), cfg);
        o.count(9);
        o.put<float>(cfg[0]);
        o.vec(cfg + 1); o.vec(cfg + 4);
        o.vec(cfg + 7); o.put<float>(cfg[10]); o.put<float>(cfg[11]);
        o.vec(cfg + 12); o.put<float>(cfg[15]); o.put<float>(cfg[16]);
        // CarState::Serialize (Car.cpp:308-314): BallHitInfo, then the fields
        bool hit_valid = (cs.flags & RSG_CF_HIT_VALID) != 0;
        o.b8(hit_valid);
        if (hit_valid) {
            o.count(5);
            o.vec(cs.hit_rel_pos); o.vec(cs.hit_ball_pos); o.vec(cs.hit_extra_vel);
            o.put<uint64_t>(cs.hit_tick); o.put<uint64_t>(cs.hit_extra_tick);
        }
        const uint32_t f = cs.flags;
        o.count(30);
        o.vec(cs.pos); o.rot(cs.rot); o.vec(cs.vel); o.vec(cs.ang_vel);
        o.b8(f & RSG_CF_ON_GROUND); o.b8(f & RSG_CF_HAS_JUMPED); o.b8(f & RSG_CF_HAS_DOUBLE_JUMPED);
        o.b8(f & RSG_CF_HAS_FLIPPED);
        o.vec(cs.flip_rel_torque);
        o.put<float>(cs.jump_time);
        o.b8(f & RSG_CF_IS_FLIPPING);
        o.put<float>(cs.flip_time);
        o.b8(f & RSG_CF_IS_JUMPING);
        o.put<float>(cs.air_time_since_jump);
        o.put<float>(cs.boost);
        o.b8(f & RSG_CF_IS_BOOSTING);
        o.put<float>(cs.boosting_time);
        o.put<float>(cs.time_since_boosted);
        o.put<float>(cs.supersonic_time);
        o.put<float>(cs.handbrake_val);
        o.b8(f & RSG_CF_AUTO_FLIPPING);
        o.put<float>(cs.auto_flip_timer);
        o.put<float>(cs.auto_flip_torque_scale);
        o.b8(f & RSG_CF_DEMOED);
        o.put<float>(cs.demo_respawn_timer);
        write_controls_raw(o, cs.last_controls);
        o.b8(f & RSG_CF_WORLD_CONTACT);
        o.vec(cs.world_contact_normal);
        o.put<uint32_t>(cs.car_contact_other_id);
        o.put<float>(cs.car_contact_cooldown);
    }

    if (mode_has_arena(s.mode)) {
        // BoostPadState::Serialize (src/Sim/BoostPad/BoostPad.cpp:21-25)
        const int np = pad_count(s.mode);
        o.put<uint32_t>((uint32_t)np);
        for (int p = 0; p < np; p++) {
            o.count(3);
            o.b8(s.pads[p].is_active != 0);
            o.put<float>(s.pads[p].cooldown);
            o.put<uint32_t>(s.pads[p].prev_locked_car_id);
        }
    }

    // BallState::Serialize (src/Sim/Ball/Ball.cpp:19-21), dropshot fields at
    // their defaults (Ball.h:34-47)
    o.count(12);
    o.vec(s.ball.pos); o.rot(s.ball.rot); o.vec(s.ball.vel); o.vec(s.ball.ang_vel);
    o.put<float>(s.ball.hs_y_target_dir);
    o.put<float>(s.ball.hs_cur_target_speed);
    o.put<float>(s.ball.hs_time_since_hit);
    o.put<int32_t>(1);
    o.put<float>(0.f);
    o.put<float>(0.f);
    o.b8(false);
    o.put<uint64_t>(0);

    write_mutators(o, s.mutators);

    // extension block: the raw records (built from the state records when the
    // snapshot carries none)
    std::vector<uint32_t> raw;
    if (s.raw_words > 0) {
        if (s.raw_layout != raw_layout_tag() || s.raw_words != raw_words_for(s.n_cars))
            return "raw records of another layout or car count";
        raw.assign(s.raw, s.raw + s.raw_words);
    } else {
        raw_from_records(s, raw);
    }
    o.put<uint32_t>(SNAP_EXT_MAGIC);
    o.put<uint32_t>(SNAP_EXT_VERSION);
    o.put<uint32_t>(raw_layout_tag());
    o.put<uint32_t>((uint32_t)s.n_cars);
    o.raw(raw.data(), raw.size() * 4);
    out.swap(o.d);
    return "";
}

// Broadphase records of a freshly created arena (proxies registered where
// Car::_BulletSetup and Ball::_BulletSetup create them, in creation order).
inline void fresh_broadphase(rsg_arena_snapshot& s, const int* creation_rank) {
    for (int k = 0; k < s.n_cars; k++) {
        CarPhys cp = make_car_phys(snapshot_car_config(s, k));
        s.cars[k].bp_cell = bp_cell_of(cp.local_aabb_center - cp.local_aabb_half);
        s.cars[k].bp_seq = (uint32_t)(1 + creation_rank[k]);
    }
    BallPhys bp = make_ball_phys();
    float ext = bp.radius + 0.08f;
    V3 c = v3(0, 0, bp.radius);
    s.ball.bp_cell = bp_cell_of(c - v3(ext, ext, ext));
    s.ball.bp_seq = 0;
}

// Arena::DeserializeNew (Arena.cpp:575-644) into a snapshot.  Only what this
// library simulates is accepted: Soccar, Heatseeker or THE_VOID, the default
// ArenaConfig, preset car configs.  States are as Car/Ball::SetState
// leave them (tick counters 0, no hidden values) unless the extension block
// is present; then they and the raw records come from it.
inline std::string snapshot_deserialize(const uint8_t* data, size_t size, rsg_arena_snapshot& s) {
    std::memset(&s, 0, sizeof(s));
    ByteIn in(data, size);
    if (in.get<uint32_t>() != 4) return "header field count is not 4";
    s.mode = in.get<uint8_t>();
    s.tick_time = in.get<float>();
    s.tick_count = in.get<uint64_t>();
    s.last_car_id = in.get<uint32_t>();
    if (s.mode != RSG_MODE_SOCCAR && s.mode != RSG_MODE_HOOPS && s.mode != RSG_MODE_HEATSEEKER &&
        s.mode != RSG_MODE_THE_VOID)
        return "game mode is not SOCCAR, HOOPS, HEATSEEKER or THE_VOID";
    {
        ByteOut want;
        write_default_arena_config(want);
        std::vector<uint8_t> got(want.d.size());
        in.raw(got.data(), got.size());
        if (in.bad) return "truncated";
        if (got != want.d) return "non-default ArenaConfig (not supported)";
    }
    uint32_t n = in.get<uint32_t>();
    if (in.bad) return "truncated";
    if (n > RSG_MAX_CARS) return "more cars than RSG_MAX_CARS";
    s.n_cars = (int32_t)n;
    int creation_rank[RSG_MAX_CARS];
    for (int k = 0; k < s.n_cars; k++) {
        rsg_controls& c = s.controls[k];
        rsg_car_state& cs = s.cars[k];
        s.car_team[k] = in.get<uint8_t>();
        s.car_id[k] = in.get<uint32_t>();
        if (s.car_team[k] > 1) return "bad team";
        if (in.get<uint32_t>() != 8) return "car controls field count is not 8";
        c.throttle = in.get<float>(); c.steer = in.get<float>();
        c.pitch = in.get<float>(); c.yaw = in.get<float>(); c.roll = in.get<float>();
        c.boost = in.b8(); c.jump = in.b8(); c.handbrake = in.b8();
        if (in.get<uint32_t>() != 9) return "car config field count is not 9";
        float cfg[17];
        cfg[0] = in.get<float>();
        in.vec(cfg + 1); in.vec(cfg + 4);
        in.vec(cfg + 7); cfg[10] = in.get<float>(); cfg[11] = in.get<float>();
        in.vec(cfg + 12); cfg[15] = in.get<float>(); cfg[16] = in.get<float>();
        {
            // Car::_Deserialize reads the serialized fields; threeWheels
            // keeps CarConfig's default (false)
            rsg_car_config& rc = s.car_config[k];
            std::memset(&rc, 0, sizeof(rc));
            const float* w = cfg;
            rc.dodge_deadzone = *w++;
            for (int i = 0; i < 3; i++) rc.hitbox_pos_offset[i] = *w++;
            for (int i = 0; i < 3; i++) rc.hitbox_size[i] = *w++;
            rsg_wheel_pair_config* wp[2] = {&rc.front_wheels, &rc.back_wheels};
            for (int q = 0; q < 2; q++) {
                for (int i = 0; i < 3; i++) wp[q]->connection_point_offset[i] = *w++;
                wp[q]->suspension_rest_length = *w++;
                wp[q]->wheel_radius = *w++;
            }
            std::string why = car_config_problem(rc);
            if (!why.empty()) return why;
            s.car_preset[k] = (uint8_t)preset_of(rc);
        }
        uint32_t f = 0;
        cs.hit_tick = ~0ull;
        cs.hit_extra_tick = ~0ull;
        if (in.b8()) {
            f |= RSG_CF_HIT_VALID;
            if (in.get<uint32_t>() != 5) return "ball hit info field count is not 5";
            in.vec(cs.hit_rel_pos); in.vec(cs.hit_ball_pos); in.vec(cs.hit_extra_vel);
            cs.hit_tick = in.get<uint64_t>();
            cs.hit_extra_tick = in.get<uint64_t>();
        }
        if (in.get<uint32_t>() != 30) return "car state field count is not 30";
        in.vec(cs.pos); in.rot(cs.rot); in.vec(cs.vel); in.vec(cs.ang_vel);
        if (in.b8()) f |= RSG_CF_ON_GROUND;
        if (in.b8()) f |= RSG_CF_HAS_JUMPED;
        if (in.b8()) f |= RSG_CF_HAS_DOUBLE_JUMPED;
        if (in.b8()) f |= RSG_CF_HAS_FLIPPED;
        in.vec(cs.flip_rel_torque);
        cs.jump_time = in.get<float>();
        if (in.b8()) f |= RSG_CF_IS_FLIPPING;
        cs.flip_time = in.get<float>();
        if (in.b8()) f |= RSG_CF_IS_JUMPING;
        cs.air_time_since_jump = in.get<float>();
        cs.boost = in.get<float>();
        if (in.b8()) f |= RSG_CF_IS_BOOSTING;
        cs.boosting_time = in.get<float>();
        cs.time_since_boosted = in.get<float>();
        cs.supersonic_time = in.get<float>();
        cs.handbrake_val = in.get<float>();
        if (in.b8()) f |= RSG_CF_AUTO_FLIPPING;
        cs.auto_flip_timer = in.get<float>();
        cs.auto_flip_torque_scale = in.get<float>();
        if (in.b8()) f |= RSG_CF_DEMOED;
        cs.demo_respawn_timer = in.get<float>();
        rsg_controls& lc = cs.last_controls;
        lc.throttle = in.get<float>(); lc.steer = in.get<float>();
        lc.pitch = in.get<float>(); lc.yaw = in.get<float>(); lc.roll = in.get<float>();
        lc.jump = in.b8(); lc.boost = in.b8(); lc.handbrake = in.b8();
        (void)in.get<uint8_t>();
        if (in.b8()) f |= RSG_CF_WORLD_CONTACT;
        in.vec(cs.world_contact_normal);
        cs.car_contact_other_id = in.get<uint32_t>();
        cs.car_contact_cooldown = in.get<float>();
        cs.flags = f;
        if (in.bad) return "truncated";
    }
    // creation order is file order; ids are forced to the written ones
    for (int k = 0; k < s.n_cars; k++) creation_rank[k] = k;

    if (mode_has_arena(s.mode)) {
        const int np = pad_count(s.mode);
        if (in.get<uint32_t>() != (uint32_t)np) return "boost pad count is not the arena's";
        for (int p = 0; p < np; p++) {
            if (in.get<uint32_t>() != 3) return "boost pad field count is not 3";
            s.pads[p].is_active = in.b8() ? 1u : 0u;
            s.pads[p].cooldown = in.get<float>();
            s.pads[p].prev_locked_car_id = in.get<uint32_t>();
        }
    }

    if (in.get<uint32_t>() != 12) return "ball state field count is not 12";
    in.vec(s.ball.pos); in.rot(s.ball.rot); in.vec(s.ball.vel); in.vec(s.ball.ang_vel);
    s.ball.hs_y_target_dir = in.get<float>();
    s.ball.hs_cur_target_speed = in.get<float>();
    s.ball.hs_time_since_hit = in.get<float>();
    {
        // dropshot fields: not simulated in these modes
        uint8_t skip[4 + 4 + 4 + 1 + 8];
        in.raw(skip, sizeof(skip));
    }
    {
        std::string err = read_mutators(in, s.mutators);
        if (!err.empty()) return err;
    }
    fresh_broadphase(s, creation_rank);

    // extension block (ignored when it does not match this snapshot)
    if (in.left() >= 16) {
        uint32_t magic = in.get<uint32_t>(), version = in.get<uint32_t>(), layout = in.get<uint32_t>(),
                 cars = in.get<uint32_t>();
        const uint32_t words = raw_words_for(s.n_cars);
        if (magic == SNAP_EXT_MAGIC && version == SNAP_EXT_VERSION && layout == raw_layout_tag() &&
            cars == (uint32_t)s.n_cars && in.left() >= (size_t)words * 4) {
            rsg_arena_snapshot keep = s;
            in.raw(s.raw, (size_t)words * 4);
            s.raw_layout = layout;
            s.raw_words = words;
            if (!records_from_raw(s)) s = keep;
        }
    }
    return "";
}

}  // namespace rsg
