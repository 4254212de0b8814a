// rsg_batch.hip -- kernels of the batched tick step and the C ABI of
// include/rsg.h.
//
// Kernel map:
//   rsg_step_arena_kernel   one lane per arena, K ticks per launch, arena state
//                           in registers/private memory between ticks
//   rsg_reset_kernel        kickoff reset of one / all arenas
//
// Host side keeps a mirror of the SoA state for the get/set API; it is
// uploaded lazily before the next step and downloaded lazily on the next get.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rsg.h"
#include "rsg_tick.h"

using namespace rsg;

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

#define HIP_OK(expr)                                                                 \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) return fail(std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

}  // namespace

// ---------------------------------------------------------------------------
// kernels

__global__ void __launch_bounds__(64) rsg_step_arena_kernel(BatchDev d, int ticks) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= d.n) return;
    ArenaLocal A;
    load_arena(d, a, A);
    TickParams P;
    P.mode = d.mode;
    P.dt = d.tick_time;
    P.seed = d.seed;
    P.arena_id = (uint32_t)a;
    P.synth_controls = d.synth_controls != 0;
    P.auto_reset = d.auto_reset != 0;
    NoMesh mesh;
    for (int t = 0; t < ticks; t++) arena_tick(A, P, mesh);
    store_arena(d, a, A);
}

__global__ void __launch_bounds__(64) rsg_reset_kernel(BatchDev d, int only_arena, uint64_t seed) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= d.n) return;
    if (only_arena >= 0 && a != only_arena) return;
    ArenaLocal A;
    load_arena(d, a, A);
    arena_kickoff_reset(A, seed, (uint32_t)a);
    store_arena(d, a, A);
}

// ---------------------------------------------------------------------------
// batch object

struct rsg_batch {
    BatchDev dev;
    std::vector<int> cars_per_arena, car_offset;
    std::vector<uint8_t> team, preset;
    int total_cars = 0;
    std::vector<uint32_t> h_car, h_ball, h_pad, h_arena;
    bool host_dirty = false;  // host mirror newer than device
    bool dev_newer = false;   // device newer than host mirror
    hipStream_t last_stream = nullptr;
};

namespace {

size_t car_words(const rsg_batch* b) { return (size_t)CAR_WORDS * b->dev.max_cars * b->dev.n; }
size_t ball_words(const rsg_batch* b) { return (size_t)BALL_WORDS * b->dev.n; }
size_t pad_words(const rsg_batch* b) { return (size_t)PAD_WORDS * PAD_NUM * b->dev.n; }
size_t arena_words(const rsg_batch* b) { return (size_t)ARENA_WORDS * b->dev.n; }

BatchDev host_view(rsg_batch* b) {
    BatchDev h = b->dev;
    h.car = b->h_car.data();
    h.ball = b->h_ball.data();
    h.pad = b->h_pad.data();
    h.arena = b->h_arena.data();
    return h;
}

int upload(rsg_batch* b) {
    if (!b->host_dirty) return 0;
    HIP_OK(hipMemcpy(b->dev.car, b->h_car.data(), car_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev.ball, b->h_ball.data(), ball_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev.pad, b->h_pad.data(), pad_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev.arena, b->h_arena.data(), arena_words(b) * 4, hipMemcpyHostToDevice));
    b->host_dirty = false;
    return 0;
}

int download(rsg_batch* b) {
    if (!b->dev_newer) return 0;
    if (b->last_stream) HIP_OK(hipStreamSynchronize(b->last_stream));
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(b->h_car.data(), b->dev.car, car_words(b) * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(b->h_ball.data(), b->dev.ball, ball_words(b) * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(b->h_pad.data(), b->dev.pad, pad_words(b) * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(b->h_arena.data(), b->dev.arena, arena_words(b) * 4, hipMemcpyDeviceToHost));
    b->dev_newer = false;
    return 0;
}

void rot_api_to_rows(const float* api, M3& m) {
    // api: forward, right, up columns (RotMat, src/Math/MathTypes/MathTypes.h:169-170)
    for (int r = 0; r < 3; r++) m.r[r] = v3(api[0 + r], api[3 + r], api[6 + r]);
}
void rot_rows_to_api(const M3& m, float* api) {
    for (int r = 0; r < 3; r++) {
        api[0 + r] = m.r[r].x;
        api[3 + r] = m.r[r].y;
        api[6 + r] = m.r[r].z;
    }
}

Ctrl ctrl_from_api(const rsg_controls& c) {
    Ctrl o;
    o.throttle = c.throttle; o.steer = c.steer; o.pitch = c.pitch; o.yaw = c.yaw; o.roll = c.roll;
    o.jump = c.jump != 0; o.boost = c.boost != 0; o.handbrake = c.handbrake != 0;
    return o;
}
rsg_controls ctrl_to_api(const Ctrl& c) {
    rsg_controls o;
    o.throttle = c.throttle; o.steer = c.steer; o.pitch = c.pitch; o.yaw = c.yaw; o.roll = c.roll;
    o.jump = c.jump; o.boost = c.boost; o.handbrake = c.handbrake; o._pad = 0;
    return o;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* rsg_last_error(void) { return g_err.c_str(); }

rsg_batch* rsg_batch_create(const rsg_batch_desc* desc) {
    if (!desc || desc->n_arenas <= 0) { fail("rsg_batch_create: n_arenas must be > 0"); return nullptr; }
    if (desc->mode != RSG_MODE_SOCCAR && desc->mode != RSG_MODE_THE_VOID) {
        fail("rsg_batch_create: only SOCCAR and THE_VOID are supported");
        return nullptr;
    }
    if (!(desc->tick_rate >= 15.f && desc->tick_rate <= 120.f)) {
        fail("rsg_batch_create: tick_rate must be within [15, 120] (src/Sim/Arena/Arena.cpp:419)");
        return nullptr;
    }
    if (!desc->cars_per_arena) { fail("rsg_batch_create: cars_per_arena is required"); return nullptr; }
    auto* b = new rsg_batch();
    b->dev.n = desc->n_arenas;
    b->dev.mode = desc->mode;
    b->dev.auto_reset = desc->auto_reset;
    b->dev.synth_controls = desc->synth_controls;
    b->dev.seed = desc->seed;
    b->dev.tick_time = 1 / desc->tick_rate;
    int maxc = 0, total = 0;
    b->cars_per_arena.resize(desc->n_arenas);
    b->car_offset.resize(desc->n_arenas);
    for (int a = 0; a < desc->n_arenas; a++) {
        int c = desc->cars_per_arena[a];
        if (c < 0 || c > RSG_MAX_CARS) {
            delete b;
            fail("rsg_batch_create: cars_per_arena out of range [0, RSG_MAX_CARS]");
            return nullptr;
        }
        b->cars_per_arena[a] = c;
        b->car_offset[a] = total;
        total += c;
        if (c > maxc) maxc = c;
    }
    b->total_cars = total;
    b->dev.max_cars = maxc > 0 ? maxc : 1;
    b->team.assign(desc->car_team, desc->car_team + total);
    b->preset.assign(desc->car_preset, desc->car_preset + total);
    for (int i = 0; i < total; i++) {
        if (b->team[i] > 1 || b->preset[i] >= RSG_PRESET_COUNT) {
            delete b;
            fail("rsg_batch_create: bad team or preset");
            return nullptr;
        }
    }
    b->h_car.assign(car_words(b), 0);
    b->h_ball.assign(ball_words(b), 0);
    b->h_pad.assign(pad_words(b), 0);
    b->h_arena.assign(arena_words(b), 0);
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&b->dev.car, car_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.ball, ball_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.pad, pad_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.arena, arena_words(b) * 4)) != hipSuccess) {
        fail(std::string("rsg_batch_create: hipMalloc: ") + hipGetErrorString(e));
        rsg_batch_destroy(b);
        return nullptr;
    }
    // Host-side initial state: cars with config, ball at rest, pads active.
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        CarRef ar = arena_ref(h, a);
        ar.u(AW_NUM_CARS) = (uint32_t)b->cars_per_arena[a];
        for (int i = 0; i < b->cars_per_arena[a]; i++) {
            int k = b->car_offset[a] + i;
            CarRef cr = car_ref(h, a, i);
            cr.u(CW_CONFIG) = (uint32_t)b->preset[k] | ((uint32_t)b->team[k] << 8);
            cr.u(CW_ID) = (uint32_t)(i + 1);
            cr.su64(CW_HIT_TICK, ~0ull);
            cr.su64(CW_HIT_EXTRA_TICK, ~0ull);
            cr.sf(CW_BOOST, BOOST_SPAWN_AMOUNT);
            store_rot(cr, CW_ROT, m3_identity());
        }
        CarRef br = ball_ref(h, a);
        br.sv(BW_POS, v3(0, 0, BALL_REST_Z) * UU_TO_BT);
        store_rot(br, BW_ROT, m3_identity());
        for (int p = 0; p < PAD_NUM; p++) pad_ref(h, a, p).u(PW_ACTIVE) = 1;
    }
    b->host_dirty = true;
    if (upload(b) != 0) { rsg_batch_destroy(b); return nullptr; }
    if (rsg_batch_reset_kickoff(b, -1, desc->seed) != 0) { rsg_batch_destroy(b); return nullptr; }
    return b;
}

void rsg_batch_destroy(rsg_batch* b) {
    if (!b) return;
    (void)hipDeviceSynchronize();
    if (b->dev.car) (void)hipFree(b->dev.car);
    if (b->dev.ball) (void)hipFree(b->dev.ball);
    if (b->dev.pad) (void)hipFree(b->dev.pad);
    if (b->dev.arena) (void)hipFree(b->dev.arena);
    delete b;
}

int rsg_batch_num_arenas(const rsg_batch* b) { return b ? b->dev.n : fail("null batch"); }
int rsg_batch_num_cars(const rsg_batch* b) { return b ? b->total_cars : fail("null batch"); }
int rsg_batch_car_offset(const rsg_batch* b, int arena) {
    if (!b || arena < 0 || arena >= b->dev.n) return fail("rsg_batch_car_offset: bad arena");
    return b->car_offset[arena];
}

int rsg_batch_step(rsg_batch* b, int ticks, void* stream) {
    if (!b) return fail("null batch");
    if (ticks < 0) return fail("rsg_batch_step: ticks must be >= 0");
    if (ticks == 0) return 0;
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    int block = 64;
    int grid = (b->dev.n + block - 1) / block;
    hipLaunchKernelGGL(rsg_step_arena_kernel, dim3(grid), dim3(block), 0, s, b->dev, ticks);
    HIP_OK(hipGetLastError());
    b->last_stream = s;
    b->dev_newer = true;
    return 0;
}

int rsg_batch_sync(rsg_batch* b) {
    if (!b) return fail("null batch");
    HIP_OK(hipStreamSynchronize(b->last_stream));
    return 0;
}

int rsg_batch_reset_kickoff(rsg_batch* b, int arena, uint64_t seed) {
    if (!b) return fail("null batch");
    if (arena >= b->dev.n) return fail("rsg_batch_reset_kickoff: bad arena");
    if (upload(b) != 0) return -1;
    if (download(b) != 0) return -1;
    int block = 64;
    int grid = (b->dev.n + block - 1) / block;
    hipLaunchKernelGGL(rsg_reset_kernel, dim3(grid), dim3(block), 0, b->last_stream, b->dev, arena, seed);
    HIP_OK(hipGetLastError());
    b->dev_newer = true;
    return 0;
}

int rsg_batch_get_cars(rsg_batch* b, rsg_car_state* out) {
    if (!b || !out) return fail("rsg_batch_get_cars: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        ArenaLocal A;
        load_arena(h, a, A);
        for (int i = 0; i < A.n_cars; i++) {
            const Car& c = A.cars[i];
            rsg_car_state& s = out[b->car_offset[a] + i];
            std::memset(&s, 0, sizeof(s));
            s.tick_count_since_update = c.ticks_since_set;
            s.hit_tick = c.hit_tick;
            s.hit_extra_tick = c.hit_extra_tick;
            V3 p = c.b.pos * BT_TO_UU, v = c.b.vel * BT_TO_UU;
            s.pos[0] = p.x; s.pos[1] = p.y; s.pos[2] = p.z;
            rot_rows_to_api(c.b.rot, s.rot);
            s.vel[0] = v.x; s.vel[1] = v.y; s.vel[2] = v.z;
            s.ang_vel[0] = c.b.ang.x; s.ang_vel[1] = c.b.ang.y; s.ang_vel[2] = c.b.ang.z;
            s.flags = c.flags;
            s.flip_rel_torque[0] = c.flip_torque.x; s.flip_rel_torque[1] = c.flip_torque.y; s.flip_rel_torque[2] = c.flip_torque.z;
            s.jump_time = c.jump_time; s.flip_time = c.flip_time;
            s.air_time = c.air_time; s.air_time_since_jump = c.air_time_since_jump;
            s.boost = c.boost; s.time_since_boosted = c.time_since_boosted; s.boosting_time = c.boosting_time;
            s.supersonic_time = c.supersonic_time; s.handbrake_val = c.handbrake;
            s.auto_flip_timer = c.af_timer; s.auto_flip_torque_scale = c.af_scale;
            s.world_contact_normal[0] = c.world_normal.x; s.world_contact_normal[1] = c.world_normal.y; s.world_contact_normal[2] = c.world_normal.z;
            s.car_contact_other_id = c.other_id;
            s.car_contact_cooldown = c.car_cooldown;
            s.demo_respawn_timer = c.demo_timer;
            s.hit_rel_pos[0] = c.hit_rel_pos.x; s.hit_rel_pos[1] = c.hit_rel_pos.y; s.hit_rel_pos[2] = c.hit_rel_pos.z;
            s.hit_ball_pos[0] = c.hit_ball_pos.x; s.hit_ball_pos[1] = c.hit_ball_pos.y; s.hit_ball_pos[2] = c.hit_ball_pos.z;
            s.hit_extra_vel[0] = c.hit_extra_vel.x; s.hit_extra_vel[1] = c.hit_extra_vel.y; s.hit_extra_vel[2] = c.hit_extra_vel.z;
            s.last_controls = ctrl_to_api(c.last);
            s.wheel_engine_force = c.engine; s.wheel_brake = c.brake; s.wheel_steer = c.steer;
            for (int w = 0; w < 4; w++) {
                s.wheel_lat_friction[w] = c.lat[w];
                s.wheel_long_friction[w] = c.lon[w];
                s.wheel_extra_pushback[w] = c.push[w];
            }
        }
    }
    return 0;
}

int rsg_batch_set_cars(rsg_batch* b, const rsg_car_state* in) {
    if (!b || !in) return fail("rsg_batch_set_cars: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        for (int i = 0; i < b->cars_per_arena[a]; i++) {
            const rsg_car_state& s = in[b->car_offset[a] + i];
            CarRef r = car_ref(h, a, i);
            Car c;
            load_car(r, c);
            c.b.pos = v3(s.pos[0], s.pos[1], s.pos[2]) * UU_TO_BT;
            rot_api_to_rows(s.rot, c.b.rot);
            c.b.vel = v3(s.vel[0], s.vel[1], s.vel[2]) * UU_TO_BT;
            c.b.ang = v3(s.ang_vel[0], s.ang_vel[1], s.ang_vel[2]);
            c.flags = s.flags;
            c.flip_torque = v3(s.flip_rel_torque[0], s.flip_rel_torque[1], s.flip_rel_torque[2]);
            c.jump_time = s.jump_time; c.flip_time = s.flip_time;
            c.air_time = s.air_time; c.air_time_since_jump = s.air_time_since_jump;
            c.boost = s.boost; c.time_since_boosted = s.time_since_boosted; c.boosting_time = s.boosting_time;
            c.supersonic_time = s.supersonic_time; c.handbrake = s.handbrake_val;
            c.af_timer = s.auto_flip_timer; c.af_scale = s.auto_flip_torque_scale;
            c.world_normal = v3(s.world_contact_normal[0], s.world_contact_normal[1], s.world_contact_normal[2]);
            c.other_id = s.car_contact_other_id;
            c.car_cooldown = s.car_contact_cooldown;
            c.demo_timer = s.demo_respawn_timer;
            c.hit_rel_pos = v3(s.hit_rel_pos[0], s.hit_rel_pos[1], s.hit_rel_pos[2]);
            c.hit_ball_pos = v3(s.hit_ball_pos[0], s.hit_ball_pos[1], s.hit_ball_pos[2]);
            c.hit_extra_vel = v3(s.hit_extra_vel[0], s.hit_extra_vel[1], s.hit_extra_vel[2]);
            c.hit_tick = s.hit_tick;
            c.hit_extra_tick = s.hit_extra_tick;
            c.ticks_since_set = s.tick_count_since_update;
            c.last = ctrl_from_api(s.last_controls);
            c.engine = s.wheel_engine_force; c.brake = s.wheel_brake; c.steer = s.wheel_steer;
            for (int w = 0; w < 4; w++) {
                c.lat[w] = s.wheel_lat_friction[w];
                c.lon[w] = s.wheel_long_friction[w];
                c.push[w] = s.wheel_extra_pushback[w];
            }
            store_car(r, c);
        }
    }
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_balls(rsg_batch* b, rsg_ball_state* out) {
    if (!b || !out) return fail("rsg_batch_get_balls: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        CarRef r = ball_ref(h, a);
        rsg_ball_state& s = out[a];
        std::memset(&s, 0, sizeof(s));
        V3 p = r.v(BW_POS) * BT_TO_UU, v = r.v(BW_VEL) * BT_TO_UU, w = r.v(BW_ANGVEL);
        s.pos[0] = p.x; s.pos[1] = p.y; s.pos[2] = p.z;
        rot_rows_to_api(load_rot(r, BW_ROT), s.rot);
        s.vel[0] = v.x; s.vel[1] = v.y; s.vel[2] = v.z;
        s.ang_vel[0] = w.x; s.ang_vel[1] = w.y; s.ang_vel[2] = w.z;
        s.tick_count_since_update = r.u64(BW_TICKS_SINCE_SET);
    }
    return 0;
}

int rsg_batch_set_balls(rsg_batch* b, const rsg_ball_state* in) {
    if (!b || !in) return fail("rsg_batch_set_balls: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        CarRef r = ball_ref(h, a);
        const rsg_ball_state& s = in[a];
        r.sv(BW_POS, v3(s.pos[0], s.pos[1], s.pos[2]) * UU_TO_BT);
        M3 m;
        rot_api_to_rows(s.rot, m);
        store_rot(r, BW_ROT, m);
        r.sv(BW_VEL, v3(s.vel[0], s.vel[1], s.vel[2]) * UU_TO_BT);
        r.sv(BW_ANGVEL, v3(s.ang_vel[0], s.ang_vel[1], s.ang_vel[2]));
        r.su64(BW_TICKS_SINCE_SET, s.tick_count_since_update);
    }
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_pads(rsg_batch* b, rsg_pad_state* out) {
    if (!b || !out) return fail("rsg_batch_get_pads: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int p = 0; p < PAD_NUM; p++) {
            CarRef r = pad_ref(h, a, p);
            rsg_pad_state& s = out[a * PAD_NUM + p];
            s.is_active = r.u(PW_ACTIVE);
            s.cooldown = r.f(PW_COOLDOWN);
            s.prev_locked_car_id = r.u(PW_PREV_LOCKED);
        }
    return 0;
}

int rsg_batch_set_pads(rsg_batch* b, const rsg_pad_state* in) {
    if (!b || !in) return fail("rsg_batch_set_pads: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int p = 0; p < PAD_NUM; p++) {
            CarRef r = pad_ref(h, a, p);
            const rsg_pad_state& s = in[a * PAD_NUM + p];
            r.u(PW_ACTIVE) = s.is_active ? 1u : 0u;
            r.sf(PW_COOLDOWN, s.cooldown);
            r.u(PW_PREV_LOCKED) = s.prev_locked_car_id;
        }
    b->host_dirty = true;
    return 0;
}

int rsg_batch_set_controls(rsg_batch* b, const rsg_controls* in) {
    if (!b || !in) return fail("rsg_batch_set_controls: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int i = 0; i < b->cars_per_arena[a]; i++)
            store_ctrl(car_ref(h, a, i), CW_CTRL, ctrl_from_api(in[b->car_offset[a] + i]));
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_controls(rsg_batch* b, rsg_controls* out) {
    if (!b || !out) return fail("rsg_batch_get_controls: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int i = 0; i < b->cars_per_arena[a]; i++)
            out[b->car_offset[a] + i] = ctrl_to_api(load_ctrl(car_ref(h, a, i), CW_CTRL));
    return 0;
}

int rsg_batch_get_tick_counts(rsg_batch* b, uint64_t* out) {
    if (!b || !out) return fail("rsg_batch_get_tick_counts: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) out[a] = arena_ref(h, a).u64(AW_TICK);
    return 0;
}

int rsg_batch_set_tick_counts(rsg_batch* b, const uint64_t* in) {
    if (!b || !in) return fail("rsg_batch_set_tick_counts: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) arena_ref(h, a).su64(AW_TICK, in[a]);
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_goals(rsg_batch* b, uint32_t* blue, uint32_t* orange, uint64_t* last_goal_tick) {
    if (!b) return fail("null batch");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        CarRef r = arena_ref(h, a);
        if (blue) blue[a] = r.u(AW_GOALS_BLUE);
        if (orange) orange[a] = r.u(AW_GOALS_ORANGE);
        if (last_goal_tick) last_goal_tick[a] = r.u64(AW_LAST_GOAL_TICK);
    }
    return 0;
}

int rsg_batch_is_ball_scored(rsg_batch* b, uint8_t* out) {
    if (!b || !out) return fail("rsg_batch_is_ball_scored: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        float y = ball_ref(h, a).v(BW_POS).y * BT_TO_UU;
        out[a] = (b->dev.mode == RSG_MODE_SOCCAR &&
                  fabsf(y) > (SOCCAR_GOAL_SCORE_BASE_THRESHOLD_Y + BALL_COLLISION_RADIUS_SOCCAR)) ? 1 : 0;
    }
    return 0;
}

int rsg_init_mesh(const void* cmf_data, uint64_t size) {
    if (!cmf_data || size < 8) return fail("rsg_init_mesh: bad .cmf data");
    int32_t nt, nv;
    std::memcpy(&nt, cmf_data, 4);
    std::memcpy(&nv, (const char*)cmf_data + 4, 4);
    if (nt <= 0 || nv <= 0 || nt > 1000000 || nv > 1000000 || size < 8 + (uint64_t)nt * 12 + (uint64_t)nv * 12)
        return fail("rsg_init_mesh: bad .cmf triangle/vertex count");
    return 0;
}

}  // extern "C"
