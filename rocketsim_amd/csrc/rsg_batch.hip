// rsg_batch.hip -- the C ABI of include/rsg.h (host code).  The kernels live
// in rsg_k_*.hip and are enqueued through rsg_kernels.h.
//
// Kernel map:
//   rsg_step_arena_kernel       one lane per arena, K ticks per launch, arena
//                               state in private (scratch) memory
//   rsg_step_arena_lds_kernel   the same with the arena state in LDS
//   rsg_step_arena_coop_kernel  4 lanes per car slot share the tick's
//                               independent parts, arena state in LDS
//   rsg_ball_pred_kernel        ball prediction trackers (rsg_ball_pred.h)
//   rsg_air_forward_kernel,     differentiable air control: rollout with
//   rsg_air_backward_kernel     checkpoints, hand-written adjoint
//   rsg_shot_forward_kernel,    differentiable aerial shot (rsg_shot.h):
//   rsg_shot_backward_kernel    rollout with checkpoints; backward with one
//                               lane per tangent input of each pair
//   rsg_game_events_kernel      GameEventTracker (rsg_game_events.h)
//   rsg_reset_kernel            kickoff reset of one / all arenas
//
// Host side keeps a mirror of the SoA state for the get/set API; it is
// uploaded lazily before the next step and downloaded lazily on the next get.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rsg.h"
#include "rsg_mesh_build.h"
#include "rsg_tick.h"
#include "rsg_api_convert.h"
#include "rsg_adjoint.h"
#include "rsg_ball_pred.h"
#include "rsg_serialize.h"
#include "rsg_game_events.h"
#include "rsg_shot.h"
#include "rsg_kernels.h"

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

// One collision mesh's device arrays.
struct DeviceMesh {
    float* verts = nullptr;
    int32_t* tris = nullptr;
    QNode* nodes = nullptr;
    QSubtree* subtrees = nullptr;
    TriInfo* tri_info = nullptr;
    uint32_t* occ = nullptr;
};
// An arena's static geometry (the Soccar arena, the Hoops arena), shared
// read-only by every batch of the process (RocketSim keeps its arena meshes
// in process-wide storage the same way, src/RocketSim.cpp:50-53,92-185).
// Device arrays are allocated once per init.
struct DeviceArena {
    bool ready = false;
    HostArena host;
    std::vector<DeviceMesh> meshes;
    MeshView* views = nullptr;      // device copies of the meshes' views
    uint32_t* cell_statics = nullptr;
    ArenaGeom geom{};
};
DeviceArena g_arenas[2];  // by arena_kind(mode)
std::mutex g_mesh_mutex;

template <class T>
int upload_array(T** dst, const std::vector<T>& src) {
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    size_t bytes = src.size() * sizeof(T);
    if (bytes == 0) return 0;
    HIP_OK(hipMalloc((void**)dst, bytes));
    HIP_OK(hipMemcpy(*dst, src.data(), bytes, hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

namespace {
// Why a MutatorConfig cannot be simulated ("" if it can): non-finite
// fields, a non-positive ball radius, ball mass or car mass, an unknown
// demo mode.
std::string mutator_config_problem(const rsg_mutator_config& m) {
    const float fields[] = {m.gravity[0], m.gravity[1], m.gravity[2], m.car_mass, m.car_world_friction,
                            m.car_world_restitution, m.ball_mass, m.ball_max_speed, m.ball_drag,
                            m.ball_world_friction, m.ball_world_restitution, m.jump_accel, m.jump_immediate_force,
                            m.boost_accel_ground, m.boost_accel_air, m.boost_used_per_second, m.respawn_delay,
                            m.bump_cooldown_time, m.boost_pad_cooldown_big, m.boost_pad_cooldown_small,
                            m.car_spawn_boost_amount, m.ball_hit_extra_force_scale, m.bump_force_scale,
                            m.ball_radius, m.recharge_boost_per_second, m.recharge_boost_delay,
                            m.goal_base_threshold_y};
    for (float f : fields)
        if (!std::isfinite(f)) return "non-finite field";
    if (!(m.ball_radius > 0.f) || !(m.ball_mass > 0.f) || !(m.car_mass > 0.f))
        return "ball_radius, ball_mass and car_mass must be positive";
    if (m.demo_mode > RSG_DEMO_DISABLED) return "unknown demo_mode";
    return "";
}

}  // namespace

// ---------------------------------------------------------------------------
// batch object

struct rsg_batch {
    BatchDev dev;
    std::vector<int> cars_per_arena, car_offset;
    std::vector<uint8_t> team, preset;  // preset: each car's index in the config table
    int total_cars = 0;
    std::vector<uint32_t> h_car, h_ball, h_pad, h_arena;
    std::vector<rsg_mutator_config> h_mut;  // per-arena MutatorConfig (the kernels only read it)
    std::vector<rsg_car_config> h_cfgs;     // car config table: the presets, then other configs
    rsg_car_config* d_cfgs = nullptr;
    std::vector<rsg_boost_pad_config> h_custom_pads;  // ArenaConfig::customBoostPads
    rsg_boost_pad_config* d_custom_pads = nullptr;
    bool host_dirty = false;  // host mirror newer than device
    bool dev_newer = false;   // device newer than host mirror
    hipStream_t last_stream = nullptr;
    int arenas_per_block = 0;  // 0: derived from the device's SIMD count
    int kernel = RSG_KERNEL_AUTO;
    bool strict = false;       // fail a step that leaves an arena with a status bit set
    uint32_t* status_partials = nullptr;  // rsg_batch_status_summary's partial results
    // captured multi-launch step (rsg_batch_graph_*)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    int graph_ticks = 0;
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
    h.mut = b->h_mut.data();
    h.cfgs = b->h_cfgs.data();
    h.custom_pads = b->h_custom_pads.empty() ? nullptr : b->h_custom_pads.data();
    return h;
}

int upload(rsg_batch* b) {
    if (!b->host_dirty) return 0;
    HIP_OK(hipMemcpy(b->dev.car, b->h_car.data(), car_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev.ball, b->h_ball.data(), ball_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev.pad, b->h_pad.data(), pad_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev.arena, b->h_arena.data(), arena_words(b) * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(const_cast<rsg_mutator_config*>(b->dev.mut), b->h_mut.data(),
                     b->h_mut.size() * sizeof(rsg_mutator_config), hipMemcpyHostToDevice));
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

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* rsg_last_error(void) { return g_err.c_str(); }

rsg_batch* rsg_batch_create(const rsg_batch_desc* desc) {
    if (!desc || desc->n_arenas <= 0) { fail("rsg_batch_create: n_arenas must be > 0"); return nullptr; }
    if (desc->mode != RSG_MODE_SOCCAR && desc->mode != RSG_MODE_HOOPS && desc->mode != RSG_MODE_HEATSEEKER &&
        desc->mode != RSG_MODE_THE_VOID) {
        fail("rsg_batch_create: only SOCCAR, HOOPS, HEATSEEKER and THE_VOID are supported");
        return nullptr;
    }
    if (!(desc->tick_rate >= 15.f && desc->tick_rate <= 120.f)) {
        fail("rsg_batch_create: tick_rate must be within [15, 120] (src/Sim/Arena/Arena.cpp:419)");
        return nullptr;
    }
    if (!desc->cars_per_arena) { fail("rsg_batch_create: cars_per_arena is required"); return nullptr; }
    if (desc->mutators) {
        std::string why = mutator_config_problem(*desc->mutators);
        if (!why.empty() || !ball_fits_broadphase(desc->mutators->ball_radius)) {
            fail("rsg_batch_create: mutators: " + (why.empty() ? std::string("the ball's AABB exceeds the broadphase cell size") : why));
            return nullptr;
        }
    }
    if (mode_has_arena(desc->mode) && !g_arenas[arena_kind(desc->mode)].ready) {
        fail("rsg_batch_create: no arena collision mesh loaded (call rsg_init_mesh first)");
        return nullptr;
    }
    auto* b = new rsg_batch();
    b->dev.n = desc->n_arenas;
    b->dev.mode = desc->mode;
    b->dev.auto_reset = desc->auto_reset;
    b->dev.synth_controls = desc->synth_controls;
    b->dev.ball_rot = desc->ball_rotation != 0;
    if (desc->n_custom_pads != 0) {
        std::string why;
        if (!mode_has_arena(desc->mode)) why = "the mode has no boost pads";
        else if (!desc->custom_pads || desc->n_custom_pads < 0 || desc->n_custom_pads > RSG_NUM_PADS)
            why = "custom_pads must hold 1.." + std::to_string(RSG_NUM_PADS) + " pads";
        else
            for (int i = 0; i < desc->n_custom_pads && why.empty(); i++)
                for (float x : desc->custom_pads[i].pos)
                    if (!std::isfinite(x)) why = "non-finite custom pad position";
        if (!why.empty()) {
            delete b;
            fail("rsg_batch_create: " + why);
            return nullptr;
        }
        b->h_custom_pads.assign(desc->custom_pads, desc->custom_pads + desc->n_custom_pads);
        for (auto& c : b->h_custom_pads) c.is_big = c.is_big != 0, c._pad[0] = c._pad[1] = c._pad[2] = 0;
    }
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
    // the config table: the presets first, then each other config once
    for (int p = 0; p < RSG_PRESET_COUNT; p++) b->h_cfgs.push_back(preset_config(p));
    b->preset.resize((size_t)total);
    for (int i = 0; i < total; i++) {
        if (b->team[i] > 1) {
            delete b;
            fail("rsg_batch_create: bad team");
            return nullptr;
        }
        if (!desc->car_configs) {
            if (!desc->car_preset || desc->car_preset[i] >= RSG_PRESET_COUNT) {
                delete b;
                fail("rsg_batch_create: bad preset");
                return nullptr;
            }
            b->preset[(size_t)i] = desc->car_preset[i];
            continue;
        }
        const rsg_car_config& c = desc->car_configs[i];
        std::string why = car_config_problem(c);
        if (!why.empty()) {
            delete b;
            fail("rsg_batch_create: car " + std::to_string(i) + ": " + why);
            return nullptr;
        }
        size_t k = 0;
        while (k < b->h_cfgs.size() && !same_car_config(b->h_cfgs[k], c)) k++;
        if (k == b->h_cfgs.size()) {
            if (k >= 255) {
                delete b;
                fail("rsg_batch_create: more than 255 distinct car configs");
                return nullptr;
            }
            rsg_car_config n = c;
            std::memset(n._pad, 0, sizeof(n._pad));
            n.three_wheels = n.three_wheels != 0;
            b->h_cfgs.push_back(n);
        }
        b->preset[(size_t)i] = (uint8_t)k;
    }
    b->h_car.assign(car_words(b), 0);
    b->h_ball.assign(ball_words(b), 0);
    b->h_pad.assign(pad_words(b), 0);
    b->h_arena.assign(arena_words(b), 0);
    b->h_mut.assign((size_t)b->dev.n, desc->mutators ? *desc->mutators : mutator_defaults(b->dev.mode));
    hipError_t e = hipSuccess;
    rsg_mutator_config* dmut = nullptr;
    if ((e = hipMalloc(&b->dev.car, car_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.ball, ball_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.pad, pad_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.arena, arena_words(b) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->dev.space, (size_t)b->dev.n * sizeof(ArenaSpace))) != hipSuccess ||
        (e = hipMalloc(&dmut, (size_t)b->dev.n * sizeof(rsg_mutator_config))) != hipSuccess ||
        (e = hipMalloc(&b->d_cfgs, b->h_cfgs.size() * sizeof(rsg_car_config))) != hipSuccess ||
        (e = hipMemcpy(b->d_cfgs, b->h_cfgs.data(), b->h_cfgs.size() * sizeof(rsg_car_config),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (!b->h_custom_pads.empty() &&
         ((e = hipMalloc(&b->d_custom_pads, b->h_custom_pads.size() * sizeof(rsg_boost_pad_config))) != hipSuccess ||
          (e = hipMemcpy(b->d_custom_pads, b->h_custom_pads.data(),
                         b->h_custom_pads.size() * sizeof(rsg_boost_pad_config), hipMemcpyHostToDevice)) != hipSuccess))) {
        b->dev.mut = dmut;
        fail(std::string("rsg_batch_create: hipMalloc: ") + hipGetErrorString(e));
        rsg_batch_destroy(b);
        return nullptr;
    }
    b->dev.mut = dmut;
    b->dev.cfgs = b->d_cfgs;
    b->dev.custom_pads = b->d_custom_pads;
    b->dev.n_custom_pads = (int)b->h_custom_pads.size();
    // Host-side initial state: cars with config, ball at rest, pads active.
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        init_arena_record(h, a, b->cars_per_arena[a], b->team.data() + b->car_offset[a],
                          b->preset.data() + b->car_offset[a]);
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
    if (b->dev.space) (void)hipFree(b->dev.space);
    if (b->dev.mut) (void)hipFree(const_cast<rsg_mutator_config*>(b->dev.mut));
    if (b->d_cfgs) (void)hipFree(b->d_cfgs);
    if (b->d_custom_pads) (void)hipFree(b->d_custom_pads);
    if (b->status_partials) (void)hipFree(b->status_partials);
    if (b->dev.ev_buf) (void)hipFree(b->dev.ev_buf);
    if (b->dev.ev_count) (void)hipFree(b->dev.ev_count);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    delete b;
}

int rsg_batch_num_arenas(const rsg_batch* b) { return b ? b->dev.n : fail("null batch"); }
int rsg_batch_num_cars(const rsg_batch* b) { return b ? b->total_cars : fail("null batch"); }
int rsg_batch_car_offset(const rsg_batch* b, int arena) {
    if (!b || arena < 0 || arena >= b->dev.n) return fail("rsg_batch_car_offset: bad arena");
    return b->car_offset[arena];
}

// Launch geometry.  One lane simulates one arena and the kernel runs at one
// wave per SIMD (its register and private-memory footprint), so a launch is
// latency bound: spreading the arenas over every SIMD (partial waves) beats
// packing 64 lanes into few waves.
//
// Two kernels: the arena's working copy in private scratch, or in LDS.  LDS
// cuts the latency of every state access but holds only ~40 KB per wave at
// one wave per SIMD (160 KB per CU), i.e. a few arenas per wave; the LDS
// kernel is chosen when that still covers the batch in one round of waves.
namespace {
constexpr size_t LDS_PER_WAVE = 40 * 1024;

int device_simds() {
    static int simds = 0;
    if (simds == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        simds = 4 * cus;
    }
    return simds;
}

int lds_stride(const rsg_batch* b) {
    size_t bytes = offsetof(ArenaLocal, cars) + (size_t)b->dev.max_cars * sizeof(Car);
    return (int)((bytes + 15) & ~(size_t)15);
}

int lds_max_per_block(const rsg_batch* b) {
    int m = (int)(LDS_PER_WAVE / (size_t)lds_stride(b));
    return m < 1 ? 0 : (m > 64 ? 64 : m);
}

// Cooperative kernel: 4 lanes per car slot, arena + ray results + flag in LDS.
int coop_lanes(const rsg_batch* b) { return 4 * b->dev.max_cars; }
int coop_stride(const rsg_batch* b) {
    size_t bytes = (size_t)lds_stride(b) + (size_t)coop_lanes(b) * sizeof(RayHit) + 2 * sizeof(int);
    return (int)((bytes + 15) & ~(size_t)15);
}
int coop_max_per_block(const rsg_batch* b) {
    int by_lanes = 64 / coop_lanes(b);
    int by_lds = (int)(LDS_PER_WAVE / (size_t)coop_stride(b));
    int m = by_lanes < by_lds ? by_lanes : by_lds;
    return m < 1 ? 0 : m;
}

struct Launch {
    int kernel;  // RSG_KERNEL_SCRATCH, RSG_KERNEL_LDS or RSG_KERNEL_COOP
    int per_block;
};

// Kernel and arenas per workgroup.  The cooperative kernel is the fastest
// wherever its per-arena block fits (2v2 at 4096 arenas: 0.48 ms per tick
// against 0.62 for the LDS kernel; at 16384: 0.98 against 1.09 and 1.16 for
// the plain kernel), so it is the default; arenas are spread one per
// workgroup until every SIMD has a wave, then packed up to the most a
// workgroup holds.
Launch choose_launch(const rsg_batch* b) {
    const int simds = device_simds();
    int spread = (b->dev.n + simds - 1) / simds;  // arenas per wave at one wave per SIMD
    spread = spread < 1 ? 1 : (spread > 64 ? 64 : spread);
    Launch L;
    L.kernel = b->kernel;
    if (L.kernel == RSG_KERNEL_AUTO) {
        if (coop_max_per_block(b) >= 1) L.kernel = RSG_KERNEL_COOP;
        else if (lds_max_per_block(b) >= spread) L.kernel = RSG_KERNEL_LDS;
        else L.kernel = RSG_KERNEL_SCRATCH;
    }
    if (b->arenas_per_block > 0) {
        L.per_block = b->arenas_per_block;
    } else {
        L.per_block = spread;
    }
    if (L.kernel == RSG_KERNEL_LDS && L.per_block > lds_max_per_block(b)) L.per_block = lds_max_per_block(b);
    if (L.kernel == RSG_KERNEL_COOP && L.per_block > coop_max_per_block(b)) L.per_block = coop_max_per_block(b);
    if (L.per_block < 1) L.per_block = 1;
    return L;
}
}  // namespace

int rsg_batch_set_launch(rsg_batch* b, int arenas_per_block) {
    if (!b) return fail("null batch");
    if (arenas_per_block < 0 || arenas_per_block > 64)
        return fail("rsg_batch_set_launch: arenas_per_block must be 0 (auto) or 1..64");
    b->arenas_per_block = arenas_per_block;
    return 0;
}

int rsg_batch_get_launch(rsg_batch* b) { return b ? choose_launch(b).per_block : fail("null batch"); }

int rsg_batch_set_kernel(rsg_batch* b, int kernel) {
    if (!b) return fail("null batch");
    if (kernel != RSG_KERNEL_AUTO && kernel != RSG_KERNEL_SCRATCH && kernel != RSG_KERNEL_LDS &&
        kernel != RSG_KERNEL_COOP)
        return fail("rsg_batch_set_kernel: unknown kernel");
    if (kernel == RSG_KERNEL_LDS && lds_max_per_block(b) < 1)
        return fail("rsg_batch_set_kernel: an arena of this batch does not fit the LDS budget");
    if (kernel == RSG_KERNEL_COOP && coop_max_per_block(b) < 1)
        return fail("rsg_batch_set_kernel: an arena of this batch does not fit the cooperative kernel");
    b->kernel = kernel;
    return 0;
}

int rsg_batch_get_kernel(rsg_batch* b) { return b ? choose_launch(b).kernel : fail("null batch"); }

namespace {
// Enqueues `ticks` ticks of the batch state `dev` with launch L.
int launch_step(const rsg_batch* b, const BatchDev& dev, const Launch& L, int ticks, hipStream_t s) {
    int block = L.per_block;
    int grid = (dev.n + block - 1) / block;
    ArenaGeom mv{};
    if (mode_has_arena(dev.mode)) {
        const DeviceArena& ga = g_arenas[arena_kind(dev.mode)];
        if (!ga.ready) return fail("rsg_batch_step: no arena mesh (call rsg_init_mesh first)");
        mv = ga.geom;
    }
    if (L.kernel == RSG_KERNEL_COOP)
        HIP_OK(launch_step_coop(dev, mv, ticks, grid, block, coop_stride(b), lds_stride(b), coop_lanes(b), s));
    else if (L.kernel == RSG_KERNEL_LDS)
        HIP_OK(launch_step_lds(dev, mv, ticks, grid, block, lds_stride(b), s));
    else
        HIP_OK(launch_step_plain(dev, mv, ticks, grid, block, s));
    return 0;
}
}  // namespace

namespace {
// The per-arena status words (RSG_STATUS_*), from wherever they are newest.
int read_status(rsg_batch* b, std::vector<uint32_t>& st) {
    const size_t n = (size_t)b->dev.n;
    st.resize(n);
    if (b->dev_newer && !b->host_dirty) {
        if (b->last_stream) HIP_OK(hipStreamSynchronize(b->last_stream));
        HIP_OK(hipMemcpy(st.data(), b->dev.arena + (size_t)AW_STATUS * n, n * 4, hipMemcpyDeviceToHost));
    } else {
        std::memcpy(st.data(), b->h_arena.data() + (size_t)AW_STATUS * n, n * 4);
    }
    return 0;
}

const char* status_text(uint32_t s) {
    if (s & RSG_STATUS_NONFINITE) return "non-finite car or ball state";
    if (s & RSG_STATUS_CONTACT_OVERFLOW) return "more contact manifolds than RSG_MAX_MANIFOLDS";
    return "more contact or friction rows than RSG_MAX_SOLVER_ROWS";
}

// Strict mode: wait for the step and fail if an arena carries a status bit,
// as the reference throws on its capacity errors.
int check_strict(rsg_batch* b, const char* who) {
    if (!b->strict) return 0;
    std::vector<uint32_t> st;
    if (read_status(b, st) != 0) return -1;
    for (size_t a = 0; a < st.size(); a++)
        if (st[a]) return fail(std::string(who) + ": arena " + std::to_string(a) + ": " + status_text(st[a]));
    return 0;
}
}  // namespace

int rsg_batch_step(rsg_batch* b, int ticks, void* stream) {
    if (!b) return fail("null batch");
    if (ticks < 0) return fail("rsg_batch_step: ticks must be >= 0");
    if (ticks == 0) return 0;
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (launch_step(b, b->dev, choose_launch(b), ticks, s) != 0) return -1;
    b->last_stream = s;
    b->dev_newer = true;
    return check_strict(b, "rsg_batch_step");
}

int rsg_batch_get_status(rsg_batch* b, uint32_t* per_arena, uint32_t* any) {
    if (!b) return fail("null batch");
    std::vector<uint32_t> st;
    if (read_status(b, st) != 0) return -1;
    uint32_t all = 0;
    int flagged = 0;
    for (uint32_t s : st) {
        all |= s;
        flagged += s ? 1 : 0;
    }
    if (per_arena) std::memcpy(per_arena, st.data(), st.size() * 4);
    if (any) *any = all;
    return flagged;
}

int rsg_batch_status_summary(rsg_batch* b, uint32_t* out, void* stream) {
    if (!b || !out) return fail("rsg_batch_status_summary: null argument");
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (b->last_stream && b->last_stream != s) HIP_OK(hipStreamSynchronize(b->last_stream));
    if (!b->status_partials)
        HIP_OK(hipMalloc(&b->status_partials, (size_t)(2 * status_blocks(b->dev.n) + 2) * sizeof(uint32_t)));
    HIP_OK(launch_status_summary(b->dev.arena + (size_t)AW_STATUS * b->dev.n, b->dev.n, b->status_partials, out, s));
    return 0;
}

// Partial results of rsg_batch_status_summary_noalloc, for batches of up to
// RSG_STATUS_NOALLOC_MAX_ARENAS arenas (2 words per 256 arenas: 4 KiB).
// Each GPU holds its own copy until the process ends; host code reaches it
// only through the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) uint32_t scratch[2 * (RSG_STATUS_NOALLOC_MAX_ARENAS / STATUS_BLOCK)];

namespace {
std::mutex g_scratch_mutex;
uint32_t* g_scratch_ptr[64] = {};  // per device

// The address of `scratch` on the current device, looked up once per device.
int scratch_on_device(uint32_t** out) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail("rsg_batch_status_summary_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

int rsg_batch_status_summary_noalloc(rsg_batch* b, uint32_t* out, void* stream) {
    if (!b || !out) return fail("rsg_batch_status_summary_noalloc: null argument");
    if (b->dev.n > RSG_STATUS_NOALLOC_MAX_ARENAS)
        return fail("rsg_batch_status_summary_noalloc: " + std::to_string(b->dev.n) +
                    " arenas exceed RSG_STATUS_NOALLOC_MAX_ARENAS (" + std::to_string(RSG_STATUS_NOALLOC_MAX_ARENAS) +
                    "); use rsg_batch_status_summary");
    uint32_t* partials = nullptr;
    if (scratch_on_device(&partials) != 0) return -1;
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (b->last_stream && b->last_stream != s) HIP_OK(hipStreamSynchronize(b->last_stream));
    // every partial slot the final kernel reads is written by this call's first kernel
    HIP_OK(launch_status_summary(b->dev.arena + (size_t)AW_STATUS * b->dev.n, b->dev.n, partials, out, s));
    return 0;
}

int rsg_batch_clear_status(rsg_batch* b) {
    if (!b) return fail("null batch");
    if (download(b) != 0) return -1;
    std::fill(b->h_arena.begin() + (size_t)AW_STATUS * b->dev.n, b->h_arena.begin() + (size_t)(AW_STATUS + 1) * b->dev.n,
              0u);
    b->host_dirty = true;
    return 0;
}

int rsg_batch_set_strict(rsg_batch* b, int strict) {
    if (!b) return fail("null batch");
    b->strict = strict != 0;
    return 0;
}

// hipGraph of `launches` step launches of `ticks_per_launch` ticks each, with
// the batch's current kernel and launch shape, captured from a private stream.
int rsg_batch_graph_create(rsg_batch* b, int ticks_per_launch, int launches) {
    if (!b) return fail("null batch");
    if (ticks_per_launch <= 0 || launches <= 0) return fail("rsg_batch_graph_create: counts must be > 0");
    if (upload(b) != 0) return -1;
    if (b->graph_exec) { (void)hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    if (!b->capture_stream) HIP_OK(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    // the step reads device state written by earlier work on other streams
    if (b->last_stream) HIP_OK(hipStreamSynchronize(b->last_stream));
    HIP_OK(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    Launch L = choose_launch(b);
    int rc = 0;
    for (int i = 0; i < launches && rc == 0; i++) rc = launch_step(b, b->dev, L, ticks_per_launch, b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc != 0) { if (g) (void)hipGraphDestroy(g); return -1; }
    if (e != hipSuccess) return fail(std::string("rsg_batch_graph_create: ") + hipGetErrorString(e));
    b->graph = g;
    HIP_OK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    b->graph_ticks = ticks_per_launch * launches;
    return 0;
}

// Replays the captured graph on `stream` (advances every arena by the
// graph's tick count).
int rsg_batch_graph_launch(rsg_batch* b, void* stream) {
    if (!b) return fail("null batch");
    if (!b->graph_exec) return fail("rsg_batch_graph_launch: no graph (call rsg_batch_graph_create)");
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    HIP_OK(hipGraphLaunch(b->graph_exec, s));
    b->last_stream = s;
    b->dev_newer = true;
    return check_strict(b, "rsg_batch_graph_launch");
}

int rsg_batch_graph_ticks(rsg_batch* b) { return b ? b->graph_ticks : fail("null batch"); }

int rsg_air_rollout(int n, int ticks, int ckpt_every, float tick_rate, const float* state0, const float* controls,
                    const float* target, float* state_out, float* loss, float* checkpoints, void* stream) {
    if (n < 0 || ticks <= 0) return fail("rsg_air_rollout: need n >= 0 and ticks > 0");
    if (ckpt_every < 1 || ckpt_every > AIR_MAX_K) return fail("rsg_air_rollout: ckpt_every must be 1..32");
    if (tick_rate <= 0.f) return fail("rsg_air_rollout: tick_rate must be > 0");
    if (!state0 || !controls || (loss && !target)) return fail("rsg_air_rollout: null argument");
    if (n == 0) return 0;
    HIP_OK(launch_air_forward(n, ticks, ckpt_every, 1.f / tick_rate, state0, controls, target, state_out, loss,
                              checkpoints, (hipStream_t)stream));
    return 0;
}

int rsg_air_rollout_backward(int n, int ticks, int ckpt_every, float tick_rate, const float* controls,
                             const float* target, const float* checkpoints, const float* dloss, float* dcontrols,
                             float* dstate0, void* stream) {
    if (n < 0 || ticks <= 0) return fail("rsg_air_rollout_backward: need n >= 0 and ticks > 0");
    if (ckpt_every < 1 || ckpt_every > AIR_MAX_K) return fail("rsg_air_rollout_backward: ckpt_every must be 1..32");
    if (tick_rate <= 0.f) return fail("rsg_air_rollout_backward: tick_rate must be > 0");
    if (!controls || !target || !checkpoints || !dloss || !dcontrols)
        return fail("rsg_air_rollout_backward: null argument");
    if (n == 0) return 0;
    HIP_OK(launch_air_backward(n, ticks, ckpt_every, 1.f / tick_rate, controls, target, checkpoints, dloss,
                               dcontrols, dstate0, (hipStream_t)stream));
    return 0;
}

int rsg_shot_rollout(int n, int ticks, int ckpt_every, float tick_rate, const float* state0, const float* controls,
                     const float* target, float* state_out, float* loss, float* checkpoints, uint32_t* flags,
                     void* stream) {
    if (n <= 0 || ticks <= 0) return fail("rsg_shot_rollout: n and ticks must be > 0");
    if (ckpt_every < 1 || ckpt_every > SHOT_MAX_K) return fail("rsg_shot_rollout: ckpt_every must be within [1, 32]");
    if (!(tick_rate >= 15.f && tick_rate <= 120.f)) return fail("rsg_shot_rollout: tick_rate must be within [15, 120]");
    if (!state0 || !controls || !target || !checkpoints) return fail("rsg_shot_rollout: null argument");
    HIP_OK(launch_shot_forward(n, ticks, ckpt_every, 1.f / tick_rate, state0, controls, target, state_out, loss,
                               checkpoints, flags, (hipStream_t)stream));
    return 0;
}

int rsg_shot_rollout_backward(int n, int ticks, int ckpt_every, float tick_rate, const float* controls,
                              const float* target, const float* checkpoints, const float* dloss, float* dcontrols,
                              float* dstate0, void* stream) {
    if (n <= 0 || ticks <= 0) return fail("rsg_shot_rollout_backward: n and ticks must be > 0");
    if (ckpt_every < 1 || ckpt_every > SHOT_MAX_K)
        return fail("rsg_shot_rollout_backward: ckpt_every must be within [1, 32]");
    if (!(tick_rate >= 15.f && tick_rate <= 120.f))
        return fail("rsg_shot_rollout_backward: tick_rate must be within [15, 120]");
    if (!controls || !target || !checkpoints || !dloss || !dcontrols)
        return fail("rsg_shot_rollout_backward: null argument");
    HIP_OK(launch_shot_backward(n, ticks, ckpt_every, 1.f / tick_rate, controls, target, checkpoints, dloss,
                                dcontrols, dstate0, (hipStream_t)stream));
    return 0;
}

// Autotuner: every feasible (kernel, arenas per workgroup) is timed on a
// device copy of the batch state for `ticks` ticks after one warm-up launch,
// and the fastest is kept.  The batch's own state is not advanced.
int rsg_batch_autotune(rsg_batch* b, int ticks, void* stream, int* out_kernel, int* out_per_block, float* out_ms) {
    if (!b) return fail("null batch");
    if (ticks <= 0) return fail("rsg_batch_autotune: ticks must be > 0");
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    BatchDev tmp = b->dev;
    tmp.car = tmp.ball = tmp.pad = tmp.arena = nullptr;
    size_t cw = car_words(b) * 4, bw = ball_words(b) * 4, pw = pad_words(b) * 4, aw = arena_words(b) * 4;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (tmp.car) (void)hipFree(tmp.car);
        if (tmp.ball) (void)hipFree(tmp.ball);
        if (tmp.pad) (void)hipFree(tmp.pad);
        if (tmp.arena) (void)hipFree(tmp.arena);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    hipError_t e;
    if ((e = hipMalloc(&tmp.car, cw ? cw : 4)) != hipSuccess || (e = hipMalloc(&tmp.ball, bw)) != hipSuccess ||
        (e = hipMalloc(&tmp.pad, pw ? pw : 4)) != hipSuccess || (e = hipMalloc(&tmp.arena, aw)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        cleanup();
        return fail(std::string("rsg_batch_autotune: ") + hipGetErrorString(e));
    }
    struct Cand { Launch L; float ms; };
    std::vector<Launch> cands;
    const int shapes[] = {1, 2, 4, 8, 16, 32, 64};
    for (int kernel : {RSG_KERNEL_COOP, RSG_KERNEL_LDS, RSG_KERNEL_SCRATCH}) {
        int maxpb = kernel == RSG_KERNEL_COOP ? coop_max_per_block(b)
                  : kernel == RSG_KERNEL_LDS ? lds_max_per_block(b) : 64;
        for (int pb : shapes) {
            if (pb > maxpb) continue;
            if ((b->dev.n + pb - 1) / pb < 1) continue;
            Launch L;
            L.kernel = kernel;
            L.per_block = pb;
            cands.push_back(L);
        }
    }
    Launch best = choose_launch(b);
    float best_ms = -1.f;
    for (const Launch& L : cands) {
        // fresh copy of the state for every candidate
        if ((e = hipMemcpyAsync(tmp.car, b->dev.car, cw, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(tmp.ball, b->dev.ball, bw, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(tmp.pad, b->dev.pad, pw, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(tmp.arena, b->dev.arena, aw, hipMemcpyDeviceToDevice, s)) != hipSuccess) {
            cleanup();
            return fail(std::string("rsg_batch_autotune: ") + hipGetErrorString(e));
        }
        if (launch_step(b, tmp, L, 1, s) != 0) { cleanup(); return -1; }  // warm-up
        (void)hipEventRecord(e0, s);
        if (launch_step(b, tmp, L, ticks, s) != 0) { cleanup(); return -1; }
        (void)hipEventRecord(e1, s);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) {
            cleanup();
            return fail(std::string("rsg_batch_autotune: ") + hipGetErrorString(e));
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (best_ms < 0.f || ms < best_ms) { best_ms = ms; best = L; }
    }
    cleanup();
    b->kernel = best.kernel;
    b->arenas_per_block = best.per_block;
    if (out_kernel) *out_kernel = best.kernel;
    if (out_per_block) *out_per_block = best.per_block;
    if (out_ms) *out_ms = best_ms / (float)ticks;
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
    HIP_OK(launch_reset(b->dev, arena, seed, b->last_stream));
    b->dev_newer = true;
    return 0;
}

int rsg_batch_get_cars(rsg_batch* b, rsg_car_state* out) {
    if (!b || (!out && b->total_cars > 0)) return fail("rsg_batch_get_cars: null argument");
    if (b->total_cars == 0) return 0;
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int i = 0; i < b->cars_per_arena[a]; i++)
            car_to_api(car_ref(h, a, i), out[b->car_offset[a] + i]);
    return 0;
}

int rsg_batch_set_cars(rsg_batch* b, const rsg_car_state* in) {
    if (!b || (!in && b->total_cars > 0)) return fail("rsg_batch_set_cars: null argument");
    if (b->total_cars == 0) return 0;
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        for (int i = 0; i < b->cars_per_arena[a]; i++) {
            car_from_api(car_ref(h, a, i), in[b->car_offset[a] + i]);
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
        ball_to_api(ball_ref(h, a), out[a]);
    }
    return 0;
}

int rsg_batch_set_balls(rsg_batch* b, const rsg_ball_state* in) {
    if (!b || !in) return fail("rsg_batch_set_balls: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        ball_from_api(ball_ref(h, a), in[a]);
    }
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_pads(rsg_batch* b, rsg_pad_state* out) {
    if (!b || !out) return fail("rsg_batch_get_pads: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int p = 0; p < PAD_NUM; p++) pad_to_api(pad_ref(h, a, p), out[a * PAD_NUM + p]);
    return 0;
}

int rsg_batch_set_pads(rsg_batch* b, const rsg_pad_state* in) {
    if (!b || !in) return fail("rsg_batch_set_pads: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int p = 0; p < PAD_NUM; p++) pad_from_api(pad_ref(h, a, p), in[a * PAD_NUM + p]);
    b->host_dirty = true;
    return 0;
}

int rsg_batch_set_controls(rsg_batch* b, const rsg_controls* in) {
    if (!b || (!in && b->total_cars > 0)) return fail("rsg_batch_set_controls: null argument");
    if (b->total_cars == 0) return 0;
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int i = 0; i < b->cars_per_arena[a]; i++)
            store_ctrl(car_ref(h, a, i), CW_CTRL, ctrl_from_api(in[b->car_offset[a] + i]));
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_controls(rsg_batch* b, rsg_controls* out) {
    if (!b || (!out && b->total_cars > 0)) return fail("rsg_batch_get_controls: null argument");
    if (b->total_cars == 0) return 0;
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++)
        for (int i = 0; i < b->cars_per_arena[a]; i++)
            out[b->car_offset[a] + i] = ctrl_to_api(load_ctrl(car_ref(h, a, i), CW_CTRL));
    return 0;
}

int rsg_batch_set_car_order(rsg_batch* b, const int32_t* order) {
    if (!b || !order) return fail("rsg_batch_set_car_order: null argument");
    for (int a = 0; a < b->dev.n; a++) {
        const int n = b->cars_per_arena[a];
        uint32_t seen = 0;
        for (int k = 0; k < n; k++) {
            int32_t v = order[b->car_offset[a] + k];
            if (v < 0 || v >= n || (seen >> v & 1u))
                return fail("rsg_batch_set_car_order: arena " + std::to_string(a) +
                            " order is not a permutation of its car indices");
            seen |= 1u << v;
        }
    }
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        uint32_t ow = 0;
        for (int k = 0; k < b->cars_per_arena[a]; k++)
            ow |= (uint32_t)order[b->car_offset[a] + k] << (4 * k);
        arena_ref(h, a).u(AW_CAR_ORDER) = ow;
    }
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_car_order(rsg_batch* b, int32_t* order) {
    if (!b || !order) return fail("rsg_batch_get_car_order: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        uint32_t ow = arena_ref(h, a).u(AW_CAR_ORDER);
        for (int k = 0; k < b->cars_per_arena[a]; k++)
            order[b->car_offset[a] + k] = (int32_t)((ow >> (4 * k)) & 0xFu);
    }
    return 0;
}

int rsg_batch_enable_events(rsg_batch* b, int capacity) {
    if (!b) return fail("null batch");
    if (capacity < 0) return fail("rsg_batch_enable_events: capacity must be >= 0");
    if (b->last_stream) HIP_OK(hipStreamSynchronize(b->last_stream));
    if (b->dev.ev_buf) { (void)hipFree(b->dev.ev_buf); b->dev.ev_buf = nullptr; }
    if (b->dev.ev_count) { (void)hipFree(b->dev.ev_count); b->dev.ev_count = nullptr; }
    b->dev.ev_cap = 0;
    if (capacity == 0) return 0;
    HIP_OK(hipMalloc((void**)&b->dev.ev_buf, (size_t)capacity * sizeof(rsg_event)));
    HIP_OK(hipMalloc((void**)&b->dev.ev_count, sizeof(uint32_t)));
    uint32_t zero = 0;
    HIP_OK(hipMemcpy(b->dev.ev_count, &zero, sizeof(zero), hipMemcpyHostToDevice));
    b->dev.ev_cap = (uint32_t)capacity;
    return 0;
}

int rsg_batch_poll_events(rsg_batch* b, rsg_event* out, int max_out, int* n_lost) {
    if (!b) return fail("null batch");
    if (n_lost) *n_lost = 0;
    if (!b->dev.ev_buf) return 0;
    if (b->last_stream) HIP_OK(hipStreamSynchronize(b->last_stream));
    uint32_t count = 0;
    HIP_OK(hipMemcpy(&count, b->dev.ev_count, sizeof(count), hipMemcpyDeviceToHost));
    uint32_t stored = count < b->dev.ev_cap ? count : b->dev.ev_cap;
    std::vector<rsg_event> ev(stored);
    if (stored) HIP_OK(hipMemcpy(ev.data(), b->dev.ev_buf, stored * sizeof(rsg_event), hipMemcpyDeviceToHost));
    uint32_t zero = 0;
    HIP_OK(hipMemcpy(b->dev.ev_count, &zero, sizeof(zero), hipMemcpyHostToDevice));
    std::sort(ev.begin(), ev.end(), [](const rsg_event& x, const rsg_event& y) {
        return x.arena != y.arena ? x.arena < y.arena : x.seq < y.seq;
    });
    int n = (int)stored < max_out ? (int)stored : (max_out > 0 ? max_out : 0);
    if (out && n) std::memcpy(out, ev.data(), (size_t)n * sizeof(rsg_event));
    if (n_lost) *n_lost = (int)(count - (uint32_t)n);
    return n;
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

int rsg_car_config_preset(int preset, rsg_car_config* out) {
    if (!out) return fail("rsg_car_config_preset: null argument");
    if (preset < 0 || preset >= RSG_PRESET_COUNT) return fail("rsg_car_config_preset: unknown preset");
    *out = preset_config(preset);
    return 0;
}

int rsg_mutator_config_default(int mode, rsg_mutator_config* out) {
    if (!out) return fail("rsg_mutator_config_default: null argument");
    if (mode != RSG_MODE_SOCCAR && mode != RSG_MODE_HOOPS && mode != RSG_MODE_HEATSEEKER && mode != RSG_MODE_THE_VOID)
        return fail("rsg_mutator_config_default: unsupported game mode");
    *out = mutator_defaults(mode);
    return 0;
}

int rsg_batch_set_mutator_config(rsg_batch* b, int arena, const rsg_mutator_config* cfg) {
    if (!b || !cfg) return fail("rsg_batch_set_mutator_config: null argument");
    if (arena < -1 || arena >= b->dev.n) return fail("rsg_batch_set_mutator_config: arena out of range");
    const rsg_mutator_config& m = *cfg;
    std::string why = mutator_config_problem(m);
    if (!why.empty()) return fail("rsg_batch_set_mutator_config: " + why);
    const int a0 = arena < 0 ? 0 : arena, a1 = arena < 0 ? b->dev.n : arena + 1;
    for (int a = a0; a < a1; a++) {
        const rsg_mutator_config& old = b->h_mut[(size_t)a];
        if ((m.ball_radius != old.ball_radius || m.ball_mass != old.ball_mass) && !ball_fits_broadphase(m.ball_radius))
            return fail("rsg_batch_set_mutator_config: the remade ball's AABB exceeds the broadphase cell size "
                        "(btRSBroadphase::createProxy refuses it)");
    }
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = a0; a < a1; a++) {
        apply_mutator_config(h, a, b->h_mut[(size_t)a], m);
        b->h_mut[(size_t)a] = m;
    }
    b->host_dirty = true;
    return 0;
}

int rsg_batch_get_mutator_config(rsg_batch* b, int arena, rsg_mutator_config* out) {
    if (!b || !out) return fail("rsg_batch_get_mutator_config: null argument");
    if (arena < 0 || arena >= b->dev.n) return fail("rsg_batch_get_mutator_config: arena out of range");
    *out = b->h_mut[(size_t)arena];
    return 0;
}

int rsg_batch_is_ball_scored(rsg_batch* b, uint8_t* out) {
    if (!b || !out) return fail("rsg_batch_is_ball_scored: null argument");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    for (int a = 0; a < b->dev.n; a++) {
        out[a] = ball_scored(b->dev.mode, ball_ref(h, a).v(BW_POS), b->h_mut[(size_t)a]) ? 1 : 0;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Game-event tracking

}  // extern "C"

struct rsg_game_events {
    int n = 0;
    rsg_game_event_config cfg{};
    float* cooldown = nullptr;
    uint32_t* flags = nullptr;
    uint64_t* last_count = nullptr;
    uint32_t* seq = nullptr;
    rsg_event* ev_buf = nullptr;
    uint32_t* ev_count = nullptr;
    uint32_t ev_cap = 0;
    hipStream_t last_stream = nullptr;
};

extern "C" {

void rsg_game_event_config_default(rsg_game_event_config* out) {
    if (!out) return;
    out->shot_min_speed = 1750.f;
    out->shot_touch_min_delay = 0.3f;
    out->pred_score_extra_margin = 0.f;
    out->shot_event_cooldown = 1.0f;
    out->shot_min_score_time = 2.0f;
    out->goal_max_touch_time = 4.0f;
    out->pass_max_touch_time = 2.0f;
    out->auto_state_set_detection = 1;
}

rsg_game_events* rsg_game_events_create(rsg_batch* b, const rsg_game_event_config* cfg, int capacity) {
    if (!b) { fail("rsg_game_events_create: null batch"); return nullptr; }
    if (b->dev.mode != RSG_MODE_SOCCAR && b->dev.mode != RSG_MODE_HOOPS) {
        fail("rsg_game_events_create: SOCCAR and HOOPS batches only");
        return nullptr;
    }
    if (capacity < 1) { fail("rsg_game_events_create: capacity must be >= 1"); return nullptr; }
    auto* t = new rsg_game_events();
    t->n = b->dev.n;
    if (cfg) t->cfg = *cfg;
    else rsg_game_event_config_default(&t->cfg);
    t->ev_cap = (uint32_t)capacity;
    const size_t n = (size_t)t->n;
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&t->cooldown, n * 4)) != hipSuccess || (e = hipMalloc(&t->flags, n * 4)) != hipSuccess ||
        (e = hipMalloc(&t->last_count, n * 8)) != hipSuccess || (e = hipMalloc(&t->seq, n * 4)) != hipSuccess ||
        (e = hipMalloc(&t->ev_buf, (size_t)capacity * sizeof(rsg_event))) != hipSuccess ||
        (e = hipMalloc(&t->ev_count, 4)) != hipSuccess || (e = hipMemset(t->cooldown, 0, n * 4)) != hipSuccess ||
        (e = hipMemset(t->flags, 0, n * 4)) != hipSuccess || (e = hipMemset(t->last_count, 0, n * 8)) != hipSuccess ||
        (e = hipMemset(t->seq, 0, n * 4)) != hipSuccess || (e = hipMemset(t->ev_count, 0, 4)) != hipSuccess) {
        fail(std::string("rsg_game_events_create: ") + hipGetErrorString(e));
        rsg_game_events_destroy(t);
        return nullptr;
    }
    return t;
}

void rsg_game_events_destroy(rsg_game_events* t) {
    if (!t) return;
    (void)hipDeviceSynchronize();
    if (t->cooldown) (void)hipFree(t->cooldown);
    if (t->flags) (void)hipFree(t->flags);
    if (t->last_count) (void)hipFree(t->last_count);
    if (t->seq) (void)hipFree(t->seq);
    if (t->ev_buf) (void)hipFree(t->ev_buf);
    if (t->ev_count) (void)hipFree(t->ev_count);
    delete t;
}

int rsg_game_events_update(rsg_game_events* t, rsg_batch* b, void* stream) {
    if (!t || !b) return fail("rsg_game_events_update: null argument");
    if (b->dev.n != t->n || (b->dev.mode != RSG_MODE_SOCCAR && b->dev.mode != RSG_MODE_HOOPS))
        return fail("rsg_game_events_update: not the batch the tracker was made for");
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (b->last_stream && b->last_stream != s) HIP_OK(hipStreamSynchronize(b->last_stream));
    GameEventDev g;
    g.cooldown = t->cooldown;
    g.flags = t->flags;
    g.last_count = t->last_count;
    g.seq = t->seq;
    g.ev_buf = t->ev_buf;
    g.ev_count = t->ev_count;
    g.ev_cap = t->ev_cap;
    HIP_OK(launch_game_events(b->dev, g, t->cfg, s));
    t->last_stream = s;
    return 0;
}

int rsg_game_events_poll(rsg_game_events* t, rsg_event* out, int max_out, int* n_lost) {
    if (!t) return fail("rsg_game_events_poll: null tracker");
    if (n_lost) *n_lost = 0;
    if (t->last_stream) HIP_OK(hipStreamSynchronize(t->last_stream));
    else HIP_OK(hipDeviceSynchronize());
    uint32_t count = 0;
    HIP_OK(hipMemcpy(&count, t->ev_count, sizeof(count), hipMemcpyDeviceToHost));
    uint32_t stored = count < t->ev_cap ? count : t->ev_cap;
    std::vector<rsg_event> ev(stored);
    if (stored) HIP_OK(hipMemcpy(ev.data(), t->ev_buf, stored * sizeof(rsg_event), hipMemcpyDeviceToHost));
    uint32_t zero = 0;
    HIP_OK(hipMemcpy(t->ev_count, &zero, sizeof(zero), hipMemcpyHostToDevice));
    std::sort(ev.begin(), ev.end(), [](const rsg_event& x, const rsg_event& y) {
        return x.arena != y.arena ? x.arena < y.arena : x.seq < y.seq;
    });
    int n = (int)stored < max_out ? (int)stored : (max_out > 0 ? max_out : 0);
    if (out && n) std::memcpy(out, ev.data(), (size_t)n * sizeof(rsg_event));
    if (n_lost) *n_lost = (int)(count - (uint32_t)n);
    return n;
}

int rsg_game_events_reset(rsg_game_events* t, int arena) {
    if (!t) return fail("rsg_game_events_reset: null tracker");
    if (arena < -1 || arena >= t->n) return fail("rsg_game_events_reset: arena out of range");
    if (t->last_stream) HIP_OK(hipStreamSynchronize(t->last_stream));
    int a0 = arena < 0 ? 0 : arena, a1 = arena < 0 ? t->n : arena + 1;
    // ResetPersistentInfo: cooldown 0, not shot, not scored last (the shot's
    // goal team is kept)
    std::vector<uint32_t> fl((size_t)(a1 - a0));
    HIP_OK(hipMemcpy(fl.data(), t->flags + a0, fl.size() * 4, hipMemcpyDeviceToHost));
    for (auto& f : fl) f &= GE_SHOT_GOAL_ORANGE;
    std::vector<float> zero((size_t)(a1 - a0), 0.f);
    HIP_OK(hipMemcpy(t->flags + a0, fl.data(), fl.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(t->cooldown + a0, zero.data(), zero.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

// ---------------------------------------------------------------------------
// Snapshots (Arena::Serialize / DeserializeNew)

int64_t rsg_snapshot_serialize(const rsg_arena_snapshot* s, uint8_t* out, uint64_t cap) {
    if (!s) return fail("rsg_snapshot_serialize: null snapshot");
    if (s->raw_words > RSG_SNAPSHOT_RAW_WORDS) return fail("rsg_snapshot_serialize: raw_words out of range");
    std::vector<uint8_t> bytes;
    std::string err = snapshot_serialize(*s, bytes);
    if (!err.empty()) return fail("rsg_snapshot_serialize: " + err);
    if (out && cap >= bytes.size()) std::memcpy(out, bytes.data(), bytes.size());
    return (int64_t)bytes.size();
}

int rsg_snapshot_deserialize(const uint8_t* data, uint64_t size, rsg_arena_snapshot* out) {
    if (!data || !out) return fail("rsg_snapshot_deserialize: null argument");
    std::string err = snapshot_deserialize(data, (size_t)size, *out);
    if (!err.empty()) return fail("rsg_snapshot_deserialize: " + err);
    return 0;
}

int rsg_batch_get_snapshot(rsg_batch* b, int arena, rsg_arena_snapshot* out) {
    if (!b || !out) return fail("rsg_batch_get_snapshot: null argument");
    if (arena < 0 || arena >= b->dev.n) return fail("rsg_batch_get_snapshot: arena out of range");
    if (b->dev.n_custom_pads) return fail("rsg_batch_get_snapshot: batches with custom boost pads have no snapshots");
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    std::memset(out, 0, sizeof(*out));
    CarRef ar = arena_ref(h, arena);
    out->mode = b->dev.mode;
    out->tick_time = b->dev.tick_time;
    out->tick_count = ar.u64(AW_TICK);
    const int n = b->cars_per_arena[arena];
    out->n_cars = n;
    uint32_t ow = ar.u(AW_CAR_ORDER), last_id = 0;
    for (int k = 0; k < n; k++) {
        int slot = (int)((ow >> (4 * k)) & 0xFu);
        CarRef cr = car_ref(h, arena, slot);
        uint32_t cfg = cr.u(CW_CONFIG);
        out->car_id[k] = cr.u(CW_ID);
        out->car_config[k] = b->h_cfgs[cfg & 0xFF];
        out->car_preset[k] = (uint8_t)preset_of(out->car_config[k]);
        out->car_team[k] = (uint8_t)((cfg >> 8) & 0xFF);
        out->controls[k] = ctrl_to_api(load_ctrl(cr, CW_CTRL));
        car_to_api(cr, out->cars[k]);
        last_id = out->car_id[k] > last_id ? out->car_id[k] : last_id;
    }
    out->last_car_id = last_id;
    out->mutators = b->h_mut[(size_t)arena];
    ball_to_api(ball_ref(h, arena), out->ball);
    if (mode_has_arena(b->dev.mode))
        for (int p = 0; p < PAD_NUM; p++) pad_to_api(pad_ref(h, arena, p), out->pads[p]);
    // raw records, cars in snapshot order
    BatchDev v = raw_view(out->raw, n);
    for (int k = 0; k < n; k++) {
        CarRef src = car_ref(h, arena, (int)((ow >> (4 * k)) & 0xFu)), dst = car_ref(v, 0, k);
        for (int w = 0; w < CAR_WORDS; w++) dst.u(w) = src.u(w);
    }
    {
        CarRef src = ball_ref(h, arena), dst = ball_ref(v, 0);
        for (int w = 0; w < BALL_WORDS; w++) dst.u(w) = src.u(w);
        CarRef sa = arena_ref(h, arena), da = arena_ref(v, 0);
        for (int w = 0; w < ARENA_WORDS; w++) da.u(w) = sa.u(w);
        uint32_t ident = 0;
        for (int k = 0; k < n; k++) ident |= (uint32_t)k << (4 * k);
        da.u(AW_CAR_ORDER) = ident;
    }
    out->raw_layout = raw_layout_tag();
    out->raw_words = raw_words_for(n);
    return 0;
}

int rsg_batch_set_snapshot(rsg_batch* b, int arena, const rsg_arena_snapshot* s) {
    if (!b || !s) return fail("rsg_batch_set_snapshot: null argument");
    if (arena < 0 || arena >= b->dev.n) return fail("rsg_batch_set_snapshot: arena out of range");
    if (b->dev.n_custom_pads) return fail("rsg_batch_set_snapshot: batches with custom boost pads have no snapshots");
    if (s->mode != b->dev.mode) return fail("rsg_batch_set_snapshot: game mode differs from the batch's");
    if (s->tick_time != b->dev.tick_time) return fail("rsg_batch_set_snapshot: tick time differs from the batch's");
    const int n = b->cars_per_arena[arena];
    if (s->n_cars != n) return fail("rsg_batch_set_snapshot: car count differs from the arena slot's");
    if (s->mutators.demo_mode > RSG_DEMO_DISABLED || !(s->mutators.ball_radius > 0.f) ||
        !(s->mutators.ball_mass > 0.f) || !(s->mutators.car_mass > 0.f))
        return fail("rsg_batch_set_snapshot: mutator config out of range");
    // Cars go to the slots in creation (ascending id) order; each must have
    // the slot's team and preset.
    int by_id[RSG_MAX_CARS], slot_of[RSG_MAX_CARS];
    for (int k = 0; k < n; k++) by_id[k] = k;
    std::sort(by_id, by_id + n, [&](int x, int y) { return s->car_id[x] < s->car_id[y]; });
    for (int i = 0; i < n; i++) {
        int k = by_id[i];
        if (i > 0 && s->car_id[k] == s->car_id[by_id[i - 1]]) return fail("rsg_batch_set_snapshot: repeated car id");
        int g = b->car_offset[arena] + i;
        const rsg_car_config want = snapshot_car_config(*s, k);
        if (s->car_team[k] != b->team[g] || !same_car_config(want, b->h_cfgs[b->preset[g]]))
            return fail("rsg_batch_set_snapshot: car " + std::to_string(s->car_id[k]) +
                        " has another team or car config than slot " + std::to_string(i));
        slot_of[k] = i;
    }
    const bool raw = s->raw_words > 0;
    if (raw && (s->raw_layout != raw_layout_tag() || s->raw_words != raw_words_for(n)))
        return fail("rsg_batch_set_snapshot: raw records of another layout or car count");
    std::vector<uint32_t> rw(s->raw, s->raw + (raw ? s->raw_words : 0));
    BatchDev v = raw_view(rw.data(), n);
    if (raw)
        for (int k = 0; k < n; k++) {
            CarRef rc = car_ref(v, 0, k);
            uint32_t cfg = rc.u(CW_CONFIG);
            if (rc.u(CW_ID) != s->car_id[k] || ((cfg >> 8) & 0xFF) != s->car_team[k])
                return fail("rsg_batch_set_snapshot: raw records disagree with the snapshot's cars");
        }
    if (download(b) != 0) return -1;
    BatchDev h = host_view(b);
    CarRef ar = arena_ref(h, arena);
    uint32_t ow = 0;
    for (int k = 0; k < n; k++) {
        ow |= (uint32_t)slot_of[k] << (4 * k);
        CarRef cr = car_ref(h, arena, slot_of[k]);
        if (raw) {
            CarRef rc = car_ref(v, 0, k);
            for (int w = 0; w < CAR_WORDS; w++) cr.u(w) = rc.u(w);
            // the config index is this batch's (the snapshot's configs matched)
            const int g = b->car_offset[arena] + slot_of[k];
            cr.u(CW_CONFIG) = (uint32_t)b->preset[g] | ((uint32_t)b->team[g] << 8);
        } else {
            car_from_api(cr, s->cars[k]);
            cr.u(CW_ID) = s->car_id[k];
            store_ctrl(cr, CW_CTRL, ctrl_from_api(s->controls[k]));
            // a new car body, then SetMutatorConfig's gravity (Arena::DeserializeNew)
            init_car_body_words(cr, s->mutators.gravity);
        }
    }
    if (raw) {
        CarRef rb = ball_ref(v, 0), db = ball_ref(h, arena);
        for (int w = 0; w < BALL_WORDS; w++) db.u(w) = rb.u(w);
        CarRef ra = arena_ref(v, 0);
        for (int w = 0; w < ARENA_WORDS; w++)
            if (w != AW_NUM_CARS && w != AW_CAR_ORDER) ar.u(w) = ra.u(w);
    } else {
        ar.su64(AW_TICK, s->tick_count);
        // the mode's default ball body, then SetMutatorConfig's gravity (the
        // deserialized config counts as unchanged: no ball remake)
        const rsg_mutator_config def = mutator_defaults(b->dev.mode);
        init_ball_body_words(ball_ref(h, arena), def.ball_radius, def.ball_mass, s->mutators.gravity);
        ball_from_api(ball_ref(h, arena), s->ball);
    }
    b->h_mut[(size_t)arena] = s->mutators;
    ar.u(AW_CAR_ORDER) = ow;
    if (mode_has_arena(b->dev.mode))
        for (int p = 0; p < PAD_NUM; p++) pad_from_api(pad_ref(h, arena, p), s->pads[p]);
    b->host_dirty = true;
    return 0;
}

// ---------------------------------------------------------------------------
// Ball prediction

}  // extern "C"

struct rsg_ball_pred {
    rsg_batch* arena = nullptr;  // the trackers' ball-only arenas (arena i = tracker i)
    int n = 0, T = 0;
    uint32_t* traj = nullptr;    // [T][RSG_PRED_WORDS][n]
    int32_t* meta = nullptr;     // head [n], size [n], last [n]
    uint32_t* cur = nullptr;     // [RSG_PRED_WORDS][n] staged states
    int32_t* ticks = nullptr;    // [n] staged ticks since the last update
    hipStream_t last_stream = nullptr;
    bool pending = false;        // an update was enqueued since the last wait
    int per_wave = 0;            // trackers per wave, 0 = derived from n and the SIMD count
};

namespace {

int pred_stride() {
    return (int)((offsetof(ArenaLocal, cars) + 15) & ~(size_t)15);
}

int pred_per_wave(int n) {
    int per = (n + device_simds() - 1) / device_simds();
    return per < 1 ? 1 : (per > 64 ? 64 : per);
}

int pred_wait(rsg_ball_pred* p) {
    if (p->pending) {
        HIP_OK(hipStreamSynchronize(p->last_stream));
        p->pending = false;
    }
    return 0;
}

int pred_launch(rsg_ball_pred* p, int kind, const BatchDev* src, hipStream_t s) {
    rsg_batch* a = p->arena;
    if (upload(a) != 0) return -1;
    ArenaGeom mv{};
    if (mode_has_arena(a->dev.mode)) {
        const DeviceArena& ga = g_arenas[arena_kind(a->dev.mode)];
        if (!ga.ready) return fail("rsg_ball_pred: no arena mesh (call rsg_init_mesh first)");
        mv = ga.geom;
    }
    PredDev d;
    d.traj = p->traj;
    d.head = p->meta;
    d.size = p->meta + p->n;
    d.last = p->meta + 2 * p->n;
    d.cur = p->cur;
    d.ticks = p->ticks;
    d.T = p->T;
    int per = p->per_wave > 0 ? p->per_wave : pred_per_wave(p->n);
    int stride = pred_stride();
    int grid = (p->n + per - 1) / per;
    HIP_OK(launch_ball_pred(a->dev, mv, d, kind, src ? *src : a->dev, stride, per, grid, s));
    a->dev_newer = true;
    p->last_stream = s;
    p->pending = true;
    return 0;
}

// Host states -> staged word planes (Ball::GetState units, as given).
int pred_stage(rsg_ball_pred* p, const rsg_ball_state* in, const int32_t* ticks) {
    const size_t n = (size_t)p->n;
    std::vector<uint32_t> w((size_t)RSG_PRED_WORDS * n);
    for (size_t i = 0; i < n; i++) {
        const rsg_ball_state& b = in[i];
        float f[18];
        std::memcpy(f, b.pos, 12);
        std::memcpy(f + 3, b.rot, 36);
        std::memcpy(f + 12, b.vel, 12);
        std::memcpy(f + 15, b.ang_vel, 12);
        for (int k = 0; k < 18; k++) std::memcpy(&w[k * n + i], &f[k], 4);
        w[18 * n + i] = (uint32_t)b.tick_count_since_update;
        w[19 * n + i] = (uint32_t)(b.tick_count_since_update >> 32);
        std::memcpy(&w[20 * n + i], &b.hs_y_target_dir, 4);
        std::memcpy(&w[21 * n + i], &b.hs_cur_target_speed, 4);
        std::memcpy(&w[22 * n + i], &b.hs_time_since_hit, 4);
    }
    if (pred_wait(p) != 0) return -1;  // the previous update may still read the staging buffers
    HIP_OK(hipMemcpy(p->cur, w.data(), w.size() * 4, hipMemcpyHostToDevice));
    if (ticks) HIP_OK(hipMemcpy(p->ticks, ticks, n * 4, hipMemcpyHostToDevice));
    return 0;
}

void pred_row_to_api(const uint32_t* r, size_t s, rsg_ball_state& o) {
    std::memset(&o, 0, sizeof(o));
    float f[18];
    for (int k = 0; k < 18; k++) std::memcpy(&f[k], &r[k * s], 4);
    std::memcpy(o.pos, f, 12);
    std::memcpy(o.rot, f + 3, 36);
    std::memcpy(o.vel, f + 12, 12);
    std::memcpy(o.ang_vel, f + 15, 12);
    o.tick_count_since_update = (uint64_t)r[18 * s] | ((uint64_t)r[19 * s] << 32);
    std::memcpy(&o.hs_y_target_dir, &r[20 * s], 4);
    std::memcpy(&o.hs_cur_target_speed, &r[21 * s], 4);
    std::memcpy(&o.hs_time_since_hit, &r[22 * s], 4);
}

// One tracker's ring -> predData order.
int pred_read(rsg_ball_pred* p, int tracker, rsg_ball_state* out, int max_out, int* size_out) {
    if (pred_wait(p) != 0) return -1;
    int32_t hs[2];
    HIP_OK(hipMemcpy(&hs[0], p->meta + tracker, 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&hs[1], p->meta + p->n + tracker, 4, hipMemcpyDeviceToHost));
    const int head = hs[0], size = hs[1];
    *size_out = size;
    int m = size < max_out ? size : max_out;
    if (m <= 0) return 0;
    // the tracker's column: T * RSG_PRED_WORDS words at a pitch of n words
    std::vector<uint32_t> col((size_t)p->T * RSG_PRED_WORDS);
    HIP_OK(hipMemcpy2D(col.data(), 4, p->traj + tracker, (size_t)p->n * 4, 4, (size_t)p->T * RSG_PRED_WORDS,
                       hipMemcpyDeviceToHost));
    for (int t = 0; t < m; t++) {
        int row = (head + t) % p->T;
        pred_row_to_api(col.data() + (size_t)row * RSG_PRED_WORDS, 1, out[t]);
    }
    return 0;
}

}  // namespace

extern "C" {

rsg_ball_pred* rsg_ball_pred_create(int mode, int n, int num_pred_ticks, float tick_rate) {
    if (n <= 0) { fail("rsg_ball_pred_create: n must be > 0"); return nullptr; }
    if (num_pred_ticks < 1) { fail("rsg_ball_pred_create: num_pred_ticks must be >= 1"); return nullptr; }
    if ((size_t)num_pred_ticks * RSG_PRED_WORDS * (size_t)n > ((size_t)1 << 36)) {
        fail("rsg_ball_pred_create: prediction buffer above 256 GiB");
        return nullptr;
    }
    std::vector<int> zeros((size_t)n, 0);
    rsg_batch_desc desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.n_arenas = n;
    desc.mode = mode;
    desc.tick_rate = tick_rate;
    desc.cars_per_arena = zeros.data();
    rsg_batch* a = rsg_batch_create(&desc);
    if (!a) return nullptr;
    auto* p = new rsg_ball_pred();
    p->arena = a;
    p->n = n;
    p->T = num_pred_ticks;
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&p->traj, (size_t)num_pred_ticks * RSG_PRED_WORDS * n * 4)) != hipSuccess ||
        (e = hipMalloc(&p->meta, (size_t)3 * n * 4)) != hipSuccess ||
        (e = hipMalloc(&p->cur, (size_t)RSG_PRED_WORDS * n * 4)) != hipSuccess ||
        (e = hipMalloc(&p->ticks, (size_t)n * 4)) != hipSuccess ||
        (e = hipMemset(p->meta, 0, (size_t)3 * n * 4)) != hipSuccess) {
        fail(std::string("rsg_ball_pred_create: ") + hipGetErrorString(e));
        rsg_ball_pred_destroy(p);
        return nullptr;
    }
    return p;
}

void rsg_ball_pred_destroy(rsg_ball_pred* p) {
    if (!p) return;
    (void)hipDeviceSynchronize();
    if (p->traj) (void)hipFree(p->traj);
    if (p->meta) (void)hipFree(p->meta);
    if (p->cur) (void)hipFree(p->cur);
    if (p->ticks) (void)hipFree(p->ticks);
    rsg_batch_destroy(p->arena);
    delete p;
}

int rsg_ball_pred_force_update(rsg_ball_pred* p, const rsg_ball_state* initial, void* stream) {
    if (!p || !initial) return fail("rsg_ball_pred_force_update: null argument");
    if (pred_stage(p, initial, nullptr) != 0) return -1;
    return pred_launch(p, PRED_FORCE, nullptr, (hipStream_t)stream);
}

int rsg_ball_pred_update_manual(rsg_ball_pred* p, const rsg_ball_state* cur, const int32_t* ticks_since_last_update,
                                void* stream) {
    if (!p || !cur || !ticks_since_last_update) return fail("rsg_ball_pred_update_manual: null argument");
    if (pred_stage(p, cur, ticks_since_last_update) != 0) return -1;
    return pred_launch(p, PRED_MANUAL, nullptr, (hipStream_t)stream);
}

int rsg_ball_pred_update_from_batch(rsg_ball_pred* p, rsg_batch* b, void* stream) {
    if (!p || !b) return fail("rsg_ball_pred_update_from_batch: null argument");
    if (b->dev.n != p->n || b->dev.mode != p->arena->dev.mode || b->dev.tick_time != p->arena->dev.tick_time)
        return fail("rsg_ball_pred_update_from_batch: the batch's arena count, mode or tick rate differs");
    if (upload(b) != 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    // the batch's state may have been written on another stream
    if (b->last_stream && b->last_stream != s) HIP_OK(hipStreamSynchronize(b->last_stream));
    return pred_launch(p, PRED_FROM_BATCH, &b->dev, s);
}

int rsg_ball_pred_get(rsg_ball_pred* p, int tracker, rsg_ball_state* out, int max_out) {
    if (!p || (!out && max_out > 0)) return fail("rsg_ball_pred_get: null argument");
    if (tracker < 0 || tracker >= p->n) return fail("rsg_ball_pred_get: tracker out of range");
    int size = 0;
    if (pred_read(p, tracker, out, max_out, &size) != 0) return -1;
    return size;
}

int rsg_ball_pred_get_all(rsg_ball_pred* p, rsg_ball_state* out) {
    if (!p || !out) return fail("rsg_ball_pred_get_all: null argument");
    if (pred_wait(p) != 0) return -1;
    const size_t n = (size_t)p->n, T = (size_t)p->T;
    std::vector<uint32_t> traj(T * RSG_PRED_WORDS * n);
    std::vector<int32_t> meta(2 * n);
    HIP_OK(hipMemcpy(traj.data(), p->traj, traj.size() * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(meta.data(), p->meta, meta.size() * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        for (size_t t = 0; t < T; t++) {
            rsg_ball_state& o = out[i * T + t];
            if ((int)t >= meta[n + i]) { std::memset(&o, 0, sizeof(o)); continue; }
            size_t row = ((size_t)meta[i] + t) % T;
            pred_row_to_api(traj.data() + row * RSG_PRED_WORDS * n + i, n, o);
        }
    }
    return 0;
}

int rsg_ball_pred_state_for_time(rsg_ball_pred* p, int tracker, float pred_time, rsg_ball_state* out) {
    if (!p || !out) return fail("rsg_ball_pred_state_for_time: null argument");
    if (tracker < 0 || tracker >= p->n) return fail("rsg_ball_pred_state_for_time: tracker out of range");
    if (pred_wait(p) != 0) return -1;
    int32_t size = 0;
    HIP_OK(hipMemcpy(&size, p->meta + p->n + tracker, 4, hipMemcpyDeviceToHost));
    if (size <= 0)
        return fail("rsg_ball_pred_state_for_time: predicted ball data is empty, update the prediction first");
    // RS_CLAMP(predTime / tickTime, 0, size - 1) in float, truncated to int
    float x = pred_time / p->arena->dev.tick_time;
    float lo = x > 0 ? x : 0;
    float hi = (float)(size - 1);
    int index = (int)(lo < hi ? lo : hi);
    std::vector<rsg_ball_state> tmp((size_t)index + 1);
    int got = 0;
    if (pred_read(p, tracker, tmp.data(), index + 1, &got) != 0) return -1;
    *out = tmp[(size_t)index];
    return 0;
}

int rsg_ball_pred_set_launch(rsg_ball_pred* p, int trackers_per_wave) {
    if (!p) return fail("null tracker");
    if (trackers_per_wave < 0 || trackers_per_wave > 64)
        return fail("rsg_ball_pred_set_launch: trackers_per_wave must be within [0, 64]");
    p->per_wave = trackers_per_wave;
    return 0;
}

int rsg_ball_pred_view(rsg_ball_pred* p, rsg_ball_pred_view_t* out) {
    if (!p || !out) return fail("rsg_ball_pred_view: null argument");
    out->traj = reinterpret_cast<const float*>(p->traj);
    out->head = p->meta;
    out->size = p->meta + p->n;
    out->n = p->n;
    out->num_pred_ticks = p->T;
    return 0;
}

int rsg_init_arena_meshes(int mode, const void* const* cmf_data, const uint64_t* sizes, int n) {
    if (!mode_has_arena(mode)) return fail("rsg_init_arena_meshes: the mode has no arena meshes");
    if (n > 0 && (!cmf_data || !sizes)) return fail("rsg_init_meshes: null argument");
    DeviceArena& g_mesh = g_arenas[arena_kind(mode)];
    std::lock_guard<std::mutex> lock(g_mesh_mutex);
    HostArena H;
    std::string err = build_host_arena(cmf_data, sizes, n, arena_kind(mode), H);
    if (!err.empty()) return fail("rsg_init_meshes: invalid collision mesh files (" + err + ")");
    // kernels of earlier launches may still read the old arrays
    (void)hipDeviceSynchronize();
    g_mesh.ready = false;
    for (DeviceMesh& m : g_mesh.meshes) {
        (void)hipFree(m.verts); (void)hipFree(m.tris); (void)hipFree(m.nodes);
        (void)hipFree(m.subtrees); (void)hipFree(m.tri_info); (void)hipFree(m.occ);
    }
    g_mesh.meshes.assign(H.meshes.size(), DeviceMesh());
    g_mesh.host = std::move(H);
    std::vector<MeshView> views;
    for (size_t i = 0; i < g_mesh.host.meshes.size(); i++) {
        const HostMesh& M = g_mesh.host.meshes[i];
        DeviceMesh& D = g_mesh.meshes[i];
        if (upload_array(&D.verts, M.verts) || upload_array(&D.tris, M.tris) || upload_array(&D.nodes, M.nodes) ||
            upload_array(&D.subtrees, M.subtrees) || upload_array(&D.tri_info, M.tri_info) ||
            upload_array(&D.occ, M.occ))
            return -1;
        MeshView v = M.view();
        v.verts = D.verts;
        v.tris = D.tris;
        v.nodes = D.nodes;
        v.subtrees = D.subtrees;
        v.tri_info = D.tri_info;
        v.occ = D.occ;
        views.push_back(v);
    }
    if (upload_array(&g_mesh.views, views) || upload_array(&g_mesh.cell_statics, g_mesh.host.cell_statics)) return -1;
    ArenaGeom g = g_mesh.host.geom();
    g.meshes = g_mesh.views;
    g.n_meshes = (int)views.size();
    g.cell_statics = g_mesh.cell_statics;
    g_mesh.geom = g;
    g_mesh.ready = true;
    return 0;
}

int rsg_init_meshes(const void* const* cmf_data, const uint64_t* sizes, int n) {
    return rsg_init_arena_meshes(RSG_MODE_SOCCAR, cmf_data, sizes, n);
}

int rsg_init_mesh(const void* cmf_data, uint64_t size) {
    const void* data[1] = {cmf_data};
    const uint64_t sizes[1] = {size};
    return rsg_init_meshes(data, sizes, 1);
}

}  // extern "C"
