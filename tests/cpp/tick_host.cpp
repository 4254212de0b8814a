// Host build of the batched tick (rocketsim_amd/csrc/rsg_tick.h, the code the
// step kernels run) for the CPU tests: a batch in the GPU's struct-of-arrays
// word layout, stepped arena by arena on the host.
//
// Build switches:
//   RSG_SSE_NORMALIZE        vector normalisation as the reference's x86-64
//                            build does it (btVector3::normalize: rsqrtss plus
//                            one Newton step, LinearMath/btVector3.h); without
//                            it, the GPU's (divide by the length);
//   RSG_TEST_MAX_MANIFOLDS,  lower contact capacities, so that ordinary play
//   RSG_TEST_MAX_SOLVER_ROWS reaches the overflow status bits;
//   RSG_TEST_SPLIT_WORLD=n   the world step as the cooperative kernel runs
//                            it: its narrowphase shared by n lanes (run here
//                            one after another, last lane first) with no
//                            penetration work space, the effects applied in
//                            pair order afterwards;
//   RSG_TEST_PAIR_EFFECTS=n  room for n recorded effects per pair.
#ifdef RSG_SSE_NORMALIZE
#include <immintrin.h>

template <class T>
static inline T sse_normalized(T a) {
    float l2 = (a.x * a.x + a.y * a.y) + a.z * a.z;
    float y = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(l2)));
    float h = l2 * 0.5f;
    h = h * y;
    h = h * y;
    y = y * (1.5f - h);
    T r = a;
    r.x = a.x * y;
    r.y = a.y * y;
    r.z = a.z * y;
    return r;
}
#define RSG_NORMALIZE_HOOK(a) sse_normalized(a)
#endif

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rsg.h"
#include "../../rocketsim_amd/csrc/rsg_mesh_build.h"
#include "../../rocketsim_amd/csrc/rsg_step.h"
#include "../../rocketsim_amd/csrc/rsg_api_convert.h"

using namespace rsg;

namespace {

struct HostBatch {
    HostArena geo;
    std::vector<int> cars_per_arena, car_offset;
    std::vector<uint32_t> car, ball, pad, arena;
    std::vector<rsg_mutator_config> mut;
    std::vector<rsg_car_config> cfgs;  // config table: presets, then the cars' other configs
    std::vector<uint8_t> cfg_index;    // each car's entry
    std::vector<rsg_boost_pad_config> custom_pads;
    BatchDev dev{};
};

}  // namespace

extern "C" {

// A host batch of n arenas (cars_per_arena[a] cars each, teams and presets
// per car, or each car's CarConfig in car_configs); cmfs = the arena's n_cmf
// mesh files in order (none for THE_VOID); mut = the MutatorConfig set
// before the cars are added (NULL: the mode's defaults).
void* th_create(int mode, int n, const int* cars_per_arena, const uint8_t* team, const uint8_t* preset,
                const void* const* cmfs, const uint64_t* cmf_sizes, int n_cmf, float tick_rate,
                const rsg_mutator_config* mut, const rsg_car_config* car_configs) {
    auto* h = new HostBatch();
    if (n_cmf > 0 && !build_host_arena(cmfs, cmf_sizes, n_cmf, arena_kind(mode), h->geo).empty()) {
        delete h;
        return nullptr;
    }
    int total = 0, maxc = 1;
    for (int a = 0; a < n; a++) {
        h->cars_per_arena.push_back(cars_per_arena[a]);
        h->car_offset.push_back(total);
        total += cars_per_arena[a];
        maxc = cars_per_arena[a] > maxc ? cars_per_arena[a] : maxc;
    }
    h->car.assign((size_t)CAR_WORDS * maxc * n, 0);
    h->ball.assign((size_t)BALL_WORDS * n, 0);
    h->pad.assign((size_t)PAD_WORDS * PAD_NUM * n, 0);
    h->arena.assign((size_t)ARENA_WORDS * n, 0);
    BatchDev& d = h->dev;
    d.car = h->car.data();
    d.ball = h->ball.data();
    d.pad = h->pad.data();
    d.arena = h->arena.data();
    d.n = n;
    d.max_cars = maxc;
    d.mode = mode;
    d.tick_time = 1 / tick_rate;
    h->mut.assign((size_t)n, mut ? *mut : mutator_defaults(mode));
    d.mut = h->mut.data();
    for (int p = 0; p < RSG_PRESET_COUNT; p++) h->cfgs.push_back(preset_config(p));
    for (int i = 0; i < total; i++) {
        if (!car_configs) { h->cfg_index.push_back(preset[i]); continue; }
        size_t k = 0;
        while (k < h->cfgs.size() && std::memcmp(&h->cfgs[k], &car_configs[i], sizeof(rsg_car_config)) != 0) k++;
        if (k == h->cfgs.size()) h->cfgs.push_back(car_configs[i]);
        h->cfg_index.push_back((uint8_t)k);
    }
    d.cfgs = h->cfgs.data();
    for (int a = 0; a < n; a++)
        init_arena_record(d, a, cars_per_arena[a], team + h->car_offset[a], h->cfg_index.data() + h->car_offset[a]);
    return h;
}

// Arena::SetMutatorConfig on every arena (rsg_batch_set_mutator_config's
// host side).
void th_set_mutators(void* p, const rsg_mutator_config* m) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) {
        apply_mutator_config(h.dev, a, h.mut[(size_t)a], *m);
        h.mut[(size_t)a] = *m;
    }
}

void th_destroy(void* p) { delete (HostBatch*)p; }

void th_set_cars(void* p, const rsg_car_state* in) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++)
        for (int i = 0; i < h.cars_per_arena[a]; i++) car_from_api(car_ref(h.dev, a, i), in[h.car_offset[a] + i]);
}

void th_get_cars(void* p, rsg_car_state* out) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++)
        for (int i = 0; i < h.cars_per_arena[a]; i++) car_to_api(car_ref(h.dev, a, i), out[h.car_offset[a] + i]);
}

void th_set_balls(void* p, const rsg_ball_state* in) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) ball_from_api(ball_ref(h.dev, a), in[a]);
}

void th_get_balls(void* p, rsg_ball_state* out) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) ball_to_api(ball_ref(h.dev, a), out[a]);
}

void th_set_pads(void* p, const rsg_pad_state* in) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++)
        for (int k = 0; k < PAD_NUM; k++) pad_from_api(pad_ref(h.dev, a, k), in[a * PAD_NUM + k]);
}

void th_get_pads(void* p, rsg_pad_state* out) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++)
        for (int k = 0; k < PAD_NUM; k++) pad_to_api(pad_ref(h.dev, a, k), out[a * PAD_NUM + k]);
}

void th_set_controls(void* p, const rsg_controls* in) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++)
        for (int i = 0; i < h.cars_per_arena[a]; i++)
            store_ctrl(car_ref(h.dev, a, i), CW_CTRL, ctrl_from_api(in[h.car_offset[a] + i]));
}

void th_set_tick_counts(void* p, const uint64_t* in) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) arena_ref(h.dev, a).su64(AW_TICK, in[a]);
}

void th_get_tick_counts(void* p, uint64_t* out) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) out[a] = arena_ref(h.dev, a).u64(AW_TICK);
}

void th_set_car_order(void* p, const int32_t* order) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) {
        uint32_t ow = 0;
        for (int k = 0; k < h.cars_per_arena[a]; k++) ow |= (uint32_t)order[h.car_offset[a] + k] << (4 * k);
        arena_ref(h.dev, a).u(AW_CAR_ORDER) = ow;
    }
}

void th_get_status(void* p, uint32_t* out) {
    HostBatch& h = *(HostBatch*)p;
    for (int a = 0; a < h.dev.n; a++) out[a] = arena_ref(h.dev, a).u(AW_STATUS);
}

#ifdef RSG_TEST_SPLIT_WORLD
// One tick in the order of rsg_step_arena_coop_kernel's phases, with the
// world step in its three parts.
static void tick_split(ArenaLocal& A, const TickParams& P, const ArenaGeom* geo, StepSpace* ws, PairLog* log) {
    tick_begin(A, P);
    {
        World world(A, geo, mode_has_arena(P.mode) && geo != nullptr);
        for (int oi = 0; oi < A.n_cars; oi++) car_pre_tick(A, P, world, A.order[oi], nullptr);
    }
    pads_pre_tick(A, P);
    tick_world_prologue(A, P, geo, ws);
    for (int lane = RSG_TEST_SPLIT_WORLD - 1; lane >= 0; lane--)
        tick_world_aabbs(A, P, geo, ws, lane, RSG_TEST_SPLIT_WORLD);
    tick_world_broadphase(A, P, geo, ws);
    for (int lane = RSG_TEST_SPLIT_WORLD - 1; lane >= 0; lane--)
        tick_world_pairs(A, P, geo, ws, log, lane, RSG_TEST_SPLIT_WORLD);
    tick_world_solve(A, P, geo, ws, log);
    for (int lane = RSG_TEST_SPLIT_WORLD - 1; lane >= 0; lane--)
        tick_world_integrate(A, P, geo, ws, lane, RSG_TEST_SPLIT_WORLD);
    for (int oi = 0; oi < A.n_cars; oi++) car_post_tick(A, P, A.order[oi]);
    if (mode_has_arena(P.mode))
        for (int p = 0; p < PAD_NUM; p++) pad_check_cars(A, p);
    tick_end(A, P);
}
#endif

// ArenaConfig::customBoostPads for every arena (the batch's custom_pads).
void th_set_custom_pads(void* p, const rsg_boost_pad_config* pads, int n) {
    HostBatch& h = *(HostBatch*)p;
    h.custom_pads.assign(pads, pads + n);
    h.dev.custom_pads = h.custom_pads.empty() ? nullptr : h.custom_pads.data();
    h.dev.n_custom_pads = n;
}

// ArenaConfig::noBallRot == false for every arena (the batch's ball_rotation).
void th_set_ball_rotation(void* p, int on) { ((HostBatch*)p)->dev.ball_rot = on != 0; }

// Steps every arena `ticks` ticks with the plain kernel's code path (or the
// split world step, RSG_TEST_SPLIT_WORLD).
void th_step(void* p, int ticks) {
    HostBatch& h = *(HostBatch*)p;
    const ArenaGeom mv = h.geo.geom();
    auto A = std::make_unique<ArenaLocal>();
    auto ws = std::make_unique<StepSpace>();
#ifdef RSG_TEST_SPLIT_WORLD
    auto log = std::make_unique<PairLog>();
    for (int a = 0; a < h.dev.n; a++) {
        load_arena(h.dev, a, *A);
        TickParams P;
        P.mode = h.dev.mode;
        P.dt = h.dev.tick_time;
        P.seed = h.dev.seed;
        P.arena_id = (uint32_t)a;
        P.synth_controls = h.dev.synth_controls != 0;
        P.auto_reset = h.dev.auto_reset != 0;
        const ArenaGeom* mp = mode_has_arena(h.dev.mode) ? &mv : nullptr;
        for (int t = 0; t < ticks; t++) tick_split(*A, P, mp, ws.get(), log.get());
        store_arena(h.dev, a, *A);
    }
#else
    for (int a = 0; a < h.dev.n; a++) step_arena(h.dev, mv, a, ticks, *A, ws.get());
#endif
}

}  // extern "C"

extern "C" {

// rsqrtss of eight fixed inputs on this CPU, the part of the reference's
// normalisation whose last bits differ between x86 implementations.
void th_rsqrt_signature(float* out) {
    const float in[8] = {1.f, 2.f, 3.f, 0.5f, 1.2345f, 7.654321f, 1e-3f, 123.456f};
    for (int i = 0; i < 8; i++) {
#ifdef RSG_SSE_NORMALIZE
        out[i] = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(in[i])));
#else
        out[i] = 0.f;
#endif
    }
}

}  // extern "C"
