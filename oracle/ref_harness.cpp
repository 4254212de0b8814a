// ref_harness.cpp -- C entry points around the reference RocketSim, compiled
// from the reference's own sources (see Makefile) into oracle/_ref/librsref.so.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle: tests, the smoke check and the
// CPU-baseline leg of the benchmark load it to check or time against; the
// product path never links or calls it.
//
// What it exposes:
//   * mesh init from an in-memory .cmf blob   (RocketSim::InitFromMem, src/RocketSim.cpp:92-185)
//   * arena create / add car / destroy        (src/Sim/Arena/Arena.cpp:50-62,416-540)
//   * full-state get/set incl. the per-wheel values btVehicleRL keeps between
//     ticks (src/Sim/btVehicleRL/btVehicleRL.h:9-27), which Car::SetState does
//     not touch (src/Sim/Car/Car.cpp:23-36)
//   * Arena::Step (src/Sim/Arena/Arena.cpp:689-756)
//   * a multi-threaded timed run of the seeded workload (one Arena per thread)
//     for the CPU baseline
//   * BallPredTracker (src/Sim/BallPredTracker/BallPredTracker.cpp) over an
//     arena of this harness.

#include "RocketSim.h"
#include "Sim/BallPredTracker/BallPredTracker.h"
#include "Sim/GameEventTracker/GameEventTracker.h"
#include "rsg_state.h"
#include "rsg_workload.h"
#include "rsg_rocketsim.hpp"
#include "../libsrc/bullet3-3.24/BulletCollision/BroadphaseCollision/btRSBroadphase.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionShapes/btBvhTriangleMeshShape.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionShapes/btBoxShape.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionShapes/btSphereShape.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionShapes/btTriangleShape.h"
#include "../libsrc/bullet3-3.24/BulletCollision/NarrowPhaseCollision/btGjkPairDetector.h"
#include "../libsrc/bullet3-3.24/BulletCollision/NarrowPhaseCollision/btVoronoiSimplexSolver.h"
#include "../libsrc/bullet3-3.24/BulletCollision/NarrowPhaseCollision/btGjkEpaPenetrationDepthSolver.h"
#include "../libsrc/bullet3-3.24/BulletCollision/NarrowPhaseCollision/btSubSimplexConvexCast.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionDispatch/btBoxBoxDetector.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <pthread.h>
#include <sched.h>
#include <thread>
#include <vector>

using namespace RocketSim;

namespace {

const CarConfig* preset_config(int p) {
    switch (p) {
        case RSG_PRESET_OCTANE: return &CAR_CONFIG_OCTANE;
        case RSG_PRESET_DOMINUS: return &CAR_CONFIG_DOMINUS;
        case RSG_PRESET_PLANK: return &CAR_CONFIG_PLANK;
        case RSG_PRESET_BREAKOUT: return &CAR_CONFIG_BREAKOUT;
        case RSG_PRESET_HYBRID: return &CAR_CONFIG_HYBRID;
        case RSG_PRESET_MERC: return &CAR_CONFIG_MERC;
        case RSG_PRESET_PSYCLOPS: return &CAR_CONFIG_PSYCLOPS;
        default: return nullptr;
    }
}

struct RefArena {
    Arena* arena = nullptr;
    std::vector<Car*> cars;  // creation order
    // Broadphase registration bookkeeping of the dynamic proxies (0 = ball,
    // 1 + i = car i): the cell each proxy is registered in, and the order of
    // its last (re-)registration, which is the order of the cell lists.
    std::vector<int> bp_cell;
    std::vector<uint32_t> bp_seq;
    uint32_t bp_counter = 0;
    // goal-score and car-bump callbacks, in the order the reference fires them
    std::vector<rsg_event> events;
    uint32_t event_seq = 0;
};

btCollisionObject* body_object(RefArena* ra, int body) {
    return body == 0 ? (btCollisionObject*)&ra->arena->ball->_rigidBody
                     : (btCollisionObject*)&ra->cars[body - 1]->_rigidBody;
}

int proxy_cell(RefArena* ra, int body) {
    return ((btRSBroadphaseProxy*)body_object(ra, body)->getBroadphaseHandle())->cellIdx;
}

// After a tick: a proxy whose cell changed was re-registered (btRSBroadphase::setAabb,
// which runs once per proxy per tick in collision-object order).
void track_bp(RefArena* ra) {
    int nb = (int)ra->bp_cell.size();
    for (int i = 0; i < nb; i++) {
        int c = proxy_cell(ra, i);
        if (c != ra->bp_cell[i]) {
            ra->bp_cell[i] = c;
            if (nb > 1) ra->bp_seq[i] = ++ra->bp_counter;
        }
    }
}

thread_local std::string g_err;

void vget(const Vec& v, float* f) { f[0] = v.x; f[1] = v.y; f[2] = v.z; }

}  // namespace

extern "C" {

const char* rsref_last_error() { return g_err.c_str(); }

// blobs: n .cmf files for SOCCAR (may be 0 for a THE_VOID-only process);
// hoops: n_hoops .cmf files for HOOPS.
int rsref_init_modes(const uint8_t* const* blobs, const uint64_t* sizes, int n, const uint8_t* const* hoops,
                     const uint64_t* hoops_sizes, int n_hoops) {
    try {
        std::map<GameMode, std::vector<FileData>> m;
        for (int i = 0; i < n; i++) {
            // Heatseeker plays in the Soccar arena
            m[GameMode::SOCCAR].push_back(FileData(blobs[i], blobs[i] + sizes[i]));
            m[GameMode::HEATSEEKER].push_back(FileData(blobs[i], blobs[i] + sizes[i]));
        }
        for (int i = 0; i < n_hoops; i++) m[GameMode::HOOPS].push_back(FileData(hoops[i], hoops[i] + hoops_sizes[i]));
        RocketSim::InitFromMem(m, true);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int rsref_init(const uint8_t* const* blobs, const uint64_t* sizes, int n) {
    return rsref_init_modes(blobs, sizes, n, nullptr, nullptr, 0);
}

// pads: n_pads custom boost pads as (x, y, z, is_big) (ArenaConfig::customBoostPads), or none.
void* rsref_arena_create_cfg(int mode, float tick_rate, int no_ball_rot, const float* pads, int n_pads) {
    try {
        auto* ra = new RefArena();
        ArenaConfig cfg;
        cfg.noBallRot = no_ball_rot != 0;
        if (n_pads > 0) {
            cfg.useCustomBoostPads = true;
            for (int i = 0; i < n_pads; i++) {
                BoostPadConfig pc;
                pc.pos = Vec(pads[4 * i], pads[4 * i + 1], pads[4 * i + 2]);
                pc.isBig = pads[4 * i + 3] != 0.f;
                cfg.customBoostPads.push_back(pc);
            }
        }
        ra->arena = Arena::Create((GameMode)mode, cfg, tick_rate);
        // The solver info's timestep is only written by the first world step
        // (btDiscreteDynamicsWorld.cpp:414); before that it holds Bullet's
        // default 1/60, which the wheel pushback of the very first tick reads
        // (btVehicleRL.cpp:186-194).  Start it at the arena's tick time so a
        // freshly created arena behaves like one that has already stepped.
        ra->arena->_bulletWorld.getSolverInfo().m_timeStep = ra->arena->tickTime;
        ra->bp_cell.push_back(proxy_cell(ra, 0));
        ra->bp_seq.push_back(0);
        ra->arena->SetCarBumpCallback([](Arena* arena, Car* bumper, Car* victim, bool is_demo, void* ui) {
            auto* r = (RefArena*)ui;
            rsg_event e{};
            e.type = is_demo ? RSG_EVENT_DEMO : RSG_EVENT_BUMP;
            e.tick = arena->tickCount;
            e.a = bumper->id;
            e.b = victim->id;
            e.seq = r->event_seq++;
            r->events.push_back(e);
        }, ra);
        if ((GameMode)mode != GameMode::THE_VOID)
            ra->arena->SetGoalScoreCallback([](Arena* arena, Team team, void* ui) {
                auto* r = (RefArena*)ui;
                rsg_event e{};
                e.type = RSG_EVENT_GOAL;
                e.tick = arena->tickCount;
                e.a = (uint32_t)team;
                e.seq = r->event_seq++;
                r->events.push_back(e);
            }, ra);
        return ra;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void* rsref_arena_create(int mode, float tick_rate) { return rsref_arena_create_cfg(mode, tick_rate, 1, nullptr, 0); }

void rsref_arena_destroy(void* h) {
    auto* ra = (RefArena*)h;
    if (!ra) return;
    delete ra->arena;
    delete ra;
}

// Returns the car id (>0), or 0 on error.
uint32_t rsref_add_car(void* h, int team, int preset) {
    auto* ra = (RefArena*)h;
    const CarConfig* cfg = preset_config(preset);
    if (!cfg) { g_err = "bad preset"; return 0; }
    Car* c = ra->arena->AddCar((Team)team, *cfg);
    ra->cars.push_back(c);
    ra->bp_cell.push_back(proxy_cell(ra, (int)ra->cars.size()));
    ra->bp_seq.push_back(++ra->bp_counter);
    return c->id;
}

// AddCar with any CarConfig.
uint32_t rsref_add_car_config(void* h, int team, const rsg_car_config* cfg) {
    auto* ra = (RefArena*)h;
    Car* c = ra->arena->AddCar((Team)team, rsg_amd::FromRsg(*cfg));
    ra->cars.push_back(c);
    ra->bp_cell.push_back(proxy_cell(ra, (int)ra->cars.size()));
    ra->bp_seq.push_back(++ra->bp_counter);
    return c->id;
}

// The reference's preset CarConfig, through the type adapter.
int rsref_car_config_preset(int preset, rsg_car_config* out) {
    const CarConfig* cfg = preset_config(preset);
    if (!cfg) return -1;
    *out = rsg_amd::ToRsg(*cfg);
    return 0;
}

int rsref_num_cars(void* h) { return (int)((RefArena*)h)->cars.size(); }
uint32_t rsref_car_id(void* h, int i) { return ((RefArena*)h)->cars[i]->id; }

void rsref_set_controls(void* h, int i, const rsg_controls* c) {
    ((RefArena*)h)->cars[i]->controls = rsg_amd::FromRsg(*c);
}

void rsref_get_car(void* h, int i, rsg_car_state* s) {
    // the public type adapter (include/rsg_rocketsim.hpp), plus the
    // broadphase record this harness tracks
    auto* ra = (RefArena*)h;
    *s = rsg_amd::ToRsg(*ra->cars[i]);
    s->bp_cell = ra->bp_cell[1 + i];
    s->bp_seq = ra->bp_seq[1 + i];
}

void rsref_set_car(void* h, int i, const rsg_car_state* s) {
    rsg_amd::ApplyRsg(*s, *((RefArena*)h)->cars[i]);
}

void rsref_get_ball(void* h, rsg_ball_state* s) {
    auto* ra = (RefArena*)h;
    *s = rsg_amd::ToRsg(ra->arena->ball->GetState());
    s->bp_cell = ra->bp_cell[0];
    s->bp_seq = ra->bp_seq[0];
}

void rsref_set_ball(void* h, const rsg_ball_state* s) {
    Ball* b = ((RefArena*)h)->arena->ball;
    b->SetState(rsg_amd::FromRsg(*s));
    b->_internalState.tickCountSinceUpdate = s->tick_count_since_update;
}

int rsref_num_pads(void* h) { return (int)((RefArena*)h)->arena->_boostPads.size(); }

void rsref_get_pad(void* h, int i, rsg_pad_state* s) {
    *s = rsg_amd::ToRsg(((RefArena*)h)->arena->_boostPads[i]->GetState());
}

void rsref_set_pad(void* h, int i, const rsg_pad_state* s) {
    ((RefArena*)h)->arena->_boostPads[i]->SetState(rsg_amd::FromRsg(*s));
}

uint64_t rsref_get_tick_count(void* h) { return ((RefArena*)h)->arena->tickCount; }
void rsref_set_tick_count(void* h, uint64_t t) { ((RefArena*)h)->arena->tickCount = t; }

int rsref_step(void* h, int ticks) {
    try {
        auto* ra = (RefArena*)h;
        for (int t = 0; t < ticks; t++) {
            ra->arena->Step(1);
            track_bp(ra);
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int rsref_is_ball_scored(void* h) { return ((RefArena*)h)->arena->IsBallScored() ? 1 : 0; }

void rsref_reset_kickoff(void* h, int seed) { ((RefArena*)h)->arena->ResetToRandomKickoff(seed); }

// Pad static config, for checking the tables restated by the product.
void rsref_pad_config(void* h, int i, float* pos3, int* is_big) {
    BoostPad* p = ((RefArena*)h)->arena->_boostPads[i];
    vget(p->config.pos, pos3);
    *is_big = p->config.isBig;
}

// Timed CPU run of the seeded workload: `threads` std::threads, each owning
// `arenas_per_thread` arenas of (n_blue, n_orange) cars of `preset`, stepped
// `ticks` ticks one tick at a time with the rsg_workload.h control process and
// a kickoff reset whenever a goal is scored.  Returns wall seconds; writes the
// total number of arena-ticks to *arena_ticks.  With `cpus` (one entry per
// thread), thread t is pinned to CPU cpus[t].
double rsref_run_workload(int threads, int arenas_per_thread, int n_blue, int n_orange,
                          int preset, int ticks, uint64_t seed, uint64_t* arena_ticks, const int* cpus) {
    std::atomic<uint64_t> total{0};
    std::atomic<int> ready{0};
    std::atomic<bool> go{false};
    std::vector<std::thread> pool;
    std::vector<std::chrono::steady_clock::time_point> finish(threads);
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            if (cpus) {
                cpu_set_t set;
                CPU_ZERO(&set);
                CPU_SET(cpus[t], &set);
                pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
            }
            std::vector<RefArena> arenas(arenas_per_thread);
            std::vector<std::vector<rsg_ctrl_gen>> gens(arenas_per_thread);
            for (int a = 0; a < arenas_per_thread; a++) {
                arenas[a].arena = Arena::Create(GameMode::SOCCAR, ArenaConfig(), 120.f);
                for (int i = 0; i < n_blue + n_orange; i++) {
                    Car* c = arenas[a].arena->AddCar(i < n_blue ? Team::BLUE : Team::ORANGE,
                                                    *preset_config(preset));
                    arenas[a].cars.push_back(c);
                }
                arenas[a].arena->ResetToRandomKickoff((int)(seed + t * 7919 + a));
                gens[a].assign(n_blue + n_orange, rsg_ctrl_gen{0, 0});
            }
            ready++;
            while (!go.load()) std::this_thread::yield();
            uint64_t done = 0;
            for (int k = 0; k < ticks; k++) {
                for (int a = 0; a < arenas_per_thread; a++) {
                    uint32_t arena_id = (uint32_t)(t * arenas_per_thread + a);
                    for (size_t i = 0; i < arenas[a].cars.size(); i++) {
                        rsg_controls c;
                        // Keep the previous draw unless a redraw is due.
                        c = rsg_amd::ToRsg(arenas[a].cars[i]->controls);
                        rsg_ctrl_next(seed, arena_id, (uint32_t)i, &gens[a][i], &c);
                        arenas[a].cars[i]->controls = rsg_amd::FromRsg(c);
                    }
                    arenas[a].arena->Step(1);
                    if (arenas[a].arena->IsBallScored())
                        arenas[a].arena->ResetToRandomKickoff((int)(seed + k));
                    done++;
                }
            }
            finish[t] = std::chrono::steady_clock::now();
            total += done;
            for (auto& ra : arenas) delete ra.arena;
        });
    }
    while (ready.load() < threads) std::this_thread::yield();
    auto t0 = std::chrono::steady_clock::now();
    go = true;
    for (auto& th : pool) th.join();
    auto t1 = *std::max_element(finish.begin(), finish.end());
    *arena_ticks = total.load();
    return std::chrono::duration<double>(t1 - t0).count();
}


// Debug view of the last tick's contact manifolds, in the dispatcher's array
// order.  Body codes: 0 = ball, 1 + i = car i (creation order), -1 - s =
// static object s (_worldCollisionRBs order: meshes, then planes).
struct rsref_point {
    float pos_a[3], pos_b[3], normal[3];
    float dist, friction, restitution, applied_impulse;
    int32_t special;
    int32_t part0, index0, part1, index1;
};

static int body_code(RefArena* ra, const btCollisionObject* o) {
    if (o == &ra->arena->ball->_rigidBody) return 0;
    for (size_t i = 0; i < ra->cars.size(); i++)
        if (o == &ra->cars[i]->_rigidBody) return 1 + (int)i;
    auto& st = ra->arena->_worldCollisionRBs;
    for (size_t i = 0; i < st.size(); i++)
        if (o == st[i]) return -1 - (int)i;
    return -1000;
}

int rsref_num_manifolds(void* h) {
    auto* ra = (RefArena*)h;
    return ra->arena->_bulletWorldParams.collisionDispatcher.getNumManifolds();
}

// Returns the number of points; writes body codes and up to 4 points.
int rsref_get_manifold(void* h, int m, int* body0, int* body1, float* threshold, rsref_point* pts) {
    auto* ra = (RefArena*)h;
    btPersistentManifold* mf = ra->arena->_bulletWorldParams.collisionDispatcher.getManifoldByIndexInternal(m);
    *body0 = body_code(ra, mf->getBody0());
    *body1 = body_code(ra, mf->getBody1());
    *threshold = mf->getContactBreakingThreshold();
    int n = mf->getNumContacts();
    for (int i = 0; i < n; i++) {
        const btManifoldPoint& cp = mf->getContactPoint(i);
        rsref_point& p = pts[i];
        for (int k = 0; k < 3; k++) {
            p.pos_a[k] = cp.m_positionWorldOnA[k];
            p.pos_b[k] = cp.m_positionWorldOnB[k];
            p.normal[k] = cp.m_normalWorldOnB[k];
        }
        p.dist = cp.m_distance1;
        p.friction = cp.m_combinedFriction;
        p.restitution = cp.m_combinedRestitution;
        p.applied_impulse = cp.m_appliedImpulse;
        p.special = cp.m_isSpecial;
        p.part0 = cp.m_partId0; p.index0 = cp.m_index0;
        p.part1 = cp.m_partId1; p.index1 = cp.m_index1;
    }
    return n;
}

// Broadphase cell (i, j, k) of a dynamic body's proxy, and its position in
// the proxy's own cell list (dynHandles order).
int rsref_get_proxy_cell(void* h, int body, int* ijk, int* order_in_cell, int* list_len) {
    auto* ra = (RefArena*)h;
    btCollisionObject* o = body == 0 ? (btCollisionObject*)&ra->arena->ball->_rigidBody
                                     : (btCollisionObject*)&ra->cars[body - 1]->_rigidBody;
    auto* bp = (btRSBroadphase*)ra->arena->_bulletWorldParams.broadphase;
    auto* px = (btRSBroadphaseProxy*)o->getBroadphaseHandle();
    ijk[0] = px->iIdx; ijk[1] = px->jIdx; ijk[2] = px->kIdx;
    auto& cell = bp->cells[px->cellIdx];
    *list_len = (int)cell.dynHandles.size();
    *order_in_cell = -1;
    for (size_t i = 0; i < cell.dynHandles.size(); i++)
        if (cell.dynHandles[i] == px) *order_in_cell = (int)i;
    return px->cellIdx;
}

// dynHandles of a given cell as body codes.
int rsref_get_cell_list(void* h, int cell_idx, int* out, int max_out) {
    auto* ra = (RefArena*)h;
    auto* bp = (btRSBroadphase*)ra->arena->_bulletWorldParams.broadphase;
    auto& cell = bp->cells[cell_idx];
    int n = 0;
    for (auto* px : cell.dynHandles) {
        if (n < max_out) out[n] = body_code(ra, (const btCollisionObject*)px->m_clientObject);
        n++;
    }
    return n;
}


// Triangle indices the reference's mesh BVH reports, in report order, for an
// AABB query (processAllTriangles) and for a ray query (performRaycast).
namespace {
struct RecordTris : public btTriangleCallback {
    int* out; int max_out; int n = 0;
    void processTriangle(btVector3*, int, int tri) override {
        if (n < max_out) out[n] = tri;
        n++;
    }
};
}

int rsref_mesh_query_aabb(const float* mn, const float* mx, int* out, int max_out) {
    auto& meshes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    if (meshes.empty()) return -1;
    RecordTris cb; cb.out = out; cb.max_out = max_out;
    meshes[0]->processAllTriangles(&cb, btVector3(mn[0], mn[1], mn[2]), btVector3(mx[0], mx[1], mx[2]));
    return cb.n;
}

int rsref_mesh_query_ray(const float* from, const float* to, int* out, int max_out) {
    auto& meshes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    if (meshes.empty()) return -1;
    RecordTris cb; cb.out = out; cb.max_out = max_out;
    meshes[0]->performRaycast(&cb, btVector3(from[0], from[1], from[2]), btVector3(to[0], to[1], to[2]));
    return cb.n;
}

// Internal-edge info of mesh triangle `tri` (btTriangleInfoMap entry): returns
// 0 and writes the three edge angles (v0v1, v1v2, v2v0) and flags, or -1 when
// the triangle has no entry.
int rsref_tri_info(int tri, float* angles, uint32_t* flags) {
    auto& meshes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    if (meshes.empty()) return -2;
    const btTriangleInfoMap* map = meshes[0]->getTriangleInfoMap();
    if (!map) return -2;
    int hash = (0 << (31 - MAX_NUM_PARTS_IN_BITS)) | tri;
    const btTriangleInfo* info = map->find(hash);
    if (!info) return -1;
    angles[0] = info->m_edgeV0V1Angle;
    angles[1] = info->m_edgeV1V2Angle;
    angles[2] = info->m_edgeV2V0Angle;
    *flags = (uint32_t)info->m_flags;
    return 0;
}

// Static objects registered in a broadphase cell, as a bit mask over
// _worldCollisionRBs order, and their list order packed 4 bits each.
int rsref_cell_statics(void* h, int cell_idx, uint32_t* order_packed) {
    auto* ra = (RefArena*)h;
    auto* bp = (btRSBroadphase*)ra->arena->_bulletWorldParams.broadphase;
    auto& cell = bp->cells[cell_idx];
    int mask = 0;
    uint32_t packed = 0;
    int n = 0;
    for (auto* px : cell.staticHandles) {
        int code = body_code(ra, (const btCollisionObject*)px->m_clientObject);
        int s = -1 - code;
        mask |= 1 << s;
        if (n < 8) packed |= (uint32_t)s << (4 * n);
        n++;
    }
    *order_packed = packed;
    return mask;
}

// The reference's own triangle test for broadphase cell (i, j, k).
int rsref_cell_has_tris(void* h, int i, int j, int k) {
    auto* ra = (RefArena*)h;
    auto* bp = (btRSBroadphase*)ra->arena->_bulletWorldParams.broadphase;
    auto& meshes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    btVector3 cmin = bp->GetCellMinPos(i, j, k);
    btVector3 cmax = cmin + btVector3(bp->cellSize, bp->cellSize, bp->cellSize);
    int out[1];
    RecordTris cb; cb.out = out; cb.max_out = 0;
    meshes[0]->processAllTriangles(&cb, cmin, cmax);
    return cb.n;
}

// Fresh broadphase with one static proxy for static object s at the given
// AABB; returns whether that proxy is registered in cell `cell_idx`.
int rsref_fresh_static_cells(void* h, int s, const float* mn, const float* mx, uint8_t* out_cells) {
    auto* ra = (RefArena*)h;
    auto* bpo = (btRSBroadphase*)ra->arena->_bulletWorldParams.broadphase;
    btHashedOverlappingPairCache cache;
    btRSBroadphase bp(bpo->minPos, bpo->maxPos, bpo->cellSize, &cache, 16);
    btCollisionObject* o = ra->arena->_worldCollisionRBs[s];
    btBroadphaseProxy* px = bp.createProxy(btVector3(mn[0], mn[1], mn[2]), btVector3(mx[0], mx[1], mx[2]),
                                           o->getCollisionShape()->getShapeType(), o, 2, -3, nullptr);
    for (int c = 0; c < bp.totalCells; c++) {
        out_cells[c] = 0;
        for (auto* q : bp.cells[c].staticHandles) if (q == px) out_cells[c] = 1;
    }
    return bp.totalCells;
}

int rsref_debug_add_loop(void* h, const float* mn, const float* mx, int* out, int max_out) {
    auto* ra = (RefArena*)h;
    auto* bp = (btRSBroadphase*)ra->arena->_bulletWorldParams.broadphase;
    btVector3 aabbMin(mn[0], mn[1], mn[2]), aabbMax(mx[0], mx[1], mx[2]);
    for (int i = 0; i < 3; i++) aabbMax[i] = btMin(aabbMax[i], bp->maxPos[i]);
    int iMin, jMin, kMin, iMax, jMax, kMax;
    bp->GetCellIndices(aabbMin, iMin, jMin, kMin);
    bp->GetCellIndices(aabbMax, iMax, jMax, kMax);
    auto& meshes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    auto* tm = (btTriangleMeshShape*)meshes[0];
    int n = 0;
    out[n++] = iMin; out[n++] = jMin; out[n++] = kMin; out[n++] = iMax; out[n++] = jMax; out[n++] = kMax;
    for (int i = iMin; i <= iMax; i++)
        for (int j = jMin; j <= jMax; j++)
            for (int k = kMin; k <= kMax; k++) {
                btVector3 cellMin = bp->GetCellMinPos(i, j, k);
                btVector3 cellMax = cellMin + btVector3(bp->cellSize, bp->cellSize, bp->cellSize);
                int o[1];
                RecordTris cb; cb.out = o; cb.max_out = 0;
                tm->processAllTriangles(&cb, cellMin, cellMax);
                if (n + 4 <= max_out) { out[n++] = i; out[n++] = j; out[n++] = k; out[n++] = cb.n; }
            }
    return n;
}

// The reference's GJK pair detector on (box, shape B): box given by its full
// half extents (btBoxShape constructor semantics), at transform xa (basis rows
// then origin); B is a sphere (kind 0, radius r) or a triangle (kind 2, verts).
// Returns 1 with normal / point on B / distance when a contact is reported.
namespace {
struct CaptureResult : public btDiscreteCollisionDetectorInterface::Result {
    bool has = false; btVector3 n, p; btScalar d = 0;
    void setShapeIdentifiersA(int, int) override {}
    void setShapeIdentifiersB(int, int) override {}
    void addContactPoint(const btVector3& normal, const btVector3& point, btScalar depth) override {
        has = true; n = normal; p = point; d = depth;
    }
};
btTransform xf_from_floats(const float* x) {
    btMatrix3x3 m(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8]);
    return btTransform(m, btVector3(x[9], x[10], x[11]));
}
}

int rsref_gjk(const float* box_half, const float* xa, int kind_b, float radius, const float* tri,
              const float* xb, float max_dist2, float* out_n, float* out_p, float* out_d, int* method) {
    btBoxShape box(btVector3(box_half[0], box_half[1], box_half[2]));
    btSphereShape sph(radius);
    btTriangleShape tr(btVector3(tri[0], tri[1], tri[2]), btVector3(tri[3], tri[4], tri[5]), btVector3(tri[6], tri[7], tri[8]));
    tr.setMargin(0);
    const btConvexShape* b = kind_b == 0 ? (const btConvexShape*)&sph : (const btConvexShape*)&tr;
    btVoronoiSimplexSolver simplex;
    btGjkEpaPenetrationDepthSolver pd;
    btGjkPairDetector det(&box, b, &simplex, &pd);
    det.setMinkowskiA(&box);
    det.setMinkowskiB(b);
    btGjkPairDetector::ClosestPointInput in;
    in.m_maximumDistanceSquared = max_dist2;
    in.m_transformA = xf_from_floats(xa);
    in.m_transformB = xf_from_floats(xb);
    CaptureResult r;
    det.getClosestPoints(in, r);
    *method = det.m_lastUsedMethod * 100 + det.m_degenerateSimplex;
    if (!r.has) return 0;
    for (int k = 0; k < 3; k++) { out_n[k] = r.n[k]; out_p[k] = r.p[k]; }
    *out_d = r.d;
    return 1;
}

// The reference's ray vs convex test (btSubsimplexConvexCast of a point, as in
// btCollisionWorld::rayTestSingleInternal): shape kind 0 sphere (radius r),
// 1 box (full half extents).  Returns 1 and fraction / normal on a hit.
int rsref_ray_convex(int kind, float radius, const float* box_half, const float* xb,
                     const float* from, const float* to, float* frac, float* normal) {
    btSphereShape point(0.f);
    point.setMargin(0.f);
    btSphereShape sph(radius);
    btBoxShape box(btVector3(box_half[0], box_half[1], box_half[2]));
    const btConvexShape* b = kind == 0 ? (const btConvexShape*)&sph : (const btConvexShape*)&box;
    btVoronoiSimplexSolver simplex;
    btSubsimplexConvexCast caster(&point, b, &simplex);
    btConvexCast::CastResult res;
    res.m_fraction = 1.f;
    btTransform ft(btMatrix3x3::getIdentity(), btVector3(from[0], from[1], from[2]));
    btTransform tt(btMatrix3x3::getIdentity(), btVector3(to[0], to[1], to[2]));
    btTransform bt = xf_from_floats(xb);
    if (!caster.calcTimeOfImpact(ft, tt, bt, bt, res)) return 0;
    if (!(res.m_normal.length2() > btScalar(0.0001))) return 0;
    res.m_normal.normalize();
    *frac = res.m_fraction;
    for (int k = 0; k < 3; k++) normal[k] = res.m_normal[k];
    return 1;
}

// The reference's box-box detector (btBoxBoxDetector::getClosestPoints, i.e.
// dBoxBox2) on two boxes of full half extents h0, h1 at transforms x0, x1
// (basis rows then origin): every reported contact, in report order, as
// normal on box 1, point on box 1 and signed distance.  Returns the count.
namespace {
struct MultiResult : public btDiscreteCollisionDetectorInterface::Result {
    int n = 0;
    float* normals; float* points; float* depths; int cap;
    void setShapeIdentifiersA(int, int) override {}
    void setShapeIdentifiersB(int, int) override {}
    void addContactPoint(const btVector3& normal, const btVector3& point, btScalar depth) override {
        if (n < cap) {
            for (int k = 0; k < 3; k++) { normals[3 * n + k] = normal[k]; points[3 * n + k] = point[k]; }
            depths[n] = depth;
        }
        n++;
    }
};
}

int rsref_box_box(const float* h0, const float* x0, const float* h1, const float* x1, float* normals,
                  float* points, float* depths, int cap) {
    btBoxShape b0(btVector3(h0[0], h0[1], h0[2]));
    btBoxShape b1(btVector3(h1[0], h1[1], h1[2]));
    btBoxBoxDetector det(&b0, &b1);
    btDiscreteCollisionDetectorInterface::ClosestPointInput in;
    in.m_transformA = xf_from_floats(x0);
    in.m_transformB = xf_from_floats(x1);
    MultiResult r;
    r.normals = normals; r.points = points; r.depths = depths; r.cap = cap;
    det.getClosestPoints(in, r, false);
    return r.n;
}

// Box margin and implicit half extents the btBoxShape constructor derives.
void rsref_box_dims(const float* half, float* implicit, float* margin) {
    btBoxShape box(btVector3(half[0], half[1], half[2]));
    btVector3 im = box.getImplicitShapeDimensions();
    for (int k = 0; k < 3; k++) implicit[k] = im[k];
    *margin = box.getMargin();
}

// Broadphase proxy AABB of static object s (_worldCollisionRBs order).
void rsref_static_aabb(void* h, int s, float* mn, float* mx) {
    auto* ra = (RefArena*)h;
    auto* px = (btRSBroadphaseProxy*)ra->arena->_worldCollisionRBs[s]->getBroadphaseHandle();
    for (int k = 0; k < 3; k++) { mn[k] = px->m_aabbMin[k]; mx[k] = px->m_aabbMax[k]; }
}

}  // extern "C"

// Wheel w of car i after the last step: ground body code (as in
// rsref_get_manifold, -2000 when not in contact) and the friction impulse.
extern "C" int rsref_wheel_debug(void* h, int car, int w, float* impulse, float* extra) {
    auto* ra = (RefArena*)h;
    Car* c = ra->cars[car];
    const btWheelInfoRL& wi = c->_bulletVehicle.m_wheelInfo[w];
    for (int k = 0; k < 3; k++) impulse[k] = wi.m_impulse[k];
    extra[0] = wi.m_wheelsSuspensionForce;
    extra[1] = wi.m_raycastInfo.m_suspensionLength;
    for (int k = 0; k < 3; k++) {
        extra[2 + k] = wi.m_raycastInfo.m_hardPointWS[k];
        extra[5 + k] = wi.m_raycastInfo.m_contactPointWS[k];
        extra[8 + k] = wi.m_raycastInfo.m_contactNormalWS[k];
    }
    const btCollisionObject* g = (const btCollisionObject*)wi.m_raycastInfo.m_groundObject;
    return g ? body_code(ra, g) : -2000;
}

// Iteration order of Arena::_cars (an unordered_set<Car*>) as indices into
// the creation order; it is the order of the per-car update loops.
extern "C" int rsref_car_order(void* h, int32_t* out) {
    auto* ra = (RefArena*)h;
    int k = 0;
    for (Car* c : ra->arena->_cars) {
        int idx = -1;
        for (size_t i = 0; i < ra->cars.size(); i++)
            if (ra->cars[i] == c) idx = (int)i;
        out[k++] = idx;
    }
    return k;
}

// World transform of car i's hitbox child (rows of the basis, then origin),
// as btCollisionWorld's compound ray test forms it.
extern "C" void rsref_car_child_xf(void* h, int car, float* out) {
    auto* ra = (RefArena*)h;
    Car* c = ra->cars[car];
    btTransform t = c->_rigidBody.getWorldTransform() * c->_compoundShape.getChildTransform(0);
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) out[i * 3 + k] = t.getBasis()[i][k];
    for (int k = 0; k < 3; k++) out[9 + k] = t.getOrigin()[k];
}

// btDefaultVehicleRaycaster::castRay through the arena's world (ignoring car
// `ignore_car`, -1 for none): returns the hit body code, -2000 on no hit.
extern "C" int rsref_world_ray(void* h, const float* from, const float* to, int ignore_car, float* frac, float* normal) {
    auto* ra = (RefArena*)h;
    btVector3 f(from[0], from[1], from[2]), t(to[0], to[1], to[2]);
    Car* c = ra->cars[ignore_car];
    ra->arena->_bulletWorld.updateAabbs();
    btVehicleRaycaster::btVehicleRaycasterResult res;
    const btCollisionObject* o = (const btCollisionObject*)c->_bulletVehicleRaycaster.castRay(f, t, &c->_rigidBody, res);
    if (!o) return -2000;
    *frac = res.m_distFraction;
    for (int k = 0; k < 3; k++) normal[k] = res.m_hitNormalInWorld[k];
    return body_code(ra, o);
}

// btCollisionWorld::rayTestSingle of one car's collision object (its
// compound shape at its world transform), as the world ray test runs it.
extern "C" int rsref_ray_single_car(void* h, const float* from, const float* to, int car, float* frac, float* normal) {
    auto* ra = (RefArena*)h;
    btVector3 f(from[0], from[1], from[2]), t(to[0], to[1], to[2]);
    Car* c = ra->cars[car];
    btTransform ft(btMatrix3x3::getIdentity(), f), tt(btMatrix3x3::getIdentity(), t);
    btCollisionWorld::ClosestRayResultCallback cb(f, t, nullptr);
    btCollisionWorld::rayTestSingle(ft, tt, &c->_rigidBody, c->_rigidBody.getCollisionShape(),
                                    c->_rigidBody.getWorldTransform(), cb);
    if (!cb.hasHit()) return 0;
    btVector3 n = cb.m_hitNormalWorld;
    n.normalize();
    *frac = cb.m_closestHitFraction;
    for (int k = 0; k < 3; k++) normal[k] = n[k];
    return 1;
}

// Events recorded from the reference's callbacks since the last call
// (rsg_event records; `arena` is 0).  Returns the number copied.
extern "C" int rsref_pop_events(void* h, rsg_event* out, int max_out) {
    auto* ra = (RefArena*)h;
    int n = (int)ra->events.size() < max_out ? (int)ra->events.size() : max_out;
    for (int i = 0; i < n; i++) out[i] = ra->events[i];
    ra->events.clear();
    return n;
}

// ---------------------------------------------------------------------------
// BallPredTracker (src/Sim/BallPredTracker/BallPredTracker.h:7-36).

static BallState ball_state_from(const rsg_ball_state* s) { return rsg_amd::FromRsg(*s); }

static void ball_state_to(const BallState& bs, rsg_ball_state* s) { *s = rsg_amd::ToRsg(bs); }

// new BallPredTracker(arena, numPredTicks): predicts from the arena's ball at once.
extern "C" void* rsref_ball_pred_create(void* h, int num_pred_ticks) {
    return new BallPredTracker(((RefArena*)h)->arena, (size_t)num_pred_ticks);
}
extern "C" void rsref_ball_pred_destroy(void* p) { delete (BallPredTracker*)p; }
extern "C" void rsref_ball_pred_update_from_arena(void* p, void* h) {
    ((BallPredTracker*)p)->UpdatePredFromArena(((RefArena*)h)->arena);
}
extern "C" void rsref_ball_pred_update_manual(void* p, const rsg_ball_state* s, int ticks_since) {
    ((BallPredTracker*)p)->UpdatePredManual(ball_state_from(s), ticks_since);
}
extern "C" void rsref_ball_pred_force(void* p, const rsg_ball_state* s) {
    ((BallPredTracker*)p)->ForceUpdateAllPred(ball_state_from(s));
}
// predData: copies min(size, max_out) states and returns the size.
extern "C" int rsref_ball_pred_get(void* p, rsg_ball_state* out, int max_out) {
    auto* t = (BallPredTracker*)p;
    int n = (int)t->predData.size();
    for (int i = 0; i < n && i < max_out; i++) ball_state_to(t->predData[i], &out[i]);
    return n;
}
extern "C" void rsref_ball_pred_for_time(void* p, float pred_time, rsg_ball_state* out) {
    ball_state_to(((BallPredTracker*)p)->GetBallStateForTime(pred_time), out);
}

// ---------------------------------------------------------------------------
// Arena::Serialize / Arena::DeserializeNew (src/Sim/Arena/Arena.cpp:546-644).

// Serialized bytes of the arena; copies them when cap suffices, returns the size.
extern "C" int64_t rsref_serialize(void* h, uint8_t* out, uint64_t cap) {
    DataStreamOut s;
    ((RefArena*)h)->arena->Serialize(s);
    if (out && cap >= s.data.size()) std::memcpy(out, s.data.data(), s.data.size());
    return (int64_t)s.data.size();
}

// A new harness arena from serialized bytes (cars in file order = creation order).
extern "C" void* rsref_deserialize(const uint8_t* data, uint64_t size) {
    try {
        DataStreamIn s;
        s.data.assign(data, data + size);
        auto* ra = new RefArena();
        ra->arena = Arena::DeserializeNew(s);
        ra->arena->_bulletWorld.getSolverInfo().m_timeStep = ra->arena->tickTime;
        ra->bp_cell.push_back(proxy_cell(ra, 0));
        ra->bp_seq.push_back(0);
        // file order: the _carIDMap is keyed by id; recover creation order by id
        std::vector<Car*> cars(ra->arena->_cars.begin(), ra->arena->_cars.end());
        std::sort(cars.begin(), cars.end(), [](Car* a, Car* b) { return a->id < b->id; });
        for (Car* c : cars) {
            ra->cars.push_back(c);
            ra->bp_cell.push_back(proxy_cell(ra, (int)ra->cars.size()));
            ra->bp_seq.push_back(++ra->bp_counter);
        }
        return ra;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

extern "C" uint32_t rsref_last_car_id(void* h) { return ((RefArena*)h)->arena->_lastCarID; }
extern "C" void rsref_get_mutators(void* h, rsg_mutator_config* out) {
    *out = rsg_amd::ToRsg(((RefArena*)h)->arena->GetMutatorConfig());
}
extern "C" void rsref_set_mutators(void* h, const rsg_mutator_config* in) {
    RefArena* ra = (RefArena*)h;
    Arena* a = ra->arena;
    const MutatorConfig& old = a->GetMutatorConfig();
    bool ball_remade = in->ball_radius != old.ballRadius || in->ball_mass != old.ballMass;
    a->SetMutatorConfig(rsg_amd::FromRsg(*in, a->gameMode));
    if (ball_remade) {
        // the new ball body's proxy is registered after every other one
        ra->bp_cell[0] = proxy_cell(ra, 0);
        ra->bp_seq[0] = ++ra->bp_counter;
    }
}
extern "C" int rsref_car_team(void* h, int i) { return (int)((RefArena*)h)->cars[i]->team; }
extern "C" void rsref_get_controls(void* h, int i, rsg_controls* c) { *c = rsg_amd::ToRsg(((RefArena*)h)->cars[i]->controls); }

// ---------------------------------------------------------------------------
// GameEventTracker (src/Sim/GameEventTracker/GameEventTracker.h:44-106) with
// its shot / goal / save callbacks recorded as rsg_event (arena 0).

struct RefTracker {
    GameEventTracker t;
    std::vector<rsg_event> events;
    uint32_t seq = 0;
};

static void tracker_push(RefTracker* rt, Arena* arena, uint32_t type, Car* x, Car* y) {
    rsg_event e{};
    e.type = type;
    e.tick = arena->tickCount;
    e.a = x ? x->id : 0;
    e.b = y ? y->id : 0;
    e.seq = rt->seq++;
    rt->events.push_back(e);
}

extern "C" void* rsref_tracker_create(const rsg_game_event_config* c) {
    auto* rt = new RefTracker();
    GameEventTrackerConfig& k = rt->t.config;
    k.shotMinSpeed = c->shot_min_speed;
    k.shotTouchMinDelay = c->shot_touch_min_delay;
    k.predScoreExtraMargin = c->pred_score_extra_margin;
    k.shotEventCooldown = c->shot_event_cooldown;
    k.shotMinScoreTime = c->shot_min_score_time;
    k.goalMaxTouchTime = c->goal_max_touch_time;
    k.passMaxTouchTime = c->pass_max_touch_time;
    rt->t.autoStateSetDetection = c->auto_state_set_detection != 0;
    rt->t.SetShotCallback([](Arena* a, Car* shooter, Car* passer, void* ui) {
        tracker_push((RefTracker*)ui, a, RSG_EVENT_SHOT, shooter, passer);
    }, rt);
    rt->t.SetGoalCallback([](Arena* a, Car* scorer, Car* passer, void* ui) {
        tracker_push((RefTracker*)ui, a, RSG_EVENT_SCORER, scorer, passer);
    }, rt);
    rt->t.SetSaveCallback([](Arena* a, Car* saver, void* ui) {
        tracker_push((RefTracker*)ui, a, RSG_EVENT_SAVE, saver, nullptr);
    }, rt);
    return rt;
}
extern "C" void rsref_tracker_destroy(void* p) { delete (RefTracker*)p; }
extern "C" void rsref_tracker_update(void* p, void* h) { ((RefTracker*)p)->t.Update(((RefArena*)h)->arena); }
extern "C" void rsref_tracker_reset(void* p) { ((RefTracker*)p)->t.ResetPersistentInfo(); }
extern "C" int rsref_tracker_pop(void* p, rsg_event* out, int max_out) {
    auto* rt = (RefTracker*)p;
    int n = (int)rt->events.size() < max_out ? (int)rt->events.size() : max_out;
    for (int i = 0; i < n; i++) out[i] = rt->events[i];
    rt->events.clear();
    return n;
}
