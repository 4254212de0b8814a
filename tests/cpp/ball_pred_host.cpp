// Host build of the ball-prediction tracker update (rocketsim_amd/csrc/
// rsg_ball_pred.h), for the CPU tests: the code the GPU kernel runs, over
// host arrays, with the vector normalisation of the reference's x86-64 build
// (Bullet's btVector3::normalize: rsqrtss plus one Newton step,
// LinearMath/btVector3.h) so that the results can be compared with the
// reference's BallPredTracker bit for bit.
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

#include <string>
#include <vector>

#include "rsg.h"
#include "../../rocketsim_amd/csrc/rsg_mesh_build.h"
#include "../../rocketsim_amd/csrc/rsg_ball_pred.h"
#include "../../rocketsim_amd/csrc/rsg_api_convert.h"

using namespace rsg;

namespace {

struct HostPred {
    HostArena geo;
    bool soccar = false;
    int n = 0, T = 0;
    std::vector<uint32_t> car, ball, pad, arena;  // the prediction arenas
    std::vector<uint32_t> traj, cur;
    std::vector<int32_t> head, size, last, ticks;
    BatchDev dev{};
};

void stage(HostPred& h, const rsg_ball_state* in) {
    const size_t n = (size_t)h.n;
    for (size_t i = 0; i < n; i++) {
        const rsg_ball_state& b = in[i];
        float f[18];
        std::memcpy(f, b.pos, 12);
        std::memcpy(f + 3, b.rot, 36);
        std::memcpy(f + 12, b.vel, 12);
        std::memcpy(f + 15, b.ang_vel, 12);
        for (int k = 0; k < 18; k++) std::memcpy(&h.cur[k * n + i], &f[k], 4);
        h.cur[18 * n + i] = (uint32_t)b.tick_count_since_update;
        h.cur[19 * n + i] = (uint32_t)(b.tick_count_since_update >> 32);
        std::memcpy(&h.cur[20 * n + i], &b.hs_y_target_dir, 4);
        std::memcpy(&h.cur[21 * n + i], &b.hs_cur_target_speed, 4);
        std::memcpy(&h.cur[22 * n + i], &b.hs_time_since_hit, 4);
    }
}

}  // namespace

extern "C" {

// A host predictor of n trackers; cmf = the Soccar mesh file (NULL for THE_VOID).
void* bp_create(const void* cmf, uint64_t cmf_size, int n, int T, float tick_rate) {
    auto* h = new HostPred();
    if (cmf) {
        const void* files[1] = {cmf};
        if (!build_host_arena(files, &cmf_size, 1, 0, h->geo).empty()) { delete h; return nullptr; }
        h->soccar = true;
    }
    h->n = n;
    h->T = T;
    h->car.assign((size_t)CAR_WORDS * n, 0);
    h->ball.assign((size_t)BALL_WORDS * n, 0);
    h->pad.assign((size_t)PAD_WORDS * PAD_NUM * n, 0);
    h->arena.assign((size_t)ARENA_WORDS * n, 0);
    h->traj.assign((size_t)T * RSG_PRED_WORDS * n, 0);
    h->cur.assign((size_t)RSG_PRED_WORDS * n, 0);
    h->head.assign(n, 0);
    h->size.assign(n, 0);
    h->last.assign(n, 0);
    h->ticks.assign(n, 0);
    BatchDev& d = h->dev;
    d.car = h->car.data();
    d.ball = h->ball.data();
    d.pad = h->pad.data();
    d.arena = h->arena.data();
    d.n = n;
    d.max_cars = 1;
    d.mode = cmf ? RSG_MODE_SOCCAR : RSG_MODE_THE_VOID;
    d.tick_time = 1 / tick_rate;
    for (int a = 0; a < n; a++) init_arena_record(d, a, 0, nullptr, nullptr);
    return h;
}

void bp_destroy(void* p) { delete (HostPred*)p; }

// kind 0 = force (cur), 1 = manual (cur, ticks), 2 = from arenas whose balls
// and tick counts are given (cur, arena_ticks).
int bp_update(void* p, int kind, const rsg_ball_state* cur, const int32_t* ticks, const uint64_t* arena_ticks) {
    HostPred& h = *(HostPred*)p;
    PredDev d;
    d.traj = h.traj.data();
    d.head = h.head.data();
    d.size = h.size.data();
    d.last = h.last.data();
    d.cur = h.cur.data();
    d.ticks = h.ticks.data();
    d.T = h.T;
    std::vector<uint32_t> sb, sa;
    BatchDev src = h.dev;
    if (kind == PRED_FROM_BATCH) {
        sb.assign((size_t)BALL_WORDS * h.n, 0);
        sa.assign((size_t)ARENA_WORDS * h.n, 0);
        src.ball = sb.data();
        src.arena = sa.data();
        for (int i = 0; i < h.n; i++) {
            ball_from_api(ball_ref(src, i), cur[i]);
            arena_ref(src, i).su64(AW_TICK, arena_ticks[i]);
        }
    } else {
        stage(h, cur);
        if (kind == PRED_MANUAL)
            for (int i = 0; i < h.n; i++) h.ticks[i] = ticks[i];
    }
    const ArenaGeom mv = h.geo.geom();
    auto* A = new ArenaLocal;
    auto* ws = new StepSpace;
    for (int i = 0; i < h.n; i++) ball_pred_update(h.dev, h.soccar ? &mv : nullptr, d, kind, src, i, *A, ws);
    delete ws;
    delete A;
    return 0;
}

// predData of one tracker; returns its size.
int bp_get(void* p, int tracker, rsg_ball_state* out) {
    HostPred& h = *(HostPred*)p;
    const size_t n = (size_t)h.n;
    for (int t = 0; t < h.size[tracker]; t++) {
        const uint32_t* r = h.traj.data() + (size_t)((h.head[tracker] + t) % h.T) * RSG_PRED_WORDS * n + tracker;
        rsg_ball_state& o = out[t];
        std::memset(&o, 0, sizeof(o));
        float f[18];
        for (int k = 0; k < 18; k++) std::memcpy(&f[k], &r[k * n], 4);
        std::memcpy(o.pos, f, 12);
        std::memcpy(o.rot, f + 3, 36);
        std::memcpy(o.vel, f + 12, 12);
        std::memcpy(o.ang_vel, f + 15, 12);
        o.tick_count_since_update = (uint64_t)r[18 * n] | ((uint64_t)r[19 * n] << 32);
        std::memcpy(&o.hs_y_target_dir, &r[20 * n], 4);
        std::memcpy(&o.hs_cur_target_speed, &r[21 * n], 4);
        std::memcpy(&o.hs_time_since_hit, &r[22 * n], 4);
    }
    return h.size[tracker];
}

}  // extern "C"
