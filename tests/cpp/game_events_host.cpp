// Host build of the game-event tracker update (rocketsim_amd/csrc/
// rsg_game_events.h) for the CPU tests: the code the GPU kernel runs, over
// host arrays, driven through recorded arena states.
#include <algorithm>
#include <vector>

#include "rsg.h"
#include "../../rocketsim_amd/csrc/rsg_game_events.h"
#include "../../rocketsim_amd/csrc/rsg_api_convert.h"

using namespace rsg;

namespace {

struct HostGE {
    int n = 0, cars = 0;
    std::vector<uint32_t> car, ball, pad, arena;
    std::vector<float> cooldown;
    std::vector<uint32_t> flags, seq;
    std::vector<uint64_t> last;
    std::vector<rsg_event> ev;
    uint32_t count = 0;
    rsg_game_event_config cfg{};
    BatchDev d{};
};

}  // namespace

extern "C" {

void* ge_create(int n, int cars, const uint8_t* teams, const uint8_t* presets, const rsg_game_event_config* cfg,
                float tick_rate, int mode) {
    auto* h = new HostGE();
    h->n = n;
    h->cars = cars;
    h->car.assign((size_t)CAR_WORDS * cars * n, 0);
    h->ball.assign((size_t)BALL_WORDS * n, 0);
    h->pad.assign((size_t)PAD_WORDS * PAD_NUM * n, 0);
    h->arena.assign((size_t)ARENA_WORDS * n, 0);
    h->cooldown.assign(n, 0.f);
    h->flags.assign(n, 0);
    h->seq.assign(n, 0);
    h->last.assign(n, 0);
    h->ev.resize(1 << 16);
    h->cfg = *cfg;
    BatchDev& d = h->d;
    d.car = h->car.data();
    d.ball = h->ball.data();
    d.pad = h->pad.data();
    d.arena = h->arena.data();
    d.n = n;
    d.max_cars = cars;
    d.mode = mode;
    d.tick_time = 1 / tick_rate;
    for (int a = 0; a < n; a++) init_arena_record(d, a, cars, teams + (size_t)a * cars, presets + (size_t)a * cars);
    return h;
}

void ge_destroy(void* p) { delete (HostGE*)p; }

// Sets every arena's state (cars in creation order, [n * cars]; balls [n];
// tick counts [n]; car update orders [n * cars]) and runs the update.
int ge_update(void* p, const rsg_car_state* cars, const rsg_ball_state* balls, const uint64_t* ticks,
              const int32_t* order) {
    HostGE& h = *(HostGE*)p;
    for (int a = 0; a < h.n; a++) {
        for (int i = 0; i < h.cars; i++) car_from_api(car_ref(h.d, a, i), cars[(size_t)a * h.cars + i]);
        ball_from_api(ball_ref(h.d, a), balls[a]);
        CarRef ar = arena_ref(h.d, a);
        ar.su64(AW_TICK, ticks[a]);
        uint32_t ow = 0;
        for (int k = 0; k < h.cars; k++) ow |= (uint32_t)order[(size_t)a * h.cars + k] << (4 * k);
        ar.u(AW_CAR_ORDER) = ow;
    }
    GameEventDev g;
    g.cooldown = h.cooldown.data();
    g.flags = h.flags.data();
    g.last_count = h.last.data();
    g.seq = h.seq.data();
    g.ev_buf = h.ev.data();
    g.ev_count = &h.count;
    g.ev_cap = (uint32_t)h.ev.size();
    for (int a = 0; a < h.n; a++) game_events_update(h.d, g, h.cfg, a);
    return 0;
}

int ge_pop(void* p, rsg_event* out, int max_out) {
    HostGE& h = *(HostGE*)p;
    int n = (int)std::min<uint32_t>(h.count, (uint32_t)h.ev.size());
    std::sort(h.ev.begin(), h.ev.begin() + n, [](const rsg_event& x, const rsg_event& y) {
        return x.arena != y.arena ? x.arena < y.arena : x.seq < y.seq;
    });
    n = n < max_out ? n : max_out;
    std::copy(h.ev.begin(), h.ev.begin() + n, out);
    h.count = 0;
    return n;
}

}  // extern "C"
