// A reference-typed program driving the batched GPU step through the C++
// facade (include/rsg.hpp) and the type adapter (include/rsg_rocketsim.hpp),
// checked against the reference it replaces.  TEST INFRASTRUCTURE: built by
// oracle/Makefile into oracle/_ref/adapter_demo, linked to the reference
// (oracle/_ref/librsref.so) and to librsg.so; run by
// tests/test_cpp_facade.py on a GPU machine.
//
//   adapter_demo <soccar.cmf> <ticks>
//
// The program keeps its own RocketSim::Arena (2v2 Soccar, kickoff, random
// controls redrawn every few ticks).  Every tick it moves the reference's
// cars, ball and boost pads into an rsg_amd::Arena with the adapter's ToRsg,
// copies the reference cars' `controls` over, steps both one tick, brings
// the GPU's records back as RocketSim::CarState / BallState with FromRsg, and
// compares them with the reference's own GetState() (positions, velocities
// 1e-3 relative, rotations 1e-4, flags and boost exactly as the parity tests
// state them).  Prints "adapter_demo: N ticks match" and exits 0, or the
// first mismatch and exits 1.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <random>
#include <vector>

#include "RocketSim.h"
#include "rsg.hpp"
#include "rsg_rocketsim.hpp"

using namespace RocketSim;

namespace {

bool close_rel(const Vec& a, const Vec& b, float floor) {
    const float da[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
    const float ra[3] = {a.x, a.y, a.z};
    for (int k = 0; k < 3; k++)
        if (!(std::fabs(da[k]) <= 1e-3f * std::fmax(std::fabs(ra[k]), floor))) return false;
    return true;
}

bool close_rot(const RotMat& a, const RotMat& b) {
    const Vec ca[3] = {a.forward, a.right, a.up}, cb[3] = {b.forward, b.right, b.up};
    for (int k = 0; k < 3; k++)
        if (!(std::fabs(ca[k].x - cb[k].x) <= 1e-4f && std::fabs(ca[k].y - cb[k].y) <= 1e-4f &&
              std::fabs(ca[k].z - cb[k].z) <= 1e-4f))
            return false;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: adapter_demo <soccar.cmf> <ticks>\n");
        return 2;
    }
    std::ifstream f(argv[1], std::ios::binary);
    std::vector<uint8_t> cmf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const int ticks = std::atoi(argv[2]);
    FileData fd(cmf.begin(), cmf.end());
    RocketSim::InitFromMem({{GameMode::SOCCAR, {fd}}}, true);
    rsg_amd::InitFromMem(cmf.data(), cmf.size());

    Arena* ref = Arena::Create(GameMode::SOCCAR);
    std::vector<Car*> cars;
    for (int i = 0; i < 4; i++) cars.push_back(ref->AddCar(i % 2 ? Team::ORANGE : Team::BLUE, CAR_CONFIG_OCTANE));
    ref->ResetToRandomKickoff(1234);
    auto gpu = rsg_amd::Arena::Create(rsg_amd::GameMode::SOCCAR,
                                      {{rsg_amd::Team::BLUE, RSG_PRESET_OCTANE}, {rsg_amd::Team::ORANGE, RSG_PRESET_OCTANE},
                                       {rsg_amd::Team::BLUE, RSG_PRESET_OCTANE}, {rsg_amd::Team::ORANGE, RSG_PRESET_OCTANE}});
    // the reference iterates its cars in heap order; the batch follows it
    std::vector<int32_t> order;
    for (Car* c : ref->_cars)
        for (int i = 0; i < 4; i++)
            if (cars[i] == c) order.push_back(i);
    rsg_amd::check(rsg_batch_set_car_order(gpu->Batch().handle(), order.data()));

    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    std::vector<int> hold(4, 0);
    for (int t = 0; t < ticks; t++) {
        for (int i = 0; i < 4; i++) {
            if (hold[i]-- <= 0) {
                CarControls c;
                c.throttle = u(rng); c.steer = u(rng); c.pitch = u(rng); c.yaw = u(rng); c.roll = u(rng);
                c.jump = u(rng) > 0.6f; c.boost = u(rng) > 0; c.handbrake = u(rng) > 0.7f;
                cars[i]->controls = c;
                hold[i] = 5 + (int)(20 * (u(rng) + 1));
            }
        }
        // reference -> GPU, through the adapter
        std::vector<rsg_car_state> cs;
        std::vector<rsg_amd::Car*> gcars = gpu->GetCars();
        for (int i = 0; i < 4; i++) {
            cs.push_back(rsg_amd::ToRsg(*cars[i]));
            gcars[i]->controls = rsg_amd::ToRsg(cars[i]->controls);  // what Step hands to the batch
        }
        // broadphase records from the batch's own view of the same poses
        auto cur = gpu->Batch().GetCarStates();
        for (int i = 0; i < 4; i++) { cs[i].bp_cell = cur[i].bp_cell; cs[i].bp_seq = cur[i].bp_seq; }
        gpu->Batch().SetCarStates(cs);
        rsg_ball_state bs = rsg_amd::ToRsg(ref->ball->GetState());
        auto cur_ball = gpu->GetBallState();
        bs.bp_cell = cur_ball.bp_cell;
        bs.bp_seq = cur_ball.bp_seq;
        gpu->SetBallState(bs);
        std::vector<rsg_pad_state> pads;
        for (BoostPad* p : ref->_boostPads) pads.push_back(rsg_amd::ToRsg(p->GetState()));
        gpu->Batch().SetPadStates(pads);
        gpu->Batch().SetTickCounts({ref->tickCount});

        ref->Step(1);
        gpu->Step(1);

        // GPU -> reference types, compared with the reference's own
        auto out = gpu->Batch().GetCarStates();
        for (int i = 0; i < 4; i++) {
            CarState a = cars[i]->GetState(), b = rsg_amd::FromRsg(out[i]);
            bool ok = close_rel(a.pos, b.pos, 1.f) && close_rel(a.vel, b.vel, 1.f) && close_rel(a.angVel, b.angVel, 1e-2f) &&
                      close_rot(a.rotMat, b.rotMat) && a.isOnGround == b.isOnGround && a.hasJumped == b.hasJumped &&
                      a.hasFlipped == b.hasFlipped && a.isDemoed == b.isDemoed && a.isSupersonic == b.isSupersonic &&
                      std::fabs(a.boost - b.boost) <= 1e-3f * std::fmax(1.f, a.boost);
            if (!ok) {
                std::printf("adapter_demo: tick %d car %d differs: ref pos (%g %g %g) vel (%g %g %g), "
                            "gpu pos (%g %g %g) vel (%g %g %g)\n", t, i, a.pos.x, a.pos.y, a.pos.z, a.vel.x, a.vel.y,
                            a.vel.z, b.pos.x, b.pos.y, b.pos.z, b.vel.x, b.vel.y, b.vel.z);
                return 1;
            }
        }
        BallState ra = ref->ball->GetState(), rb = rsg_amd::FromRsg(gpu->GetBallState());
        if (!(close_rel(ra.pos, rb.pos, 1.f) && close_rel(ra.vel, rb.vel, 1.f))) {
            std::printf("adapter_demo: tick %d ball differs\n", t);
            return 1;
        }
    }
    std::printf("adapter_demo: %d ticks match\n", ticks);
    delete ref;
    return 0;
}
