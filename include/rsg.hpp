// rsg.hpp -- C++ facade over the C ABI of librsg.so (include/rsg.h).
//
// Mirrors the reference's C++ object API for the stepping path
// (src/Sim/Arena/Arena.h, src/Sim/Car/Car.h, src/Sim/Ball/Ball.h) so that C++
// host code keeps its shape:
//   * ArenaBatch  -- N independent arenas stepped together on one GPU;
//   * Arena       -- a batch of one with the reference's method names
//                    (Create, Step, GetCarState/SetCarState, ...), so that
//                    `arena->Step(n)` code ports line for line.
// Errors from the ABI are rethrown as std::runtime_error, as the reference's
// RS_ERR_CLOSE does (src/Framework.h:67-72).  Header-only; link -lrsg.
#ifndef RSG_HPP
#define RSG_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rsg.h"

namespace rsg_amd {

inline void check(int rc) {
    if (rc < 0) throw std::runtime_error(std::string("rsg: ") + rsg_last_error());
}

enum class GameMode : int { SOCCAR = RSG_MODE_SOCCAR, THE_VOID = RSG_MODE_THE_VOID };
enum class Team : uint8_t { BLUE = RSG_TEAM_BLUE, ORANGE = RSG_TEAM_ORANGE };

using CarState = rsg_car_state;
using BallState = rsg_ball_state;
using BoostPadState = rsg_pad_state;
using CarControls = rsg_controls;

// RocketSim::InitFromMem (src/RocketSim.h:32) for the Soccar collision mesh.
inline void InitFromMem(const void* cmf_data, uint64_t size) { check(rsg_init_mesh(cmf_data, size)); }

struct CarSpec {
    Team team;
    int preset;  // RSG_PRESET_*
};

struct BatchConfig {
    GameMode mode = GameMode::SOCCAR;
    float tick_rate = 120.f;
    uint64_t seed = 0;
    bool auto_reset = false;      // kickoff reset after a goal
    bool synth_controls = false;  // draw controls on the device (include/rsg_workload.h)
};

class ArenaBatch {
public:
    // One vector of cars per arena (Arena::Create + AddCar per arena).
    ArenaBatch(const std::vector<std::vector<CarSpec>>& arenas, const BatchConfig& cfg = {}) {
        std::vector<int> counts;
        std::vector<uint8_t> teams, presets;
        for (const auto& a : arenas) {
            counts.push_back((int)a.size());
            for (const auto& c : a) {
                teams.push_back((uint8_t)c.team);
                presets.push_back((uint8_t)c.preset);
            }
        }
        rsg_batch_desc d{};
        d.n_arenas = (int)arenas.size();
        d.mode = (int)cfg.mode;
        d.tick_rate = cfg.tick_rate;
        d.seed = cfg.seed;
        d.auto_reset = cfg.auto_reset ? 1 : 0;
        d.synth_controls = cfg.synth_controls ? 1 : 0;
        d.cars_per_arena = counts.data();
        d.car_team = teams.data();
        d.car_preset = presets.data();
        h_ = rsg_batch_create(&d);
        if (!h_) throw std::runtime_error(std::string("rsg: ") + rsg_last_error());
    }
    ~ArenaBatch() { if (h_) rsg_batch_destroy(h_); }
    ArenaBatch(const ArenaBatch&) = delete;
    ArenaBatch& operator=(const ArenaBatch&) = delete;
    ArenaBatch(ArenaBatch&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }

    int NumArenas() const { return rsg_batch_num_arenas(h_); }
    int NumCars() const { return rsg_batch_num_cars(h_); }
    int CarOffset(int arena) const { int o = rsg_batch_car_offset(h_, arena); check(o); return o; }

    // Arena::Step for every arena; `stream` is a hipStream_t (nullptr = default).
    void Step(int ticks = 1, void* stream = nullptr) { check(rsg_batch_step(h_, ticks, stream)); }
    void Sync() { check(rsg_batch_sync(h_)); }

    std::vector<CarState> GetCarStates() { std::vector<CarState> v(NumCars()); check(rsg_batch_get_cars(h_, v.data())); return v; }
    void SetCarStates(const std::vector<CarState>& v) { size_check(v.size(), NumCars()); check(rsg_batch_set_cars(h_, v.data())); }
    std::vector<BallState> GetBallStates() { std::vector<BallState> v(NumArenas()); check(rsg_batch_get_balls(h_, v.data())); return v; }
    void SetBallStates(const std::vector<BallState>& v) { size_check(v.size(), NumArenas()); check(rsg_batch_set_balls(h_, v.data())); }
    std::vector<BoostPadState> GetPadStates() {
        std::vector<BoostPadState> v((size_t)NumArenas() * RSG_NUM_PADS);
        check(rsg_batch_get_pads(h_, v.data()));
        return v;
    }
    void SetPadStates(const std::vector<BoostPadState>& v) {
        size_check(v.size(), (size_t)NumArenas() * RSG_NUM_PADS);
        check(rsg_batch_set_pads(h_, v.data()));
    }
    std::vector<CarControls> GetControls() { std::vector<CarControls> v(NumCars()); check(rsg_batch_get_controls(h_, v.data())); return v; }
    void SetControls(const std::vector<CarControls>& v) { size_check(v.size(), NumCars()); check(rsg_batch_set_controls(h_, v.data())); }
    std::vector<uint64_t> GetTickCounts() { std::vector<uint64_t> v(NumArenas()); check(rsg_batch_get_tick_counts(h_, v.data())); return v; }
    void SetTickCounts(const std::vector<uint64_t>& v) { size_check(v.size(), NumArenas()); check(rsg_batch_set_tick_counts(h_, v.data())); }
    void ResetToRandomKickoff(int arena = -1, uint64_t seed = 0) { check(rsg_batch_reset_kickoff(h_, arena, seed)); }
    std::vector<uint8_t> IsBallScored() { std::vector<uint8_t> v(NumArenas()); check(rsg_batch_is_ball_scored(h_, v.data())); return v; }
    void SetLaunch(int arenas_per_block) { check(rsg_batch_set_launch(h_, arenas_per_block)); }

    // Arena::Serialize (Arena.cpp:546-573) of one arena: the reference's bytes
    // plus this library's extension block.
    std::vector<uint8_t> Serialize(int arena) {
        rsg_arena_snapshot s;
        check(rsg_batch_get_snapshot(h_, arena, &s));
        int64_t n = rsg_snapshot_serialize(&s, nullptr, 0);
        if (n < 0) check((int)n);
        std::vector<uint8_t> out((size_t)n);
        if (rsg_snapshot_serialize(&s, out.data(), out.size()) < 0) check(-1);
        return out;
    }
    // Arena::DeserializeNew (Arena.cpp:575-644) into an arena slot of the
    // same mode, tick rate, car count, teams and presets.
    void Deserialize(int arena, const std::vector<uint8_t>& bytes) {
        rsg_arena_snapshot s;
        check(rsg_snapshot_deserialize(bytes.data(), bytes.size(), &s));
        check(rsg_batch_set_snapshot(h_, arena, &s));
    }

    rsg_batch* handle() const { return h_; }

private:
    static void size_check(size_t got, size_t want) {
        if (got != want) throw std::runtime_error("rsg: array size " + std::to_string(got) + " != " + std::to_string(want));
    }
    rsg_batch* h_ = nullptr;
};

// A single arena with the reference's method names (src/Sim/Arena/Arena.h),
// backed by a batch of one.  Cars are fixed at creation, as a batch's are.
class Arena {
public:
    static std::unique_ptr<Arena> Create(GameMode mode, const std::vector<CarSpec>& cars, uint64_t seed = 0) {
        BatchConfig cfg;
        cfg.mode = mode;
        cfg.seed = seed;
        return std::unique_ptr<Arena>(new Arena(ArenaBatch({cars}, cfg)));
    }
    void Step(int ticksToSimulate = 1) { b_.Step(ticksToSimulate); b_.Sync(); }
    CarState GetCarState(int car) { return b_.GetCarStates().at((size_t)car); }
    void SetCarState(int car, const CarState& s) { auto v = b_.GetCarStates(); v.at((size_t)car) = s; b_.SetCarStates(v); }
    void SetCarControls(int car, const CarControls& c) { auto v = b_.GetControls(); v.at((size_t)car) = c; b_.SetControls(v); }
    BallState GetBallState() { return b_.GetBallStates()[0]; }
    void SetBallState(const BallState& s) { b_.SetBallStates({s}); }
    std::vector<BoostPadState> GetBoostPadStates() { return b_.GetPadStates(); }
    uint64_t TickCount() { return b_.GetTickCounts()[0]; }
    void ResetToRandomKickoff(int seed = -1) { b_.ResetToRandomKickoff(0, (uint64_t)(int64_t)seed); }
    bool IsBallScored() { return b_.IsBallScored()[0] != 0; }
    std::vector<uint8_t> Serialize() { return b_.Serialize(0); }
    void Deserialize(const std::vector<uint8_t>& bytes) { b_.Deserialize(0, bytes); }
    ArenaBatch& Batch() { return b_; }

private:
    explicit Arena(ArenaBatch&& b) : b_(std::move(b)) {}
    ArenaBatch b_;
};

}  // namespace rsg_amd

#endif  // RSG_HPP
