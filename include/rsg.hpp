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
#include <optional>
#include <vector>

#include "rsg.h"

namespace rsg_amd {

inline void check(int rc) {
    if (rc < 0) throw std::runtime_error(std::string("rsg: ") + rsg_last_error());
}

enum class GameMode : int {
    SOCCAR = RSG_MODE_SOCCAR, HOOPS = RSG_MODE_HOOPS, HEATSEEKER = RSG_MODE_HEATSEEKER, THE_VOID = RSG_MODE_THE_VOID
};
enum class Team : uint8_t { BLUE = RSG_TEAM_BLUE, ORANGE = RSG_TEAM_ORANGE };

using CarState = rsg_car_state;
using BallState = rsg_ball_state;
using BoostPadState = rsg_pad_state;
using CarControls = rsg_controls;
using MutatorConfig = rsg_mutator_config;

// MutatorConfig(GameMode) (src/Sim/MutatorConfig/MutatorConfig.cpp:5-40).
inline MutatorConfig DefaultMutatorConfig(GameMode mode) {
    MutatorConfig m;
    check(rsg_mutator_config_default((int)mode, &m));
    return m;
}

// RocketSim::InitFromMem (src/RocketSim.h:32) for the Soccar collision mesh.
inline void InitFromMem(const void* cmf_data, uint64_t size) { check(rsg_init_mesh(cmf_data, size)); }
// RocketSim::InitFromMem with a mode's list of mesh files, in order (the
// Soccar arena by default; GameMode::HOOPS for the Hoops arena).
inline void InitFromMem(const std::vector<std::vector<uint8_t>>& cmf_files, GameMode mode = GameMode::SOCCAR) {
    std::vector<const void*> data;
    std::vector<uint64_t> sizes;
    for (const auto& f : cmf_files) {
        data.push_back(f.data());
        sizes.push_back(f.size());
    }
    check(rsg_init_arena_meshes((int)mode, data.data(), sizes.data(), (int)data.size()));
}

using CarConfig = rsg_car_config;

// The reference's CAR_CONFIG_* presets (CarConfig.cpp:93-100).
inline CarConfig PresetCarConfig(int preset) {
    CarConfig c{};
    check(rsg_car_config_preset(preset, &c));
    return c;
}

struct CarSpec {
    Team team;
    int preset = RSG_PRESET_OCTANE;  // RSG_PRESET_*
    std::optional<CarConfig> config;  // any CarConfig, in place of the preset
};

struct BatchConfig {
    GameMode mode = GameMode::SOCCAR;
    float tick_rate = 120.f;
    uint64_t seed = 0;
    bool auto_reset = false;      // kickoff reset after a goal
    bool synth_controls = false;  // draw controls on the device (include/rsg_workload.h)
    bool ball_rotation = false;   // ArenaConfig::noBallRot == false: the ball's pose rotates
    std::vector<rsg_boost_pad_config> custom_pads;  // ArenaConfig::customBoostPads (empty: the mode's)
    // every arena's MutatorConfig, set before its cars are added (none: the
    // mode's defaults)
    std::optional<MutatorConfig> mutators;
};

class ArenaBatch {
public:
    // One vector of cars per arena (Arena::Create + AddCar per arena).
    ArenaBatch(const std::vector<std::vector<CarSpec>>& arenas, const BatchConfig& cfg = {}) {
        std::vector<int> counts;
        std::vector<uint8_t> teams, presets;
        std::vector<CarConfig> configs;
        bool custom = false;
        for (const auto& a : arenas) {
            counts.push_back((int)a.size());
            for (const auto& c : a) {
                teams.push_back((uint8_t)c.team);
                presets.push_back((uint8_t)c.preset);
                custom = custom || c.config.has_value();
            }
        }
        if (custom)
            for (const auto& a : arenas)
                for (const auto& c : a) configs.push_back(c.config ? *c.config : PresetCarConfig(c.preset));
        rsg_batch_desc d{};
        d.n_arenas = (int)arenas.size();
        d.mode = (int)cfg.mode;
        d.tick_rate = cfg.tick_rate;
        d.seed = cfg.seed;
        d.auto_reset = cfg.auto_reset ? 1 : 0;
        d.synth_controls = cfg.synth_controls ? 1 : 0;
        d.ball_rotation = cfg.ball_rotation ? 1 : 0;
        d.custom_pads = cfg.custom_pads.empty() ? nullptr : cfg.custom_pads.data();
        d.n_custom_pads = (int)cfg.custom_pads.size();
        d.cars_per_arena = counts.data();
        d.car_team = teams.data();
        d.car_preset = presets.data();
        d.mutators = cfg.mutators ? &*cfg.mutators : nullptr;
        d.car_configs = custom ? configs.data() : nullptr;
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
    // Arena::SetMutatorConfig / GetMutatorConfig on one arena, or all (-1).
    void SetMutatorConfig(const MutatorConfig& m, int arena = -1) { check(rsg_batch_set_mutator_config(h_, arena, &m)); }
    MutatorConfig GetMutatorConfig(int arena = 0) {
        MutatorConfig m;
        check(rsg_batch_get_mutator_config(h_, arena, &m));
        return m;
    }

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

class Arena;

// A car of an Arena (src/Sim/Car/Car.h): `controls` is read at the next Step,
// as the reference's public Car::controls is.
class Car {
public:
    CarControls controls{};
    uint32_t id = 0;
    Team team = Team::BLUE;
    int preset = RSG_PRESET_OCTANE;  // RSG_PRESET_CUSTOM when `config` is none of the presets
    CarConfig config{};
    inline CarState GetState();
    inline void SetState(const CarState& s);

private:
    friend class Arena;
    Arena* arena_ = nullptr;
};

// The ball of an Arena (src/Sim/Ball/Ball.h).
class Ball {
public:
    inline BallState GetState();
    inline void SetState(const BallState& s);

private:
    friend class Arena;
    Arena* arena_ = nullptr;
};

// A single arena with the reference's object API (src/Sim/Arena/Arena.h):
// Create, AddCar / RemoveCar / GetCar / GetCars, car->controls,
// car->GetState / SetState, ball->GetState / SetState, Step, Clone, kickoff
// and snapshots, backed by a batch of one.  A batch's car count is fixed, so
// adding or removing a car rebuilds the batch and carries every other state
// over (ids included).  The index-based methods of earlier versions stay.
class Arena {
public:
    static std::unique_ptr<Arena> Create(GameMode mode, uint64_t seed = 0) {
        return std::unique_ptr<Arena>(new Arena(mode, seed, {}));
    }
    static std::unique_ptr<Arena> Create(GameMode mode, const std::vector<CarSpec>& cars, uint64_t seed = 0) {
        return std::unique_ptr<Arena>(new Arena(mode, seed, cars));
    }

    // Arena::AddCar (Arena.cpp:50-62): the new car gets the next id and a
    // kickoff spawn slot (the reference respawns it at a random respawn slot).
    Car* AddCar(Team team, int preset = RSG_PRESET_OCTANE) {
        std::vector<CarSpec> specs = specs_;
        specs.push_back({team, preset, std::nullopt});
        rebuild(specs, -1);
        return cars_.back().get();
    }
    // Arena::AddCar(team, config) with any CarConfig (Arena.h:71).
    Car* AddCar(Team team, const CarConfig& config) {
        std::vector<CarSpec> specs = specs_;
        specs.push_back({team, RSG_PRESET_OCTANE, config});
        rebuild(specs, -1);
        return cars_.back().get();
    }
    // Arena::RemoveCar (Arena.cpp:80-100): false when no car has that id.
    bool RemoveCar(uint32_t id) {
        for (size_t i = 0; i < cars_.size(); i++) {
            if (cars_[i]->id != id) continue;
            std::vector<CarSpec> specs = specs_;
            specs.erase(specs.begin() + (long)i);
            rebuild(specs, (int)i);
            return true;
        }
        return false;
    }
    Car* GetCar(uint32_t id) {
        for (auto& c : cars_)
            if (c->id == id) return c.get();
        return nullptr;
    }
    std::vector<Car*> GetCars() {
        std::vector<Car*> v;
        for (auto& c : cars_) v.push_back(c.get());
        return v;
    }
    Ball* ball = &ball_;

    // Arena::Step: every car's `controls` go to the batch, then `ticksToSimulate` ticks.
    void Step(int ticksToSimulate = 1) {
        if (!cars_.empty()) {
            std::vector<CarControls> c;
            for (auto& car : cars_) c.push_back(car->controls);
            b_->SetControls(c);
        }
        b_->Step(ticksToSimulate);
        b_->Sync();
    }
    // Arena::Clone (Arena.cpp:646-687): a new arena in the same state, its
    // mutator config included (the reference's Clone leaves the new arena at
    // its mode's defaults).  The copy goes through the full snapshot, which
    // keeps every car's config whole (the byte format drops threeWheels).
    std::unique_ptr<Arena> Clone() {
        std::unique_ptr<Arena> a(new Arena(mode_, seed_, specs_));
        rsg_arena_snapshot snap;
        check(rsg_batch_get_snapshot(b_->handle(), 0, &snap));
        check(rsg_batch_set_snapshot(a->b_->handle(), 0, &snap));
        for (size_t i = 0; i < cars_.size(); i++) a->cars_[i]->controls = cars_[i]->controls;
        return a;
    }

    CarState GetCarState(int car) { return b_->GetCarStates().at((size_t)car); }
    void SetCarState(int car, const CarState& s) { auto v = b_->GetCarStates(); v.at((size_t)car) = s; b_->SetCarStates(v); }
    void SetCarControls(int car, const CarControls& c) { cars_.at((size_t)car)->controls = c; }
    BallState GetBallState() { return b_->GetBallStates()[0]; }
    void SetBallState(const BallState& s) { b_->SetBallStates({s}); }
    std::vector<BoostPadState> GetBoostPadStates() { return b_->GetPadStates(); }
    uint64_t TickCount() { return b_->GetTickCounts()[0]; }
    void ResetToRandomKickoff(int seed = -1) { b_->ResetToRandomKickoff(0, (uint64_t)(int64_t)seed); }
    // Arena::SetMutatorConfig / GetMutatorConfig (Arena.h:52-53).  Cars added
    // or removed later rebuild the batch with this config given at creation,
    // so every car gets its gravity and the car mass's default body (the
    // reference leaves cars that existed at SetMutatorConfig with the mass it
    // gave them).
    void SetMutatorConfig(const MutatorConfig& m) { b_->SetMutatorConfig(m, 0); }
    MutatorConfig GetMutatorConfig() { return b_->GetMutatorConfig(0); }
    bool IsBallScored() { return b_->IsBallScored()[0] != 0; }
    std::vector<uint8_t> Serialize() { return b_->Serialize(0); }
    void Deserialize(const std::vector<uint8_t>& bytes) { b_->Deserialize(0, bytes); }
    ArenaBatch& Batch() { return *b_; }

private:
    friend class Car;
    friend class Ball;
    Arena(GameMode mode, uint64_t seed, const std::vector<CarSpec>& cars) : mode_(mode), seed_(seed) {
        ball_.arena_ = this;
        rebuild(cars, -2);
    }
    int slot_of(const Car* c) const {
        for (size_t i = 0; i < cars_.size(); i++)
            if (cars_[i].get() == c) return (int)i;
        throw std::runtime_error("rsg: car does not belong to this arena");
    }
    // A new batch of one for `specs`, carrying the old batch's state over:
    // removed < 0 adds cars at the end (-1) or starts fresh (-2); otherwise
    // car `removed` is dropped and the others keep their ids (through a
    // snapshot of the old batch).
    void rebuild(const std::vector<CarSpec>& specs, int removed) {
        BatchConfig cfg;
        cfg.mode = mode_;
        cfg.seed = seed_;
        if (b_) cfg.mutators = b_->GetMutatorConfig(0);
        auto nb = std::unique_ptr<ArenaBatch>(new ArenaBatch({specs}, cfg));
        std::vector<std::unique_ptr<Car>> cars;
        if (removed == -1 && b_) {
            auto old_cars = b_->GetCarStates();
            auto cur = nb->GetCarStates();
            for (size_t i = 0; i < old_cars.size(); i++) cur[i] = old_cars[i];
            nb->SetCarStates(cur);
            nb->SetBallStates(b_->GetBallStates());
            nb->SetPadStates(b_->GetPadStates());
            nb->SetTickCounts(b_->GetTickCounts());
        } else if (removed >= 0 && b_) {
            rsg_arena_snapshot snap;
            check(rsg_batch_get_snapshot(b_->handle(), 0, &snap));
            const uint32_t gone = cars_[(size_t)removed]->id;
            int k = 0;
            for (int i = 0; i < snap.n_cars; i++) {
                if (snap.car_id[i] == gone) continue;
                snap.car_id[k] = snap.car_id[i];
                snap.car_team[k] = snap.car_team[i];
                snap.car_preset[k] = snap.car_preset[i];
                snap.car_config[k] = snap.car_config[i];
                snap.controls[k] = snap.controls[i];
                snap.cars[k] = snap.cars[i];
                k++;
            }
            snap.n_cars = k;
            snap.raw_words = 0;  // the exact records are per car slot; restore from the public state
            check(rsg_batch_set_snapshot(nb->handle(), 0, &snap));
        }
        for (size_t i = 0; i < specs.size(); i++) {
            std::unique_ptr<Car> c(new Car());
            c->arena_ = this;
            c->team = specs[i].team;
            c->config = specs[i].config ? *specs[i].config : PresetCarConfig(specs[i].preset);
            c->preset = specs[i].config ? RSG_PRESET_CUSTOM : specs[i].preset;
            cars.push_back(std::move(c));
        }
        // ids: carried over from the old cars; new cars continue the count
        uint32_t next = last_id_;
        size_t j = 0;
        for (size_t i = 0; i < cars_.size(); i++) {
            if ((int)i == removed) continue;
            cars[j]->id = cars_[i]->id;
            cars[j]->controls = cars_[i]->controls;
            j++;
        }
        for (; j < cars.size(); j++) cars[j]->id = ++next;
        last_id_ = next;
        b_ = std::move(nb);
        cars_ = std::move(cars);
        specs_ = specs;
    }

    GameMode mode_;
    uint64_t seed_;
    std::vector<CarSpec> specs_;
    std::unique_ptr<ArenaBatch> b_;
    std::vector<std::unique_ptr<Car>> cars_;
    Ball ball_;
    uint32_t last_id_ = 0;
};

inline CarState Car::GetState() { return arena_->GetCarState(arena_->slot_of(this)); }
inline void Car::SetState(const CarState& s) { arena_->SetCarState(arena_->slot_of(this), s); }
inline BallState Ball::GetState() { return arena_->GetBallState(); }
inline void Ball::SetState(const BallState& s) { arena_->SetBallState(s); }

}  // namespace rsg_amd

#endif  // RSG_HPP
