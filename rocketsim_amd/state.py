"""ctypes mirrors of include/rsg_state.h: the state records of the C ABI.

Field meanings follow the reference's public structs (CarState
src/Sim/Car/Car.h:17-120, BallState src/Sim/Ball/Ball.h:18-57, BoostPadState
src/Sim/BoostPad/BoostPad.h:23-31, CarControls src/Sim/CarControls.h:7-33),
plus the per-wheel vehicle values carried between ticks
(src/Sim/btVehicleRL/btVehicleRL.h:9-27).
"""
import ctypes as C

MODE_SOCCAR = 0
MODE_HOOPS = 1
MODE_HEATSEEKER = 2
MODE_THE_VOID = 5
ARENA_MODES = (MODE_SOCCAR, MODE_HOOPS, MODE_HEATSEEKER)  # modes played in an arena with meshes and pads

PRESET_OCTANE, PRESET_DOMINUS, PRESET_PLANK, PRESET_BREAKOUT, PRESET_HYBRID, PRESET_MERC = range(6)
PRESET_NAMES = ["octane", "dominus", "plank", "breakout", "hybrid", "merc"]
TEAM_BLUE, TEAM_ORANGE = 0, 1
NUM_PADS = 34
MAX_CARS = 8
SNAPSHOT_RAW_WORDS = 1024

# rsg_car_state.flags bits
CF_ON_GROUND = 1 << 0
CF_WHEEL = [1 << 1, 1 << 2, 1 << 3, 1 << 4]
CF_HAS_JUMPED = 1 << 5
CF_HAS_DOUBLE_JUMPED = 1 << 6
CF_HAS_FLIPPED = 1 << 7
CF_IS_FLIPPING = 1 << 8
CF_IS_JUMPING = 1 << 9
CF_IS_BOOSTING = 1 << 10
CF_SUPERSONIC = 1 << 11
CF_AUTO_FLIPPING = 1 << 12
CF_WORLD_CONTACT = 1 << 13
CF_DEMOED = 1 << 14
CF_HIT_VALID = 1 << 15
CF_BODY_DISABLED = 1 << 16

# rsg_batch_set_kernel
KERNEL_AUTO, KERNEL_SCRATCH, KERNEL_LDS, KERNEL_COOP = 0, 1, 2, 3
# per-arena status bits (include/rsg_state.h)
STATUS_NONFINITE, STATUS_CONTACT_OVERFLOW, STATUS_SOLVER_OVERFLOW = 1, 2, 4

F3 = C.c_float * 3
F4 = C.c_float * 4
F9 = C.c_float * 9


EVENT_GOAL, EVENT_BUMP, EVENT_DEMO = 1, 2, 3


class Event(C.Structure):
    """rsg_event (include/rsg_state.h)."""
    _fields_ = [("arena", C.c_uint32), ("type", C.c_uint32), ("tick", C.c_uint64), ("a", C.c_uint32),
                ("b", C.c_uint32), ("seq", C.c_uint32), ("_pad", C.c_uint32)]

    def key(self):
        return (self.arena, self.type, self.tick, self.a, self.b, self.seq)


class Controls(C.Structure):
    _fields_ = [
        ("throttle", C.c_float), ("steer", C.c_float), ("pitch", C.c_float),
        ("yaw", C.c_float), ("roll", C.c_float),
        ("jump", C.c_uint8), ("boost", C.c_uint8), ("handbrake", C.c_uint8), ("_pad", C.c_uint8),
    ]


class CarState(C.Structure):
    _fields_ = [
        ("tick_count_since_update", C.c_uint64),
        ("hit_tick", C.c_uint64),
        ("hit_extra_tick", C.c_uint64),
        ("pos", F3), ("rot", F9), ("vel", F3), ("ang_vel", F3),
        ("flags", C.c_uint32),
        ("flip_rel_torque", F3),
        ("jump_time", C.c_float), ("flip_time", C.c_float),
        ("air_time", C.c_float), ("air_time_since_jump", C.c_float),
        ("boost", C.c_float), ("time_since_boosted", C.c_float), ("boosting_time", C.c_float),
        ("supersonic_time", C.c_float), ("handbrake_val", C.c_float),
        ("auto_flip_timer", C.c_float), ("auto_flip_torque_scale", C.c_float),
        ("world_contact_normal", F3),
        ("car_contact_other_id", C.c_uint32),
        ("car_contact_cooldown", C.c_float),
        ("demo_respawn_timer", C.c_float),
        ("hit_rel_pos", F3), ("hit_ball_pos", F3), ("hit_extra_vel", F3),
        ("last_controls", Controls),
        ("wheel_engine_force", C.c_float), ("wheel_brake", C.c_float), ("wheel_steer", C.c_float),
        ("wheel_lat_friction", F4), ("wheel_long_friction", F4), ("wheel_extra_pushback", F4),
        ("bp_cell", C.c_int32), ("bp_seq", C.c_uint32),
    ]


class BallState(C.Structure):
    _fields_ = [
        ("tick_count_since_update", C.c_uint64),
        ("pos", F3), ("rot", F9), ("vel", F3), ("ang_vel", F3),
        ("bp_cell", C.c_int32), ("bp_seq", C.c_uint32),
        # Heatseeker: target net (+1 orange, -1 blue, 0 none), target speed, time since hit
        ("hs_y_target_dir", C.c_float), ("hs_cur_target_speed", C.c_float), ("hs_time_since_hit", C.c_float),
    ]


# Ball prediction record words (include/rsg_state.h RSG_PRED_*)
PRED_POS, PRED_ROT, PRED_VEL, PRED_ANG_VEL, PRED_TICKS_SINCE_UPDATE, PRED_HS, PRED_WORDS = 0, 3, 12, 15, 18, 20, 23


class BallPredView(C.Structure):
    _fields_ = [("traj", C.c_void_p), ("head", C.c_void_p), ("size", C.c_void_p),
                ("n", C.c_int), ("num_pred_ticks", C.c_int)]


class PadState(C.Structure):
    _fields_ = [("is_active", C.c_uint32), ("cooldown", C.c_float), ("prev_locked_car_id", C.c_uint32)]


DEMO_NORMAL, DEMO_ON_CONTACT, DEMO_DISABLED = 0, 1, 2
PRESET_PSYCLOPS, PRESET_COUNT, PRESET_CUSTOM = 6, 7, 255


class WheelPairConfig(C.Structure):
    _fields_ = [("wheel_radius", C.c_float), ("suspension_rest_length", C.c_float),
                ("connection_point_offset", C.c_float * 3)]


class CarConfig(C.Structure):
    """rsg_car_config: CarConfig (src/Sim/Car/CarConfig/CarConfig.h:6-43), uu."""
    _fields_ = [("hitbox_size", C.c_float * 3), ("hitbox_pos_offset", C.c_float * 3),
                ("front_wheels", WheelPairConfig), ("back_wheels", WheelPairConfig),
                ("three_wheels", C.c_uint8), ("_pad", C.c_uint8 * 3), ("dodge_deadzone", C.c_float)]


class MutatorConfig(C.Structure):
    """rsg_mutator_config: MutatorConfig (src/Sim/MutatorConfig/MutatorConfig.h:16-79)."""
    _fields_ = [
        ("gravity", C.c_float * 3), ("car_mass", C.c_float),
        ("car_world_friction", C.c_float), ("car_world_restitution", C.c_float),
        ("ball_mass", C.c_float), ("ball_max_speed", C.c_float), ("ball_drag", C.c_float),
        ("ball_world_friction", C.c_float), ("ball_world_restitution", C.c_float),
        ("jump_accel", C.c_float), ("jump_immediate_force", C.c_float),
        ("boost_accel_ground", C.c_float), ("boost_accel_air", C.c_float), ("boost_used_per_second", C.c_float),
        ("respawn_delay", C.c_float), ("bump_cooldown_time", C.c_float),
        ("boost_pad_cooldown_big", C.c_float), ("boost_pad_cooldown_small", C.c_float),
        ("car_spawn_boost_amount", C.c_float),
        ("ball_hit_extra_force_scale", C.c_float), ("bump_force_scale", C.c_float),
        ("ball_radius", C.c_float),
        ("unlimited_flips", C.c_uint8), ("unlimited_double_jumps", C.c_uint8),
        ("recharge_boost_enabled", C.c_uint8), ("demo_mode", C.c_uint8),
        ("recharge_boost_per_second", C.c_float), ("recharge_boost_delay", C.c_float),
        ("enable_team_demos", C.c_uint8), ("_pad", C.c_uint8 * 3),
        ("goal_base_threshold_y", C.c_float),
    ]


class ArenaSnapshot(C.Structure):
    """rsg_arena_snapshot (include/rsg_state.h): what Arena::Serialize writes
    for one arena (src/Sim/Arena/Arena.cpp:546-573) plus the hidden values."""
    _fields_ = [
        ("mode", C.c_int32), ("tick_time", C.c_float), ("tick_count", C.c_uint64),
        ("last_car_id", C.c_uint32), ("n_cars", C.c_int32),
        ("car_id", C.c_uint32 * MAX_CARS), ("car_team", C.c_uint8 * MAX_CARS),
        ("car_preset", C.c_uint8 * MAX_CARS), ("_pad", C.c_uint8 * 6),
        ("car_config", CarConfig * MAX_CARS),
        ("controls", Controls * MAX_CARS), ("cars", CarState * MAX_CARS),
        ("ball", BallState), ("pads", PadState * NUM_PADS), ("mutators", MutatorConfig),
        ("raw_layout", C.c_uint32), ("raw_words", C.c_uint32), ("raw", C.c_uint32 * SNAPSHOT_RAW_WORDS),
    ]


EVENT_SHOT, EVENT_SCORER, EVENT_SAVE = 4, 5, 6


class GameEventConfig(C.Structure):
    """rsg_game_event_config: GameEventTrackerConfig
    (src/Sim/GameEventTracker/GameEventTracker.h:10-41), reference defaults."""
    _fields_ = [("shot_min_speed", C.c_float), ("shot_touch_min_delay", C.c_float),
                ("pred_score_extra_margin", C.c_float), ("shot_event_cooldown", C.c_float),
                ("shot_min_score_time", C.c_float), ("goal_max_touch_time", C.c_float),
                ("pass_max_touch_time", C.c_float), ("auto_state_set_detection", C.c_int32)]

    def __init__(self, **kw):
        defaults = dict(shot_min_speed=1750.0, shot_touch_min_delay=0.3, pred_score_extra_margin=0.0,
                        shot_event_cooldown=1.0, shot_min_score_time=2.0, goal_max_touch_time=4.0,
                        pass_max_touch_time=2.0, auto_state_set_detection=1)
        defaults.update(kw)
        super().__init__(**defaults)


class BoostPadConfig(C.Structure):
    """rsg_boost_pad_config: BoostPadConfig (src/Sim/BoostPad/BoostPad.h:13-19)."""
    _fields_ = [("pos", C.c_float * 3), ("is_big", C.c_uint8), ("_pad", C.c_uint8 * 3)]


class BatchDesc(C.Structure):
    _fields_ = [
        ("n_arenas", C.c_int), ("mode", C.c_int), ("tick_rate", C.c_float),
        ("seed", C.c_uint64), ("auto_reset", C.c_int), ("synth_controls", C.c_int),
        ("cars_per_arena", C.POINTER(C.c_int)),
        ("car_team", C.POINTER(C.c_uint8)),
        ("car_preset", C.POINTER(C.c_uint8)),
        ("mutators", C.POINTER(MutatorConfig)),
        ("car_configs", C.POINTER(CarConfig)),
        ("ball_rotation", C.c_int),
        ("custom_pads", C.POINTER(BoostPadConfig)),
        ("n_custom_pads", C.c_int),
    ]


IDENTITY_ROT = (1, 0, 0, 0, 1, 0, 0, 0, 1)


def car_state_fields():
    """Names of the continuous and discrete fields compared by the parity tests."""
    return [f[0] for f in CarState._fields_]
