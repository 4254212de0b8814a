"""Writes the fixtures under tests/golden/ (teacher-forced one-tick samples
and ball-prediction tracker runs) from the
oracle (the reference RocketSim built from source, oracle/_ref/librsref.so).

    python tests/golden/make_fixtures.py

Each fixture holds inputs and expected outputs only (raw state records of
include/rsg_state.h); GPU tests replay them when the oracle library is not
present on the machine.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import ref  # noqa: E402
import ball_pred_cases  # noqa: E402
import parity  # noqa: E402


def heatseeker():
    ref.init()
    hs = parity.sample_heatseeker(ref, n_traj=12, ticks=120, seed=20261207, stride=3)
    parity.save_samples(os.path.join(HERE, "heatseeker_tf.npz"), hs)


def hoops():
    ref.init()
    hp = parity.sample_hoops(ref, n_traj=16, ticks=120, seed=20261209, stride=3)
    parity.save_samples(os.path.join(HERE, "hoops_tf.npz"), hp)
    print("hoops_tf.npz:", len(hp), "samples")


def split():
    """The six-mesh arena's samples, drawn in a process of their own (the
    oracle loads its meshes once per process)."""
    import multi_mesh_case
    s = multi_mesh_case.generate(os.path.join(HERE, "soccar_split_tf.npz"), seed=20261208, n_traj=6)
    print("soccar_split_tf.npz:", len(s), "samples")


def main():
    split()
    ref.init()
    void = parity.sample_void(ref, n_traj=24, ticks_per_traj=8, seed=20261202)
    parity.save_samples(os.path.join(HERE, "void_tf.npz"), void)
    print("void_tf.npz:", len(void), "samples")
    ball = parity.sample_soccar_ball(ref, n_traj=32, ticks_per_traj=96, seed=20261203, stride=8)
    parity.save_samples(os.path.join(HERE, "soccar_ball_tf.npz"), ball)
    print("soccar_ball_tf.npz:", len(ball), "samples")
    drive = parity.sample_soccar_drive(ref, n_traj=8, ticks_per_traj=240, seed=20261204, n_cars=4, stride=5)
    parity.save_samples(os.path.join(HERE, "soccar_drive_tf.npz"), drive)
    print("soccar_drive_tf.npz:", len(drive), "samples")
    melee = parity.sample_soccar_melee(ref, n_traj=24, ticks_per_traj=100, seed=20261205, n_cars=4, stride=2)
    parity.save_samples(os.path.join(HERE, "soccar_melee_tf.npz"), melee)
    print("soccar_melee_tf.npz:", len(melee), "samples")
    walls = parity.sample_soccar_melee(ref, n_traj=24, ticks_per_traj=100, seed=20261206, n_cars=4, stride=2,
                                       near_walls=True)
    parity.save_samples(os.path.join(HERE, "soccar_walls_tf.npz"), walls)

    heatseeker()
    hoops()
    print("soccar_walls_tf.npz:", len(walls), "samples")
    states = ball_pred_cases.initial_states(ref, 10, seed=20261207)
    run = ball_pred_cases.reference_run(ref, states, 120)
    ball_pred_cases.save(os.path.join(HERE, "soccar_ball_pred.npz"), run, 120)
    print("soccar_ball_pred.npz:", len(states), "trackers")


if __name__ == "__main__":
    if sys.argv[1:] == ["split"]:
        split()
    elif sys.argv[1:] == ["hoops"]:
        hoops()
    else:
        main()
