"""Teacher-forced samples of the oracle in the synthetic arena loaded as six
mesh files (rocketsim_amd.arena_mesh.soccar_split_cmfs).

The oracle initialises its meshes once per process (RocketSim::InitFromMem,
src/RocketSim.cpp:97-105), so samples of the split arena come from a process
of their own:

    python tests/multi_mesh_case.py OUT.npz SEED [N_TRAJ]

writes seam-crossing, ball, melee and near-wall samples to OUT.npz
(parity.save_samples).
"""
import math
import os
import random
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def sample_seams(ref, seed, n_traj, ticks=90, stride=2):
    """Play across the pieces' seams, where a body touches two meshes at once
    (their manifolds then reach the solver in the static bodies' order): the
    ball against the side-wall fillets across y = 0, the back wall above the
    goal across x = 0 and the goal posts (goal box against back wall); one car
    driving along a side wall across y = 0 and into it, one at a goal post."""
    import parity
    from rocketsim_amd.state import CF_ON_GROUND, Controls, MODE_SOCCAR
    rng = random.Random(seed)
    out = []
    for k in range(n_traj):
        A = ref.RefArena(MODE_SOCCAR)
        teams, presets = [0, 1], [rng.randint(0, 5), rng.randint(0, 5)]
        for t, p in zip(teams, presets):
            A.add_car(t, p)
        A.reset_kickoff(rng.randint(0, 1000))
        sx, sy = rng.choice([-1, 1]), rng.choice([-1, 1])
        b = A.get_ball()
        kind = k % 3
        if kind == 0:
            b.pos[:] = [sx * rng.uniform(3600, 3900), rng.uniform(-300, 300), rng.uniform(120, 600)]
            b.vel[:] = [sx * rng.uniform(500, 2000), rng.gauss(0, 800), rng.gauss(-300, 300)]
        elif kind == 1:
            b.pos[:] = [rng.uniform(-300, 300), sy * rng.uniform(4600, 4900), rng.uniform(800, 1800)]
            b.vel[:] = [rng.gauss(0, 800), sy * rng.uniform(500, 2500), rng.gauss(0, 500)]
        else:
            b.pos[:] = [sx * rng.uniform(700, 1000), sy * rng.uniform(4500, 4900), rng.uniform(100, 700)]
            b.vel[:] = [rng.gauss(0, 500), sy * rng.uniform(500, 2500), rng.gauss(0, 300)]
        b.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
        A.set_ball(b)
        d = rng.choice([-1, 1])
        a = rng.uniform(0.1, 0.5)
        cars = [(sx * rng.uniform(3650, 3950), -d * rng.uniform(150, 600), sx * math.sin(a), d * math.cos(a)),
                (-sx * rng.uniform(800, 1000), sy * rng.uniform(4200, 4500), rng.gauss(0, 0.2), sy)]
        for i, (x, y, hx, hy) in enumerate(cars):
            yaw = math.atan2(hy, hx)
            cy, syw = math.cos(yaw), math.sin(yaw)
            spd = rng.uniform(900, 2000)
            s = A.get_car(i)
            s.pos[:] = [x, y, 17.0]
            s.rot[:] = [cy, syw, 0, -syw, cy, 0, 0, 0, 1]
            s.vel[:] = [spd * cy, spd * syw, 0]
            s.ang_vel[:] = [0, 0, 0]
            s.flags = CF_ON_GROUND | 0x1E
            s.boost = 50
            A.set_car(i, s)
        A.step(2)
        ctrls = []
        for _ in teams:
            c = Controls()
            c.throttle = 1.0
            c.steer = rng.uniform(-0.3, 0.3)
            c.boost = rng.random() < 0.5
            ctrls.append(c)
        for t in range(ticks):
            if t % stride == 0:
                out.append(parity._one_tick(A, MODE_SOCCAR, teams, presets, ctrls))
            else:
                A.step(1)
        A.close()
    return out


def sample_split(ref, seed, n_traj):
    """Play across the seams, ball flights, 4-car melees and 6-car melees
    near the walls."""
    import parity
    out = sample_seams(ref, seed + 3, 3 * n_traj)
    out += parity.sample_soccar_ball(ref, n_traj=n_traj, ticks_per_traj=96, seed=seed, stride=4)
    out += parity.sample_soccar_melee(ref, n_traj=n_traj, ticks_per_traj=90, seed=seed + 1, n_cars=4, stride=2)
    out += parity.sample_soccar_melee(ref, n_traj=n_traj, ticks_per_traj=90, seed=seed + 2, n_cars=6, stride=2,
                                      near_walls=True)
    return out


def generate(path, seed, n_traj, timeout=600):
    """Runs this script in a fresh process; returns its samples."""
    import parity
    r = subprocess.run([sys.executable, os.path.abspath(__file__), path, str(seed), str(n_traj)],
                       capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    if r.returncode != 0:
        raise RuntimeError("split-arena sampling failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
    return parity.load_samples(path)


def main():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, HERE)
    from oracle import ref
    from rocketsim_amd import arena_mesh
    import parity
    path, seed = sys.argv[1], int(sys.argv[2])
    n_traj = int(sys.argv[3]) if len(sys.argv) > 3 else 8
    ref.init(meshes=arena_mesh.soccar_split_cmfs())
    samples = sample_split(ref, seed, n_traj)
    parity.save_samples(path, samples)
    print(len(samples), "samples")


if __name__ == "__main__":
    main()
