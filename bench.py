"""Throughput of the batched Soccar tick step on one MI355X.

Workload (one "step" = one tick of every arena): 16384 independent Soccar 2v2
arenas, 4 Octane cars + ball each, on the synthetic arena mesh, cars driven by
the seeded control process of include/rsg_workload.h (redraw every U{2..60}
ticks), kickoff auto-reset on goal.  State is resident in HBM for the whole
timed region.  The batch first plays a fixed pre-roll (240 ticks, 2 s of
game time) from kickoff, so that every timed window is contact-rich play
(cars on the ground and walls, ball touches), whatever --steps is.

Prints one JSON line: whole-job arena-ticks/s of the batched path, the plain
one-arena-per-thread GPU kernel on the same workload and window
(plain_gpu, with the speedup), the roofline of the dominant kernel
(HBM-bound: algorithmic bytes per arena-tick = full state read+write +
controls) and the reference CPU baseline (RocketSim compiled from source, one
Arena per pinned thread on every core this process may use, 2v2; plus the
single-thread 1v1 figure of the reference's own benchmark shape).

The batched path: kernel and launch shape from the autotuner, ticks in
8-tick launches, the whole chunks captured in one hipGraph and any remainder
launched eagerly.  --workload mixed is the uneven batch: 32768 arenas,
40/35/25 % 1v1/2v2/3v3, six hitbox presets, random boost.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload soccar2v2|mixed] [--arenas N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "arena-ticks/s, 2v2 batched vs plain GPU vs CPU; adjoint and shot rollouts/s"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per arena-tick, 2v2: 2 x (4 cars x 220 B + ball 36 B +
# pads 272 B + tick 8 B) + 4 x 24 B controls.
BYTES_PER_ARENA_TICK_2V2 = 2488
# HBM bytes per launch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE passes
# over `bench.py --profile-probe` (counter factor from a 64 MiB copy), keyed by
# (kernel, arenas, ticks per launch, arenas per workgroup).
HBM_TRAFFIC_PER_LAUNCH = {
    # profiles/rsg_step_arena_coop_kernel_soccar2v2_16384.md: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0,
    # the mean of the probe's two 8-tick launches after the 240-tick pre-roll
    ("rsg_step_arena_coop_kernel", 16384, 8, 4): (2715608 + 1446796) * 1024,
}


def _host_info():
    """Cores this process may run on, the CPU model and the oracle's compiler."""
    import platform
    import subprocess
    cpus = sorted(os.sched_getaffinity(0))
    model = platform.processor()
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            quota = None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    try:
        cc = subprocess.run(["/opt/rocm/llvm/bin/clang++", "--version"], capture_output=True,
                            text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cc = "unknown"
    return cpus, model, quota, cc


def cpu_baseline(target_seconds=10.0):
    """The reference's Arena::Step on the host: one Soccar 2v2 arena per
    std::thread, one thread pinned to each core this process may use, on a
    bounded sample of the same workload; plus one thread of a 1v1 arena for
    100k ticks (the reference's own single-thread benchmark shape)."""
    from oracle import ref
    from rocketsim_amd import arena_mesh
    if not ref.available():
        return None
    ref.init([arena_mesh.soccar_cmf()])
    cpus, model, quota, cc = _host_info()
    threads = len(cpus)
    # grow the sample until it takes at least half the target wall time
    ticks = 1000
    while True:
        n, secs = ref.run_workload(threads, 1, 2, 2, 0, ticks, 12345, cpus)
        if secs >= 0.5 * target_seconds or ticks >= 4_000_000:
            break
        ticks = int(ticks * min(50.0, max(2.0, target_seconds / max(secs, 1e-3))))
    n1, s1 = ref.run_workload(1, 1, 1, 1, 0, 100_000, 777, cpus[:1])
    return {
        "value": n / secs, "unit": "arena-ticks/s", "cores": threads, "kind": "reference",
        "sample": f"{threads} pinned threads x 1 Soccar 2v2 Octane arena x {ticks} ticks, same control process, "
                  f"kickoff reset on goal ({secs:.1f} s wall)",
        "cpu_model": model, "nproc": os.cpu_count(), "cgroup_cpu_quota": quota,
        "compiler": cc + ", -O3 -DNDEBUG -ffp-contract=off (oracle/Makefile)",
        "single_thread_1v1": {"value": n1 / s1, "unit": "ticks/s", "ticks": n1,
                              "sample": "1 thread, 1 Soccar 1v1 Octane arena, 100k ticks, seeded controls"},
    }


def make_batch(workload, n_arenas, seed):
    """The batch of a workload, its algorithmic bytes per arena-tick and a
    description.  soccar2v2: n 2v2 Octane arenas.  mixed: n arenas, 40 % 1v1 /
    35 % 2v2 / 25 % 3v3, presets uniform over the six hitboxes, boost
    U[0, 100] (seeded)."""
    import random
    from rocketsim_amd.batch import ArenaBatch
    from rocketsim_amd.state import MODE_SOCCAR
    if workload == "soccar2v2":
        b = ArenaBatch.uniform(n_arenas, 2, 2, 0, mode=MODE_SOCCAR, seed=seed, synth_controls=True, auto_reset=True)
        return b, float(BYTES_PER_ARENA_TICK_2V2), \
            f"soccar 2v2 x {n_arenas} arenas, Octane, seeded controls, auto-reset on goal", 4.0
    rng = random.Random(seed)
    counts, teams, presets = [], [], []
    for _ in range(n_arenas):
        u = rng.random()
        per_team = 1 if u < 0.40 else (2 if u < 0.75 else 3)
        counts.append(2 * per_team)
        teams += [0] * per_team + [1] * per_team
        presets += [rng.randrange(6) for _ in range(2 * per_team)]
    b = ArenaBatch(counts, teams, presets, mode=MODE_SOCCAR, seed=seed, synth_controls=True, auto_reset=True)
    cars = b.get_cars()
    for c in cars:
        c.boost = rng.uniform(0.0, 100.0)
    b.set_cars(list(cars))
    # per arena-tick: 2 x (cars x 220 + 316) + cars x 24 (see BYTES_PER_ARENA_TICK_2V2)
    bytes_avg = sum(2 * (c * 220 + 316) + c * 24 for c in counts) / n_arenas
    return b, bytes_avg, (f"soccar mixed x {n_arenas} arenas (40/35/25 % 1v1/2v2/3v3), six presets, "
                          f"boost U[0,100], seeded controls, auto-reset on goal"), sum(counts) / n_arenas


def run_timed(batch, steps, warmup, preroll, chunk, use_graph, stream):
    """Pre-roll `preroll` ticks from kickoff and warm up `warmup` ticks, then
    time exactly `steps` ticks between two synchronisations: the whole
    `chunk`-tick launches as one hipGraph replay (or eager launches) and the
    remainder as one eager launch.  The HIP events are recorded on the stream
    the kernels run on.  Returns (wall s, event ms, chunk, graph used)."""
    import torch
    sp = stream.cuda_stream
    chunk = max(1, min(chunk, steps))
    whole, rest = steps // chunk, steps % chunk
    for n in [chunk] * ((preroll + warmup) // chunk) + ([(preroll + warmup) % chunk] if (preroll + warmup) % chunk else []):
        batch.step(n, sp)
    graph = use_graph and whole > 0
    if graph:
        batch.graph_create(chunk, whole)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    if graph:
        batch.graph_launch(sp)
    else:
        for _ in range(whole):
            batch.step(chunk, sp)
    if rest:
        batch.step(rest, sp)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, start.elapsed_time(end), chunk, graph


def bench_adjoint(n_cars=65536, ticks=120, reps=5, seed=7):
    """Differentiable air control (65536 airborne cars, 120 ticks, checkpoints
    every 16 ticks): rollouts/s of forward + hand-written adjoint through the
    torch autograd Function."""
    import torch
    from rocketsim_amd.adjoint import air_rollout
    g = torch.Generator(device="cuda").manual_seed(seed)
    pos = torch.stack([torch.rand(n_cars, device="cuda", generator=g) * 4000 - 2000,
                       torch.rand(n_cars, device="cuda", generator=g) * 4000 - 2000,
                       800 + 1000 * torch.rand(n_cars, device="cuda", generator=g)], -1)
    quat = torch.randn(n_cars, 4, device="cuda", generator=g)
    quat = quat / quat.norm(dim=-1, keepdim=True)
    state0 = torch.cat([pos, quat, torch.randn(n_cars, 3, device="cuda", generator=g) * 500,
                        torch.randn(n_cars, 3, device="cuda", generator=g) * 2], -1)
    params = torch.randn(ticks, n_cars, 4, device="cuda", generator=g, requires_grad=True)
    tq = torch.randn(n_cars, 4, device="cuda", generator=g)
    target = torch.cat([tq / tq.norm(dim=-1, keepdim=True), pos + 300 * torch.randn(n_cars, 3, device="cuda",
                                                                                   generator=g)], -1)

    def one():
        controls = torch.cat([torch.tanh(params[..., :3]), torch.sigmoid(params[..., 3:])], -1)
        loss, _ = air_rollout(state0, controls, target)
        loss.sum().backward()
        return loss

    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        one()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    return {"value": n_cars / wall, "unit": "rollouts/s", "cars": n_cars, "ticks": ticks, "ckpt_every": 16,
            "ms_per_batch": wall * 1e3,
            "workload": "airborne cars, z U[800,1800], random orientation, v N(0,500), w N(0,2); "
                        "controls tanh/sigmoid of parameters; loss geodesic angle + 1e-4 |dp|^2; fwd + adjoint"}


def bench_shot(n=16384, ticks=90, reps=5, seed=13):
    """Differentiable aerial shot (16384 car+ball pairs, 90 ticks, checkpoints
    every 16 ticks): rollouts/s of forward + adjoint through the car-ball hit,
    through the LibTorch autograd Function."""
    import torch
    from rocketsim_amd.adjoint import shot_rollout
    from rocketsim_amd.workloads import shot_problem
    s0, ctrl, tgt = shot_problem(n, ticks, seed)
    state0, target = s0.float().cuda(), tgt.float().cuda()
    params = torch.atanh(ctrl[..., :3].clamp(-0.999, 0.999)).float().cuda()
    params = torch.cat([params, torch.logit(ctrl[..., 3:].clamp(1e-3, 1 - 1e-3)).float().cuda()], -1)
    params.requires_grad_(True)

    def one():
        controls = torch.cat([torch.tanh(params[..., :3]), torch.sigmoid(params[..., 3:])], -1)
        loss, _, flags = shot_rollout(state0, controls, target)
        loss.sum().backward()
        return flags

    flags = one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        one()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    fl = flags.cpu()
    return {"value": n / wall, "unit": "rollouts/s", "pairs": n, "ticks": ticks, "ckpt_every": 16,
            "ms_per_batch": wall * 1e3, "touched": int(((fl & 1) != 0).sum()),
            "outside_model": int(((fl & 6) != 0).sum()),
            "workload": "ball x U[-3000,3000], y U[-4000,4000], z U[300,1500] uu, v N(0,800); car airborne "
                        "250-600 uu from the ball, facing it; loss |ball - (0,5120,321)| at tick 90; fwd + adjoint"}


def bench_ball_pred(n=4096, T=720, reps=3, seed=11):
    """Ball prediction (BallPredTracker over the GPU): n trackers predicting
    T ticks (6 s) of the Soccar ball, from random states in the field.  Full
    re-prediction (ForceUpdateAllPred) and the incremental update of 8 ticks
    that a bot calling UpdatePredManual every 8 ticks pays."""
    import random
    import torch
    from rocketsim_amd.ball_pred import BallPredTracker
    from rocketsim_amd.state import BallState, MODE_SOCCAR
    rng = random.Random(seed)
    states = []
    for _ in range(n):
        s = BallState()
        s.pos[:] = [rng.uniform(-3800, 3800), rng.uniform(-4800, 4800), rng.uniform(100, 1900)]
        s.vel[:] = [rng.gauss(0, 1200), rng.gauss(0, 1200), rng.gauss(0, 800)]
        s.ang_vel[:] = [rng.gauss(0, 2) for _ in range(3)]
        s.rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        states.append(s)
    tr = BallPredTracker(mode=MODE_SOCCAR, n=n, num_pred_ticks=T)
    try:
        tr.force_update_all_pred(states)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            tr.force_update_all_pred(states)
        torch.cuda.synchronize()
        full = (time.perf_counter() - t0) / reps
        # incremental: the current state is the prediction 8 ticks ahead
        pd = tr.pred_data_all()
        cur = [pd[i * T + 8] for i in range(n)]
        ticks = [8] * n
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr.update_pred_manual(cur, ticks)
        torch.cuda.synchronize()
        inc = time.perf_counter() - t0
    finally:
        tr.close()
    return {"value": n * (T - 1) / full, "unit": "ball-ticks/s", "trackers": n, "num_pred_ticks": T,
            "ms_full_update": full * 1e3, "ms_update_8_ticks": inc * 1e3,
            "workload": "Soccar ball, pos U(field), vel N(0,1200/800), w N(0,2); ForceUpdateAllPred; "
                        "then UpdatePredManual 8 ticks on"}


KERNEL_NAMES = {1: "rsg_step_arena_kernel", 2: "rsg_step_arena_lds_kernel", 3: "rsg_step_arena_coop_kernel"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=512)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--preroll", type=int, default=240,
                    help="ticks played from kickoff before the warm-up, so the timed window is contact-rich play")
    ap.add_argument("--workload", choices=["soccar2v2", "mixed"], default="soccar2v2")
    ap.add_argument("--arenas", type=int, default=None, help="default 16384 (soccar2v2) or 32768 (mixed)")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--arenas-per-block", type=int, default=0,
                    help="arenas per workgroup of the step kernel (0 = autotune, or the kernel's default)")
    ap.add_argument("--kernel", choices=["auto", "scratch", "lds", "coop"], default="auto",
                    help="where the step kernel keeps each arena's working copy (auto = autotune)")
    ap.add_argument("--chunk", type=int, default=8,
                    help="ticks per kernel launch (the arena state stays on chip across a chunk)")
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of one captured hipGraph")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain one-arena-per-thread kernel (private memory, 64 arenas per workgroup, "
                         "one tick per launch) on the same workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable air-control line")
    ap.add_argument("--no-shot", action="store_true", help="skip the differentiable aerial-shot line")
    ap.add_argument("--no-ball-pred", action="store_true", help="skip the ball-prediction line")
    ap.add_argument("--profile-probe", action="store_true",
                    help="for rocprofv3 --pmc passes: fill and copy a 64 MiB buffer (calibrates the HBM byte "
                         "counters), pre-roll, then two launches of --chunk ticks; prints no JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); arenas are independent, "
              "so N GPUs would each step their own shard of the arena range with no per-tick exchange.",
              file=sys.stderr)
        sys.exit(2)
    n_arenas = args.arenas or (16384 if args.workload == "soccar2v2" else 32768)

    import torch
    from rocketsim_amd import arena_mesh, native

    native.lib()
    native.check(native.lib().rsg_init_mesh(arena_mesh.soccar_cmf(), len(arena_mesh.soccar_cmf())))
    batch, bytes_per_at, desc, cars_per_arena = make_batch(args.workload, n_arenas, args.seed)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    tuned = None
    if args.kernel == "auto" and args.arenas_per_block == 0 and not args.profile_probe:
        # kernel and launch shape from the autotuner (times each on a copy of
        # the state, after the same pre-roll as the timed run)
        for n in [args.chunk] * (args.preroll // args.chunk):
            batch.step(n, sp)
        k, pb, ms = batch.autotune(8, sp)
        tuned = {"kernel": k, "arenas_per_block": pb, "ms_per_tick": ms}
        batch.close()
        batch, _, _, _ = make_batch(args.workload, n_arenas, args.seed)
        batch.set_kernel(k)
        batch.set_launch(pb)
    else:
        batch.set_kernel({"auto": 0, "scratch": 1, "lds": 2, "coop": 3}[args.kernel])
        batch.set_launch(args.arenas_per_block)
    per_block = batch.get_launch()
    kernel_name = KERNEL_NAMES[batch.get_kernel()]

    if args.profile_probe:
        src = torch.empty(64 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        for n in [args.chunk] * (args.preroll // args.chunk):
            batch.step(n, sp)
        torch.cuda.synchronize()
        for _ in range(2):
            batch.step(args.chunk, sp)
            torch.cuda.synchronize()
        batch.close()
        print("profile probe done", flush=True)
        return

    wall, ev_ms, chunk, graph = run_timed(batch, args.steps, args.warmup, args.preroll, args.chunk,
                                          not args.no_graph, stream)
    arena_ticks = n_arenas * args.steps
    value = arena_ticks / wall
    kernel_ms = ev_ms / args.steps  # per tick (one launch covers `chunk` ticks)
    achieved = (n_arenas * bytes_per_at) / (kernel_ms * 1e-3) / 1e9
    goals = batch.get_goals()
    status_flagged = sum(1 for x in batch.get_status() if x)
    batch.close()

    plain = None
    if not args.no_plain:
        pb_batch, _, _, _ = make_batch(args.workload, n_arenas, args.seed)
        pb_batch.set_kernel(1)
        pb_batch.set_launch(64)
        p_wall, p_ev, _, _ = run_timed(pb_batch, args.steps, args.warmup, args.preroll, 1, False, stream)
        pb_batch.close()
        plain = {"kernel": "rsg_step_arena_kernel", "arenas_per_block": 64, "ticks_per_launch": 1,
                 "value": arena_ticks / p_wall, "ms_per_tick": p_ev / args.steps,
                 "speedup": (arena_ticks / wall) / (arena_ticks / p_wall)}

    adjoint = None if args.no_adjoint else bench_adjoint()
    shot = None if args.no_shot else bench_shot()
    ball_pred = None if args.no_ball_pred else bench_ball_pred()

    cpu = None if args.no_cpu_baseline else cpu_baseline()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "arena-ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "arenas": n_arenas, "cars_per_arena": cars_per_arena, "tick_rate": 120,
                   "preroll_ticks": args.preroll, "arenas_per_block": per_block, "ticks_per_launch": chunk,
                   "kernel": kernel_name, "autotuned": tuned is not None, "hip_graph": graph,
                   "goals_during_run": int(sum(goals[0]) + sum(goals[1])),
                   "arenas_with_status_bits": status_flagged},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": HBM_TRAFFIC_PER_LAUNCH.get((kernel_name, n_arenas, chunk, per_block))
                     if args.workload == "soccar2v2" else None,
                     "traffic_unit": "B per launch",
                     "kernel": kernel_name, "kernel_ms_per_tick": kernel_ms,
                     "bytes_per_arena_tick": bytes_per_at},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["vs_cpu"] = value / cpu["value"]
    if plain is not None:
        line["plain_gpu"] = plain
    if adjoint is not None:
        line["adjoint"] = adjoint
    if shot is not None:
        line["shot"] = shot
    if ball_pred is not None:
        line["ball_pred"] = ball_pred
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
