"""Seeded synthetic inputs of the differentiable-rollout workloads (the
aerial-shot configuration: 16384 car+ball pairs, 90 ticks), shared by
bench.py and the tests."""
import math

import torch

# centre of the orange goal (uu), the shot's target
GOAL = (0.0, 5120.0, 321.0)


def _basis_quat(fwd, roll):
    """Unit quaternion (x, y, z, w) of the basis whose forward is fwd, rolled
    by `roll` about it."""
    up0 = torch.tensor([0.0, 0.0, 1.0], dtype=fwd.dtype).expand_as(fwd)
    r = torch.linalg.cross(up0, fwd, dim=-1)
    r = r / r.norm(dim=-1, keepdim=True)
    u = torch.linalg.cross(fwd, r, dim=-1)
    c, s = torch.cos(roll)[..., None], torch.sin(roll)[..., None]
    r, u = r * c + u * s, u * c - r * s
    m = torch.stack([fwd, r, u], -1)  # columns
    w = torch.sqrt(torch.clamp(1 + m[:, 0, 0] + m[:, 1, 1] + m[:, 2, 2], min=1e-12)) / 2
    x = (m[:, 2, 1] - m[:, 1, 2]) / (4 * w)
    y = (m[:, 0, 2] - m[:, 2, 0]) / (4 * w)
    z = (m[:, 1, 0] - m[:, 0, 1]) / (4 * w)
    return torch.stack([x, y, z, w], -1)


def shot_problem(n, ticks, seed, dtype=torch.float64):
    """Config-4 style inputs: ball x ~U[-3000,3000], y ~U[-4000,4000],
    z ~U[300,1500] uu, velocity ~N(0,800) uu/s per axis, angular velocity
    ~N(0,1); the car airborne 250-600 uu from the ball, facing it and moving
    at it at 600-1800 uu/s relative to the ball, random roll, angular velocity
    ~N(0,1); controls pitch/yaw/roll = tanh(0.5 N(0,1)), boost =
    sigmoid(N(0,1)); target the orange goal's centre."""
    g = torch.Generator().manual_seed(seed)
    bpos = torch.stack([torch.rand(n, generator=g) * 6000 - 3000, torch.rand(n, generator=g) * 8000 - 4000,
                        300 + 1200 * torch.rand(n, generator=g)], -1).double()
    bvel = (torch.randn(n, 3, generator=g) * 800).double()
    bang = torch.randn(n, 3, generator=g).double()
    off = torch.randn(n, 3, generator=g).double()
    off = off / off.norm(dim=-1, keepdim=True)
    dist = (250 + 350 * torch.rand(n, generator=g)).double()
    cpos = bpos + off * dist[:, None]
    cpos[:, 2] = torch.clamp(cpos[:, 2], min=200.0)
    to_ball = bpos - cpos
    to_ball = to_ball / to_ball.norm(dim=-1, keepdim=True)
    speed = (600 + 1200 * torch.rand(n, generator=g)).double()
    cvel = bvel + to_ball * speed[:, None]
    cvel = cvel * torch.clamp(2290.0 / cvel.norm(dim=-1, keepdim=True), max=1.0)
    roll = (torch.rand(n, generator=g) * 2 * math.pi).double()
    quat = _basis_quat(to_ball, roll)
    cang = torch.randn(n, 3, generator=g).double()
    state0 = torch.cat([cpos, quat, cvel, cang, bpos, bvel, bang], -1).to(dtype)
    raw = torch.randn(ticks, n, 4, generator=g)
    controls = torch.cat([torch.tanh(0.5 * raw[..., :3]), torch.sigmoid(raw[..., 3:])], -1).to(dtype)
    target = torch.tensor(GOAL, dtype=dtype).expand(n, 3).clone()
    return state0, controls, target
