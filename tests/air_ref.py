"""Float64 CPU restatement of the airborne car step (test infrastructure: the
oracle for the air-control adjoint).

The step is the reference's tick for a car with no contact
(src/Sim/Car/Car.cpp:584-667 Car::_UpdateAirTorque, :486-530 Car::_UpdateBoost
with the boost input relaxed to [0, 1], the world step's velocity update and
btTransformUtil::integrateTransform, LinearMath/btTransformUtil.h:37-87, then
the speed caps of Car::_FinishPhysicsTick), written with torch operations so
torch autograd gives the exact float64 gradient.  Constants are the
reference's (src/RLConst.h).  Units: Bullet (uu / 50) inside, uu at the
interface, as the GPU entry points use.

It is pinned to the reference itself by tests/test_air_adjoint.py (oracle
trajectories of airborne cars with constant controls).
"""
import math

import torch

UU_TO_BT = 1.0 / 50.0
GRAVITY_Z = -650.0
BOOST_ACCEL_AIR = 3175.0 / 3.0
CAR_TORQUE_SCALE = 2 * math.pi / (1 << 16) * 1000
TORQUE = (130.0, 95.0, 400.0)      # pitch, yaw, roll (RLConst.h CAR_AIR_CONTROL_TORQUE)
DAMPING = (30.0, 20.0, 50.0)       # pitch, yaw, roll (RLConst.h CAR_AIR_CONTROL_DAMPING)
CAR_MAX_SPEED = 2300.0
CAR_MAX_ANG_SPEED = 5.5


def basis(q):
    x, y, z, w = q.unbind(-1)
    f = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (z * x - w * y)], -1)
    r = torch.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], -1)
    u = torch.stack([2 * (z * x + y * w), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], -1)
    return f, r, u


def qmul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx - az * by + ax * bw + ay * bz,
                        aw * by - ax * bz + ay * bw + az * bx,
                        aw * bz - ay * bx + az * bw + ax * by,
                        aw * bw - az * bz - ax * bx - ay * by], -1)


def dq(w, dt):
    th = w.norm(dim=-1, keepdim=True)
    small = th < 0.001
    th_safe = torch.where(small, torch.ones_like(th), th)
    k = torch.where(small, 0.5 * dt - dt ** 3 / 48.0 * th * th, torch.sin(0.5 * th_safe * dt) / th_safe)
    return torch.cat([w * k, torch.cos(th * dt * 0.5)], -1)


def cap(v, vmax):
    n = v.norm(dim=-1, keepdim=True)
    return torch.where(n > vmax, v * (vmax / torch.clamp(n, min=1e-30)), v)


def step(p, q, v, w, ctrl, dt):
    """One tick; p, v in Bullet units; ctrl [..., 4] = pitch, yaw, roll, boost."""
    P, Y, R, b = ctrl.unbind(-1)
    f, r, u = basis(q)
    torque = r * (-P * TORQUE[0])[..., None] + u * (Y * TORQUE[1])[..., None] + f * (-R * TORQUE[2])[..., None]
    ky = DAMPING[1] * (1 - Y.abs())
    kp = DAMPING[0] * (1 - P.abs())
    kr = DAMPING[2]
    damp = (u * (ky * (u * w).sum(-1))[..., None] + r * (kp * (r * w).sum(-1))[..., None]
            + f * (kr * (f * w).sum(-1))[..., None])
    w1 = w + (torque - damp) * (CAR_TORQUE_SCALE * dt)
    g = torch.zeros_like(v)
    g[..., 2] = GRAVITY_Z * UU_TO_BT
    v1 = v + (g + f * (b * BOOST_ACCEL_AIR * UU_TO_BT)[..., None]) * dt
    p1 = p + v1 * dt
    h = qmul(dq(w1, dt), q)
    q1 = h / h.norm(dim=-1, keepdim=True)
    return p1, q1, cap(v1, CAR_MAX_SPEED * UU_TO_BT), cap(w1, CAR_MAX_ANG_SPEED)


def rollout(state0, controls, target, tick_rate=120.0):
    """state0 [n, 13] (interface units), controls [ticks, n, 4], target [n, 7]
    -> (loss [n], final state [n, 13])."""
    dt = 1.0 / tick_rate
    p = state0[:, 0:3] * UU_TO_BT
    q = state0[:, 3:7]
    v = state0[:, 7:10] * UU_TO_BT
    w = state0[:, 10:13]
    for t in range(controls.shape[0]):
        p, q, v, w = step(p, q, v, w, controls[t], dt)
    tq, tp = target[:, 0:4], target[:, 4:7]
    d = (q * tq).sum(-1).abs().clamp(max=1 - 1e-6)
    loss = 2 * torch.acos(d) + 1e-4 * ((p / UU_TO_BT - tp) ** 2).sum(-1)
    return loss, torch.cat([p / UU_TO_BT, q, v / UU_TO_BT, w], -1)


def random_problem(n, ticks, seed, dtype=torch.float64):
    """Config-3 style inputs: z ~ U[800, 1800] uu, uniform random orientation,
    velocity ~ N(0, 500) uu/s per axis, angular velocity ~ N(0, 2) rad/s;
    controls pitch/yaw/roll = tanh(N(0,1)), boost = sigmoid(N(0,1)); random
    target orientation and a target position near the start."""
    g = torch.Generator().manual_seed(seed)
    pos = torch.stack([torch.rand(n, generator=g) * 4000 - 2000, torch.rand(n, generator=g) * 4000 - 2000,
                       800 + 1000 * torch.rand(n, generator=g)], -1)
    quat = torch.randn(n, 4, generator=g)
    quat = quat / quat.norm(dim=-1, keepdim=True)
    vel = torch.randn(n, 3, generator=g) * 500
    ang = torch.randn(n, 3, generator=g) * 2
    state0 = torch.cat([pos, quat, vel, ang], -1).to(dtype)
    raw = torch.randn(ticks, n, 4, generator=g)
    controls = torch.cat([torch.tanh(raw[..., :3]), torch.sigmoid(raw[..., 3:])], -1).to(dtype)
    tq = torch.randn(n, 4, generator=g)
    tq = tq / tq.norm(dim=-1, keepdim=True)
    tp = pos + torch.randn(n, 3, generator=g) * 300
    target = torch.cat([tq, tp], -1).to(dtype)
    return state0, controls, target
