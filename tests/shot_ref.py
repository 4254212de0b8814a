"""Float64 CPU restatement of the aerial-shot step (test infrastructure: the
oracle for the shot adjoint).

One tick of an airborne car and a flying ball with no world contact, as
rocketsim_amd/csrc/rsg_shot.h documents it: the car's air control and boost
(tests/air_ref.py), ball damping and gravity, the car hitbox vs ball contact
(closest point of the box core to the ball centre), Ball::_OnHit's extra hit
impulse (src/Sim/Ball/Ball.cpp:236-290), the sequential-impulse solve of the
contact (btSequentialImpulseConstraintSolver as RocketSim configures it:
friction 2.0, restitution 0, split impulse, 10 + 10 passes), integration and
the speed caps; written with torch operations so torch autograd gives the
exact float64 gradient.  Constants are the reference's (src/RLConst.h,
src/Sim/Car/CarConfig/CarConfig.cpp Octane, Bullet's box and sphere setup).

It is pinned to the reference itself by tests/test_shot_adjoint.py
(trajectories of the oracle's Arena::Step on the same states and controls).
"""
import math

import torch

import air_ref

UU_TO_BT = 1.0 / 50.0
BT = 50.0
F32 = lambda x: float(torch.tensor(x, dtype=torch.float32))   # noqa: E731  the reference's float constants

# Octane hitbox (CarConfig.cpp) with btBoxShape's margin handling
_HITBOX = (120.507, 86.6994, 38.6591)
_OFFSET = (13.8757, 0.0, 20.755)
_half = [F32(h) * UU_TO_BT / 2 for h in _HITBOX]
BOX_HE = [h - 0.04 for h in _half]
BOX_MARGIN = min(0.1 * min(_half), 0.04)
BOX_OFFSET = [F32(o) * UU_TO_BT for o in _OFFSET]
_he_full = [h + BOX_MARGIN for h in BOX_HE]
_l = [2 * h for h in _he_full]
CAR_INV_I = [1 / (180.0 / 12 * (_l[1] ** 2 + _l[2] ** 2)), 1 / (180.0 / 12 * (_l[0] ** 2 + _l[2] ** 2)),
             1 / (180.0 / 12 * (_l[0] ** 2 + _l[1] ** 2))]
CAR_INV_M = 1 / 180.0
RADIUS = F32(91.25) * UU_TO_BT
BALL_INV_M = 1 / 30.0
BALL_INV_I = 1 / (0.4 * 30.0 * RADIUS * RADIUS)
_car_disc = math.sqrt(sum(h * h for h in _he_full)) + math.sqrt(sum(o * o for o in BOX_OFFSET))
THRESHOLD = min(_car_disc, RADIUS + 0.08) * 0.02
FRICTION = 2.0
ERP2 = 0.8
SPLIT_TURN_ERP = 0.1
EXTRA_Z_SCALE, EXTRA_FWD_SCALE, EXTRA_MAX_DV = 0.35, 0.65, 4600.0
EXTRA_CURVE = ((0.0, 0.65), (500.0, 0.65), (2300.0, 0.55), (4600.0, 0.30))
BALL_MAX_SPEED, BALL_MAX_ANG = 6000.0, 6.0
GOAL = (0.0, 5120.0, 321.0)


def _dot(a, b):
    return (a * b).sum(-1)


def _cross(a, b):
    return torch.linalg.cross(a, b, dim=-1)


def _curve(x):
    out = torch.full_like(x, EXTRA_CURVE[-1][1])
    out = torch.where(x <= EXTRA_CURVE[0][0], torch.full_like(x, EXTRA_CURVE[0][1]), out)
    for (x0, y0), (x1, y1) in reversed(list(zip(EXTRA_CURVE[:-1], EXTRA_CURVE[1:]))):
        seg = (x > x0) & (x < x1)
        out = torch.where(seg, y0 + (y1 - y0) * (x - x0) / (x1 - x0), out)
    return out


def _rs_normalized(a):
    n = a.norm(dim=-1, keepdim=True)
    return torch.where(n > 0, a / torch.where(n > 0, n, torch.ones_like(n)), torch.zeros_like(a))


def step(s, ctrl, tick, last_extra, dt):
    """s: dict of [n, 3] / [n, 4] tensors (Bullet units); ctrl [n, 4];
    last_extra: LongTensor [n]; returns (new s, new last_extra, contact mask)."""
    cp, cq, cv, cw, bp, bv, bw = s["cp"], s["cq"], s["cv"], s["cw"], s["bp"], s["bv"], s["bw"]
    P, Y, R, B = ctrl.unbind(-1)
    f, r, u = air_ref.basis(cq)
    torque = r * (-P * air_ref.TORQUE[0])[..., None] + u * (Y * air_ref.TORQUE[1])[..., None] + \
        f * (-R * air_ref.TORQUE[2])[..., None]
    ky = air_ref.DAMPING[1] * (1 - Y.abs())
    kp = air_ref.DAMPING[0] * (1 - P.abs())
    damp = u * (ky * _dot(u, cw))[..., None] + r * (kp * _dot(r, cw))[..., None] + \
        f * (air_ref.DAMPING[2] * _dot(f, cw))[..., None]
    ext_t = (torque - damp) * (air_ref.CAR_TORQUE_SCALE * dt)
    g = torch.zeros_like(cv)
    g[..., 2] = air_ref.GRAVITY_Z * UU_TO_BT
    ext_f = (g + f * (B * air_ref.BOOST_ACCEL_AIR * UU_TO_BT)[..., None]) * dt
    bv = bv * (1.0 - 0.03) ** dt
    ext_fb = g * dt

    # contact
    box_c = cp + f * BOX_OFFSET[0] + r * BOX_OFFSET[1] + u * BOX_OFFSET[2]
    d = bp - box_c
    loc = torch.stack([_dot(f, d), _dot(r, d), _dot(u, d)], -1)
    he = torch.tensor(BOX_HE, dtype=loc.dtype)
    inside = (loc.abs() < he).all(-1)
    q = torch.maximum(torch.minimum(loc, he), -he)
    to_box = (box_c + f * q[..., 0:1] + r * q[..., 1:2] + u * q[..., 2:3]) - bp
    dist_core = to_box.norm(dim=-1)
    depth = dist_core - (BOX_MARGIN + RADIUS)
    contact = (~inside) & (depth < THRESHOLD) & (dist_core > 0)
    cm = contact[..., None]
    safe = torch.where(contact, dist_core, torch.ones_like(dist_core))
    n = to_box / safe[..., None]
    pt_b = bp + n * RADIUS
    pt_a = pt_b + n * depth[..., None]
    rel1, rel2 = pt_a - cp, pt_b - bp

    # extra hit impulse
    extra_ok = contact & ((tick > last_extra + 1) | (last_extra > tick))
    rel_pos = bp * BT - cp * BT
    rel_vel = bv * BT - cv * BT
    rel_speed = torch.minimum(rel_vel.norm(dim=-1), torch.full_like(dist_core, EXTRA_MAX_DV))
    hit_dir = _rs_normalized(rel_pos * torch.tensor([1.0, 1.0, EXTRA_Z_SCALE], dtype=bp.dtype))
    fwd_adj = (f * _dot(hit_dir, f)[..., None]) * (1 - EXTRA_FWD_SCALE)
    hit_dir = _rs_normalized(hit_dir - fwd_adj)
    added = hit_dir * rel_speed[..., None] * _curve(rel_speed)[..., None]
    cache = torch.where((extra_ok & (rel_speed > 0))[..., None], added * UU_TO_BT, torch.zeros_like(added))
    last_extra = torch.where(extra_ok, torch.full_like(last_extra, tick), last_extra)

    def inv_iw_car(x):
        return f * (CAR_INV_I[0] * _dot(f, x))[..., None] + r * (CAR_INV_I[1] * _dot(r, x))[..., None] + \
            u * (CAR_INV_I[2] * _dot(u, x))[..., None]

    # contact row
    ta0, ta1 = _cross(rel1, n), _cross(rel2, n)
    ang_a, ang_b = inv_iw_car(ta0), -ta1 * BALL_INV_I
    jac = 1 / ((CAR_INV_M + _dot(n, _cross(ang_a, rel1))) + (BALL_INV_M + _dot(n, _cross(-ang_b, rel2))))
    n1, rc1, n2, rc2 = n, ta0, -n, -ta1
    rel_v = (_dot(n1, cv + ext_f) + _dot(rc1, cw + ext_t)) + (_dot(n2, bv + ext_fb) + _dot(rc2, bw))
    rhs = -rel_v * jac
    rhs_pen = torch.where(depth > 0, torch.zeros_like(depth), -depth * ERP2 / dt) * jac
    # friction row
    vel = ((cv + ext_f) + _cross(cw + ext_t, rel1)) - ((bv + ext_fb) + _cross(bw, rel2))
    dirv = vel - n * _dot(n, vel)[..., None]
    lat = (dirv * dirv).sum(-1)
    big = lat > 1.1920928955078125e-07
    dir_main = dirv / torch.sqrt(torch.where(big, lat, torch.ones_like(lat)))[..., None]
    za = n[..., 2].abs() > 0.7071067811865475244
    a1 = n[..., 1] ** 2 + n[..., 2] ** 2
    a2 = n[..., 0] ** 2 + n[..., 1] ** 2
    k1 = 1 / torch.sqrt(torch.where(a1 > 0, a1, torch.ones_like(a1)))
    k2 = 1 / torch.sqrt(torch.where(a2 > 0, a2, torch.ones_like(a2)))
    ps_z = torch.stack([torch.zeros_like(k1), -n[..., 2] * k1, n[..., 1] * k1], -1)
    ps_x = torch.stack([-n[..., 1] * k2, n[..., 0] * k2, torch.zeros_like(k2)], -1)
    dirv = torch.where(big[..., None], dir_main, torch.where(za[..., None], ps_z, ps_x))
    f_rc1, f_rc2 = _cross(rel1, dirv), _cross(rel2, -dirv)
    f_ang_a, f_ang_b = inv_iw_car(f_rc1), f_rc2 * BALL_INV_I
    f_jac = 1 / ((CAR_INV_M + _dot(dirv, _cross(f_ang_a, rel1))) + (BALL_INV_M + _dot(dirv, _cross(-f_ang_b, rel2))))
    f_rhs = -((_dot(dirv, cv + ext_f) + _dot(f_rc1, cw)) + (_dot(-dirv, bv + ext_fb) + _dot(f_rc2, bw))) * f_jac

    zero3 = torch.zeros_like(cv)
    a_dv, a_dw, b_dv, b_dw = zero3, zero3, zero3, zero3
    a_push, a_turn, b_push, b_turn = zero3, zero3, zero3, zero3
    # split passes: until a pass moves nothing
    applied_push = torch.zeros_like(depth)
    active = contact & (rhs_pen != 0)
    for it in range(10):
        delta = rhs_pen - (_dot(n1, a_push) + _dot(rc1, a_turn)) * jac - (_dot(n2, b_push) + _dot(rc2, b_turn)) * jac
        s_ = applied_push + delta
        delta = torch.where(s_ < 0, -applied_push, delta)
        delta = torch.where(active, delta, torch.zeros_like(delta))
        applied_push = torch.where(active, torch.where(s_ < 0, torch.zeros_like(s_), s_), applied_push)
        a_push = a_push + (n1 * CAR_INV_M) * delta[..., None]
        a_turn = a_turn + ang_a * delta[..., None]
        b_push = b_push + (n2 * BALL_INV_M) * delta[..., None]
        b_turn = b_turn + ang_b * delta[..., None]
        active = active & (delta != 0)
    # velocity passes
    applied = torch.zeros_like(depth)
    f_applied = torch.zeros_like(depth)
    for it in range(10):
        delta = rhs - (_dot(n1, a_dv) + _dot(rc1, a_dw)) * jac - (_dot(n2, b_dv) + _dot(rc2, b_dw)) * jac
        s_ = applied + delta
        delta = torch.where(s_ < 0, -applied, delta)
        applied = torch.where(s_ < 0, torch.zeros_like(s_), s_)
        a_dv = a_dv + (n1 * CAR_INV_M) * delta[..., None]
        a_dw = a_dw + ang_a * delta[..., None]
        b_dv = b_dv + (n2 * BALL_INV_M) * delta[..., None]
        b_dw = b_dw + ang_b * delta[..., None]
        lo, hi = -FRICTION * applied, FRICTION * applied
        fd = f_rhs - (_dot(dirv, a_dv) + _dot(f_rc1, a_dw)) * f_jac - (_dot(-dirv, b_dv) + _dot(f_rc2, b_dw)) * f_jac
        s_ = f_applied + fd
        fd = torch.where(s_ < lo, lo - f_applied, fd)
        app = torch.where(s_ < lo, lo, s_)
        fd = torch.where(s_ < hi, fd, hi - f_applied)
        app = torch.where(s_ < hi, app, hi)
        on = applied > 0
        fd = torch.where(on, fd, torch.zeros_like(fd))
        f_applied = torch.where(on, app, f_applied)
        a_dv = a_dv + (dirv * CAR_INV_M) * fd[..., None]
        a_dw = a_dw + f_ang_a * fd[..., None]
        b_dv = b_dv + (-dirv * BALL_INV_M) * fd[..., None]
        b_dw = b_dw + f_ang_b * fd[..., None]

    car_v = torch.where(cm, (cv + a_dv) + ext_f, cv + ext_f)
    car_w = torch.where(cm, (cw + a_dw) + ext_t, cw + ext_t)
    ball_v = torch.where(cm, (bv + b_dv) + ext_fb, bv + ext_fb)
    ball_w = torch.where(cm, bw + b_dw, bw)
    pushed = cm & ((a_push != 0).any(-1, keepdim=True) | (a_turn != 0).any(-1, keepdim=True))
    cp = torch.where(pushed, cp + a_push * dt, cp)
    qs = air_ref.qmul(air_ref.dq(a_turn * SPLIT_TURN_ERP, dt), cq)
    cq = torch.where(pushed, qs / qs.norm(dim=-1, keepdim=True), cq)
    bpushed = cm & (b_push != 0).any(-1, keepdim=True)
    bp = torch.where(bpushed, bp + b_push * dt, bp)
    # integrate and cap
    cp = cp + car_v * dt
    h = air_ref.qmul(air_ref.dq(car_w, dt), cq)
    cq = h / h.norm(dim=-1, keepdim=True)
    bp = bp + ball_v * dt
    out = dict(cp=cp, cq=cq, cv=air_ref.cap(car_v, air_ref.CAR_MAX_SPEED * UU_TO_BT),
               cw=air_ref.cap(car_w, air_ref.CAR_MAX_ANG_SPEED),
               bp=bp, bv=air_ref.cap(ball_v + cache, BALL_MAX_SPEED * UU_TO_BT), bw=air_ref.cap(ball_w, BALL_MAX_ANG))
    return out, last_extra, contact


def unpack(state0):
    """[n, 22] interface units -> dict in Bullet units."""
    return dict(cp=state0[:, 0:3] * UU_TO_BT, cq=state0[:, 3:7], cv=state0[:, 7:10] * UU_TO_BT,
                cw=state0[:, 10:13], bp=state0[:, 13:16] * UU_TO_BT, bv=state0[:, 16:19] * UU_TO_BT,
                bw=state0[:, 19:22])


def pack(s):
    return torch.cat([s["cp"] * BT, s["cq"], s["cv"] * BT, s["cw"], s["bp"] * BT, s["bv"] * BT, s["bw"]], -1)


def near_world(s):
    """Ball or car near the floor, ceiling or walls, where the model stops
    (the same test as rsg_shot.h)."""
    bp, cp = s["bp"].detach(), s["cp"].detach()
    ceil_z, wall_x, wall_y, car_r = 2044.0 * UU_TO_BT, (4096.0 - 256.0) * UU_TO_BT, (5120.0 - 256.0) * UU_TO_BT, 2.5
    return ((bp[:, 2] < RADIUS + 0.5) | (bp[:, 2] > ceil_z - RADIUS - 0.5) | (bp[:, 0].abs() > wall_x - RADIUS) |
            (bp[:, 1].abs() > wall_y - RADIUS) | (cp[:, 2] < car_r) | (cp[:, 2] > ceil_z - car_r) |
            (cp[:, 0].abs() > wall_x - car_r) | (cp[:, 1].abs() > wall_y - car_r))


def rollout(state0, controls, target, tick_rate=120.0):
    """state0 [n, 22], controls [ticks, n, 4], target [n, 3] (uu) ->
    (loss [n] = |ball - target| in uu, final state [n, 22], contact [ticks, n],
    world [n]: left the modelled domain at some tick)."""
    dt = 1.0 / tick_rate
    s = unpack(state0)
    last = torch.full((state0.shape[0],), -1000, dtype=torch.long)
    contacts = []
    world = torch.zeros(state0.shape[0], dtype=torch.bool)
    for t in range(controls.shape[0]):
        s, last, c = step(s, controls[t], t, last, dt)
        contacts.append(c)
        world = world | near_world(s)
    loss = (s["bp"] * BT - target).norm(dim=-1)
    return loss, pack(s), torch.stack(contacts), world


from rocketsim_amd.workloads import _basis_quat  # noqa: E402,F401
from rocketsim_amd.workloads import shot_problem as random_problem  # noqa: E402,F401
