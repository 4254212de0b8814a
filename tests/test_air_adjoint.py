"""Differentiable air control: the forward model, the hand-written adjoint and
the torch autograd Function over the GPU kernels.

Pins (CPU):
  * the float64 restatement (tests/air_ref.py) reproduces the reference's own
    120-tick trajectories of airborne cars with constant controls (oracle);
  * the hand-written float32 adjoint (rsg_adjoint.h, host build) equals the
    float64 autograd gradient of that restatement;
  * the float64 autograd gradient equals central finite differences.
GPU:
  * the rollout kernel equals the float64 model and the plain batched step of
    THE_VOID (the forward judge);
  * the backward kernel's gradients on 512 config-3 samples are within 1e-3
    relative of the float64 gradient and of central differences (h = 1e-3).
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
import torch

import air_ref

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TICKS = 120


def _q_to_rot(q):
    """Unit quaternion -> RotMat words (forward, right, up columns)."""
    x, y, z, w = [float(a) for a in q]
    m = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    return [m[0][0], m[1][0], m[2][0], m[0][1], m[1][1], m[2][1], m[0][2], m[1][2], m[2][2]]


def _constant_controls(n, seed):
    s0, ctrl, tgt = air_ref.random_problem(n, TICKS, seed)
    c0 = ctrl[0].clone()
    c0[:, 3] = (c0[:, 3] > 0.5).double()   # the reference boosts or does not
    return s0, c0[None].repeat(TICKS, 1, 1), tgt


def test_air_model_pinned_to_reference():
    from oracle import ref
    from rocketsim_amd.state import Controls, MODE_THE_VOID
    if not ref.available():
        pytest.skip("oracle library not built")
    ref.init()
    n = 16
    s0, ctrl, tgt = _constant_controls(n, 3)
    _, fin = air_ref.rollout(s0, ctrl, tgt)
    for i in range(n):
        A = ref.RefArena(MODE_THE_VOID)
        A.add_car(0, 0)
        s = A.get_car(0)
        st = s0[i].numpy()
        s.pos[:] = list(st[0:3])
        s.rot[:] = _q_to_rot(st[3:7])
        s.vel[:] = list(st[7:10])
        s.ang_vel[:] = list(st[10:13])
        s.boost = 100.0
        s.flags = 0
        A.set_car(0, s)
        c = Controls()
        cc = ctrl[0, i].numpy()
        c.pitch, c.yaw, c.roll, c.boost = float(cc[0]), float(cc[1]), float(cc[2]), int(cc[3] > 0.5)
        A.set_controls(0, c)
        A.step(TICKS)
        g = A.get_car(0)
        f = fin[i].numpy()
        assert max(abs(g.pos[k] - f[k]) for k in range(3)) < 0.05
        assert max(abs(g.vel[k] - f[7 + k]) for k in range(3)) < 0.01
        assert max(abs(g.ang_vel[k] - f[10 + k]) for k in range(3)) < 1e-4
        R = _q_to_rot(f[3:7])
        assert max(abs(g.rot[k] - R[k]) for k in range(9)) < 1e-4
        A.close()


@pytest.fixture(scope="module")
def host_adjoint(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("air") / "airhost.so")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "cpp", "air_adjoint_host.cpp"),
                        "-o", so], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = C.CDLL(so)
    F = C.POINTER(C.c_float)
    L.air_host_rollout.restype = C.c_float
    L.air_host_rollout.argtypes = [C.c_int, C.c_int, F, F, F, F, F, F]
    return L


def test_hand_adjoint_matches_fp64_autograd(host_adjoint):
    F = C.POINTER(C.c_float)
    n = 48
    s0, ctrl, tgt = air_ref.random_problem(n, TICKS, 1)
    s0.requires_grad_(True)
    ctrl.requires_grad_(True)
    loss, _ = air_ref.rollout(s0, ctrl, tgt)
    loss.sum().backward()
    for i in range(n):
        st = np.ascontiguousarray(s0.detach().numpy()[i], dtype=np.float32)
        c = np.ascontiguousarray(ctrl.detach().numpy()[:, i, :], dtype=np.float32)
        tg = np.ascontiguousarray(tgt.numpy()[i], dtype=np.float32)
        fs = np.zeros(13, np.float32)
        dc = np.zeros((TICKS, 4), np.float32)
        ds = np.zeros(13, np.float32)
        lv = host_adjoint.air_host_rollout(TICKS, 16, st.ctypes.data_as(F), c.ctypes.data_as(F),
                                           tg.ctypes.data_as(F), fs.ctypes.data_as(F), dc.ctypes.data_as(F),
                                           ds.ctypes.data_as(F))
        assert abs(lv - loss[i].item()) <= 1e-4 * max(1.0, abs(loss[i].item()))
        gr = ctrl.grad.numpy()[:, i, :]
        assert np.linalg.norm(dc - gr) <= 1e-4 * np.linalg.norm(gr)
        gs = s0.grad.numpy()[i]
        assert np.linalg.norm(ds - gs) <= 1e-4 * np.linalg.norm(gs)


def test_fp64_autograd_matches_central_differences():
    n = 8
    s0, ctrl, tgt = air_ref.random_problem(n, TICKS, 2)
    ctrl.requires_grad_(True)
    loss, _ = air_ref.rollout(s0, ctrl, tgt)
    loss.sum().backward()
    g = torch.Generator().manual_seed(5)
    smooth = torch.ones_like(ctrl)
    smooth[..., 0:2] = (ctrl.detach()[..., 0:2].abs() > 0.05).double()  # away from the |pitch|, |yaw| kinks
    for _ in range(4):
        d = torch.randn(ctrl.shape, generator=g, dtype=torch.float64) * smooth
        h = 1e-6
        lp, _ = air_ref.rollout(s0, ctrl.detach() + h * d, tgt)
        lm, _ = air_ref.rollout(s0, ctrl.detach() - h * d, tgt)
        fd = (lp - lm).sum().item() / (2 * h)
        an = (ctrl.grad * d).sum().item()
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_torch_extension_loads():
    """The LibTorch C++ Function module links against librsg.so and torch and
    exports air_rollout (no GPU work)."""
    from rocketsim_amd.adjoint import torch_ext
    assert hasattr(torch_ext(), "air_rollout")


def _gpu_problem(n, seed):
    s0, ctrl, tgt = air_ref.random_problem(n, TICKS, seed)
    return s0, ctrl, tgt, s0.float().cuda(), ctrl.float().cuda(), tgt.float().cuda()


@pytest.mark.gpu
def test_air_rollout_forward_matches_fp64_and_plain_step():
    from rocketsim_amd.adjoint import air_rollout
    from rocketsim_amd.batch import ArenaBatch
    from rocketsim_amd.state import Controls, MODE_THE_VOID
    n = 256
    s0, ctrl, tgt = _constant_controls(n, 7)
    loss64, fin64 = air_ref.rollout(s0, ctrl, tgt)
    loss, fin = air_rollout(s0.float().cuda(), ctrl.float().cuda(), tgt.float().cuda())
    fin = fin.double().cpu()
    assert (fin[:, 0:3] - fin64[:, 0:3]).abs().max() < 0.05
    assert (fin[:, 7:10] - fin64[:, 7:10]).abs().max() < 0.05
    assert (fin[:, 10:13] - fin64[:, 10:13]).abs().max() < 1e-4
    assert ((loss.double().cpu() - loss64).abs() <= 1e-4 * loss64.abs().clamp(min=1.0)).all()
    # the plain batched step, one car per arena in THE_VOID
    b = ArenaBatch([1] * n, [0] * n, [0] * n, mode=MODE_THE_VOID)
    try:
        cars = b.get_cars()
        ctls = []
        for i in range(n):
            st = s0[i].numpy()
            cars[i].pos[:] = list(st[0:3])
            cars[i].rot[:] = _q_to_rot(st[3:7])
            cars[i].vel[:] = list(st[7:10])
            cars[i].ang_vel[:] = list(st[10:13])
            cars[i].boost = 100.0
            cars[i].flags = 0
            c = Controls()
            cc = ctrl[0, i].numpy()
            c.pitch, c.yaw, c.roll, c.boost = float(cc[0]), float(cc[1]), float(cc[2]), int(cc[3] > 0.5)
            ctls.append(c)
        b.set_cars(list(cars))
        b.set_controls(ctls)
        b.step(TICKS)
        out = b.get_cars()
        for i in range(n):
            assert max(abs(out[i].pos[k] - fin[i, k].item()) for k in range(3)) < 0.05
            assert max(abs(out[i].ang_vel[k] - fin[i, 10 + k].item()) for k in range(3)) < 1e-4
    finally:
        b.close()


@pytest.mark.gpu
def test_air_gradients_match_fp64_and_central_differences():
    from rocketsim_amd.adjoint import air_rollout
    n = 512
    s0, ctrl, tgt, s0g, ctrlg, tgtg = _gpu_problem(n, 11)
    # float64 reference gradient
    s0r = s0.clone().requires_grad_(True)
    ctrlr = ctrl.clone().requires_grad_(True)
    l64, _ = air_ref.rollout(s0r, ctrlr, tgt)
    l64.sum().backward()
    # GPU adjoint
    s0g.requires_grad_(True)
    ctrlg.requires_grad_(True)
    loss, _ = air_rollout(s0g, ctrlg, tgtg)
    loss.sum().backward()
    gc = ctrlg.grad.double().cpu()
    gs = s0g.grad.double().cpu()
    for i in range(n):
        assert (gc[:, i] - ctrlr.grad[:, i]).norm() <= 1e-3 * ctrlr.grad[:, i].norm()
        assert (gs[i] - s0r.grad[i]).norm() <= 1e-3 * s0r.grad[i].norm()
    # central differences of the float64 model, h = 1e-3, along random
    # directions.  The reference's air damping scales with 1 - |pitch| and
    # 1 - |yaw| (Car.cpp:656-658), which is not differentiable at 0: the
    # directions leave pitch and yaw entries within 0.05 of 0 alone, so the
    # difference does not straddle that kink.
    g = torch.Generator().manual_seed(3)
    smooth = torch.ones_like(ctrl)
    smooth[..., 0:2] = (ctrl[..., 0:2].abs() > 0.05).double()
    for _ in range(4):
        d = torch.randn(ctrl.shape, generator=g, dtype=torch.float64) * smooth
        h = 1e-3
        lp, _ = air_ref.rollout(s0, ctrl + h * d, tgt)
        lm, _ = air_ref.rollout(s0, ctrl - h * d, tgt)
        fd = (lp - lm) / (2 * h)
        an = (gc * d).sum(dim=(0, 2))
        assert ((fd - an).abs() <= 1e-3 * fd.abs().clamp(min=1e-2)).float().mean() >= 0.99


@pytest.mark.gpu
def test_cpp_function_equals_python_function():
    """The LibTorch C++ autograd Function and the Python one run the same
    kernels: losses, final states and gradients are identical."""
    from rocketsim_amd.adjoint import air_rollout, air_rollout_py
    n = 256
    _, _, _, s0g, ctrlg, tgtg = _gpu_problem(n, 13)
    out = []
    for fn in (air_rollout, air_rollout_py):
        s0 = s0g.clone().requires_grad_(True)
        c = ctrlg.clone().requires_grad_(True)
        loss, final = fn(s0, c, tgtg)
        (loss * torch.linspace(0.5, 1.5, n, device=loss.device)).sum().backward()
        out.append((loss.detach(), final.detach(), c.grad, s0.grad))
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)
