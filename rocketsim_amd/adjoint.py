"""Autograd over the differentiable air-control rollout of librsg.so
(rsg_air_rollout / rsg_air_rollout_backward, include/rsg.h).

    loss, final = air_rollout(state0, controls, target)

`air_rollout` goes through the LibTorch C++ torch::autograd::Function of the
rsg_torch extension (csrc/rsg_torch.cpp); `AirRollout` is the same Function
written in Python over the same two entry points.

state0 [n, 13] (position uu, quaternion x y z w, velocity uu/s, angular
velocity rad/s), controls [ticks, n, 4] (pitch, yaw, roll in [-1, 1], boost
in [0, 1]), target [n, 7] (quaternion, position uu); float32 CUDA tensors.
The forward pass keeps a checkpoint of every car's state every `ckpt_every`
ticks; the backward pass recomputes each segment from its checkpoint and runs
the hand-written adjoint kernel.  Gradients flow to controls and state0.
"""
import importlib.util
import os

import torch

from . import native

CKPT_EVERY = 16


def _soa(x):
    # [n, w] -> [w, n] contiguous (the kernels' structure-of-arrays layout)
    return x.t().contiguous()


class AirRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0, controls, target, ckpt_every=CKPT_EVERY, tick_rate=120.0):
        if not (state0.is_cuda and controls.is_cuda and target.is_cuda):
            raise RuntimeError("AirRollout: tensors must be on the GPU (the kernels are HIP only)")
        n = state0.shape[0]
        ticks = controls.shape[0]
        s0 = _soa(state0.float())
        ctrl = controls.float().permute(0, 2, 1).contiguous()   # [ticks][4][n]
        tgt = _soa(target.float())
        nseg = (ticks + ckpt_every - 1) // ckpt_every
        ckpt = torch.empty(nseg * 13 * n, dtype=torch.float32, device=state0.device)
        out = torch.empty(13, n, dtype=torch.float32, device=state0.device)
        loss = torch.empty(n, dtype=torch.float32, device=state0.device)
        stream = torch.cuda.current_stream().cuda_stream
        native.check(native.lib().rsg_air_rollout(n, ticks, ckpt_every, tick_rate, s0.data_ptr(), ctrl.data_ptr(),
                                                  tgt.data_ptr(), out.data_ptr(), loss.data_ptr(),
                                                  ckpt.data_ptr(), stream))
        ctx.save_for_backward(ctrl, tgt, ckpt)
        ctx.shape = (n, ticks, ckpt_every, tick_rate)
        final = out.t().contiguous()
        ctx.mark_non_differentiable(final)
        return loss, final

    @staticmethod
    def backward(ctx, dloss, dfinal):
        ctrl, tgt, ckpt = ctx.saved_tensors
        n, ticks, ckpt_every, tick_rate = ctx.shape
        dctrl = torch.empty_like(ctrl)
        ds0 = torch.empty(13, n, dtype=torch.float32, device=ctrl.device)
        dl = dloss.float().contiguous()
        stream = torch.cuda.current_stream().cuda_stream
        native.check(native.lib().rsg_air_rollout_backward(n, ticks, ckpt_every, tick_rate, ctrl.data_ptr(),
                                                           tgt.data_ptr(), ckpt.data_ptr(), dl.data_ptr(),
                                                           dctrl.data_ptr(), ds0.data_ptr(), stream))
        return ds0.t().contiguous(), dctrl.permute(0, 2, 1).contiguous(), None, None, None


_ext = None
TORCH_EXT_PATH = os.path.join(native.PKG, "lib", "rsg_torch.so")


def torch_ext():
    """The rsg_torch extension module (built by rocketsim_amd.build)."""
    global _ext
    if _ext is None:
        native.lib()
        if not os.path.exists(TORCH_EXT_PATH):
            raise native.NativeMissing(f"{TORCH_EXT_PATH} is not built; run `python -m rocketsim_amd.build`")
        spec = importlib.util.spec_from_file_location("rsg_torch", TORCH_EXT_PATH)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ext = mod
    return _ext


def air_rollout(state0, controls, target, ckpt_every=CKPT_EVERY, tick_rate=120.0):
    """(loss [n], final state [n, 13]) through the LibTorch C++ Function, on
    the current stream."""
    stream = torch.cuda.current_stream().cuda_stream
    loss, final = torch_ext().air_rollout(state0, controls, target, ckpt_every, tick_rate, stream)
    return loss, final


def shot_rollout(state0, controls, target, ckpt_every=CKPT_EVERY, tick_rate=120.0):
    """Differentiable aerial shot through the LibTorch C++ Function:
    state0 [n, 22] (car position uu, quaternion, velocity uu/s, angular
    velocity; ball position uu, velocity uu/s, angular velocity), controls
    [ticks, n, 4], target [n, 3] (uu) -> (loss [n] = |ball - target| (uu),
    final state [n, 22], flags [n]: 1 = the car touched the ball, 2 = the ball
    entered the hitbox core, 4 = near the floor, ceiling or walls; the last two
    are outside the model)."""
    stream = torch.cuda.current_stream().cuda_stream
    loss, final, flags = torch_ext().shot_rollout(state0, controls, target, ckpt_every, tick_rate, stream)
    return loss, final, flags


def air_rollout_py(state0, controls, target, ckpt_every=CKPT_EVERY, tick_rate=120.0):
    """The same through the Python autograd Function."""
    return AirRollout.apply(state0, controls, target, ckpt_every, tick_rate)
