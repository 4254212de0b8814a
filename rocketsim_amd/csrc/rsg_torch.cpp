// rsg_torch.cpp -- LibTorch C++ autograd Functions over the differentiable
// rollouts of librsg.so (rsg_air_rollout / rsg_air_rollout_backward and
// rsg_shot_rollout / rsg_shot_rollout_backward, include/rsg.h), built as the
// Python extension module rsg_torch.
//
//   loss, final = rsg_torch.air_rollout(state0, controls, target, ckpt_every, tick_rate, stream)
//
// state0 [n, 13] (position uu, quaternion x y z w, velocity uu/s, angular
// velocity rad/s), controls [ticks, n, 4] (pitch, yaw, roll in [-1, 1], boost
// in [0, 1]), target [n, 7] (quaternion, position uu); GPU tensors.  The
// forward keeps a checkpoint of every car's state every ckpt_every ticks; the
// backward recomputes each segment from its checkpoint and runs the
// hand-written adjoint kernel.  Gradients flow to controls and state0.
// `stream` is the hipStream_t to run on (the caller's current stream); the
// backward runs on the stream of its forward, as autograd does for every op.
#include <torch/extension.h>

#include <string>

#include "rsg.h"

namespace {

void check(int rc, const char* what) {
    TORCH_CHECK(rc >= 0, what, ": ", rsg_last_error());
}

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

struct AirRolloutFn : public torch::autograd::Function<AirRolloutFn> {
    static variable_list forward(AutogradContext* ctx, torch::Tensor state0, torch::Tensor controls,
                                 torch::Tensor target, int64_t ckpt_every, double tick_rate, int64_t stream) {
        TORCH_CHECK(state0.is_cuda() && controls.is_cuda() && target.is_cuda(),
                    "air_rollout: tensors must be on the GPU (the kernels are HIP only)");
        TORCH_CHECK(state0.dim() == 2 && state0.size(1) == 13, "air_rollout: state0 must be [n, 13]");
        TORCH_CHECK(controls.dim() == 3 && controls.size(2) == 4 && controls.size(1) == state0.size(0),
                    "air_rollout: controls must be [ticks, n, 4]");
        TORCH_CHECK(target.dim() == 2 && target.size(1) == 7 && target.size(0) == state0.size(0),
                    "air_rollout: target must be [n, 7]");
        const int64_t n = state0.size(0), ticks = controls.size(0);
        auto s0 = state0.to(torch::kFloat).t().contiguous();                  // [13][n]
        auto ctrl = controls.to(torch::kFloat).permute({0, 2, 1}).contiguous();  // [ticks][4][n]
        auto tgt = target.to(torch::kFloat).t().contiguous();                  // [7][n]
        const int64_t nseg = (ticks + ckpt_every - 1) / ckpt_every;
        auto opts = s0.options();
        auto ckpt = torch::empty({nseg * 13 * n}, opts);
        auto out = torch::empty({13, n}, opts);
        auto loss = torch::empty({n}, opts);
        check(rsg_air_rollout((int)n, (int)ticks, (int)ckpt_every, (float)tick_rate, s0.data_ptr<float>(),
                              ctrl.data_ptr<float>(), tgt.data_ptr<float>(), out.data_ptr<float>(),
                              loss.data_ptr<float>(), ckpt.data_ptr<float>(), (void*)stream),
              "rsg_air_rollout");
        ctx->save_for_backward({ctrl, tgt, ckpt});
        ctx->saved_data["n"] = n;
        ctx->saved_data["ticks"] = ticks;
        ctx->saved_data["ckpt_every"] = ckpt_every;
        ctx->saved_data["tick_rate"] = tick_rate;
        ctx->saved_data["stream"] = stream;
        auto final_state = out.t().contiguous();
        ctx->mark_non_differentiable({final_state});
        return {loss, final_state};
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        auto saved = ctx->get_saved_variables();
        auto ctrl = saved[0], tgt = saved[1], ckpt = saved[2];
        const int64_t n = ctx->saved_data["n"].toInt(), ticks = ctx->saved_data["ticks"].toInt();
        const int64_t ckpt_every = ctx->saved_data["ckpt_every"].toInt();
        const double tick_rate = ctx->saved_data["tick_rate"].toDouble();
        const int64_t stream = ctx->saved_data["stream"].toInt();
        auto dl = grads[0].defined() ? grads[0].to(torch::kFloat).contiguous() : torch::zeros({n}, ctrl.options());
        auto dctrl = torch::empty_like(ctrl);
        auto ds0 = torch::empty({13, n}, ctrl.options());
        check(rsg_air_rollout_backward((int)n, (int)ticks, (int)ckpt_every, (float)tick_rate, ctrl.data_ptr<float>(),
                                       tgt.data_ptr<float>(), ckpt.data_ptr<float>(), dl.data_ptr<float>(),
                                       dctrl.data_ptr<float>(), ds0.data_ptr<float>(), (void*)stream),
              "rsg_air_rollout_backward");
        return {ds0.t().contiguous(), dctrl.permute({0, 2, 1}).contiguous(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

// The aerial shot (rsg_shot_rollout / rsg_shot_rollout_backward): state0
// [n, 22], controls [ticks, n, 4], target [n, 3]; returns loss [n], the final
// state [n, 22] and the contact flags [n] (int32).
struct ShotRolloutFn : public torch::autograd::Function<ShotRolloutFn> {
    static variable_list forward(AutogradContext* ctx, torch::Tensor state0, torch::Tensor controls,
                                 torch::Tensor target, int64_t ckpt_every, double tick_rate, int64_t stream) {
        TORCH_CHECK(state0.is_cuda() && controls.is_cuda() && target.is_cuda(),
                    "shot_rollout: tensors must be on the GPU (the kernels are HIP only)");
        TORCH_CHECK(state0.dim() == 2 && state0.size(1) == 22, "shot_rollout: state0 must be [n, 22]");
        TORCH_CHECK(controls.dim() == 3 && controls.size(2) == 4 && controls.size(1) == state0.size(0),
                    "shot_rollout: controls must be [ticks, n, 4]");
        TORCH_CHECK(target.dim() == 2 && target.size(1) == 3 && target.size(0) == state0.size(0),
                    "shot_rollout: target must be [n, 3]");
        const int64_t n = state0.size(0), ticks = controls.size(0);
        auto s0 = state0.to(torch::kFloat).t().contiguous();
        auto ctrl = controls.to(torch::kFloat).permute({0, 2, 1}).contiguous();
        auto tgt = target.to(torch::kFloat).t().contiguous();
        const int64_t nseg = (ticks + ckpt_every - 1) / ckpt_every;
        auto opts = s0.options();
        auto ckpt = torch::empty({nseg * 23 * n}, opts);
        auto out = torch::empty({22, n}, opts);
        auto loss = torch::empty({n}, opts);
        auto flags = torch::empty({n}, opts.dtype(torch::kInt32));
        check(rsg_shot_rollout((int)n, (int)ticks, (int)ckpt_every, (float)tick_rate, s0.data_ptr<float>(),
                               ctrl.data_ptr<float>(), tgt.data_ptr<float>(), out.data_ptr<float>(),
                               loss.data_ptr<float>(), ckpt.data_ptr<float>(),
                               reinterpret_cast<uint32_t*>(flags.data_ptr<int32_t>()), (void*)stream),
              "rsg_shot_rollout");
        ctx->save_for_backward({ctrl, tgt, ckpt});
        ctx->saved_data["n"] = n;
        ctx->saved_data["ticks"] = ticks;
        ctx->saved_data["ckpt_every"] = ckpt_every;
        ctx->saved_data["tick_rate"] = tick_rate;
        ctx->saved_data["stream"] = stream;
        auto final_state = out.t().contiguous();
        ctx->mark_non_differentiable({final_state, flags});
        return {loss, final_state, flags};
    }

    static variable_list backward(AutogradContext* ctx, variable_list grads) {
        auto saved = ctx->get_saved_variables();
        auto ctrl = saved[0], tgt = saved[1], ckpt = saved[2];
        const int64_t n = ctx->saved_data["n"].toInt(), ticks = ctx->saved_data["ticks"].toInt();
        const int64_t ckpt_every = ctx->saved_data["ckpt_every"].toInt();
        const double tick_rate = ctx->saved_data["tick_rate"].toDouble();
        const int64_t stream = ctx->saved_data["stream"].toInt();
        auto dl = grads[0].defined() ? grads[0].to(torch::kFloat).contiguous() : torch::zeros({n}, ctrl.options());
        auto dctrl = torch::empty_like(ctrl);
        auto ds0 = torch::empty({22, n}, ctrl.options());
        check(rsg_shot_rollout_backward((int)n, (int)ticks, (int)ckpt_every, (float)tick_rate, ctrl.data_ptr<float>(),
                                        tgt.data_ptr<float>(), ckpt.data_ptr<float>(), dl.data_ptr<float>(),
                                        dctrl.data_ptr<float>(), ds0.data_ptr<float>(), (void*)stream),
              "rsg_shot_rollout_backward");
        return {ds0.t().contiguous(), dctrl.permute({0, 2, 1}).contiguous(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

std::vector<torch::Tensor> shot_rollout(torch::Tensor state0, torch::Tensor controls, torch::Tensor target,
                                        int64_t ckpt_every, double tick_rate, int64_t stream) {
    TORCH_CHECK(ckpt_every >= 1 && ckpt_every <= 32, "shot_rollout: ckpt_every must be within [1, 32]");
    auto r = ShotRolloutFn::apply(state0, controls, target, ckpt_every, tick_rate, stream);
    return {r[0], r[1], r[2]};
}

std::vector<torch::Tensor> air_rollout(torch::Tensor state0, torch::Tensor controls, torch::Tensor target,
                                       int64_t ckpt_every, double tick_rate, int64_t stream) {
    TORCH_CHECK(ckpt_every >= 1 && ckpt_every <= 32, "air_rollout: ckpt_every must be within [1, 32]");
    auto r = AirRolloutFn::apply(state0, controls, target, ckpt_every, tick_rate, stream);
    return {r[0], r[1]};
}

}  // namespace

PYBIND11_MODULE(rsg_torch, m) {
    m.doc() = "LibTorch autograd Function over the differentiable air-control rollout of librsg.so";
    m.def("air_rollout", &air_rollout, "loss, final = air_rollout(state0, controls, target, ckpt_every, tick_rate, stream)",
          pybind11::arg("state0"), pybind11::arg("controls"), pybind11::arg("target"), pybind11::arg("ckpt_every"),
          pybind11::arg("tick_rate"), pybind11::arg("stream"));
    m.def("shot_rollout", &shot_rollout,
          "loss, final, flags = shot_rollout(state0, controls, target, ckpt_every, tick_rate, stream)",
          pybind11::arg("state0"), pybind11::arg("controls"), pybind11::arg("target"), pybind11::arg("ckpt_every"),
          pybind11::arg("tick_rate"), pybind11::arg("stream"));
}
