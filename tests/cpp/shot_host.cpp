// Host build of the aerial-shot rollout and its backward pass
// (rocketsim_amd/csrc/rsg_shot.h) for the CPU tests: one pair, the code the
// GPU kernels run.
#include "rsg.h"
#include "../../rocketsim_amd/csrc/rsg_shot.h"

using namespace rsg;

extern "C" float shot_host_rollout(int ticks, int K, const float* state0, const float* ctrl, const float* target,
                                   float* final_state, float* dctrl, float* ds0, uint32_t* flags) {
    const float dt = 1.f / 120.f;
    ShotConsts k = shot_consts(dt);
    ShotState<float> s;
    shot_load(state0, 1, 0, s);
    int nseg = (ticks + K - 1) / K;
    float* ckpt = new float[(size_t)nseg * SHOT_CKPT_WORDS];
    uint32_t fl = 0;
    V3 tg = v3(target[0], target[1], target[2]);
    float loss = shot_forward(s, ctrl, 1, 0, ticks, K, ckpt, tg, k, fl);
    shot_store(final_state, 1, 0, s);
    shot_backward(ctrl, 1, 0, ticks, K, ckpt, tg, 1.f, k, dctrl, ds0);
    delete[] ckpt;
    *flags = fl;
    return loss;
}
