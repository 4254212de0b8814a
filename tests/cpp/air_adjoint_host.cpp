// Host build of the air-control step and its adjoint (rocketsim_amd/csrc/
// rsg_adjoint.h) for CPU tests: one car, the same checkpoint-and-recompute
// backward pass as the GPU kernels.  Arrays in interface units, array of
// structures: state[13], controls[ticks][4], target[7].
#include "../../rocketsim_amd/csrc/rsg_adjoint.h"

using namespace rsg;

static AirState load(const float* a, bool uu) {
    float k = uu ? UU_TO_BT : 1.f;
    AirState s;
    s.p = v3(a[0], a[1], a[2]) * k;
    s.q = q4(a[3], a[4], a[5], a[6]);
    s.v = v3(a[7], a[8], a[9]) * k;
    s.w = v3(a[10], a[11], a[12]);
    return s;
}

extern "C" float air_host_rollout(int ticks, int K, const float* state0, const float* ctrl, const float* target,
                                  float* final_state, float* dctrl, float* dstate0) {
    const float dt = 1.f / 120.f;
    AirState s = load(state0, true);
    AirState ck[1024];
    int nseg = (ticks + K - 1) / K;
    for (int t = 0; t < ticks; t++) {
        if (t % K == 0) ck[t / K] = s;
        AirCtrl c = {ctrl[4 * t], ctrl[4 * t + 1], ctrl[4 * t + 2], ctrl[4 * t + 3]};
        air_step(s, c, dt);
    }
    Q4 tq = q4(target[0], target[1], target[2], target[3]);
    V3 tp = v3(target[4], target[5], target[6]) * UU_TO_BT;
    float loss = air_loss(s, tq, tp);
    final_state[0] = s.p.x * BT_TO_UU; final_state[1] = s.p.y * BT_TO_UU; final_state[2] = s.p.z * BT_TO_UU;
    final_state[3] = s.q.x; final_state[4] = s.q.y; final_state[5] = s.q.z; final_state[6] = s.q.w;
    final_state[7] = s.v.x * BT_TO_UU; final_state[8] = s.v.y * BT_TO_UU; final_state[9] = s.v.z * BT_TO_UU;
    final_state[10] = s.w.x; final_state[11] = s.w.y; final_state[12] = s.w.z;
    AirState g;
    air_loss_grad(s, tq, tp, 1.f, g);
    AirState st[32];
    for (int seg = nseg - 1; seg >= 0; seg--) {
        int t0 = seg * K, t1 = t0 + K < ticks ? t0 + K : ticks;
        AirState x = ck[seg];
        for (int t = t0; t < t1; t++) {
            st[t - t0] = x;
            AirCtrl c = {ctrl[4 * t], ctrl[4 * t + 1], ctrl[4 * t + 2], ctrl[4 * t + 3]};
            air_step(x, c, dt);
        }
        for (int t = t1 - 1; t >= t0; t--) {
            AirCtrl c = {ctrl[4 * t], ctrl[4 * t + 1], ctrl[4 * t + 2], ctrl[4 * t + 3]};
            AirCtrl gc = {0.f, 0.f, 0.f, 0.f};
            air_step_adjoint(st[t - t0], c, dt, g, gc);
            dctrl[4 * t] = gc.pitch; dctrl[4 * t + 1] = gc.yaw; dctrl[4 * t + 2] = gc.roll; dctrl[4 * t + 3] = gc.boost;
        }
    }
    dstate0[0] = g.p.x * UU_TO_BT; dstate0[1] = g.p.y * UU_TO_BT; dstate0[2] = g.p.z * UU_TO_BT;
    dstate0[3] = g.q.x; dstate0[4] = g.q.y; dstate0[5] = g.q.z; dstate0[6] = g.q.w;
    dstate0[7] = g.v.x * UU_TO_BT; dstate0[8] = g.v.y * UU_TO_BT; dstate0[9] = g.v.z * UU_TO_BT;
    dstate0[10] = g.w.x; dstate0[11] = g.w.y; dstate0[12] = g.w.z;
    return loss;
}
