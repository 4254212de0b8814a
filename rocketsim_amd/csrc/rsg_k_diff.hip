// rsg_k_diff.hip -- kernels of the differentiable rollouts: the airborne
// car (rsg_adjoint.h) and the aerial shot through the car-ball hit
// (rsg_shot.h), forward with checkpoints and backward.
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"
#include "rsg_tick.h"
#include "rsg_adjoint.h"
#include "rsg_shot.h"

using namespace rsg;

// ---------------------------------------------------------------------------
// Airborne-car rollouts and their adjoint (rsg_adjoint.h).  One lane per car;
// arrays are structure-of-arrays by word: state [13][n] (position uu,
// quaternion x y z w, velocity uu/s, angular velocity rad/s), controls
// [ticks][4][n] (pitch, yaw, roll, boost), target [7][n] (quaternion,
// position uu), checkpoints [ceil(ticks / K)][13][n] in Bullet units.

RSG_FN AirState air_load(const float* a, int n, int i, bool uu) {
    float k = uu ? UU_TO_BT : 1.f;
    AirState s;
    s.p = v3(a[0 * n + i], a[1 * n + i], a[2 * n + i]) * k;
    s.q = q4(a[3 * n + i], a[4 * n + i], a[5 * n + i], a[6 * n + i]);
    s.v = v3(a[7 * n + i], a[8 * n + i], a[9 * n + i]) * k;
    s.w = v3(a[10 * n + i], a[11 * n + i], a[12 * n + i]);
    return s;
}

RSG_FN void air_store(float* a, int n, int i, const AirState& s, float kpv) {
    a[0 * n + i] = s.p.x * kpv; a[1 * n + i] = s.p.y * kpv; a[2 * n + i] = s.p.z * kpv;
    a[3 * n + i] = s.q.x; a[4 * n + i] = s.q.y; a[5 * n + i] = s.q.z; a[6 * n + i] = s.q.w;
    a[7 * n + i] = s.v.x * kpv; a[8 * n + i] = s.v.y * kpv; a[9 * n + i] = s.v.z * kpv;
    a[10 * n + i] = s.w.x; a[11 * n + i] = s.w.y; a[12 * n + i] = s.w.z;
}

RSG_FN AirCtrl air_ctrl(const float* c, int n, int i, int t) {
    const float* b = c + (size_t)t * 4 * n;
    AirCtrl k;
    k.pitch = b[i]; k.yaw = b[n + i]; k.roll = b[2 * n + i]; k.boost = b[3 * n + i];
    return k;
}

__global__ void __launch_bounds__(256) rsg_air_forward_kernel(int n, int ticks, int K, float dt, const float* s0,
                                                              const float* ctrl, const float* target, float* s_out,
                                                              float* loss, float* ckpt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    AirState s = air_load(s0, n, i, true);
    for (int t = 0; t < ticks; t++) {
        if (ckpt && t % K == 0) air_store(ckpt + (size_t)(t / K) * AIR_STATE_WORDS * n, n, i, s, 1.f);
        air_step(s, air_ctrl(ctrl, n, i, t), dt);
    }
    if (s_out) air_store(s_out, n, i, s, BT_TO_UU);
    if (loss) {
        Q4 tq = q4(target[i], target[n + i], target[2 * n + i], target[3 * n + i]);
        V3 tp = v3(target[4 * n + i], target[5 * n + i], target[6 * n + i]) * UU_TO_BT;
        loss[i] = air_loss(s, tq, tp);
    }
}

__global__ void __launch_bounds__(256) rsg_air_backward_kernel(int n, int ticks, int K, float dt, const float* ctrl,
                                                               const float* target, const float* ckpt,
                                                               const float* dloss, float* dctrl, float* ds0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nseg = (ticks + K - 1) / K;
    AirState st[AIR_MAX_K];
    // final state: recompute the last segment from its checkpoint
    AirState s = air_load(ckpt + (size_t)(nseg - 1) * AIR_STATE_WORDS * n, n, i, false);
    for (int t = (nseg - 1) * K; t < ticks; t++) air_step(s, air_ctrl(ctrl, n, i, t), dt);
    Q4 tq = q4(target[i], target[n + i], target[2 * n + i], target[3 * n + i]);
    V3 tp = v3(target[4 * n + i], target[5 * n + i], target[6 * n + i]) * UU_TO_BT;
    AirState g;
    air_loss_grad(s, tq, tp, dloss[i], g);
    for (int seg = nseg - 1; seg >= 0; seg--) {
        int t0 = seg * K, t1 = t0 + K < ticks ? t0 + K : ticks;
        AirState x = air_load(ckpt + (size_t)seg * AIR_STATE_WORDS * n, n, i, false);
        for (int t = t0; t < t1; t++) {
            st[t - t0] = x;
            air_step(x, air_ctrl(ctrl, n, i, t), dt);
        }
        for (int t = t1 - 1; t >= t0; t--) {
            AirCtrl gc = {0.f, 0.f, 0.f, 0.f};
            air_step_adjoint(st[t - t0], air_ctrl(ctrl, n, i, t), dt, g, gc);
            float* b = dctrl + (size_t)t * 4 * n;
            b[i] = gc.pitch; b[n + i] = gc.yaw; b[2 * n + i] = gc.roll; b[3 * n + i] = gc.boost;
        }
    }
    // initial-state gradient in interface units (uu for position and velocity)
    if (ds0) air_store(ds0, n, i, g, UU_TO_BT);
}

// Differentiable aerial shot (rsg_shot.h): one lane per car-ball pair.
__global__ void __launch_bounds__(256) rsg_shot_forward_kernel(int n, int ticks, int K, float dt, const float* state0,
                                                               const float* ctrl, const float* target,
                                                               float* state_out, float* loss, float* ckpt,
                                                               uint32_t* flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ShotConsts k = shot_consts(dt);
    ShotState<float> s;
    shot_load(state0, n, i, s);
    uint32_t fl = 0;
    V3 tg = v3(target[i], target[n + i], target[2 * n + i]);
    float l = shot_forward(s, ctrl, n, i, ticks, K, ckpt, tg, k, fl);
    if (state_out) shot_store(state_out, n, i, s);
    if (loss) loss[i] = l;
    if (flags) flags[i] = fl;
}

// Backward of the shot: one group of SHOT_GROUP lanes per pair, lane j < 26
// carrying the tangent along input j of every tick (22 state words, then
// pitch, yaw, roll, boost), so a pair's 26 passes per tick run side by side
// instead of one after another.  Every lane of a group recomputes the same
// segment from its checkpoint; lane 0 keeps the segment's states in LDS
// ([K][SHOT_LDS_WORDS] per group), and after each tick the group gathers the
// new state adjoint from lanes 0-21 with cross-lane shuffles.  A group spans
// half a wave, so no workgroup barrier is needed: a group whose pair is past
// n leaves at once.
constexpr int SHOT_GROUP = 32;
constexpr int SHOT_LDS_WORDS = 24;   // a checkpoint's 23 words, padded

__global__ void __launch_bounds__(256) rsg_shot_backward_kernel(int n, int ticks, int K, float dt, const float* ctrl,
                                                                const float* target, const float* ckpt,
                                                                const float* dloss, float* dctrl, float* ds0) {
    extern __shared__ float shot_seg[];
    const int lane = threadIdx.x & (SHOT_GROUP - 1);
    const int grp = threadIdx.x / SHOT_GROUP;
    const int i = blockIdx.x * (blockDim.x / SHOT_GROUP) + grp;
    if (i >= n) return;
    float* seg = shot_seg + (size_t)grp * K * SHOT_LDS_WORDS;
    ShotConsts k = shot_consts(dt);
    V3 tg = v3(target[i], target[n + i], target[2 * n + i]);
    const int nseg = (ticks + K - 1) / K;
    ShotState<float> s;
    int last_extra;
    shot_ckpt_load(ckpt + (size_t)(nseg - 1) * SHOT_CKPT_WORDS * n, n, i, s, last_extra);
    uint32_t fl = 0;
    for (int t = (nseg - 1) * K; t < ticks; t++) shot_step(s, shot_ctrl(ctrl, n, i, t), t, last_extra, fl, k);
    float g[SHOT_STATE_WORDS];
    shot_loss_grad(s, tg, dloss[i], g);
    float mine = 0.f;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int t0 = sg * K, t1 = t0 + K < ticks ? t0 + K : ticks;
        ShotState<float> x;
        int lx;
        shot_ckpt_load(ckpt + (size_t)sg * SHOT_CKPT_WORDS * n, n, i, x, lx);
        for (int t = t0; t < t1; t++) {
            if (lane == 0) {
                float* r = seg + (t - t0) * SHOT_LDS_WORDS;
                for (int w = 0; w < SHOT_STATE_WORDS; w++) r[w] = shot_word(x, w);
                r[SHOT_STATE_WORDS] = __int_as_float(lx);
            }
            shot_step(x, shot_ctrl(ctrl, n, i, t), t, lx, fl, k);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (int t = t1 - 1; t >= t0; t--) {
            const float* r = seg + (t - t0) * SHOT_LDS_WORDS;
            ShotState<float> st;
            for (int w = 0; w < SHOT_STATE_WORDS; w++) shot_word(st, w) = r[w];
            const int le = __float_as_int(r[SHOT_STATE_WORDS]);
            mine = shot_vjp_entry(st, shot_ctrl(ctrl, n, i, t), t, le, k, g, lane);
            for (int w = 0; w < SHOT_STATE_WORDS; w++) g[w] = __shfl(mine, w, SHOT_GROUP);
            if (lane >= SHOT_STATE_WORDS && lane < SHOT_INPUTS)
                dctrl[(size_t)t * 4 * n + (size_t)(lane - SHOT_STATE_WORDS) * n + i] = mine;
        }
        // the next segment's states overwrite this one's
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    if (ds0 && lane < SHOT_STATE_WORDS) ds0[(size_t)lane * n + i] = mine * shot_unit(lane);
}

namespace rsg {

hipError_t launch_air_forward(int n, int ticks, int K, float dt, const float* s0, const float* ctrl,
                              const float* target, float* s_out, float* loss, float* ckpt, hipStream_t s) {
    const int block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(rsg_air_forward_kernel, dim3(grid), dim3(block), 0, s, n, ticks, K, dt, s0, ctrl, target,
                       s_out, loss, ckpt);
    return hipGetLastError();
}

hipError_t launch_air_backward(int n, int ticks, int K, float dt, const float* ctrl, const float* target,
                               const float* ckpt, const float* dloss, float* dctrl, float* ds0, hipStream_t s) {
    const int block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(rsg_air_backward_kernel, dim3(grid), dim3(block), 0, s, n, ticks, K, dt, ctrl, target, ckpt,
                       dloss, dctrl, ds0);
    return hipGetLastError();
}

hipError_t launch_shot_forward(int n, int ticks, int K, float dt, const float* state0, const float* ctrl,
                               const float* target, float* state_out, float* loss, float* ckpt, uint32_t* flags,
                               hipStream_t s) {
    // 64-lane workgroups spread the pairs over every CU
    const int block = 64, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(rsg_shot_forward_kernel, dim3(grid), dim3(block), 0, s, n, ticks, K, dt, state0, ctrl, target,
                       state_out, loss, ckpt, flags);
    return hipGetLastError();
}

hipError_t launch_shot_backward(int n, int ticks, int K, float dt, const float* ctrl, const float* target,
                                const float* ckpt, const float* dloss, float* dctrl, float* ds0, hipStream_t s) {
    const int block = 256, per_block = block / SHOT_GROUP, grid = (n + per_block - 1) / per_block;
    const size_t lds = (size_t)per_block * K * SHOT_LDS_WORDS * sizeof(float);
    hipLaunchKernelGGL(rsg_shot_backward_kernel, dim3(grid), dim3(block), lds, s, n, ticks, K, dt, ctrl, target, ckpt,
                       dloss, dctrl, ds0);
    return hipGetLastError();
}

}  // namespace rsg
