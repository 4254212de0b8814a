// rsg_k_status.hip -- device-side summary of the per-arena status words
// (RSG_STATUS_* bits): their OR and the number of arenas with a bit set,
// written to device memory without a host round trip.  Two kernels: per
// workgroup partial results, then one workgroup combining them.
#include <hip/hip_runtime.h>

#include "rsg_kernels.h"

using namespace rsg;

// One 256-lane workgroup per STATUS_ARENAS_PER_BLOCK arenas: partial OR and
// count of flagged arenas into partials[2 * block], partials[2 * block + 1].
__global__ void __launch_bounds__(STATUS_BLOCK) rsg_status_partial_kernel(const uint32_t* status, int n,
                                                                          uint32_t* partials) {
    __shared__ uint32_t wave_or[STATUS_BLOCK / 64], wave_cnt[STATUS_BLOCK / 64];
    const int a = (int)blockIdx.x * STATUS_BLOCK + (int)threadIdx.x;
    const uint32_t s = a < n ? status[a] : 0u;
    // within the wave: flagged lanes by ballot, the OR by a butterfly of shuffles
    const uint64_t flagged = __ballot(s != 0u);
    uint32_t o = s;
    for (int off = 32; off > 0; off >>= 1) o |= __shfl_xor(o, off, 64);
    const int wave = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wave_or[wave] = o;
        wave_cnt[wave] = (uint32_t)__popcll(flagged);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ro = 0, rc = 0;
        for (int w = 0; w < STATUS_BLOCK / 64; w++) { ro |= wave_or[w]; rc += wave_cnt[w]; }
        partials[2 * blockIdx.x] = ro;
        partials[2 * blockIdx.x + 1] = rc;
    }
}

// One workgroup: combines `blocks` partial results into out[0] (OR) and
// out[1] (count).
__global__ void __launch_bounds__(STATUS_BLOCK) rsg_status_final_kernel(const uint32_t* partials, int blocks,
                                                                        uint32_t* out) {
    __shared__ uint32_t red_or[STATUS_BLOCK], red_cnt[STATUS_BLOCK];
    uint32_t o = 0, c = 0;
    for (int b = (int)threadIdx.x; b < blocks; b += STATUS_BLOCK) {
        o |= partials[2 * b];
        c += partials[2 * b + 1];
    }
    red_or[threadIdx.x] = o;
    red_cnt[threadIdx.x] = c;
    __syncthreads();
    for (int half = STATUS_BLOCK / 2; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half) {
            red_or[threadIdx.x] |= red_or[threadIdx.x + half];
            red_cnt[threadIdx.x] += red_cnt[threadIdx.x + half];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red_or[0];
        out[1] = red_cnt[0];
    }
}

namespace rsg {

hipError_t launch_status_summary(const uint32_t* status, int n, uint32_t* partials, uint32_t* out, hipStream_t s) {
    const int blocks = status_blocks(n);
    if (blocks > 0)
        hipLaunchKernelGGL(rsg_status_partial_kernel, dim3(blocks), dim3(STATUS_BLOCK), 0, s, status, n, partials);
    hipLaunchKernelGGL(rsg_status_final_kernel, dim3(1), dim3(STATUS_BLOCK), 0, s, partials, blocks, out);
    return hipGetLastError();
}

}  // namespace rsg
