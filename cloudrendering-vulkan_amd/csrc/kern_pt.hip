// Launches of the path tracer: the single-job render and the batched render.
#include "pt_kernel.h"

#include "cr_internal_ppb.h"

namespace {

__global__ void __launch_bounds__(256) order_keys_kernel(const uint32_t* cost, uint32_t n, uint32_t* keys, uint32_t* vals)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    keys[i] = ~cost[i];
    vals[i] = i;
}

} // namespace

namespace cr {

uint32_t pt_order_counts(uint32_t n) { return sort_counts_words(n) * 4u; }

hipError_t launch_tile_order(const uint32_t* cost, uint32_t n, uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp,
                             uint32_t* vals_tmp, uint32_t* counts, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(order_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, s, cost, n, keys, vals);
    return launch_radix_sort(keys, vals, keys_tmp, vals_tmp, counts, n, s);
}

hipError_t launch_pt(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s)
{
    if (a.nFrames <= PT_FEW_FRAMES) return launch_pt_few(a, dtype, layout, blocks, s);
    BatchArgs none = {};
    return launch(pick<false, K_RENDER>(dtype, layout, a.ext.trackingMode, a.ext.neeMode), a, none, blocks, 256, 0, s);
}

int pt_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee, uint32_t n_frames)
{
    if (n_frames <= PT_FEW_FRAMES) return pt_few_blocks_per_cu(dtype, layout, tracking, nee);
    return occupancy(pick<false, K_RENDER>(dtype, layout, tracking, nee), 256, 0);
}

hipError_t launch_pt_batch(const BatchArgs& b, int dtype, int layout, uint32_t tracking, uint32_t nee, int blocks,
                           int block_threads, hipStream_t s)
{
    PTArgs none = {};
    return launch(pick<true, K_RENDER>(dtype, layout, tracking, nee), none, b, blocks, block_threads,
                  (size_t)b.nJobs * sizeof(PTArgs), s);
}

int pt_batch_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee, int block_threads, uint32_t n_jobs)
{
    return occupancy(pick<true, K_RENDER>(dtype, layout, tracking, nee), block_threads, (size_t)n_jobs * sizeof(PTArgs));
}

} // namespace cr
