// Launches of the path tracer: the single-job render and the batched render.
#include "pt_kernel.h"

namespace cr {

hipError_t launch_pt(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s)
{
    BatchArgs none = {};
    return launch(pick<false, K_RENDER>(dtype, layout, a.ext.trackingMode, a.ext.neeMode), a, none, blocks, 256, 0, s);
}

int pt_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee)
{
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
