// Launches of the path tracer's learned-renderer passes: the record pass
// (camera paths up to their first real collision) and the continuation pass
// (paths resumed from recorded collisions).  Same kernel as the render.
#include "pt_kernel.h"

namespace cr {

hipError_t launch_pt_record(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s)
{
    BatchArgs none = {};
    return launch(pick<false, K_RECORD>(dtype, layout, a.ext.trackingMode, a.ext.neeMode), a, none, blocks, 256, 0, s);
}

hipError_t launch_pt_continue(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s)
{
    BatchArgs none = {};
    return launch(pick<false, K_CONTINUE>(dtype, layout, a.ext.trackingMode, a.ext.neeMode), a, none, blocks, 256, 0,
                  s);
}

int pt_learn_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee, int continuation)
{
    if (continuation) return occupancy(pick<false, K_CONTINUE>(dtype, layout, tracking, nee), 256, 0);
    return occupancy(pick<false, K_RECORD>(dtype, layout, tracking, nee), 256, 0);
}

} // namespace cr
