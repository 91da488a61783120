// Launch of the path tracer for renders of a few frames per pixel (at most
// PT_FEW_FRAMES): the render kernel with gated pixel claims, in a unit of its
// own so that it compiles beside kern_pt.hip.
#include "pt_kernel.h"

namespace cr {

hipError_t launch_pt_few(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s)
{
    BatchArgs none = {};
    return launch(pick<false, K_RENDER_FEW>(dtype, layout, a.ext.trackingMode, a.ext.neeMode), a, none, blocks, 256, 0,
                  s);
}

int pt_few_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee)
{
    return occupancy(pick<false, K_RENDER_FEW>(dtype, layout, tracking, nee), 256, 0);
}

} // namespace cr
