// step_k5_g7.hip -- step and fix-up kernels for K=5 active slots, 7-point id gather.
#include "step_kernels.h"

namespace cupf {

hipError_t launch_step_k5_g7(const StepArgs& a, hipStream_t stream, const EdgeLaunch* edge,
                             hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3,
                             bool planned, const StepArgs* next)
{
    return launch_step_t<5, false>(a, stream, edge, e0, e1, e2, e3, planned, next);
}

hipError_t launch_batch_k5_g7(const BatchArgs& b, hipStream_t stream)
{
    return launch_batch_t<5, false>(b, stream);
}

int step_slots_k5_g7(int device, int which) { return slots_t<5, false>(device, which); }

}  // namespace cupf
