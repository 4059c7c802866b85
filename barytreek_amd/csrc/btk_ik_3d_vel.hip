// Interaction kernels, 3D BTK_KERNEL_REF_VEL3D (one translation unit per kernel keeps builds parallel).
#include "btk_interact.hpp"

namespace btk {
template void launch_interact<3, BTK_KERNEL_REF_VEL3D>(const InteractLaunch &, hipStream_t);
template void launch_queue<3, BTK_KERNEL_REF_VEL3D>(const QueueLaunch &, hipStream_t);
}  // namespace btk
