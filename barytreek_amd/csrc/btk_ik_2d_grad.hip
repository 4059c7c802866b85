// Interaction kernels, 2D BTK_KERNEL_POT_GRAD (one translation unit per kernel keeps builds parallel).
#include "btk_interact.hpp"

namespace btk {
template void launch_interact<2, BTK_KERNEL_POT_GRAD>(const InteractLaunch &, hipStream_t);
template void launch_queue<2, BTK_KERNEL_POT_GRAD>(const QueueLaunch &, hipStream_t);
}  // namespace btk
