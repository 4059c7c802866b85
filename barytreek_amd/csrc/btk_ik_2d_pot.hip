// Interaction kernels, 2D BTK_KERNEL_POT (one translation unit per kernel keeps builds parallel).
#include "btk_interact.hpp"

namespace btk {
template void launch_interact<2, BTK_KERNEL_POT>(const InteractLaunch &, hipStream_t);
template void launch_queue<2, BTK_KERNEL_POT>(const QueueLaunch &, hipStream_t);
}  // namespace btk
