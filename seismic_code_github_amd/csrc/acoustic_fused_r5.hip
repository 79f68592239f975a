// One-launch tiled acoustic step, radius 5 instances.
#include "acoustic_fused_inst.h"

namespace sfd {
bool pick_fused_r5(int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b) {
    return fused_pick_r<5>(rows, nw, mode, ubf, k, b);
}
}  // namespace sfd
