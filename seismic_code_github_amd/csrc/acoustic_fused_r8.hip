// One-launch tiled acoustic step, radius 8 instances.
#include "acoustic_fused_inst.h"

namespace sfd {
bool pick_fused_r8(int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b) {
    return fused_pick_r<8>(rows, nw, mode, ubf, k, b);
}
}  // namespace sfd
