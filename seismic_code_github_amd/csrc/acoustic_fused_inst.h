// Launch table of the one-launch tiled step (acoustic_fused.h) for one
// radius; each radius is compiled in its own translation unit.
#pragma once
#include "acoustic_fused.h"
#include "acoustic_plan.h"

namespace sfd {

template <int R, int T, int NW>
bool fused_pick_shape(int mode, StepKernel* k, BatchKernel* b) {
    switch (mode) {
#define SFD_FUSED(M)                                         \
    case M:                                                  \
        if (k) *k = acoustic_fused_step<R, T, NW, M>;        \
        if (b) *b = acoustic_fused_step_batch<R, T, NW, M>;  \
        return true;
        SFD_FUSED(MODE_FAST_CPML)
        SFD_FUSED(MODE_FAST_SPONGE)
#undef SFD_FUSED
    default: return false;
    }
}

template <int R>
bool fused_pick_r(int rows, int nw, int mode, StepKernel* k, BatchKernel* b) {
    switch (rows * 100 + nw) {
    case 1604: return fused_pick_shape<R, 16, 4>(mode, k, b);
    case 1608: return fused_pick_shape<R, 16, 8>(mode, k, b);
    case 3204: return fused_pick_shape<R, 32, 4>(mode, k, b);
    case 3208: return fused_pick_shape<R, 32, 8>(mode, k, b);
    default: return false;
    }
}

}  // namespace sfd
