// Launch table of the one-launch steps for one radius: the tiled step
// (acoustic_fused.h, tiles of 16 or 32 rows) and the streaming step
// (acoustic_stream.h, chunks of 64 rows or more); each radius is compiled in
// its own translation unit.
#pragma once
#include "acoustic_fused.h"
#include "acoustic_plan.h"
#include "acoustic_stream.h"

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

template <int R, int NW, bool UBF>
bool stream_pick_kernel(int mode, StepKernel* k, BatchKernel* b) {
    switch (mode) {
#define SFD_STREAM(M)                                          \
    case M:                                                    \
        if (k) *k = acoustic_stream_step<R, NW, M, UBF>;       \
        if (b) *b = acoustic_stream_step_batch<R, NW, M, UBF>; \
        return true;
        SFD_STREAM(MODE_FAST_CPML)
        SFD_STREAM(MODE_FAST_SPONGE)
#undef SFD_STREAM
    default: return false;
    }
}

template <int R, int NW>
bool stream_pick_shape(int mode, bool ubf, StepKernel* k, BatchKernel* b) {
    return ubf ? stream_pick_kernel<R, NW, true>(mode, k, b) : stream_pick_kernel<R, NW, false>(mode, k, b);
}

template <int R>
bool fused_pick_r(int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b) {
    switch (rows * 100 + nw) {
    case 1604: return fused_pick_shape<R, 16, 4>(mode, k, b);
    case 1608: return fused_pick_shape<R, 16, 8>(mode, k, b);
    case 3204: return fused_pick_shape<R, 32, 4>(mode, k, b);
    case 3208: return fused_pick_shape<R, 32, 8>(mode, k, b);
    default: break;
    }
    // Streaming step: any chunk of kStreamMinRows rows or more.
    if (rows < kStreamMinRows) return false;
    switch (nw) {
    case 4: return stream_pick_shape<R, 4>(mode, ubf, k, b);
    case 8: return stream_pick_shape<R, 8>(mode, ubf, k, b);
    default: return false;
    }
}

}  // namespace sfd
