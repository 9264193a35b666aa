// Launchers of the z-marching step kernels.  fl_inst.hip is compiled once per tile height
// (FL_TY = 4, 8, 12, 16) and per part (FL_PART_GRID: the single-grid step; FL_PART_BATCH:
// the batched step), so the kernel instantiations build in parallel translation units.
// Each launcher returns 0, or -1 when it holds no kernel for the requested form.
#pragma once
#include "fl_step.h"

namespace fl {

// One single-grid step: grid (nblk rounded up to whole XCD rounds) x (64, TY).
int launch_grid_4(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, hipStream_t s);
int launch_grid_8(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, hipStream_t s);
int launch_grid_12(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, hipStream_t s);
int launch_grid_16(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, hipStream_t s);

// One batched step over `grid` workgroups of (64, TY).
int launch_batch_4(const BatchArgs &a, int grid, int cm, bool iso, hipStream_t s);
int launch_batch_8(const BatchArgs &a, int grid, int cm, bool iso, hipStream_t s);
int launch_batch_12(const BatchArgs &a, int grid, int cm, bool iso, hipStream_t s);
int launch_batch_16(const BatchArgs &a, int grid, int cm, bool iso, hipStream_t s);

inline int launch_grid(int ty, const StepArgs &a, int cm, bool twopass, bool iso, bool hints, hipStream_t s)
{
    switch (ty) {
    case 4: return launch_grid_4(a, cm, twopass, iso, hints, s);
    case 8: return launch_grid_8(a, cm, twopass, iso, hints, s);
    case 12: return launch_grid_12(a, cm, twopass, iso, hints, s);
    case 16: return launch_grid_16(a, cm, twopass, iso, hints, s);
    }
    return -1;
}

inline int launch_batch(int ty, const BatchArgs &a, int grid, int cm, bool iso, hipStream_t s)
{
    switch (ty) {
    case 4: return launch_batch_4(a, grid, cm, iso, s);
    case 8: return launch_batch_8(a, grid, cm, iso, s);
    case 12: return launch_batch_12(a, grid, cm, iso, s);
    case 16: return launch_batch_16(a, grid, cm, iso, s);
    }
    return -1;
}

}  // namespace fl
