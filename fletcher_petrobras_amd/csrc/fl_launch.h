// Launchers of the z-marching step kernels.  fl_inst.hip is compiled once per tile height
// (FL_TY = 4, 8, 12, 16) and per part (FL_PART_GRID: the single-grid step; FL_PART_BATCH:
// the batched step), so the kernel instantiations build in parallel translation units.
// Each launcher returns 0, or -1 when it holds no kernel for the requested form.
#pragma once
#include "fl_step_batch.h"

namespace fl {

// One single-grid step: grid (nblk rounded up to whole XCD rounds) x (64, TY), or with pair
// (both fields per lane: the 16-bit recompute forms only) x (64, TY/2).
int launch_grid_4(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, bool pair, hipStream_t s);
int launch_grid_8(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, bool pair, hipStream_t s);
int launch_grid_12(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, bool pair, hipStream_t s);
int launch_grid_16(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, bool pair, hipStream_t s);

// One batched step over `grid` workgroups of (64, TY), or (64, TY/2) with pair.
int launch_batch_4(const BatchArgs &a, int grid, int cm, bool iso, bool pair, hipStream_t s);
int launch_batch_8(const BatchArgs &a, int grid, int cm, bool iso, bool pair, hipStream_t s);
int launch_batch_12(const BatchArgs &a, int grid, int cm, bool iso, bool pair, hipStream_t s);
int launch_batch_16(const BatchArgs &a, int grid, int cm, bool iso, bool pair, hipStream_t s);

// Forms and tile heights that have the both-fields-per-lane kernel: 8-row tiles only.  At
// up to 168 VGPRs a SIMD holds three of its waves; the 4-wave workgroups of 8-row tiles then
// fill a CU with three workgroups, one wave of each per SIMD.  The 6- and 8-wave workgroups
// of 12- and 16-row tiles left a CU room for one workgroup (its waves are not spread evenly
// enough over the SIMDs for a second at 3 per SIMD) and ran 25-35% slower than the
// one-field-per-lane kernel; 4-row tiles (two halo row pairs a wave) spilled in the loop.
constexpr bool has_pair_form(int cm) { return cm == CM_F16R || cm == CM_Q16R; }
constexpr bool has_pair_tiles(int ty) { return ty == 8; }

inline int launch_grid(int ty, const StepArgs &a, int cm, bool twopass, bool iso, bool hints, bool pair,
                       hipStream_t s)
{
    switch (ty) {
    case 4: return launch_grid_4(a, cm, twopass, iso, hints, pair, s);
    case 8: return launch_grid_8(a, cm, twopass, iso, hints, pair, s);
    case 12: return launch_grid_12(a, cm, twopass, iso, hints, pair, s);
    case 16: return launch_grid_16(a, cm, twopass, iso, hints, pair, s);
    }
    return -1;
}

inline int launch_batch(int ty, const BatchArgs &a, int grid, int cm, bool iso, bool pair, hipStream_t s)
{
    switch (ty) {
    case 4: return launch_batch_4(a, grid, cm, iso, pair, s);
    case 8: return launch_batch_8(a, grid, cm, iso, pair, s);
    case 12: return launch_batch_12(a, grid, cm, iso, pair, s);
    case 16: return launch_batch_16(a, grid, cm, iso, pair, s);
    }
    return -1;
}

}  // namespace fl
