// Kernel instantiations for one tile height FL_TY and one part (FL_PART_GRID or
// FL_PART_BATCH), with their launchers (declared in fl_launch.h).
#include <hip/hip_runtime.h>

#define FL_NO_AUX_KERNELS
#include "fl_launch.h"

#ifndef FL_TY
#error "FL_TY (4, 8, 12 or 16) must be defined"
#endif
#if !defined(FL_PART_GRID) && !defined(FL_PART_BATCH)
#error "FL_PART_GRID or FL_PART_BATCH must be defined"
#endif

#define FL_CAT2(a, b) a##b
#define FL_CAT(a, b) FL_CAT2(a, b)

namespace fl {

#ifdef FL_PART_GRID
template <int CM, bool TP, bool ISO, bool HINTS>
static void zm(const StepArgs &a, bool pair, hipStream_t s)
{
    const int g = (a.nblk + 7) / 8 * 8;  // whole rounds of the 8 XCDs (xcd_remap)
    if constexpr (has_pair_form(CM) && has_pair_tiles(FL_TY) && !TP) {
        if (pair) {
            hipLaunchKernelGGL((tti_step2<FL_TY, CM, ISO, HINTS>), dim3(g), dim3(64, FL_TY / 2), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((tti_step<FL_TY, CM, TP, ISO, HINTS>), dim3(g), dim3(64, FL_TY), 0, s, a);
}

template <int CM>
static void zm_form(const StepArgs &a, bool iso, bool hints, bool pair, hipStream_t s)
{
    if (iso) {
        if (hints) zm<CM, false, true, true>(a, pair, s);
        else zm<CM, false, true, false>(a, pair, s);
    } else {
        if (hints) zm<CM, false, false, true>(a, pair, s);
        else zm<CM, false, false, false>(a, pair, s);
    }
}

int FL_CAT(launch_grid_, FL_TY)(const StepArgs &a, int cm, bool twopass, bool iso, bool hints, bool pair,
                                hipStream_t s)
{
    if (pair && (twopass || !has_pair_form(cm) || !has_pair_tiles(FL_TY))) return -1;
    if (twopass) {  // the first-derivative volumes are scaled: no isotropic form
        if (cm != CM_F32) return -1;
        if (hints) zm<CM_F32, true, false, true>(a, false, s);
        else zm<CM_F32, true, false, false>(a, false, s);
        return 0;
    }
    switch (cm) {
    case CM_F32: zm_form<CM_F32>(a, iso, hints, false, s); return 0;
    case CM_BF16: zm_form<CM_BF16>(a, iso, hints, false, s); return 0;
    case CM_F16: zm_form<CM_F16>(a, iso, hints, false, s); return 0;
    case CM_F16R: zm_form<CM_F16R>(a, iso, hints, pair, s); return 0;
    case CM_F32R: zm_form<CM_F32R>(a, iso, hints, false, s); return 0;
    case CM_Q16R: zm_form<CM_Q16R>(a, iso, hints, pair, s); return 0;
    }
    return -1;
}
#endif

#ifdef FL_PART_BATCH
template <int CM, bool P, bool PAIR>
static void bt3(const BatchArgs &a, int grid, bool iso, hipStream_t s)
{
    const dim3 blk(64, PAIR ? FL_TY / 2 : FL_TY);
    if (iso)
        hipLaunchKernelGGL((tti_step_batch<FL_TY, CM, true, P, PAIR>), dim3(grid), blk, 0, s, a, a.shots,
                           a.item_shot);
    else
        hipLaunchKernelGGL((tti_step_batch<FL_TY, CM, false, P, PAIR>), dim3(grid), blk, 0, s, a, a.shots,
                           a.item_shot);
}

template <int CM, bool PAIR>
static void bt2(const BatchArgs &a, int grid, bool iso, hipStream_t s)
{
    if (a.persist) bt3<CM, true, PAIR>(a, grid, iso, s);
    else bt3<CM, false, PAIR>(a, grid, iso, s);
}

template <int CM>
static void bt(const BatchArgs &a, int grid, bool iso, bool pair, hipStream_t s)
{
    if constexpr (has_pair_form(CM) && has_pair_tiles(FL_TY)) {
        if (pair) {
            bt2<CM, true>(a, grid, iso, s);
            return;
        }
    }
    bt2<CM, false>(a, grid, iso, s);
}

int FL_CAT(launch_batch_, FL_TY)(const BatchArgs &a, int grid, int cm, bool iso, bool pair, hipStream_t s)
{
    if (pair && (!has_pair_form(cm) || !has_pair_tiles(FL_TY))) return -1;
    switch (cm) {
    case CM_F32: bt<CM_F32>(a, grid, iso, false, s); return 0;
    case CM_BF16: bt<CM_BF16>(a, grid, iso, false, s); return 0;
    case CM_F16: bt<CM_F16>(a, grid, iso, false, s); return 0;
    case CM_F16R: bt<CM_F16R>(a, grid, iso, pair, s); return 0;
    case CM_F32R: bt<CM_F32R>(a, grid, iso, false, s); return 0;
    case CM_Q16R: bt<CM_Q16R>(a, grid, iso, pair, s); return 0;
    }
    return -1;
}
#endif

}  // namespace fl
