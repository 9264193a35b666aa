// The batched TTI step: shots over a (shot, tile) work list, each tile advanced by the
// single-grid tile function of its lane layout (step_tile in fl_step.h or, with both
// fields per lane, step_tile2 in fl_step2.h).
#pragma once
#include <type_traits>

#include "fl_step2.h"

namespace fl {

// Shot of a batched run (device copy of the host descriptor).
struct ShotDev {
    float *lv[2];       // levels of (p, q) pairs; lv[parity] is the previous one
    const void *coef;   // coefficient records
    float *trace;
    long long src_idx;  // pitched cell index
    int sx, sy, sz, px, rec_iz;
    int ntx, nty, item0, nitems;  // tiles along x, y; first work item; item count
    int xorg;                     // x of the first tile column
    float vq;                     // CM_Q16R unit of the shot
};

struct BatchArgs {
    const ShotDev *shots;
    const int *item_shot;  // shot of each work item
    int nshots, nitems, parity, zt;
    float dxxinv, dyyinv, dzzinv, dxinv, dyinv, dzinv, dt;
    int bord;
    int *run;  // run control block {k0, k1, it_base, src_n}; this step reads run[par]
    int par;
    const float *src_tab;
    int rec_every, rec_slots;
    // persistent form: 8 ticket counters (one per slice of the work list) and 8 counters of
    // workgroups done; 0 selects one workgroup per item
    int *tick;
    int persist;
};

#ifndef FL_NO_AUX_KERNELS
// Source of the first step of a batched run: every shot's pc, qc at its source cell.
__global__ void batch_insert(const ShotDev *shots, int nshots, int parity, float s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nshots) return;
    const ShotDev &sh = shots[i];
    sh.lv[1 - parity][2 * sh.src_idx] += s;
    sh.lv[1 - parity][2 * sh.src_idx + 1] += s;
}
#endif

// The batched kernel keeps a shot descriptor besides the tile state (about 10 VGPRs more
// than the single-grid kernel): held to 80 for 6 waves per SIMD, the 16-bit and recompute
// forms spill a few values, nearly all outside the steady loop (BF16 on 32 shots: 83K ->
// 94K MSamples/s against 4 waves); the FP32 form would reload a dozen per 10 planes and
// keeps 4 waves.  FL_BATCH_WAVES overrides the 8- and 12-row value at build time.
template <int TY, int CM>
constexpr int batch_waves()
{
#ifdef FL_BATCH_WAVES
    return (TY == 8 || TY == 12) ? FL_BATCH_WAVES : 4;
#else
    return ((TY == 8 || TY == 12) && CM != CM_F32) ? 6 : 4;
#endif
}

// Batched shots over the (shot, tile) items of the work list.  PERSIST false: one
// workgroup per item.  PERSIST true: the work list is cut into 8 contiguous slices; the
// workgroups with blockIdx.x % 8 == x (those the dispatcher deals to one XCD) take the items
// of slice x in order from a ticket counter until it is exhausted, and the last of them to
// finish sets the slice's counters back to zero for the next step.  No workgroup waits for
// another, so the loop ends for every workgroup whatever the number resident.
// The shot descriptors and the item table come in as read-only, non-aliased kernel
// parameters as well: loads from them are then scalar (invariant over the kernel).
// Streaming hints of the batched step (the PCPOL of step_tile): on for FP32, whose two-array
// records stream whole lines (32 shots, 12-row tiles: 73.6K against 70.9K MSamples/s); off for
// the other forms (the both-fields-per-lane FP16-recompute step lost 2.4% with them).
template <int CM>
constexpr int batch_cpol()
{
    return CM == CM_F32 ? 2 : 0;
}

// PAIR: workgroups of TY/2 waves running step_tile2 (both fields per lane).
template <int TY, int CM, bool ISO, bool PERSIST, bool PAIR>
__global__ __launch_bounds__(PAIR ? 32 * TY : 64 * TY)
__attribute__((amdgpu_waves_per_eu(PAIR ? pair_waves<TY>() : batch_waves<TY, CM>())))
void tti_step_batch(const BatchArgs b, const ShotDev *__restrict__ shots, const int *__restrict__ item_shot)
{
    __shared__ std::conditional_t<PAIR, StepLds2<TY>, StepLds<TY>> s;
    const int k = uni(b.run[b.par]);
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) b.run[1 - b.par] = k + 1;
    const int it = uni(b.run[RUN_IT_BASE]) + 1 + k;
    const int src_n = uni(b.run[RUN_SRC_N]);
    // the table holds src_n >= 1 samples: the clamped load needs no branch
    const bool do_src = (k + 1) < src_n;
    const float s_next = uni(b.src_tab[min(k + 1, src_n - 1)]);
    auto item = [&](int t) {
        const ShotDev &sh = shots[uni(item_shot[t])];
        StepArgs a;
        a.prev = uni(sh.lv[b.parity]);
        a.cur = uni(sh.lv[1 - b.parity]);
        a.coef = uni(sh.coef);
        a.dvx = a.dvy = nullptr;
        a.dxxinv = b.dxxinv; a.dyyinv = b.dyyinv; a.dzzinv = b.dzzinv;
        a.dxinv = b.dxinv; a.dyinv = b.dyinv; a.dzinv = b.dzinv; a.dt = b.dt;
        a.vq = uni(sh.vq);
        a.sx = uni(sh.sx); a.sy = uni(sh.sy); a.sz = uni(sh.sz); a.bord = b.bord; a.px = uni(sh.px);
        a.ntx = uni(sh.ntx); a.nty = uni(sh.nty); a.zt = b.zt; a.nblk = uni(sh.nitems); a.xorg = uni(sh.xorg);
        a.rec_iz = uni(sh.rec_iz);
        a.src_idx = uni(sh.src_idx);
        StepCtl ctl;
        ctl.do_src = do_src;
        ctl.src_add = uni(do_src ? s_next : 0.0f);
        ctl.trace = nullptr;
        float *const trace = uni(sh.trace);
        if (trace && b.rec_every > 0 && it % b.rec_every == 0) {
            const int slot = it / b.rec_every - 1;
            if (slot >= 0 && slot < b.rec_slots)
                ctl.trace = trace + (long long)slot * (a.sx - 2 * b.bord) * (a.sy - 2 * b.bord);
        }
        ctl.trace = uni(ctl.trace);
        const int l = t - uni(sh.item0);
        const int tix = uni(l % a.ntx);
        const int rest = uni(l / a.ntx);
        if constexpr (PAIR)
            step_tile2<TY, CM, ISO, 0>(a, ctl, tix, uni(rest % a.nty), uni(rest / a.nty), false, s);
        else
            step_tile<TY, CM, false, ISO, batch_cpol<CM>()>(a, ctl, tix, uni(rest % a.nty), uni(rest / a.nty), false,
                                                            s);
    };
    if constexpr (!PERSIST) {
        const int lb = xcd_remap(blockIdx.x, gridDim.x);
        if (lb >= b.nitems) return;
        item(step_order(lb, b.nitems, it));
    } else {
        __shared__ int tk;
        const int x = blockIdx.x & 7, per = gridDim.x >> 3;  // the grid is whole rounds of 8
        const int lo = (int)((long long)b.nitems * x / 8), hi = (int)((long long)b.nitems * (x + 1) / 8);
        const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
        for (;;) {
            __syncthreads();  // every wave has read the previous ticket
            if (lead) tk = lo + atomicAdd(&b.tick[x], 1);
            __syncthreads();
            const int t = uni(tk);
            if (t >= hi) break;
            item(t);
        }
        if (lead && atomicAdd(&b.tick[8 + x], 1) == per - 1) {  // the slice's last workgroup
            b.tick[x] = 0;
            b.tick[8 + x] = 0;
        }
    }
}

}  // namespace fl
