// Device code of the Fletcher TTI time step for gfx950 (CDNA4, wave64).
//
// One time step updates the coupled pseudo-acoustic fields p, q (reference
// original/sample.h:32-90 and der1der1/sample.h:35-123):
//   pp <- 2 pc - pp + dt^2 rhsp(pc, qc),   qp <- 2 qc - qp + dt^2 rhsq(pc, qc)
// over iz,iy,ix in [bord, s-bord) (original/OpenMP/openmp_propagate.c:27-29).
//
// Device layout of the wavefields: each time level is one array of (p, q) pairs,
// cell c = (z*sy + y)*px + x with the row pitch px >= sx a multiple of 16 cells
// (128-byte rows), p at 2c and q at 2c+1.  A step reads the current level (stencil) and
// the previous one (own cell) and overwrites the previous one; the roles then swap.
// The coefficients are records of one cell each, at the same pitched cell index.
//
// Kernels here (the z-marching step kernels are in fl_step.h, fl_step2.h and fl_step_batch.h):
//   direct_dercross  the reference's direct 64-point DerCross form
//       (original/derivatives.h:59-68), one thread per (x,y) column.
//   der1_pass        pass 1 of the two-pass form (der1der1/OpenMP/openmp_propagate.c:24-76).
//   insert_source    p[i] += s, q[i] += s (original/OpenMP/openmp_insertsource.c:7-12).
//   field_pack / field_unpack   one field between the reference layout and a level.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fl {

// 8th-order weights: the reference's float literals (original/derivatives.h:13-41).
#define FL_L1 0.8f
#define FL_L2 -0.2f
#define FL_L3 0.0380952380952381f
#define FL_L4 -0.0035714285714285713f
#define FL_L11 0.64f
#define FL_L12 -0.16f
#define FL_L13 0.03047619047619047618f
#define FL_L14 -0.00285714285714285713f
#define FL_L22 0.04f
#define FL_L23 -0.00761904761904761904f
#define FL_L24 0.00071428571428571428f
#define FL_L33 0.00145124716553287981f
#define FL_L34 -0.00013605442176870748f
#define FL_L44 0.00001275510204081632f
#define FL_K0 -2.84722222222222222222f
#define FL_K1 1.6f
#define FL_K2 -0.2f
#define FL_K3 0.02539682539682539682f
#define FL_K4 -0.00178571428571428571f

// Coefficient storage modes.
enum CoefMode : int {
    CM_F32 = 0,   // 10 FP32 volumes, precomp.h layout
    CM_BF16 = 1,  // 10 BF16 volumes; v2* pre-multiplied by dt^2
    CM_F16 = 2,   // 10 FP16 volumes; v2* pre-multiplied by dt^2
    CM_F16R = 3,  // 4 FP16 velocity factors (times dt^2) and theta, phi as u16 turns
    CM_F32R = 4,  // 4 FP32 velocity factors (precomp.h values) and theta, phi as FP32 turns
    CM_Q16R = 5,  // 4 velocity factors (times dt^2) in per-cell block floating point (15-bit
                  // mantissas, shared exponent), theta, phi as u16 turns
};

// Bytes of one cell's coefficient record.
__host__ __device__ constexpr int rec_bytes(int cm)
{
    return cm == CM_F32 ? 40 : (cm == CM_F16R || cm == CM_Q16R) ? 12 : cm == CM_F32R ? 24 : 20;
}

// Order of the precomp.h coefficients (0 ch1dxx, 1 ch1dyy, 2 ch1dzz, 3 ch1dxy, 4 ch1dyz,
// 5 ch1dxz, 6 v2px, 7 v2pz, 8 v2sz, 9 v2pn) within a stored record: slot s holds
// coefficient REC_ORDER[s].  The step kernel has the two lanes of a cell (p and q) load
// one half of the record each and broadcast it to the pair:
//   FP32 (40 B):        p lane {0 1 2 3 6}, q lane {4 5 7 8 9}; the fifth value is the
//                       lane's own v2px / v2pn.  On the device the record is held in two
//                       arrays of the pitched cells (fp32_part2): 32 bytes {0 1 2 3 4 5 7 8}
//                       at 32 c, then, after all cells, 8 bytes {6 9} at 32 np + 8 c.  A
//                       lane's two loads (16 and 4 bytes) then each read whole lines
//                       across the wave, which lets them stream non-temporally.
//   BF16 / FP16 (20 B): p lane {0 1 2 3}, q lane {4 5 7 8}, both {6 9}.
// The recompute inputs are v2px, v2pz, v2sz, v2pn, theta, phi (0..5), stored as
//   FP16 and block floating-point recompute (12 B): words (v2px, v2pn), (v2pz, v2sz),
//                          (theta, phi), loaded whole by both lanes;
//   FP32 recompute (24 B): p lane {v2px, v2pz, theta}, q lane {v2pn, v2sz, phi}.
__host__ __device__ constexpr int rec_order32(int s)
{
    constexpr int o[10] = {0, 1, 2, 3, 6, 4, 5, 7, 8, 9};
    return o[s];
}
// Byte offset of the 8-byte part of the FP32 records of a grid of np pitched cells.
__host__ __device__ constexpr long long fp32_part2(long long np) { return np * 32; }

__host__ __device__ constexpr int rec_order16(int s)
{
    constexpr int o[10] = {0, 1, 2, 3, 4, 5, 7, 8, 6, 9};
    return o[s];
}
__host__ __device__ constexpr int rec_order_r16(int s)
{
    constexpr int o[6] = {0, 3, 1, 2, 4, 5};
    return o[s];
}
__host__ __device__ constexpr int rec_order_r32(int s)
{
    constexpr int o[6] = {0, 1, 4, 3, 2, 5};
    return o[s];
}

struct StepArgs {
    float *prev;          // level read at the own cell and overwritten (pp, qp)
    const float *cur;     // level read with the stencil (pc, qc)
    const void *coef;     // coefficient records: per cell its values, consecutive
    const float *dvx, *dvy;  // TWOPASS: (pDx, qDx) and (pDy, qDy) pair volumes
    float dxxinv, dyyinv, dzzinv, dxinv, dyinv, dzinv, dt;
    float vq;             // CM_Q16R: 2^ebase, the unit of the block exponent 0
    int sx, sy, sz, bord, px;
    int ntx, nty, zt, nblk;  // tiles along x, y; z chunk depth; logical blocks
    int sw;                  // tile columns per strip of the workgroup order (tile_of)
    int xorg;                // x of the first tile column (0 or bord)
    // step index it: it_host for a plain launch.  Inside a multi-step run, run points at
    // the run's control block {k0, k1, it_base, src_n}: the step reads its counter
    // k = run[par], takes it = it_base + 1 + k and publishes run[1 - par] = k + 1 for the
    // following step, so a captured graph replays for any start step and run length.
    int it_host;
    int *run;
    int par;
    // source of the next step added to this step's output (multi-step runs): src_tab[k+1]
    // while k + 1 < src_n; src_idx is the pitched cell index
    const float *src_tab;
    long long src_idx;
    // receiver plane: slot it/rec_every-1 of trace when it % rec_every == 0
    float *trace;
    int rec_iz, rec_every, rec_slots;
};

// Per-launch scalars resolved from StepArgs at kernel entry.
struct StepCtl {
    float *trace;   // plane to write this step, or nullptr
    float src_add;
    bool do_src;
    int it;         // step index
};

// Wave-uniform copies (readfirstlane) of values the compiler computes or loads in vector
// registers although every lane holds the same value: integer divisions by runtime values
// (VALU only) and loads from memory the kernel also writes.  Without them such values
// occupy VGPRs for the whole tile.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ long long uni(long long v)
{
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
// Pointers into device memory: rebuilt as global (address space 1) pointers, so that
// accesses through them stay global_* instructions.  A generic pointer would make them
// flat_* instructions, which may complete out of order with the buffer loads and make the
// compiler drain the whole load queue (s_waitcnt vmcnt(0)) before the next use of any load.
template <class T>
__device__ __forceinline__ T *uni(T *p)
{
    typedef __attribute__((address_space(1))) T global_t;
    return (T *)(global_t *)(uintptr_t)uni((long long)(uintptr_t)p);
}

// Control block of a multi-step run (StepArgs.run, BatchArgs.run).
enum { RUN_K0 = 0, RUN_K1 = 1, RUN_IT_BASE = 2, RUN_SRC_N = 3, RUN_WORDS = 4 };

__device__ __forceinline__ StepCtl step_ctl(const StepArgs &a)
{
    StepCtl c;
    int k = 0, it = a.it_host, src_n = 0;
    if (a.run) {
        k = a.run[a.par];
        it = a.run[RUN_IT_BASE] + 1 + k;
        src_n = a.run[RUN_SRC_N];
    }
    c.do_src = a.src_tab && (k + 1) < src_n;
    c.src_add = c.do_src ? a.src_tab[k + 1] : 0.0f;
    c.trace = nullptr;
    if (a.trace && a.rec_every > 0 && it % a.rec_every == 0) {
        const int slot = it / a.rec_every - 1;
        if (slot >= 0 && slot < a.rec_slots)
            c.trace = a.trace + (long long)slot * (a.sx - 2 * a.bord) * (a.sy - 2 * a.bord);
    }
    c.trace = uni(c.trace);
    c.src_add = uni(c.src_add);
    c.it = uni(it);
    return c;
}

// The step kernels are compiled with -ffp-contract=off and fuse multiply-adds explicitly,
// so every copy of an expression rounds identically (results do not depend on tile or
// chunk shapes).  The sums keep the reference association order (derivatives.h:47,53):
// ((w1 d1 + w2 d2) + w3 d3) + w4 d4, each step one FMA.
// SC = false leaves the sum unscaled (dinv unused): the caller applies the grid spacing
// once to the combination instead.
template <bool SC = true>
__device__ __forceinline__ float fd1(float d1, float d2, float d3, float d4, float dinv)
{
    float t = FL_L1 * d1;
    t = __builtin_fmaf(FL_L2, d2, t);
    t = __builtin_fmaf(FL_L3, d3, t);
    t = __builtin_fmaf(FL_L4, d4, t);
    return SC ? t * dinv : t;
}
template <bool SC = true>
__device__ __forceinline__ float fd2(float c, float s1, float s2, float s3, float s4, float d2inv)
{
    float t = FL_K0 * c;
    t = __builtin_fmaf(FL_K1, s1, t);
    t = __builtin_fmaf(FL_K2, s2, t);
    t = __builtin_fmaf(FL_K3, s3, t);
    t = __builtin_fmaf(FL_K4, s4, t);
    return SC ? t * d2inv : t;
}

template <bool SC = true>
__device__ __forceinline__ float der1_lds(const float *p, int s, float dinv)
{
    return fd1<SC>(p[s] - p[-s], p[2 * s] - p[-2 * s], p[3 * s] - p[-3 * s], p[4 * s] - p[-4 * s], dinv);
}

template <bool SC = true>
__device__ __forceinline__ float der2_lds(const float *p, int s, float d2inv)
{
    return fd2<SC>(p[0], p[s] + p[-s], p[2 * s] + p[-2 * s], p[3 * s] + p[-3 * s], p[4 * s] + p[-4 * s],
               d2inv);
}

// q[0..8] = values at z-4..z+4
template <bool SC = true>
__device__ __forceinline__ float der1_q(const float *q, float dinv)
{
    return fd1<SC>(q[5] - q[3], q[6] - q[2], q[7] - q[1], q[8] - q[0], dinv);
}

template <bool SC = true>
__device__ __forceinline__ float der2_q(const float *q, float d2inv)
{
    return fd2<SC>(q[4], q[5] + q[3], q[6] + q[2], q[7] + q[1], q[8] + q[0], d2inv);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float f16_to_f32(uint16_t h)
{
    _Float16 v = __builtin_bit_cast(_Float16, h);
    return (float)v;
}

// The 10 per-cell coefficients, in precomp.h order; v2* include dt^2 when `scaled`.
struct Coef {
    float c[10];
};

// Coefficient combination + leapfrog, association order of original/sample.h:46-86.
template <int CM>
__device__ __forceinline__ void combine(const Coef &k, float dt, float pxx, float pyy, float pzz,
                                        float pxy, float pyz, float pxz, float qxx, float qyy,
                                        float qzz, float qxy, float qyz, float qxz, float pc,
                                        float qc, float &pp, float &qp)
{
    const float h1p = k.c[0] * pxx + k.c[1] * pyy + k.c[2] * pzz + k.c[3] * pxy + k.c[5] * pxz +
                      k.c[4] * pyz;
    const float h2p = pxx + pyy + pzz - h1p;
    const float h1q = k.c[0] * qxx + k.c[1] * qyy + k.c[2] * qzz + k.c[3] * qxy + k.c[5] * qxz +
                      k.c[4] * qyz;
    const float h2q = qxx + qyy + qzz - h1q;
    const float h1pmq = h1p - h1q;
    const float h2pmq = h2p - h2q;
    const float rhsp = k.c[6] * h2p + k.c[7] * h1q + k.c[8] * h1pmq;
    const float rhsq = k.c[9] * h2p + k.c[7] * h1q - k.c[8] * h2pmq;
    if constexpr (CM == CM_F32) {
        pp = 2.0f * pc - pp + rhsp * dt * dt;
        qp = 2.0f * qc - qp + rhsq * dt * dt;
    } else {  // dt^2 folded into the stored v2* factors
        pp = 2.0f * pc - pp + rhsp;
        qp = 2.0f * qc - qp + rhsq;
    }
}

// Logical block id from the hardware one.  Workgroups are dealt to the 8 XCDs
// round-robin (block b runs on XCD b%8); this gives each XCD a contiguous range of
// logical tiles so neighbouring tiles (shared halos) meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int g)
{
    return (b & 7) * (g >> 3) + (b >> 3);
}

// Odd steps take the logical blocks in reverse (t -> nblk-1-t) and march each z chunk from
// its top plane down: a step then starts on the tiles, and in each tile on the planes, that
// the step before ended on, whose levels were the last to pass through the memory-side
// cache.  Each tile's result does not depend on the order or the direction.  At 344^3 the
// reversed order made the FP32 step 1.2% faster (12-row tiles) and the downward march another
// 0.8%; FP16 recompute (8-row tiles, both fields per lane) 1.6% and 0.9%.
__device__ __forceinline__ int step_order(int t, int n, int it) { return (it & 1) ? n - 1 - t : t; }
__device__ __forceinline__ bool step_down(int it) { return (it & 1) != 0; }

// Tile (tx, ty, tz) of logical block t: z chunks outermost; inside a chunk, strips of sw
// tile columns (the last one narrower) walked row by row.  The tiles an XCD holds at once
// (consecutive t) then form a compact block of the plane instead of part of one long row,
// so the halos they share meet in that XCD's L2.  sw >= ntx is the plain row-major order.
__device__ __forceinline__ void tile_of(int t, int ntx, int nty, int sw, int &tx, int &ty, int &tz)
{
    const int per = ntx * nty;
    tz = t / per;
    const int r = t - tz * per;
    const int x0 = r / (sw * nty) * sw;
    const int w = min(sw, ntx - x0);
    const int u = r - x0 * nty;
    ty = u / w;
    tx = x0 + (u - ty * w);
}

// ---------------------------------------------------------------------------------
// Kernels launched by the driver translation unit only (the kernel instantiation units
// define FL_NO_AUX_KERNELS).
#ifndef FL_NO_AUX_KERNELS

struct DirectArgs {
    float *prev;
    const float *cur;
    const float *coef;    // FP32 records (the two arrays of rec_order32's note)
    long long np;         // pitched cells of the grid
    float dxxinv, dyyinv, dzzinv, dxyinv, dxzinv, dyzinv, dt;
    int sx, sy, sz, bord, px, zt;
    int ntx, nty, nblk;   // tiles of 64 x DX_TY cells along x, y; z chunks of zt planes
    float *trace;
    int rec_iz;
};

// Both fields of a cell, so that one expression evaluates the p and the q derivative with
// the same operations in the same order.
struct PQ {
    float p, q;
};
__device__ __forceinline__ PQ operator+(PQ a, PQ b) { return {a.p + b.p, a.q + b.q}; }
__device__ __forceinline__ PQ operator-(PQ a, PQ b) { return {a.p - b.p, a.q - b.q}; }
__device__ __forceinline__ PQ operator*(float s, PQ a) { return {s * a.p, s * a.q}; }
__device__ __forceinline__ PQ operator*(PQ a, float s) { return {a.p * s, a.q * s}; }

// Der2 and DerCross of derivatives.h:53,59-68 in the reference's association order, over an
// accessor: g(m) is the value m cells along the axis, h(a, b) the value a cells along the
// first axis (s11) and b along the second (s21).
template <class G>
__device__ __forceinline__ PQ der2_a(G g, float d2inv)
{
    return (FL_K0 * g(0) + FL_K1 * (g(1) + g(-1)) + FL_K2 * (g(2) + g(-2)) + FL_K3 * (g(3) + g(-3)) +
            FL_K4 * (g(4) + g(-4))) *
           d2inv;
}
template <class H>
__device__ __forceinline__ PQ cross_group_a(H h, int m, int n)
{
    PQ v = h(n, m) - h(-n, m) - h(n, -m) + h(-n, -m);
    if (m != n) v = v + h(m, n) - h(-m, n) - h(m, -n) + h(-m, -n);
    return v;
}
template <class H>
__device__ __forceinline__ PQ dercross_a(H h, float dinv)
{
    return (FL_L11 * cross_group_a(h, 1, 1) + FL_L12 * cross_group_a(h, 1, 2) + FL_L13 * cross_group_a(h, 1, 3) +
            FL_L14 * cross_group_a(h, 1, 4) + FL_L22 * cross_group_a(h, 2, 2) + FL_L23 * cross_group_a(h, 2, 3) +
            FL_L24 * cross_group_a(h, 2, 4) + FL_L33 * cross_group_a(h, 3, 3) + FL_L34 * cross_group_a(h, 3, 4) +
            FL_L44 * cross_group_a(h, 4, 4)) *
           dinv;
}

// dercross_a as a running sum over the coefficient groups (the same additions in the same
// order), with compiler barriers after the groups of L13, L23 and L44.
template <class H>
__device__ __forceinline__ PQ dercross_runs(H h, float dinv)
{
    PQ s = FL_L11 * cross_group_a(h, 1, 1);
    s = s + FL_L12 * cross_group_a(h, 1, 2);
    s = s + FL_L13 * cross_group_a(h, 1, 3);
    asm volatile("" ::: "memory");
    s = s + FL_L14 * cross_group_a(h, 1, 4);
    s = s + FL_L22 * cross_group_a(h, 2, 2);
    s = s + FL_L23 * cross_group_a(h, 2, 3);
    asm volatile("" ::: "memory");
    s = s + FL_L24 * cross_group_a(h, 2, 4);
    s = s + FL_L33 * cross_group_a(h, 3, 3);
    s = s + FL_L34 * cross_group_a(h, 3, 4);
    s = s + FL_L44 * cross_group_a(h, 4, 4);
    asm volatile("" ::: "memory");
    return s * dinv;
}

// Direct 64-point DerCross form (original/sample.h:32-90) staged through LDS.  A workgroup
// of DX_TY waves owns 64 x DX_TY cells of a z chunk; a thread computes both fields of its
// cell.  The 9 planes of the stencil (tile + 4-cell halo, (p, q) pairs) sit in a ring of
// DX_NP plane slots; the next plane is stored into the slot that no step of the current plane
// reads, so one barrier a plane suffices.  Every derivative, the coefficient combination and the
// leapfrog keep the reference's association order: the results are those of the global-
// memory form, bit for bit.
constexpr int DX_TX = 64, DX_TY = 4, DX_NP = 10;
constexpr int DX_HX = DX_TX + 8, DX_HY = DX_TY + 8, DX_PL = DX_HX * DX_HY;  // pairs per staged plane
constexpr int DX_NLD = (DX_PL + 64 * DX_TY - 1) / (64 * DX_TY);             // staging loads per thread
constexpr int DX_PER_CU = (160 * 1024) / (DX_NP * DX_PL * 8);                // resident workgroups a CU

// grid: nblk rounded up to whole XCD rounds; block (64, DX_TY).  Fields are (p, q) pairs at
// the pitched cell index.
// DX_PER_CU workgroups a CU (the LDS ring bounds them): DX_PER_CU * DX_TY / 4 waves per SIMD
__global__ __launch_bounds__(64 * DX_TY) __attribute__((amdgpu_waves_per_eu(DX_PER_CU * DX_TY / 4)))
void direct_dercross(const DirectArgs a)
{
    __shared__ float2 lds[DX_NP][DX_PL];
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    if (t >= a.nblk) return;
    const int tx = t % a.ntx, rest = t / a.ntx, ty = rest % a.nty, tz = rest / a.nty;
    const int x0 = a.bord + tx * DX_TX, y0 = a.bord + ty * DX_TY;
    const int zlo = a.bord + tz * a.zt, zhi = min(zlo + a.zt, a.sz - a.bord);
    const int f0 = zlo - 4;  // first staged plane
    const int c = threadIdx.x, r = threadIdx.y, tid = r * 64 + c;
    const int x = x0 + c, y = y0 + r;
    const bool active = x < a.sx - a.bord && y < a.sy - a.bord;
    const long long pxy = (long long)a.px * a.sy;
    const float2 *cur2 = (const float2 *)a.cur;

    // staging: pair e of a staged plane is (row e / DX_HX, column e % DX_HX) of the window
    // whose corner is (x0 - 4, y0 - 4); cells outside the array read as zero
    int goff[DX_NLD];  // pitched cell offset within a plane (planes hold < 2^31 cells)
    bool gok[DX_NLD];
#pragma unroll
    for (int k = 0; k < DX_NLD; k++) {
        const int e = k * 64 * DX_TY + tid;
        const int row = e / DX_HX, col = e - row * DX_HX;
        const int gx = x0 - 4 + col, gy = y0 - 4 + row;
        gok[k] = e < DX_PL && gx < a.sx && gy < a.sy;
        goff[k] = gy * a.px + gx;
    }
    auto fetch = [&](int zp, float2 *v) {
        const float2 *pl = cur2 + zp * pxy;
#pragma unroll
        for (int k = 0; k < DX_NLD; k++) v[k] = (gok[k] && zp < a.sz) ? pl[goff[k]] : make_float2(0.f, 0.f);
    };
    auto stage = [&](int slot, const float2 *v) {
#pragma unroll
        for (int k = 0; k < DX_NLD; k++) {
            const int e = k * 64 * DX_TY + tid;
            if (e < DX_PL) lds[slot][e] = v[k];
        }
    };
    for (int i = 0; i < 9; i++) {
        float2 v[DX_NLD];
        fetch(f0 + i, v);
        stage(i, v);
    }
    __syncthreads();

    const int own = (r + 4) * DX_HX + (c + 4);  // own cell in a staged plane
    for (int z = zlo; z < zhi; z++) {
        const int j = z - zlo;
        if (active) {
            const float2 *pl[9];
#pragma unroll
            for (int k = 0; k < 9; k++) pl[k] = &lds[(j + k) % DX_NP][own];
            // both fields at once (8-byte LDS reads), one derivative after the other and the
            // cross derivatives in three runs of coefficient groups, with compiler barriers
            // between them: at most 24 pairs of stencil values are live at once
            auto at = [](const float2 *b, int o) { const float2 v = b[o]; return PQ{v.x, v.y}; };
            const PQ xx = der2_a([&](int m) { return at(pl[4], m); }, a.dxxinv);
            const PQ yy = der2_a([&](int m) { return at(pl[4], m * DX_HX); }, a.dyyinv);
            const PQ zz = der2_a([&](int m) { return at(pl[4 + m], 0); }, a.dzzinv);
            asm volatile("" ::: "memory");
            const PQ xy = dercross_runs([&](int u, int v) { return at(pl[4], u + v * DX_HX); }, a.dxyinv);
            const PQ yz = dercross_runs([&](int u, int v) { return at(pl[4 + v], u * DX_HX); }, a.dyzinv);
            const PQ xz = dercross_runs([&](int u, int v) { return at(pl[4 + v], u); }, a.dxzinv);
            const long long cell = (long long)z * pxy + (long long)y * a.px + x;
            const float4 *r1 = (const float4 *)(a.coef + cell * 8);  // slots 0-3, 5-8
            const float2 r2 = ((const float2 *)((const char *)a.coef + fp32_part2(a.np)))[cell];  // 4, 9
            const float4 u = r1[0], w = r1[1];
            Coef k;
            k.c[rec_order32(0)] = u.x; k.c[rec_order32(1)] = u.y;
            k.c[rec_order32(2)] = u.z; k.c[rec_order32(3)] = u.w;
            k.c[rec_order32(4)] = r2.x;
            k.c[rec_order32(5)] = w.x; k.c[rec_order32(6)] = w.y;
            k.c[rec_order32(7)] = w.z; k.c[rec_order32(8)] = w.w;
            k.c[rec_order32(9)] = r2.y;
            float2 *pv = (float2 *)a.prev + cell;
            float2 o = *pv;
            const PQ cc = {pl[4][0].x, pl[4][0].y};
            combine<CM_F32>(k, a.dt, xx.p, yy.p, zz.p, xy.p, yz.p, xz.p, xx.q, yy.q, zz.q, xy.q, yz.q, xz.q,
                            cc.p, cc.q, o.x, o.y);
            if (a.trace && z == a.rec_iz)
                a.trace[(long long)(y - a.bord) * (a.sx - 2 * a.bord) + (x - a.bord)] = o.x;
            *pv = o;
        }
        if (z + 5 <= zhi + 3) {  // the next plane into the slot no step of this plane reads
            float2 nxt[DX_NLD];
            fetch(z + 5, nxt);
            stage((j + 9) % DX_NP, nxt);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// Pass 1 of the two-pass form: Dx of pc,qc over ix in [bord, sx-bord), all iy;
// Dy over iy in [bord, sy-bord), ix in [bord, sx); every plane iz in [0, sz).
// dvx, dvy: (pDx, qDx), (pDy, qDy) pairs at the pitched cell index.
// grid: (ceil(sx/64), ceil(sy/4), sz); block (64, 4)
// First derivative of both fields of the pair array along a stride of s cells, in the
// association order of derivatives.h:47.
__device__ __forceinline__ PQ der1_pq(const float2 *__restrict__ c2, long long i, long long s, float dinv)
{
    auto g = [&](long long m) {
        const float2 v = c2[i + m * s];
        return PQ{v.x, v.y};
    };
    return (FL_L1 * (g(1) - g(-1)) + FL_L2 * (g(2) - g(-2)) + FL_L3 * (g(3) - g(-3)) + FL_L4 * (g(4) - g(-4))) *
           dinv;
}

// The four volumes are written once and read by the next kernel after the whole pass, far beyond
// what the caches hold: stored non-temporally (the pass took 492 instead of 529 us at 384^3).
__global__ __launch_bounds__(256) void der1_pass(const float *__restrict__ cur, float *dvx, float *dvy,
                                                 int sx, int sy, int px, int bord, float dxinv,
                                                 float dyinv)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= sx || y >= sy) return;
    const long long c = ((long long)z * sy + y) * px + x;  // pitched cell: pairs at 2c
    const float2 *c2 = (const float2 *)cur;
    if (x >= bord && x < sx - bord) {
        const PQ d = der1_pq(c2, c, 1, dxinv);
        __builtin_nontemporal_store(d.p, dvx + 2 * c);
        __builtin_nontemporal_store(d.q, dvx + 2 * c + 1);
    }
    if (y >= bord && y < sy - bord && x >= bord) {
        const PQ d = der1_pq(c2, c, px, dyinv);
        __builtin_nontemporal_store(d.p, dvy + 2 * c);
        __builtin_nontemporal_store(d.q, dvy + 2 * c + 1);
    }
}

// Source insertion at pitched cell c of a level: p += s, q += s.
__global__ void insert_source(float *lv, long long c, float s)
{
    if (threadIdx.x == 0) {
        lv[2 * c] += s;
        lv[2 * c + 1] += s;
    }
}

// One field (fld 0 p, 1 q) of a level <-> a volume in the reference layout.
// grid: (ceil(sx/256), sy, sz); block 256
__global__ __launch_bounds__(256) void field_pack(const float *__restrict__ lin, float *lv, int fld,
                                                  int sx, int sy, int px)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= sx) return;
    const long long row = (long long)blockIdx.z * sy + blockIdx.y;
    lv[2 * (row * px + x) + fld] = lin[row * sx + x];
}
__global__ __launch_bounds__(256) void field_unpack(const float *__restrict__ lv, float *lin, int fld,
                                                    int sx, int sy, int px)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= sx) return;
    const long long row = (long long)blockIdx.z * sy + blockIdx.y;
    lin[row * sx + x] = lv[2 * (row * px + x) + fld];
}

#endif  // FL_NO_AUX_KERNELS

}  // namespace fl
