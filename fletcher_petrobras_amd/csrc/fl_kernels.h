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
// Kernels here (the z-marching step kernel is in fl_step.h):
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
//                       lane's own v2px / v2pn.
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
// Direct 64-point DerCross form (original/sample.h:32-90), global loads through L1/L2.
__device__ __forceinline__ float cross_group(const float *__restrict__ p, long long i, long long s11,
                                             long long s21, int m, int n)
{
    const long long aa = n * s11, bb = m * s21;
    float v = p[i + bb + aa] - p[i + bb - aa] - p[i - bb + aa] + p[i - bb - aa];
    if (m != n) {
        const long long a2 = m * s11, b2 = n * s21;
        v = v + p[i + b2 + a2] - p[i + b2 - a2] - p[i - b2 + a2] + p[i - b2 - a2];
    }
    return v;
}

__device__ __forceinline__ float dercross_g(const float *__restrict__ p, long long i, long long s11,
                                            long long s21, float dinv)
{
    return (FL_L11 * cross_group(p, i, s11, s21, 1, 1) + FL_L12 * cross_group(p, i, s11, s21, 1, 2) +
            FL_L13 * cross_group(p, i, s11, s21, 1, 3) + FL_L14 * cross_group(p, i, s11, s21, 1, 4) +
            FL_L22 * cross_group(p, i, s11, s21, 2, 2) + FL_L23 * cross_group(p, i, s11, s21, 2, 3) +
            FL_L24 * cross_group(p, i, s11, s21, 2, 4) + FL_L33 * cross_group(p, i, s11, s21, 3, 3) +
            FL_L34 * cross_group(p, i, s11, s21, 3, 4) + FL_L44 * cross_group(p, i, s11, s21, 4, 4)) *
           dinv;
}

__device__ __forceinline__ float der2_g(const float *__restrict__ p, long long i, long long s, float d2inv)
{
    return (FL_K0 * p[i] + FL_K1 * (p[i + s] + p[i - s]) + FL_K2 * (p[i + 2 * s] + p[i - 2 * s]) +
            FL_K3 * (p[i + 3 * s] + p[i - 3 * s]) + FL_K4 * (p[i + 4 * s] + p[i - 4 * s])) *
           d2inv;
}

__device__ __forceinline__ float der1_g(const float *__restrict__ p, long long i, long long s, float dinv)
{
    return (FL_L1 * (p[i + s] - p[i - s]) + FL_L2 * (p[i + 2 * s] - p[i - 2 * s]) +
            FL_L3 * (p[i + 3 * s] - p[i - 3 * s]) + FL_L4 * (p[i + 4 * s] - p[i - 4 * s])) *
           dinv;
}

struct DirectArgs {
    float *prev;
    const float *cur;
    const float *coef;    // FP32 records of 10 per cell
    float dxxinv, dyyinv, dzzinv, dxyinv, dxzinv, dyzinv, dt;
    int sx, sy, sz, bord, px, zt;
    float *trace;
    int rec_iz;
};

// grid: (ceil(nx/64), ceil(ny/4), nzchunks); block (64, 4).  Fields are (p, q) pairs:
// element strides 2 (x), 2 px (y), 2 px sy (z); p at even, q at odd elements.
__global__ __launch_bounds__(256) void direct_dercross(const DirectArgs a)
{
    const int x = a.bord + blockIdx.x * 64 + threadIdx.x;
    const int y = a.bord + blockIdx.y * 4 + threadIdx.y;
    if (x >= a.sx - a.bord || y >= a.sy - a.bord) return;
    const int zlo = a.bord + blockIdx.z * a.zt;
    const int zhi = min(zlo + a.zt, a.sz - a.bord);
    const long long sX = 2, sY = 2ll * a.px, sZ = 2ll * a.px * a.sy;
    const float *pc = a.cur, *qc = a.cur + 1;
    for (int z = zlo; z < zhi; z++) {
        const long long c = ((long long)z * a.sy + y) * a.px + x, i = 2 * c;
        const float pxx = der2_g(pc, i, sX, a.dxxinv);
        const float pyy = der2_g(pc, i, sY, a.dyyinv);
        const float pzz = der2_g(pc, i, sZ, a.dzzinv);
        const float pxy = dercross_g(pc, i, sX, sY, a.dxyinv);
        const float pyz = dercross_g(pc, i, sY, sZ, a.dyzinv);
        const float pxz = dercross_g(pc, i, sX, sZ, a.dxzinv);
        const float qxx = der2_g(qc, i, sX, a.dxxinv);
        const float qyy = der2_g(qc, i, sY, a.dyyinv);
        const float qzz = der2_g(qc, i, sZ, a.dzzinv);
        const float qxy = dercross_g(qc, i, sX, sY, a.dxyinv);
        const float qyz = dercross_g(qc, i, sY, sZ, a.dyzinv);
        const float qxz = dercross_g(qc, i, sX, sZ, a.dxzinv);
        Coef k;
#pragma unroll
        for (int j = 0; j < 10; j++) k.c[rec_order32(j)] = a.coef[c * 10 + j];
        float ppv = a.prev[i], qpv = a.prev[i + 1];
        combine<CM_F32>(k, a.dt, pxx, pyy, pzz, pxy, pyz, pxz, qxx, qyy, qzz, qxy, qyz, qxz,
                        pc[i], qc[i], ppv, qpv);
        if (a.trace && z == a.rec_iz)
            a.trace[(long long)(y - a.bord) * (a.sx - 2 * a.bord) + (x - a.bord)] = ppv;
        a.prev[i] = ppv;
        a.prev[i + 1] = qpv;
    }
}

// ---------------------------------------------------------------------------------
// Pass 1 of the two-pass form: Dx of pc,qc over ix in [bord, sx-bord), all iy;
// Dy over iy in [bord, sy-bord), ix in [bord, sx); every plane iz in [0, sz).
// dvx, dvy: (pDx, qDx), (pDy, qDy) pairs at the pitched cell index.
// grid: (ceil(sx/64), ceil(sy/4), sz); block (64, 4)
__global__ __launch_bounds__(256) void der1_pass(const float *__restrict__ cur, float *dvx, float *dvy,
                                                 int sx, int sy, int px, int bord, float dxinv,
                                                 float dyinv)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= sx || y >= sy) return;
    const long long i = 2 * (((long long)z * sy + y) * px + x);
    if (x >= bord && x < sx - bord) {
        dvx[i] = der1_g(cur, i, 2, dxinv);
        dvx[i + 1] = der1_g(cur + 1, i, 2, dxinv);
    }
    if (y >= bord && y < sy - bord && x >= bord) {
        dvy[i] = der1_g(cur, i, 2ll * px, dyinv);
        dvy[i + 1] = der1_g(cur + 1, i, 2ll * px, dyinv);
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
