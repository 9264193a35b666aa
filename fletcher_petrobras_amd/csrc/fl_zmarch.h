// z-marching TTI step kernels (single grid and batched shots) for gfx950.
//
// Workgroup = (64, TY) threads = one 64 x TY tile of a z chunk [zlo, zhi).
// Per front plane f (from zlo-4 to zhi+3):
//   LDS   pc, qc of plane f over the tile plus a 4-cell halo in x and y;
//         Dx of pc, qc over the tile plus the 4-row y halo (for Dxy).
//   regs  9-deep z queues of pc, qc, Dx, Dy (own column): Dzz, Dxz, Dyz of plane
//         z = f-4 come from them (the on-chip form of der1der1lm's 9-plane ring,
//         der1der1lm/zTileSize.h:4-7, sample.h:147-160);
//         4-deep queue delaying Dxx, Dyy, Dxy of plane f until z = f.
// Plane z is then combined with its 10 coefficients and the pp/qp leapfrog.
#pragma once
#include <type_traits>

#include "fl_kernels.h"

namespace fl {

// Global loads through a wave-uniform base pointer plus a 32-bit lane byte offset.
__device__ __forceinline__ float ld_f32(const void *base, uint32_t off)
{
    return *(const float *)((const char *)base + off);
}

// Buffer resource over one z plane (wave-uniform base, range = the plane's bytes):
// loads/stores take a 32-bit lane offset and keep the address in SGPRs; accesses
// past the range read 0 / are dropped by the hardware.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
// Cache-policy bits of the once-per-step streams (coefficients, pp/qp): 0 default,
// 2 = nt (streaming: not kept in L2, which then holds the pc/qc halos).
#ifndef FL_STREAM_AUX
#define FL_STREAM_AUX 0
#endif
__device__ __forceinline__ float bld_f32(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, FL_STREAM_AUX));
}
__device__ __forceinline__ uint16_t bld_u16(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, FL_STREAM_AUX);
}
__device__ __forceinline__ void bst_f32(__amdgpu_buffer_rsrc_t r, uint32_t off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, FL_STREAM_AUX);
}

// Raw coefficient words of one cell as loaded (decoded only where they are used, so
// the loads stay in flight across barriers).
struct CoefRaw {
    uint32_t w[10];
};

// Coefficients of one cell: `plane` = coefficient volume 0 at the cell's z plane,
// volumes `vb` bytes apart, each plane `pb` bytes; `off` = lane byte offset.
template <int CM>
__device__ __forceinline__ void load_coef_raw(const char *plane, long long vb, uint32_t pb, uint32_t off,
                                              CoefRaw &r)
{
    constexpr int NC = (CM == CM_F16R) ? 6 : 10;
#pragma unroll
    for (int j = 0; j < NC; j++) {
        if constexpr (CM == CM_F32)
            r.w[j] = __builtin_bit_cast(uint32_t, bld_f32(plane_rsrc(plane + j * vb, pb), off));
        else
            r.w[j] = bld_u16(plane_rsrc(plane + j * vb, pb), off);
    }
}

template <int CM>
__device__ __forceinline__ void decode_coef(const CoefRaw &r, Coef &k)
{
    if constexpr (CM == CM_F32) {
#pragma unroll
        for (int j = 0; j < 10; j++) k.c[j] = __builtin_bit_cast(float, r.w[j]);
    } else if constexpr (CM == CM_BF16 || CM == CM_F16) {
#pragma unroll
        for (int j = 0; j < 10; j++)
            k.c[j] = (CM == CM_BF16) ? bf16_to_f32((uint16_t)r.w[j]) : f16_to_f32((uint16_t)r.w[j]);
    } else {  // CM_F16R: rebuild from vpz*dt, vsv*dt, eps, delta, theta, phi-pi
        float v[6];
#pragma unroll
        for (int j = 0; j < 6; j++) v[j] = f16_to_f32((uint16_t)r.w[j]);
        const float w2 = v[0] * v[0];
        float st, ct, spc, cpc;
        __sincosf(v[4], &st, &ct);
        __sincosf(v[5], &spc, &cpc);
        const float sp = -spc, cp = -cpc;  // phi = phic + pi
        const float st2 = st * st;
        const float s2t = 2.0f * st * ct;
        const float s2p = 2.0f * sp * cp;
        k.c[0] = st2 * cp * cp;
        k.c[1] = st2 * sp * sp;
        k.c[2] = ct * ct;
        k.c[3] = st2 * s2p;
        k.c[4] = s2t * sp;
        k.c[5] = s2t * cp;
        k.c[6] = w2 * (1.0f + 2.0f * v[2]);
        k.c[7] = w2;
        k.c[8] = v[1] * v[1];
        k.c[9] = w2 * (1.0f + 2.0f * v[3]);
    }
}

template <int CM>
__device__ __forceinline__ void load_coef_b(const char *plane, long long vb, uint32_t pb, uint32_t off,
                                            Coef &k)
{
    CoefRaw r;
    load_coef_raw<CM>(plane, vb, pb, off, r);
    decode_coef<CM>(r, k);
}

template <bool B>
using flag = std::integral_constant<bool, B>;

// Unroll factor of the steady-state plane loop (a multiple of 9 turns the z-queue
// shifts into register renaming, at the price of register pressure).
#ifndef FL_ZUNROLL
#define FL_ZUNROLL 1
#endif

template <int TY, int CM, bool TWOPASS>
__device__ __forceinline__ void zmarch_tile(const StepArgs &a, const StepCtl &ctl, int tix, int tiy,
                                            int zc, float *__restrict__ sp, float *__restrict__ sq,
                                            float *__restrict__ dxp, float *__restrict__ dxq)
{
    constexpr int TX = 64, HX = TX + 8, HY = TY + 8, NT = 64 * TY;
    constexpr int NLD = (HX * HY + NT - 1) / NT;  // halo-tile loads per thread per field
    constexpr int ES = (CM == CM_F32) ? 4 : 2;    // coefficient element bytes
    // tiles start at absolute multiples of TX so every row access is line-aligned
    // (for row pitches that are); lanes left of bord or right of sx-bord only load
    const int x0 = tix * TX, y0 = a.bord + tiy * TY;
    const int zlo = a.bord + zc * a.zt;
    const int zhi = min(zlo + a.zt, a.sz - a.bord);
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = x0 + lx, y = y0 + ly;
    const bool active = (x < a.sx - a.bord) && (y < a.sy - a.bord);
    // lanes outside the update domain read a valid in-domain cell and store nothing
    const int xr = min(x, a.sx - a.bord - 1), yr = min(y, a.sy - a.bord - 1);
    const long long sxy = (long long)a.sx * a.sy;
    const uint32_t col = (uint32_t)(yr * a.sx + xr);  // (x,y) offset within a plane
    const uint32_t col4 = col * 4u, colc = col * (uint32_t)ES;
    const long long vb = a.coef_stride * ES;          // bytes between coefficient volumes
    const uint32_t pb4 = (uint32_t)(sxy * 4), pbc = (uint32_t)(sxy * ES);  // plane bytes

    // halo-tile elements handled by this thread: in-plane byte offsets
    uint32_t hoff[NLD];
    bool hok[NLD];
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int e = tid + k * NT;
        const int r = e / HX, c = e - r * HX;
        const int gy = y0 - 4 + r, gx = x0 - 4 + c;
        hok[k] = (e < HX * HY) && gx >= 0 && gx < a.sx && gy < a.sy;
        hoff[k] = hok[k] ? (uint32_t)(gy * a.sx + gx) * 4u : 0u;
    }
    float pre_p[NLD], pre_q[NLD];
    auto fetch_plane = [&](int f) {
        const float *pcf = a.pc + (long long)f * sxy;
        const float *qcf = a.qc + (long long)f * sxy;
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const float vp = ld_f32(pcf, hoff[k]);
            const float vq = ld_f32(qcf, hoff[k]);
            pre_p[k] = hok[k] ? vp : 0.0f;
            pre_q[k] = hok[k] ? vq : 0.0f;
        }
    };
    const long long nvol = sxy * a.sz;

    float qpc[9], qqc[9], qpx[9], qpy[9], qqx[9], qqy[9];
#pragma unroll
    for (int k = 0; k < 9; k++) qpc[k] = qqc[k] = qpx[k] = qpy[k] = qqx[k] = qqy[k] = 0.0f;
    float dl[6][4];  // delayed in-plane terms: pxx pyy pxy qxx qyy qxy
#pragma unroll
    for (int j = 0; j < 6; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) dl[j][k] = 0.0f;

    const int f0 = zlo - 4, f1 = zhi + 3;
    if (!TWOPASS) fetch_plane(f0);

    // One front plane f: stage it, form Dx/Dy (and, when XY, the in-plane terms of f),
    // advance the queues and, when OUT, finish plane z = f-4.
    auto plane = [&](auto XYc, auto OUTc, const int f) {
        constexpr bool XY = decltype(XYc)::value, OUT = decltype(OUTc)::value;
        const long long zoff = (long long)(f - 4) * sxy;
        const __amdgpu_buffer_rsrc_t rpp = plane_rsrc(a.pp + zoff, pb4);
        const __amdgpu_buffer_rsrc_t rqp = plane_rsrc(a.qp + zoff, pb4);
        // output-plane loads issued before the LDS work
        Coef kc;
        float ppv = 0.0f, qpv = 0.0f;
        if constexpr (OUT) {
            load_coef_b<CM>((const char *)a.coef + zoff * ES, vb, pbc, colc, kc);
            ppv = bld_f32(rpp, col4);
            qpv = bld_f32(rqp, col4);
        }
        __syncthreads();  // readers of the previous plane are done
        float npc, nqc, npx, npy, nqx, nqy;  // values entering the z queues
        float pxx = 0.f, pyy = 0.f, pxy = 0.f, qxx = 0.f, qyy = 0.f, qxy = 0.f;
        if constexpr (!TWOPASS) {
#pragma unroll
            for (int k = 0; k < NLD; k++) {
                const int e = tid + k * NT;
                if (k + 1 < NLD || e < HX * HY) {
                    sp[e] = pre_p[k];
                    sq[e] = pre_q[k];
                }
            }
            if (f < f1) fetch_plane(f + 1);
            __syncthreads();
            const float *cp = sp + (ly + 4) * HX + lx + 4;
            const float *cq = sq + (ly + 4) * HX + lx + 4;
            npc = cp[0];
            nqc = cq[0];
            npy = der1_lds(cp, HX, a.dyinv);
            nqy = der1_lds(cq, HX, a.dyinv);
            if constexpr (XY) {
                pxx = der2_lds(cp, 1, a.dxxinv);
                pyy = der2_lds(cp, HX, a.dyyinv);
                qxx = der2_lds(cq, 1, a.dxxinv);
                qyy = der2_lds(cq, HX, a.dyyinv);
#pragma unroll
                for (int rr = 0; rr < HY; rr += TY) {
                    const int r = rr + ly;
                    if (rr + TY <= HY || r < HY) {
                        dxp[r * TX + lx] = der1_lds(sp + r * HX + lx + 4, 1, a.dxinv);
                        dxq[r * TX + lx] = der1_lds(sq + r * HX + lx + 4, 1, a.dxinv);
                    }
                }
                __syncthreads();
                const float *cdp = dxp + (ly + 4) * TX + lx;
                const float *cdq = dxq + (ly + 4) * TX + lx;
                npx = cdp[0];
                nqx = cdq[0];
                pxy = der1_lds(cdp, TX, a.dyinv);
                qxy = der1_lds(cdq, TX, a.dyinv);
            } else {
                npx = der1_lds(cp, 1, a.dxinv);
                nqx = der1_lds(cq, 1, a.dxinv);
            }
        } else {
            // pc/qc tile for Der2 in x and y; the Dx tile comes from the pass-1 volumes
            const long long fo = (long long)f * sxy;
            const float *pcf = a.pc + fo, *qcf = a.qc + fo;
            const float *pDx = a.dvol + fo, *pDy = a.dvol + nvol + fo;
            const float *qDx = a.dvol + 2 * nvol + fo, *qDy = a.dvol + 3 * nvol + fo;
#pragma unroll
            for (int k = 0; k < NLD; k++) {
                const int e = tid + k * NT;
                const float vp = ld_f32(pcf, hoff[k]);
                const float vq = ld_f32(qcf, hoff[k]);
                if (k + 1 < NLD || e < HX * HY) {
                    sp[e] = hok[k] ? vp : 0.0f;
                    sq[e] = hok[k] ? vq : 0.0f;
                }
            }
            if constexpr (XY) {
#pragma unroll
                for (int rr = 0; rr < HY; rr += TY) {
                    const int r = rr + ly;
                    const int gy = y0 - 4 + r;
                    const bool ok = x < a.sx && gy < a.sy;
                    const uint32_t o = ok ? (uint32_t)(gy * a.sx + x) * 4u : 0u;
                    const float vp = ld_f32(pDx, o), vq = ld_f32(qDx, o);
                    if (rr + TY <= HY || r < HY) {
                        dxp[r * TX + lx] = ok ? vp : 0.0f;
                        dxq[r * TX + lx] = ok ? vq : 0.0f;
                    }
                }
            }
            npx = ld_f32(pDx, col4);
            nqx = ld_f32(qDx, col4);
            npy = ld_f32(pDy, col4);
            nqy = ld_f32(qDy, col4);
            __syncthreads();
            const float *cp = sp + (ly + 4) * HX + lx + 4;
            const float *cq = sq + (ly + 4) * HX + lx + 4;
            npc = cp[0];
            nqc = cq[0];
            if constexpr (XY) {
                pxx = der2_lds(cp, 1, a.dxxinv);
                pyy = der2_lds(cp, HX, a.dyyinv);
                qxx = der2_lds(cq, 1, a.dxxinv);
                qyy = der2_lds(cq, HX, a.dyyinv);
                const float *cdp = dxp + (ly + 4) * TX + lx;
                const float *cdq = dxq + (ly + 4) * TX + lx;
                pxy = der1_lds(cdp, TX, a.dyinv);
                qxy = der1_lds(cdq, TX, a.dyinv);
            }
        }
        // advance the z queues (renamed away where the loop below is unrolled)
#pragma unroll
        for (int k = 0; k < 8; k++) {
            qpc[k] = qpc[k + 1]; qqc[k] = qqc[k + 1];
            qpx[k] = qpx[k + 1]; qpy[k] = qpy[k + 1];
            qqx[k] = qqx[k + 1]; qqy[k] = qqy[k + 1];
        }
        qpc[8] = npc; qqc[8] = nqc; qpx[8] = npx; qpy[8] = npy; qqx[8] = nqx; qqy[8] = nqy;
        float o6[6];
        const float nv[6] = {pxx, pyy, pxy, qxx, qyy, qxy};
#pragma unroll
        for (int j = 0; j < 6; j++) {
            o6[j] = dl[j][0];
            dl[j][0] = dl[j][1];
            dl[j][1] = dl[j][2];
            dl[j][2] = dl[j][3];
            dl[j][3] = nv[j];
        }
        if constexpr (OUT) {
            const float pzz = der2_q(qpc, a.dzzinv);
            const float qzz = der2_q(qqc, a.dzzinv);
            const float pxz = der1_q(qpx, a.dzinv);
            const float pyz = der1_q(qpy, a.dzinv);
            const float qxz = der1_q(qqx, a.dzinv);
            const float qyz = der1_q(qqy, a.dzinv);
            combine<CM>(kc, a.dt, o6[0], o6[1], pzz, o6[2], pyz, pxz, o6[3], o6[4], qzz, o6[5],
                        qyz, qxz, qpc[4], qqc[4], ppv, qpv);
            if (active) {
                if (ctl.trace && f - 4 == a.rec_iz)
                    ctl.trace[(y - a.bord) * (a.sx - 2 * a.bord) + (x - a.bord)] = ppv;
                if (ctl.do_src && zoff + col == a.src_idx) {
                    ppv += ctl.src_add;
                    qpv += ctl.src_add;
                }
                bst_f32(rpp, col4, ppv);
                bst_f32(rqp, col4, qpv);
            }
        }
    };

    // phases: [zlo-4,zlo) fill; [zlo,zlo+4) in-plane terms; [zlo+4,zhi) full planes;
    // then the last 4 outputs.  Short chunks (< 4 planes) skip parts of the middle.
    const int fB = zlo, fC = min(zhi, zlo + 4), fD = max(zhi, zlo + 4);
    int f = f0;
    for (; f < fB; f++) plane(flag<false>{}, flag<false>{}, f);
    for (; f < fC; f++) plane(flag<true>{}, flag<false>{}, f);
    for (; f < zlo + 4; f++) plane(flag<false>{}, flag<false>{}, f);
    if constexpr (FL_ZUNROLL > 1) {
        for (; f + FL_ZUNROLL <= zhi; f += FL_ZUNROLL) {
#pragma unroll
            for (int u = 0; u < FL_ZUNROLL; u++) plane(flag<true>{}, flag<true>{}, f + u);
        }
    }
    for (; f < zhi; f++) plane(flag<true>{}, flag<true>{}, f);
    for (f = fD; f <= f1; f++) plane(flag<false>{}, flag<true>{}, f);
}

template <int TY>
struct ZmLds {
    float sp[(TY + 8) * 72];
    float sq[(TY + 8) * 72];
    float dxp[(TY + 8) * 64];
    float dxq[(TY + 8) * 64];
};

// One grid: logical block t -> (tile x, tile y, z chunk), x fastest.
template <int TY, int CM, bool TWOPASS>
__global__ __launch_bounds__(64 * TY) void zmarch(const StepArgs a)
{
    __shared__ ZmLds<TY> s;
    if (a.it_next && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0)
        *a.it_next = (a.it_dev ? *a.it_dev : 0) + 1;
    const StepCtl ctl = step_ctl(a);
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    if (t >= a.nblk) return;
    const int tix = t % a.ntx;
    const int rest = t / a.ntx;
    zmarch_tile<TY, CM, TWOPASS>(a, ctl, tix, rest % a.nty, rest / a.nty, s.sp, s.sq, s.dxp, s.dxq);
}

// ---------------------------------------------------------------- batched shots
// Shot sub-grids keep their own fields (two time levels per field; `parity` says
// which one is "previous") and coefficients.  Items are (shot, z chunk, tile y,
// tile x), shot-major; a persistent grid walks them with stride gridDim.x.
struct ShotDev {
    float *p[2], *q[2];
    const void *coef;
    long long coef_stride;
    float *trace;
    long long src_idx;
    int sx, sy, sz, rec_iz;
    int ntx, nty, item0, nitems;
};

struct BatchArgs {
    const ShotDev *shots;
    int nshots, nitems, parity, zt;
    float dxxinv, dyyinv, dzzinv, dxinv, dyinv, dzinv, dt;
    int bord;
    int it_base;
    const int *it_dev;
    int *it_next;
    const float *src_tab;
    int src_n, rec_every, rec_slots;
};

template <int TY, int CM>
__global__ __launch_bounds__(64 * TY) void zmarch_batch(const BatchArgs b)
{
    __shared__ ZmLds<TY> s;
    const int k = *b.it_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *b.it_next = k + 1;
    const int it = b.it_base + 1 + k;
    for (int t = xcd_remap(blockIdx.x, gridDim.x); t < b.nitems; t += gridDim.x) {
        int lo = 0, hi = b.nshots - 1;  // shot owning item t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (b.shots[mid].item0 <= t) lo = mid; else hi = mid - 1;
        }
        const ShotDev &sh = b.shots[lo];
        StepArgs a;
        a.pp = sh.p[b.parity];
        a.pc = sh.p[1 - b.parity];
        a.qp = sh.q[b.parity];
        a.qc = sh.q[1 - b.parity];
        a.coef = sh.coef;
        a.coef_stride = sh.coef_stride;
        a.dvol = nullptr;
        a.dxxinv = b.dxxinv; a.dyyinv = b.dyyinv; a.dzzinv = b.dzzinv;
        a.dxinv = b.dxinv; a.dyinv = b.dyinv; a.dzinv = b.dzinv; a.dt = b.dt;
        a.sx = sh.sx; a.sy = sh.sy; a.sz = sh.sz; a.bord = b.bord;
        a.ntx = sh.ntx; a.nty = sh.nty; a.zt = b.zt; a.nblk = sh.nitems;
        a.rec_iz = sh.rec_iz;
        a.src_idx = sh.src_idx;
        StepCtl ctl;
        ctl.do_src = b.src_tab && (k + 1) < b.src_n;
        ctl.src_add = ctl.do_src ? b.src_tab[k + 1] : 0.0f;
        ctl.trace = nullptr;
        if (sh.trace && b.rec_every > 0 && it % b.rec_every == 0) {
            const int slot = it / b.rec_every - 1;
            if (slot >= 0 && slot < b.rec_slots)
                ctl.trace = sh.trace + (long long)slot * (sh.sx - 2 * b.bord) * (sh.sy - 2 * b.bord);
        }
        const int l = t - sh.item0;
        const int tix = l % sh.ntx;
        const int rest = l / sh.ntx;
        zmarch_tile<TY, CM, false>(a, ctl, tix, rest % sh.nty, rest / sh.nty, s.sp, s.sq, s.dxp, s.dxq);
    }
}

// Source of the first step of a batched run: every shot's pc, qc at its source cell.
__global__ void batch_insert(const ShotDev *shots, int nshots, int parity, float s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nshots) return;
    const ShotDev &sh = shots[i];
    sh.p[1 - parity][sh.src_idx] += s;
    sh.q[1 - parity][sh.src_idx] += s;
}

}  // namespace fl
