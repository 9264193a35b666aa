// Split-field z-marching TTI step (the drop-in step kernel) for gfx950.
//
// A wave covers 32 columns of one row for BOTH fields: lanes 0-31 carry p, lanes 32-63
// carry q of the same 32 cells.  A workgroup of TY waves is a 32 x TY tile of one z chunk.
// Each lane keeps only its own field's state, so the register footprint is half that
// of a one-lane-per-cell layout:
//   LDS   plane f = z+4 of pc and qc over the tile plus a 4-cell halo (per field);
//         Dx of each field over the tile plus the 4-row y halo (for Dxy).
//   regs  9-deep z queues of the own field's c, Dx, Dy (Dzz, Dxz, Dyz of plane z:
//         der1der1lm's 9-plane ring, der1der1lm/zTileSize.h:4-7, held on chip);
//         a 4-deep queue delaying Dxx, Dyy, Dxy of plane f until z = f.
// The coupling terms of plane z (original/sample.h:75-81) need h1q on the p lane and
// h2p on the q lane: one cross-lane swap (lane l <-> l^32).  Coefficients shared by
// p and q are loaded by both halves from the same 32 addresses (one 128-byte segment).
//
// TWOPASS reads Dx, Dy from the four first-derivative volumes of der1_pass
// (der1der1/OpenMP/openmp_propagate.c:24-76) instead of forming them on chip.
#pragma once
#include "fl_zmarch.h"

namespace fl {

// 1: the output-plane streams (coefficients, pp/qp) are loaded one plane before use
// (measured slower on gfx950: the rotated registers force waits at the barriers)
#ifndef FL_PREFETCH_AHEAD
#define FL_PREFETCH_AHEAD 0
#endif

template <int TY>
struct SfLds {
    float t[2][(TY + 8) * 40];  // pc, qc tile with halo
    float dx[2][(TY + 8) * 32]; // Dx of pc, qc: tile columns, rows with y halo
};

template <int TY, int CM, bool TWOPASS>
__device__ __forceinline__ void sf_tile(const StepArgs &a, const StepCtl &ctl, int tix, int tiy, int zc,
                                        SfLds<TY> &s)
{
    constexpr int TX = 32, HX = TX + 8, HY = TY + 8, NT = 64 * TY;
    constexpr int HT = HX * HY;                  // halo-tile elements per field
    constexpr int NLD = (2 * HT + NT - 1) / NT;  // staging loads per thread (both fields)
    constexpr int ES = (CM == CM_F32) ? 4 : 2;
    // tiles start at absolute multiples of TX so every row access is line-aligned
    // (for row pitches that are); lanes left of bord or right of sx-bord only load
    const int x0 = tix * TX, y0 = a.bord + tiy * TY;
    const int zlo = a.bord + zc * a.zt;
    const int zhi = min(zlo + a.zt, a.sz - a.bord);
    const int lane = threadIdx.x, ly = threadIdx.y, tid = ly * 64 + lane;
    const int fld = lane >> 5, cx = lane & 31;  // field (0 p, 1 q) and tile column
    const int x = x0 + cx, y = y0 + ly;
    const bool active = (x >= a.bord) && (x < a.sx - a.bord) && (y < a.sy - a.bord);
    const int xr = min(max(x, a.bord), a.sx - a.bord - 1), yr = min(y, a.sy - a.bord - 1);
    const long long sxy = (long long)a.sx * a.sy;
    const uint32_t col = (uint32_t)(yr * a.sx + xr);
    const uint32_t col4 = col * 4u, colc = col * (uint32_t)ES;
    const long long vb = a.coef_stride * ES;
    const uint32_t pb4 = (uint32_t)(sxy * 4), pbc = (uint32_t)(sxy * ES);

    // staging elements of this thread: field, in-plane byte offset, validity
    uint32_t hoff[NLD];
    bool hok[NLD], hq[NLD];
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int e = tid + k * NT;
        const int ef = e >= HT ? 1 : 0, ee = e - ef * HT;
        const int r = ee / HX, c = ee - r * HX;
        const int gy = y0 - 4 + r, gx = x0 - 4 + c;
        hq[k] = ef;
        hok[k] = (e < 2 * HT) && gx >= 0 && gx < a.sx && gy < a.sy;
        hoff[k] = hok[k] ? (uint32_t)(gy * a.sx + gx) * 4u : 0u;
    }
    float pre[NLD];  // raw staging loads; halo cells outside the grid are zeroed at the LDS store
    auto fetch_plane = [&](int f) {
        const float *pcf = a.pc + (long long)f * sxy;
        const float *qcf = a.qc + (long long)f * sxy;
#pragma unroll
        for (int k = 0; k < NLD; k++) pre[k] = ld_f32(hq[k] ? qcf : pcf, hoff[k]);
    };
    const long long nvol = sxy * a.sz;
    float *const tflat = &s.t[0][0];  // both field tiles, p then q
    float *const tl = s.t[fld];
    float *const dl_lds = s.dx[fld];

    float qc_[9], qx_[9], qy_[9];  // own field: c, Dx, Dy at z-4..z+4
#pragma unroll
    for (int k = 0; k < 9; k++) qc_[k] = qx_[k] = qy_[k] = 0.0f;
    float dl[3][4];  // delayed Dxx, Dyy, Dxy of the own field
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) dl[j][k] = 0.0f;

    const int f0 = zlo - 4, f1 = zhi + 3;
    if (!TWOPASS) fetch_plane(f0);

    // coefficients and previous fields of the next output plane, kept raw while in flight
    CoefRaw kr_n;
    float pp_n = 0.0f, qp_n = 0.0f;
    auto load_out = [&](int z, CoefRaw &kr, float &vp, float &vq) {
        const long long zo = (long long)z * sxy;
        load_coef_raw<CM>((const char *)a.coef + zo * ES, vb, pbc, colc, kr);
        vp = bld_f32(plane_rsrc(a.pp + zo, pb4), col4);
        vq = bld_f32(plane_rsrc(a.qp + zo, pb4), col4);
    };

    auto plane = [&](auto XYc, auto OUTc, const int f) {
        constexpr bool XY = decltype(XYc)::value, OUT = decltype(OUTc)::value;
        const long long zoff = (long long)(f - 4) * sxy;
        const __amdgpu_buffer_rsrc_t rpp = plane_rsrc(a.pp + zoff, pb4);
        const __amdgpu_buffer_rsrc_t rqp = plane_rsrc(a.qp + zoff, pb4);
        CoefRaw kr;
        float vpp = 0.0f, vqp = 0.0f;
#if FL_PREFETCH_AHEAD
        // streams of plane z were issued one plane earlier; issue those of z+1 now
        if constexpr (OUT) {
            kr = kr_n;
            vpp = pp_n;
            vqp = qp_n;
        }
        if (f - 3 >= zlo && f - 3 < zhi) load_out(f - 3, kr_n, pp_n, qp_n);
#else
        if constexpr (OUT) load_out(f - 4, kr, vpp, vqp);
#endif
        __syncthreads();  // readers of the previous plane are done
        float nc, nx, ny;  // own field values entering the z queues
        float xx = 0.f, yy = 0.f, xy = 0.f;
        if constexpr (!TWOPASS) {
#pragma unroll
            for (int k = 0; k < NLD; k++) {
                const int e = tid + k * NT;
                if (k + 1 < NLD || e < 2 * HT) tflat[e] = hok[k] ? pre[k] : 0.0f;
            }
            if (f < f1) fetch_plane(f + 1);
        } else {
            const long long fo = (long long)f * sxy;
            const float *pcf = a.pc + fo, *qcf = a.qc + fo;
#pragma unroll
            for (int k = 0; k < NLD; k++) {
                const int e = tid + k * NT;
                const float v = ld_f32(hq[k] ? qcf : pcf, hoff[k]);
                if (k + 1 < NLD || e < 2 * HT) tflat[e] = hok[k] ? v : 0.0f;
            }
            const float *Dx = a.dvol + (fld ? 2 * nvol : 0) + fo;
            const float *Dy = a.dvol + (fld ? 3 * nvol : nvol) + fo;
            if constexpr (XY) {
#pragma unroll
                for (int rr = 0; rr < HY; rr += TY) {
                    const int r = rr + ly;
                    const int gy = y0 - 4 + r;
                    const bool ok = x < a.sx && gy < a.sy;
                    const float v = ld_f32(Dx, ok ? (uint32_t)(gy * a.sx + x) * 4u : 0u);
                    if (rr + TY <= HY || r < HY) dl_lds[r * TX + cx] = ok ? v : 0.0f;
                }
            }
            nx = ld_f32(Dx, col4);
            ny = ld_f32(Dy, col4);
        }
        __syncthreads();
        const float *cp = tl + (ly + 4) * HX + cx + 4;
        nc = cp[0];
        if constexpr (!TWOPASS) ny = der1_lds(cp, HX, a.dyinv);
        if constexpr (XY) {
            xx = der2_lds(cp, 1, a.dxxinv);
            yy = der2_lds(cp, HX, a.dyyinv);
            if constexpr (!TWOPASS) {
#pragma unroll
                for (int rr = 0; rr < HY; rr += TY) {
                    const int r = rr + ly;
                    if (rr + TY <= HY || r < HY) dl_lds[r * TX + cx] = der1_lds(tl + r * HX + cx + 4, 1, a.dxinv);
                }
                __syncthreads();
            }
            const float *cd = dl_lds + (ly + 4) * TX + cx;
            if constexpr (!TWOPASS) nx = cd[0];
            xy = der1_lds(cd, TX, a.dyinv);
        } else if constexpr (!TWOPASS) {
            nx = der1_lds(cp, 1, a.dxinv);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            qc_[k] = qc_[k + 1];
            qx_[k] = qx_[k + 1];
            qy_[k] = qy_[k + 1];
        }
        qc_[8] = nc; qx_[8] = nx; qy_[8] = ny;
        float o3[3];
        const float nv[3] = {xx, yy, xy};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            o3[j] = dl[j][0];
            dl[j][0] = dl[j][1];
            dl[j][1] = dl[j][2];
            dl[j][2] = dl[j][3];
            dl[j][3] = nv[j];
        }
        if constexpr (OUT) {
            Coef kc;
            decode_coef<CM>(kr, kc);
            const float prev = fld ? vqp : vpp;
            const float zz = der2_q(qc_, a.dzzinv);
            const float xz = der1_q(qx_, a.dzinv);
            const float yz = der1_q(qy_, a.dzinv);
            // H1, H2 of the own field, association order of original/sample.h:46-53
            float h1 = kc.c[0] * o3[0];
            h1 = __builtin_fmaf(kc.c[1], o3[1], h1);
            h1 = __builtin_fmaf(kc.c[2], zz, h1);
            h1 = __builtin_fmaf(kc.c[3], o3[2], h1);
            h1 = __builtin_fmaf(kc.c[5], xz, h1);
            h1 = __builtin_fmaf(kc.c[4], yz, h1);
            const float h2 = o3[0] + o3[1] + zz - h1;
            // p lane sends h2p and receives h1q; q lane sends h1q and receives h2p
            const float other = __shfl_xor(fld ? h1 : h2, 32);
            float rhs;
            if (fld == 0) {  // rhsp = v2px h2p + v2pz h1q + v2sz (h1p - h1q)
                rhs = kc.c[6] * h2;
                rhs = __builtin_fmaf(kc.c[7], other, rhs);
                rhs = __builtin_fmaf(kc.c[8], h1 - other, rhs);
            } else {         // rhsq = v2pn h2p + v2pz h1q - v2sz (h2p - h2q)
                rhs = kc.c[9] * other;
                rhs = __builtin_fmaf(kc.c[7], h1, rhs);
                rhs = __builtin_fmaf(-kc.c[8], other - h2, rhs);
            }
            const float cur = qc_[4];
            const float base = 2.0f * cur - prev;
            float nw;
            if constexpr (CM == CM_F32) nw = __builtin_fmaf(rhs * a.dt, a.dt, base);
            else nw = base + rhs;
            if (active) {
                if (fld == 0 && ctl.trace && f - 4 == a.rec_iz)
                    ctl.trace[(y - a.bord) * (a.sx - 2 * a.bord) + (x - a.bord)] = nw;
                if (ctl.do_src && zoff + col == a.src_idx) nw += ctl.src_add;
                if (fld == 0) bst_f32(rpp, col4, nw);
                else bst_f32(rqp, col4, nw);
            }
        }
    };

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

// Minimum waves per SIMD requested from the register allocator (0: compiler default).
#ifndef FL_WAVES_PER_EU
#define FL_WAVES_PER_EU 0
#endif
#if FL_WAVES_PER_EU > 0
#define FL_WPE_ATTR __attribute__((amdgpu_waves_per_eu(FL_WAVES_PER_EU)))
#else
#define FL_WPE_ATTR
#endif

template <int TY, int CM, bool TWOPASS>
__global__ __launch_bounds__(64 * TY) FL_WPE_ATTR void zmarch_sf(const StepArgs a)
{
    __shared__ SfLds<TY> s;
    if (a.it_next && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0)
        *a.it_next = (a.it_dev ? *a.it_dev : 0) + 1;
    const StepCtl ctl = step_ctl(a);
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    if (t >= a.nblk) return;
    const int tix = t % a.ntx;
    const int rest = t / a.ntx;
    sf_tile<TY, CM, TWOPASS>(a, ctl, tix, rest % a.nty, rest / a.nty, s);
}

// Batched shots on the split-field tile (items as for zmarch_batch, 32-column tiles).
template <int TY, int CM>
__global__ __launch_bounds__(64 * TY) void zmarch_batch_sf(const BatchArgs b)
{
    __shared__ SfLds<TY> s;
    const int k = *b.it_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *b.it_next = k + 1;
    const int it = b.it_base + 1 + k;
    for (int t = xcd_remap(blockIdx.x, gridDim.x); t < b.nitems; t += gridDim.x) {
        int lo = 0, hi = b.nshots - 1;
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
        sf_tile<TY, CM, false>(a, ctl, tix, rest % sh.nty, rest / sh.nty, s);
    }
}

}  // namespace fl
