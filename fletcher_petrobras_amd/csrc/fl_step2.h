// The TTI step with both fields of a cell in one lane, for the 16-bit recompute forms
// (CM_F16R, CM_Q16R): the same z-marching as fl_step.h, the same expressions in the same
// association order, so the results are bitwise those of tti_step.
//
// Why a second lane layout.  The recompute forms are bound by VALU issue and LDS traffic,
// not by HBM.  With one field per lane (tti_step) both lanes of a cell decode the record
// and form the six angular coefficients, and the pair exchanges records and H1/H2 through
// DPP moves and per-lane selects.  With one cell per lane each is done once per cell, and
// the two fields' stencils share every LDS read (pairs, 8 bytes a lane): about 23% fewer
// VALU instructions and 30% less LDS traffic per cell.
//
// Work decomposition.  A workgroup of TY/2 waves owns a 32 x TY tile of the xy plane and a
// z chunk; wave w covers tile rows 2w (lanes 0-31) and 2w+1 (lanes 32-63), lane l the
// column l % 32 of its row, both fields.
//
// On chip, per lane:
//   registers  10-slot queues of c and Dy (pairs) over planes f-8 .. f, as tti_step;
//              5-slot queues of Dxx, Dyy: formed at the plane f they belong to, from the
//              neighbours Dx and Dy read anyway, and used four planes later when f is the
//              output plane (the staged plane is read once per step instead of twice);
//   LDS        2 staged planes of (pc, qc) pairs (tile + 4-cell halo): plane f, read in the
//              step, and plane f+1, written in it; 10 planes of Dx (tile rows + 4-row y
//              halo): Dxy(z) = Der1_y(Dx(z)) and Dxz(z) = Der1_z(Dx) (the own column of
//              planes z-4 .. z+3; z+4 is the step's own) read from them.  Keeping Dx there
//              instead of in a register queue frees the 20 VGPRs that keep the steady loop
//              free of spills.
// Every ring position is a compile-time constant in the ten unrolled steps of the steady
// loop (2 and 10 divide 10, and so does the 5 of the Dxx, Dyy queues), so no LDS address is
// formed at run time there.
#pragma once
#include "fl_step.h"

namespace fl {

// (p, q) of one cell; 8-byte aligned so that one LDS access moves both.
struct __attribute__((aligned(8))) F2 {
    float p, q;
};

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// (the elements are copied out first: clang 22 lowers __builtin_bit_cast of an element of
// an ext_vector_type value as a read of element 0)
__device__ __forceinline__ F2 f2_of(u32x2 v)
{
    const uint32_t lo = v[0], hi = v[1];
    return F2{__builtin_bit_cast(float, lo), __builtin_bit_cast(float, hi)};
}
__device__ __forceinline__ void fl_pin(F2 &v)
{
    fl_pin(v.p);
    fl_pin(v.q);
}

// First and second derivative of both fields along one axis from the same neighbours
// (stride s in pairs), each sum as der1_lds / der2_lds forms it.
template <bool SC, bool D2>
__device__ __forceinline__ void der12_pr(const F2 *c, int s, F2 ctr, float dinv, float d2inv, F2 &d1, F2 &d2)
{
    // neighbours 1, 2 then 3, 4: the sums advance in their association order, and at most
    // four pairs of LDS reads are in flight
    F2 a0 = c[s], b0 = c[-s], a1 = c[2 * s], b1 = c[-2 * s];
    float tp = FL_L1 * (a0.p - b0.p), tq = FL_L1 * (a0.q - b0.q);
    tp = __builtin_fmaf(FL_L2, a1.p - b1.p, tp);
    tq = __builtin_fmaf(FL_L2, a1.q - b1.q, tq);
    float up = 0.f, uq = 0.f;
    if constexpr (D2) {
        up = FL_K0 * ctr.p;
        uq = FL_K0 * ctr.q;
        up = __builtin_fmaf(FL_K1, a0.p + b0.p, up);
        uq = __builtin_fmaf(FL_K1, a0.q + b0.q, uq);
        up = __builtin_fmaf(FL_K2, a1.p + b1.p, up);
        uq = __builtin_fmaf(FL_K2, a1.q + b1.q, uq);
    }
    __builtin_amdgcn_sched_barrier(0);
    const F2 a2 = c[3 * s], b2 = c[-3 * s], a3 = c[4 * s], b3 = c[-4 * s];
    tp = __builtin_fmaf(FL_L3, a2.p - b2.p, tp);
    tq = __builtin_fmaf(FL_L3, a2.q - b2.q, tq);
    tp = __builtin_fmaf(FL_L4, a3.p - b3.p, tp);
    tq = __builtin_fmaf(FL_L4, a3.q - b3.q, tq);
    d1.p = SC ? tp * dinv : tp;
    d1.q = SC ? tq * dinv : tq;
    if constexpr (D2) {
        up = __builtin_fmaf(FL_K3, a2.p + b2.p, up);
        uq = __builtin_fmaf(FL_K3, a2.q + b2.q, uq);
        up = __builtin_fmaf(FL_K4, a3.p + b3.p, up);
        uq = __builtin_fmaf(FL_K4, a3.q + b3.q, uq);
        d2.p = SC ? up * d2inv : up;
        d2.q = SC ? uq * d2inv : uq;
    }
}
template <bool SC>
__device__ __forceinline__ F2 der1_pr(const F2 *c, int s, float dinv)
{
    F2 d1, d2;
    der12_pr<SC, false>(c, s, F2{0.f, 0.f}, dinv, 0.f, d1, d2);
    return d1;
}

// The decoded record of one cell: the six angular coefficients (precomp.h order 0-5), the
// velocity factors and, for CM_Q16R, 2^e.  Each value is the product tti_step forms in one
// of the two lanes (decode_rec), with the same operands.  zs = -1 negates c[4] and c[5]
// exactly (the coefficients of the z first derivatives, for a downward march).
struct CoefPair {
    float c[6];
    float own_p, own_q, vpz, vsz, e2;
};

template <int CM>
__device__ __forceinline__ void decode_pair(const CoefRec<CM> &r, CoefPair &k, float zs)
{
    static_assert(CM == CM_F16R || CM == CM_Q16R, "recompute records only");
    if constexpr (CM == CM_F16R) {  // words (v2px, v2pn), (v2pz, v2sz), (theta, phi)
        k.own_p = f16_to_f32((uint16_t)r.w[0]);
        k.own_q = f16_to_f32((uint16_t)(r.w[0] >> 16));
        k.vpz = f16_to_f32((uint16_t)r.w[1]);
        k.vsz = f16_to_f32((uint16_t)(r.w[1] >> 16));
        k.e2 = 1.0f;
    } else {  // 15-bit mantissas and the cell's exponent e (bits 30-31 of the two words)
        const uint32_t e = (r.w[0] >> 30) | ((r.w[1] >> 28) & 0xcu);
        k.e2 = __builtin_bit_cast(float, (e + 127u) << 23);
        k.own_p = (float)(r.w[0] & 0x7fffu);
        k.own_q = (float)((r.w[0] >> 15) & 0x7fffu);
        k.vpz = (float)(r.w[1] & 0x7fffu);
        k.vsz = (float)((r.w[1] >> 15) & 0x7fffu);
    }
    const float tt = (float)(r.w[2] & 0xffffu) * (1.0f / 65536.0f);  // theta, turns
    const float tp = (float)(r.w[2] >> 16) * (1.0f / 65536.0f);      // phi
    const float st = __builtin_amdgcn_sinf(tt), ct = __builtin_amdgcn_cosf(tt);
    const float sp = __builtin_amdgcn_sinf(tp), cp = __builtin_amdgcn_cosf(tp);
    const float st2 = st * st, ct2 = ct * ct, s2t = (st + st) * (ct * zs);  // sin^2, cos^2, +-sin 2 theta
    const float sp2 = sp * sp, cp2 = cp * cp, s2p = (sp + sp) * cp;  // the same of phi
    k.c[0] = st2 * cp2;  // sin^2 theta cos^2 phi
    k.c[1] = st2 * sp2;  // sin^2 theta sin^2 phi
    k.c[2] = ct2;        // cos^2 theta
    k.c[3] = st2 * s2p;  // sin^2 theta sin 2phi
    k.c[4] = s2t * sp;   // sin 2theta sin phi
    k.c[5] = s2t * cp;   // sin 2theta cos phi
}

template <int TY>
struct StepLds2 {  // 51 KB for 8-row tiles
    static constexpr int HX = 40, HY = TY + 8;
    F2 t[2][HY * HX];  // staged planes: (row, column) of the tile + halo
    F2 x[10][HY * 32];  // Dx: (row, tile column), tile rows + y halo
};

// One tile of one z chunk (the work of step_tile, fields paired per lane), upwards or, with
// down, from the top plane down as step_tile does it.
template <int TY, int CM, bool ISO, int PCPOL>
__device__ __forceinline__ void step_tile2(const StepArgs &a, const StepCtl &ctl, int tix, int tiy, int zc,
                                           bool down, StepLds2<TY> &s)
{
    static_assert(TY % 2 == 0, "two tile rows per wave");
    constexpr int TX = 32, HX = 40, HY = TY + 8, W = TY / 2, NT = 64 * W;
    constexpr int NSP = HX * HY;              // pairs per staged plane
    constexpr int NLD = (NSP + NT - 1) / NT;  // staging loads per lane
    constexpr int REC = rec_bytes(CM);
    constexpr int NH = (4 + W - 1) / W;       // y-halo row pairs of Dx per wave (at most)
    constexpr int PD = 2;  // records and previous level two steps ahead, in two register sets
    const int x0 = a.xorg + tix * TX, y0 = a.bord + tiy * TY;
    const int zlo = a.bord + zc * a.zt;
    const int zhi = min(zlo + a.zt, a.sz - a.bord);
    const int nz = zhi - zlo, f0 = zlo - 4;
    const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int cx = lane & 31, rs = lane >> 5;  // column, row of the pair of rows
    const int x = x0 + cx, y = y0 + 2 * wv + rs;
    const bool active = (x >= a.bord) && (x < a.sx - a.bord) && (y < a.sy - a.bord);
    const float rq = ISO ? a.dxxinv * a.vq : a.vq;
    const uint32_t pxy = (uint32_t)a.px * (uint32_t)a.sy;
    const uint32_t pb8 = pxy * 8u, pbr = pxy * (uint32_t)REC;
    const uint32_t col = (uint32_t)(y * a.px + x);
    const uint32_t vo8 = active ? col * 8u : FL_OOB;  // own (p, q) pair
    RecOff vor;
    vor.voff = active ? col * (uint32_t)REC : FL_OOB;
    vor.voff2 = 0;

    const long long zin = (long long)f0 * pxy, zout = (long long)zlo * pxy;
    const __amdgpu_buffer_rsrc_t r_cur = span_rsrc(a.cur + 2 * zin, (uint32_t)(nz + 8) * pb8);
    const __amdgpu_buffer_rsrc_t r_prev = span_rsrc(a.prev + 2 * zout, (uint32_t)nz * pb8);
    // input plane of step i (from f0) and its output plane (from zlo), in either direction
    const int pin0 = down ? nz + 7 : 0, pout0 = down ? nz + 7 : -8, pdir = down ? -1 : 1;
    auto pin = [&](int i) { return (uint32_t)(pin0 + pdir * i); };
    auto pout = [&](int i) { return pout0 + pdir * i; };
    const float dzs = down ? -a.dzinv : a.dzinv;
    const float zsc = (ISO && down) ? -1.0f : 1.0f;  // sign of c[4], c[5] (decode_pair)
    // records of the next output plane, and the step to the one after
    const char *cpl = (const char *)a.coef + (down ? zout + (long long)(nz - 1) * pxy : zout) * REC;
    const long long cst = down ? -(long long)pbr : (long long)pbr;

    int src_z = -1;
    uint32_t src_col = 0;
    if (ctl.do_src && a.src_idx >= 0) {
        src_z = (int)(a.src_idx / (long long)pxy);
        src_col = (uint32_t)(a.src_idx - (long long)src_z * pxy);
    }
    const int rec_z = ctl.trace ? a.rec_iz : -1;

    // staging: pair e of a staged plane = row*HX + column, the same index in LDS; wave w
    // loads e = w*64 + k*NT + lane
    uint32_t hoff[NLD];
    bool hk[NLD];
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int ew = wv * 64 + k * NT;
        hk[k] = ew < NSP;
        const int e = ew + lane;
        const int r = e / HX, w = e - r * HX;
        const int gy = y0 - 4 + r, gx = x0 - 4 + w;
        const bool ok = e < NSP && gx >= 0 && gx < a.sx && gy < a.sy;
        hoff[k] = ok ? (uint32_t)(gy * a.px + gx) * 8u : FL_OOB;
    }
    F2 pre[NLD];
    auto fetch = [&](int i) {  // the input plane of step i
        const uint32_t so = pin(i) * pb8;
#pragma unroll
        for (int k = 0; k < NLD; k++)
            if (hk[k]) pre[k] = f2_of(__builtin_amdgcn_raw_buffer_load_b64(r_cur, hoff[k], so, 0));
    };
    auto stage = [&](F2 *T) {
#pragma unroll
        for (int k = 0; k < NLD; k++)
            if (hk[k]) {
                const int e = wv * 64 + k * NT + lane;
                if (NSP % 64 == 0 || e < NSP) T[e] = pre[k];
            }
    };

    // window corners in the LDS planes (every stencil read an immediate offset from them)
    const int ow = (2 * wv + rs) * HX + cx;   // staged plane, 9 x 9 window of the own cell
    const int oxw = (2 * wv + rs) * 32 + cx;  // Dx plane, 9-row column of the own cell
    constexpr int CW = 4 * HX + 4, CX = 4 * 32;
    int ohr[NH], ohx[NH];  // y-halo rows of Dx done by this wave, relative to ow / oxw
    bool hrow[NH];
#pragma unroll
    for (int m = 0; m < NH; m++) {
        const int j = wv + m * W;  // halo row pair 0..3: rows 0-1, 2-3, TY+4..5, TY+6..7
        hrow[m] = j < 4;
        const int r = j < 2 ? 2 * j : TY + 2 * j;
        ohr[m] = (r - 2 * wv) * HX;
        ohx[m] = (r - 2 * wv) * 32;
    }

    F2 qc_[10], qy_[10], qa_[5], qb_[5];  // c, Dy; Dxx, Dyy (Dx: the LDS ring)
#pragma unroll
    for (int k = 0; k < 10; k++) qc_[k] = qy_[k] = F2{0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 5; k++) qa_[k] = qb_[k] = F2{0.f, 0.f};
    CoefRec<CM> kp[2];
    F2 ppv[2] = {};

    fetch(0);
    stage(s.t[0]);
    fetch(1);
    __syncthreads();

    int r10 = 0;  // Dx slot of plane i (i % 10) in the generic steps
    // one input plane (step i); template arguments as in step_tile:
    //   PH  queue slot of plane i (static steps: i % 10)
    //   XW  the y-halo rows of Dx of plane i: 0 never / 1 always / 2 if i < nz+4 (f is an
    //       output plane); Dxx, Dyy of plane i whenever XW != 0.  The tile rows of Dx go to
    //       the ring on every plane (Dxz reads them)
    //   OUT output plane z = f-4, GEN generic copy (queues shifted afterwards), PF prefetch
    //   bits, PAR i % 2 (staged slot), PS prefetch set
    auto step = [&](auto PHc, auto XWc, auto OUTc, auto GENc, auto PFc, auto PARc, auto PSc, const int i,
                    const bool pf_next) {
        constexpr int PH = decltype(PHc)::value;
        constexpr int XW = decltype(XWc)::value;
        constexpr bool OUT = decltype(OUTc)::value != 0, GEN = decltype(GENc)::value != 0;
        constexpr int PF = decltype(PFc)::value;
        constexpr int PAR = decltype(PARc)::value;
        constexpr int PS = decltype(PSc)::value;
        const int z = zlo + pout(i);  // output plane
        const F2 *const Tf = s.t[PAR];
        F2 *const Xf = s.x[GEN ? r10 : PH];
        const F2 *const Xz = s.x[GEN ? (r10 >= 4 ? r10 - 4 : r10 + 6) : (PH + 6) % 10];  // plane i-4
        const uint32_t so_o = (uint32_t)pout(i) * pb8;

        CoefRec<CM> kr;
        F2 vprev = {0.f, 0.f};
        if constexpr (OUT && (PF & 1)) {
            kr = kp[PS];
            vprev = ppv[PS];
        } else if constexpr (OUT) {
            load_rec<CM, PCPOL>(span_rsrc(cpl, pbr), vor, kr);
            cpl += cst;
            vprev = f2_of(__builtin_amdgcn_raw_buffer_load_b64(r_prev, vo8, so_o, PCPOL));
        }
        const bool xw = (XW == 1) || (XW == 2 && i < nz + 4);
        const F2 *cp = Tf + ow + CW;
        F2 nc = cp[0], nx, ny, nxx = {0.f, 0.f}, nyy = {0.f, 0.f};
        // one axis at a time: the scheduler would otherwise start every LDS read of the
        // step at once, 60-odd registers of pairs in flight
        der12_pr<!ISO, XW != 0>(cp, 1, nc, a.dxinv, a.dxxinv, nx, nxx);
        __builtin_amdgcn_sched_barrier(0);
        der12_pr<!ISO, XW != 0>(cp, HX, nc, a.dyinv, a.dyyinv, ny, nyy);
        __builtin_amdgcn_sched_barrier(0);
        Xf[oxw + CX] = nx;  // every plane: Dxz reads it
        if (xw) {
#pragma unroll
            for (int m = 0; m < NH; m++)
                if (hrow[m]) Xf[oxw + ohx[m]] = der1_pr<!ISO>(Tf + ow + ohr[m] + 4, 1, a.dxinv);
        }
        fl_pin(nc);
        fl_pin(nx);
        fl_pin(ny);
        constexpr int ins = GEN ? 8 : PH, ins5 = GEN ? 4 : PH % 5;
        qc_[ins] = nc;
        qy_[ins] = ny;
        if constexpr (XW != 0) {
            fl_pin(nxx);
            fl_pin(nyy);
            qa_[ins5] = nxx;
            qb_[ins5] = nyy;
        }
        __builtin_amdgcn_sched_barrier(0);
        F2 dxy = {0.f, 0.f};
        if constexpr (OUT) dxy = der1_pr<!ISO>(Xz + oxw + CX, 32, a.dyinv);
        __builtin_amdgcn_sched_barrier(0);
        // next plane into the other staged slot, then the plane after it into registers
        if (i + 1 <= nz + 7) stage(s.t[1 - PAR]);
        if (i + 2 <= nz + 7) fetch(i + 2);

        if constexpr (OUT) {
            constexpr int b = GEN ? 0 : (PH + 2) % 10;  // slot of plane i-8
            constexpr int b5 = GEN ? 0 : (PH + 1) % 5;  // slot of plane i-4
            float zp[9], zq[9], xp[9], xq[9], yp[9], yq[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                zp[k] = qc_[(b + k) % 10].p; zq[k] = qc_[(b + k) % 10].q;
                yp[k] = qy_[(b + k) % 10].p; yq[k] = qy_[(b + k) % 10].q;
                // Dx of plane i-8+k from its ring slot (the centre plane is not used), plane
                // i from the register
                F2 v = nx;
                if (k < 8 && k != 4) v = s.x[GEN ? (r10 + 2 + k) % 10 : (PH + 2 + k) % 10][oxw + CX];
                xp[k] = v.p; xq[k] = v.q;
            }
            const F2 dxx = qa_[b5], dyy = qb_[b5];
            const float dzzp = der2_q<!ISO>(zp, a.dzzinv), dzzq = der2_q<!ISO>(zq, a.dzzinv);
            // the z first derivatives of a downward march come out negated: dzs < 0 restores
            // their sign, or (ISO, unscaled sums) the negated c[4], c[5]
            const float dxzp = der1_q<!ISO>(xp, dzs), dxzq = der1_q<!ISO>(xq, dzs);
            const float dyzp = der1_q<!ISO>(yp, dzs), dyzq = der1_q<!ISO>(yq, dzs);
            CoefPair kc;
            decode_pair<CM>(kr, kc, zsc);
            // H1, H2 of both fields, association order of original/sample.h:46-53
            float h1p = kc.c[0] * dxx.p;
            h1p = __builtin_fmaf(kc.c[1], dyy.p, h1p);
            h1p = __builtin_fmaf(kc.c[2], dzzp, h1p);
            h1p = __builtin_fmaf(kc.c[3], dxy.p, h1p);
            h1p = __builtin_fmaf(kc.c[5], dxzp, h1p);
            h1p = __builtin_fmaf(kc.c[4], dyzp, h1p);
            const float h2p = dxx.p + dyy.p + dzzp - h1p;
            float h1q = kc.c[0] * dxx.q;
            h1q = __builtin_fmaf(kc.c[1], dyy.q, h1q);
            h1q = __builtin_fmaf(kc.c[2], dzzq, h1q);
            h1q = __builtin_fmaf(kc.c[3], dxy.q, h1q);
            h1q = __builtin_fmaf(kc.c[5], dxzq, h1q);
            h1q = __builtin_fmaf(kc.c[4], dyzq, h1q);
            const float h2q = dxx.q + dyy.q + dzzq - h1q;
            // rhsp = v2px h2p + v2pz h1q + v2sz (h1p - h1q)
            // rhsq = v2pn h2p + v2pz h1q + v2sz (h2q - h2p)   (original/sample.h:75-81)
            float rp = kc.own_p * h2p;
            rp = __builtin_fmaf(kc.vpz, h1q, rp);
            rp = __builtin_fmaf(kc.vsz, h1p - h1q, rp);
            float rqq = kc.own_q * h2p;
            rqq = __builtin_fmaf(kc.vpz, h1q, rqq);
            rqq = __builtin_fmaf(kc.vsz, h2q - h2p, rqq);
            const F2 cur = qc_[(b + 4) % 10];
            const float bp = 2.0f * cur.p - vprev.p, bq = 2.0f * cur.q - vprev.q;
            float np, nq;
            if constexpr (CM == CM_Q16R) {
                const float u = rq * kc.e2;
                np = __builtin_fmaf(rp, u, bp);
                nq = __builtin_fmaf(rqq, u, bq);
            } else if constexpr (ISO) {
                np = __builtin_fmaf(rp, a.dxxinv, bp);
                nq = __builtin_fmaf(rqq, a.dxxinv, bq);
            } else {
                np = bp + rp;
                nq = bq + rqq;
            }
            if (z == rec_z) {
                if (active) ctl.trace[(y - a.bord) * (a.sx - 2 * a.bord) + (x - a.bord)] = np;
            }
            if (__builtin_expect(z == src_z, 0)) {
                if (col == src_col && active) {
                    np += ctl.src_add;
                    nq += ctl.src_add;
                }
            }
            const u32x2 v = {__builtin_bit_cast(uint32_t, np), __builtin_bit_cast(uint32_t, nq)};
            __builtin_amdgcn_raw_buffer_store_b64(v, r_prev, vo8, so_o, PCPOL);
        }
        if constexpr ((PF & 2) != 0 || (PF & 4) != 0) {
            if ((PF & 2) || pf_next) {
                constexpr int NS2 = (PS + PD) % 2;
                load_rec<CM, PCPOL>(span_rsrc(cpl, pbr), vor, kp[NS2]);
                cpl += cst;
                ppv[NS2] = f2_of(
                    __builtin_amdgcn_raw_buffer_load_b64(r_prev, vo8, (uint32_t)pout(i + PD) * pb8, PCPOL));
            }
        }
        if constexpr (GEN) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                qc_[k] = qc_[k + 1];
                qy_[k] = qy_[k + 1];
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                qa_[k] = qa_[k + 1];
                qb_[k] = qb_[k + 1];
            }
            r10 = r10 == 9 ? 0 : r10 + 1;
        }
        __syncthreads();
    };

    constexpr int PFB = 4;
    const bool grp0 = 8 + 9 < nz + 4;
    auto pro = [&](auto Jc, auto XWc) {
        constexpr int j = decltype(Jc)::value;
        step(ic<j>{}, XWc, ic<0>{}, ic<0>{}, ic<(j + PD >= 8 ? PFB : 0)>{}, ic<j & 1>{}, ic<(j + 18) % 2>{}, j,
             j + PD >= 8 ? grp0 : false);
    };
    pro(ic<0>{}, ic<0>{});
    pro(ic<1>{}, ic<0>{});
    pro(ic<2>{}, ic<0>{});
    pro(ic<3>{}, ic<0>{});
    pro(ic<4>{}, ic<2>{});
    pro(ic<5>{}, ic<2>{});
    pro(ic<6>{}, ic<2>{});
    pro(ic<7>{}, ic<2>{});
    int i = 8;
    for (; i + 9 < nz + 4; i += 10) {
        const bool more = i + 10 + 9 < nz + 4;
        unroll_phases(
            [&](auto P) {
                constexpr int p = decltype(P)::value;
                constexpr int pf = 1 | (p + PD <= 9 ? 2 : 4);
                step(ic<(8 + p) % 10>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<pf>{}, ic<p & 1>{}, ic<p % 2>{}, i + p, more);
            },
            std::make_integer_sequence<int, 10>{});
    }
    // the generic steps: planes i-4 .. i-1 of the Dxx, Dyy queues (slots (i-4) % 5 ..,
    // i = 8 mod 10 here: 4, 0, 1, 2) into slots 0-3, and the Dx slot of plane i
    {
        const F2 a4 = qa_[4], b4 = qb_[4];
        qa_[4] = qa_[3]; qb_[4] = qb_[3];  // slot 3 holds plane i-5: unused
        qa_[3] = qa_[2]; qb_[3] = qb_[2];
        qa_[2] = qa_[1]; qb_[2] = qb_[1];
        qa_[1] = qa_[0]; qb_[1] = qb_[0];
        qa_[0] = a4; qb_[0] = b4;
        r10 = 8;  // (8 + 10 g) % 10
    }
    for (; i + 1 <= nz + 7; i += 2) {
        step(ic<0>{}, ic<2>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<0>{}, ic<0>{}, i, false);
        step(ic<0>{}, ic<2>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<1>{}, ic<0>{}, i + 1, false);
    }
    if (i <= nz + 7) step(ic<0>{}, ic<2>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<0>{}, ic<0>{}, i, false);
}

// Waves per SIMD asked of the register allocator: 3 (168 VGPRs; the steady loop uses 150
// with no spills).  The 8-row workgroups (4 waves, 51 KB of LDS) then run three to a CU,
// one wave of each per SIMD.
template <int TY>
constexpr int pair_waves()
{
#ifdef FL_PAIR_WAVES
    return FL_PAIR_WAVES;
#else
    return 3;
#endif
}

template <int TY, int CM, bool ISO, bool HINTS>
__global__ __launch_bounds__(32 * TY) __attribute__((amdgpu_waves_per_eu(pair_waves<TY>())))
void tti_step2(const StepArgs a)
{
    __shared__ StepLds2<TY> s;
    const StepCtl ctl = step_ctl(a);
    if (a.run && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0)
        a.run[1 - a.par] = a.run[a.par] + 1;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    if (lb >= a.nblk) return;
    const int t = step_order(lb, a.nblk, ctl.it);
    int tx, ty, tz;
    tile_of(t, a.ntx, a.nty, a.sw, tx, ty, tz);
    step_tile2<TY, CM, ISO, HINTS ? 2 : 0>(a, ctl, uni(tx), uni(ty), uni(tz), step_down(ctl.it), s);
}

}  // namespace fl
