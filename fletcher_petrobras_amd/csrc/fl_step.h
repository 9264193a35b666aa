// The TTI step kernel for gfx950: split-field z-marching with LDS plane rings.
//
// Work decomposition.  A workgroup of TY waves owns a 32 x TY tile of the xy plane and a
// chunk [zlo, zhi) of z; it streams the input planes f = zlo-4 .. zhi+3 of pc, qc through
// LDS once (on odd steps from zhi+3 down to zlo-4: the planes below are then those above,
// and "f-8 .. f" below reads "f+8 .. f").  A wave covers 32 columns of one tile row for
// BOTH fields: lane 2c carries p and lane 2c+1 q of column c, so each lane keeps only its
// own field's state.
//
// On chip, per lane (own field, own column):
//   registers  9-deep z queues of c = pc|qc, Dx and Dy over planes f-8 .. f (Dzz, Dxz, Dyz
//              of the output plane z = f-4: der1der1lm's 9-plane ring,
//              der1der1lm/zTileSize.h:4-7 and sample.h:147-160, held in registers).  The
//              ring (10 slots) is addressed statically: the steady loop is unrolled 10
//              times, one copy per ring phase, so planes advance without register moves.
//   LDS        a ring of 6 staged planes of pc, qc (tile + 4-cell halo): plane f for Dx, Dy
//              and the queues, plane z = f-4 for Dxx, Dyy, plane f+1 being written;
//              a ring of 5 planes of Dx over the tile rows plus the 4-row y halo, from which
//              Dxy(z) = Der1_y(Dx) is read (der1der1/sample.h:57-62, 83-88).
// With the rings every plane needs a single barrier: nothing written in a step is read
// in it, and nothing read in it is overwritten before the next barrier.
//
// Memory.  The levels hold (p, q) pairs per cell (pitched rows), so a wave's 32 cells x 2
// fields are 64 consecutive floats; the coefficients are one record per cell, whose two
// halves the p and q lanes of the cell load and exchange (DPP broadcast).  Every
// stream goes through a buffer resource over the chunk's span with a
// wave-uniform plane offset (soffset): one SGPR add per plane, no 64-bit address math.
// Lanes outside the grid carry an offset past the range, so their loads return 0 and
// their stores are dropped by the hardware (no masks, no clamps).
//
// The combination and leapfrog follow original/sample.h:46-86 in its association order;
// p and q exchange h1q / h2p with one swap between neighbouring lanes (DPP).
//
// TWOPASS takes Dx, Dy from the four first-derivative volumes of der1_pass
// (der1der1/OpenMP/openmp_propagate.c:24-76) instead of forming them on chip.
//
// fl_step2.h holds the same step with both fields of a cell in one lane (the 16-bit
// recompute forms on 8-row tiles), fl_step_batch.h the batched step over a (shot, tile)
// work list built on either.
#pragma once
#include <type_traits>
#include <utility>

#include "fl_kernels.h"

namespace fl {

// Global load through a wave-uniform base pointer plus a 32-bit lane byte offset.
__device__ __forceinline__ float ld_f32(const void *base, uint32_t off)
{
    return *(const float *)((const char *)base + off);
}

// Prefetch distances in steps: the output plane's coefficient record and previous-level
// value (prefetch_steps, 0 = loaded in their own step) and the staged input plane
// (FL_STAGE_AHEAD, 1 or 2).  Both use two register sets chosen by step parity.
// Measured at 344^3: two steps ahead gain 3% (FP32) and 7% (FP16 recompute); the stored
// 16-bit forms, at 80 VGPRs, lose 6% and keep one; staging two planes ahead gains nothing.
#ifndef FL_STAGE_AHEAD
#define FL_STAGE_AHEAD 1
#endif
template <int CM>
constexpr int prefetch_steps()
{
#ifdef FL_PREFETCH
    return FL_PREFETCH;
#else
    return (CM == CM_BF16 || CM == CM_F16) ? 1 : 2;
#endif
}

// Streaming hints (PCPOL 2, off 0): the previous level, read and written once per step, and
// the coefficient records whose loads read whole lines across a wave (the recompute forms,
// and FP32 in its two arrays) are loaded and stored non-temporal.  Records whose lines two
// loads share (BF16, FP16) keep the default policy: the hint on either load made those
// steps slower.  Whether the hints pay depends
// on the grid (faster on a 384^3 array, slower on a 1096x1096x840 one), so the host
// selects them per grid.
template <int CM, int PCPOL>
constexpr int coef_cpol()
{
    return (PCPOL && (CM == CM_F32 || CM == CM_F16R || CM == CM_F32R || CM == CM_Q16R)) ? 2 : 0;
}

// Lane byte offset beyond every buffer range used here (spans are < 2^31 bytes).
constexpr uint32_t FL_OOB = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t span_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ uint32_t bload32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}

// Coefficient record as loaded by one lane: FP32 its half of the record (5 values, 20 B);
// BF16 / FP16 its half (2 words) and the shared word of v2px, v2pn; FP32 recompute its
// half (own velocity factor, v2pz or v2sz, own angle: 12 B); FP16 recompute the whole
// record (4 FP16 velocity factors, 2 u16 angles).  Kept raw until used, so the
// loads stay in flight across barriers.
template <int CM>
struct CoefRec {
    static constexpr int NW = (CM == CM_F32) ? 5 : 3;
    uint32_t w[NW];
};

// Lane byte offsets of a cell's record: the own half (voff) and, for the 16-bit forms, the
// shared last word (voff2); FP32: the own 16 bytes in the first array and the own 4 bytes
// in the second (rec_order32).
struct RecOff {
    uint32_t voff, voff2;
};
template <int CM>
__device__ __forceinline__ RecOff rec_offsets(uint32_t col, int fld, bool active)
{
    RecOff o;
    if constexpr (CM == CM_F32) {
        o.voff = active ? col * 32u + (uint32_t)fld * 16u : FL_OOB;
        o.voff2 = active ? col * 8u + (uint32_t)fld * 4u : FL_OOB;
    } else if constexpr (CM == CM_F16R || CM == CM_Q16R) {
        o.voff = active ? col * 12u : FL_OOB;
        o.voff2 = 0;
    } else if constexpr (CM == CM_F32R) {
        o.voff = active ? col * 24u + (uint32_t)fld * 12u : FL_OOB;
        o.voff2 = 0;
    } else {
        o.voff = active ? col * 20u + (uint32_t)fld * 8u : FL_OOB;
        o.voff2 = active ? col * 20u + 16u : FL_OOB;
    }
    return o;
}

// Loads of one record: the recompute forms whole (one 12- or 24-byte load), BF16 and FP16
// the lane's half and the shared word.
template <int CM, int PCPOL>
__device__ __forceinline__ void load_rec(__amdgpu_buffer_rsrc_t r, const RecOff &o, CoefRec<CM> &c)
{
    static_assert(CM != CM_F32, "FP32 records: load_next_rec");
    if constexpr (CM == CM_F16R || CM == CM_F32R || CM == CM_Q16R) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b96(r, o.voff, 0, coef_cpol<CM, PCPOL>());
        c.w[0] = a[0]; c.w[1] = a[1]; c.w[2] = a[2];
    } else {
        const auto a = __builtin_amdgcn_raw_buffer_load_b64(r, o.voff, 0, 0);
        c.w[0] = a[0]; c.w[1] = a[1];
        c.w[2] = __builtin_amdgcn_raw_buffer_load_b32(r, o.voff2, 0, 0);
    }
}

// The records of the output planes of a chunk, one plane after the other, upwards or downwards
// in z: p the records (FP32: its first array), p2 the second FP32 array, d1 and d2 their
// signed steps from one plane to the next.
struct RecStream {
    const char *p, *p2;
    long long d1, d2;
};
template <int CM>
__device__ __forceinline__ RecStream rec_stream(const void *coef, long long np, long long cell0, uint32_t pxy,
                                                bool down)
{
    RecStream r;
    const long long sg = down ? -1 : 1;
    if constexpr (CM == CM_F32) {
        r.p = (const char *)coef + cell0 * 32;
        r.p2 = (const char *)coef + fp32_part2(np) + cell0 * 8;
        r.d1 = sg * (long long)pxy * 32;
        r.d2 = sg * (long long)pxy * 8;
    } else {
        r.p = (const char *)coef + cell0 * rec_bytes(CM);
        r.p2 = nullptr;
        r.d1 = sg * (long long)pxy * rec_bytes(CM);
        r.d2 = 0;
    }
    return r;
}

// Loads the next plane's record of the lane's cell and advances the stream (pxy cells a plane).
template <int CM, int PCPOL>
__device__ __forceinline__ void load_next_rec(RecStream &rs, uint32_t pxy, const RecOff &o, CoefRec<CM> &c)
{
    if constexpr (CM == CM_F32) {
        constexpr int cp = coef_cpol<CM, PCPOL>();
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(span_rsrc(rs.p, pxy * 32u), o.voff, 0, cp);
        c.w[0] = a[0]; c.w[1] = a[1]; c.w[2] = a[2]; c.w[3] = a[3];
        c.w[4] = __builtin_amdgcn_raw_buffer_load_b32(span_rsrc(rs.p2, pxy * 8u), o.voff2, 0, cp);
        rs.p += rs.d1;
        rs.p2 += rs.d2;
    } else {
        const uint32_t pbr = pxy * (uint32_t)rec_bytes(CM);
        load_rec<CM, PCPOL>(span_rsrc(rs.p, pbr), o, c);
        rs.p += rs.d1;
    }
}

// Value of the p lane (even) or of the q lane (odd) of each lane pair, in both lanes
// (DPP quad_perm [0,0,2,2] and [1,1,3,3]).
__device__ __forceinline__ uint32_t from_p(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xA0, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t from_q(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xF5, 0xF, 0xF, true);
}
__device__ __forceinline__ float from_pf(float v)
{
    return __builtin_bit_cast(float, from_p(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ float from_qf(float v)
{
    return __builtin_bit_cast(float, from_q(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

// The coefficients of the own cell in precomp.h order.  For the stored forms c[6] and c[9]
// both hold the lane's own factor (v2px on p lanes, v2pn on q lanes), the only one the
// update of the lane's field reads.
template <int CM>
__device__ __forceinline__ void decode_rec(const CoefRec<CM> &r, int fld, Coef &k)
{
    if constexpr (CM == CM_F32) {
        k.c[0] = __builtin_bit_cast(float, from_p(r.w[0]));
        k.c[1] = __builtin_bit_cast(float, from_p(r.w[1]));
        k.c[2] = __builtin_bit_cast(float, from_p(r.w[2]));
        k.c[3] = __builtin_bit_cast(float, from_p(r.w[3]));
        k.c[4] = __builtin_bit_cast(float, from_q(r.w[0]));
        k.c[5] = __builtin_bit_cast(float, from_q(r.w[1]));
        k.c[7] = __builtin_bit_cast(float, from_q(r.w[2]));
        k.c[8] = __builtin_bit_cast(float, from_q(r.w[3]));
        k.c[6] = k.c[9] = __builtin_bit_cast(float, r.w[4]);
    } else if constexpr (CM == CM_BF16 || CM == CM_F16) {
        const uint32_t a0 = from_p(r.w[0]), a1 = from_p(r.w[1]);
        const uint32_t b0 = from_q(r.w[0]), b1 = from_q(r.w[1]);
        // the own half by a select on the lane parity (a mask in SGPRs), not a per-lane
        // shift count, which would hold a VGPR for the whole kernel
        const uint32_t own = fld ? (r.w[2] >> 16) : (r.w[2] & 0xffffu);
        if constexpr (CM == CM_BF16) {
            k.c[0] = bf16_lo(a0); k.c[1] = bf16_hi(a0);
            k.c[2] = bf16_lo(a1); k.c[3] = bf16_hi(a1);
            k.c[4] = bf16_lo(b0); k.c[5] = bf16_hi(b0);
            k.c[7] = bf16_lo(b1); k.c[8] = bf16_hi(b1);
            k.c[6] = k.c[9] = bf16_lo(own);
        } else {
            k.c[0] = f16_to_f32((uint16_t)a0); k.c[1] = f16_to_f32((uint16_t)(a0 >> 16));
            k.c[2] = f16_to_f32((uint16_t)a1); k.c[3] = f16_to_f32((uint16_t)(a1 >> 16));
            k.c[4] = f16_to_f32((uint16_t)b0); k.c[5] = f16_to_f32((uint16_t)(b0 >> 16));
            k.c[7] = f16_to_f32((uint16_t)b1); k.c[8] = f16_to_f32((uint16_t)(b1 >> 16));
            k.c[6] = k.c[9] = f16_to_f32((uint16_t)own);
        }
    } else {
        // recomputed forms: the lane's own velocity factor and angle (p lanes v2px and
        // theta, q lanes v2pn and phi), v2pz and v2sz from the pair
        float own, vpz, vsz, t;
        if constexpr (CM == CM_F16R) {  // words (v2px, v2pn), (v2pz, v2sz), (theta, phi)
            own = fld ? f16_to_f32((uint16_t)(r.w[0] >> 16)) : f16_to_f32((uint16_t)r.w[0]);
            vpz = f16_to_f32((uint16_t)r.w[1]);
            vsz = f16_to_f32((uint16_t)(r.w[1] >> 16));
            t = (float)(fld ? (r.w[2] >> 16) : (r.w[2] & 0xffffu)) * (1.0f / 65536.0f);
        } else if constexpr (CM == CM_Q16R) {
            // as CM_F16R, the factors as 15-bit mantissas of the cell's exponent e (bits
            // 30-31 of the two words) in units of 2^ebase: the mantissas here, 2^e in c[9],
            // both powers of two applied once in the leapfrog (the same products: scaling by
            // a power of two is exact)
            const uint32_t e = (r.w[0] >> 30) | ((r.w[1] >> 28) & 0xcu);
            k.c[9] = __builtin_bit_cast(float, (e + 127u) << 23);  // 2^e
            own = (float)((r.w[0] >> (fld ? 15 : 0)) & 0x7fffu);
            vpz = (float)(r.w[1] & 0x7fffu);
            vsz = (float)((r.w[1] >> 15) & 0x7fffu);
            t = (float)(fld ? (r.w[2] >> 16) : (r.w[2] & 0xffffu)) * (1.0f / 65536.0f);
        } else {  // CM_F32R: lane words (own factor, v2pz | v2sz, theta | phi turns)
            own = __builtin_bit_cast(float, r.w[0]);
            vpz = __builtin_bit_cast(float, from_p(r.w[1]));
            vsz = __builtin_bit_cast(float, from_q(r.w[1]));
            t = __builtin_bit_cast(float, r.w[2]);
        }
        // the angular coefficients of precomp.h:25-38; v_sin / v_cos take turns.  Each
        // lane evaluates the sine and cosine of its own angle (theta on p lanes, phi on q
        // lane, and the products of one angle too (sin^2, cos^2, sin 2a); each coefficient
        // is then one product of a p-lane and a q-lane value
        const float sn = __builtin_amdgcn_sinf(t), cs = __builtin_amdgcn_cosf(t);
        const float s2 = sn * sn, c2 = cs * cs, sc2 = (sn + sn) * cs;
        const float st2 = from_pf(s2), s2t = from_pf(sc2);  // sin^2 theta, sin 2theta
        k.c[0] = st2 * from_qf(c2);   // sin^2 theta cos^2 phi
        k.c[1] = st2 * from_qf(s2);   // sin^2 theta sin^2 phi
        k.c[2] = from_pf(c2);         // cos^2 theta
        k.c[3] = st2 * from_qf(sc2);  // sin^2 theta sin 2phi
        k.c[4] = s2t * from_qf(sn);   // sin 2theta sin phi
        k.c[5] = s2t * from_qf(cs);   // sin 2theta cos phi
        k.c[6] = own;
        if constexpr (CM != CM_Q16R) k.c[9] = own;
        k.c[7] = vpz;
        k.c[8] = vsz;
    }
}

template <int TY>
struct StepLds {
    static constexpr int HX = 40, HY = TY + 8;
    float t[6][HY * HX * 2];  // staged levels: (row, column, field) of the tile + halo
    float x[5][HY * 32 * 2];  // Dx: (row, tile column, field), tile rows + y halo
};

template <int V>
using ic = std::integral_constant<int, V>;

// An empty register constraint: the value is materialised in a VGPR at this point.
__device__ __forceinline__ void fl_pin(float &x) { asm volatile("" : "+v"(x)); }

template <class F, int... P>
__device__ __forceinline__ void unroll_phases(F &&fn, std::integer_sequence<int, P...>)
{
    (fn(ic<P>{}), ...);
}

// One tile of one z chunk.  Output planes z in [zlo, zhi) of the previous level are
// overwritten once.
// ISO (dx = dy = dz = h): the derivatives stay unscaled sums and 1/h^2 multiplies the
// right-hand side once, in the leapfrog (the same expression up to rounding: every term
// of H1, H2 carries 1/h^2).
// down: the chunk is marched from its top plane to its bottom one.  Step i then takes input
// plane zhi+3-i instead of zlo-4+i, the z queues hold the planes in reverse order, and the
// z second derivative, a sum of symmetric pairs, comes out bit for bit the same; each z
// first derivative comes out exactly negated (every operation of the sum is odd in its
// inputs) and its sign is restored by a negative dz scale (or a multiply by -1 when ISO).
template <int TY, int CM, bool TWOPASS, bool ISO, int PCPOL>
__device__ __forceinline__ void step_tile(const StepArgs &a, const StepCtl &ctl, int tix, int tiy, int zc,
                                          bool down, StepLds<TY> &s)
{
    constexpr int TX = 32, HX = 40, HY = TY + 8, NT = 64 * TY;
    constexpr int RW = 2 * HX;                 // floats per staged row (pairs)
    constexpr int NS = RW * HY;                // floats per staged plane (a multiple of 64)
    constexpr int NLD = (NS + NT - 1) / NT;    // staging loads per lane
    constexpr int XW2 = 2 * TX;                // floats per Dx row
    constexpr int NH = (8 + TY - 1) / TY;      // y-halo rows of Dx per wave (at most)
    static_assert(NS % 64 == 0, "staged plane must be whole waves");
    const int x0 = a.xorg + tix * TX, y0 = a.bord + tiy * TY;
    const int zlo = a.bord + zc * a.zt;
    const int zhi = min(zlo + a.zt, a.sz - a.bord);
    const int nz = zhi - zlo, f0 = zlo - 4;
    // lane id (= threadIdx.x, blocks are 64 wide) from mbcnt: a value the register
    // allocator can recompute instead of keeping the work-item id alive or spilling it
    const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.y);  // tile row (wave-uniform)
    // lane = 2 * column + field: a half wave reads 32 consecutive floats (16 cells x 2
    // fields) from global memory and LDS alike, free of LDS bank conflicts
    const int fld = lane & 1, cx = lane >> 1;
    const int x = x0 + cx, y = y0 + wv;
    const bool active = (x >= a.bord) && (x < a.sx - a.bord) && (y < a.sy - a.bord);
    static_assert(!(ISO && TWOPASS), "two-pass derivative volumes are scaled");
    const float dts = ISO ? a.dxxinv * a.dt : a.dt;  // scale of rhs in the FP32 leapfrog
    const float rq = ISO ? a.dxxinv * a.vq : a.vq;   // and in the block floating-point form
    const uint32_t pxy = (uint32_t)a.px * (uint32_t)a.sy;  // cells per pitched plane
    const uint32_t pb8 = pxy * 8u;
    const uint32_t col = (uint32_t)(y * a.px + x);
    const uint32_t vo8 = active ? col * 8u + fld * 4u : FL_OOB;  // own (cell, field) float
    const RecOff vor = rec_offsets<CM>(col, fld, active);

    const long long zin = (long long)f0 * pxy, zout = (long long)zlo * pxy;
    const __amdgpu_buffer_rsrc_t r_cur = span_rsrc(a.cur + 2 * zin, (uint32_t)(nz + 8) * pb8);
    const __amdgpu_buffer_rsrc_t r_prev = span_rsrc(a.prev + 2 * zout, (uint32_t)nz * pb8);
    // input plane of step i (from f0) and its output plane (from zlo), in either direction
    const int pin0 = down ? nz + 7 : 0, pout0 = down ? nz + 7 : -8, pdir = down ? -1 : 1;
    auto pin = [&](int i) { return (uint32_t)(pin0 + pdir * i); };
    auto pout = [&](int i) { return pout0 + pdir * i; };
    const float dzs = down ? -a.dzinv : a.dzinv, zsg = down ? -1.0f : 1.0f;
    // next output plane's records
    RecStream recs = rec_stream<CM>(a.coef, (long long)pxy * a.sz, down ? zout + (long long)(nz - 1) * pxy : zout,
                                    pxy, down);

    // source cell of this step (added to both fields of the output) and receiver plane
    int src_z = -1;
    uint32_t src_col = 0;
    if (ctl.do_src && a.src_idx >= 0) {
        src_z = (int)(a.src_idx / (long long)pxy);
        src_col = (uint32_t)(a.src_idx - (long long)src_z * pxy);
    }
    const int rec_z = ctl.trace ? a.rec_iz : -1;

    // staging: float e of a staged plane = row*RW + 2*column + field, the same index in
    // LDS; wave w loads e = w*64 + k*NT + lane (whole waves, contiguous per wave)
    uint32_t hoff[NLD];
    bool hk[NLD];
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int ew = wv * 64 + k * NT;
        hk[k] = ew < NS;
        const int e = ew + lane;
        const int r = e / RW, w = e - r * RW;
        const int gy = y0 - 4 + r, gx = x0 - 4 + (w >> 1);
        const bool ok = hk[k] && gx >= 0 && gx < a.sx && gy < a.sy;
        hoff[k] = ok ? ((uint32_t)(gy * a.px + gx) * 2u + (uint32_t)(w & 1)) * 4u : FL_OOB;
    }
    constexpr int SA = FL_STAGE_AHEAD;
    float pre[SA][NLD];  // raw loads of plane j in set j % SA
    auto fetch = [&](auto Sc, int i) {  // loads of the input plane of step i into set Sc
        constexpr int S = decltype(Sc)::value;
        const uint32_t so = pin(i) * pb8;
#pragma unroll
        for (int k = 0; k < NLD; k++)
            if (hk[k]) pre[S][k] = __builtin_bit_cast(float, bload32(r_cur, hoff[k], so));
    };
    auto stage = [&](auto Sc, float *T) {
        constexpr int S = decltype(Sc)::value;
#pragma unroll
        for (int k = 0; k < NLD; k++)
            if (hk[k]) T[wv * 64 + k * NT + lane] = pre[S][k];
    };

    // lane offsets in the LDS planes (window corners: every stencil read is a
    // non-negative immediate offset)
    const int ow = wv * RW + lane;   // 9 x 9 window of the own (cell, field)
    const int oxw = wv * XW2 + lane; // 9-row column of the own (cell, field) in Dx
    constexpr int CW = 4 * RW + 8, CX = 4 * XW2;  // the own cell within them
    int ohr[NH], ohx[NH];  // y-halo rows of Dx done by this wave, relative to ow / oxw
                           // (uniform: no lane offsets kept besides ow, oxw)
    bool hrow[NH];
#pragma unroll
    for (int m = 0; m < NH; m++) {
        const int j = wv + m * TY;  // halo row index 0..7 (rows 0-3 and TY+4..TY+7)
        hrow[m] = j < 8;
        const int r = j < 4 ? j : TY + j;
        ohr[m] = (r - wv) * RW;
        ohx[m] = (r - wv) * XW2;
    }

    // TWOPASS: first-derivative pair volumes of pass 1, over the chunk's input planes
    const __amdgpu_buffer_rsrc_t r_dx = span_rsrc(TWOPASS ? a.dvx + 2 * zin : a.cur, (uint32_t)(nz + 8) * pb8);
    const __amdgpu_buffer_rsrc_t r_dy = span_rsrc(TWOPASS ? a.dvy + 2 * zin : a.cur, (uint32_t)(nz + 8) * pb8);

    // own field c, Dx, Dy of planes f-8..f: ring of 10 slots (plane j in slot j % 10 in
    // the statically addressed steps; 10 keeps the step parity of every slot fixed)
    float qc_[10], qx_[10], qy_[10];
#pragma unroll
    for (int k = 0; k < 10; k++) qc_[k] = qx_[k] = qy_[k] = 0.0f;
    // output-plane streams loaded prefetch_steps() ahead, in NSETS register sets (the set of
    // step i is (i - 8) % NSETS; NSETS divides the 10 steps of an unrolled group)
    constexpr int PDK = prefetch_steps<CM>() > 0 ? prefetch_steps<CM>() : 1;
    constexpr int NSETS = PDK <= 2 ? 2 : 5;
    static_assert(PDK < NSETS || PDK == 2, "prefetch distance beyond the register sets");
    CoefRec<CM> kp[NSETS];
    float ppv[NSETS] = {};

    // prologue: the first input plane into ring slot 0, the next SA planes in flight
    fetch(ic<0>{}, 0);
    stage(ic<0>{}, s.t[0]);
    fetch(ic<1 % SA>{}, 1);
    if constexpr (SA == 2) fetch(ic<0>{}, 2);
    __syncthreads();

    int r6 = 0, r5 = 0;  // LDS ring slots of the input plane of step i: i % 6, i % 5
    // one input plane (step i; plane f0 + i upwards, zhi + 3 - i downwards):
    //   PH   queue slot of plane i (static steps: i % 10)
    //   XW   Dx halo planes 0 never / 1 always / 2 if i < nz+4 (f is an output plane)
    //   OUT  output plane z = f-4
    //   GEN  generic copy: plane i-8+k in slot k, queues shifted afterwards
    //   PF   bit 0: this step's output streams were loaded PD steps earlier into set PS;
    //        bits 1-2: load those of step i + PD into set (PS + PD) % NSETS (2 always, 4 if
    //        `pf_next`)
    //   PAR  i % 2 (the staging register sets of the step)
    //   PS   (i - 8) % NSETS (the prefetch register set of the step)
    auto step = [&](auto PHc, auto XWc, auto OUTc, auto GENc, auto PFc, auto PARc, auto PSc, const int i,
                    const bool pf_next) {
        constexpr int PH = decltype(PHc)::value;
        constexpr int XW = decltype(XWc)::value;
        constexpr bool OUT = decltype(OUTc)::value != 0, GEN = decltype(GENc)::value != 0;
        constexpr int PF = decltype(PFc)::value;
        constexpr int SET = decltype(PARc)::value;
        constexpr int PS = decltype(PSc)::value;
        constexpr int PD = prefetch_steps<CM>() > 0 ? prefetch_steps<CM>() : 1;
        const int z = zlo + pout(i);  // output plane
        const int r6n = r6 == 5 ? 0 : r6 + 1, r5n = r5 == 4 ? 0 : r5 + 1;
        float *const Tf = s.t[r6];
        float *const Tz = s.t[r6 >= 4 ? r6 - 4 : r6 + 2];
        float *const Xf = s.x[r5];
        float *const Xz = s.x[r5n];
        const uint32_t so_o = (uint32_t)pout(i) * pb8;

        // streams of the output plane, kept raw until used
        CoefRec<CM> kr;
        float vprev = 0.0f;
        if constexpr (OUT && (PF & 1)) {
            kr = kp[PS];
            vprev = ppv[PS];
        } else if constexpr (OUT) {
            load_next_rec<CM, PCPOL>(recs, pxy, vor, kr);
            vprev = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r_prev, vo8, so_o, PCPOL));
        }
        const bool xw = (XW == 1) || (XW == 2 && i < nz + 4);
        float nx, ny, gxh[NH];
        if constexpr (TWOPASS) {
            const uint32_t so = pin(i) * pb8;
            nx = __builtin_bit_cast(float, bload32(r_dx, vo8, so));
            ny = __builtin_bit_cast(float, bload32(r_dy, vo8, so));
            if (xw) {
#pragma unroll
                for (int m = 0; m < NH; m++)
                    if (hrow[m]) {
                        const int j = wv + m * TY;
                        const int gy = y0 - 4 + (j < 4 ? j : TY + j);
                        const uint32_t o = (x < a.sx && gy < a.sy)
                                               ? ((uint32_t)(gy * a.px + x) * 2u + fld) * 4u : FL_OOB;
                        gxh[m] = __builtin_bit_cast(float, bload32(r_dx, o, so));
                    }
            }
        }
        const float *cp = Tf + ow + CW;
        float nc = cp[0];
        if constexpr (!TWOPASS) {
            nx = der1_lds<!ISO>(cp, 2, a.dxinv);
            ny = der1_lds<!ISO>(cp, RW, a.dyinv);
        }
        if (xw) {
            Xf[oxw + CX] = nx;
#pragma unroll
            for (int m = 0; m < NH; m++)
                if (hrow[m]) {
                    if constexpr (TWOPASS) Xf[oxw + ohx[m]] = gxh[m];
                    else Xf[oxw + ohx[m]] = der1_lds<!ISO>(Tf + ow + ohr[m] + 8, 2, a.dxinv);
                }
        }
        // formed here, in this step: keeps the compiler from carrying the raw LDS values to
        // the step that first uses the queue entry instead
        fl_pin(nc);
        fl_pin(nx);
        fl_pin(ny);
        constexpr int ins = GEN ? 8 : PH;
        qc_[ins] = nc;
        qx_[ins] = nx;
        qy_[ins] = ny;

        float dxx = 0.f, dyy = 0.f, dxy = 0.f;
        if constexpr (OUT) {
            const float *zp = Tz + ow + CW;
            dxx = der2_lds<!ISO>(zp, 2, a.dxxinv);
            dyy = der2_lds<!ISO>(zp, RW, a.dyyinv);
            dxy = der1_lds<!ISO>(Xz + oxw + CX, XW2, a.dyinv);
        }
        // next plane into the ring from its set, then the set refilled SA planes ahead
        if (i + 1 <= nz + 7) stage(ic<(SET + 1) % SA>{}, s.t[r6n]);
        if (i + 1 + SA <= nz + 7) fetch(ic<(SET + 1) % SA>{}, i + 1 + SA);

        if constexpr (OUT) {
            constexpr int b = GEN ? 0 : (PH + 2) % 10;  // slot of plane i-8
            float qz[9], qzx[9], qzy[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                qz[k] = qc_[(b + k) % 10];
                qzx[k] = qx_[(b + k) % 10];
                qzy[k] = qy_[(b + k) % 10];
            }
            const float dzz = der2_q<!ISO>(qz, a.dzzinv);
            float dxz = der1_q<!ISO>(qzx, dzs);
            float dyz = der1_q<!ISO>(qzy, dzs);
            if constexpr (ISO) {
                dxz *= zsg;
                dyz *= zsg;
            }
            Coef kc;
            decode_rec<CM>(kr, fld, kc);
            // H1, H2 of the own field, association order of original/sample.h:46-53
            float h1 = kc.c[0] * dxx;
            h1 = __builtin_fmaf(kc.c[1], dyy, h1);
            h1 = __builtin_fmaf(kc.c[2], dzz, h1);
            h1 = __builtin_fmaf(kc.c[3], dxy, h1);
            h1 = __builtin_fmaf(kc.c[5], dxz, h1);
            h1 = __builtin_fmaf(kc.c[4], dyz, h1);
            const float h2 = dxx + dyy + dzz - h1;
            // p lanes receive h1q, q lanes h2p: swap within lane pairs (DPP quad_perm 1,0,3,2)
            const float send = fld ? h1 : h2;
            const float other = __builtin_bit_cast(
                float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, send), 0xB1, 0xF, 0xF, true));
            // rhsp = v2px h2p + v2pz h1q + v2sz (h1p - h1q)
            // rhsq = v2pn h2p + v2pz h1q - v2sz (h2p - h2q) = ... + v2sz (h2q - h2p)
            // with h2p, h1q selected per half (original/sample.h:75-81, same roundings)
            const float h2p = fld ? other : h2, h1q = fld ? h1 : other;
            float rhs = ((fld && CM != CM_Q16R) ? kc.c[9] : kc.c[6]) * h2p;
            rhs = __builtin_fmaf(kc.c[7], h1q, rhs);
            rhs = __builtin_fmaf(kc.c[8], (fld ? h2 : h1) - other, rhs);
            const float cur = qz[4];
            const float base = 2.0f * cur - vprev;
            float nw;
            if constexpr (CM == CM_F32 || CM == CM_F32R) nw = __builtin_fmaf(rhs * dts, a.dt, base);
            else if constexpr (CM == CM_Q16R) nw = __builtin_fmaf(rhs, rq * kc.c[9], base);
            else if constexpr (ISO) nw = __builtin_fmaf(rhs, a.dxxinv, base);
            else nw = base + rhs;
            if (z == rec_z) {
                if (fld == 0 && active)
                    ctl.trace[(y - a.bord) * (a.sx - 2 * a.bord) + (x - a.bord)] = nw;
            }
            if (__builtin_expect(z == src_z, 0)) {  // uniform: the source plane only
                if (col == src_col && active) nw += ctl.src_add;
            }
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, nw), r_prev, vo8, so_o, PCPOL);
        }
        if constexpr ((PF & 2) != 0 || (PF & 4) != 0) {
            if ((PF & 2) || pf_next) {
                constexpr int NS2 = (PS + PD) % NSETS;
                load_next_rec<CM, PCPOL>(recs, pxy, vor, kp[NS2]);
                ppv[NS2] = __builtin_bit_cast(
                    float, __builtin_amdgcn_raw_buffer_load_b32(r_prev, vo8, (uint32_t)pout(i + PD) * pb8, PCPOL));
            }
        }
        if constexpr (GEN) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                qc_[k] = qc_[k + 1];
                qx_[k] = qx_[k + 1];
                qy_[k] = qy_[k + 1];
            }
        }
        r6 = r6n;
        r5 = r5n;
        __syncthreads();
    };

    // planes 0-3: queues only; 4-7: queues + Dx halo planes (when an output plane); planes
    // 8-PD .. 7 load the streams of the first output planes when a static group follows
    constexpr int PD = prefetch_steps<CM>() > 0 ? prefetch_steps<CM>() : 1;
    constexpr int PFB = prefetch_steps<CM>() ? 4 : 0;
    const bool grp0 = 8 + 9 < nz + 4;
    // step j < 8 loads the streams of step j + PD when that is an output step of the first
    // group; its prefetch set is (j - 8) mod NSETS
    auto pro = [&](auto Jc, auto XWc) {
        constexpr int j = decltype(Jc)::value;
        step(ic<j>{}, XWc, ic<0>{}, ic<0>{}, ic<(j + PD >= 8 ? PFB : 0)>{}, ic<j & 1>{},
             ic<(j + 10 * NSETS - 8) % NSETS>{}, j, j + PD >= 8 ? grp0 : false);
    };
    pro(ic<0>{}, ic<0>{});
    pro(ic<1>{}, ic<0>{});
    pro(ic<2>{}, ic<0>{});
    pro(ic<3>{}, ic<0>{});
    pro(ic<4>{}, ic<2>{});
    pro(ic<5>{}, ic<2>{});
    pro(ic<6>{}, ic<2>{});
    pro(ic<7>{}, ic<2>{});
    // steady state: groups of 10 planes (i = 8 mod 10 at their start), each writing Dx
    // halos and an output plane, with the output streams PD steps ahead
    int i = 8;
    for (; i + 9 < nz + 4; i += 10) {
        const bool more = i + 10 + 9 < nz + 4;
        unroll_phases(
            [&](auto P) {
                constexpr int p = decltype(P)::value;
                constexpr int pf = prefetch_steps<CM>() ? (1 | (p + PD <= 9 ? 2 : 4)) : 0;
                step(ic<(8 + p) % 10>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<pf>{}, ic<p & 1>{}, ic<p % NSETS>{}, i + p,
                     more);
            },
            std::make_integer_sequence<int, 10>{});
    }
    // the rest (i even here), queues shifted after each plane, two steps per round so that
    // the register sets keep their parity
    for (; i + 1 <= nz + 7; i += 2) {
        step(ic<0>{}, ic<2>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<0>{}, ic<0>{}, i, false);
        step(ic<0>{}, ic<2>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<1>{}, ic<0>{}, i + 1, false);
    }
    if (i <= nz + 7) step(ic<0>{}, ic<2>{}, ic<1>{}, ic<1>{}, ic<0>{}, ic<0>{}, ic<0>{}, i, false);
}

// Waves per SIMD the register allocator is asked for.  With 8-row tiles the LDS rings
// (51 KB per workgroup) allow three workgroups per CU, with 12-row tiles (63 KB) two:
// 6 waves per SIMD, <= 80 VGPRs, which the 16-bit and recompute forms meet.  The FP32
// records keep 4 waves (83 VGPRs): a 6-wave build spills only outside the steady loop but
// measured no faster at 344^3.  4- and 16-row tiles are held to 4 waves per SIMD by their LDS.
// FL_STEP_WAVES overrides the 8- and 12-row value at build time.
template <int TY, int CM>
constexpr int step_waves()
{
#ifdef FL_STEP_WAVES
    return (TY == 8 || TY == 12) ? FL_STEP_WAVES : 4;
#else
    return ((TY == 8 || TY == 12) && CM != CM_F32) ? 6 : 4;
#endif
}

template <int TY, int CM, bool TWOPASS, bool ISO, bool HINTS>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(step_waves<TY, CM>())))
void tti_step(const StepArgs a)
{
    __shared__ StepLds<TY> s;
    const StepCtl ctl = step_ctl(a);
    if (a.run && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0)
        a.run[1 - a.par] = a.run[a.par] + 1;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    if (lb >= a.nblk) return;
    const int t = step_order(lb, a.nblk, ctl.it);
    int tx, ty, tz;
    tile_of(t, a.ntx, a.nty, a.sw, tx, ty, tz);
    step_tile<TY, CM, TWOPASS, ISO, HINTS ? 2 : 0>(a, ctl, uni(tx), uni(ty), uni(tz), step_down(ctl.it), s);
}

}  // namespace fl
