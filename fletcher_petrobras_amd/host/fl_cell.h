/*
 * Per-cell coefficient formulas and record packings of the Fletcher TTI model, shared by
 * the host C code (fl_model.c) and the device precompute kernels (fl_driver.hip): the same
 * statements on both sides, so host and device records agree bit for bit wherever the two
 * libraries' double sine and cosine agree.  Every mathematical call takes an explicit
 * double argument (C++ would otherwise pick the float overloads on the device).
 *
 *   fl_cell_f32     precomp.h:25-38, 53-58: trig in double rounded to float, float products
 *   fl_cell_f64     the same formulas entirely in double (the FP64 reference's coefficients)
 *   fl_cell_pack16  the 10 coefficients (v2* times dt^2) in BF16 or FP16
 *   fl_cell_recompute16 / _q16 / 32   the recompute inputs (velocity factors, angle turns)
 */
#ifndef FLETCHER_FL_CELL_H
#define FLETCHER_FL_CELL_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define FL_CELL __host__ __device__ static inline
#else
#define FL_CELL static inline
#endif

#define FL_CELL_PI 3.14159265358979323846

/* precomp.h order: 0 ch1dxx, 1 ch1dyy, 2 ch1dzz, 3 ch1dxy, 4 ch1dyz, 5 ch1dxz, 6 v2px, 7 v2pz,
 * 8 v2sz, 9 v2pn */
FL_CELL void fl_cell_f32(float vpz, float vsv, float eps, float del, float phi, float theta, float *c)
{
    const float sinTheta = (float)sin((double)theta);
    const float cosTheta = (float)cos((double)theta);
    const float sin2Theta = (float)sin(2.0 * (double)theta);
    const float sinPhi = (float)sin((double)phi);
    const float cosPhi = (float)cos((double)phi);
    const float sin2Phi = (float)sin(2.0 * (double)phi);
    c[0] = sinTheta * sinTheta * cosPhi * cosPhi;
    c[1] = sinTheta * sinTheta * sinPhi * sinPhi;
    c[2] = cosTheta * cosTheta;
    c[3] = sinTheta * sinTheta * sin2Phi;
    c[4] = sin2Theta * sinPhi;
    c[5] = sin2Theta * cosPhi;
    const float v2pz = vpz * vpz;
    c[8] = vsv * vsv;
    c[7] = v2pz;
    c[6] = (float)((double)v2pz * (1.0 + 2.0 * (double)eps));
    c[9] = (float)((double)v2pz * (1.0 + 2.0 * (double)del));
}

FL_CELL void fl_cell_f64(float vpz, float vsv, float eps, float del, float phi, float theta, double *c)
{
    const double st = sin((double)theta), ct = cos((double)theta);
    const double s2t = sin(2.0 * (double)theta);
    const double sp = sin((double)phi), cp = cos((double)phi);
    const double s2p = sin(2.0 * (double)phi);
    c[0] = st * st * cp * cp;
    c[1] = st * st * sp * sp;
    c[2] = ct * ct;
    c[3] = st * st * s2p;
    c[4] = s2t * sp;
    c[5] = s2t * cp;
    const double v2pz = (double)vpz * (double)vpz;
    c[8] = (double)vsv * (double)vsv;
    c[7] = v2pz;
    c[6] = v2pz * (1.0 + 2.0 * (double)eps);
    c[9] = v2pz * (1.0 + 2.0 * (double)del);
}

/* Round to nearest even into IEEE binary16; 65520 and above round to infinity. */
FL_CELL uint16_t fl_cell_to_f16(double v)
{
    const uint16_t sign = signbit(v) ? 0x8000 : 0;
    if (isnan(v)) return 0x7e00;
    const double a = fabs(v);
    if (a >= 65520.0) return (uint16_t)(sign | 0x7c00);
    int e;
    const double m = frexp(a, &e); /* a = m 2^e, 0.5 <= m < 1 */
    if (a != 0.0 && e - 1 >= -14) {
        double q = nearbyint(ldexp(m, 11)); /* 11-bit significand, ties to even */
        int E = e - 1;
        if (q >= 2048.0) {
            q = 1024.0;
            E += 1;
        }
        return (uint16_t)(sign | ((E + 15) << 10) | ((int)q - 1024));
    }
    /* subnormal range: multiples of 2^-24 (1024 * 2^-24 encodes the smallest normal) */
    return (uint16_t)(sign | (uint16_t)nearbyint(ldexp(a, 24)));
}

/* Round to nearest even into bfloat16 (8 significant bits). */
FL_CELL uint16_t fl_cell_to_bf16(double v)
{
    if (isnan(v)) return 0x7fc0;
    double r = v;
    if (v != 0.0 && isfinite(v)) {
        int e;
        const double m = frexp(v, &e);
        r = ldexp(nearbyint(ldexp(m, 8)), e - 8); /* 8 significant bits, ties to even */
    }
    union {
        float f;
        uint32_t u;
    } b;
    b.f = (float)r; /* exact within the float range */
    return (uint16_t)(b.u >> 16);
}

/* Angle in turns as 16-bit fixed point, rounded to nearest, modulo one turn. */
FL_CELL uint16_t fl_cell_angle_u16(double turns)
{
    const double t = turns - floor(turns); /* [0, 1) */
    return (uint16_t)((uint32_t)nearbyint(t * 65536.0) & 0xffffu);
}

/* The 10 coefficients of fl_cell_f64 with v2* times dt^2, rounded once to BF16 or FP16. */
FL_CELL void fl_cell_pack16(float vpz, float vsv, float eps, float del, float phi, float theta, double dt,
                            int bf16, uint16_t *out)
{
    double c[10];
    fl_cell_f64(vpz, vsv, eps, del, phi, theta, c);
    const double dt2 = dt * dt;
    for (int k = 6; k < 10; k++) c[k] *= dt2;
    for (int k = 0; k < 10; k++) out[k] = bf16 ? fl_cell_to_bf16(c[k]) : fl_cell_to_f16(c[k]);
}

/* Recompute inputs: v2px, v2pz, v2sz, v2pn times dt^2 in FP16, theta and phi as u16 turns. */
FL_CELL void fl_cell_recompute16(float vpz, float vsv, float eps, float del, float phi, float theta, double dt,
                                 uint16_t *out)
{
    const double dt2 = dt * dt;
    const double v2pz = (double)vpz * (double)vpz;
    out[0] = fl_cell_to_f16(v2pz * (1.0 + 2.0 * (double)eps) * dt2);
    out[1] = fl_cell_to_f16(v2pz * dt2);
    out[2] = fl_cell_to_f16((double)vsv * (double)vsv * dt2);
    out[3] = fl_cell_to_f16(v2pz * (1.0 + 2.0 * (double)del) * dt2);
    out[4] = fl_cell_angle_u16((double)theta / (2.0 * FL_CELL_PI));
    out[5] = fl_cell_angle_u16((double)phi / (2.0 * FL_CELL_PI));
}

/* The four FP64 velocity factors times dt^2 of a cell (v2px, v2pz, v2sz, v2pn). */
FL_CELL void fl_cell_q16_factors(float vpz, float vsv, float eps, float del, double dt2, double *f)
{
    const double v2pz = (double)vpz * (double)vpz;
    f[0] = v2pz * (1.0 + 2.0 * (double)eps) * dt2;
    f[1] = v2pz * dt2;
    f[2] = (double)vsv * (double)vsv * dt2;
    f[3] = v2pz * (1.0 + 2.0 * (double)del) * dt2;
}

/* Block floating point of the four factors: 15-bit mantissas q * 2^E of one exponent E shared
 * by the cell, the smallest that holds the four, clamped to [emin, emin + 15]; returns
 * E - emin. */
FL_CELL int fl_cell_q16_block(const double *f, int emin, uint32_t *q)
{
    double r = 0.0;
    for (int k = 0; k < 4; k++) r = f[k] / 32767.5 > r ? f[k] / 32767.5 : r;
    int e = emin;
    if (r > 0.0) {
        int ex;
        frexp(r, &ex); /* r = fr * 2^ex, fr in [0.5, 1): 2^ex is the first power above r */
        e = ex;
        while (e > emin) { /* the smallest E whose rounded mantissas still fit */
            int fits = 1;
            for (int k = 0; k < 4; k++)
                if (nearbyint(ldexp(f[k], -(e - 1))) > 32767.0) fits = 0;
            if (!fits) break;
            e--;
        }
        if (e < emin) e = emin;
        if (e > emin + 15) e = emin + 15;
    }
    for (int k = 0; k < 4; k++) {
        const double u = nearbyint(ldexp(f[k], -e));
        q[k] = (uint32_t)(u < 0.0 ? 0.0 : (u > 32767.0 ? 32767.0 : u));
    }
    return e - emin;
}

/* Largest f / 32767.5 over a cell's factors (the exponent base is taken from the grid's max). */
FL_CELL double fl_cell_q16_need(float vpz, float vsv, float eps, float del, double dt2)
{
    double f[4], m = 0.0;
    fl_cell_q16_factors(vpz, vsv, eps, del, dt2, f);
    for (int k = 0; k < 4; k++)
        if (f[k] / 32767.5 > m) m = f[k] / 32767.5;
    return m;
}

/* Exponent base of the grid from its largest need: that cell gets E = 15. */
FL_CELL int fl_cell_q16_base_of(double m)
{
    if (!(m > 0.0)) return 0;
    int ex;
    frexp(m, &ex);
    return ex - 15;
}

/* Block floating-point recompute inputs: record words (v2px, v2pn) and (v2pz, v2sz) with the
 * mantissas in bits 0-14 and 15-29, the exponent's bits 0-1 in bits 30-31 of the first word
 * and bits 2-3 in those of the second; halfwords out[0] = low (v2px), out[3] = high of word 0,
 * out[1] = low (v2pz), out[2] = high of word 1; theta, phi as u16 turns. */
FL_CELL void fl_cell_recompute_q16(float vpz, float vsv, float eps, float del, float phi, float theta,
                                   double dt, int ebase, uint16_t *out)
{
    double f[4];
    uint32_t q[4];
    fl_cell_q16_factors(vpz, vsv, eps, del, dt * dt, f);
    const uint32_t e = (uint32_t)fl_cell_q16_block(f, ebase, q);
    const uint32_t w0 = q[0] | (q[3] << 15) | ((e & 3u) << 30);
    const uint32_t w1 = q[1] | (q[2] << 15) | ((e >> 2) << 30);
    out[0] = (uint16_t)(w0 & 0xffffu);
    out[3] = (uint16_t)(w0 >> 16);
    out[1] = (uint16_t)(w1 & 0xffffu);
    out[2] = (uint16_t)(w1 >> 16);
    out[4] = fl_cell_angle_u16((double)theta / (2.0 * FL_CELL_PI));
    out[5] = fl_cell_angle_u16((double)phi / (2.0 * FL_CELL_PI));
}

/* FP32 recompute inputs: the four velocity factors of fl_cell_f32 bit for bit, theta and phi
 * in turns reduced to [-1/2, 1/2] in FP64 and rounded once. */
FL_CELL void fl_cell_recompute32(float vpz, float vsv, float eps, float del, float phi, float theta, float *out)
{
    const float v2pz = vpz * vpz;
    out[0] = (float)((double)v2pz * (1.0 + 2.0 * (double)eps));
    out[1] = v2pz;
    out[2] = vsv * vsv;
    out[3] = (float)((double)v2pz * (1.0 + 2.0 * (double)del));
    const double tt = (double)theta / (2.0 * FL_CELL_PI), tp = (double)phi / (2.0 * FL_CELL_PI);
    out[4] = (float)(tt - nearbyint(tt));
    out[5] = (float)(tp - nearbyint(tp));
}

#endif
