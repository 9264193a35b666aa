/*
 * Host model setup for the Fletcher TTI path: the once-per-run work of the
 * reference front-end (model arrays, random boundary, CFL, Ricker source,
 * coefficient precompute), restated with the reference's float/double
 * promotion order so the FP32 results are bitwise those of the reference.
 * Reference citations are per function.
 */
#include "fletcher/fl_model.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FL_SIGMA 0.75 /* fletcher.h:10 */

/* ------------------------------------------------------------------ rand()
 * glibc random(): state r[0]=seed, r[i]=16807*r[i-1] mod (2^31-1) for i<31,
 * r[i]=r[i-31] for 31<=i<34, then r[i]=r[i-31]+r[i-3] (mod 2^32); the k-th
 * output is r[k+344]>>1.  Kept as a 34-entry ring of the most recent values. */
void fl_rand_seed(fl_rand_t *s, uint32_t seed)
{
    int32_t r[34];
    r[0] = (int32_t)(seed == 0 ? 1u : seed);
    for (int i = 1; i < 31; i++) {
        const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int64_t w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        r[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    memcpy(s->r, r, sizeof(r));
    s->k = 34; /* index of the next value of the recurrence */
    for (int i = 34; i < 344; i++) (void)fl_rand_next(s);
}

int32_t fl_rand_next(fl_rand_t *s)
{
    const int k = s->k;
    const uint32_t v = (uint32_t)s->r[(k - 31) % 34] + (uint32_t)s->r[(k - 3) % 34];
    s->r[k % 34] = (int32_t)v;
    s->k = (k + 1 >= 68) ? k + 1 - 34 : k + 1; /* k stays in [34,68), k mod 34 preserved */
    return (int32_t)(v >> 1);
}

/* ------------------------------------------------------------------ models */

int fl_model_homogeneous(int form, int64_t n, float *vpz, float *vsv, float *epsilon,
                         float *delta, float *phi, float *theta)
{
    /* main.c:168-221 */
    for (int64_t i = 0; i < n; i++) {
        switch (form) {
        case FL_FORM_ISO:
            vpz[i] = 3000.0;
            epsilon[i] = 0.0;
            delta[i] = 0.0;
            phi[i] = 0.0;
            theta[i] = 0.0;
            vsv[i] = 0.0;
            break;
        case FL_FORM_VTI:
            vpz[i] = 3000.0;
            epsilon[i] = 0.24;
            delta[i] = 0.1;
            phi[i] = 0.0;
            theta[i] = 0.0;
            vsv[i] = vpz[i] * sqrtf(fabsf(epsilon[i] - delta[i]) / FL_SIGMA);
            break;
        case FL_FORM_TTI:
            vpz[i] = 3000.0;
            epsilon[i] = 0.24;
            delta[i] = 0.1;
            phi[i] = 1.0; /* the reference avoids null coefficients with phi=1 */
            theta[i] = atanf(1.0);
            vsv[i] = vpz[i] * sqrtf(fabsf(epsilon[i] - delta[i]) / FL_SIGMA);
            break;
        default:
            return -1;
        }
    }
    return 0;
}

float fl_recommended_dt(int64_t n, const float *vpz, const float *epsilon, float dx, float dy,
                        float dz, double mi)
{
    /* main.c:225-236: maxvel over vpz*sqrt(1+2eps), recdt = MI*min(d)/maxvel */
    float maxvel = vpz[0] * sqrt(1.0 + 2 * epsilon[0]);
    for (int64_t i = 1; i < n; i++) maxvel = fmaxf(maxvel, vpz[i] * sqrt(1.0 + 2 * epsilon[i]));
    float mindelta = dx;
    if (dy < mindelta) mindelta = dy;
    if (dz < mindelta) mindelta = dz;
    float recdt = (mi * mindelta) / maxvel;
    return recdt;
}

int fl_num_steps(float tmax, float dt)
{
    return (int)ceil(tmax / dt); /* main.c:101, float quotient */
}

void fl_random_boundary(uint32_t seed, int sx, int sy, int sz, int nx, int ny, int nz, int bord,
                        int absorb, float *vpz, float *vsv)
{
    /* boundary.c:7-98 */
#define IND(ix, iy, iz) ((((int64_t)(iz)) * sy + (iy)) * (int64_t)sx + (ix))
    fl_rand_t rs;
    fl_rand_seed(&rs, seed);
    float maxP = 0.0, maxS = 0.0;
    for (int iz = bord + absorb; iz < nz + bord + absorb; iz++)
        for (int iy = bord + absorb; iy < ny + bord + absorb; iy++)
            for (int64_t i = IND(bord + absorb, iy, iz); i < IND(nx + bord + absorb, iy, iz); i++) {
                maxP = fmaxf(vpz[i], maxP);
                maxS = fmaxf(vsv[i], maxS);
            }
    const int bordLen = bord + absorb - 1; /* last index of the low absorbing zone */
    const int firstIn = bordLen + 1;       /* first interior index */
    const float frac = 1.0 / (float)(absorb);
    for (int iz = 0; iz < sz; iz++)
        for (int iy = 0; iy < sy; iy++)
            for (int ix = 0; ix < sx; ix++) {
                const int64_t i = IND(ix, iy, iz);
                if ((iz >= firstIn && iz <= bordLen + nz) && (iy >= firstIn && iy <= bordLen + ny) &&
                    (ix >= firstIn && ix <= bordLen + nx))
                    continue; /* interior untouched */
                if ((iz >= bord && iz <= 2 * bordLen + nz) && (iy >= bord && iy <= 2 * bordLen + ny) &&
                    (ix >= bord && ix <= 2 * bordLen + nx)) {
                    int distz, disty, distx, ivelz, ively, ivelx;
                    if (iz > bordLen + nz) { distz = iz - bordLen - nz; ivelz = bordLen + nz; }
                    else if (iz < firstIn) { distz = firstIn - iz; ivelz = firstIn; }
                    else { distz = 0; ivelz = iz; }
                    if (iy > bordLen + ny) { disty = iy - bordLen - ny; ively = bordLen + ny; }
                    else if (iy < firstIn) { disty = firstIn - iy; ively = firstIn; }
                    else { disty = 0; ively = iy; }
                    if (ix > bordLen + nx) { distx = ix - bordLen - nx; ivelx = bordLen + nx; }
                    else if (ix < firstIn) { distx = firstIn - ix; ivelx = firstIn; }
                    else { distx = 0; ivelx = ix; }
                    int dist = (disty > distz) ? disty : distz;
                    dist = (dist > distx) ? dist : distx;
                    const float bordDist = (float)(dist)*frac;
                    const float rfac = (float)fl_rand_next(&rs) / (float)2147483647;
                    const int64_t j = IND(ivelx, ively, ivelz);
                    vpz[i] = vpz[j] * (1.0 - bordDist) + maxP * rfac * bordDist;
                    vsv[i] = vsv[j] * (1.0 - bordDist) + maxS * rfac * bordDist;
                } else {
                    vpz[i] = 0.0;
                    vsv[i] = 0.0;
                }
            }
#undef IND
}

float fl_source(float dt, int it)
{
    /* source.c:7-15 with source.h:11-14; the double constants promote exactly as there */
    const double FCUT = 40.0, PICUBE = 31.00627668029982017537;
    const double TWOSQRTPI = 3.54490770181103205458, THREESQRTPI = 5.31736155271654808184;
    float tf, fc, fct, expo;
    tf = TWOSQRTPI / FCUT;
    fc = FCUT / THREESQRTPI;
    fct = fc * (((float)it) * dt - tf);
    expo = PICUBE * fct * fct;
    return ((1.0f - 2.0f * expo) * expf(-expo));
}

void fl_precompute_f32(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                       const float *delta, const float *phi, const float *theta, float *const *c)
{
    /* precomp.h:25-38 and :53-58 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const float sinTheta = sin(theta[i]);
        const float cosTheta = cos(theta[i]);
        const float sin2Theta = sin(2.0 * theta[i]);
        const float sinPhi = sin(phi[i]);
        const float cosPhi = cos(phi[i]);
        const float sin2Phi = sin(2.0 * phi[i]);
        c[0][i] = sinTheta * sinTheta * cosPhi * cosPhi;
        c[1][i] = sinTheta * sinTheta * sinPhi * sinPhi;
        c[2][i] = cosTheta * cosTheta;
        c[3][i] = sinTheta * sinTheta * sin2Phi;
        c[4][i] = sin2Theta * sinPhi;
        c[5][i] = sin2Theta * cosPhi;
        const float v2pz = vpz[i] * vpz[i];
        c[8][i] = vsv[i] * vsv[i];
        c[7][i] = v2pz;
        c[6][i] = v2pz * (1.0 + 2.0 * epsilon[i]);
        c[9][i] = v2pz * (1.0 + 2.0 * delta[i]);
    }
}

void fl_precompute_f64(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                       const float *delta, const float *phi, const float *theta, double *const *c)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double st = sin((double)theta[i]), ct = cos((double)theta[i]);
        const double s2t = sin(2.0 * theta[i]);
        const double sp = sin((double)phi[i]), cp = cos((double)phi[i]);
        const double s2p = sin(2.0 * phi[i]);
        c[0][i] = st * st * cp * cp;
        c[1][i] = st * st * sp * sp;
        c[2][i] = ct * ct;
        c[3][i] = st * st * s2p;
        c[4][i] = s2t * sp;
        c[5][i] = s2t * cp;
        const double v2pz = (double)vpz[i] * vpz[i];
        c[8][i] = (double)vsv[i] * vsv[i];
        c[7][i] = v2pz;
        c[6][i] = v2pz * (1.0 + 2.0 * epsilon[i]);
        c[9][i] = v2pz * (1.0 + 2.0 * delta[i]);
    }
}

float fl_cfl_dt(int64_t n, const float *vpz, const float *epsilon, float h, float courant)
{
    double vmax = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const double v = vpz[i] * sqrt(1.0 + 2.0 * epsilon[i]);
        if (v > vmax) vmax = v;
    }
    return vmax > 0.0 ? (float)(courant * h / vmax) : 0.0f;
}

void fl_extract_window(const float *src, int sx, int sy, int sz, int x0, int y0, int z0, int wx,
                       int wy, int wz, float *dst)
{
    (void)sz;
    for (int iz = 0; iz < wz; iz++)
        for (int iy = 0; iy < wy; iy++)
            memcpy(dst + ((int64_t)iz * wy + iy) * wx,
                   src + (((int64_t)(z0 + iz)) * sy + (y0 + iy)) * sx + x0, sizeof(float) * wx);
}

/* ------------------------------------------------------------------ 16-bit packs */

uint16_t fl_f64_to_f16(double v)
{
    const uint16_t sign = signbit(v) ? 0x8000 : 0;
    if (isnan(v)) return 0x7e00;
    const double a = fabs(v);
    if (a >= 65520.0) return sign | 0x7c00; /* rounds to infinity */
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

uint16_t fl_f64_to_bf16(double v)
{
    if (isnan(v)) return 0x7fc0;
    double r = v;
    if (v != 0.0 && isfinite(v)) {
        int e;
        const double m = frexp(v, &e);
        r = ldexp(nearbyint(ldexp(m, 8)), e - 8); /* 8 significant bits, ties to even */
    }
    const float f = (float)r; /* exact within the float range */
    uint32_t u;
    memcpy(&u, &f, 4);
    return (uint16_t)(u >> 16);
}

void fl_pack_coefs16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                     const float *delta, const float *phi, const float *theta, double dt, int bf16,
                     uint16_t *const *out)
{
    const double dt2 = dt * dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double st = sin((double)theta[i]), ct = cos((double)theta[i]);
        const double s2t = sin(2.0 * theta[i]);
        const double sp = sin((double)phi[i]), cp = cos((double)phi[i]);
        const double s2p = sin(2.0 * phi[i]);
        const double vp2 = (double)vpz[i] * vpz[i];
        const double c[10] = {st * st * cp * cp, st * st * sp * sp, ct * ct, st * st * s2p,
                              s2t * sp, s2t * cp, vp2 * (1.0 + 2.0 * epsilon[i]) * dt2, vp2 * dt2,
                              (double)vsv[i] * vsv[i] * dt2, vp2 * (1.0 + 2.0 * delta[i]) * dt2};
        for (int k = 0; k < 10; k++) out[k][i] = bf16 ? fl_f64_to_bf16(c[k]) : fl_f64_to_f16(c[k]);
    }
}

uint16_t fl_angle_u16(double turns)
{
    const double t = turns - floor(turns);           /* [0, 1) */
    return (uint16_t)((uint32_t)nearbyint(t * 65536.0) & 0xffffu);
}

void fl_pack_recompute16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                         const float *delta, const float *phi, const float *theta, double dt,
                         uint16_t *const *out)
{
    const double pi = 3.14159265358979323846, dt2 = dt * dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        /* the four velocity factors of fl_precompute_f64, times dt^2 */
        const double v2pz = (double)vpz[i] * vpz[i];
        out[0][i] = fl_f64_to_f16(v2pz * (1.0 + 2.0 * epsilon[i]) * dt2);
        out[1][i] = fl_f64_to_f16(v2pz * dt2);
        out[2][i] = fl_f64_to_f16((double)vsv[i] * vsv[i] * dt2);
        out[3][i] = fl_f64_to_f16(v2pz * (1.0 + 2.0 * delta[i]) * dt2);
        /* angles as 16-bit fixed point in revolutions, round to nearest, modulo one turn */
        out[4][i] = fl_angle_u16((double)theta[i] / (2.0 * pi));
        out[5][i] = fl_angle_u16((double)phi[i] / (2.0 * pi));
    }
}

/* Block floating point of the four FP64 velocity factors times dt^2 of one cell: 15-bit
 * mantissas q * 2^E of one exponent E shared by the cell, the smallest that holds the four,
 * clamped to [emin, emin + 15]. */
static int q16_block(const double f[4], int emin, uint32_t q[4])
{
    double r = 0.0;
    for (int k = 0; k < 4; k++) r = f[k] / 32767.5 > r ? f[k] / 32767.5 : r;
    int e = emin;
    if (r > 0.0) {
        int ex;
        frexp(r, &ex);  /* r = fr * 2^ex, fr in [0.5, 1): 2^ex is the first power above r */
        e = ex;
        while (e > emin) {  /* the smallest E whose rounded mantissas still fit */
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

static void q16_factors(int64_t i, const float *vpz, const float *vsv, const float *epsilon,
                        const float *delta, double dt2, double f[4])
{
    const double v2pz = (double)vpz[i] * vpz[i];
    f[0] = v2pz * (1.0 + 2.0 * epsilon[i]) * dt2;
    f[1] = v2pz * dt2;
    f[2] = (double)vsv[i] * vsv[i] * dt2;
    f[3] = v2pz * (1.0 + 2.0 * delta[i]) * dt2;
}

int fl_recompute_q16_base(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                          const float *delta, double dt)
{
    /* the exponent of the cell that needs the largest one, less 15 */
    double m = 0.0;
    const double dt2 = dt * dt;
#pragma omp parallel for schedule(static) reduction(max : m)
    for (int64_t i = 0; i < n; i++) {
        double f[4];
        q16_factors(i, vpz, vsv, epsilon, delta, dt2, f);
        for (int k = 0; k < 4; k++)
            if (f[k] / 32767.5 > m) m = f[k] / 32767.5;
    }
    if (!(m > 0.0)) return 0;
    int ex;
    frexp(m, &ex);
    return ex - 15;
}

void fl_pack_recompute_q16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                           const float *delta, const float *phi, const float *theta, double dt,
                           int ebase, uint16_t *const *out)
{
    const double pi = 3.14159265358979323846, dt2 = dt * dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double f[4];
        uint32_t q[4];
        q16_factors(i, vpz, vsv, epsilon, delta, dt2, f);
        const uint32_t e = (uint32_t)q16_block(f, ebase, q);
        /* record words (v2px, v2pn) and (v2pz, v2sz): mantissas in bits 0-14 and 15-29,
         * e bits 0-1 in bits 30-31 of the first word, e bits 2-3 in those of the second */
        const uint32_t w0 = q[0] | (q[3] << 15) | ((e & 3u) << 30);
        const uint32_t w1 = q[1] | (q[2] << 15) | ((e >> 2) << 30);
        out[0][i] = (uint16_t)(w0 & 0xffffu);
        out[3][i] = (uint16_t)(w0 >> 16);
        out[1][i] = (uint16_t)(w1 & 0xffffu);
        out[2][i] = (uint16_t)(w1 >> 16);
        out[4][i] = fl_angle_u16((double)theta[i] / (2.0 * pi));
        out[5][i] = fl_angle_u16((double)phi[i] / (2.0 * pi));
    }
}

void fl_pack_recompute32(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                         const float *delta, const float *phi, const float *theta, float *const *out)
{
    const double pi = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        /* the FP32 velocity factors of precomp.h:53-58, as fl_precompute_f32 forms them */
        const float v2pz = vpz[i] * vpz[i];
        out[0][i] = v2pz * (1.0 + 2.0 * epsilon[i]);
        out[1][i] = v2pz;
        out[2][i] = vsv[i] * vsv[i];
        out[3][i] = v2pz * (1.0 + 2.0 * delta[i]);
        /* angles in turns, reduced to [-1/2, 1/2] in FP64 and rounded once */
        const double tt = (double)theta[i] / (2.0 * pi), tp = (double)phi[i] / (2.0 * pi);
        out[4][i] = (float)(tt - nearbyint(tt));
        out[5][i] = (float)(tp - nearbyint(tp));
    }
}

void fl_interleave32(int64_t n, int nk, const float *const *src, float *dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < nk; k++) dst[i * nk + k] = src[k][i];
}

void fl_interleave16(int64_t n, int nk, const uint16_t *const *src, uint16_t *dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < nk; k++) dst[i * nk + k] = src[k][i];
}
