/*
 * Host model setup for the Fletcher TTI path: the once-per-run work of the
 * reference front-end (model arrays, random boundary, CFL, Ricker source,
 * coefficient precompute), restated with the reference's float/double
 * promotion order so the FP32 results are bitwise those of the reference.
 * Reference citations are per function.
 */
#include "fletcher/fl_model.h"

#include <math.h>

#include "fl_cell.h"
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

void fl_boundary_planes(fl_rand_t *rs, int sx, int sy, int nx, int ny, int nz, int bord, int absorb,
                        float maxP, float maxS, const float *elo_vpz, const float *elo_vsv,
                        const float *ehi_vpz, const float *ehi_vsv, int z0, int nzs, float *vpz, float *vsv)
{
    /* boundary.c:26-98 for planes z0 .. z0+nzs-1, in the reference's iz -> iy -> ix order */
#define PIX(ix, iy) ((int64_t)(iy) * sx + (ix))
    const int bordLen = bord + absorb - 1; /* last index of the low absorbing zone */
    const int firstIn = bordLen + 1;       /* first interior index */
    const float frac = 1.0 / (float)(absorb);
    const int64_t plane = (int64_t)sx * sy;
    for (int iz = z0; iz < z0 + nzs; iz++) {
        float *pv = vpz + (iz - z0) * plane, *ps = vsv + (iz - z0) * plane;
        /* the plane that holds the interior cell j of this plane's absorbing cells */
        const float *jv = iz < firstIn ? elo_vpz : iz > bordLen + nz ? ehi_vpz : pv;
        const float *js = iz < firstIn ? elo_vsv : iz > bordLen + nz ? ehi_vsv : ps;
        for (int iy = 0; iy < sy; iy++)
            for (int ix = 0; ix < sx; ix++) {
                const int64_t i = PIX(ix, iy);
                if ((iz >= firstIn && iz <= bordLen + nz) && (iy >= firstIn && iy <= bordLen + ny) &&
                    (ix >= firstIn && ix <= bordLen + nx))
                    continue; /* interior untouched */
                if ((iz >= bord && iz <= 2 * bordLen + nz) && (iy >= bord && iy <= 2 * bordLen + ny) &&
                    (ix >= bord && ix <= 2 * bordLen + nx)) {
                    int distz, disty, distx, ively, ivelx;
                    if (iz > bordLen + nz) distz = iz - bordLen - nz;
                    else if (iz < firstIn) distz = firstIn - iz;
                    else distz = 0;
                    if (iy > bordLen + ny) { disty = iy - bordLen - ny; ively = bordLen + ny; }
                    else if (iy < firstIn) { disty = firstIn - iy; ively = firstIn; }
                    else { disty = 0; ively = iy; }
                    if (ix > bordLen + nx) { distx = ix - bordLen - nx; ivelx = bordLen + nx; }
                    else if (ix < firstIn) { distx = firstIn - ix; ivelx = firstIn; }
                    else { distx = 0; ivelx = ix; }
                    int dist = (disty > distz) ? disty : distz;
                    dist = (dist > distx) ? dist : distx;
                    const float bordDist = (float)(dist)*frac;
                    const float rfac = (float)fl_rand_next(rs) / (float)2147483647;
                    const int64_t j = PIX(ivelx, ively);
                    pv[i] = jv[j] * (1.0 - bordDist) + maxP * rfac * bordDist;
                    ps[i] = js[j] * (1.0 - bordDist) + maxS * rfac * bordDist;
                } else {
                    pv[i] = 0.0;
                    ps[i] = 0.0;
                }
            }
    }
#undef PIX
}

void fl_boundary_max(int sx, int sy, int nx, int ny, int bord, int absorb, int z0, int nzs,
                     const float *vpz, const float *vsv, float *maxP, float *maxS)
{
    /* boundary.c:12-24: the interior maxima, over the interior cells of planes z0 .. z0+nzs-1 */
    const int b = bord + absorb;
    const int64_t plane = (int64_t)sx * sy;
    for (int iz = z0; iz < z0 + nzs; iz++)
        for (int iy = b; iy < ny + b; iy++)
            for (int64_t i = (iz - z0) * plane + (int64_t)iy * sx + b; i < (iz - z0) * plane + (int64_t)iy * sx + nx + b;
                 i++) {
                *maxP = fmaxf(vpz[i], *maxP);
                *maxS = fmaxf(vsv[i], *maxS);
            }
}

void fl_random_boundary(uint32_t seed, int sx, int sy, int sz, int nx, int ny, int nz, int bord,
                        int absorb, float *vpz, float *vsv)
{
    /* boundary.c:7-98 */
    fl_rand_t rs;
    fl_rand_seed(&rs, seed);
    const int b = bord + absorb;
    const int64_t plane = (int64_t)sx * sy;
    float maxP = 0.0, maxS = 0.0;
    fl_boundary_max(sx, sy, nx, ny, bord, absorb, b, nz, vpz + b * plane, vsv + b * plane, &maxP, &maxS);
    const int64_t lo = (int64_t)b * plane, hi = (int64_t)(b + nz - 1) * plane;
    fl_boundary_planes(&rs, sx, sy, nx, ny, nz, bord, absorb, maxP, maxS, vpz + lo, vsv + lo, vpz + hi, vsv + hi,
                       0, sz, vpz, vsv);
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
    /* precomp.h:25-38 and :53-58 (fl_cell_f32) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float v[10];
        fl_cell_f32(vpz[i], vsv[i], epsilon[i], delta[i], phi[i], theta[i], v);
        for (int k = 0; k < 10; k++) c[k][i] = v[k];
    }
}

void fl_precompute_f64(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                       const float *delta, const float *phi, const float *theta, double *const *c)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double v[10];
        fl_cell_f64(vpz[i], vsv[i], epsilon[i], delta[i], phi[i], theta[i], v);
        for (int k = 0; k < 10; k++) c[k][i] = v[k];
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

uint16_t fl_f64_to_f16(double v) { return fl_cell_to_f16(v); }

uint16_t fl_f64_to_bf16(double v) { return fl_cell_to_bf16(v); }

uint16_t fl_angle_u16(double turns) { return fl_cell_angle_u16(turns); }

void fl_pack_coefs16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                     const float *delta, const float *phi, const float *theta, double dt, int bf16,
                     uint16_t *const *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint16_t v[10];
        fl_cell_pack16(vpz[i], vsv[i], epsilon[i], delta[i], phi[i], theta[i], dt, bf16, v);
        for (int k = 0; k < 10; k++) out[k][i] = v[k];
    }
}

void fl_pack_recompute16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                         const float *delta, const float *phi, const float *theta, double dt,
                         uint16_t *const *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint16_t v[6];
        fl_cell_recompute16(vpz[i], vsv[i], epsilon[i], delta[i], phi[i], theta[i], dt, v);
        for (int k = 0; k < 6; k++) out[k][i] = v[k];
    }
}

int fl_recompute_q16_base(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                          const float *delta, double dt)
{
    /* the exponent of the cell that needs the largest one, less 15 */
    double m = 0.0;
    const double dt2 = dt * dt;
#pragma omp parallel for schedule(static) reduction(max : m)
    for (int64_t i = 0; i < n; i++) {
        const double c = fl_cell_q16_need(vpz[i], vsv[i], epsilon[i], delta[i], dt2);
        if (c > m) m = c;
    }
    return fl_cell_q16_base_of(m);
}

void fl_pack_recompute_q16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                           const float *delta, const float *phi, const float *theta, double dt,
                           int ebase, uint16_t *const *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint16_t v[6];
        fl_cell_recompute_q16(vpz[i], vsv[i], epsilon[i], delta[i], phi[i], theta[i], dt, ebase, v);
        for (int k = 0; k < 6; k++) out[k][i] = v[k];
    }
}

void fl_pack_recompute32(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                         const float *delta, const float *phi, const float *theta, float *const *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float v[6];
        fl_cell_recompute32(vpz[i], vsv[i], epsilon[i], delta[i], phi[i], theta[i], v);
        for (int k = 0; k < 6; k++) out[k][i] = v[k];
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
