/*
 * Seeded heterogeneous TTI models (no reference counterpart: the reference
 * front-end only builds homogeneous models, main.c:168-221).  Smooth random
 * fields are sums of separable Gaussians normalised to [0,1]; the random
 * generator is splitmix64, so a (seed, shape) pair fixes every value.  After
 * the model fill, the reference random-velocity boundary is applied
 * (fl_random_boundary, boundary.c:7-98) and vsv follows main.c:218.
 */
#include "fletcher/fl_model.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static double unif(uint64_t *s, double a, double b)
{
    return a + (b - a) * ((double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0));
}

/* Separable Gaussian sum: f(x,y,z) = sum_k a_k gx_k(x) gy_k(y) gz_k(z). */
typedef struct {
    int ng, sx, sy, sz;
    float *gx, *gy, *gz, *amp; /* ng*sx, ng*sy, ng*sz, ng */
} gsum_t;

static int gsum_init(gsum_t *g, uint64_t *rng, int ng, int sx, int sy, int sz, double smin,
                     double smax, double amin, double amax)
{
    g->ng = ng; g->sx = sx; g->sy = sy; g->sz = sz;
    g->gx = (float *)malloc(sizeof(float) * (size_t)ng * sx);
    g->gy = (float *)malloc(sizeof(float) * (size_t)ng * sy);
    g->gz = (float *)malloc(sizeof(float) * (size_t)ng * sz);
    g->amp = (float *)malloc(sizeof(float) * (size_t)ng);
    if (!g->gx || !g->gy || !g->gz || !g->amp) return -1;
    for (int k = 0; k < ng; k++) {
        const double cx = unif(rng, 0, sx), cy = unif(rng, 0, sy), cz = unif(rng, 0, sz);
        const double s = unif(rng, smin, smax);
        g->amp[k] = (float)unif(rng, amin, amax);
        const double inv = 1.0 / (2.0 * s * s);
        for (int i = 0; i < sx; i++) g->gx[(size_t)k * sx + i] = (float)exp(-(i - cx) * (i - cx) * inv);
        for (int i = 0; i < sy; i++) g->gy[(size_t)k * sy + i] = (float)exp(-(i - cy) * (i - cy) * inv);
        for (int i = 0; i < sz; i++) g->gz[(size_t)k * sz + i] = (float)exp(-(i - cz) * (i - cz) * inv);
    }
    return 0;
}

static void gsum_free(gsum_t *g)
{
    free(g->gx); free(g->gy); free(g->gz); free(g->amp);
}

/* out = sum of Gaussians, then (if normalise) min-max scaled into [0,1]. */
static void gsum_eval(const gsum_t *g, float *out, int normalise)
{
    const int sx = g->sx, sy = g->sy, sz = g->sz, ng = g->ng;
    double vmin = 1e300, vmax = -1e300;
#pragma omp parallel for schedule(static) reduction(min : vmin) reduction(max : vmax)
    for (int iz = 0; iz < sz; iz++) {
        float *row = (float *)malloc(sizeof(float) * sx);
        for (int iy = 0; iy < sy; iy++) {
            for (int ix = 0; ix < sx; ix++) row[ix] = 0.0f;
            for (int k = 0; k < ng; k++) {
                const float w = g->amp[k] * g->gz[(size_t)k * sz + iz] * g->gy[(size_t)k * sy + iy];
                const float *gx = g->gx + (size_t)k * sx;
                for (int ix = 0; ix < sx; ix++) row[ix] += w * gx[ix];
            }
            float *o = out + ((int64_t)iz * sy + iy) * sx;
            for (int ix = 0; ix < sx; ix++) {
                o[ix] = row[ix];
                if (row[ix] < vmin) vmin = row[ix];
                if (row[ix] > vmax) vmax = row[ix];
            }
        }
        free(row);
    }
    if (!normalise) return;
    const double span = vmax > vmin ? vmax - vmin : 1.0;
    const int64_t n = (int64_t)sx * sy * sz;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = (float)((out[i] - vmin) / span);
}

static int smooth_field(uint64_t *rng, int ng, int sx, int sy, int sz, double smin, double smax,
                        float *out)
{
    gsum_t g;
    if (gsum_init(&g, rng, ng, sx, sy, sz, smin, smax, -1.0, 1.0)) { gsum_free(&g); return -1; }
    gsum_eval(&g, out, 1);
    gsum_free(&g);
    return 0;
}

static void finish_vsv(int64_t n, const float *vpz, const float *eps, const float *delta, float *vsv)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) vsv[i] = vpz[i] * sqrtf(fabsf(eps[i] - delta[i]) / 0.75);
}

void fl_gradient_model_defaults(fl_gradient_model_t *m, uint64_t seed)
{
    m->seed = seed;
    m->vmin = 1500.0f; m->vmax = 4500.0f;
    m->nblobs = 24; m->blob_amp = 0.1f; m->blob_smin = 12.0f; m->blob_smax = 40.0f;
    m->eps_max = 0.3f; m->delta_max = 0.15f;
    m->theta_max = (float)(M_PI / 3.0); m->phi_max = (float)(2.0 * M_PI);
    m->nsmooth = 16; m->smooth_smin = 24.0f; m->smooth_smax = 96.0f;
    m->boundary_seed = 1;
}

int fl_model_gradient(const fl_gradient_model_t *m, int sx, int sy, int sz, int nx, int ny, int nz,
                      int bord, int absorb, float *vpz, float *vsv, float *epsilon, float *delta,
                      float *phi, float *theta)
{
    uint64_t rng = m->seed;
    const int64_t n = (int64_t)sx * sy * sz;
    const int z0 = bord + absorb;
    /* blobs: perturbation factor in [1-amp, 1+amp] */
    gsum_t g;
    if (gsum_init(&g, &rng, m->nblobs, sx, sy, sz, m->blob_smin, m->blob_smax, -m->blob_amp,
                  m->blob_amp)) { gsum_free(&g); return -1; }
    gsum_eval(&g, vpz, 0);
    gsum_free(&g);
    const double den = nz > 1 ? (double)(nz - 1) : 1.0;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < sz; iz++) {
        double t = (iz - z0) / den;
        t = t < 0 ? 0 : (t > 1 ? 1 : t);
        const double vz = m->vmin + (m->vmax - m->vmin) * t;
        float *p = vpz + (int64_t)iz * sy * sx;
        for (int64_t k = 0; k < (int64_t)sy * sx; k++) {
            double f = 1.0 + p[k];
            if (f < 1.0 - m->blob_amp) f = 1.0 - m->blob_amp;
            if (f > 1.0 + m->blob_amp) f = 1.0 + m->blob_amp;
            p[k] = (float)(vz * f);
        }
    }
    if (smooth_field(&rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, epsilon)) return -1;
    if (smooth_field(&rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, delta)) return -1;
    if (smooth_field(&rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, theta)) return -1;
    if (smooth_field(&rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, phi)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const float e = (float)(m->eps_max * epsilon[i]);
        const float dcap = e < m->delta_max ? e : m->delta_max;
        epsilon[i] = e;
        delta[i] = (float)(dcap * delta[i]);
        theta[i] = (float)(m->theta_max * theta[i]);
        float ph = (float)(m->phi_max * phi[i]);
        if (ph >= m->phi_max) ph = nextafterf(m->phi_max, 0.0f);
        phi[i] = ph;
    }
    finish_vsv(n, vpz, epsilon, delta, vsv);
    fl_random_boundary(m->boundary_seed, sx, sy, sz, nx, ny, nz, bord, absorb, vpz, vsv);
    return 0;
}

void fl_layered_model_defaults(fl_layered_model_t *m, uint64_t seed)
{
    m->seed = seed;
    m->nlayers = 12;
    m->vmin = 1500.0f; m->vmax = 5000.0f; m->max_dip_deg = 15.0f;
    m->eps_max = 0.3f; m->delta_max = 0.2f;
    m->theta_max = (float)(M_PI / 3.0); m->phi_max = (float)(2.0 * M_PI);
    m->nsmooth = 16; m->smooth_smin = 48.0f; m->smooth_smax = 192.0f;
    m->boundary_seed = 1;
}

int fl_model_layered(const fl_layered_model_t *m, int sx, int sy, int sz, int nx, int ny, int nz,
                     int bord, int absorb, float *vpz, float *vsv, float *epsilon, float *delta,
                     float *phi, float *theta)
{
    uint64_t rng = m->seed;
    const int L = m->nlayers < 1 ? 1 : (m->nlayers > 64 ? 64 : m->nlayers);
    const int64_t n = (int64_t)sx * sy * sz;
    const int z0 = bord + absorb;
    double depth[64], gx[64], gy[64], lv[64], le[64], ld[64];
    /* interface k (k=1..L-1) at depth z0 + k*nz/L, dipping along a random azimuth */
    for (int k = 0; k < L; k++) {
        const double dip = unif(&rng, 0.0, m->max_dip_deg) * M_PI / 180.0;
        const double az = unif(&rng, 0.0, 2.0 * M_PI);
        depth[k] = z0 + (double)k * nz / L + unif(&rng, -0.2, 0.2) * nz / L;
        gx[k] = tan(dip) * cos(az);
        gy[k] = tan(dip) * sin(az);
        const double frac = L > 1 ? (double)k / (L - 1) : 0.0;
        lv[k] = m->vmin + (m->vmax - m->vmin) * frac * unif(&rng, 0.9, 1.0);
        le[k] = unif(&rng, 0.0, m->eps_max);
        const double dmax = le[k] < m->delta_max ? le[k] : m->delta_max;
        ld[k] = unif(&rng, 0.0, dmax);
    }
    if (lv[0] < m->vmin) lv[0] = m->vmin;
    const double xc = sx * 0.5, yc = sy * 0.5;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < sz; iz++)
        for (int iy = 0; iy < sy; iy++)
            for (int ix = 0; ix < sx; ix++) {
                int layer = 0;
                for (int k = 1; k < L; k++) {
                    const double zk = depth[k] + gx[k] * (ix - xc) + gy[k] * (iy - yc);
                    if (iz >= zk) layer = k;
                }
                const int64_t i = ((int64_t)iz * sy + iy) * sx + ix;
                vpz[i] = (float)lv[layer];
                epsilon[i] = (float)le[layer];
                delta[i] = (float)ld[layer];
            }
    if (smooth_field(&rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, theta)) return -1;
    if (smooth_field(&rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, phi)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        theta[i] = (float)(m->theta_max * theta[i]);
        float ph = (float)(m->phi_max * phi[i]);
        if (ph >= m->phi_max) ph = nextafterf(m->phi_max, 0.0f);
        phi[i] = ph;
    }
    finish_vsv(n, vpz, epsilon, delta, vsv);
    fl_random_boundary(m->boundary_seed, sx, sy, sz, nx, ny, nz, bord, absorb, vpz, vsv);
    return 0;
}
