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

/* Plane iz of the Gaussian sum into out (sy*sx floats), not normalised. */
static void gsum_plane(const gsum_t *g, int iz, float *out)
{
    const int sx = g->sx, sy = g->sy, sz = g->sz, ng = g->ng;
    for (int iy = 0; iy < sy; iy++) {
        float *row = out + (int64_t)iy * sx;
        for (int ix = 0; ix < sx; ix++) row[ix] = 0.0f;
        for (int k = 0; k < ng; k++) {
            const float w = g->amp[k] * g->gz[(size_t)k * sz + iz] * g->gy[(size_t)k * sy + iy];
            const float *gx = g->gx + (size_t)k * sx;
            for (int ix = 0; ix < sx; ix++) row[ix] += w * gx[ix];
        }
    }
}

/* Minimum and maximum of the Gaussian sum over the whole volume. */
static int gsum_minmax(const gsum_t *g, double *vmin_out, double *vmax_out)
{
    const int sx = g->sx, sy = g->sy, sz = g->sz;
    double vmin = 1e300, vmax = -1e300;
    int bad = 0;
#pragma omp parallel reduction(min : vmin) reduction(max : vmax) reduction(| : bad)
    {
        float *pl = (float *)malloc(sizeof(float) * (size_t)sx * sy);
        if (!pl) bad = 1;
#pragma omp for schedule(static)
        for (int iz = 0; iz < sz; iz++) {
            if (!pl) continue;
            gsum_plane(g, iz, pl);
            for (int64_t i = 0; i < (int64_t)sx * sy; i++) {
                if (pl[i] < vmin) vmin = pl[i];
                if (pl[i] > vmax) vmax = pl[i];
            }
        }
        free(pl);
    }
    *vmin_out = vmin;
    *vmax_out = vmax;
    return bad ? -1 : 0;
}

/* Plane iz min-max scaled into [0,1] by the volume's range. */
static void gsum_plane_unit(const gsum_t *g, int iz, double vmin, double vmax, float *out)
{
    gsum_plane(g, iz, out);
    const double span = vmax > vmin ? vmax - vmin : 1.0;
    for (int64_t i = 0; i < (int64_t)g->sx * g->sy; i++) out[i] = (float)((out[i] - vmin) / span);
}

static void finish_vsv(int64_t n, const float *vpz, const float *eps, const float *delta, float *vsv)
{
    for (int64_t i = 0; i < n; i++) vsv[i] = vpz[i] * sqrtf(fabsf(eps[i] - delta[i]) / 0.75);
}

/* ---------------------------------------------------------------- generators by planes
 * A generator holds everything drawn from the seed (Gaussians, layers) and the ranges of its
 * smooth fields, and fills any range of z planes of the six arrays, before the boundary.  The
 * whole-volume generators and the streamed ones below share it. */
enum { GEN_GRADIENT = 0, GEN_LAYERED = 1 };
typedef struct {
    int kind, sx, sy, sz, nx, ny, nz, bord, absorb;
    fl_gradient_model_t gm;
    fl_layered_model_t lm;
    gsum_t g[5];      /* gradient: blobs, eps, delta, theta, phi; layered: theta, phi */
    int ng;
    double lo[5], hi[5];  /* volume range of each smooth field */
    int L;
    double depth[64], gx[64], gy[64], lv[64], le[64], ld[64];
} gen_t;

static void gen_free(gen_t *G)
{
    for (int k = 0; k < G->ng; k++) gsum_free(&G->g[k]);
    G->ng = 0;
}

static int gen_gradient(gen_t *G, const fl_gradient_model_t *m, int sx, int sy, int sz, int nx, int ny, int nz,
                        int bord, int absorb)
{
    memset(G, 0, sizeof(*G));
    G->kind = GEN_GRADIENT;
    G->sx = sx; G->sy = sy; G->sz = sz; G->nx = nx; G->ny = ny; G->nz = nz; G->bord = bord; G->absorb = absorb;
    G->gm = *m;
    uint64_t rng = m->seed;
    /* the draws in the order of the seed's stream: blobs, then eps, delta, theta, phi */
    G->ng = 1;
    if (gsum_init(&G->g[0], &rng, m->nblobs, sx, sy, sz, m->blob_smin, m->blob_smax, -m->blob_amp, m->blob_amp))
        return -1;
    for (int k = 1; k < 5; k++) {
        G->ng = k + 1;
        if (gsum_init(&G->g[k], &rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, -1.0, 1.0) ||
            gsum_minmax(&G->g[k], &G->lo[k], &G->hi[k]))
            return -1;
    }
    return 0;
}

static int gen_layered(gen_t *G, const fl_layered_model_t *m, int sx, int sy, int sz, int nx, int ny, int nz,
                       int bord, int absorb)
{
    memset(G, 0, sizeof(*G));
    G->kind = GEN_LAYERED;
    G->sx = sx; G->sy = sy; G->sz = sz; G->nx = nx; G->ny = ny; G->nz = nz; G->bord = bord; G->absorb = absorb;
    G->lm = *m;
    uint64_t rng = m->seed;
    const int L = m->nlayers < 1 ? 1 : (m->nlayers > 64 ? 64 : m->nlayers);
    G->L = L;
    const int z0 = bord + absorb;
    /* interface k (k=1..L-1) at depth z0 + k*nz/L, dipping along a random azimuth */
    for (int k = 0; k < L; k++) {
        const double dip = unif(&rng, 0.0, m->max_dip_deg) * M_PI / 180.0;
        const double az = unif(&rng, 0.0, 2.0 * M_PI);
        G->depth[k] = z0 + (double)k * nz / L + unif(&rng, -0.2, 0.2) * nz / L;
        G->gx[k] = tan(dip) * cos(az);
        G->gy[k] = tan(dip) * sin(az);
        const double frac = L > 1 ? (double)k / (L - 1) : 0.0;
        G->lv[k] = m->vmin + (m->vmax - m->vmin) * frac * unif(&rng, 0.9, 1.0);
        G->le[k] = unif(&rng, 0.0, m->eps_max);
        const double dmax = G->le[k] < m->delta_max ? G->le[k] : m->delta_max;
        G->ld[k] = unif(&rng, 0.0, dmax);
    }
    if (G->lv[0] < m->vmin) G->lv[0] = m->vmin;
    for (int k = 0; k < 2; k++) {
        G->ng = k + 1;
        if (gsum_init(&G->g[k], &rng, m->nsmooth, sx, sy, sz, m->smooth_smin, m->smooth_smax, -1.0, 1.0) ||
            gsum_minmax(&G->g[k], &G->lo[k], &G->hi[k]))
            return -1;
    }
    return 0;
}

/* Planes z0 .. z0+nzs-1 of the model before the boundary, into slab arrays of nzs planes. */
static void gen_planes(const gen_t *G, int z0, int nzs, float *vpz, float *vsv, float *epsilon, float *delta,
                       float *phi, float *theta)
{
    const int sx = G->sx, sy = G->sy;
    const int64_t plane = (int64_t)sx * sy;
#pragma omp parallel for schedule(static)
    for (int iz = z0; iz < z0 + nzs; iz++) {
        const int64_t o = (int64_t)(iz - z0) * plane;
        float *p = vpz + o, *e = epsilon + o, *d = delta + o, *t = theta + o, *f = phi + o;
        if (G->kind == GEN_GRADIENT) {
            const fl_gradient_model_t *m = &G->gm;
            gsum_plane(&G->g[0], iz, p);
            const double den = G->nz > 1 ? (double)(G->nz - 1) : 1.0;
            double tt = (iz - (G->bord + G->absorb)) / den;
            tt = tt < 0 ? 0 : (tt > 1 ? 1 : tt);
            const double vz = m->vmin + (m->vmax - m->vmin) * tt;
            for (int64_t k = 0; k < plane; k++) {
                double fct = 1.0 + p[k];
                if (fct < 1.0 - m->blob_amp) fct = 1.0 - m->blob_amp;
                if (fct > 1.0 + m->blob_amp) fct = 1.0 + m->blob_amp;
                p[k] = (float)(vz * fct);
            }
            gsum_plane_unit(&G->g[1], iz, G->lo[1], G->hi[1], e);
            gsum_plane_unit(&G->g[2], iz, G->lo[2], G->hi[2], d);
            gsum_plane_unit(&G->g[3], iz, G->lo[3], G->hi[3], t);
            gsum_plane_unit(&G->g[4], iz, G->lo[4], G->hi[4], f);
            for (int64_t k = 0; k < plane; k++) {
                const float ev = (float)(m->eps_max * e[k]);
                const float dcap = ev < m->delta_max ? ev : m->delta_max;
                e[k] = ev;
                d[k] = (float)(dcap * d[k]);
                t[k] = (float)(m->theta_max * t[k]);
                float ph = (float)(m->phi_max * f[k]);
                if (ph >= m->phi_max) ph = nextafterf(m->phi_max, 0.0f);
                f[k] = ph;
            }
        } else {
            const fl_layered_model_t *m = &G->lm;
            const double xc = sx * 0.5, yc = sy * 0.5;
            for (int iy = 0; iy < sy; iy++)
                for (int ix = 0; ix < sx; ix++) {
                    int layer = 0;
                    for (int k = 1; k < G->L; k++) {
                        const double zk = G->depth[k] + G->gx[k] * (ix - xc) + G->gy[k] * (iy - yc);
                        if (iz >= zk) layer = k;
                    }
                    const int64_t i = (int64_t)iy * sx + ix;
                    p[i] = (float)G->lv[layer];
                    e[i] = (float)G->le[layer];
                    d[i] = (float)G->ld[layer];
                }
            gsum_plane_unit(&G->g[0], iz, G->lo[0], G->hi[0], t);
            gsum_plane_unit(&G->g[1], iz, G->lo[1], G->hi[1], f);
            for (int64_t k = 0; k < plane; k++) {
                t[k] = (float)(m->theta_max * t[k]);
                float ph = (float)(m->phi_max * f[k]);
                if (ph >= m->phi_max) ph = nextafterf(m->phi_max, 0.0f);
                f[k] = ph;
            }
        }
        finish_vsv(plane, p, e, d, vsv + o);
    }
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
    gen_t G;
    if (gen_gradient(&G, m, sx, sy, sz, nx, ny, nz, bord, absorb)) {
        gen_free(&G);
        return -1;
    }
    gen_planes(&G, 0, sz, vpz, vsv, epsilon, delta, phi, theta);
    gen_free(&G);
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
    gen_t G;
    if (gen_layered(&G, m, sx, sy, sz, nx, ny, nz, bord, absorb)) {
        gen_free(&G);
        return -1;
    }
    gen_planes(&G, 0, sz, vpz, vsv, epsilon, delta, phi, theta);
    gen_free(&G);
    fl_random_boundary(m->boundary_seed, sx, sy, sz, nx, ny, nz, bord, absorb, vpz, vsv);
    return 0;
}

/* ---------------------------------------------------------------- streamed generation */
struct fl_model_stream {
    gen_t G;
    uint32_t boundary_seed;
    float maxP, maxS;         /* interior maxima of vpz, vsv before the boundary */
    float *edge[4];           /* vpz, vsv of the planes bord+absorb and bord+absorb+nz-1 */
    double maxvel;            /* max vpz*sqrt(1+2 eps) over the whole array after the boundary */
    fl_rand_t rs;
    int next_z;               /* first plane of the next fill */
};

void fl_model_stream_close(fl_model_stream_t *s)
{
    if (!s) return;
    gen_free(&s->G);
    for (int k = 0; k < 4; k++) free(s->edge[k]);
    free(s);
}

/* Planes of one internal pass: about 256 MiB of the six arrays. */
static int stream_slab(const gen_t *G)
{
    const int64_t plane = (int64_t)G->sx * G->sy;
    const int64_t k = (256ll << 20) / (24 * plane);
    return (int)(k < 1 ? 1 : (k > G->sz ? G->sz : k));
}

int fl_model_stream_fill(void *user, int z0, int nzs, float *vpz, float *vsv, float *epsilon, float *delta,
                         float *phi, float *theta)
{
    fl_model_stream_t *s = (fl_model_stream_t *)user;
    if (!s) return -1;
    const gen_t *G = &s->G;
    if (z0 != s->next_z || nzs < 1 || z0 + nzs > G->sz) return -1;
    gen_planes(G, z0, nzs, vpz, vsv, epsilon, delta, phi, theta);
    fl_boundary_planes(&s->rs, G->sx, G->sy, G->nx, G->ny, G->nz, G->bord, G->absorb, s->maxP, s->maxS, s->edge[0],
                       s->edge[1], s->edge[2], s->edge[3], z0, nzs, vpz, vsv);
    s->next_z = z0 + nzs;
    if (s->next_z == G->sz) { /* a complete pass: the next one starts again from plane 0 */
        fl_rand_seed(&s->rs, s->boundary_seed);
        s->next_z = 0;
    }
    return 0;
}

static fl_model_stream_t *stream_open(fl_model_stream_t *s)
{
    gen_t *G = &s->G;
    const int64_t plane = (int64_t)G->sx * G->sy;
    const int slab = stream_slab(G);
    float *buf = (float *)malloc(sizeof(float) * 6 * (size_t)slab * plane);
    for (int k = 0; k < 4; k++) s->edge[k] = (float *)malloc(sizeof(float) * (size_t)plane);
    if (!buf || !s->edge[0] || !s->edge[1] || !s->edge[2] || !s->edge[3]) {
        free(buf);
        fl_model_stream_close(s);
        return NULL;
    }
    float *a[6];
    for (int k = 0; k < 6; k++) a[k] = buf + (size_t)k * slab * plane;
    /* pass 1: interior maxima and the two edge planes, before the boundary */
    const int b = G->bord + G->absorb, lo = b, hi = b + G->nz - 1;
    s->maxP = 0.0f;
    s->maxS = 0.0f;
    for (int z0 = b; z0 <= hi; z0 += slab) {
        const int n = z0 + slab <= hi + 1 ? slab : hi + 1 - z0;
        gen_planes(G, z0, n, a[0], a[1], a[2], a[3], a[4], a[5]);
        fl_boundary_max(G->sx, G->sy, G->nx, G->ny, G->bord, G->absorb, z0, n, a[0], a[1], &s->maxP, &s->maxS);
        if (lo >= z0 && lo < z0 + n) {
            memcpy(s->edge[0], a[0] + (size_t)(lo - z0) * plane, sizeof(float) * plane);
            memcpy(s->edge[1], a[1] + (size_t)(lo - z0) * plane, sizeof(float) * plane);
        }
        if (hi >= z0 && hi < z0 + n) {
            memcpy(s->edge[2], a[0] + (size_t)(hi - z0) * plane, sizeof(float) * plane);
            memcpy(s->edge[3], a[1] + (size_t)(hi - z0) * plane, sizeof(float) * plane);
        }
    }
    /* pass 2: the finished model once, for its CFL velocity (as fl_cfl_dt over the array) */
    fl_rand_seed(&s->rs, s->boundary_seed);
    s->next_z = 0;
    s->maxvel = 0.0;
    for (int z0 = 0; z0 < G->sz; z0 += slab) {
        const int n = z0 + slab <= G->sz ? slab : G->sz - z0;
        fl_model_stream_fill(s, z0, n, a[0], a[1], a[2], a[3], a[4], a[5]);
        for (int64_t i = 0; i < (int64_t)n * plane; i++) {
            const double v = a[0][i] * sqrt(1.0 + 2.0 * a[2][i]);
            if (v > s->maxvel) s->maxvel = v;
        }
    }
    free(buf);
    return s;
}

fl_model_stream_t *fl_model_stream_gradient(const fl_gradient_model_t *m, int sx, int sy, int sz, int nx, int ny,
                                            int nz, int bord, int absorb)
{
    fl_model_stream_t *s = (fl_model_stream_t *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    if (gen_gradient(&s->G, m, sx, sy, sz, nx, ny, nz, bord, absorb)) {
        fl_model_stream_close(s);
        return NULL;
    }
    s->boundary_seed = m->boundary_seed;
    return stream_open(s);
}

fl_model_stream_t *fl_model_stream_layered(const fl_layered_model_t *m, int sx, int sy, int sz, int nx, int ny,
                                           int nz, int bord, int absorb)
{
    fl_model_stream_t *s = (fl_model_stream_t *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    if (gen_layered(&s->G, m, sx, sy, sz, nx, ny, nz, bord, absorb)) {
        fl_model_stream_close(s);
        return NULL;
    }
    s->boundary_seed = m->boundary_seed;
    return stream_open(s);
}

float fl_model_stream_cfl_dt(const fl_model_stream_t *s, float h, float courant)
{
    return s->maxvel > 0.0 ? (float)(courant * h / s->maxvel) : 0.0f;
}
