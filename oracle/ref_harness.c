/*
 * ORACLE / REFERENCE HARNESS — test infrastructure only.
 *
 * Linked, by oracle/Makefile, against the UNMODIFIED reference C/OpenMP
 * sources of one variant directory (original/, der1der1/ or der1der1lm/):
 * main.c, model.c, boundary.c, source.c, utils.c, map.c, walltime.c and
 * OpenMP/openmp_*.c.  The result is oracle/_ref/libref_<variant>.so, which
 * is never committed.  main.c is compiled with -Dmain=ref_main and
 * -DModel=harness_Model, so running ref_main(argc, argv) executes the
 * reference front-end (argument parsing, model, CFL, RandomVelocityBoundary,
 * main.c:15-298) and hands every array to harness_Model below, which copies
 * the model, runs the reference Model() (model.c:32-164: precompute, time
 * loop over DRIVER_InsertSource / DRIVER_Propagate / SwapArrays) and copies
 * the final fields and the precomputed coefficient globals.
 *
 * A second entry set (ref_setup / ref_steps / ref_fetch) drives the
 * reference DRIVER_* back-end on caller-provided model arrays, timed like
 * model.c:82-88, for the CPU baseline and for parity on models the
 * reference front-end cannot build.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "utils.h"
#include "source.h"
#include "driver.h"
#include "model.h"
#include "walltime.h"

/* coefficient globals live in the reference model.c (MODEL_GLOBALVARS) */
extern float *ch1dxx, *ch1dyy, *ch1dzz, *ch1dxy, *ch1dyz, *ch1dxz, *v2px, *v2pz, *v2sz, *v2pn;

static float *cap_model[6];  /* vpz vsv epsilon delta phi theta after the boundary */
static float *cap_coef[10];  /* precomp order */
static float *cap_field[4];  /* pp pc qp qc at the end, pc/qc = current */
static long cap_n;
static int cap_dims[5];      /* sx sy sz st iSource */
static float cap_dt;
static int cap_run_model = 1;

static float *dup(const float *a, long n)
{
    float *r = (float *)malloc(sizeof(float) * n);
    memcpy(r, a, sizeof(float) * n);
    return r;
}

static void free_captures(void)
{
    for (int k = 0; k < 6; k++) { free(cap_model[k]); cap_model[k] = NULL; }
    for (int k = 0; k < 10; k++) { free(cap_coef[k]); cap_coef[k] = NULL; }
    for (int k = 0; k < 4; k++) { free(cap_field[k]); cap_field[k] = NULL; }
    cap_n = 0;
}

/* address of coefficient global k, precomp order */
static float **coef_global(int k)
{
    float **g[10] = {&ch1dxx, &ch1dyy, &ch1dzz, &ch1dxy, &ch1dyz,
                     &ch1dxz, &v2px,   &v2pz,   &v2sz,   &v2pn};
    return g[k];
}

void harness_Model(const int st, const int iSource, const float dtOutput, SlicePtr sPtr,
                   const int sx, const int sy, const int sz, const int bord,
                   const float dx, const float dy, const float dz, const float dt, const int it,
                   float *restrict pp, float *restrict pc, float *restrict qp, float *restrict qc,
                   float *restrict vpz, float *restrict vsv, float *restrict epsilon,
                   float *restrict delta, float *restrict phi, float *restrict theta)
{
    free_captures();
    const long n = (long)sx * sy * sz;
    cap_n = n;
    cap_dims[0] = sx; cap_dims[1] = sy; cap_dims[2] = sz; cap_dims[3] = st; cap_dims[4] = iSource;
    cap_dt = dt;
    cap_model[0] = dup(vpz, n); cap_model[1] = dup(vsv, n); cap_model[2] = dup(epsilon, n);
    cap_model[3] = dup(delta, n); cap_model[4] = dup(phi, n); cap_model[5] = dup(theta, n);
    if (!cap_run_model) return;
    Model(st, iSource, dtOutput, sPtr, sx, sy, sz, bord, dx, dy, dz, dt, it, pp, pc, qp, qc, vpz,
          vsv, epsilon, delta, phi, theta);
    /* Model swaps its local pointers st times: the current field is in pc's buffer
     * when st is even and in pp's buffer when st is odd. */
    const int odd = st & 1;
    cap_field[0] = dup(odd ? pc : pp, n);
    cap_field[1] = dup(odd ? pp : pc, n);
    cap_field[2] = dup(odd ? qc : qp, n);
    cap_field[3] = dup(odd ? qp : qc, n);
    for (int k = 0; k < 10; k++) {
        float **gp = coef_global(k);
        cap_coef[k] = *gp;  /* take ownership of the reference's malloc'd volume */
        *gp = NULL;
    }
}

/* glibc's initial rand() state is srand(1): restores a fresh process's boundary sequence. */
void ref_reset_rand(void) { srand(1); }

/* 1: ref_main runs the reference Model(); 0: stop after model + boundary. */
void ref_set_run_model(int flag) { cap_run_model = flag; }

void ref_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

long ref_capture_size(int *dims5, float *dt)
{
    memcpy(dims5, cap_dims, sizeof(cap_dims));
    *dt = cap_dt;
    return cap_n;
}

/* which: 0..5 model arrays, 10..19 coefficients, 20..23 pp pc qp qc. Returns 0 if present. */
int ref_capture_get(int which, float *out)
{
    float *src = NULL;
    if (which >= 0 && which < 6) src = cap_model[which];
    else if (which >= 10 && which < 20) src = cap_coef[which - 10];
    else if (which >= 20 && which < 24) src = cap_field[which - 20];
    if (!src) return -1;
    memcpy(out, src, sizeof(float) * cap_n);
    return 0;
}

/* ---------------- direct back-end driving (CPU baseline, arbitrary models) ---------------- */

static float *b_f[4];
static int b_dims[4];
static float b_d[4];

void ref_free(void)
{
    for (int k = 0; k < 4; k++) { free(b_f[k]); b_f[k] = NULL; }
    for (int k = 0; k < 10; k++) {
        float **gp = coef_global(k);
        free(*gp);
        *gp = NULL;
    }
    DRIVER_Finalize();
}

static void ref_precompute(const int sx, const int sy, const int sz, const int bord,
                           float *vpz, float *vsv, float *epsilon, float *delta, float *phi,
                           float *theta)
{
#define MODEL_INITIALIZE
#include "precomp.h"
#undef MODEL_INITIALIZE
}

/* Precompute with the reference precomp.h fragment and initialise the reference back-end. */
int ref_setup(int sx, int sy, int sz, int bord, float dx, float dy, float dz, float dt,
              float *vpz, float *vsv, float *epsilon, float *delta, float *phi, float *theta)
{
    ref_free();
    const long n = (long)sx * sy * sz;
    for (int k = 0; k < 4; k++) {
        b_f[k] = (float *)calloc(n, sizeof(float));
        if (!b_f[k]) return -1;
    }
    b_dims[0] = sx; b_dims[1] = sy; b_dims[2] = sz; b_dims[3] = bord;
    b_d[0] = dx; b_d[1] = dy; b_d[2] = dz; b_d[3] = dt;
    ref_precompute(sx, sy, sz, bord, vpz, vsv, epsilon, delta, phi, theta);
    DRIVER_Initialize(sx, sy, sz, bord, dx, dy, dz, dt, vpz, vsv, epsilon, delta, phi, theta,
                      b_f[0], b_f[1], b_f[2], b_f[3]);
    return 0;
}

/* Steps it0+1 .. it0+nsteps of model.c:71-88; returns the summed Propagate+Swap wall time. */
double ref_steps(int it0, int nsteps, int iSource)
{
    double walltime = 0.0;
    for (int it = it0 + 1; it <= it0 + nsteps; it++) {
        float src = Source(b_d[3], it - 1);
        DRIVER_InsertSource(b_d[3], it - 1, iSource, b_f[1], b_f[3], src);
        const double t0 = wtime();
        DRIVER_Propagate(b_dims[0], b_dims[1], b_dims[2], b_dims[3], b_d[0], b_d[1], b_d[2],
                         b_d[3], it, b_f[0], b_f[1], b_f[2], b_f[3]);
        SwapArrays(&b_f[0], &b_f[1], &b_f[2], &b_f[3]);
        walltime += wtime() - t0;
    }
    return walltime;
}

/* which: 0 pp, 1 pc, 2 qp, 3 qc (current fields after the last swap). */
int ref_fetch(int which, float *out)
{
    if (which < 0 || which > 3 || !b_f[which]) return -1;
    memcpy(out, b_f[which], sizeof(float) * (long)b_dims[0] * b_dims[1] * b_dims[2]);
    return 0;
}

/* Coefficient volume k (precomp order) after ref_setup. */
int ref_fetch_coef(int k, float *out)
{
    if (k < 0 || k > 9) return -1;
    float *src = *coef_global(k);
    if (!src) return -1;
    memcpy(out, src, sizeof(float) * (long)b_dims[0] * b_dims[1] * b_dims[2]);
    return 0;
}

/* The reference Source(), for closed-form checks of the oracle's restatement. */
float ref_source(float dt, int it) { return Source(dt, it); }

/* The reference RandomVelocityBoundary (boundary.c:7-98), after srand(seed). */
void RandomVelocityBoundary(int sx, int sy, int sz, int nx, int ny, int nz, int bord, int absorb,
                            float *vpz, float *vsv);
void ref_random_boundary(unsigned seed, int sx, int sy, int sz, int nx, int ny, int nz, int bord,
                         int absorb, float *vpz, float *vsv)
{
    srand(seed);
    RandomVelocityBoundary(sx, sy, sz, nx, ny, nz, bord, absorb, vpz, vsv);
}
