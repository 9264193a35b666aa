/*
 * ORACLE — test infrastructure only; parity unpinned for nothing listed here
 * (every function is checked in tests/ against the reference build, the
 * archived reference logs or closed forms).
 *
 * CPU restatement, in plain C + OpenMP, of the reference's TTI propagate path:
 *   original/derivatives.h:13-68   FD weights, Der1, Der2, DerCross
 *   original/sample.h:1-90         fused per-cell update (direct DerCross)
 *   der1der1/sample.h:1-123        first-of-first two-pass update
 *   original/OpenMP/openmp_propagate.c:9-34, der1der1/OpenMP/openmp_propagate.c:9-95  loop bounds
 *   original/OpenMP/openmp_insertsource.c:7-12, original/source.c:7-15  source
 *   original/precomp.h:14-69       coefficient precompute
 *   original/model.c:71-88         time loop
 *
 * Two instantiations: *_f32 keeps the reference's float arithmetic and
 * association order (built with -ffp-contract=off it is bitwise equal to the
 * reference OpenMP binary); *_f64 is the same code in double with exact
 * rational FD weights — the FP64 accuracy oracle.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------- FP32 instantiation: the reference's float literals ---------------- */
#define REAL float
#define SFX f32
#define L1 0.8f
#define L2 -0.2f
#define L3 0.0380952380952381f
#define L4 -0.0035714285714285713f
#define L11 0.64f
#define L12 -0.16f
#define L13 0.03047619047619047618f /* value of L1*L3 (the reference comment says L1*L2) */
#define L14 -0.00285714285714285713f
#define L22 0.04f
#define L23 -0.00761904761904761904f
#define L24 0.00071428571428571428f
#define L33 0.00145124716553287981f
#define L34 -0.00013605442176870748f
#define L44 0.00001275510204081632f
#define K0 -2.84722222222222222222f
#define K1 1.6f
#define K2 -0.2f
#define K3 0.02539682539682539682f
#define K4 -0.00178571428571428571f
#include "fl_oracle_body.h"
#undef REAL
#undef SFX
#undef L1
#undef L2
#undef L3
#undef L4
#undef L11
#undef L12
#undef L13
#undef L14
#undef L22
#undef L23
#undef L24
#undef L33
#undef L34
#undef L44
#undef K0
#undef K1
#undef K2
#undef K3
#undef K4

/* ---------------- FP64 instantiation: exact rational weights ---------------- */
#define REAL double
#define SFX f64
#define L1 (4.0 / 5.0)
#define L2 (-1.0 / 5.0)
#define L3 (4.0 / 105.0)
#define L4 (-1.0 / 280.0)
#define L11 (L1 * L1)
#define L12 (L1 * L2)
#define L13 (L1 * L3)
#define L14 (L1 * L4)
#define L22 (L2 * L2)
#define L23 (L2 * L3)
#define L24 (L2 * L4)
#define L33 (L3 * L3)
#define L34 (L3 * L4)
#define L44 (L4 * L4)
#define K0 (-205.0 / 72.0)
#define K1 (8.0 / 5.0)
#define K2 (-1.0 / 5.0)
#define K3 (8.0 / 315.0)
#define K4 (-1.0 / 560.0)
#include "fl_oracle_body.h"

/* ---------------- scalar helpers ---------------- */

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Ricker wavelet — original/source.c:7-15 with the constants of original/source.h:11-14.
 * Mixed float/double promotion follows the reference statement by statement. */
float oracle_source(float dt, int it)
{
    const double FCUT = 40.0, PICUBE = 31.00627668029982017537;
    const double TWOSQRTPI = 3.54490770181103205458, THREESQRTPI = 5.31736155271654808184;
    float tf, fc, fct, expo;
    tf = TWOSQRTPI / FCUT;
    fc = FCUT / THREESQRTPI;
    fct = fc * (((float)it) * dt - tf);
    expo = PICUBE * fct * fct;
    return ((1.0f - 2.0f * expo) * expf(-expo));
}

/* Coefficient precompute — original/precomp.h:25-38 (trig in double, products in float,
 * left to right) and :53-58 (velocity factors with double literals). */
void oracle_precompute_f32(long n, const float *vpz, const float *vsv, const float *epsilon,
                           const float *delta, const float *phi, const float *theta,
                           float *const *c)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        float sinTheta = sin(theta[i]);
        float cosTheta = cos(theta[i]);
        float sin2Theta = sin(2.0 * theta[i]);
        float sinPhi = sin(phi[i]);
        float cosPhi = cos(phi[i]);
        float sin2Phi = sin(2.0 * phi[i]);
        c[0][i] = sinTheta * sinTheta * cosPhi * cosPhi;
        c[1][i] = sinTheta * sinTheta * sinPhi * sinPhi;
        c[2][i] = cosTheta * cosTheta;
        c[3][i] = sinTheta * sinTheta * sin2Phi;
        c[4][i] = sin2Theta * sinPhi;
        c[5][i] = sin2Theta * cosPhi;
        c[8][i] = vsv[i] * vsv[i];
        c[7][i] = vpz[i] * vpz[i];
        c[6][i] = c[7][i] * (1.0 + 2.0 * epsilon[i]);
        c[9][i] = c[7][i] * (1.0 + 2.0 * delta[i]);
    }
}

/* Same formulas entirely in double from the float model arrays (FP64 oracle input). */
void oracle_precompute_f64(long n, const float *vpz, const float *vsv, const float *epsilon,
                           const float *delta, const float *phi, const float *theta,
                           double *const *c)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
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
        const double vp2 = (double)vpz[i] * vpz[i];
        c[8][i] = (double)vsv[i] * vsv[i];
        c[7][i] = vp2;
        c[6][i] = vp2 * (1.0 + 2.0 * epsilon[i]);
        c[9][i] = vp2 * (1.0 + 2.0 * delta[i]);
    }
}

/* Single first/second/cross derivative at one index, for closed-form tests. */
float oracle_der1_f32(const float *p, long i, long s, float dinv) { return der1_f32(p, i, s, dinv); }
float oracle_der2_f32(const float *p, long i, long s, float d2inv) { return der2_f32(p, i, s, d2inv); }
float oracle_dercross_f32(const float *p, long i, long s11, long s21, float dinv)
{
    return dercross_f32(p, i, s11, s21, dinv);
}
double oracle_der1_f64(const double *p, long i, long s, double dinv) { return der1_f64(p, i, s, dinv); }
double oracle_der2_f64(const double *p, long i, long s, double d2inv) { return der2_f64(p, i, s, d2inv); }
double oracle_dercross_f64(const double *p, long i, long s11, long s21, double dinv)
{
    return dercross_f64(p, i, s11, s21, dinv);
}
