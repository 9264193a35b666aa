/*
 * Host-side model setup of the Fletcher TTI path: the parts of the reference
 * front-end that run once before the time loop.  C ABI, plain pointers.
 *
 * Arrays are FP32 volumes of sx*sy*sz cells, x fastest:
 *   ind(ix,iy,iz) = ((iz*sy + iy)*sx + ix)        (reference map.h:14)
 * Each declaration cites the reference code it restates.
 */
#ifndef FLETCHER_FL_MODEL_H
#define FLETCHER_FL_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem formulations of main.c:13 */
enum { FL_FORM_ISO = 0, FL_FORM_VTI = 1, FL_FORM_TTI = 2 };

/* glibc-compatible rand() stream (TYPE_3 additive feedback generator, the
 * algorithm behind the reference's unseeded rand() in boundary.c:82), kept
 * re-entrant so that every shot can own its own stream. */
typedef struct {
    int32_t r[34];
    int k;
} fl_rand_t;
void fl_rand_seed(fl_rand_t *s, uint32_t seed);  /* srand(seed) */
int32_t fl_rand_next(fl_rand_t *s);              /* rand(), in [0, 2^31-1] */

/* Homogeneous ISO / VTI / TTI model — main.c:168-221 (SIGMA=0.75, fletcher.h:10).
 * Returns 0, or -1 for an unknown form. */
int fl_model_homogeneous(int form, int64_t n, float *vpz, float *vsv, float *epsilon,
                         float *delta, float *phi, float *theta);

/* Recommended time step MI*min(dx,dy,dz)/max(vpz*sqrt(1+2*eps)) — main.c:225-236 (MI=0.2). */
float fl_recommended_dt(int64_t n, const float *vpz, const float *epsilon, float dx, float dy,
                        float dz, double mi);

/* Random-velocity absorbing boundary — boundary.c:7-98, with the rand() stream
 * started by srand(seed) (the reference never seeds: seed 1). */
void fl_random_boundary(uint32_t seed, int sx, int sy, int sz, int nx, int ny, int nz, int bord,
                        int absorb, float *vpz, float *vsv);

/* The same boundary plane by plane, for models produced in z slabs:
 * fl_boundary_max accumulates the interior maxima (boundary.c:12-24) over the interior
 * planes z0 .. z0+nzs-1 held in vpz, vsv (slab arrays of nzs planes; start with 0);
 * fl_boundary_planes applies boundary.c:26-98 to the slab of planes z0 .. z0+nzs-1,
 * continuing the rand() stream rs (seeded with fl_rand_seed, slabs in increasing z).
 * elo_*, ehi_* are the unmodified planes iz = bord+absorb and bord+absorb+nz-1, from which
 * the cells of the low and high z absorbing zones take their edge values. */
void fl_boundary_max(int sx, int sy, int nx, int ny, int bord, int absorb, int z0, int nzs,
                     const float *vpz, const float *vsv, float *maxP, float *maxS);
void fl_boundary_planes(fl_rand_t *rs, int sx, int sy, int nx, int ny, int nz, int bord, int absorb,
                        float maxP, float maxS, const float *elo_vpz, const float *elo_vsv,
                        const float *ehi_vpz, const float *ehi_vsv, int z0, int nzs, float *vpz, float *vsv);

/* Ricker source sample at time it*dt — source.c:7-15, source.h:11-14. */
float fl_source(float dt, int it);

/* Coefficient precompute into the 10 FP32 volumes of precomp.h:1-12, in the
 * order ch1dxx ch1dyy ch1dzz ch1dxy ch1dyz ch1dxz v2px v2pz v2sz v2pn —
 * precomp.h:25-38 (trig in double, products in float) and :53-58. */
void fl_precompute_f32(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                       const float *delta, const float *phi, const float *theta, float *const *coef);

/* Same 10 coefficients evaluated entirely in FP64 from the FP32 model. */
void fl_precompute_f64(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                       const float *delta, const float *phi, const float *theta,
                       double *const *coef);

/* 16-bit coefficient storage (no reference counterpart).  The 10 coefficients of
 * fl_precompute_f64 with v2px, v2pz, v2sz, v2pn multiplied by dt^2, each rounded once
 * (round-to-nearest-even) from FP64 to BF16 (bf16 != 0) or FP16; out[k] hold n values. */
void fl_pack_coefs16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                     const float *delta, const float *phi, const float *theta, double dt, int bf16,
                     uint16_t *const *out);
/* Inputs of the in-kernel recompute, each rounded once from FP64: the velocity factors
 * v2px, v2pz, v2sz, v2pn of fl_precompute_f64 times dt^2 in FP16, and theta, phi as
 * 16-bit fixed point in revolutions (fl_angle_u16); the six angular coefficients are
 * rebuilt from the angles per cell. */
void fl_pack_recompute16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                         const float *delta, const float *phi, const float *theta, double dt,
                         uint16_t *const *out);
/* Inputs of the 16-bit block floating-point recompute: per cell, the four FP64 velocity
 * factors times dt^2 as 15-bit mantissas of one shared exponent 2^(ebase + e), e in [0, 15]
 * the smallest that holds them; as 32-bit words out[0] | out[3] << 16 = v2px | v2pn << 15 |
 * (e & 3) << 30 and out[1] | out[2] << 16 = v2pz | v2sz << 15 | (e >> 2) << 30.  ebase =
 * fl_recompute_q16_base over the grid.  theta, phi as fl_angle_u16 turns. */
int fl_recompute_q16_base(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                          const float *delta, double dt);
void fl_pack_recompute_q16(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                           const float *delta, const float *phi, const float *theta, double dt,
                           int ebase, uint16_t *const *out);
/* Inputs of the in-kernel FP32 recompute: v2px, v2pz, v2sz, v2pn exactly as
 * fl_precompute_f32 (precomp.h:53-58) forms them, and theta, phi in turns reduced to
 * [-1/2, 1/2] (FP64, rounded once to FP32); the six angular coefficients are rebuilt
 * from the angles per cell in FP32. */
void fl_pack_recompute32(int64_t n, const float *vpz, const float *vsv, const float *epsilon,
                         const float *delta, const float *phi, const float *theta, float *const *out);
/* Cell records: dst[i*nk + k] = src[k][i] (the device keeps the coefficients of a cell
 * together, one record per cell). */
void fl_interleave32(int64_t n, int nk, const float *const *src, float *dst);
void fl_interleave16(int64_t n, int nk, const uint16_t *const *src, uint16_t *dst);
/* Single conversions (round to nearest even), exported for tests. */
uint16_t fl_f64_to_f16(double v);
uint16_t fl_f64_to_bf16(double v);
/* An angle given in turns as round(frac(turns) * 65536) mod 65536. */
uint16_t fl_angle_u16(double turns);

/* Number of time steps ceil(tmax/dt) in float, main.c:101. */
int fl_num_steps(float tmax, float dt);

/* ---------------- seeded heterogeneous models (no reference counterpart) ----------------
 * All generators fill the whole sx*sy*sz array, interior = [bord+absorb, bord+absorb+n)
 * on each axis, then apply fl_random_boundary(boundary_seed, ...).
 * vsv = vpz*sqrt(|eps-delta|/0.75) as main.c:218. */

/* Linear vpz gradient vmin->vmax with depth over the interior, times (1 + sum of nblobs
 * Gaussian blobs of amplitude in [-amp,amp], widths in [smin,smax] cells) clipped to
 * [1-amp, 1+amp]; eps a smooth field in [0,eps_max]; delta = smooth*min(eps, delta_max)
 * (so delta<=eps); theta smooth in [0,theta_max]; phi smooth in [0,phi_max). */
typedef struct {
    uint64_t seed;
    float vmin, vmax;
    int nblobs;
    float blob_amp, blob_smin, blob_smax;
    float eps_max, delta_max, theta_max, phi_max;
    int nsmooth;            /* Gaussians per smooth field */
    float smooth_smin, smooth_smax;
    uint32_t boundary_seed;
} fl_gradient_model_t;
void fl_gradient_model_defaults(fl_gradient_model_t *m, uint64_t seed);
int fl_model_gradient(const fl_gradient_model_t *m, int sx, int sy, int sz, int nx, int ny, int nz,
                      int bord, int absorb, float *vpz, float *vsv, float *epsilon, float *delta,
                      float *phi, float *theta);

/* Dipping-layer model: nlayers planar layers with normals tilted up to max_dip_deg,
 * vpz rising from vmin to vmax by layer, eps/delta piecewise per layer (delta<=eps),
 * theta/phi smooth random fields. */
typedef struct {
    uint64_t seed;
    int nlayers;
    float vmin, vmax, max_dip_deg;
    float eps_max, delta_max, theta_max, phi_max;
    int nsmooth;
    float smooth_smin, smooth_smax;
    uint32_t boundary_seed;
} fl_layered_model_t;
void fl_layered_model_defaults(fl_layered_model_t *m, uint64_t seed);
int fl_model_layered(const fl_layered_model_t *m, int sx, int sy, int sz, int nx, int ny, int nz,
                     int bord, int absorb, float *vpz, float *vsv, float *epsilon, float *delta,
                     float *phi, float *theta);

/* The same models produced in z slabs, for grids whose full host arrays would not fit:
 * opening draws everything from the seed and makes two passes over the volume (interior
 * maxima and edge planes of the boundary, then the CFL velocity of the finished model);
 * fl_model_stream_fill then writes planes z0 .. z0+nzs-1 of the six arrays (slab arrays of
 * nzs planes), boundary applied, bit for bit the planes of fl_model_gradient /
 * fl_model_layered.  Slabs come in increasing z from plane 0; after the last plane the
 * stream starts over.  The fill has the signature of fl_model_slab_fn (fl_hip.h), with the
 * stream as its user pointer.  NULL when a host allocation fails. */
typedef struct fl_model_stream fl_model_stream_t;
fl_model_stream_t *fl_model_stream_gradient(const fl_gradient_model_t *m, int sx, int sy, int sz, int nx, int ny,
                                            int nz, int bord, int absorb);
fl_model_stream_t *fl_model_stream_layered(const fl_layered_model_t *m, int sx, int sy, int sz, int nx, int ny,
                                           int nz, int bord, int absorb);
int fl_model_stream_fill(void *stream, int z0, int nzs, float *vpz, float *vsv, float *epsilon, float *delta,
                         float *phi, float *theta);
/* fl_cfl_dt of the finished model. */
float fl_model_stream_cfl_dt(const fl_model_stream_t *s, float h, float courant);
void fl_model_stream_close(fl_model_stream_t *s);

/* dt from a Courant number on max(vpz*sqrt(1+2*eps)) over the array. */
float fl_cfl_dt(int64_t n, const float *vpz, const float *epsilon, float h, float courant);

/* Copy the window [x0,x0+wx)x[y0,y0+wy)x[z0,z0+wz) of a volume (sub-grid extraction
 * for shots and oracle windows). */
void fl_extract_window(const float *src, int sx, int sy, int sz, int x0, int y0, int z0, int wx,
                       int wy, int wz, float *dst);

#ifdef __cplusplus
}
#endif
#endif
