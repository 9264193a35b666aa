/*
 * Extensions of the HIP back-end beyond the reference driver.h: variant and
 * precision selection, tile shapes and their autotuner, receiver traces,
 * multi-step runs captured in HIP graphs, device timing, and the batched
 * multi-shot engine.  C ABI, plain pointers and sizes.
 *
 * Every int-returning call returns FL_OK (0) or a negative FL_ERR_* code and
 * records a message for fl_last_error().
 */
#ifndef FLETCHER_FL_HIP_H
#define FLETCHER_FL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FL_OK 0
#define FL_ERR_ARG -1     /* invalid argument (shape, index, mode) */
#define FL_ERR_STATE -2   /* call out of order (e.g. Propagate before Initialize) */
#define FL_ERR_HIP -3     /* HIP runtime error */
#define FL_ERR_NOMEM -4   /* device or host allocation failed */

/* Update schemes for the cross derivatives:
 *  FUSED   one z-marching kernel, mixed derivatives as Der1 of Der1 formed on chip
 *          (the arithmetic of der1der1/sample.h:57-62 in a single pass);
 *  DIRECT  the 64-point DerCross of original/derivatives.h:59-68;
 *  TWOPASS the reference der1der1 structure: first derivatives written to four
 *          volumes, then the update (der1der1/OpenMP/openmp_propagate.c:24-93). */
enum { FL_VARIANT_FUSED = 0, FL_VARIANT_DIRECT = 1, FL_VARIANT_TWOPASS = 2 };

/* Coefficient storage.  FP32 keeps the precomp.h volumes bit for bit; BF16/FP16
 * store the same 10 volumes (v2* pre-multiplied by dt^2, evaluated in FP64 and
 * rounded once); FP16_RECOMPUTE stores the 4 velocity factors (times dt^2) in FP16
 * and theta, phi as 16-bit fixed point turns, and rebuilds the 6 angular coefficients
 * per cell; Q16_RECOMPUTE stores the same inputs with the velocity factors of a cell in
 * block floating point (15-bit mantissas and one shared exponent), 16 times finer than
 * FP16 for the cell's largest factor; FP32_RECOMPUTE stores the 4 precomp.h velocity factors bit for bit and
 * theta, phi as FP32 turns, and rebuilds the angular coefficients in FP32 (24 bytes per
 * cell instead of 40).  Wavefields stay FP32.  The recompute forms and the 16-bit forms
 * run with the fused variant. */
enum {
    FL_PREC_FP32 = 0,
    FL_PREC_BF16 = 1,
    FL_PREC_FP16 = 2,
    FL_PREC_FP16_RECOMPUTE = 3,
    FL_PREC_FP32_RECOMPUTE = 4,
    FL_PREC_Q16_RECOMPUTE = 5
};

const char *fl_last_error(void);
int fl_status(void);      /* code of the last failed call, FL_OK after fl_clear_error */
void fl_clear_error(void);

/* Front-ends that call only driver.h select with the environment instead, read by
 * DRIVER_Initialize: FL_VARIANT=fused|direct|twopass,
 * FL_PRECISION=fp32|bf16|fp16|fp16r|fp32r|q16r,
 * FL_STREAM_HINTS=0|1 (fl_set_stream_hints), FL_DEVICE_PRECOMPUTE=0|1 (fl_set_device_precompute),
 * FL_TUNE=<cache file> (autotune right after initialisation). */
int fl_set_variant(int variant);   /* takes effect at the next step */
int fl_get_variant(void);
int fl_set_precision(int precision); /* used by the next DRIVER_Initialize */
int fl_get_precision(void);
int fl_set_sync(int on);           /* 1 (default): DRIVER_Propagate waits for the step */

/* Tile shape of the z-marching kernel: ty rows (4, 8, 12 or 16) of 32 cells, z chunks of
 * zt planes.  zt 0 selects the chunk depth that fits whole rounds of resident workgroups
 * to the grid (the default, with 12-row tiles). */
int fl_set_tiles(int ty, int zt);
int fl_get_tiles(int *ty, int *zt);

/* Streaming cache hints of the z-marching step (fused and two-pass variants): 1 loads and
 * stores the previous level, and loads the recompute forms' coefficient records,
 * non-temporal; 0 (the default) uses the default policy.  Results are identical either way;
 * which is faster depends on the grid, and fl_autotune times both. */
int fl_set_stream_hints(int on);
int fl_get_stream_hints(void);

/* Workgroup order of the z-marching step: strips of sw tile columns, each walked row by row,
 * so that the tiles one XCD runs at once form a compact block whose shared halos meet in its
 * L2.  0 (the default) takes 8 columns (or all of them on narrower grids); a width at or
 * above the tile columns is the plain row-major order.  Results
 * are identical for every width.  fl_get_tile_strip returns the width in use. */
int fl_set_tile_strip(int sw);
int fl_get_tile_strip(void);

/* Lane layout of the z-marching step for the 16-bit recompute forms (fp16r, q16r) on 8-row
 * tiles: 1 (the default) runs both fields of a cell in one lane (workgroups of 4 waves, each
 * wave two tile rows), 0 one field per lane.  Other forms and tile heights always run one
 * field per lane.  Results are identical either way; single-grid and batched steps both
 * follow it.  No reference counterpart (the reference has one thread per cell). */
int fl_set_lane_pairs(int on);
int fl_get_lane_pairs(void);

/* Time every candidate (ty, zt), with and without stream hints, for
 * `steps` steps on the initialised grid and keep the fastest; fields are restored
 * afterwards.  When cache_path is not NULL the choice is looked up / appended there, keyed
 * by device name, grid, variant and precision. */
int fl_autotune(int steps, const char *cache_path);

/* Model source of fl_initialize_streamed: writes planes z0 .. z0+nzs-1 of the six model
 * arrays (slab arrays of sx*sy*nzs floats, reference layout) and returns 0, non-zero to
 * abort.  Called with z0 = 0, then each slab after the previous one. */
typedef int (*fl_model_slab_fn)(void *user, int z0, int nzs, float *vpz, float *vsv, float *epsilon,
                                float *delta, float *phi, float *theta);

/* DRIVER_Initialize for models too large for whole host arrays (no reference counterpart:
 * model setup of precomp.h:14-69 on the device).  The model arrives in z slabs of about
 * 64 MiB per array through pinned staging; the coefficient records of the selected precision
 * are computed on the device in FP64 with the per-cell code of the host path.  Fields start
 * at zero.  fl_model_stream_fill (fl_model.h) is such a source for the seeded models. */
int fl_initialize_streamed(int sx, int sy, int sz, int bord, float dx, float dy, float dz, float dt,
                           fl_model_slab_fn fill, void *user);

/* 1: DRIVER_Initialize and fl_batch_create compute the coefficient records on the device
 * as well (FP64 trig of the device library; records equal to the host's wherever the two
 * libraries' sine and cosine agree).  0 (default): host precompute, bit for bit precomp.h. */
int fl_set_device_precompute(int on);

/* Device fields after the last step: which = 0 pp, 1 pc, 2 qp, 3 qc (n floats). */
int fl_get_field(int which, float *host);
int fl_set_field(int which, const float *host);

/* Receiver traces: after every step it with it % every == 0, pc on plane iz over
 * [bord, sx-bord) x [bord, sy-bord) is stored in slot it/every - 1 (< max_traces).
 * iz < 0 disables.  fl_get_traces copies up to max slots and returns the count. */
int fl_record_traces(int iz, int every, int max_traces);
int fl_get_traces(float *host, int max_traces);

/* nsteps steps it0+1 .. it0+nsteps of the reference loop (model.c:71-88): before step
 * it0+1+k, src[k] is added to pc and qc at iSource.  Steps are captured in HIP graphs
 * of graph_chunk steps (0: no graph) and replayed.  The instantiated graph is kept and
 * replayed by later calls with the same chunk, tiles, traces and source cell (any it0 and
 * nsteps): only the first such call pays for capture and instantiation. */
int fl_run_steps(int it0, int nsteps, int iSource, const float *src, int graph_chunk);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The steps
 * of fl_run_steps (fused variant) launched one by one, allocating nothing: the run's step
 * counters and its source samples live in a fixed scratch array of the library on each GPU,
 * at most 65536 samples per call (FL_ERR_ARG beyond).  src is host memory.  The array keeps
 * the last call's contents; calls that overlap on one GPU share it and race. */
int fl_run_steps_static(int it0, int nsteps, int iSource, const float *src);

/* All four fields (and the recorded traces) back to zero, levels in their initial roles,
 * device state and graphs kept: a run can be repeated from rest. */
int fl_reset_fields(void);

int fl_synchronize(void);

/* Device timing of the propagate kernels: while on, every step kernel launched by
 * DRIVER_Propagate is bracketed by HIP events on the library stream.
 * fl_kernel_times returns the summed milliseconds and the count, then clears. */
int fl_set_timing(int on);
int fl_kernel_times(double *total_ms, int *count);
/* HIP events around a region on the library stream. */
int fl_region_begin(void);
int fl_region_end(double *ms);

/* Diagnostics: the two-pass variant's first-derivative volume k (0 pDx, 1 pDy, 2 qDx,
 * 3 qDy; der1der1/OpenMP/openmp_propagate.c:24-76) of the last step, n floats. */
int fl_get_derivative_volume(int k, float *host);
/* The device coefficient records of the initialised grid: one record per pitched cell
 * (row pitch sx rounded up to 16 cells), fl_record_bytes() bytes in all.  The device keeps
 * the FP32 records as two arrays (32 and 8 bytes a cell); this call returns them cell by cell
 * in the 10-slot record order all the same. */
long long fl_record_bytes(void);
int fl_get_records(void *host);

int fl_device_name(char *buf, int len);
long long fl_device_bytes(void);   /* device memory held by the back-end */
/* Free and total memory of the device (hipMemGetInfo). */
int fl_device_memory(long long *free_bytes, long long *total_bytes);
/* Algorithmic HBM bytes per updated cell per step of the current configuration. */
int fl_bytes_per_cell(void);

/* ---------------- batched shots: one launch per step over (shot, tile) items ---------- */
typedef struct {
    int sx, sy, sz;   /* array extents of the shot's sub-grid, halo and absorbing zone included */
    int isource;      /* source cell index within the sub-grid */
    int rec_iz;       /* receiver plane (-1: none) */
    const float *vpz, *vsv, *epsilon, *delta, *phi, *theta; /* host model of the sub-grid */
} fl_shot_t;

int fl_batch_create(int nshots, const fl_shot_t *shots, int bord, float dx, float dy, float dz,
                    float dt, int precision);
int fl_batch_set_tiles(int ty, int zt);
/* Launch form of the batched step (no reference counterpart).  0 (default): one workgroup
 * per (shot, tile) item.  Otherwise the persistent form: the resident workgroups (at most
 * wgs of them when wgs > 1) take the items of their slice of the work list from a ticket
 * counter.  Results are the same bits either way. */
int fl_batch_set_persistent(int wgs);
int fl_batch_get_tiles(int *ty, int *zt);  /* zt as used (0 resolved to the default) */
/* Steps it0+1..it0+nsteps for every shot; src as in fl_run_steps; traces every
 * rec_every steps (0: none) into max_traces slots per shot. */
int fl_batch_run(int it0, int nsteps, const float *src, int graph_chunk, int rec_every,
                 int max_traces);
int fl_batch_get_field(int shot, int which, float *host);
int fl_batch_get_traces(int shot, float *host, int max_traces);
int fl_batch_autotune(int steps, const char *cache_path);
/* Device time of the last fl_batch_run: the source upload, first insertion and the steps,
 * measured after the work list and the graph are in place. */
int fl_batch_kernel_ms(double *ms);
/* Every shot's fields and traces back to zero (as fl_reset_fields). */
int fl_batch_reset(void);
void fl_batch_destroy(void);

/* ---------------- FP64 accuracy reference on the device ----------------
 * No reference counterpart as a back-end: it is the reference's der1der1 time loop
 * (der1der1/OpenMP/openmp_propagate.c:9-95, der1der1/sample.h:35-123, model.c:71-88)
 * evaluated in FP64 with exact rational weights and FP64 coefficients
 * (fl_precompute_f64), so that full-size runs of the FP32 and 16-bit forms have an
 * FP64 result to be measured against.  Reference layout, one state per process,
 * independent of the DRIVER_* and fl_batch_* state. */
int fl_ref64_create(int sx, int sy, int sz, int bord, float dx, float dy, float dz, float dt,
                    const float *vpz, const float *vsv, const float *epsilon, const float *delta,
                    const float *phi, const float *theta);
/* Steps it0+1..it0+nsteps: before step it, src[it-1-it0] (FP32 Ricker samples, promoted)
 * is added to pc and qc at iSource; pc on plane rec_iz is kept after every step with
 * it % rec_every == 0 (rec_every 0: none) in up to max_traces slots. */
int fl_ref64_run(int it0, int nsteps, int iSource, const float *src, int rec_iz, int rec_every,
                 int max_traces);
int fl_ref64_get_field(int which, double *host);   /* 0 pp, 1 pc, 2 qp, 3 qc */
int fl_ref64_get_traces(double *host, int max_traces);  /* returns the slot count filled */
void fl_ref64_destroy(void);

#ifdef __cplusplus
}
#endif
#endif
