/*
 * Host time loop and output of the Fletcher path, written against driver.h
 * only, so it runs on any back-end that implements the reference plug-in API.
 */
#ifndef FLETCHER_FL_RUN_H
#define FLETCHER_FL_RUN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double walltime;          /* sum of Propagate + pointer swap seconds (model.c:82-88) */
    double msamples;          /* 1e-6 * (sx-2bord)(sy-2bord)(sz-2bord) * st / walltime (model.c:44-45,117) */
    long long samples_per_step;
    int steps;
} fl_report_t;

typedef struct fl_rsf fl_rsf_t;

/* The reference Model() (model.c:32-164): DRIVER_Initialize, then for it=1..st
 * Source(dt,it-1) -> DRIVER_InsertSource -> timed DRIVER_Propagate + swap ->
 * DRIVER_Update_pointers every dtOutput of simulated time (appending pc to `rsf`
 * when not NULL), then DRIVER_Finalize unless keep_device.  On return pc/qc
 * name the buffers that hold the last field copied back (host swap order). */
int fl_model_run(int st, int iSource, float dtOutput, int sx, int sy, int sz, int bord, float dx,
                 float dy, float dz, float dt, float **pp, float **pc, float **qp, float **qc,
                 float *vpz, float *vsv, float *epsilon, float *delta, float *phi, float *theta,
                 fl_rsf_t *rsf, int keep_device, fl_report_t *rep);

/* SEP/Madagascar RSF output in the layout of utils.c:71-205: header `<name>.rsf`
 * (in=, data_format="native_float", esize=4, n1.. d1..) and binary `<name>.rsf@`.
 * A slice file accumulates one window per call; n of the last axis = number of calls. */
fl_rsf_t *fl_rsf_open(const char *name, int ixStart, int ixEnd, int iyStart, int iyEnd, int izStart,
                      int izEnd, float dx, float dy, float dz, float dt);
int fl_rsf_append(fl_rsf_t *f, int sx, int sy, int sz, const float *arr);
int fl_rsf_close(fl_rsf_t *f);
/* One-shot dump of a window (utils.c:13-48 DumpFieldToFile). */
int fl_rsf_write_field(const char *name, int sx, int sy, int sz, int ixStart, int ixEnd, int iyStart,
                       int iyEnd, int izStart, int izEnd, float d1, float d2, float d3,
                       const float *arr);
/* Receiver traces as a 3-D RSF cube n1=nx, n2=ny, n3=ntraces (d3 = dt*every). */
int fl_rsf_write_traces(const char *name, int nx, int ny, int ntr, float dx, float dy, float dtr,
                        const float *traces);

#ifdef __cplusplus
}
#endif
#endif
