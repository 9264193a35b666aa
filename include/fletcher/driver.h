/*
 * The reference back-end plug-in API, unchanged: original/driver.h:8-22
 * (identical in der1der1/, der1der1hm/, der1der1lm/).  A front-end written
 * against the reference (model.c:63,76,83,100,162) links against
 * libfletcher_amd.so without source changes.
 *
 * Semantics of this HIP back-end (the reference CUDA back-end's, cuda_driver.c):
 *  - DRIVER_Initialize precomputes the 10 coefficient volumes from the model
 *    (precomp.h:14-69), or their 16-bit or recompute forms under another precision
 *    (fl_hip.h), and uploads them with the fields to the GPU.  It reads
 *    pp/pc/qp/qc once (the reference front-end passes zeros).
 *  - DRIVER_InsertSource adds src to the device pc and qc at iSource
 *    (OpenMP/openmp_insertsource.c:7-12, CUDA/cuda_insertsource.cu:28-32).
 *  - DRIVER_Propagate advances one time step on the device and swaps the device
 *    time levels internally (CUDA/cuda_propagate.cu:37-80); the host pointers are
 *    not touched.  It returns after the step completes, unless fl_set_sync(0).
 *  - DRIVER_Update_pointers copies the device pc to the host pc (NULL tolerated,
 *    CUDA/cuda_stuff.cu:195-201).
 *  - Errors never exit(): the call becomes a no-op and fl_last_error() (fl_hip.h)
 *    reports why.
 */
#ifndef __driver_h__
#define __driver_h__

#ifdef __cplusplus
extern "C" {
#define FL_RESTRICT __restrict__
#else
#define FL_RESTRICT restrict
#endif

void DRIVER_Initialize(const int sx, const int sy, const int sz, const int bord, float dx,
                       float dy, float dz, float dt, float *FL_RESTRICT vpz,
                       float *FL_RESTRICT vsv, float *FL_RESTRICT epsilon,
                       float *FL_RESTRICT delta, float *FL_RESTRICT phi,
                       float *FL_RESTRICT theta, float *FL_RESTRICT pp, float *FL_RESTRICT pc,
                       float *FL_RESTRICT qp, float *FL_RESTRICT qc);

void DRIVER_Finalize();

void DRIVER_Propagate(const int sx, const int sy, const int sz, const int bord, const float dx,
                      const float dy, const float dz, const float dt, const int it, float *pp,
                      float *pc, float *qp, float *qc);

void DRIVER_Update_pointers(const int sx, const int sy, const int sz, float *pc);

void DRIVER_InsertSource(float dt, int it, int iSource, float *p, float *q, float src);

#ifdef __cplusplus
}
#endif
#endif
