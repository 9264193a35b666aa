/*
 * Time loop (reference model.c:32-164) and RSF output (reference utils.c:71-243)
 * of the Fletcher TTI path.  Uses only the driver.h plug-in API.
 */
#include "fletcher/fl_run.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "fletcher/driver.h"
#include "fletcher/fl_model.h"

static double wtime(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1.0e-6 * tv.tv_usec;
}

struct fl_rsf {
    int dir; /* 0 x slice, 1 y slice, 2 z slice, 3 full */
    int ix0, ix1, iy0, iy1, iz0, iz1, count;
    float dx, dy, dz, dt;
    FILE *head, *bin;
    char binname[512];
};

fl_rsf_t *fl_rsf_open(const char *name, int ixStart, int ixEnd, int iyStart, int iyEnd, int izStart,
                      int izEnd, float dx, float dy, float dz, float dt)
{
    fl_rsf_t *f = (fl_rsf_t *)calloc(1, sizeof(*f));
    if (!f) return NULL;
    f->dir = (ixStart == ixEnd) ? 0 : (iyStart == iyEnd) ? 1 : (izStart == izEnd) ? 2 : 3;
    f->ix0 = ixStart; f->ix1 = ixEnd; f->iy0 = iyStart; f->iy1 = iyEnd; f->iz0 = izStart; f->iz1 = izEnd;
    f->dx = dx; f->dy = dy; f->dz = dz; f->dt = dt;
    char hname[512];
    snprintf(hname, sizeof(hname), "%s.rsf", name);
    snprintf(f->binname, sizeof(f->binname), "%s.rsf@", name);
    f->head = fopen(hname, "w");
    f->bin = fopen(f->binname, "wb");
    if (!f->head || !f->bin) {
        if (f->head) fclose(f->head);
        if (f->bin) fclose(f->bin);
        free(f);
        return NULL;
    }
    return f;
}

int fl_rsf_append(fl_rsf_t *f, int sx, int sy, int sz, const float *arr)
{
    (void)sz;
    const int n1 = f->ix1 - f->ix0 + 1;
    for (int iz = f->iz0; iz <= f->iz1; iz++)
        for (int iy = f->iy0; iy <= f->iy1; iy++)
            if (fwrite(arr + ((int64_t)iz * sy + iy) * sx + f->ix0, sizeof(float), n1, f->bin) != (size_t)n1)
                return -1;
    f->count++;
    return 0;
}

int fl_rsf_close(fl_rsf_t *f)
{
    if (!f) return -1;
    const char *b = strrchr(f->binname, '/');
    fprintf(f->head, "in=\"%s%s\"\n", b ? "" : "./", f->binname);
    fprintf(f->head, "data_format=\"native_float\"\n");
    fprintf(f->head, "esize=%d\n", (int)sizeof(float));
    const int nx = f->ix1 - f->ix0 + 1, ny = f->iy1 - f->iy0 + 1, nz = f->iz1 - f->iz0 + 1;
    switch (f->dir) {
    case 0:
        fprintf(f->head, "n1=%d\nn2=%d\nn3=%d\nd1=%f\nd2=%f\nd3=%f\n", ny, nz, f->count, f->dy, f->dz, f->dt);
        break;
    case 1:
        fprintf(f->head, "n1=%d\nn2=%d\nn3=%d\nd1=%f\nd2=%f\nd3=%f\n", nx, nz, f->count, f->dx, f->dz, f->dt);
        break;
    case 2:
        fprintf(f->head, "n1=%d\nn2=%d\nn3=%d\nd1=%f\nd2=%f\nd3=%f\n", nx, ny, f->count, f->dx, f->dy, f->dt);
        break;
    default:
        fprintf(f->head, "n1=%d\nn2=%d\nn3=%d\nn4=%d\nd1=%f\nd2=%f\nd3=%f\nd4=%f\n", nx, ny, nz, f->count,
                f->dx, f->dy, f->dz, f->dt);
    }
    fclose(f->head);
    fclose(f->bin);
    free(f);
    return 0;
}

int fl_rsf_write_field(const char *name, int sx, int sy, int sz, int ixStart, int ixEnd, int iyStart,
                       int iyEnd, int izStart, int izEnd, float d1, float d2, float d3, const float *arr)
{
    (void)sz;
    char hname[512], bname[512];
    snprintf(hname, sizeof(hname), "%s.rsf", name);
    snprintf(bname, sizeof(bname), "%s.rsf@", name);
    FILE *fp = fopen(hname, "w");
    if (!fp) return -1;
    const char *base = strrchr(bname, '/');
    fprintf(fp, "in=\"./%s\"\n", base ? base + 1 : bname);
    fprintf(fp, "data_format=\"native_float\"\n");
    fprintf(fp, "esize=%d\n", (int)sizeof(float));
    fprintf(fp, "n1=%d\nd1=%f\nn2=%d\nd2=%f\nn3=%d\nd3=%f\n", ixEnd - ixStart + 1, d1, iyEnd - iyStart + 1, d2,
            izEnd - izStart + 1, d3);
    fclose(fp);
    fp = fopen(bname, "wb");
    if (!fp) return -1;
    for (int iz = izStart; iz <= izEnd; iz++)
        for (int iy = iyStart; iy <= iyEnd; iy++)
            fwrite(arr + ((int64_t)iz * sy + iy) * sx + ixStart, sizeof(float), ixEnd - ixStart + 1, fp);
    fclose(fp);
    return 0;
}

int fl_rsf_write_traces(const char *name, int nx, int ny, int ntr, float dx, float dy, float dtr,
                        const float *traces)
{
    char hname[512], bname[512];
    snprintf(hname, sizeof(hname), "%s.rsf", name);
    snprintf(bname, sizeof(bname), "%s.rsf@", name);
    FILE *fp = fopen(hname, "w");
    if (!fp) return -1;
    const char *base = strrchr(bname, '/');
    fprintf(fp, "in=\"./%s\"\ndata_format=\"native_float\"\nesize=4\n", base ? base + 1 : bname);
    fprintf(fp, "n1=%d\nd1=%f\nn2=%d\nd2=%f\nn3=%d\nd3=%f\n", nx, dx, ny, dy, ntr, dtr);
    fclose(fp);
    fp = fopen(bname, "wb");
    if (!fp) return -1;
    const size_t n = (size_t)nx * ny * ntr;
    const size_t w = fwrite(traces, sizeof(float), n, fp);
    fclose(fp);
    return w == n ? 0 : -1;
}

int fl_model_run(int st, int iSource, float dtOutput, int sx, int sy, int sz, int bord, float dx,
                 float dy, float dz, float dt, float **pp, float **pc, float **qp, float **qc,
                 float *vpz, float *vsv, float *epsilon, float *delta, float *phi, float *theta,
                 fl_rsf_t *rsf, int keep_device, fl_report_t *rep)
{
    float tSim = 0.0;
    int nOut = 1;
    float tOut = nOut * dtOutput;
    const long long samplesPropagate = (long long)(sx - 2 * bord) * (sy - 2 * bord) * (sz - 2 * bord);
    DRIVER_Initialize(sx, sy, sz, bord, dx, dy, dz, dt, vpz, vsv, epsilon, delta, phi, theta, *pp, *pc,
                      *qp, *qc);
    double walltime = 0.0;
    for (int it = 1; it <= st; it++) {
        const float src = fl_source(dt, it - 1);
        DRIVER_InsertSource(dt, it - 1, iSource, *pc, *qc, src);
        const double t0 = wtime();
        DRIVER_Propagate(sx, sy, sz, bord, dx, dy, dz, dt, it, *pp, *pc, *qp, *qc);
        float *t = *pp; *pp = *pc; *pc = t;
        t = *qp; *qp = *qc; *qc = t;
        walltime += wtime() - t0;
        tSim = it * dt;
        if (tSim >= tOut) {
            DRIVER_Update_pointers(sx, sy, sz, *pc);
            if (rsf) fl_rsf_append(rsf, sx, sy, sz, *pc);
            tOut = (++nOut) * dtOutput;
        }
    }
    DRIVER_Update_pointers(sx, sy, sz, *pc);
    if (rep) {
        rep->walltime = walltime;
        rep->samples_per_step = samplesPropagate;
        rep->steps = st;
        rep->msamples = walltime > 0 ? (1.0e-6 * (double)samplesPropagate * st) / walltime : 0.0;
    }
    if (!keep_device) DRIVER_Finalize();
    return 0;
}
