/*
 * `fletcher` — command-line front-end with the reference grammar
 * (original/main.c:47-60):   fletcher FORM nx ny nz absorb dx dy dz dt tmax
 * FORM = ISO | VTI | TTI.  Builds the reference model (main.c:93-260), runs the
 * reference time loop (model.c) on the linked back-end and prints the
 * reference's report lines and Report.csv (main.c:112-119, model.c:117-158).
 *
 * Back-end options come from the environment (the reference has none):
 *   FL_VARIANT=fused|direct|twopass   FL_PRECISION=fp32|bf16|fp16|fp16r|fp32r|q16r
 *   FL_RSF=<name>  dump pc at the start and every 0.01 s of simulated time as <name>.rsf, as
 *                  der1der1/main.c:273-300 and model.c:97-102 do (utils.c:71-205 layout)
 *   FL_TUNE=<cache file>  autotune tiles for this grid at DRIVER_Initialize
 * Errors are reported on stderr with a non-zero exit status.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fletcher/driver.h"
#include "fletcher/fl_hip.h"
#include "fletcher/fl_model.h"
#include "fletcher/fl_run.h"

/* The process's memory high-water mark (VmHWM of /proc/self/status), as model.c:119-126 reads it. */
static long high_water_mark(char *unit, int unit_len)
{
    long hwm = 0;
    snprintf(unit, unit_len, "kB");
    FILE *fp = fopen("/proc/self/status", "r");
    if (!fp) return 0;
    char line[256];
    while (fgets(line, sizeof line, fp))
        if (strncmp(line, "VmHWM", 5) == 0) {
            char u[8] = "kB";
            if (sscanf(line + 6, "%ld %7s", &hwm, u) >= 1) snprintf(unit, unit_len, "%s", u);
            break;
        }
    fclose(fp);
    return hwm;
}

/* Report.csv in the layout of model.c:15-28,135-158: problem size, then wall time,
 * MSamples/s and the high-water mark. */
static int write_report(int sx, int sy, int sz, int bord, int st, double walltime, double msamples, long hwm,
                        const char *unit)
{
    FILE *fr = fopen("Report.csv", "w");
    if (!fr) return -1;
    fprintf(fr, "sx; %d; sy; %d; sz; %d; bord; %d;  st; %d; \n", sx, sy, sz, bord, st);
    fprintf(fr, "walltime; %lf; MSamples; %lf; HWM;  %ld; HWMUnit;  %s;\n", walltime, msamples, hwm, unit);
    return fclose(fr);
}

static int env_choice(const char *name, const char *const *opts, int n, int dflt)
{
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    for (int i = 0; i < n; i++)
        if (strcmp(v, opts[i]) == 0) return i;
    fprintf(stderr, "%s=%s is not one of the accepted values\n", name, v);
    exit(2);
}

int main(int argc, char **argv)
{
    const int bord = 4; /* stencil halo, main.c:21 */
    if (argc < 11) {
        printf("program requires %d input arguments; execution halted\n", 10);
        return 2;
    }
    int form;
    if (strcmp(argv[1], "ISO") == 0) form = FL_FORM_ISO;
    else if (strcmp(argv[1], "VTI") == 0) form = FL_FORM_VTI;
    else if (strcmp(argv[1], "TTI") == 0) form = FL_FORM_TTI;
    else {
        printf("Input problem formulation (%s) is unknown\n", argv[1]);
        return 2;
    }
    const int nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]), absorb = atoi(argv[5]);
    const float dx = atof(argv[6]), dy = atof(argv[7]), dz = atof(argv[8]), dt = atof(argv[9]);
    const float tmax = atof(argv[10]);
    if (nx <= 0 || ny <= 0 || nz <= 0 || absorb < 1) {
        fprintf(stderr, "grid sizes must be positive and absorb >= 1\n");
        return 2;
    }
    static const char *variants[] = {"fused", "direct", "twopass"};
    static const char *precs[] = {"fp32", "bf16", "fp16", "fp16r", "fp32r", "q16r"};
    fl_set_variant(env_choice("FL_VARIANT", variants, 3, 0));
    fl_set_precision(env_choice("FL_PRECISION", precs, 6, 0));

    const int sx = nx + 2 * bord + 2 * absorb, sy = ny + 2 * bord + 2 * absorb, sz = nz + 2 * bord + 2 * absorb;
    const int st = fl_num_steps(tmax, dt);
    const int iSource = ((sz / 2) * sy + sy / 2) * sx + sx / 2;
    const float dtOutput = 0.01;
    static const char *names[] = {"isotropic", "anisotropic with vertical transversely isotropy",
                                  "anisotropic with tilted transversely isotropy"};
    printf("Problem is %s\n", names[form]);
    printf("Grid size is (%d,%d,%d) with spacing (%.2f,%.2f,%.2f); simulated area (%.2f,%.2f,%.2f) \n", nx,
           ny, nz, dx, dy, dz, (nx - 1) * dx, (ny - 1) * dy, (nz - 1) * dz);
    printf("Grid is extended by %d absortion points and %d border points at each extreme\n", absorb, bord);
    printf("Wave is propagated at internal+absortion points of size (%d,%d,%d)\n", nx + 2 * absorb,
           ny + 2 * absorb, nz + 2 * absorb);
    printf("Source at coordinates (%d,%d,%d)\n", sx / 2, sy / 2, sz / 2);
    printf("Will run %d time steps of %f to reach time %f\n", st, dt, st * dt);

    const int64_t n = (int64_t)sx * sy * sz;
    float *m[6], *f[4];
    for (int k = 0; k < 6; k++) m[k] = (float *)malloc(sizeof(float) * n);
    for (int k = 0; k < 4; k++) f[k] = (float *)calloc(n, sizeof(float));
    for (int k = 0; k < 6; k++)
        if (!m[k]) { fprintf(stderr, "host allocation failed\n"); return 3; }
    for (int k = 0; k < 4; k++)
        if (!f[k]) { fprintf(stderr, "host allocation failed\n"); return 3; }
    fl_model_homogeneous(form, n, m[0], m[1], m[2], m[3], m[4], m[5]);
    const float recdt = fl_recommended_dt(n, m[0], m[2], dx, dy, dz, 0.2);
    printf("Recomended maximum time step is %f; used time step is %f\n", recdt, dt);
    fl_random_boundary(1, sx, sy, sz, nx, ny, nz, bord, absorb, m[0], m[1]);

    fl_rsf_t *rsf = NULL;
    const char *rsfname = getenv("FL_RSF");
    if (rsfname && *rsfname) {
        rsf = fl_rsf_open(rsfname, 0, sx - 1, 0, sy - 1, 0, sz - 1, dx, dy, dz, dt);
        if (!rsf) { fprintf(stderr, "cannot open %s.rsf\n", rsfname); return 3; }
        fl_rsf_append(rsf, sx, sy, sz, f[1]); /* the initial pc, der1der1/main.c:279 */
    }
    /* FL_TUNE is applied by DRIVER_Initialize itself (the library reads the environment) */
    fl_report_t rep;
    float *pp = f[0], *pc = f[1], *qp = f[2], *qc = f[3];
    fl_model_run(st, iSource, dtOutput, sx, sy, sz, bord, dx, dy, dz, dt, &pp, &pc, &qp, &qc, m[0], m[1],
                 m[2], m[3], m[4], m[5], rsf, 0, &rep);
    if (rsf) fl_rsf_close(rsf);
    if (fl_status() != FL_OK) {
        fprintf(stderr, "fletcher: %s\n", fl_last_error());
        return 1;
    }
    char unit[8];
    const long hwm = high_water_mark(unit, sizeof unit);
    printf("Execution time (s) is %lf\n", rep.walltime);
    printf("MSamples/s %.0lf\n", rep.msamples);
    printf("Memory High Water Mark is %ld %s\n", hwm, unit);
    if (write_report(sx, sy, sz, bord, st, rep.walltime, rep.msamples, hwm, unit) != 0)
        fprintf(stderr, "fletcher: cannot write Report.csv\n");
    for (int k = 0; k < 6; k++) free(m[k]);
    for (int k = 0; k < 4; k++) free(f[k]);
    return 0;
}
