/*
 * ORACLE BODY — test infrastructure, never product code.
 *
 * Included twice by fl_oracle.c: once with REAL=float (bitwise restatement of
 * the reference OpenMP path) and once with REAL=double (FP64 accuracy oracle).
 * Every expression keeps the reference's association order so the FP32
 * instantiation, built without FMA contraction, reproduces the reference
 * bit for bit.
 *
 * Macros expected: REAL, SFX (name suffix), K0..K4, L1..L4, L11..L44 (REAL literals).
 */

#define OR_CAT2(a, b) a##b
#define OR_CAT(a, b) OR_CAT2(a, b)
#define OR_FN(name) OR_CAT(name, SFX)

/* Der2 — original/derivatives.h:53. Sum order: K0*c, K1 pair, K2 pair, K3 pair, K4 pair. */
static inline REAL OR_FN(der2_)(const REAL *p, long i, long s, REAL d2inv)
{
    return (K0 * p[i] + K1 * (p[i + s] + p[i - s]) + K2 * (p[i + 2 * s] + p[i - 2 * s]) +
            K3 * (p[i + 3 * s] + p[i - 3 * s]) + K4 * (p[i + 4 * s] + p[i - 4 * s])) *
           d2inv;
}

/* Der1 — original/derivatives.h:47. */
static inline REAL OR_FN(der1_)(const REAL *p, long i, long s, REAL dinv)
{
    return (L1 * (p[i + s] - p[i - s]) + L2 * (p[i + 2 * s] - p[i - 2 * s]) +
            L3 * (p[i + 3 * s] - p[i - 3 * s]) + L4 * (p[i + 4 * s] - p[i - 4 * s])) *
           dinv;
}

/* One coefficient group of DerCross — original/derivatives.h:59-68.
 * Group (m,n), m<=n: first quadruple at offsets (+-m*s21, +-n*s11), then, when
 * m!=n, the mirrored quadruple (+-n*s21, +-m*s11), accumulated left to right
 * exactly as the reference macro chains its 8 terms. */
static inline REAL OR_FN(cross_group_)(const REAL *p, long i, long s11, long s21, int m, int n)
{
    const long a = n * s11, b = m * s21;
    REAL v = p[i + b + a] - p[i + b - a] - p[i - b + a] + p[i - b - a];
    if (m != n) {
        const long a2 = m * s11, b2 = n * s21;
        v = v + p[i + b2 + a2] - p[i + b2 - a2] - p[i - b2 + a2] + p[i - b2 - a2];
    }
    return v;
}

/* DerCross — 64-point mixed derivative, groups summed L11,L12,L13,L14,L22,L23,L24,L33,L34,L44. */
static inline REAL OR_FN(dercross_)(const REAL *p, long i, long s11, long s21, REAL dinv)
{
    return (L11 * OR_FN(cross_group_)(p, i, s11, s21, 1, 1) +
            L12 * OR_FN(cross_group_)(p, i, s11, s21, 1, 2) +
            L13 * OR_FN(cross_group_)(p, i, s11, s21, 1, 3) +
            L14 * OR_FN(cross_group_)(p, i, s11, s21, 1, 4) +
            L22 * OR_FN(cross_group_)(p, i, s11, s21, 2, 2) +
            L23 * OR_FN(cross_group_)(p, i, s11, s21, 2, 3) +
            L24 * OR_FN(cross_group_)(p, i, s11, s21, 2, 4) +
            L33 * OR_FN(cross_group_)(p, i, s11, s21, 3, 3) +
            L34 * OR_FN(cross_group_)(p, i, s11, s21, 3, 4) +
            L44 * OR_FN(cross_group_)(p, i, s11, s21, 4, 4)) *
           dinv;
}

/* Coefficient combination and leapfrog shared by both variants — original/sample.h:46-86.
 * c[] order: ch1dxx ch1dyy ch1dzz ch1dxy ch1dyz ch1dxz v2px v2pz v2sz v2pn (precomp.h:1-12). */
static inline void OR_FN(update_cell_)(long i, const REAL *const *c, REAL dt,
                                        REAL pxx, REAL pyy, REAL pzz, REAL pxy, REAL pyz, REAL pxz,
                                        REAL qxx, REAL qyy, REAL qzz, REAL qxy, REAL qyz, REAL qxz,
                                        REAL *pp, const REAL *pc, REAL *qp, const REAL *qc)
{
    const REAL cpxx = c[0][i] * pxx;
    const REAL cpyy = c[1][i] * pyy;
    const REAL cpzz = c[2][i] * pzz;
    const REAL cpxy = c[3][i] * pxy;
    const REAL cpxz = c[5][i] * pxz;
    const REAL cpyz = c[4][i] * pyz;
    const REAL h1p = cpxx + cpyy + cpzz + cpxy + cpxz + cpyz;
    const REAL h2p = pxx + pyy + pzz - h1p;

    const REAL cqxx = c[0][i] * qxx;
    const REAL cqyy = c[1][i] * qyy;
    const REAL cqzz = c[2][i] * qzz;
    const REAL cqxy = c[3][i] * qxy;
    const REAL cqxz = c[5][i] * qxz;
    const REAL cqyz = c[4][i] * qyz;
    const REAL h1q = cqxx + cqyy + cqzz + cqxy + cqxz + cqyz;
    const REAL h2q = qxx + qyy + qzz - h1q;

    const REAL h1pmq = h1p - h1q;
    const REAL h2pmq = h2p - h2q;

    const REAL rhsp = c[6][i] * h2p + c[7][i] * h1q + c[8][i] * h1pmq;
    const REAL rhsq = c[9][i] * h2p + c[7][i] * h1q - c[8][i] * h2pmq;

    pp[i] = (REAL)2.0 * pc[i] - pp[i] + rhsp * dt * dt;
    qp[i] = (REAL)2.0 * qc[i] - qp[i] + rhsq * dt * dt;
}

/* OPENMP_Propagate — original/OpenMP/openmp_propagate.c:9-34 with original/sample.h:32-90.
 * Update domain iz,iy,ix in [bord, s-bord); halo untouched. */
void OR_FN(oracle_propagate_original_)(int sx, int sy, int sz, int bord, REAL dx, REAL dy, REAL dz,
                                       REAL dt, const REAL *const *c, REAL *pp, const REAL *pc,
                                       REAL *qp, const REAL *qc)
{
    const long strideX = 1, strideY = sx, strideZ = (long)sx * sy;
    const REAL dxxinv = (REAL)1.0 / (dx * dx);
    const REAL dyyinv = (REAL)1.0 / (dy * dy);
    const REAL dzzinv = (REAL)1.0 / (dz * dz);
    const REAL dxyinv = (REAL)1.0 / (dx * dy);
    const REAL dxzinv = (REAL)1.0 / (dx * dz);
    const REAL dyzinv = (REAL)1.0 / (dy * dz);
#pragma omp parallel for schedule(static)
    for (int iz = bord; iz < sz - bord; iz++)
        for (int iy = bord; iy < sy - bord; iy++)
            for (int ix = bord; ix < sx - bord; ix++) {
                const long i = ((long)iz * sy + iy) * sx + ix;
                const REAL pxx = OR_FN(der2_)(pc, i, strideX, dxxinv);
                const REAL pyy = OR_FN(der2_)(pc, i, strideY, dyyinv);
                const REAL pzz = OR_FN(der2_)(pc, i, strideZ, dzzinv);
                const REAL pxy = OR_FN(dercross_)(pc, i, strideX, strideY, dxyinv);
                const REAL pyz = OR_FN(dercross_)(pc, i, strideY, strideZ, dyzinv);
                const REAL pxz = OR_FN(dercross_)(pc, i, strideX, strideZ, dxzinv);
                const REAL qxx = OR_FN(der2_)(qc, i, strideX, dxxinv);
                const REAL qyy = OR_FN(der2_)(qc, i, strideY, dyyinv);
                const REAL qzz = OR_FN(der2_)(qc, i, strideZ, dzzinv);
                const REAL qxy = OR_FN(dercross_)(qc, i, strideX, strideY, dxyinv);
                const REAL qyz = OR_FN(dercross_)(qc, i, strideY, strideZ, dyzinv);
                const REAL qxz = OR_FN(dercross_)(qc, i, strideX, strideZ, dxzinv);
                OR_FN(update_cell_)(i, c, dt, pxx, pyy, pzz, pxy, pyz, pxz, qxx, qyy, qzz, qxy,
                                    qyz, qxz, pp, pc, qp, qc);
            }
}

/* OPENMP_PropagateDer1Der1ORIG — der1der1/OpenMP/openmp_propagate.c:9-95 with
 * der1der1/sample.h:35-123. Pass 1 ranges follow :24-76 exactly:
 *   pDx,qDx: iz in [0,sz), iy in [0,sy), ix in [bord, sx-bord)
 *   pDy,qDy: iz in [0,sz), iy in [bord, sy-bord), ix in [bord, sx)
 * Pass 2 replaces DerCross by Der1 of the first-derivative volumes. */
void OR_FN(oracle_propagate_der1der1_)(int sx, int sy, int sz, int bord, REAL dx, REAL dy,
                                       REAL dz, REAL dt, const REAL *const *c, REAL *pp,
                                       const REAL *pc, REAL *qp, const REAL *qc, REAL *pDx,
                                       REAL *pDy, REAL *qDx, REAL *qDy)
{
    const long strideX = 1, strideY = sx, strideZ = (long)sx * sy;
    const REAL dxxinv = (REAL)1.0 / (dx * dx);
    const REAL dyyinv = (REAL)1.0 / (dy * dy);
    const REAL dzzinv = (REAL)1.0 / (dz * dz);
    const REAL dxinv = (REAL)1.0 / dx;
    const REAL dyinv = (REAL)1.0 / dy;
    const REAL dzinv = (REAL)1.0 / dz;
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (int iz = 0; iz < sz; iz++)
            for (int iy = 0; iy < sy; iy++) {
                const long row = ((long)iz * sy + iy) * sx;
                for (int ix = bord; ix < sx - bord; ix++) {
                    pDx[row + ix] = OR_FN(der1_)(pc, row + ix, strideX, dxinv);
                    qDx[row + ix] = OR_FN(der1_)(qc, row + ix, strideX, dxinv);
                }
            }
#pragma omp for schedule(static)
        for (int iz = 0; iz < sz; iz++)
            for (int iy = bord; iy < sy - bord; iy++) {
                const long row = ((long)iz * sy + iy) * sx;
                for (int ix = bord; ix < sx; ix++) {
                    pDy[row + ix] = OR_FN(der1_)(pc, row + ix, strideY, dyinv);
                    qDy[row + ix] = OR_FN(der1_)(qc, row + ix, strideY, dyinv);
                }
            }
#pragma omp for schedule(static)
        for (int iz = bord; iz < sz - bord; iz++)
            for (int iy = bord; iy < sy - bord; iy++)
                for (int ix = bord; ix < sx - bord; ix++) {
                    const long i = ((long)iz * sy + iy) * sx + ix;
                    const REAL pxy = OR_FN(der1_)(pDx, i, strideY, dyinv);
                    const REAL pxz = OR_FN(der1_)(pDx, i, strideZ, dzinv);
                    const REAL pyz = OR_FN(der1_)(pDy, i, strideZ, dzinv);
                    const REAL pxx = OR_FN(der2_)(pc, i, strideX, dxxinv);
                    const REAL pyy = OR_FN(der2_)(pc, i, strideY, dyyinv);
                    const REAL pzz = OR_FN(der2_)(pc, i, strideZ, dzzinv);
                    const REAL qxy = OR_FN(der1_)(qDx, i, strideY, dyinv);
                    const REAL qxz = OR_FN(der1_)(qDx, i, strideZ, dzinv);
                    const REAL qyz = OR_FN(der1_)(qDy, i, strideZ, dzinv);
                    const REAL qxx = OR_FN(der2_)(qc, i, strideX, dxxinv);
                    const REAL qyy = OR_FN(der2_)(qc, i, strideY, dyyinv);
                    const REAL qzz = OR_FN(der2_)(qc, i, strideZ, dzzinv);
                    OR_FN(update_cell_)(i, c, dt, pxx, pyy, pzz, pxy, pyz, pxz, qxx, qyy, qzz,
                                        qxy, qyz, qxz, pp, pc, qp, qc);
                }
    }
}

/* Pass-2 update of one plane iz of der1der1 from first-derivative planes reached through
 * per-plane base pointers: dx[k], dy[k] hold the (pDx|qDx, pDy|qDy) plane iz-4+k, indexed
 * by the in-plane offset iy*sx+ix (der1der1/sample.h:55-123; in der1der1hm the planes sit
 * in full volumes, in der1der1lm in the 9-plane ring of zTileSize.h:4-7).  The z
 * derivatives of the first-derivative planes keep the association order of Der1. */
static inline REAL OR_FN(der1_planes_)(const REAL *const *pl, long o, REAL dinv)
{
    return (L1 * (pl[5][o] - pl[3][o]) + L2 * (pl[6][o] - pl[2][o]) +
            L3 * (pl[7][o] - pl[1][o]) + L4 * (pl[8][o] - pl[0][o])) *
           dinv;
}

static void OR_FN(update_plane_)(int sx, int sy, int sz, int bord, int iz, REAL dxxinv, REAL dyyinv,
                                 REAL dzzinv, REAL dyinv, REAL dzinv, REAL dt, const REAL *const *c,
                                 REAL *pp, const REAL *pc, REAL *qp, const REAL *qc,
                                 const REAL *const *pdx, const REAL *const *pdy,
                                 const REAL *const *qdx, const REAL *const *qdy)
{
    const long strideX = 1, strideY = sx, strideZ = (long)sx * sy;
    (void)sz;
#pragma omp for schedule(static)
    for (int iy = bord; iy < sy - bord; iy++)
        for (int ix = bord; ix < sx - bord; ix++) {
            const long o = (long)iy * sx + ix, i = (long)iz * strideZ + o;
            const REAL pxy = OR_FN(der1_)(pdx[4], o, strideY, dyinv);
            const REAL pxz = OR_FN(der1_planes_)(pdx, o, dzinv);
            const REAL pyz = OR_FN(der1_planes_)(pdy, o, dzinv);
            const REAL pxx = OR_FN(der2_)(pc, i, strideX, dxxinv);
            const REAL pyy = OR_FN(der2_)(pc, i, strideY, dyyinv);
            const REAL pzz = OR_FN(der2_)(pc, i, strideZ, dzzinv);
            const REAL qxy = OR_FN(der1_)(qdx[4], o, strideY, dyinv);
            const REAL qxz = OR_FN(der1_planes_)(qdx, o, dzinv);
            const REAL qyz = OR_FN(der1_planes_)(qdy, o, dzinv);
            const REAL qxx = OR_FN(der2_)(qc, i, strideX, dxxinv);
            const REAL qyy = OR_FN(der2_)(qc, i, strideY, dyyinv);
            const REAL qzz = OR_FN(der2_)(qc, i, strideZ, dzzinv);
            OR_FN(update_cell_)(i, c, dt, pxx, pyy, pzz, pxy, pyz, pxz, qxx, qyy, qzz, qxy, qyz,
                                qxz, pp, pc, qp, qc);
        }
}

/* First derivatives of input plane fz into the planes dst (pDx, pDy, qDx, qDy in that
 * order), over the ranges of der1der1hm/OpenMP/openmp_propagate.c:24-76 and 88-130:
 * Dx rows iy in [0, sy), Dy rows iy in [bord, sy-bord), both ix in [bord, sx-bord).
 * Each loop is a work-shared loop of its own, as in the reference (nowait except the last). */
static void OR_FN(first_derivs_plane_)(int sx, int sy, int bord, int fz, REAL dxinv, REAL dyinv,
                                       const REAL *pc, const REAL *qc, REAL *const *dst)
{
    const long strideX = 1, strideY = sx, plane = (long)fz * sx * sy;
#pragma omp for schedule(static) nowait
    for (int iy = 0; iy < sy; iy++)
        for (int ix = bord; ix < sx - bord; ix++)
            dst[0][(long)iy * sx + ix] = OR_FN(der1_)(pc, plane + (long)iy * sx + ix, strideX, dxinv);
#pragma omp for schedule(static) nowait
    for (int iy = bord; iy < sy - bord; iy++)
        for (int ix = bord; ix < sx - bord; ix++)
            dst[1][(long)iy * sx + ix] = OR_FN(der1_)(pc, plane + (long)iy * sx + ix, strideY, dyinv);
#pragma omp for schedule(static) nowait
    for (int iy = 0; iy < sy; iy++)
        for (int ix = bord; ix < sx - bord; ix++)
            dst[2][(long)iy * sx + ix] = OR_FN(der1_)(qc, plane + (long)iy * sx + ix, strideX, dxinv);
#pragma omp for schedule(static)
    for (int iy = bord; iy < sy - bord; iy++)
        for (int ix = bord; ix < sx - bord; ix++)
            dst[3][(long)iy * sx + ix] = OR_FN(der1_)(qc, plane + (long)iy * sx + ix, strideY, dyinv);
}

/* OPENMP_PropagateDer1Der1HM — der1der1hm/OpenMP/openmp_propagate.c:9-160: the same
 * arithmetic as der1der1 with the first derivatives kept in full volumes but formed plane
 * by plane in z order: planes [0, 2 bord) first, then before updating plane iz the plane
 * iz + bord.  dv = pDx, pDy, qDx, qDy volumes (sx*sy*sz each). */
void OR_FN(oracle_propagate_der1der1hm_)(int sx, int sy, int sz, int bord, REAL dx, REAL dy,
                                         REAL dz, REAL dt, const REAL *const *c, REAL *pp,
                                         const REAL *pc, REAL *qp, const REAL *qc, REAL *const *dv)
{
    const long sxy = (long)sx * sy;
    const REAL dxxinv = (REAL)1.0 / (dx * dx);
    const REAL dyyinv = (REAL)1.0 / (dy * dy);
    const REAL dzzinv = (REAL)1.0 / (dz * dz);
    const REAL dxinv = (REAL)1.0 / dx;
    const REAL dyinv = (REAL)1.0 / dy;
    const REAL dzinv = (REAL)1.0 / dz;
#pragma omp parallel
    {
        for (int fz = 0; fz < 2 * bord; fz++) {
            REAL *dst[4] = {dv[0] + fz * sxy, dv[1] + fz * sxy, dv[2] + fz * sxy, dv[3] + fz * sxy};
            OR_FN(first_derivs_plane_)(sx, sy, bord, fz, dxinv, dyinv, pc, qc, dst);
        }
        for (int iz = bord; iz < sz - bord; iz++) {
            const int fz = iz + bord;
            REAL *dst[4] = {dv[0] + fz * sxy, dv[1] + fz * sxy, dv[2] + fz * sxy, dv[3] + fz * sxy};
            OR_FN(first_derivs_plane_)(sx, sy, bord, fz, dxinv, dyinv, pc, qc, dst);
            const REAL *pl[4][9];
            for (int k = 0; k < 9; k++)
                for (int v = 0; v < 4; v++) pl[v][k] = dv[v] + (long)(iz - 4 + k) * sxy;
            OR_FN(update_plane_)(sx, sy, sz, bord, iz, dxxinv, dyyinv, dzzinv, dyinv, dzinv, dt, c,
                                 pp, pc, qp, qc, pl[0], pl[1], pl[2], pl[3]);
        }
    }
}

/* OPENMP_PropagateDer1Der1LM — der1der1lm/OpenMP/openmp_propagate.c:9-156 with
 * der1der1lm/sample.h:35-160 and zTileSize.h:4-7: the first derivatives live in a ring of
 * 2*bord+1 = 9 planes; input plane f sits in ring slot f % 9.  ring = pDx, pDy, qDx, qDy
 * rings (sx*sy*9 each). */
void OR_FN(oracle_propagate_der1der1lm_)(int sx, int sy, int sz, int bord, REAL dx, REAL dy,
                                         REAL dz, REAL dt, const REAL *const *c, REAL *pp,
                                         const REAL *pc, REAL *qp, const REAL *qc, REAL *const *ring)
{
    const long sxy = (long)sx * sy;
    const int nr = 2 * bord + 1;
    const REAL dxxinv = (REAL)1.0 / (dx * dx);
    const REAL dyyinv = (REAL)1.0 / (dy * dy);
    const REAL dzzinv = (REAL)1.0 / (dz * dz);
    const REAL dxinv = (REAL)1.0 / dx;
    const REAL dyinv = (REAL)1.0 / dy;
    const REAL dzinv = (REAL)1.0 / dz;
#pragma omp parallel
    {
        for (int fz = 0; fz < 2 * bord; fz++) {
            REAL *dst[4] = {ring[0] + (fz % nr) * sxy, ring[1] + (fz % nr) * sxy,
                            ring[2] + (fz % nr) * sxy, ring[3] + (fz % nr) * sxy};
            OR_FN(first_derivs_plane_)(sx, sy, bord, fz, dxinv, dyinv, pc, qc, dst);
        }
        for (int iz = bord; iz < sz - bord; iz++) {
            const int fz = iz + bord;
            REAL *dst[4] = {ring[0] + (fz % nr) * sxy, ring[1] + (fz % nr) * sxy,
                            ring[2] + (fz % nr) * sxy, ring[3] + (fz % nr) * sxy};
            OR_FN(first_derivs_plane_)(sx, sy, bord, fz, dxinv, dyinv, pc, qc, dst);
            const REAL *pl[4][9];
            for (int k = 0; k < 9; k++)
                for (int v = 0; v < 4; v++) pl[v][k] = ring[v] + (long)((iz - 4 + k) % nr) * sxy;
            OR_FN(update_plane_)(sx, sy, sz, bord, iz, dxxinv, dyyinv, dzzinv, dyinv, dzinv, dt, c,
                                 pp, pc, qp, qc, pl[0], pl[1], pl[2], pl[3]);
        }
    }
}

/* Time loop — original/model.c:71-88 (+ OpenMP/openmp_insertsource.c:7-12, utils.c:233-243).
 * For it=1..st: src=Source(dt,it-1) added to pc and qc at iSource, Propagate, swap.
 * src[] holds the FP32 Ricker samples (src[k] = Source(dt,k)), promoted for FP64.
 * On return pc/qc hold the field at t=st*dt and pp/qp the field one step earlier.
 * Traces (no reference counterpart): after step it with it%rec_every==0, pc on plane
 * rec_iz over [bord,sx-bord)x[bord,sy-bord) is appended to traces. Returns #traces. */
/* variant: 0 original, 1 der1der1, 2 der1der1hm, 3 der1der1lm */
int OR_FN(oracle_run_)(int variant, int sx, int sy, int sz, int bord, REAL dx, REAL dy, REAL dz,
                       REAL dt, int st, long iSource, const float *src, const REAL *const *c,
                       REAL *pp, REAL *pc, REAL *qp, REAL *qc, int rec_iz, int rec_every,
                       REAL *traces)
{
    const long n = (long)sx * sy * sz;
    /* first-derivative storage: 4 volumes (der1der1, der1der1hm) or 4 rings of 2*bord+1
     * planes (der1der1lm) */
    const long nd = variant == 3 ? (long)sx * sy * (2 * bord + 1) : n;
    REAL *scratch = NULL;
    if (variant >= 1 && variant <= 3) {
        scratch = (REAL *)calloc((size_t)4 * nd, sizeof(REAL));
        if (!scratch) return -1;
    }
    REAL *dv[4] = {scratch, scratch ? scratch + nd : NULL, scratch ? scratch + 2 * nd : NULL,
                   scratch ? scratch + 3 * nd : NULL};
    REAL *a_pp = pp, *a_pc = pc, *a_qp = qp, *a_qc = qc;
    const int nxr = sx - 2 * bord, nyr = sy - 2 * bord;
    int ntr = 0;
    for (int it = 1; it <= st; it++) {
        const REAL s = (REAL)src[it - 1];
        a_pc[iSource] += s;
        a_qc[iSource] += s;
        if (variant == 1)
            OR_FN(oracle_propagate_der1der1_)(sx, sy, sz, bord, dx, dy, dz, dt, c, a_pp, a_pc,
                                              a_qp, a_qc, dv[0], dv[1], dv[2], dv[3]);
        else if (variant == 2)
            OR_FN(oracle_propagate_der1der1hm_)(sx, sy, sz, bord, dx, dy, dz, dt, c, a_pp, a_pc,
                                                a_qp, a_qc, dv);
        else if (variant == 3)
            OR_FN(oracle_propagate_der1der1lm_)(sx, sy, sz, bord, dx, dy, dz, dt, c, a_pp, a_pc,
                                                a_qp, a_qc, dv);
        else
            OR_FN(oracle_propagate_original_)(sx, sy, sz, bord, dx, dy, dz, dt, c, a_pp, a_pc,
                                              a_qp, a_qc);
        REAL *t = a_pp; a_pp = a_pc; a_pc = t;
        t = a_qp; a_qp = a_qc; a_qc = t;
        if (traces && rec_every > 0 && it % rec_every == 0) {
            REAL *dst = traces + (long)ntr * nxr * nyr;
            for (int iy = 0; iy < nyr; iy++)
                memcpy(dst + (long)iy * nxr, a_pc + ((long)rec_iz * sy + iy + bord) * sx + bord,
                       sizeof(REAL) * nxr);
            ntr++;
        }
    }
    if (a_pc != pc) { /* odd number of swaps: bring the current fields back into pc/qc */
        for (long k = 0; k < n; k++) {
            REAL t = pc[k]; pc[k] = pp[k]; pp[k] = t;
            t = qc[k]; qc[k] = qp[k]; qp[k] = t;
        }
    }
    free(scratch);
    return ntr;
}

#undef OR_FN
#undef OR_CAT
#undef OR_CAT2
