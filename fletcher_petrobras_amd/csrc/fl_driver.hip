// HIP back-end of the Fletcher TTI path behind the reference driver.h API
// (original/driver.h:8-22) plus the fl_hip.h extensions.  One process-wide
// single-GPU state, like the reference back-ends' device singletons
// (original/CUDA/cuda_driver.c:11-34); every entry point reports failures
// through fl_last_error() instead of exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../host/fl_cell.h"
#include "fl_launch.h"
#include "fletcher/driver.h"
#include "fletcher/fl_hip.h"
#include "fletcher/fl_model.h"

using namespace fl;

extern "C" int fl_autotune(int steps, const char *cache_path);

// ------------------------------------------------------------------ errors
static char g_err[512] = "";
static int g_status = FL_OK;

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    g_status = code;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(FL_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

extern "C" const char *fl_last_error(void) { return g_err; }
extern "C" int fl_status(void) { return g_status; }
extern "C" void fl_clear_error(void)
{
    g_err[0] = 0;
    g_status = FL_OK;
}

// ------------------------------------------------------------------ helpers
// ------------------------------------------------------------------ state
struct Tiles {
    int ty = 12, zt = 0;  // zt 0: chunk_depth's choice for the grid
    int hints = 0;        // streaming cache hints of the z-marching step (fl_set_stream_hints)
    int sw = 0;           // tile columns per strip (fl_set_tile_strip); 0: strip_width's choice
    int pair = 1;         // both fields per lane for the forms that have it (fl_set_lane_pairs)
};

struct Device {
    bool init = false;
    int sx = 0, sy = 0, sz = 0, bord = 0, px = 0;
    float dx = 0, dy = 0, dz = 0, dt = 0;
    long long n = 0;        // cells of the reference layout, sx*sy*sz
    long long np = 0;       // cells of the pitched device layout, px*sy*sz
    float *lv[2] = {nullptr, nullptr};  // time levels of (p, q) pairs
    int lc = 1;             // level holding pc, qc; lv[1 - lc] holds pp, qp
    void *coef = nullptr;   // coefficient records, pitched like the levels
    int cm = CM_F32;
    float vq = 0;           // CM_Q16R: 2^ebase
    float *dvx = nullptr, *dvy = nullptr;  // two-pass (pDx, qDx), (pDy, qDy) pairs
    float *trace = nullptr;
    int rec_iz = -1, rec_every = 0, rec_slots = 0;
    int *run = nullptr;     // control block of multi-step runs (RUN_WORDS ints)
    float *src_tab = nullptr;  // source samples of the current multi-step run
    int src_cap = 0;
    long long bytes = 0;
};

static Device D;
static hipStream_t g_stream = nullptr;
static int g_variant = FL_VARIANT_FUSED;
static int g_precision = FL_PREC_FP32;
static bool g_sync = true;
static Tiles g_tiles;
static bool g_timing = false;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_ev_used, g_ev_free;
static hipEvent_t g_reg0 = nullptr, g_reg1 = nullptr;
// tile width of the step kernel (columns per wave and field)
static int tile_x() { return 32; }
// First tile column at x = 0, so that every tile row starts on a 128-byte line, or at x =
// bord when that needs one tile column fewer over the update domain [bord, sx - bord) and
// the column saved is at least an eighth of them.  Misaligned rows cost more than a
// partly idle last column on wide grids: the FP32 step on the 1096^2 x 840 array ran 6.4%
// faster from x = 0 (35 columns) than from x = 4 (34), and took the streaming hints there.
static int tile_xorg(int sx, int bord)
{
    const int t = tile_x();
    const int n0 = (sx - bord + t - 1) / t, nb = (sx - 2 * bord + t - 1) / t;
    return (nb < n0 && n0 <= 8) ? bord : 0;
}
static int tiles_x(int sx, int bord) { return (sx - bord - tile_xorg(sx, bord) + tile_x() - 1) / tile_x(); }

// Row pitch of the device layout: 16 cells = one 128-byte line of (p, q) pairs.
static int pitch_of(int sx) { return (sx + 15) / 16 * 16; }

// Deepest z chunk whose level span (zt + 8 pitched planes of pairs) fits a buffer range
// (< 2^31 bytes); 0 when not even one chunk does or a record plane does not.
static int max_zt(int px, int sy)
{
    const long long pb = (long long)px * sy * 8;
    if ((long long)px * sy * 40 >= 0x7fffffffll) return 0;
    return (int)std::max<long long>(0, (0x7fffffffll / pb) - 8);
}

// Reference-layout index -> pitched cell index.
static long long pitched_index(long long i, int sx, int sy, int px)
{
    const long long sxy = (long long)sx * sy;
    const long long iz = i / sxy, r = i - iz * sxy, iy = r / sx, ix = r - iy * sx;
    return (iz * sy + iy) * px + ix;
}

static int g_cus = 0;  // compute units of the device

static int ensure_stream()
{
    if (g_stream) return FL_OK;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(FL_ERR_HIP, "no HIP device visible");
    HIPCHK(hipSetDevice(0));
    HIPCHK(hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, 0));
    // default flags: the stream also orders after legacy null-stream work (hipMemcpy etc.)
    HIPCHK(hipStreamCreate(&g_stream));
    return FL_OK;
}

static int dev_alloc(void **p, size_t bytes, long long *acct)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(FL_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    *acct += (long long)bytes;
    return FL_OK;
}

static int coef_count(int cm) { return (cm == CM_F16R || cm == CM_F32R || cm == CM_Q16R) ? 6 : 10; }
static int coef_elem(int cm) { return (cm == CM_F32 || cm == CM_F32R) ? 4 : 2; }

extern "C" int fl_bytes_per_cell(void)
{
    const int cm = D.init ? D.cm : g_precision;
    const int c = coef_count(cm) * coef_elem(cm);
    if (g_variant == FL_VARIANT_TWOPASS) return 24 + c + 40;  // + write/read 4 volumes, re-read pc,qc
    return 24 + c;
}

// Host precompute into the device coefficient records: per cell, its coef_count(cm)
// values (FP32 or the 16-bit forms, or the recompute inputs) in the slot order
// rec_order32 / rec_order16 / rec_order_r32 / rec_order_r16 (record of
// coef_count * coef_elem bytes; FP32 in its two arrays, slots 0-3, 5-8 and slots 4, 9);
// z slabs of at most 64 MiB per array on the host.
// ebase: the CM_Q16R exponent base (fl_recompute_q16_base over the whole grid).
static int upload_coefs(const float *vpz, const float *vsv, const float *eps, const float *del,
                        const float *phi, const float *theta, float dt, int cm, void *dst, int sx,
                        int sy, int sz, int px, int ebase = 0)
{
    const long long sxy = (long long)sx * sy;
    const int slab = std::max(1, (int)std::min<long long>(sz, (64ll << 20) / (sxy * 4)));
    const int nk = coef_count(cm);
    const size_t rec = (size_t)nk * coef_elem(cm);
    HIPCHK(hipMemset(dst, 0, rec * px * sy * sz));  // pitch padding
    std::vector<float> f32, r32;
    std::vector<uint16_t> h16, r16;
    for (int z0 = 0; z0 < sz; z0 += slab) {
        const int z1 = std::min(sz, z0 + slab);
        const long long off = (long long)z0 * sxy, m = (long long)(z1 - z0) * sxy;
        if (cm == CM_F32 || cm == CM_F32R) {
            f32.resize((size_t)nk * m);
            r32.resize((size_t)nk * m);
            float *c[10];
            for (int k = 0; k < nk; k++) c[k] = f32.data() + k * m;
            if (cm == CM_F32)
                fl_precompute_f32(m, vpz + off, vsv + off, eps + off, del + off, phi + off, theta + off, c);
            else
                fl_pack_recompute32(m, vpz + off, vsv + off, eps + off, del + off, phi + off, theta + off, c);
            const float *co[10];
            for (int k = 0; k < nk; k++) co[k] = c[cm == CM_F32 ? rec_order32(k) : rec_order_r32(k)];
            if (cm == CM_F32) {  // slots 0-3, 5-8 at 32 B a cell, then slots 4, 9 at 8 B a cell
                const float *c1[8] = {co[0], co[1], co[2], co[3], co[5], co[6], co[7], co[8]};
                const float *c2[2] = {co[4], co[9]};
                const long long np = (long long)px * sy * sz;
                fl_interleave32(m, 8, c1, r32.data());
                HIPCHK(hipMemcpy2D((char *)dst + (size_t)z0 * sy * px * 32, px * 32, r32.data(), sx * 32, sx * 32,
                                   (size_t)sy * (z1 - z0), hipMemcpyHostToDevice));
                fl_interleave32(m, 2, c2, r32.data());
                HIPCHK(hipMemcpy2D((char *)dst + fp32_part2(np) + (size_t)z0 * sy * px * 8, px * 8, r32.data(),
                                   sx * 8, sx * 8, (size_t)sy * (z1 - z0), hipMemcpyHostToDevice));
            } else {
                fl_interleave32(m, nk, co, r32.data());
                HIPCHK(hipMemcpy2D((char *)dst + (size_t)z0 * sy * px * rec, px * rec, r32.data(), sx * rec,
                                   sx * rec, (size_t)sy * (z1 - z0), hipMemcpyHostToDevice));
            }
        } else {
            h16.resize((size_t)nk * m);
            r16.resize((size_t)nk * m);
            uint16_t *c[10];
            for (int k = 0; k < nk; k++) c[k] = h16.data() + k * m;
            if (cm == CM_F16R)
                fl_pack_recompute16(m, vpz + off, vsv + off, eps + off, del + off, phi + off, theta + off, dt, c);
            else if (cm == CM_Q16R)
                fl_pack_recompute_q16(m, vpz + off, vsv + off, eps + off, del + off, phi + off, theta + off, dt,
                                      ebase, c);
            else
                fl_pack_coefs16(m, vpz + off, vsv + off, eps + off, del + off, phi + off, theta + off, dt,
                                cm == CM_BF16, c);
            if (cm == CM_F16 || cm == CM_F16R) {
                // FP16 velocity factors v^2 dt^2 (coefficients 6-9, or recompute inputs 0-3)
                // overflow to infinity at 65520: v dt above about 256 m
                const int k0 = cm == CM_F16 ? 6 : 0;
                for (int k = k0; k < k0 + 4; k++)
                    for (long long i = 0; i < m; i++)
                        if ((c[k][i] & 0x7c00u) == 0x7c00u)
                            return fail(FL_ERR_ARG,
                                        "velocity factor v^2 dt^2 of cell %lld exceeds the FP16 range "
                                        "(v dt > 256 m): use FP32, BF16 or block floating-point storage",
                                        (long long)off + i);
            }
            const uint16_t *co[10];
            for (int k = 0; k < nk; k++)
                co[k] = c[(cm == CM_F16R || cm == CM_Q16R) ? rec_order_r16(k) : rec_order16(k)];
            fl_interleave16(m, nk, co, r16.data());
            HIPCHK(hipMemcpy2D((char *)dst + (size_t)z0 * sy * px * rec, px * rec, r16.data(), sx * rec,
                               sx * rec, (size_t)sy * (z1 - z0), hipMemcpyHostToDevice));
        }
    }
    return FL_OK;
}

// ------------------------------------------------------------------ device precompute
// The coefficient records computed on the device in FP64 from the six FP32 model volumes
// (pitched like the levels), with the per-cell code of the host path (fl_cell.h): the same
// records wherever the device's double sine and cosine equal the host library's.
struct ModelVols {
    const float *m[6];  // vpz, vsv, epsilon, delta, phi, theta
};

// Largest velocity factor need of the grid (the block floating-point exponent base).
__global__ __launch_bounds__(256) void q16_need_max(ModelVols v, long long np, int px, int sx, double dt2,
                                                    unsigned long long *out)
{
    __shared__ double red[256];
    double m = 0.0;
    for (long long c = blockIdx.x * 256ll + threadIdx.x; c < np; c += (long long)gridDim.x * 256) {
        if ((int)(c % px) >= sx) continue;
        const double need = fl_cell_q16_need(v.m[0][c], v.m[1][c], v.m[2][c], v.m[3][c], dt2);
        if (need > m) m = need;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    // non-negative doubles order as their bit patterns
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

// Records of every pitched cell (zero in the row padding); *bad receives 1 + the index of a
// cell whose FP16 velocity factor overflows.
template <int CM>
__global__ __launch_bounds__(256) void precompute_records(ModelVols v, long long np, int px, int sx, double dt,
                                                          int ebase, void *rec, unsigned long long *bad)
{
    for (long long c = blockIdx.x * 256ll + threadIdx.x; c < np; c += (long long)gridDim.x * 256) {
        const bool pad = (int)(c % px) >= sx;
        const float vp = pad ? 0.f : v.m[0][c], vs = pad ? 0.f : v.m[1][c], e = pad ? 0.f : v.m[2][c];
        const float d = pad ? 0.f : v.m[3][c], ph = pad ? 0.f : v.m[4][c], th = pad ? 0.f : v.m[5][c];
        if constexpr (CM == CM_F32) {  // the two arrays: slots 0-3, 5-8 and slots 4, 9
            float k[10];
            fl_cell_f32(vp, vs, e, d, ph, th, k);
            float *r1 = (float *)rec + c * 8;
            float *r2 = (float *)((char *)rec + fp32_part2(np)) + c * 2;
            for (int s = 0; s < 10; s++) {
                const float v = pad ? 0.f : k[rec_order32(s)];
                if (s == 4 || s == 9) r2[s == 9] = v;
                else r1[s < 4 ? s : s - 1] = v;
            }
        } else if constexpr (CM == CM_F32R) {
            float k[6];
            fl_cell_recompute32(vp, vs, e, d, ph, th, k);
            float *r = (float *)rec + c * 6;
            for (int s = 0; s < 6; s++) r[s] = pad ? 0.f : k[rec_order_r32(s)];
        } else if constexpr (CM == CM_BF16 || CM == CM_F16) {
            uint16_t k[10];
            fl_cell_pack16(vp, vs, e, d, ph, th, dt, CM == CM_BF16, k);
            uint16_t *r = (uint16_t *)rec + c * 10;
            for (int s = 0; s < 10; s++) r[s] = pad ? 0 : k[rec_order16(s)];
            if (CM == CM_F16 && !pad)
                for (int s = 6; s < 10; s++)
                    if ((k[s] & 0x7c00u) == 0x7c00u) atomicMax(bad, (unsigned long long)c + 1);
        } else {
            uint16_t k[6];
            if constexpr (CM == CM_F16R) fl_cell_recompute16(vp, vs, e, d, ph, th, dt, k);
            else fl_cell_recompute_q16(vp, vs, e, d, ph, th, dt, ebase, k);
            uint16_t *r = (uint16_t *)rec + c * 6;
            for (int s = 0; s < 6; s++) r[s] = pad ? 0 : k[rec_order_r16(s)];
            if (CM == CM_F16R && !pad)
                for (int s = 0; s < 4; s++)
                    if ((k[s] & 0x7c00u) == 0x7c00u) atomicMax(bad, (unsigned long long)c + 1);
        }
    }
}

// Where DRIVER_Initialize and fl_initialize_streamed take the model from: six whole host
// arrays (reference layout), or slabs from a callback.
struct ModelSrc {
    const float *a[6];
    fl_model_slab_fn fill;
    void *user;
    int sx, sy;
};

static int model_slab(const ModelSrc &src, int z0, int nzs, float *const *out)
{
    if (src.fill) return src.fill(src.user, z0, nzs, out[0], out[1], out[2], out[3], out[4], out[5]);
    const size_t plane = (size_t)src.sx * src.sy;
    for (int k = 0; k < 6; k++) memcpy(out[k], src.a[k] + (size_t)z0 * plane, plane * nzs * 4);
    return 0;
}

static bool g_device_precompute = false;

extern "C" int fl_set_device_precompute(int on)
{
    g_device_precompute = on != 0;
    return FL_OK;
}

// The model streamed to the device in z slabs (pinned staging of at most 6 x 64 MiB), the
// records of form cm computed there into dst; *vq receives 2^ebase for CM_Q16R.
static int device_records(const ModelSrc &src, int cm, int sx, int sy, int sz, int px, float dt, void *dst, float *vq)
{
    const long long np = (long long)px * sy * sz;
    const long long plane = (long long)sx * sy;
    const int slab = (int)std::max<long long>(1, std::min<long long>(sz, (64ll << 20) / (plane * 4)));
    float *dv[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float *stage = nullptr;
    unsigned long long *scal = nullptr;  // [0] q16 need, [1] overflow cell + 1
    long long dummy = 0;
    int rc = FL_OK;
    for (int k = 0; k < 6 && rc == FL_OK; k++) rc = dev_alloc((void **)&dv[k], (size_t)np * 4, &dummy);
    if (rc == FL_OK) rc = dev_alloc((void **)&scal, 2 * sizeof(unsigned long long), &dummy);
    if (rc == FL_OK && hipHostMalloc((void **)&stage, (size_t)6 * slab * plane * 4, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        stage = nullptr;
        rc = fail(FL_ERR_NOMEM, "pinned staging of %lld bytes failed", 6ll * slab * plane * 4);
    }
    hipError_t e = hipSuccess;
    for (int k = 0; k < 6 && rc == FL_OK && e == hipSuccess; k++) e = hipMemsetAsync(dv[k], 0, (size_t)np * 4, g_stream);
    if (rc == FL_OK && e == hipSuccess) e = hipMemsetAsync(scal, 0, 2 * sizeof(unsigned long long), g_stream);
    float *h[6];
    for (int k = 0; k < 6; k++) h[k] = stage + (size_t)k * slab * plane;
    for (int z0 = 0; z0 < sz && rc == FL_OK && e == hipSuccess; z0 += slab) {
        const int n = std::min(slab, sz - z0);
        e = hipStreamSynchronize(g_stream);  // the staging buffer is free again
        if (e != hipSuccess) break;
        if (model_slab(src, z0, n, h)) {
            rc = fail(FL_ERR_ARG, "the model source failed for planes %d..%d", z0, z0 + n - 1);
            break;
        }
        for (int k = 0; k < 6 && e == hipSuccess; k++)
            e = hipMemcpy2DAsync(dv[k] + (size_t)z0 * sy * px, (size_t)px * 4, h[k], (size_t)sx * 4, (size_t)sx * 4,
                                 (size_t)sy * n, hipMemcpyHostToDevice, g_stream);
    }
    ModelVols mv;
    for (int k = 0; k < 6; k++) mv.m[k] = dv[k];
    const int grid = (int)std::min<long long>((np + 255) / 256, 65536);
    int ebase = 0;
    if (rc == FL_OK && e == hipSuccess && cm == CM_Q16R) {
        hipLaunchKernelGGL(q16_need_max, dim3(grid), dim3(256), 0, g_stream, mv, np, px, sx, (double)dt * dt, scal);
        unsigned long long bits = 0;
        e = hipMemcpyAsync(&bits, scal, 8, hipMemcpyDeviceToHost, g_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
        double m;
        memcpy(&m, &bits, 8);
        ebase = fl_cell_q16_base_of(m);
        *vq = std::ldexp(1.0f, ebase);
    }
    if (rc == FL_OK && e == hipSuccess) {
        switch (cm) {
#define FL_PRE(CMV) \
    case CMV: hipLaunchKernelGGL(precompute_records<CMV>, dim3(grid), dim3(256), 0, g_stream, mv, np, px, sx, (double)dt, ebase, dst, scal + 1); break;
            FL_PRE(CM_F32) FL_PRE(CM_BF16) FL_PRE(CM_F16) FL_PRE(CM_F16R) FL_PRE(CM_F32R) FL_PRE(CM_Q16R)
#undef FL_PRE
        }
        unsigned long long badc = 0;
        e = hipMemcpyAsync(&badc, scal + 1, 8, hipMemcpyDeviceToHost, g_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
        if (e == hipSuccess && badc)
            rc = fail(FL_ERR_ARG,
                      "velocity factor v^2 dt^2 of pitched cell %llu exceeds the FP16 range (v dt > 256 m): "
                      "use FP32, BF16 or block floating-point storage", badc - 1);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    for (int k = 0; k < 6; k++)
        if (dv[k]) (void)hipFree(dv[k]);
    if (scal) (void)hipFree(scal);
    if (stage) (void)hipHostFree(stage);
    if (rc == FL_OK && e != hipSuccess) rc = fail(FL_ERR_HIP, "device precompute: %s", hipGetErrorString(e));
    return rc;
}

// ------------------------------------------------------------------ launch plumbing
// Workgroups of the step kernel resident on one CU: the LDS rings (6 staged planes and 5
// Dx planes of ty+8 rows) against 160 KB, and the waves per SIMD the kernel is built for
// (step_waves: 4 for FP32, 6 for the other forms with 8- or 12-row tiles).
// Workgroups of the step kernel resident on one CU: limited by their LDS and by the waves
// per SIMD their registers allow (step_waves, pair_waves in the kernels).
static int resident_per_cu(int ty, int cm, bool pair)
{
    if (pair) {  // use_pair(cm, ty, ..): StepLds2
        const long long lds = (2ll * (ty + 8) * 40 + 10ll * (ty + 8) * 32) * 8;
        return std::max(1, (int)std::min<long long>((160ll << 10) / lds, 4 * 3 / (ty / 2)));  // 3 for ty 8
    }
    const long long lds = (6ll * (ty + 8) * 80 + 5ll * (ty + 8) * 64) * 4;
    const int waves = (cm != CM_F32 && (ty == 8 || ty == 12)) ? 6 : 4;
    return std::max(1, (int)std::min<long long>((160ll << 10) / lds, 4 * waves / ty));
}

// The both-fields-per-lane kernel runs a form when it has one for the tile height and the
// knob asks for it.
static bool use_pair(int cm, int ty, bool twopass)
{
    return g_tiles.pair != 0 && has_pair_form(cm) && has_pair_tiles(ty) && !twopass;
}

// z chunk depth for `tiles` tile columns per plane over nzu planes: the chunk count that
// minimises (rounds of resident workgroups) x (planes a workgroup streams, zt + 8).  At
// 384^3 it gives 32x12x188 for FP32 (768 tiles, three rounds of 256) and 32x12x94 for the
// other forms (1536 tiles, three rounds of 512), the autotuner's choices.
static int chunk_depth(long long tiles, int nzu, int ty, int cm, int ztmax)
{
    const long long slots = (long long)std::max(g_cus, 1) * resident_per_cu(ty, cm, use_pair(cm, ty, false));
    double best = 1e300;
    int bz = nzu;
    for (int nzc = 1; nzc <= 64; nzc++) {
        const int zt = (nzu + nzc - 1) / nzc;
        if (zt < 8 && nzc > 1) break;
        const int n = (nzu + zt - 1) / zt;
        const long long rounds = (tiles * n + slots - 1) / slots;
        const double cost = (double)rounds * (zt + 8);
        if (cost < best * 0.999) { best = cost; bz = zt; }
    }
    return std::max(1, std::min(std::min(bz, nzu), ztmax));
}

static int default_zt(int sx, int sy, int sz, int bord, int ty)
{
    const long long tiles = (long long)tiles_x(sx, bord) * ((sy - 2 * bord + ty - 1) / ty);
    const int cm = g_variant == FL_VARIANT_FUSED ? D.cm : CM_F32;
    return chunk_depth(tiles, sz - 2 * bord, ty, cm, max_zt(pitch_of(sx), sy));
}

// Tile columns per strip of the workgroup order (tile_of in the kernels).  Tiles that run
// at the same time on one XCD share their halo columns and rows through its L2; a halo
// shared with a tile that ran earlier comes from HBM again, a whole 128-byte line for each
// 4-cell x halo.  Row-major order puts part of one tile row on an XCD when the plane is
// wide; strips of 8 columns keep a block of 8 x 4..12 tiles there.  Timed on MI355X for FP32
// (12-row tiles) and the 16-bit recompute forms (8-row tiles) over 384^2 .. 1600^2 planes
// with widths 4 .. 16: 8 is within run-to-run noise of the fastest width on every grid and
// 7-13% faster than row-major order on planes of 1096 cells and more.
static int strip_width(int ntx) { return std::min(ntx, 8); }

static void fill_geom(StepArgs &a, int sx, int sy, int sz, int bord, int ty, int zt)
{
    a.sx = sx; a.sy = sy; a.sz = sz; a.bord = bord;
    a.xorg = tile_xorg(sx, bord);  // tiles [xorg, xorg+TX), ... up to sx-bord
    a.ntx = tiles_x(sx, bord);
    a.nty = (sy - 2 * bord + ty - 1) / ty;
    a.zt = zt;
    const int nzc = (sz - 2 * bord + zt - 1) / zt;
    a.nblk = a.ntx * a.nty * nzc;
}

// Equal spacings select the kernels that scale the right-hand side once (ISO).
static bool iso_spacing(float dx, float dy, float dz) { return dx == dy && dy == dz; }

static int launch_steps(const StepArgs &a, int ty, int cm, bool tp, hipStream_t s)
{
    if (launch_grid(ty, a, cm, tp, iso_spacing(D.dx, D.dy, D.dz), g_tiles.hints != 0, use_pair(cm, ty, tp), s) == 0)
        return FL_OK;
    return fail(FL_ERR_ARG, "unsupported tile rows %d / coefficient mode %d", ty, cm);
}

static StepArgs base_args()
{
    StepArgs a;
    memset(&a, 0, sizeof(a));
    a.prev = D.lv[1 - D.lc];
    a.cur = D.lv[D.lc];
    a.coef = D.coef;
    a.dvx = D.dvx;
    a.dvy = D.dvy;
    a.px = D.px;
    a.dxxinv = 1.0f / (D.dx * D.dx);
    a.dyyinv = 1.0f / (D.dy * D.dy);
    a.dzzinv = 1.0f / (D.dz * D.dz);
    a.dxinv = 1.0f / D.dx;
    a.dyinv = 1.0f / D.dy;
    a.dzinv = 1.0f / D.dz;
    a.dt = D.dt;
    a.vq = D.vq;
    const int zt = g_tiles.zt > 0 ? std::min(std::min(g_tiles.zt, D.sz - 2 * D.bord), max_zt(D.px, D.sy))
                                  : default_zt(D.sx, D.sy, D.sz, D.bord, g_tiles.ty);
    fill_geom(a, D.sx, D.sy, D.sz, D.bord, g_tiles.ty, zt);
    a.sw = g_tiles.sw > 0 ? std::min(g_tiles.sw, a.ntx) : strip_width(a.ntx);
    a.src_idx = -1;
    a.trace = D.trace;
    a.rec_iz = D.rec_iz;
    a.rec_every = D.rec_every;
    a.rec_slots = D.rec_slots;
    return a;
}

static void swap_levels() { D.lc ^= 1; }

static int ensure_dvol()
{
    if (D.dvx) return FL_OK;
    int rc = dev_alloc((void **)&D.dvx, (size_t)2 * D.np * 4, &D.bytes);
    if (!rc) rc = dev_alloc((void **)&D.dvy, (size_t)2 * D.np * 4, &D.bytes);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(D.dvx, 0, (size_t)2 * D.np * 4, g_stream));
    HIPCHK(hipMemsetAsync(D.dvy, 0, (size_t)2 * D.np * 4, g_stream));
    return FL_OK;
}

// Staging buffer of host transfers: z slabs of one field, at most 64 MiB, kept between
// transfers (released by DRIVER_Finalize and fl_batch_destroy).
static float *g_xfer = nullptr;
static size_t g_xfer_bytes = 0;

static void free_xfer()
{
    if (g_xfer) (void)hipFree(g_xfer);
    g_xfer = nullptr;
    g_xfer_bytes = 0;
}

// One field (fld 0 p, 1 q) of a level from / to a reference-layout host volume, slab by
// slab through the staging buffer.
static int level_xfer(float *lv, int fld, float *host, bool put, int sx, int sy, int sz, int px)
{
    const long long plane = (long long)sx * sy;
    const int zs = (int)std::max<long long>(1, std::min<long long>(sz, (64ll << 20) / (plane * 4)));
    const size_t need = (size_t)zs * plane * 4;
    if (g_xfer_bytes < need) {
        free_xfer();
        long long dummy = 0;
        int rc = dev_alloc((void **)&g_xfer, need, &dummy);
        if (rc) return rc;
        g_xfer_bytes = need;
    }
    hipError_t e = hipStreamSynchronize(g_stream);
    for (int z0 = 0; z0 < sz && e == hipSuccess; z0 += zs) {
        const int nz = std::min(zs, sz - z0);
        float *slab = lv + 2ll * z0 * sy * px;
        float *h = host + z0 * plane;
        const size_t bytes = (size_t)nz * plane * 4;
        const dim3 g((sx + 255) / 256, sy, nz);
        if (put) {
            e = hipMemcpy(g_xfer, h, bytes, hipMemcpyHostToDevice);
            if (e == hipSuccess) hipLaunchKernelGGL(field_pack, g, dim3(256), 0, g_stream, g_xfer, slab, fld, sx, sy, px);
            if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
        } else {
            hipLaunchKernelGGL(field_unpack, g, dim3(256), 0, g_stream, slab, g_xfer, fld, sx, sy, px);
            e = hipStreamSynchronize(g_stream);
            if (e == hipSuccess) e = hipMemcpy(h, g_xfer, bytes, hipMemcpyDeviceToHost);
        }
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(FL_ERR_HIP, "field transfer: %s", hipGetErrorString(e));
    return FL_OK;
}

// One step on the current levels (no swap).  it_host = step index for traces.
static int launch_step(StepArgs a, int it_host, hipStream_t s)
{
    a.it_host = it_host;
    if (g_variant == FL_VARIANT_FUSED) {
        return launch_steps(a, g_tiles.ty, D.cm, false, s);
    }
    if (D.cm != CM_F32)
        return fail(FL_ERR_ARG, "variant %d needs FP32 coefficients", g_variant);
    if (g_variant == FL_VARIANT_TWOPASS) {
        int rc = ensure_dvol();
        if (rc) return rc;
        a.dvx = D.dvx;
        a.dvy = D.dvy;
        dim3 g1((D.sx + 63) / 64, (D.sy + 3) / 4, D.sz);
        hipLaunchKernelGGL(der1_pass, g1, dim3(64, 4), 0, s, a.cur, D.dvx, D.dvy, D.sx, D.sy, D.px,
                           D.bord, a.dxinv, a.dyinv);
        return launch_steps(a, g_tiles.ty, CM_F32, true, s);
    }
    if (g_variant == FL_VARIANT_DIRECT) {
        DirectArgs d;
        d.prev = a.prev;
        d.cur = a.cur;
        d.coef = (const float *)D.coef;
        d.np = D.np;
        d.px = D.px;
        d.dxxinv = a.dxxinv; d.dyyinv = a.dyyinv; d.dzzinv = a.dzzinv;
        d.dxyinv = 1.0f / (D.dx * D.dy);
        d.dxzinv = 1.0f / (D.dx * D.dz);
        d.dyzinv = 1.0f / (D.dy * D.dz);
        d.dt = D.dt;
        d.sx = D.sx; d.sy = D.sy; d.sz = D.sz; d.bord = D.bord;
        // z chunks: about four rounds of the resident workgroups, chunks of 16 planes or more
        d.ntx = (D.sx - 2 * D.bord + DX_TX - 1) / DX_TX;
        d.nty = (D.sy - 2 * D.bord + DX_TY - 1) / DX_TY;
        const int nzu = D.sz - 2 * D.bord;
        const long long layer = (long long)d.ntx * d.nty;
        const int nzc = (int)std::max<long long>(
            1, std::min<long long>((4ll * DX_PER_CU * std::max(g_cus, 1) + layer - 1) / layer, nzu / 16));
        d.zt = (nzu + nzc - 1) / nzc;
        d.nblk = d.ntx * d.nty * ((nzu + d.zt - 1) / d.zt);
        d.trace = nullptr;
        d.rec_iz = D.rec_iz;
        if (D.trace && D.rec_every > 0 && it_host % D.rec_every == 0) {
            const int slot = it_host / D.rec_every - 1;
            if (slot >= 0 && slot < D.rec_slots)
                d.trace = D.trace + (long long)slot * (D.sx - 2 * D.bord) * (D.sy - 2 * D.bord);
        }
        hipLaunchKernelGGL(direct_dercross, dim3((d.nblk + 7) / 8 * 8), dim3(64, DX_TY), 0, s, d);
        return FL_OK;
    }
    return fail(FL_ERR_ARG, "unknown variant %d", g_variant);
}

static int take_events(hipEvent_t *e0, hipEvent_t *e1)
{
    if (!g_ev_free.empty()) {
        *e0 = g_ev_free.back().first;
        *e1 = g_ev_free.back().second;
        g_ev_free.pop_back();
        return FL_OK;
    }
    HIPCHK(hipEventCreate(e0));
    HIPCHK(hipEventCreate(e1));
    return FL_OK;
}

// ------------------------------------------------------------------ graphs
// Instantiated graph of K captured steps.  It is replayed while the launch arguments of
// its first two steps (levels, tiles, coefficient records, traces, source cell, control
// block) and the kernel selection are those it was captured with; two are kept, one per
// level parity at the start of a run.
struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int K = 0, ty = 0, cm = 0, hints = 0, variant = 0, pair = 0;
    StepArgs a0, a1;
    unsigned long long used = 0;
};
static GraphSlot g_graphs[2];
static unsigned long long g_graph_tick = 0;

static void drop_graphs()
{
    for (auto &g : g_graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g = GraphSlot();
    }
}

__global__ void set_run(int *run, int it_base, int src_n)
{
    if (threadIdx.x == 0) {
        run[RUN_K0] = 0;
        run[RUN_K1] = 0;
        run[RUN_IT_BASE] = it_base;
        run[RUN_SRC_N] = src_n;
    }
}

// ------------------------------------------------------------------ allocation-free runs
// Scratch of fl_run_steps_static: the run control block (RUN_WORDS ints) at offset 0 and up
// to FL_STATIC_MAX_STEPS source samples at offset 256.  Each GPU holds its own copy at its
// own address until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returned for that device (g_scratch_dev), never through the name.
#define FL_STATIC_MAX_STEPS 65536
#define FL_STATIC_SRC_OFF 256
static __device__ __align__(256) unsigned char scratch[FL_STATIC_SRC_OFF + 4 * FL_STATIC_MAX_STEPS];
#define FL_MAX_DEVICES 64
static char *g_scratch_dev[FL_MAX_DEVICES];

// Resolves the scratch array of the current device once (called by DRIVER_Initialize).
static int resolve_scratch()
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= FL_MAX_DEVICES) return fail(FL_ERR_HIP, "device %d beyond the scratch table", dev);
    if (g_scratch_dev[dev]) return FL_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(FL_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = (char *)p;
    return FL_OK;
}

// ------------------------------------------------------------------ driver.h
static void free_device()
{
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    for (int k = 0; k < 2; k++)
        if (D.lv[k]) (void)hipFree(D.lv[k]);
    if (D.coef) (void)hipFree(D.coef);
    if (D.dvx) (void)hipFree(D.dvx);
    if (D.dvy) (void)hipFree(D.dvy);
    if (D.trace) (void)hipFree(D.trace);
    if (D.run) (void)hipFree(D.run);
    if (D.src_tab) (void)hipFree(D.src_tab);
    drop_graphs();
    free_xfer();
    D = Device();
}

// Device state for a grid: levels (pp, qp, pc, qc from host arrays or zero), coefficient
// records from the model source, computed on the host (whole arrays, unless device
// precompute is selected) or on the device (always for slab sources).
static int initialize_src(int sx, int sy, int sz, int bord, float dx, float dy, float dz, float dt,
                          const ModelSrc &src, const float *pp, const float *pc, const float *qp, const float *qc)
{
    if (bord < 4) return fail(FL_ERR_ARG, "bord=%d: the 8th-order stencil needs bord >= 4", bord);
    if (sx <= 2 * bord || sy <= 2 * bord || sz <= 2 * bord)
        return fail(FL_ERR_ARG, "grid %dx%dx%d has no interior for bord=%d", sx, sy, sz, bord);
    if (!(dx > 0 && dy > 0 && dz > 0 && dt > 0))
        return fail(FL_ERR_ARG, "spacings and dt must be positive");
    if (!src.fill && (!src.a[0] || !src.a[1] || !src.a[2] || !src.a[3] || !src.a[4] || !src.a[5]))
        return fail(FL_ERR_ARG, "model arrays must not be NULL");
    if (max_zt(pitch_of(sx), sy) < 1)
        return fail(FL_ERR_ARG, "xy plane %dx%d too large for 32-bit plane offsets", sx, sy);
    int rc = ensure_stream();
    if (!rc) rc = resolve_scratch();
    if (rc) return rc;
    free_device();
    const long long n = (long long)sx * sy * sz;
    D.sx = sx; D.sy = sy; D.sz = sz; D.bord = bord;
    D.dx = dx; D.dy = dy; D.dz = dz; D.dt = dt;
    D.n = n;
    D.px = pitch_of(sx);
    D.np = (long long)D.px * sy * sz;
    D.lc = 1;
    D.cm = g_precision;
    for (int k = 0; k < 2; k++) {
        if ((rc = dev_alloc((void **)&D.lv[k], (size_t)2 * D.np * 4, &D.bytes))) { free_device(); return rc; }
        hipError_t e = hipMemsetAsync(D.lv[k], 0, (size_t)2 * D.np * 4, g_stream);
        if (e != hipSuccess) { free_device(); return fail(FL_ERR_HIP, "field clear: %s", hipGetErrorString(e)); }
    }
    // pp, qp into the previous level, pc, qc into the current one
    const float *host[4] = {pp, pc, qp, qc};
    for (int k = 0; k < 4; k++) {
        if (!host[k]) continue;
        float *lv = (k & 1) ? D.lv[D.lc] : D.lv[1 - D.lc];
        if ((rc = level_xfer(lv, k >> 1, const_cast<float *>(host[k]), true, sx, sy, sz, D.px))) {
            free_device();
            return rc;
        }
    }
    const size_t cbytes = (size_t)coef_count(D.cm) * D.np * coef_elem(D.cm);
    if ((rc = dev_alloc(&D.coef, cbytes, &D.bytes))) { free_device(); return rc; }
    if (src.fill || g_device_precompute) {
        rc = device_records(src, D.cm, sx, sy, sz, D.px, dt, D.coef, &D.vq);
    } else {
        int ebase = 0;
        if (D.cm == CM_Q16R) {
            ebase = fl_recompute_q16_base(D.n, src.a[0], src.a[1], src.a[2], src.a[3], dt);
            D.vq = std::ldexp(1.0f, ebase);
        }
        rc = upload_coefs(src.a[0], src.a[1], src.a[2], src.a[3], src.a[4], src.a[5], dt, D.cm, D.coef, sx, sy, sz,
                          D.px, ebase);
    }
    if (rc) {
        free_device();
        return rc;
    }
    if ((rc = dev_alloc((void **)&D.run, RUN_WORDS * sizeof(int), &D.bytes))) { free_device(); return rc; }
    D.init = true;
    return FL_OK;
}

static int initialize(int sx, int sy, int sz, int bord, float dx, float dy, float dz, float dt,
                      const float *vpz, const float *vsv, const float *eps, const float *del,
                      const float *phi, const float *theta, const float *pp, const float *pc,
                      const float *qp, const float *qc)
{
    ModelSrc src;
    memset(&src, 0, sizeof(src));
    const float *a[6] = {vpz, vsv, eps, del, phi, theta};
    for (int k = 0; k < 6; k++) src.a[k] = a[k];
    src.sx = sx;
    src.sy = sy;
    return initialize_src(sx, sy, sz, bord, dx, dy, dz, dt, src, pp, pc, qp, qc);
}

extern "C" int fl_initialize_streamed(int sx, int sy, int sz, int bord, float dx, float dy, float dz, float dt,
                                      fl_model_slab_fn fill, void *user)
{
    if (!fill) return fail(FL_ERR_ARG, "fl_initialize_streamed needs a model source");
    ModelSrc src;
    memset(&src, 0, sizeof(src));
    src.fill = fill;
    src.user = user;
    src.sx = sx;
    src.sy = sy;
    int rc = initialize_src(sx, sy, sz, bord, dx, dy, dz, dt, src, nullptr, nullptr, nullptr, nullptr);
    if (rc == FL_OK) fl_clear_error();
    return rc;
}

// Selection through the environment for front-ends that only call driver.h (the
// reference's main.c/model.c): FL_VARIANT = fused|direct|twopass, FL_PRECISION =
// fp32|bf16|fp16|fp16r|fp32r|q16r, FL_STREAM_HINTS = 0|1, FL_DEVICE_PRECOMPUTE = 0|1, applied at
// DRIVER_Initialize when set;
// FL_TUNE = cache file, the autotuner runs (or reads the cache) right after it.
static int env_pick(const char *name, const char *const *opts, int n)
{
    const char *v = getenv(name);
    if (!v || !*v) return -1;
    for (int i = 0; i < n; i++)
        if (strcmp(v, opts[i]) == 0) return i;
    return -2;
}

extern "C" void DRIVER_Initialize(const int sx, const int sy, const int sz, const int bord, float dx,
                                  float dy, float dz, float dt, float *__restrict__ vpz,
                                  float *__restrict__ vsv, float *__restrict__ epsilon,
                                  float *__restrict__ delta, float *__restrict__ phi,
                                  float *__restrict__ theta, float *__restrict__ pp,
                                  float *__restrict__ pc, float *__restrict__ qp,
                                  float *__restrict__ qc)
{
    static const char *const variants[] = {"fused", "direct", "twopass"};
    static const char *const precs[] = {"fp32", "bf16", "fp16", "fp16r", "fp32r", "q16r"};
    static const char *const onoff[] = {"0", "1"};
    const int v = env_pick("FL_VARIANT", variants, 3), p = env_pick("FL_PRECISION", precs, 6);
    const int h = env_pick("FL_STREAM_HINTS", onoff, 2), dp = env_pick("FL_DEVICE_PRECOMPUTE", onoff, 2);
    if (v == -2 || p == -2 || h == -2 || dp == -2) {
        fail(FL_ERR_ARG, "FL_VARIANT / FL_PRECISION / FL_STREAM_HINTS / FL_DEVICE_PRECOMPUTE hold an unknown value");
        return;
    }
    if (v >= 0) g_variant = v;
    if (p >= 0) g_precision = p;
    if (h >= 0) g_tiles.hints = h;
    if (dp >= 0) g_device_precompute = dp != 0;
    if (initialize(sx, sy, sz, bord, dx, dy, dz, dt, vpz, vsv, epsilon, delta, phi, theta, pp, pc, qp, qc) != FL_OK)
        return;
    fl_clear_error();
    const char *tune = getenv("FL_TUNE");
    if (tune && *tune) fl_autotune(5, tune);
}

extern "C" void DRIVER_Finalize() { free_device(); }

static int propagate(int sx, int sy, int sz, int bord, int it)
{
    if (!D.init) return fail(FL_ERR_STATE, "DRIVER_Propagate before DRIVER_Initialize");
    if (sx != D.sx || sy != D.sy || sz != D.sz || bord != D.bord)
        return fail(FL_ERR_ARG, "grid %dx%dx%d/%d differs from the initialised %dx%dx%d/%d", sx, sy, sz,
                    bord, D.sx, D.sy, D.sz, D.bord);
    StepArgs a = base_args();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_timing) {
        int rc = take_events(&e0, &e1);
        if (rc) return rc;
        HIPCHK(hipEventRecord(e0, g_stream));
    }
    int rc = launch_step(a, it, g_stream);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    if (g_timing) {
        HIPCHK(hipEventRecord(e1, g_stream));
        g_ev_used.push_back({e0, e1});
    }
    swap_levels();
    if (g_sync) HIPCHK(hipStreamSynchronize(g_stream));
    return FL_OK;
}

extern "C" void DRIVER_Propagate(const int sx, const int sy, const int sz, const int bord, const float dx,
                                 const float dy, const float dz, const float dt, const int it, float *pp,
                                 float *pc, float *qp, float *qc)
{
    (void)dx; (void)dy; (void)dz; (void)dt; (void)pp; (void)pc; (void)qp; (void)qc;
    propagate(sx, sy, sz, bord, it);
}

static int insert_src(int iSource, float src)
{
    if (!D.init) return fail(FL_ERR_STATE, "DRIVER_InsertSource before DRIVER_Initialize");
    if (iSource < 0 || (long long)iSource >= D.n)
        return fail(FL_ERR_ARG, "iSource=%d outside the %lld-cell grid", iSource, D.n);
    hipLaunchKernelGGL(insert_source, dim3(1), dim3(64), 0, g_stream, D.lv[D.lc],
                       pitched_index(iSource, D.sx, D.sy, D.px), src);
    HIPCHK(hipGetLastError());
    if (g_sync) HIPCHK(hipStreamSynchronize(g_stream));
    return FL_OK;
}

extern "C" void DRIVER_InsertSource(float dt, int it, int iSource, float *p, float *q, float src)
{
    (void)dt; (void)it; (void)p; (void)q;
    insert_src(iSource, src);
}

static int update_pointers(int sx, int sy, int sz, float *pc)
{
    if (!D.init) return fail(FL_ERR_STATE, "DRIVER_Update_pointers before DRIVER_Initialize");
    if ((long long)sx * sy * sz != D.n) return fail(FL_ERR_ARG, "grid size differs from the initialised one");
    if (!pc) return FL_OK;
    return level_xfer(D.lv[D.lc], 0, pc, false, D.sx, D.sy, D.sz, D.px);
}

extern "C" void DRIVER_Update_pointers(const int sx, const int sy, const int sz, float *pc)
{
    update_pointers(sx, sy, sz, pc);
}

// ------------------------------------------------------------------ extensions
extern "C" int fl_set_variant(int v)
{
    if (v < FL_VARIANT_FUSED || v > FL_VARIANT_TWOPASS) return fail(FL_ERR_ARG, "unknown variant %d", v);
    g_variant = v;
    return FL_OK;
}
extern "C" int fl_get_variant(void) { return g_variant; }

extern "C" int fl_set_precision(int p)
{
    if (p < FL_PREC_FP32 || p > FL_PREC_Q16_RECOMPUTE) return fail(FL_ERR_ARG, "unknown precision %d", p);
    g_precision = p;
    return FL_OK;
}
extern "C" int fl_get_precision(void) { return D.init ? D.cm : g_precision; }

extern "C" int fl_set_sync(int on)
{
    g_sync = on != 0;
    return FL_OK;
}

extern "C" int fl_set_tiles(int ty, int zt)
{
    if (ty == 0) ty = 8;
    if (ty != 4 && ty != 8 && ty != 12 && ty != 16)
        return fail(FL_ERR_ARG, "tile rows must be 4, 8, 12 or 16 (got %d)", ty);
    if (zt < 0) return fail(FL_ERR_ARG, "zt must be >= 0");
    g_tiles.ty = ty;
    g_tiles.zt = zt;
    return FL_OK;
}

extern "C" int fl_set_stream_hints(int on)
{
    if (on != 0 && on != 1) return fail(FL_ERR_ARG, "stream hints must be 0 or 1 (got %d)", on);
    g_tiles.hints = on;
    return FL_OK;
}

extern "C" int fl_get_stream_hints(void) { return g_tiles.hints; }

extern "C" int fl_set_tile_strip(int sw)
{
    if (sw < 0) return fail(FL_ERR_ARG, "tile strip width must be >= 0 (got %d)", sw);
    g_tiles.sw = sw;
    return FL_OK;
}

extern "C" int fl_get_tile_strip(void)
{
    if (!D.init) return g_tiles.sw;
    return base_args().sw;
}

extern "C" int fl_set_lane_pairs(int on)
{
    if (on != 0 && on != 1) return fail(FL_ERR_ARG, "lane pairs must be 0 or 1 (got %d)", on);
    g_tiles.pair = on;
    return FL_OK;
}

extern "C" int fl_get_lane_pairs(void) { return g_tiles.pair; }

extern "C" int fl_get_tiles(int *ty, int *zt)
{
    if (ty) *ty = g_tiles.ty;
    if (zt) *zt = (g_tiles.zt > 0 || !D.init) ? g_tiles.zt : default_zt(D.sx, D.sy, D.sz, D.bord, g_tiles.ty);
    return FL_OK;
}

extern "C" int fl_get_field(int which, float *host)
{
    if (!D.init) return fail(FL_ERR_STATE, "no initialised grid");
    if (which < 0 || which > 3 || !host) return fail(FL_ERR_ARG, "bad field request %d", which);
    float *lv = (which & 1) ? D.lv[D.lc] : D.lv[1 - D.lc];  // pp pc qp qc
    return level_xfer(lv, which >> 1, host, false, D.sx, D.sy, D.sz, D.px);
}

extern "C" int fl_set_field(int which, const float *host)
{
    if (!D.init) return fail(FL_ERR_STATE, "no initialised grid");
    if (which < 0 || which > 3 || !host) return fail(FL_ERR_ARG, "bad field request %d", which);
    float *lv = (which & 1) ? D.lv[D.lc] : D.lv[1 - D.lc];
    return level_xfer(lv, which >> 1, const_cast<float *>(host), true, D.sx, D.sy, D.sz, D.px);
}

extern "C" long long fl_record_bytes(void)
{
    if (!D.init) return fail(FL_ERR_STATE, "no initialised grid");
    return (long long)coef_count(D.cm) * coef_elem(D.cm) * D.np;
}

extern "C" int fl_get_records(void *host)
{
    if (!D.init) return fail(FL_ERR_STATE, "no initialised grid");
    if (!host) return fail(FL_ERR_ARG, "host buffer must not be NULL");
    HIPCHK(hipStreamSynchronize(g_stream));
    HIPCHK(hipMemcpy(host, D.coef, (size_t)fl_record_bytes(), hipMemcpyDeviceToHost));
    if (D.cm == CM_F32) {  // the two device arrays back to one 10-slot record a cell
        const long long np = D.np;
        std::vector<float> dev((const float *)host, (const float *)host + np * 10);
        float *out = (float *)host;
        const float *a1 = dev.data(), *a2 = dev.data() + np * 8;
        for (long long c = 0; c < np; c++) {
            for (int s = 0; s < 4; s++) out[c * 10 + s] = a1[c * 8 + s];
            out[c * 10 + 4] = a2[c * 2];
            for (int s = 5; s < 9; s++) out[c * 10 + s] = a1[c * 8 + s - 1];
            out[c * 10 + 9] = a2[c * 2 + 1];
        }
    }
    return FL_OK;
}

extern "C" int fl_get_derivative_volume(int k, float *host)
{
    if (!D.init || !D.dvx) return fail(FL_ERR_STATE, "no two-pass step has run");
    if (k < 0 || k > 3 || !host) return fail(FL_ERR_ARG, "bad derivative volume %d", k);
    // 0 pDx, 1 pDy, 2 qDx, 3 qDy
    return level_xfer((k & 1) ? D.dvy : D.dvx, k >> 1, host, false, D.sx, D.sy, D.sz, D.px);
}

extern "C" int fl_record_traces(int iz, int every, int max_traces)
{
    if (!D.init) return fail(FL_ERR_STATE, "no initialised grid");
    if (D.trace) {
        HIPCHK(hipStreamSynchronize(g_stream));
        (void)hipFree(D.trace);
        D.bytes -= (long long)D.rec_slots * (D.sx - 2 * D.bord) * (D.sy - 2 * D.bord) * 4;
        D.trace = nullptr;
    }
    D.rec_iz = -1; D.rec_every = 0; D.rec_slots = 0;
    if (iz < 0) return FL_OK;
    if (iz < D.bord || iz >= D.sz - D.bord || every <= 0 || max_traces <= 0)
        return fail(FL_ERR_ARG, "receiver plane %d / every %d / slots %d invalid", iz, every, max_traces);
    const size_t bytes = (size_t)max_traces * (D.sx - 2 * D.bord) * (D.sy - 2 * D.bord) * 4;
    int rc = dev_alloc((void **)&D.trace, bytes, &D.bytes);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(D.trace, 0, bytes, g_stream));
    D.rec_iz = iz; D.rec_every = every; D.rec_slots = max_traces;
    return FL_OK;
}

extern "C" int fl_get_traces(float *host, int max_traces)
{
    if (!D.init || !D.trace) return fail(FL_ERR_STATE, "traces are not being recorded");
    const int m = std::min(max_traces, D.rec_slots);
    HIPCHK(hipStreamSynchronize(g_stream));
    HIPCHK(hipMemcpy(host, D.trace, (size_t)m * (D.sx - 2 * D.bord) * (D.sy - 2 * D.bord) * 4,
                     hipMemcpyDeviceToHost));
    return m;
}

// The graph of K steps starting from the current levels, captured and instantiated on
// first use.  args(k) gives the launch arguments of step k.
template <class F>
static int step_graph(int K, F &&args, GraphSlot **out)
{
    const StepArgs a0 = args(0), a1 = args(1);
    for (auto &g : g_graphs)
        if (g.exec && g.K == K && g.ty == g_tiles.ty && g.cm == D.cm && g.hints == g_tiles.hints &&
            g.variant == g_variant && g.pair == g_tiles.pair && memcmp(&g.a0, &a0, sizeof(StepArgs)) == 0 &&
            memcmp(&g.a1, &a1, sizeof(StepArgs)) == 0) {
            g.used = ++g_graph_tick;
            *out = &g;
            return FL_OK;
        }
    GraphSlot *g = &g_graphs[0];
    for (auto &c : g_graphs)
        if (c.used < g->used) g = &c;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    *g = GraphSlot();
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    const int lc0 = D.lc;
    int rc = FL_OK;
    for (int k = 0; k < K && rc == FL_OK; k++) {
        rc = launch_steps(args(k), g_tiles.ty, D.cm, false, g_stream);
        swap_levels();
    }
    D.lc = lc0;  // K is even: the levels are back in their roles after a replay
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(g_stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(FL_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ec != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(FL_ERR_HIP, "graph instantiation: %s", hipGetErrorString(ec));
    }
    g->exec = exec;
    g->graph = graph;
    g->K = K;
    g->ty = g_tiles.ty;
    g->cm = D.cm;
    g->hints = g_tiles.hints;
    g->variant = g_variant;
    g->pair = g_tiles.pair;
    g->a0 = a0;
    g->a1 = a1;
    g->used = ++g_graph_tick;
    *out = g;
    return FL_OK;
}

extern "C" int fl_run_steps(int it0, int nsteps, int iSource, const float *src, int graph_chunk)
{
    if (!D.init) return fail(FL_ERR_STATE, "fl_run_steps before DRIVER_Initialize");
    if (nsteps <= 0) return FL_OK;
    if (!src) return fail(FL_ERR_ARG, "src must not be NULL");
    if (iSource < 0 || (long long)iSource >= D.n) return fail(FL_ERR_ARG, "iSource outside the grid");
    // the captured form adds the source when the kernel stores its cell: only cells of
    // the update domain are stored, so a source elsewhere takes the call-by-call path
    const long long sxy = (long long)D.sx * D.sy;
    const int siz = (int)(iSource / sxy), siy = (int)((iSource % sxy) / D.sx), six = (int)(iSource % D.sx);
    const bool src_in_domain = siz >= D.bord && siz < D.sz - D.bord && siy >= D.bord &&
                               siy < D.sy - D.bord && six >= D.bord && six < D.sx - D.bord;
    if (g_variant != FL_VARIANT_FUSED || graph_chunk <= 0 || !src_in_domain) {
        // plain launches: InsertSource + Propagate per step
        const bool sync = g_sync;
        g_sync = false;
        int rc = FL_OK;
        for (int k = 0; k < nsteps && rc == FL_OK; k++) {
            rc = insert_src(iSource, src[k]);
            if (rc == FL_OK) rc = propagate(D.sx, D.sy, D.sz, D.bord, it0 + 1 + k);
        }
        g_sync = sync;
        if (rc) return rc;
        if (g_sync) HIPCHK(hipStreamSynchronize(g_stream));
        return FL_OK;
    }
    // source samples: a device table grown in powers of two, so that graphs stay valid
    // across runs of up to its capacity
    if (D.src_cap < nsteps) {
        int cap = 4096;
        while (cap < nsteps) cap *= 2;
        if (D.src_tab) {
            HIPCHK(hipStreamSynchronize(g_stream));
            (void)hipFree(D.src_tab);
            D.bytes -= (long long)D.src_cap * 4;
        }
        D.src_tab = nullptr;
        D.src_cap = 0;
        int rc = dev_alloc((void **)&D.src_tab, (size_t)cap * 4, &D.bytes);
        if (rc) return rc;
        D.src_cap = cap;
    }
    HIPCHK(hipMemcpyAsync(D.src_tab, src, (size_t)nsteps * 4, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(set_run, dim3(1), dim3(64), 0, g_stream, D.run, it0, nsteps);
    const long long psrc = pitched_index(iSource, D.sx, D.sy, D.px);
    hipLaunchKernelGGL(insert_source, dim3(1), dim3(64), 0, g_stream, D.lv[D.lc], psrc, src[0]);
    const int K = std::min(nsteps, graph_chunk) & ~1;  // even: levels return to their roles
    auto step_args = [&](int k) {
        StepArgs a = base_args();
        a.run = D.run;
        a.par = k & 1;
        a.src_tab = D.src_tab;
        a.src_idx = psrc;
        return a;
    };
    int done = 0;
    if (K >= 2 && nsteps >= K) {
        GraphSlot *g = nullptr;
        int rc = step_graph(K, step_args, &g);
        if (rc) return rc;
        const int reps = nsteps / K;
        for (int r = 0; r < reps; r++) HIPCHK(hipGraphLaunch(g->exec, g_stream));
        done = reps * K;
    }
    for (int k = done; k < nsteps; k++) {
        int rc = launch_steps(step_args(k), g_tiles.ty, D.cm, false, g_stream);
        if (rc) return rc;
        swap_levels();
    }
    HIPCHK(hipGetLastError());
    if (g_sync) HIPCHK(hipStreamSynchronize(g_stream));
    return FL_OK;
}

// fl_run_steps without any allocation: the fused variant's steps launched one by one, the
// run control block and the source samples in the device's scratch array.  Calls that
// overlap on one GPU share that array and race.
extern "C" int fl_run_steps_static(int it0, int nsteps, int iSource, const float *src)
{
    if (!D.init) return fail(FL_ERR_STATE, "fl_run_steps_static before DRIVER_Initialize");
    if (g_variant != FL_VARIANT_FUSED) return fail(FL_ERR_ARG, "fl_run_steps_static runs the fused variant only");
    if (nsteps <= 0) return FL_OK;
    if (nsteps > FL_STATIC_MAX_STEPS)
        return fail(FL_ERR_ARG, "fl_run_steps_static: %d steps exceed the scratch limit of %d per call", nsteps,
                    FL_STATIC_MAX_STEPS);
    if (!src) return fail(FL_ERR_ARG, "src must not be NULL");
    if (iSource < 0 || (long long)iSource >= D.n) return fail(FL_ERR_ARG, "iSource outside the grid");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    char *const scr = (dev >= 0 && dev < FL_MAX_DEVICES) ? g_scratch_dev[dev] : nullptr;
    if (!scr) return fail(FL_ERR_STATE, "no scratch address for device %d", dev);
    int *const run = (int *)scr;
    float *const tab = (float *)(scr + FL_STATIC_SRC_OFF);
    const long long sxy = (long long)D.sx * D.sy;
    const int siz = (int)(iSource / sxy), siy = (int)((iSource % sxy) / D.sx), six = (int)(iSource % D.sx);
    const bool src_in_domain = siz >= D.bord && siz < D.sz - D.bord && siy >= D.bord &&
                               siy < D.sy - D.bord && six >= D.bord && six < D.sx - D.bord;
    const long long psrc = pitched_index(iSource, D.sx, D.sy, D.px);
    if (!src_in_domain) {
        // the source cell is never stored by the step kernel: insert before every step
        for (int k = 0; k < nsteps; k++) {
            hipLaunchKernelGGL(insert_source, dim3(1), dim3(64), 0, g_stream, D.lv[D.lc], psrc, src[k]);
            StepArgs a = base_args();
            a.it_host = it0 + 1 + k;
            int rc = launch_steps(a, g_tiles.ty, D.cm, false, g_stream);
            if (rc) return rc;
            swap_levels();
        }
    } else {
        HIPCHK(hipMemcpyAsync(tab, src, (size_t)nsteps * 4, hipMemcpyHostToDevice, g_stream));
        hipLaunchKernelGGL(set_run, dim3(1), dim3(64), 0, g_stream, run, it0, nsteps);  // counters from 0
        hipLaunchKernelGGL(insert_source, dim3(1), dim3(64), 0, g_stream, D.lv[D.lc], psrc, src[0]);
        for (int k = 0; k < nsteps; k++) {
            StepArgs a = base_args();
            a.run = run;
            a.par = k & 1;
            a.src_tab = tab;
            a.src_idx = psrc;
            int rc = launch_steps(a, g_tiles.ty, D.cm, false, g_stream);
            if (rc) return rc;
            swap_levels();
        }
    }
    HIPCHK(hipGetLastError());
    if (g_sync) HIPCHK(hipStreamSynchronize(g_stream));
    return FL_OK;
}

extern "C" int fl_reset_fields(void)
{
    if (!D.init) return fail(FL_ERR_STATE, "fl_reset_fields before DRIVER_Initialize");
    for (int k = 0; k < 2; k++) HIPCHK(hipMemsetAsync(D.lv[k], 0, (size_t)2 * D.np * 4, g_stream));
    if (D.trace)
        HIPCHK(hipMemsetAsync(D.trace, 0, (size_t)D.rec_slots * (D.sx - 2 * D.bord) * (D.sy - 2 * D.bord) * 4,
                              g_stream));
    D.lc = 1;
    if (g_sync) HIPCHK(hipStreamSynchronize(g_stream));
    return FL_OK;
}

extern "C" int fl_synchronize(void)
{
    if (!g_stream) return FL_OK;
    HIPCHK(hipStreamSynchronize(g_stream));
    return FL_OK;
}

extern "C" int fl_set_timing(int on)
{
    g_timing = on != 0;
    return FL_OK;
}

extern "C" int fl_kernel_times(double *total_ms, int *count)
{
    double tot = 0.0;
    int cnt = 0;
    if (!g_ev_used.empty()) {
        HIPCHK(hipStreamSynchronize(g_stream));
        for (auto &p : g_ev_used) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, p.first, p.second));
            tot += ms;
            cnt++;
            g_ev_free.push_back(p);
        }
        g_ev_used.clear();
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = cnt;
    return FL_OK;
}

extern "C" int fl_region_begin(void)
{
    int rc = ensure_stream();
    if (rc) return rc;
    if (!g_reg0) {
        HIPCHK(hipEventCreate(&g_reg0));
        HIPCHK(hipEventCreate(&g_reg1));
    }
    HIPCHK(hipEventRecord(g_reg0, g_stream));
    return FL_OK;
}

extern "C" int fl_region_end(double *ms)
{
    if (!g_reg0) return fail(FL_ERR_STATE, "fl_region_end without fl_region_begin");
    HIPCHK(hipEventRecord(g_reg1, g_stream));
    HIPCHK(hipEventSynchronize(g_reg1));
    float m = 0.0f;
    HIPCHK(hipEventElapsedTime(&m, g_reg0, g_reg1));
    if (ms) *ms = m;
    return FL_OK;
}

extern "C" int fl_device_name(char *buf, int len)
{
    int rc = ensure_stream();
    if (rc) return rc;
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, 0));
    snprintf(buf, len, "%s %s", p.name, p.gcnArchName);
    return FL_OK;
}

extern "C" long long fl_device_bytes(void) { return D.bytes; }

extern "C" int fl_device_memory(long long *free_bytes, long long *total_bytes)
{
    int rc = ensure_stream();
    if (rc) return rc;
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (long long)f;
    if (total_bytes) *total_bytes = (long long)t;
    return FL_OK;
}

// ------------------------------------------------------------------ autotuner
static bool cache_lookup(const char *path, const std::string &key, int *ty, int *zt,
                         int *hints = nullptr)
{
    if (!path) return false;
    FILE *f = fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        char k[400];
        int a, b, h = 0;
        double ms;
        const int n = sscanf(line, "%399s %d %d %lf %d", k, &a, &b, &ms, &h);
        if (n >= 3 && key == k) {
            *ty = a; *zt = b;
            if (hints) *hints = n == 5 ? h : 0;
            found = true;  // last entry wins
        }
    }
    fclose(f);
    return found;
}

static void cache_store(const char *path, const std::string &key, int ty, int zt, double ms,
                        int hints = 0)
{
    if (!path) return;
    FILE *f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %d %d %.4f %d\n", key.c_str(), ty, zt, ms, hints);
    fclose(f);
}

static std::string tune_key(const char *kind, int sx, int sy, int sz, int cm)
{
    char name[256] = "";
    fl_device_name(name, sizeof(name));
    for (char *c = name; *c; c++)
        if (*c == ' ') *c = '_';
    char buf[400];
    // the lane layout (fl_set_lane_pairs) changes which tiles are fastest for the forms that
    // have both: it is part of the key for them
    const int lp = has_pair_form(cm) ? g_tiles.pair : 0;
    snprintf(buf, sizeof(buf), "%s|%s|%dx%dx%d|v%d|p%d%s", name, kind, sx, sy, sz, g_variant, cm, lp ? "|lp" : "");
    return buf;
}

static std::vector<int> zt_candidates(int nzu)
{
    // fixed depths (4 + a multiple of 10 runs every output plane of a chunk on the
    // statically addressed steps) plus even splits of the column into 2..16 chunks
    std::vector<int> c = {14, 24, 34, 44, 54, 64, 74, 84, 94, 124, 154, 194};
    for (int k : {2, 4, 6, 8, 12, 16}) c.push_back((nzu + k - 1) / k);
    std::vector<int> r;
    for (int z : c)
        if (z >= 8 && z < nzu && std::find(r.begin(), r.end(), z) == r.end()) r.push_back(z);
    r.push_back(nzu);
    std::sort(r.begin(), r.end());
    return r;
}

extern "C" int fl_autotune(int steps, const char *cache_path)
{
    if (!D.init) return fail(FL_ERR_STATE, "fl_autotune before DRIVER_Initialize");
    if (g_variant == FL_VARIANT_DIRECT) return FL_OK;  // no tile parameters
    const std::string key = tune_key("grid", D.sx, D.sy, D.sz, D.cm);
    int ty, zt, hints = 0;
    if (cache_lookup(cache_path, key, &ty, &zt, &hints)) {
        g_tiles.hints = hints;
        return fl_set_tiles(ty, zt);
    }
    steps = std::max(1, steps);
    // keep the fields: tuning steps run on them and are undone afterwards
    float *save = nullptr;
    long long dummy = 0;
    const size_t lvn = (size_t)2 * D.np;
    int rc = dev_alloc((void **)&save, 2 * lvn * 4, &dummy);
    if (rc) return rc;
    const int lc0 = D.lc;
    for (int k = 0; k < 2; k++)
        HIPCHK(hipMemcpyAsync(save + k * lvn, D.lv[k], lvn * 4, hipMemcpyDeviceToDevice, g_stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const Tiles keep = g_tiles;
    double best = 1e30;
    Tiles bt = keep;
    // two passes over the candidates, the fastest single measurement wins: one pass
    // leaves the early candidates timed while the clocks still ramp.  Each is timed with
    // and without the streaming hints.
    const int nh = 2;
    for (int pass = 0; pass < 2 && rc == FL_OK; pass++)
    for (int ch = 0; ch < nh && rc == FL_OK; ch++)
    for (int cty : {4, 8, 12, 16}) {
        for (int czt : zt_candidates(std::min(D.sz - 2 * D.bord, max_zt(D.px, D.sy)))) {
            g_tiles.ty = cty;
            g_tiles.zt = czt;
            g_tiles.hints = ch;
            StepArgs a = base_args();
            a.trace = nullptr;
            rc = launch_step(a, 0, g_stream);  // warm-up
            if (rc) break;
            HIPCHK(hipEventRecord(e0, g_stream));
            for (int s = 0; s < steps && rc == FL_OK; s++) {
                StepArgs b = base_args();
                b.trace = nullptr;
                rc = launch_step(b, s & 1, g_stream);  // both tile orders (step_order)
                swap_levels();
            }
            HIPCHK(hipEventRecord(e1, g_stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rc) break;
            if (ms < best) {
                best = ms;
                bt = g_tiles;
            }
        }
        if (rc) break;
    }
    D.lc = lc0;
    for (int k = 0; k < 2; k++)
        HIPCHK(hipMemcpyAsync(D.lv[k], save + k * lvn, lvn * 4, hipMemcpyDeviceToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    (void)hipFree(save);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        g_tiles = keep;
        return rc;
    }
    g_tiles = bt;
    cache_store(cache_path, key, bt.ty, bt.zt, best / steps, bt.hints);
    return FL_OK;
}

#include "fl_batch.inc"
#include "fl_ref64.inc"
