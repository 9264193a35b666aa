// fl_stream — HBM calibration for the roofline of the TTI step kernel.
//
//   fl_stream [MiB] [reps] [mode]
//
// Runs plain copy kernels (read N bytes, write N bytes) on cuda:0 and prints the
// achieved GB/s (HIP events, best and mean of reps).  Modes:
//   0  float4 grid-stride copy, grid = 8 blocks of 256 per CU   (the PMC reference)
//   1  float4 copy, one element per thread, full grid
//   2  float4 copy, 4 elements in flight per thread, full grid
//   3  float4 copy with non-temporal (streaming) loads and stores, 4 in flight
//   4  read only: N bytes read (float4 per thread, full grid), nothing written
//   5  7:1 mix: 7 streams of N/8 bytes read, one of N/8 written (the step kernel's
//      read/write proportion); the rate counts all N bytes
//   6  3:1 mix (3 read streams, 1 written)    7  1:1 (1 read stream, 1 written)
//   8  7:1 mix as 5, two float4 of each stream per thread (twice the loads in flight)
//   9  7:1 mix as 5 with non-temporal loads and stores
//  10  the step kernel's bytes as one plain stream per cell pair: an 80-byte record pair
//      (5 float4, consecutive), 16 bytes of the current level, 16 of the previous read and
//      16 written back in place (112 read : 16 written); the rate counts all 128 bytes
//  11  as 10 with non-temporal record loads and previous-level loads and stores (the
//      current level, which the step re-reads for its halos, cached as usual)
//  12  as 10 with non-temporal record loads only
//  13  the step kernel's bytes per lane from records in blocks: each run of 32 cells keeps
//      its 10 coefficients as 10 rows of 32 floats (1280 bytes); the lanes of a cell pair
//      (p, q) read 5 rows each (every load instruction reads two whole 128-byte lines), 4
//      bytes of the current and of the previous level, and write 4 bytes of the previous
//      level back (64 bytes a cell).  The step itself keeps one 40-byte record per cell: in
//      the step, this layout with non-temporal loads (mode 14's gain here) ran 9% slower.
//  14  as 13 with non-temporal record loads and previous-level loads and stores
//  15  as 13 with non-temporal record loads only
// Under rocprofv3 --pmc, mode 0 is the known-size reference that converts FETCH_SIZE /
// WRITE_SIZE values into bytes: a copy of N MiB moves exactly N MiB each way.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e = (x);                                                         \
        if (e != hipSuccess) {                                                      \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
            return 1;                                                               \
        }                                                                           \
    } while (0)

__global__ __launch_bounds__(256) void copy_gs(const float4 *__restrict__ a, float4 *__restrict__ b,
                                               long long n4)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(256) void copy_1(const float4 *__restrict__ a, float4 *__restrict__ b,
                                              long long n4)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}

template <bool NT>
__global__ __launch_bounds__(256) void copy_4(const float *__restrict__ a, float *__restrict__ b,
                                              long long n)
{
    // 4 float4 per thread, strided by the block so each wave instruction stays contiguous
    const long long base = ((long long)blockIdx.x * 4 * 256) * 4;
    float v[4][4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long i = base + ((long long)k * 256 + threadIdx.x) * 4;
#pragma unroll
        for (int c = 0; c < 4; c++)
            v[k][c] = i + c < n ? (NT ? __builtin_nontemporal_load(a + i + c) : a[i + c]) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long i = base + ((long long)k * 256 + threadIdx.x) * 4;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (i + c < n) {
                if (NT) __builtin_nontemporal_store(v[k][c], b + i + c);
                else b[i + c] = v[k][c];
            }
    }
}

__global__ __launch_bounds__(256) void read_1(const float4 *__restrict__ a, float4 *__restrict__ b,
                                              long long n4)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    const float4 v = a[i];
    if (v.x + v.y + v.z + v.w == 1234.5f) b[i] = v;  // never true for the zeroed input
}
template <int K>
__global__ __launch_bounds__(256) void mix_k1(const float4 *__restrict__ a, float4 *__restrict__ b,
                                              long long m4)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m4) return;
    float4 s = a[i];
#pragma unroll
    for (int k = 1; k < K; k++) {
        const float4 v = a[k * m4 + i];
        s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
    b[i] = s;
}
template <int K>
__global__ __launch_bounds__(256) void mix_k2(const float4 *__restrict__ a, float4 *__restrict__ b,
                                              long long m4)
{
    const long long i = (long long)blockIdx.x * 512 + threadIdx.x;
    float4 s0 = {0.f, 0.f, 0.f, 0.f}, s1 = s0;
    const bool ok0 = i < m4, ok1 = i + 256 < m4;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const float4 v0 = ok0 ? a[k * m4 + i] : s0;
        const float4 v1 = ok1 ? a[k * m4 + i + 256] : s1;
        s0.x += v0.x; s0.y += v0.y; s0.z += v0.z; s0.w += v0.w;
        s1.x += v1.x; s1.y += v1.y; s1.z += v1.z; s1.w += v1.w;
    }
    if (ok0) b[i] = s0;
    if (ok1) b[i + 256] = s1;
}

template <int K>
__global__ __launch_bounds__(256) void mix_nt(const float4 *__restrict__ a, float4 *__restrict__ b,
                                              long long m4)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m4) return;
    const float *af = (const float *)a;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) s[c] += __builtin_nontemporal_load(af + (k * m4 + i) * 4 + c);
    float *bf = (float *)b;
#pragma unroll
    for (int c = 0; c < 4; c++) __builtin_nontemporal_store(s[c], bf + i * 4 + c);
}

// One thread per cell pair: record pair rec[5i..5i+4], current level cur[i], previous
// level prev[i] read and overwritten.  NTR: non-temporal record loads; NTP: non-temporal
// previous-level load and store.
typedef float v4f __attribute__((ext_vector_type(4)));
template <bool NTR, bool NTP>
__global__ __launch_bounds__(256) void step_bytes(const float4 *rec4, const float4 *cur4, float4 *prev4,
                                                  long long np)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const v4f *rec = (const v4f *)rec4, *cur = (const v4f *)cur4;
    v4f *prev = (v4f *)prev4;
    v4f r[5];
#pragma unroll
    for (int k = 0; k < 5; k++) r[k] = NTR ? __builtin_nontemporal_load(rec + 5 * i + k) : rec[5 * i + k];
    const v4f c = cur[i];
    v4f p = NTP ? __builtin_nontemporal_load(prev + i) : prev[i];
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 5; k++) t += r[k].x + r[k].y + r[k].z + r[k].w;
    p = 2.f * c - p + t;
    if (NTP) __builtin_nontemporal_store(p, prev + i);
    else prev[i] = p;
}

template <bool NTR, bool NTP>
__global__ __launch_bounds__(256) void step_blocked(const float *rec, const float *cur, float *prev, long long ncell)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long cell = g >> 1;
    if (cell >= ncell) return;
    const int fld = (int)(g & 1), cx = (int)(cell & 31);
    const float *rb = rec + (cell >> 5) * 320 + fld * 32 + cx;
    float t = 0.f;
#pragma unroll
    for (int j = 0; j < 5; j++) t += NTR ? __builtin_nontemporal_load(rb + 64 * j) : rb[64 * j];
    const float c = cur[g];
    float p = NTP ? __builtin_nontemporal_load(prev + g) : prev[g];
    p = 2.f * c - p + t;
    if (NTP) __builtin_nontemporal_store(p, prev + g);
    else prev[g] = p;
}

int main(int argc, char **argv)
{
    const long long mib = argc > 1 ? atoll(argv[1]) : 2048;
    const int reps = argc > 2 ? atoi(argv[2]) : 20;
    const int mode = argc > 3 ? atoi(argv[3]) : 0;
    if (mode < 0 || mode > 15) {
        fprintf(stderr, "mode must be 0..15\n");
        return 2;
    }
    const long long bytes = mib << 20, n4 = bytes / 16;
    float4 *a, *b;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    double best = 0, sum = 0;
    for (int r = 0; r < reps; r++) {
        CHECK(hipEventRecord(e0, 0));
        if (mode == 0)
            hipLaunchKernelGGL(copy_gs, dim3(p.multiProcessorCount * 8), dim3(256), 0, 0, a, b, n4);
        else if (mode == 1)
            hipLaunchKernelGGL(copy_1, dim3((n4 + 255) / 256), dim3(256), 0, 0, a, b, n4);
        else if (mode == 2)
            hipLaunchKernelGGL(copy_4<false>, dim3((n4 + 1023) / 1024), dim3(256), 0, 0, (const float *)a,
                               (float *)b, n4 * 4);
        else if (mode == 3)
            hipLaunchKernelGGL(copy_4<true>, dim3((n4 + 1023) / 1024), dim3(256), 0, 0, (const float *)a,
                               (float *)b, n4 * 4);
        else if (mode == 4)
            hipLaunchKernelGGL(read_1, dim3((n4 + 255) / 256), dim3(256), 0, 0, a, b, n4);
        else if (mode == 5)
            hipLaunchKernelGGL(mix_k1<7>, dim3((n4 / 8 + 255) / 256), dim3(256), 0, 0, a, b, n4 / 8);
        else if (mode == 6)
            hipLaunchKernelGGL(mix_k1<3>, dim3((n4 / 4 + 255) / 256), dim3(256), 0, 0, a, b, n4 / 4);
        else if (mode == 7)
            hipLaunchKernelGGL(mix_k1<1>, dim3((n4 / 2 + 255) / 256), dim3(256), 0, 0, a, b, n4 / 2);
        else if (mode == 8)
            hipLaunchKernelGGL(mix_k2<7>, dim3((n4 / 8 + 511) / 512), dim3(256), 0, 0, a, b, n4 / 8);
        else if (mode == 9)
            hipLaunchKernelGGL(mix_nt<7>, dim3((n4 / 8 + 255) / 256), dim3(256), 0, 0, a, b, n4 / 8);
        else if (mode == 10)  // a holds the records (5/7 of the bytes) and the current level, b the previous
            hipLaunchKernelGGL((step_bytes<false, false>), dim3((n4 / 7 + 255) / 256), dim3(256), 0, 0, a,
                               a + 5 * (n4 / 7), b, n4 / 7);
        else if (mode == 11)
            hipLaunchKernelGGL((step_bytes<true, true>), dim3((n4 / 7 + 255) / 256), dim3(256), 0, 0, a,
                               a + 5 * (n4 / 7), b, n4 / 7);
        else if (mode == 12)
            hipLaunchKernelGGL((step_bytes<true, false>), dim3((n4 / 7 + 255) / 256), dim3(256), 0, 0, a,
                               a + 5 * (n4 / 7), b, n4 / 7);
        else {  // a: records (10 floats a cell) then the current level (2), b: the previous level
            const long long nc = n4 * 4 / 12 / 32 * 32;
            const float *rec = (const float *)a, *cur = rec + 10 * nc;
            float *prev = (float *)b;
            const dim3 grid((unsigned)((2 * nc + 255) / 256));
            if (mode == 13) hipLaunchKernelGGL((step_blocked<false, false>), grid, dim3(256), 0, 0, rec, cur, prev, nc);
            else if (mode == 14) hipLaunchKernelGGL((step_blocked<true, true>), grid, dim3(256), 0, 0, rec, cur, prev, nc);
            else hipLaunchKernelGGL((step_blocked<true, false>), grid, dim3(256), 0, 0, rec, cur, prev, nc);
        }
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double moved = mode <= 3 ? 2.0 * bytes : mode >= 13 ? 64.0 * (n4 * 4 / 12 / 32 * 32) : mode >= 10 ? 128.0 * (n4 / 7) : (double)bytes;
        const double gbs = moved / (ms * 1e-3) / 1e9;
        if (r > 0) sum += gbs;
        if (gbs > best) best = gbs;
    }
    printf("{\"mode\": %d, \"copy_mib\": %lld, \"reps\": %d, \"best_gbs\": %.1f, \"mean_gbs\": %.1f, "
           "\"cus\": %d}\n",
           mode, mib, reps, best, reps > 1 ? sum / (reps - 1) : best, p.multiProcessorCount);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
