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
int main(int argc, char **argv)
{
    const long long mib = argc > 1 ? atoll(argv[1]) : 2048;
    const int reps = argc > 2 ? atoi(argv[2]) : 20;
    const int mode = argc > 3 ? atoi(argv[3]) : 0;
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
        else
            hipLaunchKernelGGL(mix_k1<1>, dim3((n4 / 2 + 255) / 256), dim3(256), 0, 0, a, b, n4 / 2);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double moved = mode <= 3 ? 2.0 * bytes : (double)bytes;
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
