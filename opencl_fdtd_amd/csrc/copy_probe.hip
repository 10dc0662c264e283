// HBM probe: plain copy kernels of known size, an E-step-shaped stream
// (9 arrays read, 3 written, no neighbours) and a stream shaped like the
// fused H+E pass (9 read, 9 written: E, H, eps in; E, H, S out).  Used to measure the practical
// bandwidth ceiling of the GPU and to calibrate the rocprofv3 FETCH_SIZE /
// WRITE_SIZE counters (bytes per KiB reported) against a known byte count.
//   copy_probe [floats per buffer, default 2^28] [repetitions, default 20]
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

__global__ void copy1(const float* __restrict__ a, float* __restrict__ b, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) b[i] = a[i];
}
__global__ void copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += s) b[i] = a[i];
}
// 9 streams read, 3 written (E-step byte pattern without neighbours)
__global__ void e_like(const float* __restrict__ h, const float* __restrict__ eps, float* __restrict__ e, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) {
        float a0 = h[i], a1 = h[i + n], a2 = h[i + 2 * n];
        float q0 = eps[i], q1 = eps[i + n], q2 = eps[i + 2 * n];
        e[i] += a0 / q0; e[i + n] += a1 / q1; e[i + 2 * n] += a2 / q2;
    }
}

// 9 streams read, 9 written (the fused H+E pass's byte pattern without
// neighbours or psi)
__global__ void he_like(const float* __restrict__ e, const float* __restrict__ h, const float* __restrict__ eps,
                        float* __restrict__ eo, float* __restrict__ ho, float* __restrict__ so, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) {
        float e0 = e[i], e1 = e[i + n], e2 = e[i + 2 * n];
        float h0 = h[i], h1 = h[i + n], h2 = h[i + 2 * n];
        float q0 = eps[i], q1 = eps[i + n], q2 = eps[i + 2 * n];
        ho[i] = h0 - e1; ho[i + n] = h1 - e2; ho[i + 2 * n] = h2 - e0;
        so[i] = e1 * h2; so[i + n] = e2 * h0; so[i + 2 * n] = e0 * h1;
        eo[i] = e0 + h0 / q0; eo[i + n] = e1 + h1 / q1; eo[i + 2 * n] = e2 + h2 / q2;
    }
}

int main(int argc, char** argv) {
    size_t n = (argc > 1) ? atoll(argv[1]) : (size_t)256 << 20;   // floats per buffer
    int reps = (argc > 2) ? atoi(argv[2]) : 20;
    float *a, *b, *c;
    CK(hipMalloc(&a, n * 4 * 3));
    CK(hipMalloc(&b, n * 4 * 3));
    CK(hipMalloc(&c, n * 4 * 3));
    CK(hipMemset(a, 0, n * 12)); CK(hipMemset(b, 0, n * 12)); CK(hipMemset(c, 0, n * 12));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    int grids[] = {1024, 4096, 16384};
    for (int g : grids) {
        float ms;
        hipLaunchKernelGGL(copy1, dim3(g), dim3(256), 0, 0, a, b, n);
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(copy1, dim3(g), dim3(256), 0, 0, a, b, n);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1)); CK(hipEventElapsedTime(&ms, e0, e1));
        printf("copy1 grid %6d: %.1f GB/s (%.1f us per launch, %zu bytes moved)\n", g, 8.0 * n * reps / ms / 1e6, ms * 1e3 / reps, 8 * n);
        hipLaunchKernelGGL(copy4, dim3(g), dim3(256), 0, 0, (const float4*)a, (float4*)b, n / 4);
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(copy4, dim3(g), dim3(256), 0, 0, (const float4*)a, (float4*)b, n / 4);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1)); CK(hipEventElapsedTime(&ms, e0, e1));
        printf("copy4 grid %6d: %.1f GB/s\n", g, 8.0 * n * reps / ms / 1e6);
        size_t m = n;   // 3 components of m floats in each of a (h), b (eps), c (e)
        hipLaunchKernelGGL(e_like, dim3(g), dim3(256), 0, 0, a, b, c, m);
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(e_like, dim3(g), dim3(256), 0, 0, a, b, c, m);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1)); CK(hipEventElapsedTime(&ms, e0, e1));
        printf("e_like grid %6d: %.1f GB/s\n", g, 48.0 * m * reps / ms / 1e6);
    }
    // he_like on six buffers of 3m floats (m = 16 Mi cells: 1.2 GB moved per launch)
    {
        const size_t m = (size_t)16 << 20;
        float* f[6];
        for (int k = 0; k < 6; ++k) { CK(hipMalloc(&f[k], m * 12)); CK(hipMemset(f[k], 0, m * 12)); }
        CK(hipMemset(f[2], 0x3f, m * 12));   // eps ~ 0.75, nonzero
        for (int g : grids) {
            float ms;
            hipLaunchKernelGGL(he_like, dim3(g), dim3(256), 0, 0, f[0], f[1], f[2], f[3], f[4], f[5], m);
            CK(hipEventRecord(e0));
            for (int r = 0; r < reps; r++)
                hipLaunchKernelGGL(he_like, dim3(g), dim3(256), 0, 0, f[0], f[1], f[2], f[3], f[4], f[5], m);
            CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1)); CK(hipEventElapsedTime(&ms, e0, e1));
            printf("he_like grid %6d: %.1f GB/s\n", g, 72.0 * m * reps / ms / 1e6);
        }
        for (int k = 0; k < 6; ++k) CK(hipFree(f[k]));
    }
    CK(hipDeviceSynchronize());
    printf("done\n");
    return 0;
}
