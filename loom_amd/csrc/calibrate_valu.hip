// -*- C++ -*-
/// @file calibrate_valu.hip
/// @brief Vector-ALU rates of the device for reading the roofline of the
/// VALU-bound path: eight independent FP32 fma chains per lane (non-packed:
/// the library is built without packed FP32, so this is its FMA ceiling) and
/// the same shape with 32-bit integer xor/add.  Standalone executable; prints
/// the rates and exits.  The instruction counts are known exactly (8 per loop
/// trip per wave plus a few), which also calibrates SQ_INSTS_VALU.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace
{
constexpr int kBlocks  = 256 * 64; // 64 workgroups of 4 waves per CU
constexpr int kThreads = 256;
constexpr int kIters   = 4096;

__global__ __launch_bounds__(256) void
fma_chains(float* out, int iters, float b, float c)
{
    const int t  = blockIdx.x * blockDim.x + threadIdx.x;
    float     a0 = t, a1 = t + 1, a2 = t + 2, a3 = t + 3, a4 = t + 4, a5 = t + 5, a6 = t + 6,
          a7 = t + 7;
    for(int i = 0; i < iters; ++i)
    {
        a0 = __builtin_fmaf(a0, b, c);
        a1 = __builtin_fmaf(a1, b, c);
        a2 = __builtin_fmaf(a2, b, c);
        a3 = __builtin_fmaf(a3, b, c);
        a4 = __builtin_fmaf(a4, b, c);
        a5 = __builtin_fmaf(a5, b, c);
        a6 = __builtin_fmaf(a6, b, c);
        a7 = __builtin_fmaf(a7, b, c);
    }
    out[t] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

__global__ __launch_bounds__(256) void
int_chains(std::uint32_t* out, int iters, std::uint32_t b)
{
    const int     t  = blockIdx.x * blockDim.x + threadIdx.x;
    std::uint32_t a0 = t, a1 = t ^ 1u, a2 = t ^ 2u, a3 = t ^ 3u, a4 = t ^ 4u, a5 = t ^ 5u,
                  a6 = t ^ 6u, a7 = t ^ 7u;
    for(int i = 0; i < iters; ++i)
    {
        a0 = (a0 ^ b) + i;
        a1 = (a1 ^ b) + i;
        a2 = (a2 ^ b) + i;
        a3 = (a3 ^ b) + i;
        a4 = (a4 ^ b) + i;
        a5 = (a5 ^ b) + i;
        a6 = (a6 ^ b) + i;
        a7 = (a7 ^ b) + i;
    }
    out[t] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

#define CHECK(x)                                                                     \
    do                                                                               \
    {                                                                                \
        const hipError_t e_ = (x);                                                   \
        if(e_ != hipSuccess)                                                         \
        {                                                                            \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
            return 1;                                                                \
        }                                                                            \
    } while(0)
} // namespace

int
main()
{
    const std::size_t n   = std::size_t(kBlocks) * kThreads;
    void*             out = nullptr;
    CHECK(hipMalloc(&out, n * sizeof(float)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const double lane_ops = double(n) * kIters * 8;
    for(int rep = 0; rep < 3; ++rep)
    {
        float ms = 0;
        CHECK(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL(fma_chains, dim3(kBlocks), dim3(kThreads), 0, nullptr,
                           static_cast<float*>(out), kIters, 0.999f, 0.001f);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, nullptr));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("fp32 fma (non-packed): %.3f ms  %.1f TFLOP/s  %.1f T lane-fma/s\n", ms,
                    2 * lane_ops / ms / 1e9, lane_ops / ms / 1e9);
        CHECK(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL(int_chains, dim3(kBlocks), dim3(kThreads), 0, nullptr,
                           static_cast<std::uint32_t*>(out), kIters, 0x9e3779b9u);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, nullptr));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("int32 xor, add: %.3f ms  %.1f T lane-instructions/s\n", ms,
                    2 * lane_ops / ms / 1e9);
    }
    CHECK(hipFree(out));
    return 0;
}
