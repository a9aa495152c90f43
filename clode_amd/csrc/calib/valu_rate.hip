// Vector-ALU fused multiply-add throughput of the device, FP32 and FP64.
//
// Each lane runs eight independent FMA chains, so with a few waves per SIMD
// the dependent-issue latency is hidden and the rate approaches the issue
// limit.  Built, like all device code of this project, with packed FP32
// disabled: the FP32 figure is the scalar (non-packed) rate that the fp32c
// integration kernels run at.  Prints one JSON line per waves-per-SIMD
// setting.  Used by `bench.py --calibrate-fma` to state the achievable peak
// beside the nominal one.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                                   \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "valu_rate: %s: %s\n", #x, hipGetErrorString(e_));               \
            std::exit(1);                                                                          \
        }                                                                                          \
    } while (0)

__device__ inline float fma_t(float x, float a, float b) { return __builtin_fmaf(x, a, b); }
__device__ inline double fma_t(double x, double a, double b) { return __builtin_fma(x, a, b); }

template <typename T>
__global__ void __launch_bounds__(256) chains(T *out, int iters, T a, T b)
{
    T x0 = (T)threadIdx.x, x1 = x0 + (T)1, x2 = x0 + (T)2, x3 = x0 + (T)3;
    T x4 = x0 + (T)4, x5 = x0 + (T)5, x6 = x0 + (T)6, x7 = x0 + (T)7;
    for (int i = 0; i < iters; ++i) {
        x0 = fma_t(x0, a, b);
        x1 = fma_t(x1, a, b);
        x2 = fma_t(x2, a, b);
        x3 = fma_t(x3, a, b);
        x4 = fma_t(x4, a, b);
        x5 = fma_t(x5, a, b);
        x6 = fma_t(x6, a, b);
        x7 = fma_t(x7, a, b);
    }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}

// TFLOP/s of one timed launch of `blocks` workgroups of 256 lanes
template <typename T>
double rate(int blocks, int iters)
{
    const size_t n = (size_t)blocks * 256;
    T *d = nullptr;
    CHECK(hipMalloc(&d, n * sizeof(T)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    chains<T><<<blocks, 256>>>(d, 100, (T)0.999, (T)0.001);  // first launch, untimed
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    chains<T><<<blocks, 256>>>(d, iters, (T)0.999, (T)0.001);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    CHECK(hipFree(d));
    return (double)n * iters * 8 * 2 / (ms * 1e-3) / 1e12;
}

int main()
{
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const int iters = 20000;
    for (int waves : {1, 2, 4, 8}) {
        const int blocks = cus * waves;  // a workgroup is 4 waves, one per SIMD of a CU
        const double f32 = rate<float>(blocks, iters);
        const double f64 = rate<double>(blocks, iters);
        std::printf("{\"waves_per_simd\": %d, \"fp32_fma_tflops\": %.2f, \"fp64_fma_tflops\": %.2f}\n", waves, f32,
                    f64);
        std::fflush(stdout);
    }
    return 0;
}
