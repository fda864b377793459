// stream_ceiling.hip -- practical HBM ceiling of the SOR pass's access shape.
//
// Measures, with HIP events, three pure-streaming kernels on MI355X:
//   copy   : 1 FP64 load + 1 store per element (plain copy)
//   shape  : per point 9 doubles from operand blocks of 64 points, 2 doubles
//            from two plain arrays, 1 double stored -- the 96 bytes per colour
//            point a SOR pass must move, with no stencil and no arithmetic
//   read   : the same 11 loads, no store
// at the size of one colour of a 256x256x128x16 volume (67,108,864 points).
// The "shape" figure is the bandwidth a perfect SOR pass could reach.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                             \
        }                                                                         \
    } while (0)

__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ a, double* __restrict__ b,
                                              long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_shape(const double* __restrict__ ob,
                                               const double* __restrict__ a,
                                               const double* __restrict__ y,
                                               double* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* o = ob + (i >> 6) * 576 + (i & 63);
    double s = a[i] + y[i];
#pragma unroll
    for (int f = 0; f < 9; f++) s += o[f * 64];
    out[i] = s;
}

__global__ __launch_bounds__(256) void k_read(const double* __restrict__ ob,
                                              const double* __restrict__ a,
                                              const double* __restrict__ y,
                                              double* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    double s = 0;
    if (i < n) {
        const double* o = ob + (i >> 6) * 576 + (i & 63);
        s = a[i] + y[i];
#pragma unroll
        for (int f = 0; f < 9; f++) s += o[f * 64];
    }
    if (s == 12345.678) out[0] = s;  // data-dependent store keeps the loads
}

int main() {
    const long n = 67108864L;
    double *ob, *a, *y, *out;
    CHECK(hipMalloc(&ob, n * 9 * sizeof(double)));
    CHECK(hipMalloc(&a, n * sizeof(double)));
    CHECK(hipMalloc(&y, n * sizeof(double)));
    CHECK(hipMalloc(&out, n * sizeof(double)));
    CHECK(hipMemset(ob, 0, n * 9 * sizeof(double)));
    CHECK(hipMemset(a, 0, n * sizeof(double)));
    CHECK(hipMemset(y, 0, n * sizeof(double)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const char* names[3] = {"copy", "shape", "read"};
    const double bytes[3] = {16.0 * n, 96.0 * n, 88.0 * n};
    const unsigned grid = (unsigned)((n + 255) / 256);
    for (int kind = 0; kind < 3; kind++) {
        float best = 1e30f;
        for (int rep = 0; rep < 8; rep++) {
            CHECK(hipEventRecord(e0));
            if (kind == 0) k_copy<<<grid, 256>>>(a, out, n);
            else if (kind == 1) k_shape<<<grid, 256>>>(ob, a, y, out, n);
            else k_read<<<grid, 256>>>(ob, a, y, out, n);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) best = ms;
        }
        std::printf("%-6s %8.4f ms  %7.0f GB/s  (%.0f bytes)\n", names[kind], best,
                    bytes[kind] / best / 1e6, bytes[kind]);
    }
    CHECK(hipFree(ob));
    CHECK(hipFree(a));
    CHECK(hipFree(y));
    CHECK(hipFree(out));
    return 0;
}
