// stream_ceiling.hip -- practical HBM ceilings of the SOR pass's access shape.
//
// Measures, with HIP events, pure-streaming kernels on MI355X at the size of
// one colour of a 256x256x128x16 volume (67,108,864 points):
//   copy          1 FP64 load + 1 store per element (plain copy)
//   shape         per point 9 doubles from operand blocks of 64 points, 2
//                 doubles from two plain arrays, 1 double stored -- the 96
//                 bytes per colour point a SOR pass must move -- over
//                 contiguous arrays, no stencil, no arithmetic
//   read          the same 11 loads, no store
//   layout-flat   the shape in the solver's own layout: operand blocks per
//                 interior (k,l) plane with the row pitch RP, colour arrays
//                 that keep the ghost rows and planes; one thread per point
//   layout-lmarch the same layout in the SOR pass's order: one thread per
//                 (k, packed point), marching l two steps per load phase
// "layout-lmarch" is the bandwidth a SOR pass of this layout and order could
// reach; "shape" what a hole-free layout could.
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

// colour-layout geometry of 256x256x128x16: H packed columns, padded J x K x L,
// RP points per operand plane, NT interior packed points per plane
constexpr int kH = 129, kJ = 258, kK = 130, kL = 18, kRP = 33024, kNT = 256 * 129;
constexpr long kNPL = (long)(kK - 2) * (kL - 2);

__device__ __forceinline__ long oidx(int k, int l, int t) {
    const long plane = (long)(l - 1) * (kK - 2) + (k - 1);
    return plane * kRP * 9 + (long)(t >> 6) * 576 + (t & 63);
}

__device__ __forceinline__ long cidx(int k, int l, int t) {
    const int j = t / kH + 1, m = t % kH;
    return (((long)l * kK + k) * kJ + j) * kH + m;
}

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

__global__ __launch_bounds__(256) void k_layout_flat(const double* __restrict__ ob,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ out) {
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= kNPL * kNT) return;
    const int t = (int)(p % kNT);
    const long pl = p / kNT;
    const int k = (int)(pl % (kK - 2)) + 1, l = (int)(pl / (kK - 2)) + 1;
    const double* o = ob + oidx(k, l, t);
    const long c = cidx(k, l, t);
    double s = x[c] + y[c];
#pragma unroll
    for (int f = 0; f < 9; f++) s += o[f * 64];
    out[c] = s;
}

__global__ __launch_bounds__(256) void k_layout_lmarch(const double* __restrict__ ob,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       double* __restrict__ out) {
    const int runs = (kNT + 255) / 256;
    const int k = blockIdx.x / runs + 1, t = (blockIdx.x % runs) * 256 + threadIdx.x;
    if (t >= kNT) return;
    for (int l = 1; l <= kL - 2; l += 2) {
        const double* o0 = ob + oidx(k, l, t);
        const double* o1 = ob + oidx(k, l + 1, t);
        const long c0 = cidx(k, l, t), c1 = cidx(k, l + 1, t);
        double s0 = x[c0] + y[c0], s1 = x[c1] + y[c1];
#pragma unroll
        for (int f = 0; f < 9; f++) {
            s0 += o0[f * 64];
            s1 += o1[f * 64];
        }
        out[c0] = s0;
        out[c1] = s1;
    }
}

int main() {
    const long n = 67108864L;
    const long nob = kNPL * kRP * 9, nc = (long)kL * kK * kJ * kH;
    double *ob, *a, *y, *out;
    CHECK(hipMalloc(&ob, nob * sizeof(double)));  // >= n * 9
    CHECK(hipMalloc(&a, nc * sizeof(double)));    // >= n
    CHECK(hipMalloc(&y, nc * sizeof(double)));
    CHECK(hipMalloc(&out, nc * sizeof(double)));
    CHECK(hipMemset(ob, 0, nob * sizeof(double)));
    CHECK(hipMemset(a, 0, nc * sizeof(double)));
    CHECK(hipMemset(y, 0, nc * sizeof(double)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const char* names[5] = {"copy", "shape", "read", "layout-flat", "layout-lmarch"};
    const double npts = (double)kNPL * kNT;  // packed slots, incl. the half column
    const double bytes[5] = {16.0 * n, 96.0 * n, 88.0 * n, 96.0 * npts, 96.0 * npts};
    const unsigned grid = (unsigned)((n + 255) / 256);
    const unsigned gflat = (unsigned)((kNPL * kNT + 255) / 256);
    const unsigned gmarch = (unsigned)(((kNT + 255) / 256) * (kK - 2));
    for (int kind = 0; kind < 5; kind++) {
        float best = 1e30f;
        for (int rep = 0; rep < 10; rep++) {
            CHECK(hipEventRecord(e0));
            if (kind == 0) k_copy<<<grid, 256>>>(a, out, n);
            else if (kind == 1) k_shape<<<grid, 256>>>(ob, a, y, out, n);
            else if (kind == 2) k_read<<<grid, 256>>>(ob, a, y, out, n);
            else if (kind == 3) k_layout_flat<<<gflat, 256>>>(ob, a, y, out);
            else k_layout_lmarch<<<gmarch, 256>>>(ob, a, y, out);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (rep >= 2 && ms < best) best = ms;
        }
        std::printf("%-14s %8.4f ms  %7.0f GB/s  %5.1f%% of 8 TB/s  (%.0f bytes)\n", names[kind],
                    best, bytes[kind] / best / 1e6, bytes[kind] / best / 1e6 / 80.0, bytes[kind]);
    }
    CHECK(hipFree(ob));
    CHECK(hipFree(a));
    CHECK(hipFree(y));
    CHECK(hipFree(out));
    return 0;
}
