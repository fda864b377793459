// seg4d_net.hip -- the learned warm start (seg4d_net.h): parameters, the
// training pool, forward / backward passes on the kernels of
// seg4d_net_kernels.h, Adam, and the prediction used by Engine::setup.
#include "seg4d_net.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "seg4d_engine.h"
#include "seg4d_net_kernels.h"

namespace s4d {

using net::bf16;
using net::bf16x8;
using net::kCh;
using net::NGeo;
using net::Taps;

namespace {

constexpr int kPatchCore = 64;  // loss points per axis of a training patch
constexpr int kHalo = 4;        // receptive field of the four spatial convs
constexpr int kPatch = kPatchCore + 2 * kHalo + 2;  // padded patch extent (74)
constexpr int kSpatialTaps = 27, kTemporalTaps = 3;

long cdivl(long a, long b) { return (a + b - 1) / b; }

// Blocks of k_pred_moments over nv voxels: at most 2048, each a range of *per
// voxels (a multiple of 256).
int sumsq_blocks(long nv, long* per) {
    const long nb = std::min<long>(2048, cdivl(nv, 256));
    *per = cdivl(cdivl(nv, nb), 256) * 256;
    return (int)cdivl(nv, *per);
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// 3x3x3 taps, t = 9 (dz+1) + 3 (dy+1) + (dx+1); rows q = 3 (dz+1) + (dy+1)
Taps spatial_taps(const NGeo& G, int sign) {
    Taps T;
    std::memset(&T, 0, sizeof T);
    T.n = kSpatialTaps;
    T.nrow = 9;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            const int q = 3 * (dz + 1) + (dy + 1);
            T.rowoff[q] = sign * (G.I * (dy + G.J * dz));
            for (int dx = -1; dx <= 1; dx++) {
                const int t = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1);
                T.off[t] = sign * (dx + G.I * (dy + G.J * dz));
                T.trow[t] = (signed char)q;
                T.tdx[t] = (signed char)(sign * dx);
            }
        }
    return T;
}

// 3-tap temporal conv, t = dl + 1
Taps temporal_taps(const NGeo& G, int sign) {
    Taps T;
    std::memset(&T, 0, sizeof T);
    T.n = kTemporalTaps;
    T.nrow = 3;
    for (int dl = -1; dl <= 1; dl++) {
        const int t = dl + 1;
        T.off[t] = sign * (G.I * G.J * G.K * dl);
        T.rowoff[t] = T.off[t];
        T.trow[t] = (signed char)t;
        T.tdx[t] = 0;
    }
    return T;
}

// ---- small kernels -------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gmean(NGeo G, net::bf16* __restrict__ out,
                                               const double* g0, const double* g1,
                                               const double* g2, const double* g3,
                                               const double* g4, const double* g5,
                                               const double* g6, const double* g7) {
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    const long nv = (long)G.I * G.J * G.K * G.L;
    if (v >= nv) return;
    const double s = ((((((g0[v] + g1[v]) + g2[v]) + g3[v]) + g4[v]) + g5[v]) + g6[v]) + g7[v];
    out[v] = (bf16)(float)(s * 0.125);
}

// (i, j, k, l) of voxel v; 32-bit arithmetic (the grids here hold fewer than
// 2^31 voxels, checked on the host).
__device__ __forceinline__ bool point_of(const NGeo& G, long v, int& i, int& j, int& k, int& l) {
    const unsigned uv = (unsigned)v;
    const unsigned r1 = uv / (unsigned)G.I;
    i = (int)(uv - r1 * (unsigned)G.I);
    const unsigned r2 = r1 / (unsigned)G.J;
    j = (int)(r1 - r2 * (unsigned)G.J);
    l = (int)(r2 / (unsigned)G.K);
    k = (int)(r2 - (unsigned)l * (unsigned)G.K);
    return net::interior(G, i, j, k, l);
}

// sigma = sqrt(sum / n) of the increment u - prev over the interior, from the
// device sum *dsum (0 if there is no increment)
__device__ __forceinline__ double step_sigma(const double* dsum, double inv_n) {
    return dsum ? sqrt(*dsum * inv_n) : 0.0;
}

// Third difference u - 3 h1 + 3 h2 - h3 of four consecutive levels: the
// error the quadratic extrapolation made at the last step, whose rms kappa
// scales the cubic base's correction.
__device__ __forceinline__ double third_diff(double uc, double p1, double p2, double p3) {
    return ((uc - 3.0 * p1) + 3.0 * p2) - p3;
}

// Features of the prediction, unscaled: [u^n, d, g_mean, t (order 3)] as 4
// BF16 channels per point (every point: ghosts zero), with d = u^n - u^(n-1)
// and t the third difference (the network's first layer applies s_d / sigma
// and s_e / kappa through its weights, see k_scale_w0), and per-block partial
// sums of d^2 and t^2 over the interior (partial[2b], partial[2b + 1]): one
// pass over the four levels.
__global__ __launch_bounds__(256) void k_features(NGeo G, int order, const double* __restrict__ u,
                                                  const double* __restrict__ h1,
                                                  const double* __restrict__ h2,
                                                  const double* __restrict__ h3,
                                                  const bf16* __restrict__ gm,
                                                  bf16* __restrict__ feat,
                                                  double* __restrict__ partial) {
    __shared__ double sm[2][256];
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    double s = 0.0, s3 = 0.0;
    int i, j, k, l;
    if (v < (long)G.I * G.J * G.K * G.L) {
        net::bf16x4 f;
#pragma unroll
        for (int q = 0; q < 4; q++) f[q] = (bf16)0.0f;
        if (point_of(G, v, i, j, k, l)) {
            // order 0: no history, the increment feature is zero
            const double uc = u[v], p1 = order >= 1 ? h1[v] : uc;
            const double d = uc - p1;
            f[0] = (bf16)(float)uc;
            f[1] = (bf16)(float)d;
            f[2] = gm[v];
            s = d * d;
            if (order == 3) {
                const double t = third_diff(uc, p1, h2[v], h3[v]);
                f[3] = (bf16)(float)t;
                s3 = t * t;
            }
        }
        *reinterpret_cast<net::bf16x4*>(feat + v * 4) = f;
    }
    sm[0][threadIdx.x] = s;
    sm[1][threadIdx.x] = s3;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + w];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sm[0][0];
        partial[2 * blockIdx.x + 1] = sm[1][0];
    }
}

// First level of a fixed-order two-column reduction: out[2c + q] = sum of
// partial[2b + q] over the chunk b in [c chunk, (c + 1) chunk).
__global__ __launch_bounds__(256) void k_chunk_sums2(const double* __restrict__ partial, long nb,
                                                     long chunk, double* __restrict__ out) {
    __shared__ double sm[2][256];
    const long b0 = (long)blockIdx.x * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
    double s = 0.0, s3 = 0.0;
    for (long b = b0 + threadIdx.x; b < b1; b += 256) {
        s += partial[2 * b];
        s3 += partial[2 * b + 1];
    }
    sm[0][threadIdx.x] = s;
    sm[1][threadIdx.x] = s3;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + w];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sm[0][0];
        out[2 * blockIdx.x + 1] = sm[1][0];
    }
}

// BF16 spatial weights of the first layer for the prediction's 4 input
// channels, the feature scales folded in: out[t][ci][co] = w[t][ci][co] sc(ci)
// (ci < 4 of the 8 stored), sc(1) = s_d / sigma, sc(3) = s_e / kappa (0 where
// that rms is 0), 1 otherwise; sigma and kappa from the device sums
// (sums[0], sums[1]) over the interior points (inv_n = 1 / their count).
__global__ __launch_bounds__(256) void k_scale_w0(const float* __restrict__ w, int n,
                                                  const double* __restrict__ sums, double inv_n,
                                                  float sd, float se, bf16* __restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int co = e % kCh, ci = (e / kCh) % 4, t = e / (kCh * 4);
    float sc = 1.f;
    if (ci == 1 || ci == 3) {
        const double r = sqrt(sums[ci == 1 ? 0 : 1] * inv_n);
        sc = r > 0.0 ? (float)((ci == 1 ? sd : se) / r) : 0.f;
    }
    out[e] = (bf16)(w[(t * net::kCin0 + ci) * kCh + co] * sc);
}

// Per-block partial sums (FP64) over the interior of d^2 = (u - h1)^2 and
// (with h2, h3) of the third difference squared; block b covers the voxels
// [b per, (b + 1) per) (per a multiple of 256), so that a volume gives at
// most a few thousand partials: partial[2b], partial[2b + 1].
__global__ __launch_bounds__(256) void k_pred_moments(NGeo G, const double* __restrict__ u,
                                                      const double* __restrict__ h1,
                                                      const double* __restrict__ h2,
                                                      const double* __restrict__ h3, long per,
                                                      double* __restrict__ partial) {
    __shared__ double sm[2][256];
    const long nv = (long)G.I * G.J * G.K * G.L;
    const long v0 = (long)blockIdx.x * per, v1 = v0 + per < nv ? v0 + per : nv;
    double s = 0, s3 = 0;
    for (long v = v0 + threadIdx.x; v < v1; v += 256) {
        int i, j, k, l;
        if (point_of(G, v, i, j, k, l)) {
            const double uc = u[v], p1 = h1[v];
            s += (uc - p1) * (uc - p1);
            if (h3) {
                const double t = third_diff(uc, p1, h2[v], h3[v]);
                s3 += t * t;
            }
        }
    }
    sm[0][threadIdx.x] = s;
    sm[1][threadIdx.x] = s3;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + w];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sm[0][0];
        partial[2 * blockIdx.x + 1] = sm[1][0];
    }
}

// One patch of a training batch: patch point (pi, pj, pk, pl) of batch slot
// b reads run point (ax + pi, ay + pj, az + pk, pl).  Patch ghosts are zero;
// loss points are the interior points at least kHalo + 1 from a patch
// ghost that is not also a ghost of the run.  um1..um3: levels n-1..n-3
// (um2, um3 only for the cubic base), u0 level n, up1 level n+1.
__global__ __launch_bounds__(256) void k_gather_patch(
    NGeo PG, int b, int RI, int RJ, int RK, int order, const double* __restrict__ um3,
    const double* __restrict__ um2, const double* __restrict__ um1, const double* __restrict__ u0,
    const double* __restrict__ up1, const bf16* __restrict__ gm, int ax, int ay, int az, float sd,
    float se, float st, double inv_sig, double inv_kap, bf16* __restrict__ feat,
    float* __restrict__ target, unsigned char* __restrict__ lmask) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    const long per = (long)PG.I * PG.J * PG.K * PG.P;
    if (q >= per) return;
    const int pi = (int)(q % PG.I);
    const long r1 = q / PG.I;
    const int pj = (int)(r1 % PG.J);
    const long r2 = r1 / PG.J;
    const int pk = (int)(r2 % PG.K);
    const int pl = (int)(r2 / PG.K);
    const long v = (((long)(b * PG.P + pl) * PG.K + pk) * PG.J + pj) * PG.I + pi;
    const bool ghost = pi == 0 || pi == PG.I - 1 || pj == 0 || pj == PG.J - 1 || pk == 0 ||
                       pk == PG.K - 1 || pl == 0 || pl == PG.P - 1;
    bf16x8 f;
#pragma unroll
    for (int c = 0; c < 8; c++) f[c] = (bf16)0.0f;
    float tg = 0.f;
    unsigned char lm = 0;
    if (!ghost) {
        const long rv = (((long)pl * RK + (az + pk)) * RJ + (ay + pj)) * RI + (ax + pi);
        const double uc = u0[rv], um = order >= 1 ? um1[rv] : uc, up = up1[rv];
        f[0] = (bf16)(float)uc;
        f[1] = (bf16)((float)((uc - um) * inv_sig) * sd);
        f[2] = gm[rv];
        if (order == 0) {
            tg = (float)(up - uc) * st;  // the increment itself (scale q = 1)
        } else if (order == 3) {
            const double p2 = um2[rv], p3 = um3[rv];
            const double base = ((4.0 * uc - 6.0 * um) + 4.0 * p2) - p3;
            f[3] = (bf16)((float)(third_diff(uc, um, p2, p3) * inv_kap) * se);
            tg = (float)((up - base) * inv_kap) * st;
        } else {
            tg = (float)((up - (uc + (uc - um))) * inv_sig) * st;
        }
        auto ok = [](int p, int n, int a, int rn) {
            return (p >= kHalo + 1 || a == 0) && (p <= n - kHalo - 2 || a + n == rn);
        };
        lm = ok(pi, PG.I, ax, RI) && ok(pj, PG.J, ay, RJ) && ok(pk, PG.K, az, RK) ? 1 : 0;
    }
    *reinterpret_cast<bf16x8*>(feat + v * 8) = f;
    target[v] = tg;
    lmask[v] = lm;
}

// Per-block partial sums (FP64) over the interior of a run level n of d^2,
// e^2 and c^2: d = u^n - u^(n-1), e = the third difference (order 3; 0 for
// the linear base), c = u^(n+1) - base.
__global__ __launch_bounds__(256) void k_moments(NGeo G, int order, const double* __restrict__ um3,
                                                 const double* __restrict__ um2,
                                                 const double* __restrict__ um1,
                                                 const double* __restrict__ u0,
                                                 const double* __restrict__ up1,
                                                 double* __restrict__ partial) {
    __shared__ double s1[256], s2[256], s3[256];
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    double a = 0, e2 = 0, c = 0;
    int i, j, k, l;
    if (v < (long)G.I * G.J * G.K * G.L && point_of(G, v, i, j, k, l)) {
        const double uc = u0[v], um = order >= 1 ? um1[v] : uc;
        const double d = uc - um;
        double base = order >= 1 ? uc + d : uc, t3 = 0.0;
        if (order == 3) {
            base = ((4.0 * uc - 6.0 * um) + 4.0 * um2[v]) - um3[v];
            t3 = third_diff(uc, um, um2[v], um3[v]);
        }
        const double e = up1[v] - base;
        a = d * d;
        e2 = t3 * t3;
        c = e * e;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = e2;
    s3[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
            s3[threadIdx.x] += s3[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = s1[0];
        partial[3 * blockIdx.x + 1] = s2[0];
        partial[3 * blockIdx.x + 2] = s3[0];
    }
}

// out[e] = sum_b partial[b * stride + col(e)], one block per output, fixed
// order (strided lane sums, then a tree).  col(e) maps the output index to
// the partial layout: for a weight gradient of `cin` input channels the
// partial rows are 16 per tap.
__global__ __launch_bounds__(256) void k_reduce_cols(const float* __restrict__ partial, int nb,
                                                     long stride, int cin, int n, float scale,
                                                     float* __restrict__ out) {
    __shared__ float sm[256];
    const int e = blockIdx.x;
    if (e >= n) return;
    int col = e;
    if (cin > 0) {
        const int co = e % kCh, r = e / kCh, ci = r % cin, t = r / cin;
        col = (t * 16 + ci) * kCh + co;
    }
    float s = 0.f;
    for (int b = threadIdx.x; b < nb; b += 256) s += partial[(long)b * stride + col];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = scale * sm[0];
}

__global__ __launch_bounds__(256) void k_reduce_d(const double* __restrict__ partial, int nb,
                                                  int ncol, double* __restrict__ out) {
    __shared__ double sm[256];
    for (int c = 0; c < ncol; c++) {
        double s = 0;
        for (int b = threadIdx.x; b < nb; b += 256) s += partial[(long)b * ncol + c];
        sm[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = sm[0];
        __syncthreads();
    }
}

// wt[t][co][ci] = w[t][ci][co] for 16 x 16 taps
__global__ __launch_bounds__(256) void k_transpose_taps(const bf16* __restrict__ w, int ntaps,
                                                        bf16* __restrict__ wt) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ntaps * 256) return;
    const int t = e / 256, ci = (e / 16) % 16, co = e % 16;
    wt[(t * 16 + co) * 16 + ci] = w[e];
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float* __restrict__ a, int n,
                                                     bf16* __restrict__ b) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < n) b[e] = (bf16)a[e];
}

}  // namespace

// ---- Net -------------------------------------------------------------------------
Net::~Net() { release(); }

int Net::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return -1;
}

int Net::layout() {
    int o = 0;
    for (int L = 0; L < nlayers_; L++) {
        const int cin = L == 0 ? net::kCin0 : kCh;
        off_ws[L] = o;
        o += net::kMaxTaps * cin * kCh;
        off_wt[L] = o;
        o += 4 * kCh * kCh;
        off_b[L] = o;
        o += kCh;
    }
    off_wp = o;
    o += kCh;
    off_bp = o;
    o += 1;
    nparam = o;
    return o;
}

void Net::release() {
    auto fr = [](auto*& p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    fr(wm_);
    fr(wb_);
    fr(wtr_);
    fr(m1_);
    fr(m2_);
    fr(grad_);
    fr(w0s_);
    for (Run& r : runs_) {
        fr(r.gmean);
        for (double*& s : r.snaps) fr(s);
    }
    runs_.clear();
    fr(feat_);
    for (int L = 0; L < kMaxLayers; L++) {
        fr(S_[L]);
        fr(Y_[L]);
    }
    fr(dA_);
    fr(dB_);
    fr(out_);
    fr(target_);
    fr(lmask_);
    fr(partial_);
    fr(dscal_);
    fr(gmean_);
    if (hscal_) (void)hipHostFree(hscal_);
    hscal_ = nullptr;
    aI_ = aJ_ = aK_ = aL_ = 0;
    g_version_ = -1;
    gpts_ = 0;
    adam_t_ = 0;
}

int Net::init(uint64_t seed, int layers, int base) {
    release();
    if (layers < 1 || layers > kMaxLayers || (base != 0 && base != 1 && base != 3)) {
        err = "net: layers must be 1..4 and the base order 0, 1 or 3";
        return -1;
    }
    nlayers_ = layers;
    base_ = base;
    layout();
    const size_t nb = (size_t)nparam;
    if (check(hipMalloc((void**)&wm_, nb * 4), "hipMalloc net") ||
        check(hipMalloc((void**)&wb_, nb * 2), "hipMalloc net") ||
        check(hipMalloc((void**)&wtr_, nb * 2), "hipMalloc net") ||
        check(hipMalloc((void**)&m1_, nb * 4), "hipMalloc net") ||
        check(hipMalloc((void**)&m2_, nb * 4), "hipMalloc net") ||
        check(hipMalloc((void**)&grad_, nb * 4), "hipMalloc net") ||
        check(hipMalloc((void**)&w0s_, (size_t)net::kMaxTaps * net::kCin0 * kCh * 2),
              "hipMalloc net") ||
        check(hipMalloc((void**)&dscal_, 64 * sizeof(double)), "hipMalloc net") ||
        check(hipHostMalloc((void**)&hscal_, 64 * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc net"))
        return -1;
    // He initialisation: N(0, 2 / fan_in) as a sum of uniforms (Irwin-Hall 12)
    std::vector<float> w(nb, 0.f);
    uint64_t s = seed;
    auto normal = [&]() {
        double a = 0;
        for (int q = 0; q < 12; q++) a += uniform01(s);
        return a - 6.0;
    };
    for (int L = 0; L < nlayers_; L++) {
        const int cin = L == 0 ? net::kCin0 : kCh;
        const int cin_used = L == 0 ? (base_ == 3 ? 4 : 3) : kCh;
        const double ss = std::sqrt(2.0 / (kSpatialTaps * cin_used));
        for (int t = 0; t < kSpatialTaps; t++)
            for (int ci = 0; ci < cin_used; ci++)
                for (int co = 0; co < kCh; co++)
                    w[off_ws[L] + (t * cin + ci) * kCh + co] = (float)(ss * normal());
        const double st = std::sqrt(2.0 / (kTemporalTaps * kCh));
        for (int t = 0; t < kTemporalTaps; t++)
            for (int ci = 0; ci < kCh; ci++)
                for (int co = 0; co < kCh; co++)
                    w[off_wt[L] + (t * kCh + ci) * kCh + co] = (float)(st * normal());
    }
    // the projection starts at zero: the initial prediction is the base
    // extrapolation itself
    if (check(hipMemcpy(wm_, w.data(), nb * 4, hipMemcpyHostToDevice), "net weights") ||
        check(hipMemset(m1_, 0, nb * 4), "net memset") ||
        check(hipMemset(m2_, 0, nb * 4), "net memset") ||
        check(hipMemset(grad_, 0, nb * 4), "net memset") ||
        check(hipMemset(wb_, 0, nb * 2), "net memset") ||
        check(hipMemset(wtr_, 0, nb * 2), "net memset"))
        return -1;
    {
        // dscal_[10] = 1: the target scale q of the base order 0
        const double one = 1.0;
        if (check(hipMemcpy(dscal_ + 10, &one, sizeof one, hipMemcpyHostToDevice), "net const"))
            return -1;
    }
    s_d = s_t = s_e = 1.f;
    adam_t_ = 0;
    return refresh();
}

int Net::ensure_acts(int I, int J, int K, int L, bool training) {
    if (I == aI_ && J == aJ_ && K == aK_ && L == aL_ && training == a_training_ && feat_)
        return 0;
    auto fr = [](auto*& p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    fr(feat_);
    for (int q = 0; q < kMaxLayers; q++) {
        fr(S_[q]);
        fr(Y_[q]);
    }
    fr(dA_);
    fr(dB_);
    fr(out_);
    fr(target_);
    fr(lmask_);
    fr(partial_);
    aI_ = aJ_ = aK_ = aL_ = 0;
    const size_t pts = (size_t)I * J * K * L;
    auto al = [&](auto** p, size_t bytes) {
        if (check(hipMalloc((void**)p, bytes), "hipMalloc net activations")) return -1;
        return check(hipMemset(*p, 0, bytes), "net memset");
    };
    if (al(&feat_, pts * 8 * 2)) return -1;
    const int nS = training ? nlayers_ : 1, nY = training ? nlayers_ : 2;
    for (int q = 0; q < nS; q++)
        if (al(&S_[q], pts * kCh * 2)) return -1;
    for (int q = 0; q < nY; q++)
        if (al(&Y_[q], pts * kCh * 2)) return -1;
    if (al(&out_, pts * 4)) return -1;
    if (training) {
        if (al(&dA_, pts * kCh * 2) || al(&dB_, pts * kCh * 2) || al(&target_, pts * 4) ||
            al(&lmask_, pts))
            return -1;
    }
    // partials (floats): weight gradients (2048 blocks x 28 x 16 x 16), loss
    // rows (18 per block of 256 points), or the prediction's two FP64 sums per
    // block of 256 points and per chunk of 1024 blocks
    const long nb_loss = cdivl((long)pts, 256);
    npartial_ = std::max<long>(2048L * net::kMaxTaps * 16 * kCh, nb_loss * 18);
    npartial_ = std::max<long>(npartial_, (nb_loss * 2 + cdivl(nb_loss, 1024) * 2) * 2);
    if (al(&partial_, (size_t)npartial_ * 4)) return -1;
    aI_ = I;
    aJ_ = J;
    aK_ = K;
    aL_ = L;
    a_training_ = training;
    return 0;
}

// Blocks of k_conv_s: 16-point x tiles x 4 x 4 row groups x interior frames.
static unsigned spatial_blocks(const NGeo& G) {
    const long n = (long)((G.I - 2 + 15) / 16) * ((G.J - 2 + 3) / 4) * ((G.K - 2 + 3) / 4) *
                   ((G.L / G.P) * (G.P - 2));
    return (unsigned)n;
}

static int conv_waves(long nrow, int* rows_per_wave) {
    const long waves = std::min<long>(nrow, 256L * 16);
    *rows_per_wave = (int)cdivl(nrow, waves);
    return (int)cdivl(nrow, *rows_per_wave);
}

int Net::forward(const NGeo& G, hipStream_t st, bool training, const bf16* w0) {
    const long nrow = (long)(G.J - 2) * (G.K - 2) * ((G.L / G.P) * (G.P - 2));
    int rpw = 1;
    const int waves = conv_waves(nrow, &rpw);
    const dim3 grid((unsigned)cdivl(waves, 4));
    const Taps Tt = temporal_taps(G, 1);
    const dim3 sgrid(spatial_blocks(G));
    const bf16* X = feat_;
    if (!training) {
        // inference: one fused kernel per layer, S kept on chip
        const long nlb = (long)((G.I - 2 + 15) / 16) * ((G.J - 2 + 3) / 4) * ((G.K - 2 + 3) / 4) *
                         (G.L / G.P);
        // the last layer's epilogue is the 1x1 projection (out_, FP32)
        for (int L = 0; L < nlayers_; L++) {
            bf16* Y = Y_[L & 1];
            const bool last = L == nlayers_ - 1;
            const dim3 lg((unsigned)nlb);
            const bf16 *ws = (L == 0 && w0) ? w0 : wb_ + off_ws[L], *wt = wb_ + off_wt[L];
            const float* b = wm_ + off_b[L];
            const float *wp = wm_ + off_wp, *bp = wm_ + off_bp;
            if (L == 0 && w0 && last)  // the prediction's 4-channel features
                net::k_layer_fwd<4, true><<<lg, 256, 0, st>>>(G, X, ws, wt, b, Y, wp, bp, out_);
            else if (L == 0 && w0)
                net::k_layer_fwd<4, false><<<lg, 256, 0, st>>>(G, X, ws, wt, b, Y, wp, bp, out_);
            else if (L == 0 && last)
                net::k_layer_fwd<8, true><<<lg, 256, 0, st>>>(G, X, ws, wt, b, Y, wp, bp, out_);
            else if (L == 0)
                net::k_layer_fwd<8, false><<<lg, 256, 0, st>>>(G, X, ws, wt, b, Y, wp, bp, out_);
            else if (last)
                net::k_layer_fwd<16, true><<<lg, 256, 0, st>>>(G, X, ws, wt, b, Y, wp, bp, out_);
            else
                net::k_layer_fwd<16, false><<<lg, 256, 0, st>>>(G, X, ws, wt, b, Y, wp, bp, out_);
            X = Y;
        }
        return check(hipGetLastError(), "net forward");
    }
    for (int L = 0; L < nlayers_; L++) {
        bf16* S = training ? S_[L] : S_[0];
        bf16* Y = training ? Y_[L] : Y_[L & 1];
        if (L == 0)
            net::k_conv_s<8, net::EPI_NONE><<<sgrid, 256, 0, st>>>(G, 1, X, wb_ + off_ws[L],
                                                                  nullptr, nullptr, S);
        else
            net::k_conv_s<16, net::EPI_NONE><<<sgrid, 256, 0, st>>>(G, 1, X, wb_ + off_ws[L],
                                                                   nullptr, nullptr, S);
        net::k_conv<16, kTemporalTaps, net::EPI_BIAS_RELU><<<grid, 256, 0, st>>>(
            G, Tt, S, wb_ + off_wt[L], wm_ + off_b[L], nullptr, Y, rpw);
        X = Y;
    }
    const long nv = (long)G.I * G.J * G.K * G.L;
    net::k_project<<<(unsigned)cdivl(nv, 256), 256, 0, st>>>(G, X, wm_ + off_wp,
                                                             wm_ + off_bp, out_);
    return check(hipGetLastError(), "net forward");
}

int Net::backward(const NGeo& G, hipStream_t st, float inv_n) {
    const long nv = (long)G.I * G.J * G.K * G.L;
    const long nrow = (long)(G.J - 2) * (G.K - 2) * ((G.L / G.P) * (G.P - 2));
    int rpw = 1;
    const int waves = conv_waves(nrow, &rpw);
    const dim3 grid((unsigned)cdivl(waves, 4));
    const Taps Ts = spatial_taps(G, 1), Tt = temporal_taps(G, 1);
    const Taps Ttb = temporal_taps(G, -1);
    const int nbl = (int)cdivl(nv, 256);
    // loss and projection
    net::k_loss_project_bwd<<<nbl, 256, 0, st>>>(G, out_, target_, lmask_, Y_[nlayers_ - 1],
                                                 wm_ + off_wp, 2.f * inv_n, dA_, partial_);
    // partial columns: 0 = sum d^2, 1..16 = sum d a_c, 17 = sum d; the loss
    // is sum d^2 / N, so the projection's gradients are (2/N) times the sums
    k_reduce_cols<<<kCh, 256, 0, st>>>(partial_ + 1, nbl, 18, 0, kCh, 2.f * inv_n, grad_ + off_wp);
    k_reduce_cols<<<1, 256, 0, st>>>(partial_ + 17, nbl, 18, 0, 1, 2.f * inv_n, grad_ + off_bp);
    k_reduce_cols<<<1, 256, 0, st>>>(partial_, nbl, 18, 0, 1, 1.f, (float*)dscal_);
    const int nblk = (int)std::min<long>(2048, nrow);
    const int rpb = (int)cdivl(nrow, nblk);
    const long dw_stride = (long)net::kMaxTaps * 16 * kCh;
    for (int L = nlayers_ - 1; L >= 0; L--) {
        // dA_ holds the gradient at the pre-activation of layer L
        net::k_bias_grad<<<nblk, 256, 0, st>>>(G, dA_, partial_, rpb);
        k_reduce_cols<<<kCh, 256, 0, st>>>(partial_, nblk, kCh, 0, kCh, 1.f, grad_ + off_b[L]);
        net::k_conv_dw<16, kTemporalTaps><<<nblk, 64, 0, st>>>(G, Tt, S_[L], dA_, partial_, rpb);
        k_reduce_cols<<<kTemporalTaps * 256, 256, 0, st>>>(partial_, nblk, dw_stride, 16,
                                                          kTemporalTaps * 256, 1.f,
                                                          grad_ + off_wt[L]);
        net::k_conv<16, kTemporalTaps, net::EPI_NONE><<<grid, 256, 0, st>>>(G, Ttb, dA_, wtr_ + off_wt[L],
                                                             nullptr, nullptr, dB_, rpw);
        if (L == 0) {
            net::k_conv_dw<8, kSpatialTaps><<<nblk, 64, 0, st>>>(G, Ts, feat_, dB_, partial_, rpb);
            k_reduce_cols<<<kSpatialTaps * 8 * kCh, 256, 0, st>>>(
                partial_, nblk, dw_stride, 8, kSpatialTaps * 8 * kCh, 1.f, grad_ + off_ws[L]);
        } else {
            net::k_conv_dw<16, kSpatialTaps><<<nblk, 64, 0, st>>>(G, Ts, Y_[L - 1], dB_, partial_, rpb);
            k_reduce_cols<<<kSpatialTaps * 256, 256, 0, st>>>(
                partial_, nblk, dw_stride, 16, kSpatialTaps * 256, 1.f, grad_ + off_ws[L]);
            net::k_conv_s<16, net::EPI_MASK><<<dim3(spatial_blocks(G)), 256, 0, st>>>(
                G, -1, dB_, wtr_ + off_ws[L], nullptr, Y_[L - 1], dA_);
        }
    }
    return check(hipGetLastError(), "net backward");
}

int Net::record(Engine& E, bool new_run) {
    if (!ready()) {
        err = "net: not initialised";
        return -1;
    }
    const long pts = (long)E.I_ * E.J_ * E.K_ * E.L_;
    for (int f = 0; f < 8; f++)
        if (!E.full(ArrayId(A_G0 + f))) {
            err = "net: the volume has no edge-indicator arrays";
            return -1;
        }
    if (pts >= (1L << 31)) {
        err = "net: volumes of 2^31 or more padded points are not supported";
        return -1;
    }
    if (new_run || runs_.empty()) {
        Run r;
        r.I = E.I_;
        r.J = E.J_;
        r.K = E.K_;
        r.L = E.L_;
        if (check(hipMalloc((void**)&r.gmean, pts * 2), "hipMalloc pool")) return -1;
        const NGeo G{E.I_, E.J_, E.K_, E.L_, E.L_};
        const double* g[8];
        for (int f = 0; f < 8; f++) g[f] = E.full(ArrayId(A_G0 + f));
        k_gmean<<<(unsigned)cdivl(pts, 256), 256, 0, E.st_>>>(G, r.gmean, g[0], g[1], g[2], g[3],
                                                              g[4], g[5], g[6], g[7]);
        runs_.push_back(r);
    }
    Run& r = runs_.back();
    if (r.I != E.I_ || r.J != E.J_ || r.K != E.K_ || r.L != E.L_) {
        err = "net: record with other dimensions than the run";
        return -1;
    }
    double* s = nullptr;
    if (check(hipMalloc((void**)&s, pts * sizeof(double)), "hipMalloc pool") ||
        check(hipMemcpyAsync(s, E.full(A_U), pts * sizeof(double), hipMemcpyDeviceToDevice,
                             E.st_), "pool copy"))
        return -1;
    r.snaps.push_back(s);
    // rms increment and rms third difference of this level over the interior
    // (0 while the run has too few levels)
    double sig = 0.0, kap = 0.0;
    if (r.snaps.size() >= 2) {
        const NGeo G{r.I, r.J, r.K, r.L, r.L};
        long per = 0;
        const int nb = sumsq_blocks(pts, &per);
        double* part = nullptr;
        if (check(hipMalloc((void**)&part, (size_t)nb * 2 * sizeof(double)), "hipMalloc"))
            return -1;
        const size_t n = r.snaps.size() - 1;  // this level
        const double* h2 = n >= 3 ? r.snaps[n - 2] : nullptr;
        const double* h3 = n >= 3 ? r.snaps[n - 3] : nullptr;
        k_pred_moments<<<nb, 256, 0, E.st_>>>(G, s, r.snaps[n - 1], h2, h3, per, part);
        k_reduce_d<<<1, 256, 0, E.st_>>>(part, nb, 2, dscal_);
        const int rc = check(hipMemcpyAsync(hscal_, dscal_, 2 * sizeof(double),
                                            hipMemcpyDeviceToHost, E.st_), "record") ||
                       check(hipStreamSynchronize(E.st_), "record");
        (void)hipFree(part);
        if (rc) return -1;
        const double nint = (double)(r.I - 2) * (r.J - 2) * (r.K - 2) * (r.L - 2);
        sig = std::sqrt(hscal_[0] / nint);
        kap = std::sqrt(hscal_[1] / nint);
    }
    r.sigma.push_back(sig);
    r.kappa.push_back(kap);
    return check(hipStreamSynchronize(E.st_), "net record");
}

long Net::pool_pairs() const {
    // levels n with the base's history n-1..n-base and a successor n+1
    long n = 0;
    for (const Run& r : runs_) {
        const long m = (long)r.snaps.size() - 1 - base_;
        n += m > 0 ? m : 0;
    }
    return n;
}

int Net::compute_scales(hipStream_t st) {
    double sd = 0, se = 0, sc = 0;
    long npts = 0;
    for (const Run& r : runs_) {
        const NGeo G{r.I, r.J, r.K, r.L, r.L};
        const long pts = (long)r.I * r.J * r.K * r.L;
        const int nb = (int)cdivl(pts, 256);
        double* part = nullptr;
        if (check(hipMalloc((void**)&part, (size_t)nb * 3 * sizeof(double)), "hipMalloc")) return -1;
        for (size_t n = (size_t)base_; n + 1 < r.snaps.size(); n++) {
            const double* um3 = base_ == 3 ? r.snaps[n - 3] : nullptr;
            const double* um2 = base_ == 3 ? r.snaps[n - 2] : nullptr;
            k_moments<<<nb, 256, 0, st>>>(G, base_, um3, um2,
                                          base_ >= 1 ? r.snaps[n - 1] : nullptr, r.snaps[n],
                                          r.snaps[n + 1], part);
            k_reduce_d<<<1, 256, 0, st>>>(part, nb, 3, dscal_);
            if (check(hipMemcpyAsync(hscal_, dscal_, 3 * sizeof(double), hipMemcpyDeviceToHost, st),
                      "moments") ||
                check(hipStreamSynchronize(st), "moments")) {
                (void)hipFree(part);
                return -1;
            }
            // features by the level's rms increment / third difference, targets
            // by the latter (cubic base) or the former (linear base)
            const double s2 = r.sigma[n] > 0 ? r.sigma[n] * r.sigma[n] : 1.0;
            const double k2 = r.kappa[n] > 0 ? r.kappa[n] * r.kappa[n] : 1.0;
            sd += hscal_[0] / s2;
            se += hscal_[1] / k2;
            sc += hscal_[2] / (base_ == 3 ? k2 : base_ == 1 ? s2 : 1.0);
            npts += (long)(r.I - 2) * (r.J - 2) * (r.K - 2) * (r.L - 2);
        }
        (void)hipFree(part);
    }
    if (npts == 0 || (base_ >= 1 && !(sd > 0)) || !(sc > 0)) {
        err = "net: training pool needs runs with motion and enough time levels for the base";
        return -1;
    }
    s_d = sd > 0 ? (float)(1.0 / std::sqrt(sd / npts)) : 1.f;
    s_e = se > 0 ? (float)(1.0 / std::sqrt(se / npts)) : 1.f;
    s_t = (float)(1.0 / std::sqrt(sc / npts));
    return 0;
}

int Net::train(int iters, int batch, float lr, uint64_t seed, int log_every, double* loss) {
    if (!ready()) {
        err = "net: not initialised";
        return -1;
    }
    if (runs_.empty() || pool_pairs() == 0) {
        err = "net: empty training pool";
        return -1;
    }
    const int P = runs_[0].L;
    for (const Run& r : runs_)
        if (r.L != P || r.I < kPatch || r.J < kPatch || r.K < kPatch) {
            err = "net: runs need equal frame counts and at least 74 padded points per axis";
            return -1;
        }
    if (batch < 1 || iters < 1) {
        err = "net: batch and iterations must be positive";
        return -1;
    }
    hipStream_t st = nullptr;
    if (check(hipStreamCreate(&st), "stream")) return -1;
    int rc = 0;
    if (adam_t_ == 0) rc = compute_scales(st);
    const NGeo G{kPatch, kPatch, kPatch, batch * P, P};
    if (!rc) rc = ensure_acts(G.I, G.J, G.K, G.L, true);
    // (run, level) pairs with the base's history and a successor
    std::vector<std::pair<int, int>> pairs;
    for (int q = 0; q < (int)runs_.size(); q++)
        for (int n = base_; n + 1 < (int)runs_[q].snaps.size(); n++) pairs.push_back({q, n});
    uint64_t s = seed;
    double tail = 0;
    int ntail = 0;
    const int tail_from = iters - std::max(1, iters / 10);
    const long per = (long)G.I * G.J * G.K * P;
    for (int it = 0; it < iters && !rc; it++) {
        long nloss = 0;
        for (int b = 0; b < batch; b++) {
            const auto pr = pairs[(size_t)(uniform01(s) * pairs.size())];
            const Run& r = runs_[pr.first];
            const int n = pr.second;
            const int ax = (int)(uniform01(s) * (r.I - kPatch + 1));
            const int ay = (int)(uniform01(s) * (r.J - kPatch + 1));
            const int az = (int)(uniform01(s) * (r.K - kPatch + 1));
            const double* um3 = base_ == 3 ? r.snaps[n - 3] : nullptr;
            const double* um2 = base_ == 3 ? r.snaps[n - 2] : nullptr;
            k_gather_patch<<<(unsigned)cdivl(per, 256), 256, 0, st>>>(
                G, b, r.I, r.J, r.K, base_, um3, um2, base_ >= 1 ? r.snaps[n - 1] : nullptr,
                r.snaps[n], r.snaps[n + 1],
                r.gmean, ax, ay, az, s_d, s_e, s_t, r.sigma[n] > 0 ? 1.0 / r.sigma[n] : 0.0,
                r.kappa[n] > 0 ? 1.0 / r.kappa[n] : 0.0, feat_, target_, lmask_);
            auto cnt = [&](int a, int rn) {
                const int lo = a == 0 ? 1 : kHalo + 1;
                const int hi = a + kPatch == rn ? kPatch - 2 : kPatch - kHalo - 2;
                return (long)(hi - lo + 1);
            };
            nloss += cnt(ax, r.I) * cnt(ay, r.J) * cnt(az, r.K) * (P - 2);
        }
        const float inv_n = 1.f / (float)nloss;
        rc = forward(G, st, true);
        if (!rc) rc = backward(G, st, inv_n);
        if (rc) break;
        adam_t_++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)adam_t_);
        const float c2 = 1.f - std::pow(b2, (float)adam_t_);
        net::k_adam<<<(unsigned)cdivl(nparam, 256), 256, 0, st>>>(wm_, m1_, m2_, grad_, nparam, lr,
                                                                  b1, b2, 1e-8f, c1, c2, wb_);
        for (int L = 0; L < nlayers_; L++) {
            if (L > 0)
                k_transpose_taps<<<(kSpatialTaps * 256 + 255) / 256, 256, 0, st>>>(
                    wb_ + off_ws[L], kSpatialTaps, wtr_ + off_ws[L]);
            k_transpose_taps<<<(kTemporalTaps * 256 + 255) / 256, 256, 0, st>>>(
                wb_ + off_wt[L], kTemporalTaps, wtr_ + off_wt[L]);
        }
        rc = check(hipMemcpyAsync(hscal_, dscal_, sizeof(float), hipMemcpyDeviceToHost, st), "loss");
        if (!rc) rc = check(hipStreamSynchronize(st), "train step");
        if (rc) break;
        const double lval = (double)(*(float*)hscal_) * inv_n;
        if (it >= tail_from) {
            tail += lval;
            ntail++;
        }
        if (log_every > 0 && (it % log_every == 0 || it == iters - 1)) {
            std::fprintf(stderr, "[net] step %d loss %.5f\n", it, lval);
        }
    }
    if (!rc) rc = check(hipStreamSynchronize(st), "train");
    (void)hipStreamDestroy(st);
    if (loss) *loss = ntail ? tail / ntail : 0.0;
    return rc;
}

int Net::forward_batch(int I, int J, int K, int L, int P, const void* feat_dev, float* out_dev) {
    if (!ready()) {
        err = "net: not initialised";
        return -1;
    }
    const NGeo G{I, J, K, L, P};
    if (ensure_acts(I, J, K, L, false)) return -1;
    const size_t pts = (size_t)I * J * K * L;
    if (check(hipMemcpy(feat_, feat_dev, pts * 16, hipMemcpyDeviceToDevice), "feat copy") ||
        forward(G, nullptr, false) ||
        check(hipMemcpy(out_dev, out_, pts * 4, hipMemcpyDeviceToDevice), "out copy"))
        return -1;
    return check(hipDeviceSynchronize(), "forward_batch");
}

int Net::grad_batch(int I, int J, int K, int L, int P, const void* feat_dev,
                    const float* target_dev, const unsigned char* lmask_dev, double* loss) {
    if (!ready()) {
        err = "net: not initialised";
        return -1;
    }
    const NGeo G{I, J, K, L, P};
    if (ensure_acts(I, J, K, L, true)) return -1;
    const size_t pts = (size_t)I * J * K * L;
    std::vector<unsigned char> lm(pts);
    if (check(hipMemcpy(lm.data(), lmask_dev, pts, hipMemcpyDeviceToHost), "lmask copy"))
        return -1;
    long n = 0;
    for (unsigned char c : lm) n += c ? 1 : 0;
    if (n == 0) {
        err = "net: no loss points";
        return -1;
    }
    const float inv_n = 1.f / (float)n;
    if (check(hipMemcpy(feat_, feat_dev, pts * 16, hipMemcpyDeviceToDevice), "feat copy") ||
        check(hipMemcpy(target_, target_dev, pts * 4, hipMemcpyDeviceToDevice), "target copy") ||
        check(hipMemcpy(lmask_, lmask_dev, pts, hipMemcpyDeviceToDevice), "lmask copy") ||
        check(hipMemset(grad_, 0, (size_t)nparam * 4), "grad clear") ||
        forward(G, nullptr, true) || backward(G, nullptr, inv_n) ||
        check(hipDeviceSynchronize(), "grad_batch"))
        return -1;
    float lsum = 0.f;
    if (check(hipMemcpy(&lsum, dscal_, 4, hipMemcpyDeviceToHost), "loss")) return -1;
    if (loss) *loss = (double)lsum * inv_n;
    return 0;
}

int Net::predict(Engine& E, const double* h1, const double* h2, const double* h3,
                 LearnedStart* ls) {
    if (!ready()) {
        err = "net: not initialised";
        return -1;
    }
    if ((base_ >= 1 && !h1) || (base_ == 3 && (!h2 || !h3))) {
        err = "net: the prediction needs the base's history levels";
        return -1;
    }
    const NGeo G{E.I_, E.J_, E.K_, E.L_, E.L_};
    const long pts = (long)E.I_ * E.J_ * E.K_ * E.L_;
    for (int f = 0; f < 8; f++)
        if (!E.full(ArrayId(A_G0 + f))) {
            err = "net: the volume has no edge-indicator arrays";
            return -1;
        }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, E.st_);
    if (ensure_acts(E.I_, E.J_, E.K_, E.L_, false)) return -1;
    if (g_version_ != E.g_version_ || gpts_ != pts) {
        if (gmean_) (void)hipFree(gmean_);
        gmean_ = nullptr;
        if (check(hipMalloc((void**)&gmean_, pts * 2), "hipMalloc gmean")) return -1;
        gpts_ = pts;
        const double* g[8];
        for (int f = 0; f < 8; f++) g[f] = E.full(ArrayId(A_G0 + f));
        k_gmean<<<(unsigned)cdivl(pts, 256), 256, 0, E.st_>>>(G, gmean_, g[0], g[1], g[2], g[3],
                                                              g[4], g[5], g[6], g[7]);
        g_version_ = E.g_version_;
    }
    if (pts >= (1L << 31)) {
        err = "net: volumes of 2^31 or more padded points are not supported";
        return -1;
    }
    const double* u = E.full(A_U);
    const double inv_n = 1.0 / ((double)(E.I_ - 2) * (E.J_ - 2) * (E.K_ - 2) * (E.L_ - 2));
    // one pass: the features and the block sums of the increment and of the
    // third difference squared, reduced in two fixed-order levels into
    // dscal_[8], dscal_[9] (left on the device for the set-up kernel); the
    // first layer's weights take the feature scales s_d / sigma, s_e / kappa
    const unsigned grid = (unsigned)cdivl(pts, 256);
    const long chunk = 1024;
    const long nc = cdivl((long)grid, chunk);
    if ((long)grid * 2 + nc * 2 > npartial_ / 2) {
        err = "net: prediction partials";
        return -1;
    }
    double* part = (double*)partial_;
    k_features<<<grid, 256, 0, E.st_>>>(G, base_, u, h1, base_ == 3 ? h2 : nullptr,
                                        base_ == 3 ? h3 : nullptr, gmean_, feat_, part);
    k_chunk_sums2<<<(unsigned)nc, 256, 0, E.st_>>>(part, (long)grid, chunk, part + 2 * grid);
    k_reduce_d<<<1, 256, 0, E.st_>>>(part + 2 * grid, (int)nc, 2, dscal_ + 8);
    const double* dsum = dscal_ + 8;
    const int nw0 = net::kMaxTaps * 4 * kCh;
    k_scale_w0<<<(unsigned)cdivl(nw0, 256), 256, 0, E.st_>>>(wm_ + off_ws[0], nw0, dsum, inv_n,
                                                              s_d, s_e, w0s_);
    if (forward(G, E.st_, false, w0s_)) return -1;
    ls->out = out_;
    // the target's scale: kappa (order 3), sigma (order 1), or 1 (order 0:
    // dscal_[10] holds 1.0, set at init)
    ls->dsum = base_ == 3 ? dsum + 1 : base_ == 1 ? dsum : dscal_ + 10;
    ls->inv_n = base_ == 0 ? 1.0 : inv_n;
    ls->inv_st = 1.f / s_t;
    ls->order = base_;
    (void)hipEventRecord(e1, E.st_);
    if (check(hipEventSynchronize(e1), "net predict")) return -1;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    predict_ms += ms;
    predicts++;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return check(hipGetLastError(), "net predict");
}

int Net::save(const char* path) {
    if (!ready()) {
        err = "net: not initialised";
        return -1;
    }
    std::vector<float> w((size_t)nparam);
    if (check(hipMemcpy(w.data(), wm_, w.size() * 4, hipMemcpyDeviceToHost), "net save")) return -1;
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("net: cannot create ") + path;
        return -1;
    }
    // S4DNET02: magic, parameter count, layers, base order, s_d, s_t, s_e,
    // Adam step, FP32 parameters (little-endian)
    const char magic[8] = {'S', '4', 'D', 'N', 'E', 'T', '0', '2'};
    const int32_t np = nparam, nl = nlayers_, bo = base_;
    const int64_t t = adam_t_;
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&np, 4, 1, f) == 1 &&
              std::fwrite(&nl, 4, 1, f) == 1 && std::fwrite(&bo, 4, 1, f) == 1 &&
              std::fwrite(&s_d, 4, 1, f) == 1 && std::fwrite(&s_t, 4, 1, f) == 1 &&
              std::fwrite(&s_e, 4, 1, f) == 1 && std::fwrite(&t, 8, 1, f) == 1 &&
              std::fwrite(w.data(), 4, w.size(), f) == w.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = "net: write failed";
    return ok ? 0 : -1;
}

int Net::load(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("net: cannot open ") + path;
        return -1;
    }
    // S4DNET01 (four layers, linear base) or S4DNET02 (shape in the header)
    char magic[8];
    int32_t np = 0, nl = 4, bo = 1;
    float sd = 1, stv = 1, se = 1;
    int64_t t = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "S4DNET0", 7) == 0 &&
              (magic[7] == '1' || magic[7] == '2') && std::fread(&np, 4, 1, f) == 1;
    const bool v2 = ok && magic[7] == '2';
    if (ok && v2) ok = std::fread(&nl, 4, 1, f) == 1 && std::fread(&bo, 4, 1, f) == 1;
    ok = ok && std::fread(&sd, 4, 1, f) == 1 && std::fread(&stv, 4, 1, f) == 1;
    if (ok && v2) ok = std::fread(&se, 4, 1, f) == 1;
    ok = ok && std::fread(&t, 8, 1, f) == 1;
    if (!ok) {
        std::fclose(f);
        err = "net: not a weights file";
        return -1;
    }
    if (init(0, nl, bo)) {
        std::fclose(f);
        return -1;
    }
    if (np != nparam) {
        std::fclose(f);
        err = "net: weights file of another network shape";
        return -1;
    }
    std::vector<float> w((size_t)nparam);
    ok = std::fread(w.data(), 4, w.size(), f) == w.size();
    std::fclose(f);
    if (!ok) {
        err = "net: truncated weights file";
        return -1;
    }
    if (check(hipMemcpy(wm_, w.data(), w.size() * 4, hipMemcpyHostToDevice), "net load")) return -1;
    s_d = sd;
    s_t = stv;
    s_e = se;
    adam_t_ = t;
    return refresh();
}

int Net::refresh() {
    k_f32_to_bf16<<<(unsigned)cdivl(nparam, 256), 256>>>(wm_, nparam, wb_);
    for (int L = 0; L < nlayers_; L++) {
        if (L > 0)
            k_transpose_taps<<<(kSpatialTaps * 256 + 255) / 256, 256>>>(wb_ + off_ws[L],
                                                                     kSpatialTaps,
                                                                     wtr_ + off_ws[L]);
        k_transpose_taps<<<(kTemporalTaps * 256 + 255) / 256, 256>>>(wb_ + off_wt[L],
                                                                  kTemporalTaps,
                                                                  wtr_ + off_wt[L]);
    }
    return check(hipDeviceSynchronize(), "net refresh");
}

}  // namespace s4d
