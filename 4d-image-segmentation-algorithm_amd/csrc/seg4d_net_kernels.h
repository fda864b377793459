// seg4d_net_kernels.h -- HIP kernels of the learned warm start: a separable
// 4D convolutional network (3x3x3 spatial conv, then 3-tap temporal conv,
// bias, ReLU; one to four such layers of 16 channels and a 1x1 projection)
// that predicts the correction of an extrapolated SOR start from u^n, its
// recent increments and g.
//
// Activations live on the solver's padded grid (I x J x K x L points, x
// fastest), channel-last in BF16: element (v, c) at v * C + c with C = 8 for
// the input features and 16 otherwise.  Ghost points hold zeros and are
// never written, so they are the convolutions' zero padding.  A training
// batch stacks patches along the frame axis: frames with l % P in {0, P-1}
// are ghosts (P = L for one volume).
//
// The convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16 (gfx950):
// one wave computes a tile of 16 consecutive x points by 16 output channels
// with FP32 accumulation.  For that instruction, lane ln holds
//   A[m = ln % 16][k = 8 (ln / 16) .. + 7]   (8 consecutive k, BF16)
//   B[k = 8 (ln / 16) .. + 7][n = ln % 16]
//   C[m = 4 (ln / 16) + r][n = ln % 16], r = 0..3  (FP32)
// With channel-last data, the 8 k of an A operand are 8 channels of one
// point (16 contiguous bytes).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace s4d {
namespace net {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCh = 16;       // hidden channels
constexpr int kCin0 = 8;      // input feature channels (3 used, 4 with the cubic base)
constexpr int kMaxTaps = 28;  // 27 spatial taps padded to an even count
constexpr int kWave = 64;

struct NGeo {
    int I, J, K, L;  // padded dims of the (batched) grid
    int P;           // frame period: l % P in {0, P-1} are ghost frames
};

// Taps of one convolution as linear point offsets (dx + I (dy + J (dz +
// K dl))), built on the host for the grid at hand.  For the weight gradient
// the taps are also grouped by row: `rowoff` lists the distinct (dy, dz, dl)
// rows (dx = 0) and tap t reads row trow[t] shifted by tdx[t].
struct Taps {
    int n;
    int off[kMaxTaps];
    int nrow;
    int rowoff[9];
    signed char trow[kMaxTaps];
    signed char tdx[kMaxTaps];
};

__device__ __forceinline__ long nvox(const NGeo& G, int i, int j, int k, int l) {
    return ((long)(l * G.K + k) * G.J + j) * G.I + i;
}

__device__ __forceinline__ bool interior(const NGeo& G, int i, int j, int k, int l) {
    const int lp = l % G.P;
    return i >= 1 && i <= G.I - 2 && j >= 1 && j <= G.J - 2 && k >= 1 && k <= G.K - 2 &&
           lp >= 1 && lp <= G.P - 2;
}

// Row r of the interior rows (j, k, l) of the grid, l counting interior
// frames of every period.
__device__ __forceinline__ void row_of(const NGeo& G, long r, int& j, int& k, int& l) {
    const int nj = G.J - 2, nk = G.K - 2, nlp = G.P - 2;
    j = (int)(r % nj) + 1;
    r /= nj;
    k = (int)(r % nk) + 1;
    r /= nk;
    const int lq = (int)r;
    l = (lq / nlp) * G.P + (lq % nlp) + 1;
}

__device__ __forceinline__ long n_rows(const NGeo& G) {
    return (long)(G.J - 2) * (G.K - 2) * ((G.L / G.P) * (G.P - 2));
}

// Epilogue of a convolution: 0 = store, 1 = + bias and ReLU,
// 2 = multiply by (mask > 0) (ReLU derivative in the backward pass).
enum { EPI_NONE = 0, EPI_BIAS_RELU = 1, EPI_MASK = 2 };

// ----------------------------------------------------------------------------
// Convolution Y = sum_tap W[tap] X[v + d(tap)], output tile 16 x-points by 16
// channels per wave.  W is BF16 [tap][ci][co] (CIN x 16 per tap, taps padded
// with zeros to a multiple of 32 / CIN).  Each wave walks `rows_per_wave`
// rows; the weights stay in registers.
// ----------------------------------------------------------------------------
template <int CIN, int NT, int EPI>
__global__ __launch_bounds__(256) void k_conv(NGeo G, Taps T, const bf16* __restrict__ X,
                                               const bf16* __restrict__ W,
                                               const float* __restrict__ bias,
                                               const bf16* __restrict__ mask,
                                               bf16* __restrict__ Y, int rows_per_wave) {
    constexpr int TPS = 32 / CIN;               // taps per MFMA K-step
    constexpr int NS = (NT + TPS - 1) / TPS;     // K-steps
    const int ln = threadIdx.x & 63;
    const int g = ln >> 4, m = ln & 15;
    // B operands: lane holds W[tap][ci][co = m] for k = 8 g .. 8 g + 7
    bf16x8 wreg[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        bf16x8 b;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int kk = s * 32 + 8 * g + q;
            const int tap = kk / CIN, ci = kk % CIN;
            b[q] = tap < NT ? W[(tap * CIN + ci) * kCh + m] : (bf16)0.0f;
        }
        wreg[s] = b;
    }
    const int tap_in_step = (8 * g) / CIN, ci0 = (8 * g) % CIN;
    // this lane's tap offsets, one per K-step (uniform loads, lane selects)
    int loff[NS];
    bool lok[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int o = 0;
#pragma unroll
        for (int q = 0; q < TPS; q++)
            if (tap_in_step == q && s * TPS + q < NT) o = T.off[s * TPS + q];
        loff[s] = o;
        lok[s] = s * TPS + tap_in_step < NT;
    }
    const float bco = (EPI == EPI_BIAS_RELU) ? bias[m] : 0.0f;
    const long nrow = n_rows(G);
    const int tiles = (G.I - 2 + 15) / 16;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    for (int rr = 0; rr < rows_per_wave; rr++) {
        const long r = wave * rows_per_wave + rr;
        if (r >= nrow) break;
        int j, k, l;
        row_of(G, r, j, k, l);
        for (int tl = 0; tl < tiles; tl++) {
            const int i = 1 + tl * 16 + m;  // this lane's A row (x point)
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < NS; s++) {
                bf16x8 a;
                const bool ok = lok[s] && i <= G.I - 2;
                if (ok) {
                    const long src = (nvox(G, i, j, k, l) + loff[s]) * CIN + ci0;
                    a = *reinterpret_cast<const bf16x8*>(X + src);
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++) a[q] = (bf16)0.0f;
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wreg[s], acc, 0, 0, 0);
            }
            // C[x = 4 g + rr2][co = m]
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int io = 1 + tl * 16 + 4 * g + q;
                if (io > G.I - 2) continue;
                const long dst = nvox(G, io, j, k, l) * kCh + m;
                float v = acc[q];
                if (EPI == EPI_BIAS_RELU) v = fmaxf(v + bco, 0.0f);
                if (EPI == EPI_MASK) v = ((float)mask[dst] > 0.0f) ? v : 0.0f;
                Y[dst] = (bf16)v;
            }
        }
    }
}

// ----------------------------------------------------------------------------
// 3x3x3 spatial convolution with the input patch staged in LDS.  A block of 4
// waves computes a 16-point x-tile of 4 x 4 rows (j, k) of one frame: it loads
// the (16 + 2) x (4 + 2) x (4 + 2) points around them (zero outside the grid)
// into LDS once, so each input point is fetched from L2 once per block
// instead of once per tap; wave w then computes rows (j0 + jj, k0 + w), jj =
// 0..3, with four independent accumulators.  The MFMA computes C = W^T X^T:
// A = weights (lane: out channel m, 8 consecutive k), B = inputs (8 channels
// of one point), so a lane's four results are four consecutive channels of
// one point, stored as one 8-byte write.  sign = +1: taps d(t) =
// (t % 3 - 1, t / 3 % 3 - 1, t / 9 - 1) in (x, y, z); -1: taps -d(t) (the
// backward-data pass, with per-tap transposed weights).
// ----------------------------------------------------------------------------
constexpr int kSpX = 18, kSpY = 6, kSpZ = 6;  // staged patch (points)

template <int CIN, int EPI>
__global__ __launch_bounds__(256) void k_conv_s(NGeo G, int sign, const bf16* __restrict__ X,
                                                 const bf16* __restrict__ W,
                                                 const float* __restrict__ bias,
                                                 const bf16* __restrict__ mask,
                                                 bf16* __restrict__ Y) {
    constexpr int NT = 27;
    constexpr int TPS = 32 / CIN;                // taps per MFMA K-step
    constexpr int NS = (NT + TPS - 1) / TPS;     // K-steps
    constexpr int C8 = CIN / 8;                  // 16-byte chunks per point
    constexpr int NP = kSpZ * kSpY * kSpX;
    __shared__ bf16x8 sp[NP * C8];
    const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = ln >> 4, m = ln & 15;
    // block -> (x tile, j group, k group, interior frame)
    const int ntx = (G.I - 2 + 15) / 16, nj = (G.J - 2 + 3) / 4, nk = (G.K - 2 + 3) / 4;
    long b = blockIdx.x;
    const int tx = (int)(b % ntx);
    b /= ntx;
    const int tj = (int)(b % nj);
    b /= nj;
    const int tk = (int)(b % nk);
    const int lq = (int)(b / nk);
    const int nlp = G.P - 2;
    const int l = (lq / nlp) * G.P + (lq % nlp) + 1;
    const int i0 = 1 + 16 * tx, j0 = 1 + 4 * tj, k0 = 1 + 4 * tk;
    for (int e = threadIdx.x; e < NP * C8; e += 256) {
        const int c = e % C8;
        int r = e / C8;
        const int x = r % kSpX;
        r /= kSpX;
        const int y = r % kSpY, z = r / kSpY;
        const int i = i0 - 1 + x, j = j0 - 1 + y, k = k0 - 1 + z;
        bf16x8 v;
        if (i <= G.I - 1 && j <= G.J - 1 && k <= G.K - 1) {
            v = *reinterpret_cast<const bf16x8*>(X + nvox(G, i, j, k, l) * CIN + 8 * c);
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = (bf16)0.0f;
        }
        sp[e] = v;
    }
    // A operands: lane holds W[tap][ci][co = m] for k = 8 g .. 8 g + 7
    bf16x8 wreg[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        bf16x8 a;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int kk = s * 32 + 8 * g + q;
            const int tap = kk / CIN, ci = kk % CIN;
            a[q] = tap < NT ? W[(tap * CIN + ci) * kCh + m] : (bf16)0.0f;
        }
        wreg[s] = a;
    }
    // this lane's tap per K-step as an LDS point offset, and its chunk
    const int tap_in_step = (8 * g) / CIN, chunk = ((8 * g) % CIN) / 8;
    int toff[NS];
    bool tok[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int t = s * TPS + tap_in_step;
        const int dx = sign * (t % 3 - 1), dy = sign * (t / 3 % 3 - 1), dz = sign * (t / 9 - 1);
        tok[s] = t < NT;
        toff[s] = (dz * kSpY + dy) * kSpX + dx;
    }
    __syncthreads();
    f32x4 acc[4];
#pragma unroll
    for (int jj = 0; jj < 4; jj++) acc[jj] = f32x4{0.f, 0.f, 0.f, 0.f};
    // patch position of this lane's output point in row jj: (m+1, jj+1, wv+1)
    const int base = ((wv + 1) * kSpY + 1) * kSpX + (m + 1);
#pragma unroll
    for (int s = 0; s < NS; s++) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            bf16x8 bv;
            if (tok[s]) {
                bv = sp[(base + jj * kSpX + toff[s]) * C8 + chunk];
            } else {
#pragma unroll
                for (int q = 0; q < 8; q++) bv[q] = (bf16)0.0f;
            }
            acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[s], bv, acc[jj], 0, 0, 0);
        }
    }
    const int i = i0 + m, k = k0 + wv;
    if (i > G.I - 2 || k > G.K - 2) return;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int j = j0 + jj;
        if (j > G.J - 2) continue;
        const long dst = nvox(G, i, j, k, l) * kCh + 4 * g;
        bf16 o[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float v = acc[jj][r];
            if (EPI == EPI_BIAS_RELU) v = fmaxf(v + bias[4 * g + r], 0.0f);
            if (EPI == EPI_MASK) v = ((float)mask[dst + r] > 0.0f) ? v : 0.0f;
            o[r] = (bf16)v;
        }
        *reinterpret_cast<uint2*>(Y + dst) = *reinterpret_cast<const uint2*>(o);
    }
}

// ----------------------------------------------------------------------------
// One inference layer fused: S(t) = spatial conv of X(t) (k_conv_s's tiling
// and staging), then Y(t) = ReLU(bias + sum_dt Wt[dt] S(t + dt)), marching t
// through each frame period.  A block keeps S of the frames t-2, t-1, t for
// its 256 points in an LDS ring (ghost frames zero), so S never goes to
// memory; the next frame's input patch is loaded into registers while the
// current one is computed.  Each wave reads and writes only its own part of
// the ring; the patch is shared by the 4 waves.
// ----------------------------------------------------------------------------
// PROJ (the network's last layer): instead of Y, the 1x1 projection
// out[v] = sum_c wp[c] y[v][c] + bp[0] (FP32) of the layer output rounded to
// BF16 as Y would hold it; the 16 channels of a point are spread over the 4
// lanes m, m + 16, m + 32, m + 48 and summed with two lane exchanges.
template <int CIN, bool PROJ>
__global__ __launch_bounds__(256) void k_layer_fwd(NGeo G, const bf16* __restrict__ X,
                                                    const bf16* __restrict__ Ws,
                                                    const bf16* __restrict__ Wt,
                                                    const float* __restrict__ bias,
                                                    bf16* __restrict__ Y,
                                                    const float* __restrict__ wp,
                                                    const float* __restrict__ bp,
                                                    float* __restrict__ out) {
    // CIN = 4 (the prediction's unscaled features): a point is one 8-byte
    // bf16x4; a lane's 8 K values are 2 taps x 4 channels, two 8-byte reads
    static_assert(CIN == 4 || CIN % 8 == 0, "input channels: 4 or a multiple of 8");
    constexpr bool C4 = CIN == 4;
    constexpr int NT = 27;
    constexpr int TPS = 32 / CIN;
    constexpr int NS = (NT + TPS - 1) / TPS;
    constexpr int C8 = C4 ? 1 : CIN / 8;
    constexpr int NP = kSpZ * kSpY * kSpX;
    constexpr int NE = C4 ? NP : NP * C8;        // staged elements (C4: 8-byte points)
    constexpr int PER = (NE + 255) / 256;        // patch elements per thread
    __shared__ bf16x8 sp[C4 ? (NP + 1) / 2 : NE];
    bf16x4* sp4 = reinterpret_cast<bf16x4*>(sp);
    __shared__ bf16x8 ring[3][4][4][16][2];      // [t % 3][wave][row tile][point][8 channels]
    const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = ln >> 4, m = ln & 15;
    const int ntx = (G.I - 2 + 15) / 16, nj = (G.J - 2 + 3) / 4, nk = (G.K - 2 + 3) / 4;
    long b = blockIdx.x;
    const int tx = (int)(b % ntx);
    b /= ntx;
    const int tj = (int)(b % nj);
    b /= nj;
    const int tk = (int)(b % nk);
    const int period = (int)(b / nk);
    const int lbase = period * G.P;  // ghost frame 0 of this period
    const int i0 = 1 + 16 * tx, j0 = 1 + 4 * tj, k0 = 1 + 4 * tk;
    // spatial and temporal A operands, bias of this lane's 4 channels
    bf16x8 ws[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        bf16x8 a;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int kk = s * 32 + 8 * g + q;
            const int tap = kk / CIN, ci = kk % CIN;
            a[q] = tap < NT ? Ws[(tap * CIN + ci) * kCh + m] : (bf16)0.0f;
        }
        ws[s] = a;
    }
    bf16x8 wt[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        bf16x8 a;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int kk = s * 32 + 8 * g + q;
            const int tap = kk / kCh, ci = kk % kCh;
            a[q] = tap < 3 ? Wt[(tap * kCh + ci) * kCh + m] : (bf16)0.0f;
        }
        wt[s] = a;
    }
    float bco[4], wpr[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        bco[r] = bias[4 * g + r];
        wpr[r] = PROJ ? wp[4 * g + r] : 0.f;
    }
    const float bpo = PROJ ? bp[0] : 0.f;
    const int tap_in_step = (8 * g) / CIN, chunk = C4 ? 0 : ((8 * g) % CIN) / 8;
    auto tap_off = [](int t) { return ((t / 9 - 1) * kSpY + (t / 3 % 3 - 1)) * kSpX + (t % 3 - 1); };
    int toff[NS], toff2[NS];
    bool tok[NS], tok2[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int t = s * TPS + tap_in_step;
        toff[s] = tap_off(t);
        tok[s] = t < NT;
        toff2[s] = tap_off(t + 1);  // C4: the lane's second tap
        tok2[s] = C4 && t + 1 < NT;
    }
    const int base = ((wv + 1) * kSpY + 1) * kSpX + (m + 1);
    bf16x8 zero8;
#pragma unroll
    for (int q = 0; q < 8; q++) zero8[q] = (bf16)0.0f;
    // S of the ghost frame 0
    if (g < 2) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) ring[0][wv][jj][m][g] = zero8;
    }
    bf16x4 zero4;
#pragma unroll
    for (int q = 0; q < 4; q++) zero4[q] = (bf16)0.0f;
    auto fetch = [&](int l, bf16x8* pre, bf16x4* pre4) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + 256 * q;
            bf16x8 v = zero8;
            bf16x4 v4 = zero4;
            if (e < NE) {
                const int c = e % C8;
                int r = e / C8;
                const int x = r % kSpX;
                r /= kSpX;
                const int y = r % kSpY, z = r / kSpY;
                const int i = i0 - 1 + x, j = j0 - 1 + y, k = k0 - 1 + z;
                if (i <= G.I - 1 && j <= G.J - 1 && k <= G.K - 1) {
                    if constexpr (C4)
                        v4 = *reinterpret_cast<const bf16x4*>(X + nvox(G, i, j, k, lbase + l) * 4);
                    else
                        v = *reinterpret_cast<const bf16x8*>(X + nvox(G, i, j, k, lbase + l) * CIN +
                                                             8 * c);
                }
            }
            if constexpr (C4) pre4[q] = v4;
            else pre[q] = v;
        }
    };
    bf16x8 pre[C4 ? 1 : PER];
    bf16x4 pre4[C4 ? PER : 1];
    fetch(1, pre, pre4);
    const int i = i0 + m, k = k0 + wv;
    const bool outok = i <= G.I - 2 && k <= G.K - 2;
    for (int l = 1; l <= G.P - 1; l++) {
        const bool real = l <= G.P - 2;  // frame P-1 is a ghost: S = 0
        // (every wave finished reading the previous patch before the last barrier)
        if (real) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int e = threadIdx.x + 256 * q;
                if (e < NE) {
                    if constexpr (C4) sp4[e] = pre4[q];
                    else sp[e] = pre[q];
                }
            }
        }
        if (l + 1 <= G.P - 2) fetch(l + 1, pre, pre4);
        __syncthreads();
        f32x4 acc[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) acc[jj] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (real) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    bf16x8 bv;
                    if constexpr (C4) {
                        const int pt = base + jj * kSpX;
                        const bf16x4 lo = tok[s] ? sp4[pt + toff[s]] : zero4;
                        const bf16x4 hi = tok2[s] ? sp4[pt + toff2[s]] : zero4;
                        bv = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
                    } else {
                        bv = tok[s] ? sp[(base + jj * kSpX + toff[s]) * C8 + chunk] : zero8;
                    }
                    acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ws[s], bv, acc[jj], 0, 0, 0);
                }
            }
        }
        // S(l) as BF16 into the ring: this lane's channels 4g..4g+3 of point m
        const int sl = l % 3;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            bf16 o[4];
#pragma unroll
            for (int r = 0; r < 4; r++) o[r] = (bf16)acc[jj][r];
            *reinterpret_cast<uint2*>(reinterpret_cast<bf16*>(&ring[sl][wv][jj][m][0]) + 4 * g) =
                *reinterpret_cast<const uint2*>(o);
        }
        __syncthreads();
        // temporal conv of frame l - 1 from S(l-2), S(l-1), S(l)
        if (l >= 2) {
            f32x4 acc2[4];
#pragma unroll
            for (int jj = 0; jj < 4; jj++) acc2[jj] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int tap = 2 * s + g / 2;  // K = tap * 16 + channel
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const bf16x8 bv = tap < 3 ? ring[(l - 2 + tap) % 3][wv][jj][m][g & 1] : zero8;
                    acc2[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wt[s], bv, acc2[jj], 0, 0, 0);
                }
            }
            if constexpr (PROJ) {
                // every lane takes part in the exchanges; only in-range points store
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    float part = 0.f;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        part += wpr[r] * (float)(bf16)fmaxf(acc2[jj][r] + bco[r], 0.0f);
                    part += __shfl_xor(part, 16, 64);
                    part += __shfl_xor(part, 32, 64);
                    const int j = j0 + jj;
                    if (outok && g == 0 && j <= G.J - 2)
                        out[nvox(G, i, j, k, lbase + l - 1)] = part + bpo;
                }
            } else if (outok) {
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const int j = j0 + jj;
                    if (j > G.J - 2) continue;
                    const long dst = nvox(G, i, j, k, lbase + l - 1) * kCh + 4 * g;
                    bf16 o[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) o[r] = (bf16)fmaxf(acc2[jj][r] + bco[r], 0.0f);
                    *reinterpret_cast<uint2*>(Y + dst) = *reinterpret_cast<const uint2*>(o);
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Weight gradient dW[tap][ci][co] = sum_v X[v + d(tap)][ci] dY[v][co] over the
// interior points of a range of rows, as per-block partials (FP32, summed by
// k_reduce_rows).  K of the MFMA is 32 x points of one row; X and dY are
// staged in LDS so that each lane can gather 8 consecutive points of one
// channel.  One wave per block; each block handles `rows_per_block` rows.
// ----------------------------------------------------------------------------
template <int CIN, int NT>
__global__ __launch_bounds__(64) void k_conv_dw(NGeo G, Taps T, const bf16* __restrict__ X,
                                                 const bf16* __restrict__ dY,
                                                 float* __restrict__ partial,
                                                 int rows_per_block) {
    // staged rows: the distinct (dy, dz, dl) rows of the taps, 34 points each
    // (x offsets -1..+1 around a 32-point chunk)
    constexpr int kRowsMax = 9;
    __shared__ bf16 sx[kRowsMax][34][CIN];
    __shared__ bf16 sy[32][kCh];
    const int ln = threadIdx.x;
    const int g = ln >> 4, m = ln & 15;
    // accumulators: per tap, a 16 (ci) x 16 (co) tile; CIN = 8 uses rows 0..7
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const long nrow = n_rows(G);
    for (int rr = 0; rr < rows_per_block; rr++) {
        const long r = (long)blockIdx.x * rows_per_block + rr;
        if (r >= nrow) break;
        int j, k, l;
        row_of(G, r, j, k, l);
        for (int x0 = 1; x0 <= G.I - 2; x0 += 32) {
            __syncthreads();
            // stage dY[x0 .. x0+31][0..15] (zero past the interior)
            for (int e = ln; e < 32 * kCh; e += 64) {
                const int xi = e / kCh, c = e % kCh;
                const int x = x0 + xi;
                sy[xi][c] = x <= G.I - 2 ? dY[nvox(G, x, j, k, l) * kCh + c] : (bf16)0.0f;
            }
            // stage X rows, points x0-1 .. x0+32 (within the padded row)
            const long rbase = nvox(G, 0, j, k, l);
            for (int q = 0; q < T.nrow; q++) {
                const long base = rbase + T.rowoff[q];
                for (int e = ln; e < 34 * CIN; e += 64) {
                    const int xi = e / CIN, c = e % CIN;
                    const int x = x0 - 1 + xi;
                    sx[q][xi][c] = x <= G.I - 1 ? X[(base + x) * CIN + c] : (bf16)0.0f;
                }
            }
            __syncthreads();
            // B = dY chunk: lane holds dY[x = 8 g + q][co = m]
            bf16x8 b;
#pragma unroll
            for (int q = 0; q < 8; q++) b[q] = sy[8 * g + q][m];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                // A = X^T: lane holds X[x = 8 g + q + dx][ci = m]
                bf16x8 a;
                const int dx = T.tdx[t];
                const int tr = T.trow[t];
#pragma unroll
                for (int q = 0; q < 8; q++)
                    a[q] = m < CIN ? sx[tr][8 * g + q + dx + 1][m < CIN ? m : 0] : (bf16)0.0f;
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
    // partial[block][tap][ci][co]; C[ci = 4 g + q][co = m]
    float* out = partial + (long)blockIdx.x * kMaxTaps * 16 * kCh;
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
        for (int q = 0; q < 4; q++) out[(t * 16 + 4 * g + q) * kCh + m] = acc[t][q];
    }
}

// Sum of per-block partials: out[e] = sum_b partial[b][e], fixed order.
__global__ __launch_bounds__(256) void k_reduce_rows(const float* __restrict__ partial, int nb,
                                                     int n, float* __restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float s = 0.f;
    for (int b = 0; b < nb; b++) s += partial[(long)b * n + e];
    out[e] = s;
}

// Bias gradient partials: sum over interior points of dY[v][co].
__global__ __launch_bounds__(256) void k_bias_grad(NGeo G, const bf16* __restrict__ dY,
                                                   float* __restrict__ partial,
                                                   int rows_per_block) {
    __shared__ float sm[256];
    const int co = threadIdx.x & 15, lane16 = threadIdx.x >> 4;
    float s = 0.f;
    const long nrow = n_rows(G);
    for (int rr = 0; rr < rows_per_block; rr++) {
        const long r = (long)blockIdx.x * rows_per_block + rr;
        if (r >= nrow) break;
        int j, k, l;
        row_of(G, r, j, k, l);
        for (int x = 1 + lane16; x <= G.I - 2; x += 16)
            s += (float)dY[nvox(G, x, j, k, l) * kCh + co];
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 16) {
        float t = 0.f;
        for (int q = 0; q < 16; q++) t += sm[q * 16 + threadIdx.x];
        partial[(long)blockIdx.x * kCh + threadIdx.x] = t;
    }
}

// 1x1 projection to one channel: out[v] = sum_c w[c] A[v][c] + b[0] (FP32).
__global__ __launch_bounds__(256) void k_project(NGeo G, const bf16* __restrict__ A,
                                                 const float* __restrict__ w,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ out) {
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    const long nv = (long)G.I * G.J * G.K * G.L;
    if (v >= nv) return;
    const int i = (int)(v % G.I);
    const long r1 = v / G.I;
    const int j = (int)(r1 % G.J);
    const long r2 = r1 / G.J;
    const int k = (int)(r2 % G.K), l = (int)(r2 / G.K);
    if (!interior(G, i, j, k, l)) return;
    const bf16x8* a = reinterpret_cast<const bf16x8*>(A + v * kCh);
    const bf16x8 a0 = a[0], a1 = a[1];
    float s = b[0];
#pragma unroll
    for (int q = 0; q < 8; q++) s += w[q] * (float)a0[q] + w[8 + q] * (float)a1[q];
    out[v] = s;
}

// Loss and the gradients of the projection: d = out - target on the loss
// points (lmask > 0); per-block partials of sum d^2, of sum d A[v][c] and of
// sum d; dA[v][c] = scale d w[c] where A[v][c] > 0 (the last ReLU), zero
// elsewhere (BF16).
__global__ __launch_bounds__(256) void k_loss_project_bwd(
    NGeo G, const float* __restrict__ out, const float* __restrict__ target,
    const unsigned char* __restrict__ lmask, const bf16* __restrict__ A,
    const float* __restrict__ w, float scale, bf16* __restrict__ dA,
    float* __restrict__ partial) {
    // partial[block][0] = sum d^2, [1..16] = sum d A_c, [17] = sum d
    __shared__ float sm[18][256];
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    const long nv = (long)G.I * G.J * G.K * G.L;
    float d = 0.f;
    float ac[16];
#pragma unroll
    for (int c = 0; c < 16; c++) ac[c] = 0.f;
    if (v < nv && lmask[v]) {
        d = out[v] - target[v];
        const bf16x8* a = reinterpret_cast<const bf16x8*>(A + v * kCh);
        const bf16x8 a0 = a[0], a1 = a[1];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            ac[q] = (float)a0[q];
            ac[8 + q] = (float)a1[q];
        }
        // A is the ReLU output of the last layer: its derivative masks dA
        bf16x8 g0, g1;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            g0[q] = (bf16)(ac[q] > 0.f ? scale * d * w[q] : 0.f);
            g1[q] = (bf16)(ac[8 + q] > 0.f ? scale * d * w[8 + q] : 0.f);
        }
        bf16x8* ga = reinterpret_cast<bf16x8*>(dA + v * kCh);
        ga[0] = g0;
        ga[1] = g1;
    } else if (v < nv) {
        // no loss here: zero gradient (ghost points stay zero as well)
        bf16x8 z;
#pragma unroll
        for (int q = 0; q < 8; q++) z[q] = (bf16)0.0f;
        bf16x8* ga = reinterpret_cast<bf16x8*>(dA + v * kCh);
        ga[0] = z;
        ga[1] = z;
    }
    sm[0][threadIdx.x] = d * d;
#pragma unroll
    for (int c = 0; c < 16; c++) sm[1 + c][threadIdx.x] = d * ac[c];
    sm[17][threadIdx.x] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 18; q++) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 18) partial[(long)blockIdx.x * 18 + threadIdx.x] = sm[threadIdx.x][0];
}

// Adam step on FP32 master weights; BF16 copy for the kernels (wb may be
// null for parameters used in FP32).
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ wm, float* __restrict__ m1,
                                              float* __restrict__ m2,
                                              const float* __restrict__ grad, int n, float lr,
                                              float b1, float b2, float eps, float c1, float c2,
                                              bf16* __restrict__ wb) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float gr = grad[e];
    const float a = b1 * m1[e] + (1.f - b1) * gr;
    const float bq = b2 * m2[e] + (1.f - b2) * gr * gr;
    m1[e] = a;
    m2[e] = bq;
    const float w = wm[e] - lr * (a / c1) / (sqrtf(bq / c2) + eps);
    wm[e] = w;
    if (wb) wb[e] = (bf16)w;
}

}  // namespace net
}  // namespace s4d
