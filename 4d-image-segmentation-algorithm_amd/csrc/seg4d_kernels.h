// seg4d_kernels.h -- HIP kernels (gfx950, wave64) for the semi-implicit
// subjective-surface / MCF time step.  Included by exactly one translation
// unit per shared library (seg4d_engine.hip).
//
// Arithmetic contract: every FP64 expression below is evaluated in the same
// grouping as the reference C (segmentation4D.c, mcf_4D.c), the build uses
// -ffp-contract=off, and FP64 division / sqrt are the correctly rounded IEEE
// operations, so a colour pass produces bitwise the values of the
// reference's pass on the same inputs.  Only reductions differ in summation
// order (they feed the stopping test, never the solution).
//
// Device data layout
//   full  : padded x-fastest, idx = ((l*K + k)*J + j)*I + i  (the reference's)
//   colour: the red (i+j+k+l even) and black points of every padded row are
//           stored in separate arrays, row length H = (I+1)/2, packed index
//           m = i >> 1: pidx = ((l*K + k)*J + j)*H + m.  A colour pass then
//           streams only its own colour's coefficients/prev/u and reads the
//           other colour at column m (rows j+-1, k+-1, l+-1) or m-1+2s (row).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "seg4d_net.h"  // LearnedStart

namespace s4d {

constexpr int kBlock = 256;

struct Geo {
    int I, J, K, L;   // padded dims
    int sy, sz, sw;   // full-layout strides
    int H;            // colour row length
    int py, pz, pw;   // colour-layout strides
    int c0;           // colour-layout offset of plane (k, l) = (0, 0)
    int RP;           // operand-block plane length: (J-2)*H rounded up to 64
};

// Colour-array index of packed column m of row (j, k, l).  The plane pitch
// pz is a multiple of 64 elements and c0 = -H mod 64, so the first interior
// row (j = 1) of every (k, l) plane starts on a 64-element boundary: the
// wavefronts of the passes, which walk the interior rows of a plane as one
// flat run of points, then load whole 128-byte lines.
__host__ __device__ __forceinline__ int cidx(const Geo& G, int j, int k, int l, int m) {
    return G.c0 + l * G.pw + k * G.pz + j * G.py + m;
}

// Operand blocks.  The 8 coefficients and prev of every interior point of
// one colour are stored interleaved per group of 64 consecutive points of an
// interior (k,l) row plane (flat index t = (j-1)*H + m):
//   [plane][t/64][component 0..8][t%64],  components e w n s t b fw bw prev.
// A wavefront of a SOR pass then reads its 9 operands as one contiguous
// 4.5 KB stream instead of 9 separate arrays (DRAM efficiency falls with the
// number of concurrent streams).
constexpr int kOpnd = 9;

// Order of the interior (k,l) planes in the operand blocks: frame pairs
// (1,2), (3,4), ... one after the other; within a pair k-major with the two
// frames of one k adjacent (a last unpaired frame: k-major alone).  The
// passes' automatic launch order works through chunks of 2 frames k plane by
// k plane, so each XCD then reads one contiguous operand stream instead of
// two (red pass at 256x256x128x16: 1.256 -> 1.238 ms in a timing copy).
__host__ __device__ __forceinline__ long opnd_plane(const Geo& G, int k, int l) {
    const int l0 = (l - 1) & ~1;                            // first frame of the pair - 1
    const int w = (G.L - 2 - l0) < 2 ? (G.L - 2 - l0) : 2;  // frames in the pair
    return (long)l0 * (G.K - 2) + (long)(k - 1) * w + ((l - 1) & 1);
}

__device__ __forceinline__ long opnd_index(const Geo& G, int k, int l, int t) {
    return opnd_plane(G, k, l) * G.RP * kOpnd + (long)(t >> 6) * (64 * kOpnd) + (t & 63);
}

// Split operands of the FP32-refined form (written by its set-up kernel): the
// FP32 operand blocks hold the HIGH 32-bit word of each FP64 coefficient
// (components 0..7; the FP32 passes use it as a 20-bit-mantissa coefficient,
// converted exactly) and the FP32 right-hand side (component 8); the FP64
// operand buffer holds, per 64-point group, the LOW words of the 8
// coefficients and the two words of prev as 32-bit words [10][64]
// (kSplit components, in the same allocation).  High and low word give the
// FP64 coefficient back exactly, so the FP64 defect reads the same 72 bytes
// per point as from full FP64 blocks, while the set-up writes 32 bytes per
// point less.
constexpr int kSplit = 10;

__device__ __forceinline__ long split_index(const Geo& G, int k, int l, int t) {
    return opnd_plane(G, k, l) * G.RP * kSplit + (long)(t >> 6) * (64 * kSplit) + (t & 63);
}

// FP32 coefficient of a split high word (exact: 20 mantissa bits fit in 23).
__device__ __forceinline__ float hiword_coef(float bits) {
    return (float)__hiloint2double(__float_as_int(bits), 0);
}

struct Ptr8 {
    double* p[8];
};
struct CPtr8 {
    const double* p[8];
};

// Per-solve control block in device memory.  `stop` is read by every black
// pass so that a pass enqueued ahead of the host's decision becomes a no-op.
struct SorCtl {
    double R;       // residual of the last reduced iteration
    double tol;     // absolute bound the residual is compared with
    double sumsq;   // sum over the interior of prev^2 h^4 (relative mode)
    int z;          // last reduced iteration
    int stop;       // 1 once !(R > tol && z < cap)
    int zfinal;     // iteration at which the solve stopped
    int pad;
};

// ----------------------------------------------------------------------------
// reductions (fixed order => run-to-run deterministic)
// ----------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over a 256-thread block; result valid in thread 0.  Every call of a
// block shares `sm`: the leading barrier keeps a call's writes behind thread
// 0's reads of the previous call.
__device__ __forceinline__ double block_sum256(double v) {
    __shared__ double sm[kBlock / 64];
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sm[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = ((sm[0] + sm[1]) + sm[2]) + sm[3];
    return r;
}

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X.  Remap the
// hardware block id so that each XCD receives one contiguous range of
// logical blocks: neighbouring k-planes then share an XCD (and its L2).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int nx = 8;
    const int per = nb / nx, rem = nb % nx;
    const int x = b % nx;
    return x * per + (x < rem ? x : rem) + b / nx;
}

// ----------------------------------------------------------------------------
// face gradients (segmentation4D.c:456-560 for u, :596-820 for the image)
// ----------------------------------------------------------------------------
struct Scale {
    double h, q;          // h and quotient 4h
    double inv_h, inv_q;  // exact reciprocals when h is a power of two
};

template <bool POW2>
__device__ __forceinline__ double scl(double d, double s, double inv) {
    // When s is a power of two, d * (1/s) is the correctly rounded d / s.
    if constexpr (POW2) return d * inv;
    else return d / s;
}

// Sum of squares of the 4 gradient components of face F at full index p:
// ((ux^2 + uy^2) + uz^2) + uw^2, components as in the reference.
template <int F, bool POW2>
__device__ __forceinline__ double face_sq(const double* __restrict__ u, int p,
                                          const Geo& G, const Scale& S) {
    const int st[4] = {1, G.sy, G.sz, G.sw};
    constexpr int A = F >> 1;
    const int dA = (F & 1) ? -st[A] : st[A];
    double acc = 0.0;
#pragma unroll
    for (int B = 0; B < 4; B++) {
        double c;
        if (B == A) {
            c = scl<POW2>(u[p + dA] - u[p], S.h, S.inv_h);
        } else {
            c = scl<POW2>((u[p + dA + st[B]] + u[p + st[B]]) -
                          (u[p + dA - st[B]] + u[p - st[B]]), S.q, S.inv_q);
        }
        acc = (B == 0) ? c * c : acc + c * c;
    }
    return acc;
}

// The 33-point stencil of u around p (centre, 8 faces, 24 edges) loaded into
// registers before any arithmetic, so that all its loads are in flight
// together: s[dw+1][dz+1][dy+1][dx+1] (only points with at most two nonzero
// offsets are loaded).  The face sums are then formed exactly as face_sq does.
struct Stencil33 {
    double s[3][3][3][3];
};

__device__ __forceinline__ void load_stencil33(const double* __restrict__ u, int p,
                                               const Geo& G, Stencil33& T) {
    const int st[4] = {1, G.sy, G.sz, G.sw};
#pragma unroll
    for (int w = 0; w < 3; w++)
#pragma unroll
        for (int z = 0; z < 3; z++)
#pragma unroll
            for (int y = 0; y < 3; y++)
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    const int nz = (x != 1) + (y != 1) + (z != 1) + (w != 1);
                    if (nz <= 2)
                        T.s[w][z][y][x] = u[p + (x - 1) * st[0] + (y - 1) * st[1] +
                                            (z - 1) * st[2] + (w - 1) * st[3]];
                }
}

template <int F, bool POW2>
__device__ __forceinline__ double face_sq_reg(const Stencil33& T, const Scale& S) {
    constexpr int A = F >> 1;
    constexpr int sa = (F & 1) ? -1 : 1;
    // value at offset o (o[axis] in {-1, 0, 1})
    auto at = [&](int o0, int o1, int o2, int o3) -> double {
        return T.s[o3 + 1][o2 + 1][o1 + 1][o0 + 1];
    };
    auto off = [&](int axis, int v, int q) -> int { return axis == q ? v : 0; };
    double acc = 0.0;
#pragma unroll
    for (int B = 0; B < 4; B++) {
        double c;
        if (B == A) {
            c = scl<POW2>(at(off(A, sa, 0), off(A, sa, 1), off(A, sa, 2), off(A, sa, 3)) -
                              at(0, 0, 0, 0),
                          S.h, S.inv_h);
        } else {
            const double pApB = at(off(A, sa, 0) + off(B, 1, 0), off(A, sa, 1) + off(B, 1, 1),
                                   off(A, sa, 2) + off(B, 1, 2), off(A, sa, 3) + off(B, 1, 3));
            const double pB = at(off(B, 1, 0), off(B, 1, 1), off(B, 1, 2), off(B, 1, 3));
            const double pAmB = at(off(A, sa, 0) - off(B, 1, 0), off(A, sa, 1) - off(B, 1, 1),
                                   off(A, sa, 2) - off(B, 1, 2), off(A, sa, 3) - off(B, 1, 3));
            const double mB = at(-off(B, 1, 0), -off(B, 1, 1), -off(B, 1, 2), -off(B, 1, 3));
            c = scl<POW2>((pApB + pB) - (pAmB + mB), S.q, S.inv_q);
        }
        acc = (B == 0) ? c * c : acc + c * c;
    }
    return acc;
}

template <bool POW2>
__device__ __forceinline__ void all_face_sq_reg(const Stencil33& T, const Scale& S, double a[8]) {
    a[0] = face_sq_reg<0, POW2>(T, S);
    a[1] = face_sq_reg<1, POW2>(T, S);
    a[2] = face_sq_reg<2, POW2>(T, S);
    a[3] = face_sq_reg<3, POW2>(T, S);
    a[4] = face_sq_reg<4, POW2>(T, S);
    a[5] = face_sq_reg<5, POW2>(T, S);
    a[6] = face_sq_reg<6, POW2>(T, S);
    a[7] = face_sq_reg<7, POW2>(T, S);
}

template <bool POW2>
__device__ __forceinline__ void all_face_sq(const double* __restrict__ u, int p,
                                            const Geo& G, const Scale& S,
                                            double a[8]) {
    a[0] = face_sq<0, POW2>(u, p, G, S);
    a[1] = face_sq<1, POW2>(u, p, G, S);
    a[2] = face_sq<2, POW2>(u, p, G, S);
    a[3] = face_sq<3, POW2>(u, p, G, S);
    a[4] = face_sq<4, POW2>(u, p, G, S);
    a[5] = face_sq<5, POW2>(u, p, G, S);
    a[6] = face_sq<6, POW2>(u, p, G, S);
    a[7] = face_sq<7, POW2>(u, p, G, S);
}

__device__ __forceinline__ bool is_interior(const Geo& G, int i, int j, int k, int l) {
    return i >= 1 && i <= G.I - 2 && j >= 1 && j <= G.J - 2 && k >= 1 &&
           k <= G.K - 2 && l >= 1 && l <= G.L - 2;
}

// ----------------------------------------------------------------------------
// Step set-up: coefficients + colour packing + prev copy + sum prev^2.
// segmentation4D.c:224 (copy), :446-580 (coefficients); mcf_4D.c:208-342.
// Grid: x over the padded (j,i) plane, y = k, z = l (all padded).
// ----------------------------------------------------------------------------
struct SplitBufs {
    double* opnd[2];     // [colour] operand blocks (coefficients + prev)
    double* ured[2];     // red ping-pong buffers
    double* ublack;
    // FP32-refined form with the first defect fused (MIXED): FP32 operand
    // blocks (coefficients + defect) and the initial corrections per colour
    float* opnd32[2];
    float* dred;
    float* dblack;
};

// Block -> (run bx, plane k, frame l) map of full-layout kernels whose
// stencil reaches the k+-1 and l+-1 planes: every XCD processes whole tiles
// of tx runs x tk planes x tl frames per residency round, so the planes a
// 33-point stencil re-reads are fetched once into that XCD's L2.
struct PlaneMap {
    int gx, gk, gl;       // runs of kBlock points per (j,i) plane, K, L
    int tx, tk, tl;
    int ntx, ntk, ntiles;
};

__device__ __forceinline__ bool map_plane_block(const PlaneMap& M, int b, int& bx, int& k,
                                                int& l) {
    const int x = b % 8, s = b / 8;
    const int per = M.tx * M.tk * M.tl;
    const int r = s / per, w = s - r * per;
    const int T = r * 8 + x;
    if (T >= M.ntiles) return false;
    const int tix = T % M.ntx, tik = (T / M.ntx) % M.ntk, til = T / (M.ntx * M.ntk);
    bx = tix * M.tx + w % M.tx;
    k = tik * M.tk + (w / M.tx) % M.tk;
    l = til * M.tl + w / (M.tx * M.tk);
    return bx < M.gx && k < M.gk && l < M.gl;
}

// warm: 0 = SOR starts from u^n (the reference); 1 = it starts from the
// linear extrapolation 2 u^n - u^(n-1) (prev_full holds u^(n-1) on entry);
// 2 = from the learned start (interior points, LearnedStart: the network's
// correction of u^n, of the linear or of the cubic extrapolation); 3 = from the
// quadratic extrapolation 3 u^n - 3 u^(n-1) + u^(n-2) (prev2 holds u^(n-2));
// 4 = from the cubic one 4 u^n - 6 u^(n-1) + 4 u^(n-2) - u^(n-3) (prev3
// holds u^(n-3)).  u^n is not copied: the host rotates the full-layout
// buffers after the launch (u's buffer becomes the history's newest level
// and the oldest level's buffer receives u^(n+1) from k_unpack).
// MIXED: the set-up of the FP32-refined form, fused with its first FP64
// defect.  The colour arrays receive u^n and the FP32 correction arrays the
// start's offset d0 = start - u^n (the correction solve starts from it); the
// FP32 operand blocks receive the coefficients and -r for r = b - A u^n,
// formed as the defect pass forms it (b = prev = u^n, the 8 face neighbours
// from the stencil), and partials2 the block sums of r^2 h^4.  c = tau / h^2.
//
// Lanes and points: wavefront w of block bx takes the 32 packed columns
// t2 = 32 q + (lane & 31), q = q0 + 4 bx + w, of its (k, l) plane, where
// t2 = (j - 1) H + m runs over the plane's rows (row 0 from t2 = -H,
// q0 = -ceil(H / 32)); lanes 0-31 take those columns' red points, lanes
// 32-63 their black points.  Every store of a wavefront -- the 32 operand
// words of each component and colour, the 32 colour-array values of each
// colour -- then starts on a 32-element boundary (operand blocks of 64, the
// colour layout's aligned rows), and the loads of the full layout still
// cover one contiguous run of 64 points.  (Lanes on consecutive points of
// the full layout instead: MIXED 6.81 against 6.15 ms at 256x256x128x16.)
template <bool WITH_G, bool POW2, bool MIXED = false>
static __global__ __launch_bounds__(kBlock) void k_setup_step(
    Geo G, PlaneMap PM, Scale S, double eps2, double h4, const double* __restrict__ u,
    CPtr8 g, SplitBufs sb, const double* hist1, const double* hist2, const double* hist3,
    LearnedStart ls, double* __restrict__ partials,
    int warm, double cc = 0.0, double* __restrict__ partials2 = nullptr) {
    int bx, k, l;
    if (!map_plane_block(PM, blockIdx.x, bx, k, l)) return;
    const int lane = (int)(threadIdx.x & 63);
    const int q = 4 * bx + (int)(threadIdx.x >> 6) - (G.H + 31) / 32;
    const int s = 32 * q + (lane & 31) + G.H;  // j H + m
    const int jq = s >= 0 ? s / G.H : 0;
    const int iq = 2 * (s - jq * G.H) + ((lane >> 5) ^ ((jq + k + l) & 1));
    double acc = 0.0, acc2 = 0.0;
    if (s >= 0 && jq < G.J && iq < G.I) {
        const int j = jq, i = iq;
        const int p = ((l * G.K + k) * G.J + j) * G.I + i;
        const int c = (i + j + k + l) & 1;
        const int pp = cidx(G, j, k, l, i >> 1);
        const double uc = u[p];
        double start = uc;
        const bool in = is_interior(G, i, j, k, l);
        // history u^(n-1), u^(n-2), u^(n-3) as the start needs it
        const bool cub = warm == 4 || (warm == 2 && ls.order == 3);
        const double pf = warm >= 1 ? hist1[p] : 0.0;
        const double p2 = (warm == 3 || cub) ? hist2[p] : 0.0;
        const double p3 = cub ? hist3[p] : 0.0;
        if (warm == 1 && in) start = 2.0 * uc - pf;
        if (warm == 3 && in) start = (3.0 * uc - 3.0 * pf) + p2;
        if (cub && in) start = ((4.0 * uc - 6.0 * pf) + 4.0 * p2) - p3;
        if (warm == 2 && in) {
            // the extrapolation the network corrects: cubic, linear or none
            const double base = ls.order == 3 ? start : ls.order == 1 ? uc + (uc - pf) : uc;
            start = base + sqrt(*ls.dsum * ls.inv_n) * (double)(ls.out[p] * ls.inv_st);
        }
        if constexpr (MIXED) {
            // (at ghosts start = u^n, so d0 = 0 there)
            if (c == 0) {
                sb.ured[0][pp] = uc;
                sb.dred[pp] = (float)(start - uc);
            } else {
                sb.ublack[pp] = uc;
                sb.dblack[pp] = (float)(start - uc);
            }
        } else if (c == 0) {
            sb.ured[0][pp] = start;
            // the second red buffer's interior is written by the first red
            // pass before anything reads it: only its ghosts are set here
            if (!in) sb.ured[1][pp] = start;
        } else {
            sb.ublack[pp] = start;
        }
        if (is_interior(G, i, j, k, l)) {
            const int t2 = (j - 1) * G.H + (i >> 1);
            const long oi = opnd_index(G, k, l, t2);
            double* ob = (c ? sb.opnd[1] : sb.opnd[0]) + oi;
            // MIXED: split words (split_index) instead of full FP64 blocks
            unsigned* ow = reinterpret_cast<unsigned*>(c ? sb.opnd[1] : sb.opnd[0]) +
                           split_index(G, k, l, t2);
            if constexpr (MIXED) {
                ow[8 * 64] = (unsigned)__double2hiint(uc);
                ow[9 * 64] = (unsigned)__double2loint(uc);
            } else {
                ob[8 * 64] = uc;
            }
            Stencil33 T;
            load_stencil33(u, p, G, T);
            double gl[8];
            if constexpr (WITH_G) {
#pragma unroll
                for (int f = 0; f < 8; f++) gl[f] = g.p[f][p];
            }
            double a[8];
            all_face_sq_reg<POW2>(T, S, a);
            double N[8];
#pragma unroll
            for (int f = 0; f < 8; f++) N[f] = sqrt(a[f] + eps2);
            const double avg =
                (((((((N[0] + N[1]) + N[2]) + N[3]) + N[4]) + N[5]) + N[6]) + N[7]) / 8.0;
            const double vc = sqrt(avg * avg + eps2);
            double cf[8];
#pragma unroll
            for (int f = 0; f < 8; f++) {
                if constexpr (WITH_G) cf[f] = vc * gl[f] * (1.0 / N[f]);
                else cf[f] = vc * (1.0 / N[f]);
                if constexpr (MIXED) ow[f * 64] = (unsigned)__double2loint(cf[f]);
                else ob[f * 64] = cf[f];
            }
            acc = uc * uc * h4;
            if constexpr (MIXED) {
                float* ob32 = (c ? sb.opnd32[1] : sb.opnd32[0]) + oi;
#pragma unroll
                for (int f = 0; f < 8; f++) ob32[f * 64] = __int_as_float(__double2hiint(cf[f]));
                // numerator order e, w, s, n, b, t, bw, fw; diagonal e, w, n, s, t, b, fw, bw
                double num = cf[0] * T.s[1][1][1][2];
                num = num + cf[1] * T.s[1][1][1][0];
                num = num + cf[3] * T.s[1][1][0][1];
                num = num + cf[2] * T.s[1][1][2][1];
                num = num + cf[5] * T.s[1][0][1][1];
                num = num + cf[4] * T.s[1][2][1][1];
                num = num + cf[7] * T.s[0][1][1][1];
                num = num + cf[6] * T.s[2][1][1][1];
                double dg = cf[0] + cf[1];
                dg = dg + cf[2];
                dg = dg + cf[3];
                dg = dg + cf[4];
                dg = dg + cf[5];
                dg = dg + cf[6];
                dg = dg + cf[7];
                const double D = 1 + cc * dg;
                const double r = uc * D - cc * num - uc;
                ob32[8 * 64] = (float)(-r);
                acc2 = r * r * h4;
            }
        }
    }
    const double bs = block_sum256(acc);
    if (threadIdx.x == 0) partials[(l * PM.gk + k) * PM.gx + bx] = bs;
    if constexpr (MIXED) {
        const double bs2 = block_sum256(acc2);
        if (threadIdx.x == 0) partials2[(l * PM.gk + k) * PM.gx + bx] = bs2;
    }
}

// ----------------------------------------------------------------------------
// Red-black SOR colour pass on the colour layout, marching along l.
// segmentation4D.c:233-285.  One thread owns packed column m of rows (j,k)
// and walks l = 1..L-2, keeping the other colour's column-m values of rows
// l-1, l, l+1 in registers (they are the bw / in-row / fw neighbours).
//
// RESID = 0: no residual; 1: residual of the point BEFORE its update (the
// red pass of iteration z+1 yields the red residual terms of iteration z:
// red values and black neighbours are then exactly those at the end of z);
// 2: residual AFTER the update (black pass of iteration z: its red
// neighbours are final for z).  segmentation4D.c:287-308.
// ----------------------------------------------------------------------------
// Block -> (run bx, plane k) map of the SOR passes (1-D grid).
//  linear: XCD-contiguous ranges of the run-major order (xcd_remap);
//  tiled : every XCD processes whole tiles of tb runs x tk planes, one tile
//          per residency round, so that the blocks sharing the other colour's
//          k+-1 planes are co-resident on one XCD (one L2) and march l in step.
struct SweepMap {
    int gx, gk;      // block grid: runs of kBlock / kp points per row plane,
                     // (groups of kp consecutive interior k planes) x (l chunks)
    int kp;          // k planes per block (1, 2, 4): the block's planes share
                     // their other-colour neighbour planes through L1/L2
    int gk0, lc;     // k groups; l steps per block (interior frames split into
                     // gk / gk0 chunks of lc, so that small volumes fill the GPU)
    int tiled;
    int tb, tk;      // tile shape; tb * tk blocks per tile
    int ntb, ntiles;
    int nband, bpb;  // linear order: the runs of a plane in nband bands of bpb runs;
                     // consecutive blocks walk one band through the k planes
};

// Blocks of a SOR launch with map M (some return at once when the shape
// does not divide evenly).
__host__ __device__ inline int sweep_grid(const SweepMap& M) {
    if (M.tiled) return ((M.ntiles + 7) / 8) * 8 * M.tb * M.tk;
    if (M.nband > 1) return M.nband * M.bpb * M.gk;
    return M.gx * M.gk;
}

__device__ __forceinline__ bool map_block(const SweepMap& M, int b, int nb, int& bx, int& k) {
    if (!M.tiled) {
        const int lb = xcd_remap(b, nb);
        if (M.nband <= 1) {
            bx = lb % M.gx;
            k = lb / M.gx + 1;
            return true;
        }
        // lb = ((chunk * nband + band) * gk0 + k group) * bpb + run in band
        const int r = lb % M.bpb;
        int q = lb / M.bpb;
        const int kg = q % M.gk0;
        q /= M.gk0;
        const int band = q % M.nband, ch = q / M.nband;
        bx = band * M.bpb + r;
        k = ch * M.gk0 + kg + 1;
        return bx < M.gx && k <= M.gk;
    }
    const int x = b % 8, s = b / 8;
    const int per = M.tb * M.tk;
    const int r = s / per, w = s - r * per;
    const int T = r * 8 + x;
    if (T >= M.ntiles) return false;
    const int tbi = T % M.ntb, tki = T / M.ntb;
    bx = tbi * M.tb + w % M.tb;
    k = tki * M.tk + w / M.tb + 1;
    return bx < M.gx && k <= M.gk;
}

// Operands of one colour point, gathered before any arithmetic so that the
// loads of U consecutive l-steps are all in flight together.  T = double for
// the reference system, float for the correction equation of the FP32-refined
// form (there component 8 of the operand block holds the right-hand side).
template <typename T>
struct PointIn {
    T c[8];          // e w n s t b fw bw
    T pv, uo;        // right-hand side (prev), own value
    T yo;            // in-row neighbour that is not the column partner
    T yS, yN, yB, yT;
    int pp;
    int sft;         // own i = 2m + sft
    bool act;
};

// NTO: the operand-block loads carry the nontemporal hint (the stream is read
// once per pass; the hint keeps it from displacing the other colour's planes,
// which neighbouring blocks re-read, from L2).
// A value of the other colour; CORR: plus its FP32 correction dy when given
// (the corrected state u + d of the FP32-refined form, formed as u[q] + d[q]
// exactly like the separate correction update).
template <typename T, bool CORR>
__device__ __forceinline__ T other(const T* __restrict__ y, const float* __restrict__ dy, int q) {
    if constexpr (CORR) return dy ? y[q] + (T)dy[q] : y[q];
    else return y[q];
}

// SPLIT (FP64 only): the coefficients from the split operands, high words in
// hi (the FP32 operand block of the point) and low words and prev in ow.
// FP32: the operand block's components 0..7 are split high words.
template <typename T, bool NTO = false, bool CORR = false, bool SPLIT = false>
__device__ __forceinline__ void load_point(const Geo& G, const T* __restrict__ ob,
                                           const T* __restrict__ y, const T* xin, int pp,
                                           int sft, bool act, PointIn<T>& P,
                                           const float* __restrict__ dx = nullptr,
                                           const float* __restrict__ dy = nullptr,
                                           const float* __restrict__ hi = nullptr,
                                           const unsigned* __restrict__ ow = nullptr) {
    P.pp = pp;
    P.sft = sft;
    P.act = act;
    if (act) {
        P.yo = other<T, CORR>(y, dy, pp + (sft ? 1 : -1));
        P.yS = other<T, CORR>(y, dy, pp - G.py);
        P.yN = other<T, CORR>(y, dy, pp + G.py);
        P.yB = other<T, CORR>(y, dy, pp - G.pz);
        P.yT = other<T, CORR>(y, dy, pp + G.pz);
        if constexpr (SPLIT) {
#pragma unroll
            for (int f = 0; f < 8; f++)
                P.c[f] = __hiloint2double(__float_as_int(hi[f * 64]), (int)ow[f * 64]);
            P.pv = __hiloint2double((int)ow[8 * 64], (int)ow[9 * 64]);
        } else if constexpr (NTO) {
#pragma unroll
            for (int f = 0; f < 8; f++) P.c[f] = __builtin_nontemporal_load(ob + f * 64);
            P.pv = __builtin_nontemporal_load(ob + 8 * 64);
        } else {
#pragma unroll
            for (int f = 0; f < 8; f++) P.c[f] = ob[f * 64];
            P.pv = ob[8 * 64];
        }
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int f = 0; f < 8; f++) P.c[f] = (T)hiword_coef((float)P.c[f]);
        }
        if constexpr (CORR) P.uo = xin[pp] + (T)dx[pp];
        else P.uo = xin[pp];
    }
}

// One point update, reference operand order (segmentation4D.c:242-256); the
// residual term r = u (1 + c sum) - c sum coef u_nb - prev (segmentation4D.c:
// 290-307) before (RESID 1) or after (RESID 2) the update.  RESID 3 does not
// update: it writes -r (the defect, FP32) to *defect and accumulates r^2 h^4;
// RESID 4 also writes the high words of the coefficients (split operands),
// 8..1 components before *defect; RESID 5 is RESID 3 on the corrected state
// u + d, whose own value it stores (the defect-correction update fused into
// the next defect); RESID 6 is RESID 5 on split operands.
template <typename T, int RESID>
__device__ __forceinline__ void update_point(const PointIn<T>& P, T ybw, T ypart, T yfw, T c,
                                             T om, double h4, T* xout, float* defect,
                                             double& acc) {
    if (!P.act) return;
    const T uW = P.sft ? ypart : P.yo;
    const T uE = P.sft ? P.yo : ypart;
    // numerator order e, w, s, n, b, t, bw, fw
    T num = P.c[0] * uE;
    num = num + P.c[1] * uW;
    num = num + P.c[3] * P.yS;
    num = num + P.c[2] * P.yN;
    num = num + P.c[5] * P.yB;
    num = num + P.c[4] * P.yT;
    num = num + P.c[7] * ybw;
    num = num + P.c[6] * yfw;
    // diagonal order e, w, n, s, t, b, fw, bw
    T dg = P.c[0] + P.c[1];
    dg = dg + P.c[2];
    dg = dg + P.c[3];
    dg = dg + P.c[4];
    dg = dg + P.c[5];
    dg = dg + P.c[6];
    dg = dg + P.c[7];
    const T D = 1 + c * dg;
    if constexpr (RESID == 1 || RESID >= 3) {
        const T r = P.uo * D - c * num - P.pv;
        const double rd = (double)r;
        acc += rd * rd * h4;
        if constexpr (RESID >= 3) {
            *defect = (float)(-r);
            if constexpr (RESID == 4) {  // high words of the 8 coefficients
#pragma unroll
                for (int f = 0; f < 8; f++)
                    defect[(f - 8) * 64] = __int_as_float(__double2hiint((double)P.c[f]));
            }
            if constexpr (RESID >= 5) __builtin_nontemporal_store(P.uo, xout + P.pp);
            return;
        }
    }
    const T gs = (P.pv + c * num) / D;
    const T un = P.uo + om * (gs - P.uo);
    // streaming store: the new value is read again only by the next pass, so
    // it need not displace the other colour's planes from L2 (measured 1%
    // faster than a plain store at 256x256x128x16)
    __builtin_nontemporal_store(un, xout + P.pp);
    if constexpr (RESID == 2) {
        const T r = un * D - c * num - P.pv;
        const double rd = (double)r;
        acc += rd * rd * h4;
    }
}

// ----------------------------------------------------------------------------
// Red-black SOR colour pass on the colour layout, marching along l, U steps
// per group (all loads of a group are issued before its arithmetic).
// segmentation4D.c:233-285.  One thread owns packed column m of rows (j,k)
// and walks l = 1..L-2, keeping the other colour's column-m values of rows
// l-1 .. l+U in registers (they are the bw / in-row / fw neighbours).
//
// RESID = 0: no residual; 1: residual of the point BEFORE its update (the
// red pass of iteration z+1 yields the red residual terms of iteration z:
// red values and black neighbours are then exactly those at the end of z);
// 2: residual AFTER the update (black pass of iteration z: its red
// neighbours are final for z).  segmentation4D.c:287-308.  3: residual only,
// the defect -r written as FP32 into component 8 of the operand blocks
// `defect_ob` (the right-hand side of the FP32 correction equation); 4: as 3,
// and the high words of the 8 coefficients into components 0..7; 5: as 3 on
// u + d (own correction dx, the other colour's dy), storing u + d; 6: as 5
// with the coefficients and prev from the split operands (high words in
// `defect_ob`, low words and prev in `opnd` as 32-bit words).
// ----------------------------------------------------------------------------
// The work of one block (run bx, k group / l chunk kb), shared by the
// single-volume and the batched kernels; every thread of the block calls it.
template <typename T, int RESID, int U, bool NTO>
__device__ __forceinline__ void sor_block(
    const Geo& G, const SweepMap& M, int bx, int kb, T c, T om, double h4, int colour,
    const T* xin, T* xout, const T* __restrict__ y,
    const T* __restrict__ opnd, double* __restrict__ partials,
    float* __restrict__ defect_ob, const float* __restrict__ dx = nullptr,
    const float* __restrict__ dy = nullptr) {
    constexpr bool CORR = RESID >= 5;  // dx: own correction, dy: the other colour's (or none)
    constexpr bool SPLIT = RESID == 6;
    const int lb = (kb - 1) * M.gx + bx;
    // block index kb - 1 = l chunk * gk0 + k group; the block's kp planes:
    // threads [sub * per, (sub + 1) * per) take plane kgrp kp + sub + 1, run bx
    // of `per` points, frames l0 .. l1
    const int kgrp = (kb - 1) % M.gk0, lch = (kb - 1) / M.gk0;
    const int l0 = 1 + lch * M.lc;
    const int per = kBlock / M.kp;
    const int sub = threadIdx.x / per;
    const int kk = kgrp * M.kp + sub + 1;
    const bool kok = kk <= G.K - 2;
    const int k = kok ? kk : G.K - 2;  // indices stay in range for idle threads
    const int t = bx * per + (threadIdx.x - sub * per);
    const int jr = t / G.H;
    const int m = t - jr * G.H;
    const int j = jr + 1;
    const bool rowok = kok && j <= G.J - 2;
    double acc = 0.0;
    const int base = cidx(G, j, k, 0, m);  // l = 0
    // operand block of (k, l, t): the plane's offset (opnd_plane) plus the
    // point's place in its 64-point group; split operands: high words in the
    // FP32 block (same index), low words and prev in the FP64 buffer's
    // [10][64] word groups
    const long og = (long)(t >> 6) * (64 * kOpnd) + (t & 63);
    const long sg = (long)(t >> 6) * (64 * kSplit) + (t & 63);
    const long opl = (long)G.RP * kOpnd, spl = (long)G.RP * kSplit;
    auto ooff = [&](int lq) { return opnd_plane(G, k, lq) * opl + og; };
    auto soff = [&](int lq) { return opnd_plane(G, k, lq) * spl + sg; };
    const unsigned* ow = SPLIT ? reinterpret_cast<const unsigned*>(opnd) : nullptr;
    // window w[q] = other colour at column m, row l-1+q
    T w[U + 2];
    w[0] = rowok ? other<T, CORR>(y, dy, base + (l0 - 1) * G.pw) : T(0);
    w[1] = rowok ? other<T, CORR>(y, dy, base + l0 * G.pw) : T(0);
    const int lmax = min(l0 + M.lc - 1, G.L - 2);
    int l = l0;
    for (; l + U - 1 <= lmax; l += U) {
        PointIn<T> P[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int lq = l + q;
            const int pp = base + lq * G.pw;
            w[q + 2] = rowok ? other<T, CORR>(y, dy, pp + G.pw) : T(0);
            const int sft = colour ^ ((j + k + lq) & 1);
            const int i = 2 * m + sft;
            load_point<T, NTO, CORR, SPLIT>(G, opnd + ooff(lq), y, xin, pp, sft,
                                            rowok && i >= 1 && i <= G.I - 2, P[q], dx, dy,
                                            SPLIT ? defect_ob + ooff(lq) : nullptr,
                                            SPLIT ? ow + soff(lq) : nullptr);
        }
#pragma unroll
        for (int q = 0; q < U; q++)
            update_point<T, RESID>(P[q], w[q], w[q + 1], w[q + 2], c, om, h4, xout,
                                   RESID >= 3 ? defect_ob + ooff(l + q) + 8 * 64 : nullptr, acc);
        w[0] = w[U];
        w[1] = w[U + 1];
    }
    for (; l <= lmax; l++) {  // remainder, one step at a time
        PointIn<T> P;
        const int pp = base + l * G.pw;
        w[2] = rowok ? other<T, CORR>(y, dy, pp + G.pw) : T(0);
        const int sft = colour ^ ((j + k + l) & 1);
        const int i = 2 * m + sft;
        load_point<T, NTO, CORR, SPLIT>(G, opnd + ooff(l), y, xin, pp, sft,
                                        rowok && i >= 1 && i <= G.I - 2, P, dx, dy,
                                        SPLIT ? defect_ob + ooff(l) : nullptr,
                                        SPLIT ? ow + soff(l) : nullptr);
        update_point<T, RESID>(P, w[0], w[1], w[2], c, om, h4, xout,
                               RESID >= 3 ? defect_ob + ooff(l) + 8 * 64 : nullptr, acc);
        w[0] = w[1];
        w[1] = w[2];
    }
    if constexpr (RESID != 0) {
        const double bs = block_sum256(acc);
        if (threadIdx.x == 0) partials[lb] = bs;
    }
}

template <typename T, int RESID, int U, bool NTO = false>
static __global__ __launch_bounds__(kBlock) void k_sor_sweep(
    Geo G, SweepMap M, T c, T om, double h4, int colour,
    const T* xin, T* xout, const T* __restrict__ y,
    const T* __restrict__ opnd, double* __restrict__ partials,
    const SorCtl* __restrict__ ctl, float* __restrict__ defect_ob,
    const float* __restrict__ dx = nullptr, const float* __restrict__ dy = nullptr) {
    if (ctl != nullptr && ctl->stop) return;
    int bx, kb;
    if (!map_block(M, blockIdx.x, gridDim.x, bx, kb)) return;
    sor_block<T, RESID, U, NTO>(G, M, bx, kb, c, om, h4, colour, xin, xout, y, opnd, partials,
                                defect_ob, dx, dy);
}

// Sets the control block of one FP32 correction solve: stop when its residual
// is below max(floor2 * R, frac * tol) of the outer FP64 system.
static __global__ void k_inner_ctl(const SorCtl* __restrict__ outer, SorCtl* inner, double frac,
                            double floor2) {
    if (threadIdx.x == 0) {
        const double a = frac * outer->tol, b2 = floor2 * outer->R;
        inner->tol = a > b2 ? a : b2;
        inner->R = 0.0;
        inner->z = 0;
        inner->stop = 0;
        inner->zfinal = 0;
    }
}

// Fixed-order sum of the red and black block partials of one iteration by a
// block of 1024 threads; the result is valid in thread 0.
__device__ __forceinline__ double resid_sum(const double* __restrict__ pa, int na,
                                            const double* __restrict__ pb, int nbp) {
    double s = 0.0;
    for (int q = threadIdx.x; q < na; q += 1024) s += pa[q];
    double s2 = 0.0;
    for (int q = threadIdx.x; q < nbp; q += 1024) s2 += pb[q];
    s = s + s2;
    s = wave_sum(s);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    double R = sm[0];
    for (int q = 1; q < 16; q++) R += sm[q];
    return R;
}

// First level of the residual reduction of one SOR iteration: block q sums
// chunk (q mod nc) of the red (q < nc) or black partials, ch partials each,
// into out[q]; the stopping test then reduces the 2 nc chunk sums.  A queued
// reduction of a stopped solve (ctl->stop) returns at once.
__device__ __forceinline__ double chunk_sum(const double* __restrict__ pa,
                                           const double* __restrict__ pb, int nb, int nc,
                                           int ch, int q) {
    const double* src = q < nc ? pa : pb;
    const int c = q < nc ? q : q - nc;
    const int lo = c * ch, hi = lo + ch < nb ? lo + ch : nb;
    double s = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += 1024) s += src[i];
    s = wave_sum(s);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    double S = sm[0];
    for (int w = 1; w < 16; w++) S += sm[w];
    return S;
}

static __global__ __launch_bounds__(1024) void k_resid_chunks(const double* __restrict__ pa,
                                                             const double* __restrict__ pb,
                                                             int nb, int nc, int ch,
                                                             double* __restrict__ out,
                                                             const SorCtl* __restrict__ ctl) {
    if (ctl != nullptr && ctl->stop) return;  // (no control block: always)
    const double S = chunk_sum(pa, pb, nb, nc, ch, blockIdx.x);
    if (threadIdx.x == 0) out[blockIdx.x] = S;
}

// Final residual reduction of one SOR iteration and the reference's stopping
// test `while (R > gsTol && z < maxNoGSIteration)` (segmentation4D.c:309).
// `host` (device alias of page-locked host memory, or null): the control
// block is also written there and made visible to the host, so the host
// reads the decision without a copy in the stream.  A reduce that returns
// at once after a stop leaves it as the stopping reduce wrote it.
static __global__ __launch_bounds__(1024) void k_reduce_resid(
    const double* __restrict__ pa, int na, const double* __restrict__ pb, int nbp,
    SorCtl* ctl, double* __restrict__ hist, int z, int cap, int queued,
    SorCtl* host = nullptr) {
    // a reduce enqueued ahead of the host's decision (queued != 0) keeps a
    // stopped state: the passes after the stop left stale partials
    if (queued && ctl->stop) return;
    const double R = resid_sum(pa, na, pb, nbp);
    if (threadIdx.x == 0) {
        ctl->R = R;
        ctl->z = z;
        if (hist) hist[z - 1] = R;
        const bool cont = (R > ctl->tol) && (z < cap);
        if (!cont) {
            ctl->stop = 1;
            ctl->zfinal = z;
        }
        if (host) {
            *host = *ctl;
            __threadfence_system();
        }
    }
}

// ----------------------------------------------------------------------------
// Fused sweeps: every iteration of one red-black SOR solve in a single
// cooperative launch (all blocks co-resident), the passes separated by
// grid-wide barriers.  Opt-in: each barrier makes every block release and
// acquire its writes at agent scope (the L2 of each XCD written back and
// invalidated), which on an MI355X costs more than the launch gaps between
// single-pass launches it removes.
// ----------------------------------------------------------------------------
struct GridBar {
    // each field on a 256-byte line of its own: arrivals, polling and the
    // release do not contend for one line
    unsigned xcnt[8][64];  // blocks of XCD group x (blockIdx % 8) arrived
    unsigned count;        // groups whose last block arrived
    unsigned pad0[63];
    unsigned gen;          // barriers completed
    unsigned pad1[63];
    int abort;             // a block waited beyond kBarrierSpins: every block leaves
    int pad2[63];
};

// about a second of polling: far beyond any barrier of a healthy solve
constexpr long kBarrierSpins = 1L << 20;

static __global__ void k_bar_reset(GridBar* gb) {
    unsigned* w = reinterpret_cast<unsigned*>(gb);
    for (unsigned q = threadIdx.x; q < sizeof(GridBar) / sizeof(unsigned); q += blockDim.x)
        w[q] = 0u;
}

// Grid-wide barrier for a cooperative launch of gridDim.x blocks, in two
// levels: thread 0 of each block releases the block's writes (one agent-scope
// fence) and arrives on its group's counter (group = blockIdx % 8, the XCD
// the block runs on); the last block of a group arrives on the group counter,
// and the last group advances the generation every block waits for.  The
// counters and the generation are accessed with relaxed agent-scope atomics,
// the memory order coming from the fence before the arrival and one fence
// after the wait (an acquire on every poll would invalidate the XCD's L2 at
// each of them, under the blocks still working).  The wait is bounded: a
// timed-out wait raises the abort flag, which every waiting block sees, so
// that all blocks leave the kernel instead of hanging.  Returns false after
// an abort.  `gen` is the block's count of barriers passed.
__device__ __forceinline__ bool grid_barrier(GridBar* gb, unsigned& gen) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int good = 1;
        const unsigned ng = gridDim.x < 8 ? gridDim.x : 8u;
        const unsigned x = blockIdx.x % 8;
        const unsigned gsize = (gridDim.x - x + 7) / 8;  // blocks of group x
        __threadfence();  // release: this block's writes before its arrival
        bool last = false;
        const unsigned a =
            __hip_atomic_fetch_add(&gb->xcnt[x][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a == gsize - 1) {
            // every block of the group has arrived (and released its writes)
            __hip_atomic_store(&gb->xcnt[x][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            const unsigned b =
                __hip_atomic_fetch_add(&gb->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = b == ng - 1;
        }
        if (last) {
            __hip_atomic_store(&gb->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            __hip_atomic_store(&gb->gen, gen + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long spins = 0;
            while (__hip_atomic_load(&gb->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(&gb->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    ++spins > kBarrierSpins) {
                    __hip_atomic_store(&gb->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
        }
        __threadfence();  // acquire: the other blocks' writes after the wait
        gen++;
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// chunk_sum and resid_sum as a block of 1024 threads forms them, computed by
// a block of kBlock = 256: thread t plays threads t, t + 256, t + 512 and
// t + 768 (same elements, same order; lane 0 of wave w holds the wave sum of
// waves w, w + 4, w + 8, w + 12), so the sums are bitwise equal.  The result
// is valid in every thread.
__device__ __forceinline__ double sum1024_finish(const double (&s)[4]) {
    __shared__ double sm[16];
    __syncthreads();  // a previous call's reads of sm are done
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int r = 0; r < 4; r++) sm[(threadIdx.x >> 6) + 4 * r] = s[r];
    __syncthreads();
    double S = sm[0];
    for (int w = 1; w < 16; w++) S += sm[w];
    return S;
}

__device__ __attribute__((noinline)) double chunk_sum_by256(const double* __restrict__ pa,
                                                  const double* __restrict__ pb, int nb, int nc,
                                                  int ch, int q) {
    const double* src = q < nc ? pa : pb;
    const int c = q < nc ? q : q - nc;
    const int lo = c * ch, hi = lo + ch < nb ? lo + ch : nb;
    double s[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        double a = 0.0;
        for (int i = lo + (int)threadIdx.x + 256 * r; i < hi; i += 1024) a += src[i];
        s[r] = wave_sum(a);
    }
    return sum1024_finish(s);
}

__device__ __attribute__((noinline)) double resid_sum_by256(const double* __restrict__ pa, int na,
                                                  const double* __restrict__ pb, int nbp) {
    double s[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int t = (int)threadIdx.x + 256 * r;
        double a = 0.0;
        for (int q = t; q < na; q += 1024) a += pa[q];
        double a2 = 0.0;
        for (int q = t; q < nbp; q += 1024) a2 += pb[q];
        a = a + a2;
        s[r] = wave_sum(a);
    }
    return sum1024_finish(s);
}

// One colour pass of the fused solve: this block's work items.
template <typename T, int RESID, int U>
__device__ __forceinline__ void fused_pass(const Geo& G, const SweepMap& M, T c, T om, double h4,
                                           int colour, const T* xin, T* xout, const T* y,
                                           const T* opnd, double* partials) {
    const int nitems = sweep_grid(M);
    for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
        int bx, kb;
        if (map_block(M, b, nitems, bx, kb))
            sor_block<T, RESID, U, false>(G, M, bx, kb, c, om, h4, colour, xin, xout, y, opnd,
                                          partials, nullptr);
    }
}

// The whole solve: red(1), black(1), then per iteration z red(z+1) (the red
// residual terms of z), the chunk sums, the stopping test of z -- every
// block forms the final sum itself, in the fixed order, so all take the same
// decision -- and black(z+1).  red(1) runs the same residual-forming pass as
// the later red passes (its terms are overwritten by red(2) before they are
// summed; the update is the same), so that the kernel holds two pass bodies
// only: with a third, the compiler's allocation grew to 155 VGPRs (3 waves
// per SIMD) against 112-121.  Work items are the blocks of the single-pass
// launch (same map, same partial index), block b of this grid taking items
// b, b + gridDim.x, ... (gridDim.x a multiple of 8 unless the pass has fewer
// than 8 blocks: an item runs on the XCD its own launch would give it).  Hence u, every residual and the sweep
// count are bitwise those of the pass-per-launch solve.  ctl: tol set, stop
// clear on entry; on exit R, z, stop and zfinal as k_reduce_resid leaves
// them (gb->abort set instead if a barrier timed out).
template <typename T, int U>
static __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_sor_fused(
    Geo G, SweepMap M, T c, T om, double h4, T* red0, T* red1, T* black,
    const T* __restrict__ op0, const T* __restrict__ op1, double* part_r, double* part_b, int nb,
    int nc, int ch, double* red2, SorCtl* ctl, int cap, GridBar* gb) {
    unsigned gen = 0;
    const double tol = ctl->tol;
    for (int z = 0;; z++) {
        // red(z + 1) reads red[z & 1] and writes red[(z + 1) & 1]
        T* rin = (z & 1) ? red1 : red0;
        T* rout = (z & 1) ? red0 : red1;
        fused_pass<T, 1, U>(G, M, c, om, h4, 0, rin, rout, black, op0, part_r);
        if (!grid_barrier(gb, gen)) return;
        if (z > 0) {
            for (int q = blockIdx.x; q < 2 * nc; q += gridDim.x) {
                const double S = chunk_sum_by256(part_r, part_b, nb, nc, ch, q);
                if (threadIdx.x == 0) red2[q] = S;
            }
            if (!grid_barrier(gb, gen)) return;
            const double R = resid_sum_by256(red2, nc, red2 + nc, nc);
            const bool cont = (R > tol) && (z < cap);
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                ctl->R = R;
                ctl->z = z;
                if (!cont) {
                    ctl->stop = 1;
                    ctl->zfinal = z;
                }
            }
            if (!cont) return;
        }
        fused_pass<T, 2, U>(G, M, c, om, h4, 1, black, black, rout, op1, part_b);
        if (!grid_barrier(gb, gen)) return;
    }
}

// First level of a two-level fixed-order sum: block b of 1024 threads sums
// in[b * chunk .. min(n, (b + 1) * chunk)) into out[b].
static __global__ __launch_bounds__(1024) void k_sum_chunks(const double* __restrict__ in, long n,
                                                           int chunk, double* __restrict__ out) {
    const long q0 = (long)blockIdx.x * chunk;
    const long q1 = q0 + chunk < n ? q0 + chunk : n;
    double s = 0.0;
    for (long q = q0 + threadIdx.x; q < q1; q += 1024) s += in[q];
    s = wave_sum(s);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = sm[0];
        for (int q = 1; q < 16; q++) S += sm[q];
        out[blockIdx.x] = S;
    }
}

// Sets the per-solve tolerance: absolute gsTol, or rel^2 * sum(prev^2 h^4).
static __global__ __launch_bounds__(1024) void k_init_ctl(const double* __restrict__ partials,
                                                  int np, SorCtl* ctl, double gsTol,
                                                  double rel2) {
    double s = 0.0;
    for (int q = threadIdx.x; q < np; q += 1024) s += partials[q];
    s = wave_sum(s);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = sm[0];
        for (int q = 1; q < 16; q++) S += sm[q];
        ctl->sumsq = S;
        ctl->tol = rel2 > 0.0 ? rel2 * S : gsTol;
        ctl->R = 0.0;
        ctl->z = 0;
        ctl->stop = 0;
        ctl->zfinal = 0;
    }
}

// The FP64 residual of the set-up's fused defect: ctl->R = sum of np
// partials (after k_init_ctl).
static __global__ __launch_bounds__(1024) void k_set_r0(const double* __restrict__ partials,
                                                       int np, SorCtl* ctl) {
    double s = 0.0;
    for (int q = threadIdx.x; q < np; q += 1024) s += partials[q];
    s = wave_sum(s);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double R = sm[0];
        for (int q = 1; q < 16; q++) R += sm[q];
        ctl->R = R;
    }
}

// Generic fixed-order sum of `np` partials into *out (SQRT: its square root,
// correctly rounded like the host's sqrt).
template <bool SQRT = false>
static __global__ __launch_bounds__(1024) void k_reduce_sum(const double* __restrict__ partials,
                                                    int np, double* out) {
    double s = 0.0;
    for (int q = threadIdx.x; q < np; q += 1024) s += partials[q];
    s = wave_sum(s);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = sm[0];
        for (int q = 1; q < 16; q++) S += sm[q];
        *out = SQRT ? sqrt(S) : S;
    }
}

// Blocks of the _l2norm_ reduction over n elements: one partial per 4096
// elements (16 per thread, grid-stride), at most kL2MaxPartials; shared by
// the engine's and the allocation-free entry point so both sum in one order.
constexpr int kL2MaxPartials = 65536;
constexpr long kL2PerPartial = 16L * kBlock;
__host__ __device__ inline int l2_blocks(long n) {
    const long b = (n + kL2PerPartial - 1) / kL2PerPartial;
    return b < 1 ? 1 : (b > kL2MaxPartials ? kL2MaxPartials : (int)b);
}

// ----------------------------------------------------------------------------
// Colour layout -> full layout after the solve (interior points only), with
// the reference's Mmask pass (segmentation4D.c:1093-1101) fused in when
// there are no balls to rescale first.
// ----------------------------------------------------------------------------
// Every padded point: the ghosts come from the colour arrays too (the set-up
// wrote u^n's ghosts there, the MCF step its boundary values), because u's
// buffer is a recycled history level; the mask applies to the interior only.
__device__ __forceinline__ void unpack_point(const Geo& G, const double* __restrict__ ur,
                                             const double* __restrict__ ub,
                                             double* __restrict__ u,
                                             const double* __restrict__ mask, long q) {
    if (q >= (long)G.I * G.J * G.K * G.L) return;
    const int p = (int)q;
    int r = p / G.I;
    const int i = p - r * G.I;
    const int j = r % G.J;
    r /= G.J;
    const int k = r % G.K, l = r / G.K;
    const int pp = cidx(G, j, k, l, i >> 1);
    double v = ((i + j + k + l) & 1) ? ub[pp] : ur[pp];
    if (mask != nullptr && mask[p] == 0 && is_interior(G, i, j, k, l)) v = 0;
    u[p] = v;
}

static __global__ __launch_bounds__(kBlock) void k_unpack(Geo G, const double* __restrict__ ur,
                                                  const double* __restrict__ ub,
                                                  double* __restrict__ u,
                                                  const double* __restrict__ mask) {
    // one thread per padded point in array order (a 1-D grid of
    // cdiv(I J K L, kBlock) blocks), so that a wavefront's stores of u cover
    // whole lines whatever the row length
    unpack_point(G, ur, ub, u, mask, (long)blockIdx.x * kBlock + threadIdx.x);
}

// k_unpack behind a device gate: nothing is written unless gate[1] is set
// (the epilogue of a step captured as one graph, whose solve ends on the
// device).  sel: the FP64 solve's control block, whose final iteration's
// parity names the red buffer holding the accepted values (null: ur0, as
// after a correction).
static __global__ __launch_bounds__(kBlock) void k_unpack_gated(
    Geo G, const double* __restrict__ ur0, const double* __restrict__ ur1,
    const double* __restrict__ ub, double* __restrict__ u, const double* __restrict__ mask,
    const SorCtl* __restrict__ sel, const int* __restrict__ gate) {
    if (gate[1] == 0) return;
    const double* ur = (sel != nullptr && (sel->zfinal & 1)) ? ur1 : ur0;
    unpack_point(G, ur, ub, u, mask, (long)blockIdx.x * kBlock + threadIdx.x);
}

// ----------------------------------------------------------------------------
// Implicit heat step (presmoothing, segmentation4D.c:337-409) on the colour
// layout.  k_heat_pack copies the reflected input into the colour arrays
// (both red buffers, every padded point: the ghosts stay frozen) and its
// interior into hp, the right-hand side prev in the interior-plane order of
// the operand blocks: hp[colour][plane][t], t = (j-1) H + m.
// ----------------------------------------------------------------------------
__device__ __forceinline__ long heat_prev_index(const Geo& G, int c, int k, int l, int t) {
    const long nplane = (long)(G.K - 2) * (G.L - 2);
    return ((long)c * nplane + opnd_plane(G, k, l)) * G.RP + t;
}

static __global__ __launch_bounds__(kBlock) void k_heat_pack(Geo G, const double* __restrict__ d,
                                                     double* __restrict__ ur0,
                                                     double* __restrict__ ur1,
                                                     double* __restrict__ ub,
                                                     double* __restrict__ hp) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y, l = blockIdx.z;
    if (t >= G.I * G.J) return;
    const int j = t / G.I, i = t - j * G.I;
    const int c = (i + j + k + l) & 1;
    const int pp = cidx(G, j, k, l, i >> 1);
    const double v = d[((l * G.K + k) * G.J + j) * G.I + i];
    if (c == 0) {
        ur0[pp] = v;
        ur1[pp] = v;
    } else {
        ub[pp] = v;
    }
    if (is_interior(G, i, j, k, l)) hp[heat_prev_index(G, c, k, l, (j - 1) * G.H + (i >> 1))] = v;
}

// One colour pass of the heat step, one thread per packed column m of rows
// (j, k), marching l with the other colour at l-1, l, l+1 in registers (the
// launch map of the SOR passes).  Neighbour sum in the reference's order
// +x, -x, -y, +y, -z, +z, -t, +t (segmentation4D.c:362-369); residual
// u (1 + 8c) - c sum - prev (:392-401) before (RESID 1: red pass z+1 yields
// the red terms of iteration z) or after (RESID 2: black pass of z) the
// update, summed per block.
// Loads of one heat-pass point (issued for two l-steps before their arithmetic).
struct HeatIn {
    double yo, uS, uN, uB, uT, pv, uo;
    int pp, sft;
    bool act;
};

template <int RESID>
__device__ __forceinline__ void heat_update(const HeatIn& P, double wb, double wc, double wf,
                                            double c, double om, double h4, double* xout,
                                            double& acc) {
    if (!P.act) return;
    const double uW = P.sft ? wc : P.yo;
    const double uE = P.sft ? P.yo : wc;
    double nb = uE + uW;
    nb = nb + P.uS;
    nb = nb + P.uN;
    nb = nb + P.uB;
    nb = nb + P.uT;
    nb = nb + wb;
    nb = nb + wf;
    const double D = 1 + c * 8;
    if constexpr (RESID == 1) {
        const double r = P.uo * D - c * nb - P.pv;
        acc += r * r * h4;
    }
    const double gs = (P.pv + c * nb) / D;
    const double un = P.uo + om * (gs - P.uo);
    __builtin_nontemporal_store(un, xout + P.pp);
    if constexpr (RESID == 2) {
        const double r = un * D - c * nb - P.pv;
        acc += r * r * h4;
    }
}

template <int RESID>
static __global__ __launch_bounds__(kBlock) void k_heat_sweep(
    Geo G, SweepMap M, double c, double om, double h4, int colour, const double* xin,
    double* xout, const double* __restrict__ y, const double* __restrict__ hp,
    double* __restrict__ partials) {
    int bx, kb;
    if (!map_block(M, blockIdx.x, gridDim.x, bx, kb)) return;
    const int lb = (kb - 1) * M.gx + bx;
    const int kgrp = (kb - 1) % M.gk0, lch = (kb - 1) / M.gk0;
    const int l0 = 1 + lch * M.lc;
    const int per = kBlock / M.kp;
    const int sub = threadIdx.x / per;
    const int kk = kgrp * M.kp + sub + 1;
    const bool kok = kk <= G.K - 2;
    const int k = kok ? kk : G.K - 2;
    const int t = bx * per + (threadIdx.x - sub * per);
    const int jr = t / G.H;
    const int m = t - jr * G.H;
    const int j = jr + 1;
    const bool rowok = kok && j <= G.J - 2;
    const int base = cidx(G, j, k, 0, m);
    double acc = 0.0;
    // w[q]: the other colour at column m, frame l - 1 + q
    double w[4];
    w[0] = rowok ? y[base + (l0 - 1) * G.pw] : 0.0;
    w[1] = rowok ? y[base + l0 * G.pw] : 0.0;
    auto load = [&](int l, HeatIn& P) {
        const int pp = base + l * G.pw;
        P.pp = pp;
        P.sft = colour ^ ((j + k + l) & 1);
        const int i = 2 * m + P.sft;
        P.act = rowok && i >= 1 && i <= G.I - 2;
        if (P.act) {
            P.yo = y[pp + (P.sft ? 1 : -1)];
            P.uS = y[pp - G.py];
            P.uN = y[pp + G.py];
            P.uB = y[pp - G.pz];
            P.uT = y[pp + G.pz];
            P.pv = hp[heat_prev_index(G, colour, k, l, t)];
            P.uo = xin[pp];
        }
    };
    const int lmax = min(l0 + M.lc - 1, G.L - 2);
    int l = l0;
    for (; l + 1 <= lmax; l += 2) {
        HeatIn P0, P1;
        w[2] = rowok ? y[base + (l + 1) * G.pw] : 0.0;
        w[3] = rowok ? y[base + (l + 2) * G.pw] : 0.0;
        load(l, P0);
        load(l + 1, P1);
        heat_update<RESID>(P0, w[0], w[1], w[2], c, om, h4, xout, acc);
        heat_update<RESID>(P1, w[1], w[2], w[3], c, om, h4, xout, acc);
        w[0] = w[2];
        w[1] = w[3];
    }
    if (l <= lmax) {
        HeatIn P0;
        w[2] = rowok ? y[base + (l + 1) * G.pw] : 0.0;
        load(l, P0);
        heat_update<RESID>(P0, w[0], w[1], w[2], c, om, h4, xout, acc);
    }
    if constexpr (RESID != 0) {
        const double bs = block_sum256(acc);
        if (threadIdx.x == 0) partials[lb] = bs;
    }
}

// ----------------------------------------------------------------------------
// Ghost shell by clamped-index copy (closed form of the sequential X,Y,Z,W
// passes of segmentation4D.c:826-864).  Grid x over padded (j,i), y=k, z=l.
// ----------------------------------------------------------------------------
__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

static __global__ __launch_bounds__(kBlock) void k_reflect(Geo G, double* u) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y, l = blockIdx.z;
    if (t >= G.I * G.J) return;
    const int j = t / G.I, i = t - j * G.I;
    if (is_interior(G, i, j, k, l)) return;
    const int ii = clampi(i, 1, G.I - 2), jj = clampi(j, 1, G.J - 2);
    const int kk = clampi(k, 1, G.K - 2), ll = clampi(l, 1, G.L - 2);
    u[((l * G.K + k) * G.J + j) * G.I + i] = u[((ll * G.K + kk) * G.J + jj) * G.I + ii];
}

// ----------------------------------------------------------------------------
// Edge indicator
// ----------------------------------------------------------------------------
// The 27 points of the 33-point stencil that the four forward faces (+x, +y,
// +z, +t) use: the centre, p + e_A, p - e_B, p + e_A + e_B, p + e_A - e_B.
__device__ __forceinline__ void load_stencil_fwd(const double* __restrict__ u, int p,
                                                 const Geo& G, Stencil33& T) {
    const int st[4] = {1, G.sy, G.sz, G.sw};
#pragma unroll
    for (int w = 0; w < 3; w++)
#pragma unroll
        for (int z = 0; z < 3; z++)
#pragma unroll
            for (int y = 0; y < 3; y++)
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    const int o[4] = {x - 1, y - 1, z - 1, w - 1};
                    const int nz = (o[0] != 0) + (o[1] != 0) + (o[2] != 0) + (o[3] != 0);
                    const int neg = (o[0] < 0) + (o[1] < 0) + (o[2] < 0) + (o[3] < 0);
                    // at most two nonzero offsets, and never two negative ones
                    if (nz <= 2 && !(nz == 2 && neg == 2))
                        T.s[w][z][y][x] = u[p + o[0] * st[0] + o[1] * st[1] + o[2] * st[2] +
                                            o[3] * st[3]];
                }
}

// Edge indicator g on 8 faces (segmentation4D.c:582-823 after presmoothing):
// g_f = 1/(1 + K (theta |grad I0|_f^2 + rho |grad Ithr|_f^2)).
//
// A thread forms only the four forward faces of its point from registers
// (27-point stencil) and writes each twice: as face +A of p and as face -A
// of p + e_A, whose gradient components are the same sums with the addends
// swapped or negated (squares and commuted additions are bitwise equal), so
// g_w(p + x) = g_e(p) etc. bit for bit.  Points on the low boundary of an
// axis form their backward face themselves; ghost points get no g (the
// reference loops over the interior only).  BOTH = false drops the second
// image: when rho == 0 and that image is finite, rho * vb is a signed zero
// and theta * va + (+-0) gives g bit for bit (1 + K * (+-0) = 1).
// Blocks are mapped to XCD-resident tiles (PlaneMap) over the interior
// (run, k - 1, l - 1), so the neighbour rows and planes a tile re-reads stay
// in one L2.
template <bool POW2, bool BOTH>
static __global__ __launch_bounds__(kBlock) void k_gfun(Geo G, PlaneMap PM, Scale S, double K_,
                                                double theta, double rho,
                                                const double* __restrict__ a,
                                                const double* __restrict__ b, Ptr8 g) {
    int bx, kk, ll;
    if (!map_plane_block(PM, blockIdx.x, bx, kk, ll)) return;
    // lanes walk whole padded rows (lanes on ghost columns and rows idle), so
    // that a wavefront's stores cover whole 128-byte lines of the g arrays
    // (3.47 -> 3.14 ms against lanes on the interior points only, 256x256x128x16)
    const int t = bx * kBlock + threadIdx.x;
    if (t >= G.I * G.J) return;
    const int j = t / G.I, i = t - j * G.I;
    if (i < 1 || i > G.I - 2 || j < 1 || j > G.J - 2) return;
    const int k = kk + 1, l = ll + 1;
    const int p = ((l * G.K + k) * G.J + j) * G.I + i;
    const int st[4] = {1, G.sy, G.sz, G.sw};
    const int co[4] = {i, j, k, l};
    const int hi[4] = {G.I - 2, G.J - 2, G.K - 2, G.L - 2};
    double va[4], vb[4];
    {
        Stencil33 T;
        load_stencil_fwd(a, p, G, T);
        va[0] = face_sq_reg<0, POW2>(T, S);
        va[1] = face_sq_reg<2, POW2>(T, S);
        va[2] = face_sq_reg<4, POW2>(T, S);
        va[3] = face_sq_reg<6, POW2>(T, S);
    }
    if constexpr (BOTH) {
        Stencil33 T;
        load_stencil_fwd(b, p, G, T);
        vb[0] = face_sq_reg<0, POW2>(T, S);
        vb[1] = face_sq_reg<2, POW2>(T, S);
        vb[2] = face_sq_reg<4, POW2>(T, S);
        vb[3] = face_sq_reg<6, POW2>(T, S);
    }
    // streaming stores (the 64 bytes of g per point are the bulk of the
    // traffic; kept out of L2 they leave it to the stencil's re-reads:
    // 4.38 -> 3.48 ms at 256x256x128x16)
#pragma unroll
    for (int A = 0; A < 4; A++) {
        double gv;
        if constexpr (BOTH) gv = 1.0 / (1 + K_ * (theta * va[A] + rho * vb[A]));
        else gv = 1.0 / (1 + K_ * (theta * va[A]));
        __builtin_nontemporal_store(gv, g.p[2 * A] + p);
        if (co[A] < hi[A]) __builtin_nontemporal_store(gv, g.p[2 * A + 1] + p + st[A]);
    }
    // backward faces of the low boundary (their west-type neighbour is a ghost)
#pragma unroll
    for (int A = 0; A < 4; A++) {
        if (co[A] != 1) continue;
        double wa, wb = 0.0;
        if (A == 0) wa = face_sq<1, POW2>(a, p, G, S);
        else if (A == 1) wa = face_sq<3, POW2>(a, p, G, S);
        else if (A == 2) wa = face_sq<5, POW2>(a, p, G, S);
        else wa = face_sq<7, POW2>(a, p, G, S);
        if constexpr (BOTH) {
            if (A == 0) wb = face_sq<1, POW2>(b, p, G, S);
            else if (A == 1) wb = face_sq<3, POW2>(b, p, G, S);
            else if (A == 2) wb = face_sq<5, POW2>(b, p, G, S);
            else wb = face_sq<7, POW2>(b, p, G, S);
            __builtin_nontemporal_store(1.0 / (1 + K_ * (theta * wa + rho * wb)),
                                        g.p[2 * A + 1] + p);
        } else {
            __builtin_nontemporal_store(1.0 / (1 + K_ * (theta * wa)), g.p[2 * A + 1] + p);
        }
    }
}

// ----------------------------------------------------------------------------
// Flat helpers over the full padded array
// ----------------------------------------------------------------------------
// Partial sums of (a-b)^2 h^4 (segmentation4D.c:194-213); grid-stride with a
// fixed grid so that the summation order is deterministic.
static __global__ __launch_bounds__(kBlock) void k_sqdiff(const double* __restrict__ a,
                                                  const double* __restrict__ b, long n,
                                                  double h4, double* __restrict__ partials) {
    double acc = 0.0;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < n;
         q += (long)gridDim.x * kBlock) {
        const double d = a[q] - b[q];
        acc += d * d * h4;
    }
    const double bs = block_sum256(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// MCF exact solution ((x^2+y^2+z^2+w^2-1)/6)+t at cell centres
// x = ((i-1) h + h/2) - 1.25 (mcf_4D.c:444-462).  With `ghosts_only` the
// exact-value Dirichlet boundary of mcf_4D.c:344-443 is written instead; its
// eight sequential face passes leave on every ghost the value of the last
// pass that touches it (priority w > z > y > x, high side written after low
// side), and that pass uses the fixed coordinate -(h/2)-1.25 or (h/2)+1.25.
static __global__ __launch_bounds__(kBlock) void k_mcf_exact(Geo G, double h, double tt,
                                                     double* __restrict__ out,
                                                     int ghosts_only) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y, l = blockIdx.z;
    if (t >= G.I * G.J) return;
    const int j = t / G.I, i = t - j * G.I;
    double x = ((i - 1) * h + (h / 2.)) - 1.25;
    double y = ((j - 1) * h + (h / 2.)) - 1.25;
    double z = ((k - 1) * h + (h / 2.)) - 1.25;
    double w = ((l - 1) * h + (h / 2.)) - 1.25;
    if (ghosts_only) {
        if (is_interior(G, i, j, k, l)) return;
        const double lo = -(h / 2.) - 1.25, hi = (h / 2.) + 1.25;
        if (l == 0 || l == G.L - 1) w = (l == 0) ? lo : hi;
        else if (k == 0 || k == G.K - 1) z = (k == 0) ? lo : hi;
        else if (j == 0 || j == G.J - 1) y = (j == 0) ? lo : hi;
        else x = (i == 0) ? lo : hi;
    }
    double v = x * x + y * y;
    v = v + z * z;
    v = v + w * w;
    out[((l * G.K + k) * G.J + j) * G.I + i] = ((v - 1.0) / 6.0) + tt;
}

// Re-pack the ghost shell of the full array into the colour arrays (both red
// buffers).  Used after the MCF boundary fill.
static __global__ __launch_bounds__(kBlock) void k_pack_ghosts(Geo G, const double* __restrict__ u,
                                                       double* ur0, double* ur1, double* ub) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y, l = blockIdx.z;
    if (t >= G.I * G.J) return;
    const int j = t / G.I, i = t - j * G.I;
    if (is_interior(G, i, j, k, l)) return;
    const int p = ((l * G.K + k) * G.J + j) * G.I + i;
    const int pp = cidx(G, j, k, l, i >> 1);
    const double v = u[p];
    if ((i + j + k + l) & 1) {
        ub[pp] = v;
    } else {
        ur0[pp] = v;
        ur1[pp] = v;
    }
}

// Colour layout -> full layout for one coefficient face (interior only),
// used when the host asks for the coefficient arrays.
// Component f (0..7 coefficient, 8 prev) of the operand blocks -> full layout.
// With split operands (hr, hb non-null: the FP32 blocks with the high words)
// the value is put together from its high and low words.
static __global__ __launch_bounds__(kBlock) void k_unpack_coef(Geo G, const double* __restrict__ orr,
                                                       const double* __restrict__ obl, int f,
                                                       double* __restrict__ out,
                                                       const float* __restrict__ hr = nullptr,
                                                       const float* __restrict__ hb = nullptr) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y, l = blockIdx.z;
    if (t >= G.I * G.J) return;
    const int j = t / G.I, i = t - j * G.I;
    const int p = ((l * G.K + k) * G.J + j) * G.I + i;
    double v = 0.0;
    if (is_interior(G, i, j, k, l)) {
        const int t2 = (j - 1) * G.H + (i >> 1);
        const bool blk = (i + j + k + l) & 1;
        if (hr != nullptr) {
            const unsigned* ow = reinterpret_cast<const unsigned*>(blk ? obl : orr) +
                                 split_index(G, k, l, t2);
            if (f < 8) {
                const float* h = (blk ? hb : hr) + opnd_index(G, k, l, t2) + f * 64;
                v = __hiloint2double(__float_as_int(*h), (int)ow[f * 64]);
            } else {
                v = __hiloint2double((int)ow[8 * 64], (int)ow[9 * 64]);
            }
        } else {
            const long q = opnd_index(G, k, l, t2) + f * 64;
            v = blk ? obl[q] : orr[q];
        }
    }
    out[p] = v;
}

// ----------------------------------------------------------------------------
// Ball kernels of the segmentation's initialisation and epilogue.  A frame's
// balls are processed by one block in the reference's order c = 0..count-1
// (segmentation4D.c:984, :1037, :1117): later balls see the earlier balls'
// writes, and no two blocks write the same frame.  Ball boxes follow
// segmentation4D.c:907-912 -- (int) truncation of centre -/+ (R + 0.5) --
// clamped to the padded array (the reference leaves them unchecked).
// ----------------------------------------------------------------------------
struct BallList {                // centres of frame l: c[offset[l] + q], q < count[l]
    const int* count;
    const int* offset;
    const double* cx;
    const double* cy;
    const double* cz;
    double R;
};

struct BallBox {
    int i0, j0, k0, ni, nj, nk, tot;
    __device__ __forceinline__ void point(int q, int& i, int& j, int& k) const {
        i = i0 + q % ni;
        j = j0 + (q / ni) % nj;
        k = k0 + q / (ni * nj);
    }
};

__device__ __forceinline__ BallBox ball_box(const Geo& G, double x0, double y0, double z0,
                                            double R) {
    // lower ends raised to 0, upper ends lowered to the last index; a box
    // entirely outside the array comes out empty
    BallBox b;
    b.k0 = max((int)(z0 - R - 0.5), 0);
    b.j0 = max((int)(y0 - R - 0.5), 0);
    b.i0 = max((int)(x0 - R - 0.5), 0);
    b.nk = min((int)(z0 + R + 0.5), G.K - 1) - b.k0 + 1;
    b.nj = min((int)(y0 + R + 0.5), G.J - 1) - b.j0 + 1;
    b.ni = min((int)(x0 + R + 0.5), G.I - 1) - b.i0 + 1;
    b.tot = (b.ni > 0 && b.nj > 0 && b.nk > 0) ? b.ni * b.nj * b.nk : 0;
    return b;
}

__device__ __forceinline__ double ball_dist(int i, int j, int k, double x0, double y0,
                                            double z0) {
    const double dx = i - x0, dy = j - y0, dz = k - z0;
    return sqrt((dx * dx) + (dy * dy) + (dz * dz));
}

// Block-wide min (from 100000) and max (from 0) of frame l of `a` over the
// points of box `bb` within distance R of the centre (segmentation4D.c:
// 897-969).  Every thread receives both; the block is synchronised on exit,
// so writes that follow cannot overtake a read of this reduction.
__device__ void ball_minmax_block(const Geo& G, const double* a, int l, const BallBox& bb,
                                  double x0, double y0, double z0, double R, double* mn_out,
                                  double* mx_out) {
    __shared__ double smn[kBlock / 64], smx[kBlock / 64];
    double mn = 100000, mx = 0;
    for (int q = threadIdx.x; q < bb.tot; q += kBlock) {
        int i, j, k;
        bb.point(q, i, j, k);
        if (ball_dist(i, j, k, x0, y0, z0) <= R) {
            const double v = a[((l * G.K + k) * G.J + j) * G.I + i];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, off, 64));
        mx = fmax(mx, __shfl_xor(mx, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        smn[threadIdx.x >> 6] = mn;
        smx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    mn = smn[0];
    mx = smx[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; w++) {
        mn = fmin(mn, smn[w]);
        mx = fmax(mx, smx[w]);
    }
    *mn_out = mn;
    *mx_out = mx;
    __syncthreads();
}

// Per-ball rescale to [0, 1] (segmentation4D.c:1020-1092): ball 0 of a frame
// overwrites, later balls keep the larger value.
static __global__ __launch_bounds__(kBlock) void k_rescale_balls(Geo G, double* u, BallList B) {
    const int l = blockIdx.x + 1;
    for (int c = 0; c < B.count[l]; c++) {
        const int qc = B.offset[l] + c;
        const double x0 = B.cx[qc], y0 = B.cy[qc], z0 = B.cz[qc];
        const BallBox bb = ball_box(G, x0, y0, z0, B.R);
        double mn, mx;
        ball_minmax_block(G, u, l, bb, x0, y0, z0, B.R, &mn, &mx);
        const double quot = 1. / (mx - mn);
        const double offs = mn * quot;
        for (int q = threadIdx.x; q < bb.tot; q += kBlock) {
            int i, j, k;
            bb.point(q, i, j, k);
            if (ball_dist(i, j, k, x0, y0, z0) <= B.R) {
                const int p = ((l * G.K + k) * G.J + j) * G.I + i;
                const double nv = (quot * u[p]) - offs;
                if (c == 0) u[p] = nv;
                else if (u[p] < nv) u[p] = nv;
            }
        }
        __syncthreads();
    }
}

// Local thresholding (segmentation4D.c:971-1018): inside each ball, points of
// `in` brighter than alpha*min + beta*max of that ball are set to the
// image's interior maximum *gmax in `thr` (which may alias `in`).
static __global__ __launch_bounds__(kBlock) void k_threshold_balls(Geo G, const double* in,
                                                           double* thr, BallList B,
                                                           double alpha, double beta,
                                                           const double* gmax_p) {
    const int l = blockIdx.x + 1;
    const double gmax = *gmax_p;
    for (int c = 0; c < B.count[l]; c++) {
        const int qc = B.offset[l] + c;
        const double x0 = B.cx[qc], y0 = B.cy[qc], z0 = B.cz[qc];
        const BallBox bb = ball_box(G, x0, y0, z0, B.R);
        double mn, mx;
        ball_minmax_block(G, in, l, bb, x0, y0, z0, B.R, &mn, &mx);
        const double tv = alpha * mn + beta * mx;
        for (int q = threadIdx.x; q < bb.tot; q += kBlock) {
            int i, j, k;
            bb.point(q, i, j, k);
            const int p = ((l * G.K + k) * G.J + j) * G.I + i;
            if (ball_dist(i, j, k, x0, y0, z0) <= B.R && in[p] > tv) thr[p] = gmax;
        }
        __syncthreads();
    }
}

// Seed of the segmentation function (segmentation4D.c:1106-1148): within the
// ball box and distance d <= 12, u = max(u, |1/(d + 1/2) - 1/(12 + 1/2)|)
// and Mmask = 1.
static __global__ __launch_bounds__(kBlock) void k_seed_balls(Geo G, double* u, double* mask,
                                                      BallList B) {
    const int l = blockIdx.x + 1;
    const double v = 0.5, Rs = 12.;
    for (int c = 0; c < B.count[l]; c++) {
        const int qc = B.offset[l] + c;
        const double x0 = B.cx[qc], y0 = B.cy[qc], z0 = B.cz[qc];
        const BallBox bb = ball_box(G, x0, y0, z0, B.R);
        for (int q = threadIdx.x; q < bb.tot; q += kBlock) {
            int i, j, k;
            bb.point(q, i, j, k);
            const double d = ball_dist(i, j, k, x0, y0, z0);
            if (d <= Rs) {
                const int p = ((l * G.K + k) * G.J + j) * G.I + i;
                const double nv = fabs((1. / (d + v)) - (1. / (Rs + v)));
                if (u[p] < nv) u[p] = nv;
                mask[p] = 1.;
            }
        }
        __syncthreads();
    }
}

// One ball's min and max (segmentation4D.c:897-969) into out[0], out[1].
static __global__ __launch_bounds__(kBlock) void k_ball_minmax(Geo G, const double* a, int l,
                                                       double x0, double y0, double z0,
                                                       double R, double* out) {
    const BallBox bb = ball_box(G, x0, y0, z0, R);
    double mn, mx;
    ball_minmax_block(G, a, l, bb, x0, y0, z0, R, &mn, &mx);
    if (threadIdx.x == 0) {
        out[0] = mn;
        out[1] = mx;
    }
}

// Interior maximum from 0 (segmentation4D.c:876-895), per-block partials;
// the order of a max does not change its value.
static __global__ __launch_bounds__(kBlock) void k_max_interior(Geo G, const double* __restrict__ a,
                                                        double* __restrict__ partials) {
    __shared__ double sm[kBlock / 64];
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y + 1, l = blockIdx.z + 1;
    const int w = G.I - 2;
    double mx = 0;
    if (t < w * (G.J - 2)) {
        const int j = t / w + 1, i = t - (j - 1) * w + 1;
        const double v = a[((l * G.K + k) * G.J + j) * G.I + i];
        if (v > mx) mx = v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sm[0];
        for (int q = 1; q < kBlock / 64; q++) r = fmax(r, sm[q]);
        partials[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = r;
    }
}

static __global__ __launch_bounds__(1024) void k_reduce_max(const double* __restrict__ partials,
                                                    int np, double* out) {
    __shared__ double sm[16];
    double m = 0;
    for (int q = threadIdx.x; q < np; q += 1024) m = fmax(m, partials[q]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sm[0];
        for (int q = 1; q < 16; q++) r = fmax(r, sm[q]);
        *out = r;
    }
}

// Mmask pass alone (segmentation4D.c:1093-1101).
static __global__ __launch_bounds__(kBlock) void k_apply_mask(Geo G, double* __restrict__ u,
                                                      const double* __restrict__ mask) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y + 1, l = blockIdx.z + 1;
    const int w = G.I - 2;
    if (t >= w * (G.J - 2)) return;
    const int j = t / w + 1, i = t - (j - 1) * w + 1;
    const int p = ((l * G.K + k) * G.J + j) * G.I + i;
    if (mask[p] == 0) u[p] = 0;
}

}  // namespace s4d
