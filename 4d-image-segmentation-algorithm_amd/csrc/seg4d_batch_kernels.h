// seg4d_batch_kernels.h -- kernels of the batched form: many volumes of
// different sizes in one launch per colour pass.
//
// Each volume keeps its own buffers, geometry, launch order and control
// block; a descriptor per volume (BatchJob) holds them in device memory.
// Block b of a batched launch belongs to the volume whose block range
// [block0, block0 + nblocks) contains it, and does exactly the work block
// b - block0 does in that volume's own launch (same kernel body, same
// partial index, same fixed-order reduction), so every volume's result is
// bitwise that of its single-volume solve.  Volume ranges start at multiples
// of 8, so a block's XCD (b mod 8) is the one the volume's own launch order
// assumes, and every group of 8 consecutive blocks belongs to one volume: a
// table of one entry per group (vmap[b / 8]) names it with a single load
// instead of a search over the ranges.
#pragma once
#include "seg4d_kernels.h"

namespace s4d {

template <typename T>
struct BatchJob {
    Geo G;
    SweepMap M;
    int block0, nblocks;    // SOR launch: this volume's blocks
    int cblock0, cnblocks;  // element-wise launches: this volume's blocks
    T c, om;
    double h4;
    const T* opnd[2];       // operand blocks per colour
    T* red[2];              // red ping-pong
    T* black;
    double* part_r;         // block partials per colour
    double* part_b;
    int nb;                 // partials per colour (gx * gk)
    long np;                // colour-array elements
    SorCtl* ctl;            // control block of this job's solve
    float* defect_ob[2];    // FP64 jobs: FP32 operand blocks (defect, copies)
    double* red2;           // chunk sums of the two-level residual reduction
    int nc, ch;             // nc chunks of ch partials per colour
    int r1block0;           // this volume's first block of the chunk-sum launch
};

// Pointers read from a descriptor in memory are generic to the compiler, and
// generic accesses become FLAT instructions (aperture checks, and every
// access counted on both the vector-memory and the LDS/scalar counters, which
// serialises the waits).  The buffers are device memory: casting through the
// global address space lets the compiler emit global loads and stores, as for
// pointers passed as kernel arguments.
template <typename P>
__device__ __forceinline__ P* glob(P* p) {
    // through an integer: a generic -> global -> generic cast pair folds away
    return (P*)(__attribute__((address_space(1))) P*)(uintptr_t)p;
}

template <typename T>
__device__ __forceinline__ int batch_cvolume(const BatchJob<T>* __restrict__ J, int nvol, int b) {
    int v = 0;
    while (v + 1 < nvol && b >= J[v + 1].cblock0) v++;
    return v;
}

// One SOR colour pass of every running volume.  colour 0: red, reads
// red[zr], writes red[zw]; colour 1: black in place, its red neighbours in
// red[zr].  Volumes whose control block says stop return at once.  ctl is
// the control-block array (volume v's is ctl[v], the one J[v].ctl names), so
// that the stop flag and the descriptor load together once v is known.
template <typename T, int RESID, int U>
static __global__ __launch_bounds__(kBlock) void k_sor_batch(const BatchJob<T>* __restrict__ J,
                                                      const int* __restrict__ vmap,
                                                      const SorCtl* __restrict__ ctl, int colour,
                                                      int zr, int zw) {
    const int v = vmap[blockIdx.x >> 3];
    const int stop = ctl[v].stop;
    const BatchJob<T>& B = J[v];
    const int lb = blockIdx.x - B.block0;
    if ((stop != 0) | (lb >= B.nblocks)) return;
    int bx, kb;
    if (!map_block(B.M, lb, B.nblocks, bx, kb)) return;
    if (colour == 0)
        sor_block<T, RESID, U, false>(B.G, B.M, bx, kb, B.c, B.om, B.h4, 0, glob(B.red[zr]),
                                      glob(B.red[zw]), glob(B.black), glob(B.opnd[0]),
                                      glob(B.part_r), nullptr);
    else
        sor_block<T, RESID, U, false>(B.G, B.M, bx, kb, B.c, B.om, B.h4, 1, glob(B.black),
                                      glob(B.black), glob(B.red[zr]), glob(B.opnd[1]),
                                      glob(B.part_b), nullptr);
}

// FP64 defect of the corrected state u + d of every running volume (the
// correction update fused into the defect, as in the single-volume form):
// red stores u_r + d_r and reads black neighbours u_b + d_b; black stores
// u_b + d_b with its red neighbours already corrected.
static __global__ __launch_bounds__(kBlock) void k_defect_corr_batch(
    const BatchJob<double>* __restrict__ J64, const BatchJob<float>* __restrict__ J32,
    const int* __restrict__ vmap, const SorCtl* __restrict__ ctl, int colour) {
    const int v = vmap[blockIdx.x >> 3];
    const int stop = ctl[v].stop;
    const BatchJob<double>& B = J64[v];
    const int lb = blockIdx.x - B.block0;
    if ((stop != 0) | (lb >= B.nblocks)) return;
    int bx, kb;
    if (!map_block(B.M, lb, B.nblocks, bx, kb)) return;
    const BatchJob<float>& D = J32[v];
    // split operands: the batch's set-up is the fused FP32-refined one
    if (colour == 0)
        sor_block<double, 6, 2, false>(B.G, B.M, bx, kb, B.c, B.om, B.h4, 0, glob(B.red[0]),
                                       glob(B.red[0]), glob(B.black), glob(B.opnd[0]),
                                       glob(B.part_r), glob(B.defect_ob[0]),
                                       glob(D.red[D.ctl->zfinal & 1]), glob(D.black));
    else
        sor_block<double, 6, 2, false>(B.G, B.M, bx, kb, B.c, B.om, B.h4, 1, glob(B.black),
                                       glob(B.black), glob(B.red[0]), glob(B.opnd[1]),
                                       glob(B.part_b), glob(B.defect_ob[1]), glob(D.black),
                                       nullptr);
}

// First level of the residual reduction of every running volume: the chunk
// sums of k_resid_chunks (same chunks, same order).
template <typename T>
static __global__ __launch_bounds__(1024) void k_resid_chunks_batch(const BatchJob<T>* __restrict__ J,
                                                                   int nvol) {
    int v = 0;
    while (v + 1 < nvol && (int)blockIdx.x >= J[v + 1].r1block0) v++;
    const BatchJob<T>& B = J[v];
    if (B.ctl->stop) return;
    const int q = blockIdx.x - B.r1block0;
    const double S = chunk_sum(glob(B.part_r), glob(B.part_b), B.nb, B.nc, B.ch, q);
    if (threadIdx.x == 0) glob(B.red2)[q] = S;
}

// Residual reduction and stopping test of every running volume (one block
// per volume, over its chunk sums); the iteration index is the control
// block's own counter.
template <typename T>
static __global__ __launch_bounds__(1024) void k_reduce_batch(const BatchJob<T>* __restrict__ J, int cap) {
    const BatchJob<T>& B = J[blockIdx.x];
    SorCtl* ctl = B.ctl;
    if (ctl->stop) return;
    const double R = resid_sum(B.red2, B.nc, B.red2 + B.nc, B.nc);
    if (threadIdx.x == 0) {
        const int z = ctl->z + 1;
        ctl->R = R;
        ctl->z = z;
        if (!((R > ctl->tol) && (z < cap))) {
            ctl->stop = 1;
            ctl->zfinal = z;
        }
    }
}

// FP64 residual of the FP32-refined form (over the chunk sums of
// k_resid_chunks_batch, the single-volume defect's two levels) and its
// decision: a volume stops once R <= tol (or R is not a number) or after
// max_rounds corrections; the control block's z counts the corrections.
static __global__ __launch_bounds__(1024) void k_defect_decide(const BatchJob<double>* __restrict__ J,
                                                        int max_rounds) {
    const BatchJob<double>& B = J[blockIdx.x];
    SorCtl* ctl = B.ctl;
    if (ctl->stop) return;
    const double R = resid_sum(B.red2, B.nc, B.red2 + B.nc, B.nc);
    if (threadIdx.x == 0) {
        ctl->R = R;
        if (!(R > ctl->tol) || ctl->z >= max_rounds) {
            ctl->stop = 1;
            ctl->zfinal = ctl->z;
        }
    }
}

// Control blocks of the correction solves: bound max(frac tol, floor2 R) of
// the volume's FP64 system; volumes that stopped skip the round.
static __global__ void k_inner_ctl_batch(const BatchJob<double>* __restrict__ J64,
                                  const BatchJob<float>* __restrict__ J32, int nvol, double frac,
                                  double floor2) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvol) return;
    SorCtl* outer = J64[v].ctl;
    SorCtl* inner = J32[v].ctl;
    const double a = frac * outer->tol, b2 = floor2 * outer->R;
    inner->tol = a > b2 ? a : b2;
    inner->R = 0.0;
    inner->z = 0;
    inner->zfinal = 0;
    inner->stop = outer->stop;
    if (!outer->stop) outer->z = outer->z + 1;
}

// The first decision of the FP32-refined form from the set-up's fused
// defect (R already in the control block): stop once R <= tol.
static __global__ void k_decide_batch(const BatchJob<double>* __restrict__ J64, int nvol) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvol) return;
    SorCtl* ctl = J64[v].ctl;
    if (!(ctl->R > ctl->tol)) {
        ctl->stop = 1;
        ctl->zfinal = 0;
    }
}

// d = 0 (ghosts included; the second red buffer is written before it is
// read) for running volumes from their second round on: the first round
// starts from the set-up's d0 = start - u^n.
static __global__ __launch_bounds__(kBlock) void k_zero_batch(const BatchJob<double>* __restrict__ J64,
                                                       const BatchJob<float>* __restrict__ J32,
                                                       int nvol) {
    const int v = batch_cvolume(J32, nvol, blockIdx.x);
    if (J64[v].ctl->stop || J64[v].ctl->z == 0) return;
    const BatchJob<float>& B = J32[v];
    const long stride = (long)B.cnblocks * kBlock;
    float* dr = glob(B.red[0]);
    float* db = glob(B.black);
    for (long q = (long)(blockIdx.x - B.cblock0) * kBlock + threadIdx.x; q < B.np; q += stride) {
        dr[q] = 0.0f;
        db[q] = 0.0f;
    }
}

// ---- the captured time step (graph mode 2) ---------------------------------
// A step captured as one graph holds a fixed number of iterations per solve
// and of correction rounds.  gate[0] is set when a volume's solve needs more
// than the graph holds; gate[1] = 1 when every volume finished inside it, and
// only then does the epilogue (k_unpack_gated) write u.  gate[2 + v]
// accumulates volume v's FP32 sweeps over the captured rounds.

static __global__ void k_gate_begin(int* gate, int nvol) {
    for (int q = (int)threadIdx.x; q < nvol + 2; q += (int)blockDim.x) gate[q] = 0;
}

// After the captured FP32 iterations of a round: a running volume whose
// correction solve has not stopped overran.  It is marked stopped, so that
// the rest of the graph skips it (its defect, later rounds), and the overrun
// is flagged; the step is then redone eagerly.  Volumes that finished add
// their sweeps (0 for volumes that stopped before the round).
static __global__ void k_gate_inner(const BatchJob<double>* __restrict__ J64,
                                    const BatchJob<float>* __restrict__ J32, int nvol, int* gate) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvol) return;
    SorCtl* outer = J64[v].ctl;
    const SorCtl* inner = J32[v].ctl;
    if (!outer->stop && !inner->stop) {
        outer->stop = 1;
        gate[0] = 1;
    } else {
        gate[2 + v] += inner->zfinal;
    }
}

// gate[1] = 1 iff no volume overran and every volume's solve stopped.
static __global__ void k_gate_done(const BatchJob<double>* __restrict__ J64, int nvol, int* gate) {
    __shared__ int any_running;
    if (threadIdx.x == 0) any_running = 0;
    __syncthreads();
    for (int v = (int)threadIdx.x; v < nvol; v += (int)blockDim.x)
        if (!J64[v].ctl->stop) any_running = 1;
    __syncthreads();
    if (threadIdx.x == 0) gate[1] = (gate[0] == 0 && any_running == 0) ? 1 : 0;
}

}  // namespace s4d
