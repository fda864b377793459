// seg4d_engine.hip -- host driver of the HIP kernels for one 4D volume.
#include "seg4d_engine.h"

#include <cmath>
#include <cstdio>
#include <cstring>

#include "seg4d_kernels.h"
#include "seg4d_net.h"

namespace s4d {

namespace {
thread_local std::string g_last_error;

bool is_pow2(double h) {
    if (!(h > 0.0) || !std::isfinite(h)) return false;
    int e = 0;
    return std::frexp(h, &e) == 0.5;
}

Scale make_scale(double h) {
    Scale S;
    S.h = h;
    S.q = 4.0 * h;
    S.inv_h = 1.0 / h;
    S.inv_q = 1.0 / S.q;
    return S;
}

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Block map of the full-layout stencil kernels (set-up, edge indicator): gx
// runs of kBlock points per (j,i) plane, gk planes, gl frames, in tiles of 32
// runs x 8 planes x 1 frame per XCD residency round (4% faster than 8 x 8 x 4
// for the set-up at 256x256x128x16); planes of fewer than 32 runs keep
// 8 x 8 x 4, which leaves fewer idle blocks.
PlaneMap plane_map(int gx, int gk, int gl) {
    PlaneMap PM;
    PM.gx = gx;
    PM.gk = gk;
    PM.gl = gl;
    const bool wide = PM.gx >= 32;
    PM.tx = wide ? 32 : 8;
    PM.tk = 8;
    PM.tl = wide ? 1 : 4;
    PM.ntx = (PM.gx + PM.tx - 1) / PM.tx;
    PM.ntk = (PM.gk + PM.tk - 1) / PM.tk;
    PM.ntiles = PM.ntx * PM.ntk * ((PM.gl + PM.tl - 1) / PM.tl);
    return PM;
}

unsigned plane_grid(const PlaneMap& PM) {
    return (unsigned)(((PM.ntiles + 7) / 8) * 8 * PM.tx * PM.tk * PM.tl);
}
}  // namespace

// Chunking of the two-level residual reduction over nb partials per colour:
// nc chunks of ch partials (2 nc <= 4096 chunk sums).
void resid_chunking(int nb, int* nc, int* ch) {
    int n = (int)cdiv(nb, 4096);
    if (n < 1) n = 1;
    if (n > 2048) n = 2048;
    *nc = n;
    *ch = (int)cdiv(nb, n);
}

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

Engine::~Engine() { release(); }

int Engine::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    set_error(err);
    return -1;
}

int Engine::launch_check(const char* what) { return check(hipGetLastError(), what); }

// Device alias of a page-locked host control block (null if the runtime
// gives none: the block is then copied in the stream).
static SorCtl* host_alias(SorCtl* h) {
    void* d = nullptr;
    if (!h || hipHostGetDevicePointer(&d, h, 0) != hipSuccess || !d) {
        (void)hipGetLastError();
        return nullptr;
    }
    return static_cast<SorCtl*>(d);
}

int Engine::init(int I, int J, int K, int L, bool seg, bool mcf) {
    release();
    if (I < 3 || J < 3 || K < 3 || L < 3) {
        err = "padded dims must be >= 3 in every direction";
        set_error(err);
        return -1;
    }
    const long n = (long)I * J * K * L;
    if (n >= (1L << 31)) {
        err = "padded volume exceeds 2^31 elements (32-bit index, as the reference)";
        set_error(err);
        return -1;
    }
    int ndev = 0;
    if (check(hipGetDeviceCount(&ndev), "hipGetDeviceCount")) return -1;
    if (ndev < 1) {
        err = "no HIP device visible";
        set_error(err);
        return -1;
    }
    I_ = I; J_ = J; K_ = K; L_ = L;
    H_ = (I + 1) / 2;
    n_ = n;
    {
        const long pz = (((long)H_ * J + 63) / 64) * 64;
        np_ = (64 - H_ % 64) % 64 + pz * K * L;  // cidx's layout (make_geo)
    }
    if (check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate")) return -1;
    if (check(hipEventCreateWithFlags(&ev_, hipEventDisableTiming), "hipEventCreate")) return -1;
    auto alloc = [&](double** p, long cnt) -> int {
        if (check(hipMalloc((void**)p, (size_t)cnt * sizeof(double)), "hipMalloc")) return -1;
        return check(hipMemsetAsync(*p, 0, (size_t)cnt * sizeof(double), st_), "hipMemset");
    };
    if (alloc(&full_[A_U], n) || alloc(&full_[A_PREV], n)) return -1;
    if (seg) {
        if (alloc(&full_[A_I0], n) || alloc(&full_[A_MASK], n)) return -1;
        for (int f = 0; f < 8; f++)
            if (alloc(&full_[A_G0 + f], n)) return -1;
    }
    if (mcf && alloc(&full_[A_EXACT], n)) return -1;
    {
        const long rp = (((long)(J - 2) * H_ + 63) / 64) * 64;
        nop_ = (long)(K - 2) * (L - 2) * rp * kOpnd;
    }
    for (int c = 0; c < 2; c++) {
        if (alloc(&op_[c], nop_) || alloc(&ur_[c], np_)) return -1;
    }
    if (alloc(&ub_, np_)) return -1;
    // partial-sum capacity: the largest grid of any reducing kernel
    const long g_full = (long)cdiv((long)I * J, kBlock) * K * L;
    // (a SOR pass has at most cdiv(points per row plane, kBlock / 4) runs
    // per k plane and one frame per block)
    const long g_sweep = (long)cdiv((long)H_ * (J - 2), kBlock / 4) * (K - 2) * (L - 2);
    // (the set-up kernel: blocks of 128 packed columns of a plane's rows)
    const long g_setup =
        cdiv(cdiv((long)(J - 1) * H_, 32) + (H_ + 31) / 32, 4) * K * L;
    long np = g_full > g_sweep ? g_full : g_sweep;
    if (g_setup > np) np = g_setup;
    if (np < 4096) np = 4096;
    npart_ = np;
    if (alloc(&part_, 2 * np)) return -1;
    if (alloc(&dscal_, 8) || alloc(&red2_, 4096)) return -1;
    if (check(hipMalloc((void**)&ctl_, sizeof(SorCtl)), "hipMalloc ctl")) return -1;
    if (check(hipHostMalloc((void**)&hctl_, sizeof(SorCtl), hipHostMallocDefault),
              "hipHostMalloc"))
        return -1;
    if (check(hipHostMalloc((void**)&hscal_, 8 * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc"))
        return -1;
    std::memset(hctl_, 0, sizeof(SorCtl));
    hctl_dev_ = host_alias(hctl_);
    if (check(hipStreamSynchronize(st_), "init sync")) return -1;
    ready_ = true;
    return 0;
}

void Engine::release() {
    if (st_) (void)hipStreamSynchronize(st_);
    for (auto& p : full_) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    for (int c = 0; c < 2; c++) {
        if (op_[c]) (void)hipFree(op_[c]);
        if (ur_[c]) (void)hipFree(ur_[c]);
        op_[c] = ur_[c] = nullptr;
    }
    if (ub_) (void)hipFree(ub_);
    if (part_) (void)hipFree(part_);
    if (dscal_) (void)hipFree(dscal_);
    if (red2_) (void)hipFree(red2_);
    red2_ = nullptr;
    if (ctl_) (void)hipFree(ctl_);
    if (hctl_) (void)hipHostFree(hctl_);
    if (hscal_) (void)hipHostFree(hscal_);
    ub_ = part_ = dscal_ = hscal_ = nullptr;
    for (int c = 0; c < 2; c++) {
        if (op32_[c]) (void)hipFree(op32_[c]);
        if (dr_[c]) (void)hipFree(dr_[c]);
        op32_[c] = dr_[c] = nullptr;
    }
    if (db_) (void)hipFree(db_);
    if (hp_) (void)hipFree(hp_);
    hp_ = nullptr;
    if (prev2_) (void)hipFree(prev2_);
    if (prev3_) (void)hipFree(prev3_);
    prev2_ = prev3_ = nullptr;
    hist_ok_ = hist2_ok_ = hist3_ok_ = false;
    if (balls_buf_) (void)hipFree(balls_buf_);
    balls_buf_ = nullptr;
    balls_cap_ = 0;
    balls_valid_ = false;
    if (ctl2_) (void)hipFree(ctl2_);
    if (hctl2_) (void)hipHostFree(hctl2_);
    if (gbar_) (void)hipFree(gbar_);
    if (hgbar_) (void)hipHostFree(hgbar_);
    gbar_ = nullptr;
    hgbar_ = nullptr;
    db_ = nullptr;
    ctl2_ = hctl2_ = hctl2_dev_ = nullptr;
    ctl_ = hctl_ = hctl_dev_ = nullptr;
    for (auto e : pev_) (void)hipEventDestroy(e);
    for (auto e : evq_) (void)hipEventDestroy(e);
    evq_.clear();
    pev_.clear();
    if (ev_) (void)hipEventDestroy(ev_);
    ev_ = nullptr;
    if (side_) {
        (void)hipStreamSynchronize(side_);
        (void)hipStreamDestroy(side_);
    }
    if (side_ev_) (void)hipEventDestroy(side_ev_);
    side_ = nullptr;
    side_ev_ = nullptr;
    side_pending_ = false;
    if (st_ && own_st_) (void)hipStreamDestroy(st_);
    st_ = nullptr;
    own_st_ = true;
    ready_ = false;
}

int Engine::adopt_stream(hipStream_t s) {
    if (check(hipStreamSynchronize(st_), "adopt_stream sync")) return -1;
    if (own_st_ && st_) (void)hipStreamDestroy(st_);
    st_ = s;
    own_st_ = false;
    return 0;
}

void Engine::undo_rotation(bool history_intact) {
    if (!rotated_) return;
    full_[A_U] = rot_[0];
    full_[A_PREV] = rot_[1];
    prev2_ = rot_[2];
    prev3_ = rot_[3];
    rotated_ = false;
    if (history_intact) return;
    // The buffer that became u's may already hold (part of) u^(n+1): k_unpack
    // writes it at the end of the step, and a later failure (the host wait, a
    // batch's later volume) cannot tell whether it ran.  That buffer is a
    // history level again (prev_full, prev2_ or prev3_), so no level can be
    // trusted: the next step starts from u^n until new history exists.
    hist_ok_ = hist2_ok_ = hist3_ok_ = false;
}

void Engine::step_done() {
    rotated_ = false;
    hist3_ok_ = shifted3_;  // prev3 = u^(n-2)
    hist2_ok_ = shifted2_;  // prev2 = u^(n-1)
    hist_ok_ = true;        // prev_full = u^n and u = u^(n+1)
}

static Geo make_geo(int I, int J, int K, int L, int H) {
    Geo G;
    G.I = I; G.J = J; G.K = K; G.L = L;
    G.sy = I; G.sz = I * J; G.sw = I * J * K;
    G.H = H;
    G.py = H;
    G.pz = ((H * J + 63) / 64) * 64;
    G.pw = G.pz * K;
    G.c0 = (64 - H % 64) % 64;
    G.RP = (((J - 2) * H + 63) / 64) * 64;
    return G;
}

int Engine::upload_side(ArrayId a, const double* host) {
    if (!full_[a]) return check(hipErrorInvalidValue, "upload: array not allocated");
    if (!side_ && check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "side stream"))
        return -1;
    if (!side_ev_ &&
        check(hipEventCreateWithFlags(&side_ev_, hipEventDisableTiming), "side event"))
        return -1;
    if (wait_side()) return -1;
    // the array may still be read by work queued on the engine's stream
    if (check(hipEventRecord(ev_, st_), "event") || check(hipStreamWaitEvent(side_, ev_, 0), "wait"))
        return -1;
    if (a == A_U || a == A_PREV) hist_ok_ = hist2_ok_ = hist3_ok_ = false;
    if (a >= A_G0 && a <= A_G7) g_version_++;
    if (check(hipMemcpyAsync(full_[a], host, (size_t)n_ * sizeof(double), hipMemcpyHostToDevice,
                             side_), "side upload") ||
        check(hipEventRecord(side_ev_, side_), "side event"))
        return -1;
    side_pending_ = true;
    return 0;
}

int Engine::wait_side() {
    if (!side_pending_) return 0;
    side_pending_ = false;
    return check(hipStreamWaitEvent(st_, side_ev_, 0), "wait for the side copy");
}

int Engine::finish_side() {
    if (!side_) return 0;
    const int rc = wait_side();
    return check(hipStreamSynchronize(side_), "side copy") | rc;
}

int Engine::upload(ArrayId a, const double* host) {
    if (!full_[a]) return check(hipErrorInvalidValue, "upload: array not allocated");
    if (wait_side()) return -1;
    if (a == A_U || a == A_PREV) hist_ok_ = hist2_ok_ = hist3_ok_ = false;
    if (a >= A_G0 && a <= A_G7) g_version_++;
    if (check(hipMemcpyAsync(full_[a], host, (size_t)n_ * sizeof(double),
                             hipMemcpyHostToDevice, st_), "upload"))
        return -1;
    return check(hipStreamSynchronize(st_), "upload sync");
}

int Engine::download(ArrayId a, double* host) {
    if (!full_[a]) return check(hipErrorInvalidValue, "download: array not allocated");
    if (wait_side()) return -1;
    if (check(hipMemcpyAsync(host, full_[a], (size_t)n_ * sizeof(double),
                             hipMemcpyDeviceToHost, st_), "download"))
        return -1;
    return check(hipStreamSynchronize(st_), "download sync");
}

namespace {
// Element copy over the padded array (segmentation4D.c:866-874): one FP64
// load and store per lane, the same access width as the SOR passes, so it
// doubles as the known-size reference for hardware byte counters.
__global__ __launch_bounds__(kBlock) void k_copy_f64(const double* __restrict__ src,
                                                    double* __restrict__ dst, long n) {
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < n;
         q += (long)gridDim.x * kBlock)
        dst[q] = src[q];
}
}  // namespace

int Engine::copy(ArrayId src, ArrayId dst) {
    if (!full_[src] || !full_[dst]) return check(hipErrorInvalidValue, "copy: array not allocated");
    if (src == dst) return 0;
    if (dst == A_U || dst == A_PREV) hist_ok_ = hist2_ok_ = hist3_ok_ = false;
    if (dst >= A_G0 && dst <= A_G7) g_version_++;
    const long nb = (n_ + kBlock - 1) / kBlock;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof.on) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st_);
    }
    k_copy_f64<<<(unsigned)(nb < 65536 ? nb : 65536), kBlock, 0, st_>>>(full_[src], full_[dst], n_);
    const int rc = launch_check("k_copy_f64");
    if (prof.on) {
        (void)hipEventRecord(e1, st_);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        prof.copy_ms += ms;
        prof.copy_launches++;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return rc;
}

int Engine::download_coef(int f, double* host) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    if (!scratch(A_TMP)) return -1;
    dim3 grid(cdiv((long)I_ * J_, kBlock), K_, L_);
    k_unpack_coef<<<grid, kBlock, 0, st_>>>(G, op_[0], op_[1], f, full_[A_TMP],
                                            split_ ? op32_[0] : nullptr,
                                            split_ ? op32_[1] : nullptr);
    if (launch_check("k_unpack_coef")) return -1;
    return download(A_TMP, host);
}

Geo Engine::geometry() const { return make_geo(I_, J_, K_, L_, H_); }

int Engine::setup(const StepParams& P, bool with_g, bool write_prev, SorCtl* ctl_target,
                  bool fuse_defect) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    const Scale S = make_scale(P.h);
    const bool p2 = is_pow2(P.h);
    const double h4 = P.h * P.h * P.h * P.h;
    CPtr8 g{};
    if (with_g)
        for (int f = 0; f < 8; f++) g.p[f] = full_[A_G0 + f];
    if (precision_ == 1 && ensure_mixed()) return -1;
    const bool mixed = fuse_defect && precision_ == 1;
    SplitBufs sb;
    for (int c = 0; c < 2; c++) {
        sb.opnd[c] = op_[c];
        sb.ured[c] = ur_[c];
        sb.opnd32[c] = mixed ? op32_[c] : nullptr;
    }
    sb.ublack = ub_;
    sb.dred = mixed ? dr_[0] : nullptr;
    sb.dblack = mixed ? db_ : nullptr;
    const double csor = P.tau / (P.h * P.h);
    double* part2 = part_ + npart_;
    double* prevf = write_prev ? full_[A_PREV] : nullptr;  // u^(n-1) (read by warm starts)
    // the extrapolated and the learned start need prev_full = u^(n-1) from a
    // completed step; without it the solve starts from u^n.  The quadratic
    // and cubic starts need one / two more levels (prev2_, prev3_); with
    // fewer, the highest order available is used.  The learned start
    // corrects the extrapolation of its network's base order (linear or
    // cubic) and keeps the history that base needs; until that history
    // exists it starts like mode 4.
    const bool use_net = warm_ == 2 && net_ && net_->ready();
    const int nbase = use_net ? net_->base() : -1;
    const bool keep2 = (warm_ >= 3 || nbase == 3) && write_prev;  // history 2 / 3 levels deep
    const bool keep3 = (warm_ == 4 || nbase == 3) && write_prev;
    int warm = 0;
    if ((warm_ == 1 || keep2) && hist_ok_ && write_prev) warm = 1;
    if (keep2 && hist_ok_ && hist2_ok_) warm = 3;
    if (warm_ == 4 && keep3 && hist_ok_ && hist2_ok_ && hist3_ok_) warm = 4;
    // History ring: prev_full (u^(n-1)), prev2_ (u^(n-2)), prev3_ (u^(n-3)).
    // No level is copied: after the launch the buffers rotate (prev_full =
    // u's buffer holding u^n, prev2_ = u^(n-1), ...) and u's buffer becomes
    // the oldest level the mode keeps, which k_unpack fills with u^(n+1).
    const double* h2 = nullptr;
    const double* h3 = nullptr;
    shifted2_ = shifted3_ = false;
    const bool h2_valid = hist2_ok_, h3_valid = hist3_ok_;  // before this step shifts them
    if (keep2) {
        if (!prev2_ && check(hipMalloc((void**)&prev2_, (size_t)n_ * sizeof(double)),
                             "hipMalloc prev2"))
            return -1;
        h2 = prev2_;
        shifted2_ = hist_ok_;  // prev2 becomes a valid u^(n-1)
        if (keep3) {
            if (!prev3_ && check(hipMalloc((void**)&prev3_, (size_t)n_ * sizeof(double)),
                                 "hipMalloc prev3"))
                return -1;
            h3 = prev3_;
            shifted3_ = hist_ok_ && hist2_ok_;  // prev3 becomes a valid u^(n-2)
        }
    }
    LearnedStart ls{nullptr, nullptr, 0.0, 1.f, 1};
    // order 0 needs no history (start u^n + c / s_t); orders 1 and 3 their
    // levels u^(n-1) .. u^(n-base)
    if (use_net && write_prev &&
        (nbase == 0 || (hist_ok_ && (nbase == 1 || (h2_valid && h3_valid))))) {
        // the ring is read before the set-up kernel rotates it; the kernel
        // forms base + q c / s_t from the network's output
        if (net_->predict(*this, nbase >= 1 ? full_[A_PREV] : nullptr,
                          nbase == 3 ? prev2_ : nullptr, nbase == 3 ? prev3_ : nullptr, &ls))
            return check(hipErrorUnknown, net_->err.c_str());
        warm = 2;
    }
    // blocks of 4 wavefronts, each on 32 packed columns of the plane's rows
    // (k_setup_step): chunks q from -ceil(H/32) to the last row's end
    const long nq = cdiv((long)(J_ - 1) * H_, 32) + (H_ + 31) / 32;
    const PlaneMap PM = plane_map((int)cdiv(nq, 4), K_, L_);
    const dim3 grid(plane_grid(PM));
    if ((long)PM.gx * PM.gk * PM.gl > npart_)
        return check(hipErrorInvalidValue, "setup: partials exceed their buffer");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof.on) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st_);
    }
    if (mixed && with_g) {
        if (p2) k_setup_step<true, true, true><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm, csor, part2);
        else k_setup_step<true, false, true><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm, csor, part2);
    } else if (mixed) {
        if (p2) k_setup_step<false, true, true><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm, csor, part2);
        else k_setup_step<false, false, true><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm, csor, part2);
    } else if (with_g) {
        if (p2) k_setup_step<true, true><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm);
        else k_setup_step<true, false><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm);
    } else {
        if (p2) k_setup_step<false, true><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm);
        else k_setup_step<false, false><<<grid, kBlock, 0, st_>>>(G, PM, S, P.eps2, h4, full_[A_U], g, sb, prevf, h2, h3, ls, part_, warm);
    }
    if (launch_check("k_setup_step")) return -1;
    split_ = mixed;
    if (write_prev) {
        rot_[0] = full_[A_U];
        rot_[1] = full_[A_PREV];
        rot_[2] = prev2_;
        rot_[3] = prev3_;
        rotated_ = true;
        double* freed = keep3 ? prev3_ : (keep2 ? prev2_ : full_[A_PREV]);
        if (keep3) prev3_ = prev2_;
        if (keep2) prev2_ = full_[A_PREV];
        full_[A_PREV] = full_[A_U];
        full_[A_U] = freed;  // written whole by k_unpack at the end of the step
    }
    if (prof.on) {
        (void)hipEventRecord(e1, st_);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        prof.setup_ms += ms;
        prof.setup_launches++;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    // one partial per (run, k, l), summed in two fixed-order levels
    const long np = (long)PM.gx * PM.gk * PM.gl;
    const int chunk = 4096;
    const int nb2 = (int)cdiv(np, chunk);
    if (nb2 > 2048) return check(hipErrorInvalidValue, "setup partials");
    SorCtl* tctl = ctl_target ? ctl_target : ctl_;
    k_sum_chunks<<<nb2, 1024, 0, st_>>>(part_, np, chunk, red2_);
    k_init_ctl<<<1, 1024, 0, st_>>>(red2_, nb2, tctl, P.gsTol, P.rel2);
    if (mixed) {  // R of the fused first defect, same two-level order
        k_sum_chunks<<<nb2, 1024, 0, st_>>>(part2, np, chunk, red2_ + 2048);
        k_set_r0<<<1, 1024, 0, st_>>>(red2_ + 2048, nb2, tctl);
    }
    defect_fused_ = mixed;
    return launch_check("k_init_ctl");
}

int Engine::autotune(int reps, float* best_ms) {
    // candidates: {variant, tile runs, tile planes, frame chunk}
    static const int cand[][4] = {
        {35, 16, 8, 0}, {1, 16, 8, 2}, {1, 16, 8, 0}, {3, 16, 8, 2}, {3, 16, 8, 0},
        {3, 8, 8, 0},   {3, 8, 16, 0}, {3, 8, 32, 0}, {3, 16, 4, 0}, {3, 16, 16, 0},
        {3, 32, 4, 0},  {3, 32, 8, 0}, {3, 4, 32, 0}, {3, 64, 4, 0}};
    const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    const int keep_v = sweep_variant_, keep_tb = sweep_tb_, keep_tk = sweep_tk_;
    const int keep_lc = sweep_lc_;
    hipEvent_t e0, e1;
    if (check(hipEventCreate(&e0), "event")) return -1;
    if (check(hipEventCreate(&e1), "event")) {
        (void)hipEventDestroy(e0);
        return -1;
    }
    float best = 1e30f;
    int bq = -1;
    int rc = 0;
    for (int q = 0; q < ncand && rc == 0; q++) {
        sweep_variant_ = cand[q][0];
        sweep_tb_ = cand[q][1];
        sweep_tk_ = cand[q][2];
        sweep_lc_ = cand[q][3];
        const SweepMap M = sweep_map();
        if ((long)M.gx * M.gk > npart_) continue;
        const dim3 grid((unsigned)sweep_grid(M));
        float cand_ms = 1e30f;
        for (int r = 0; r <= reps && rc == 0; r++) {
            rc |= check(hipEventRecord(e0, st_), "event");
            k_sor_sweep<double, 0, 2><<<grid, kBlock, 0, st_>>>(G, M, 0.5, 1.0, 1.0, 0, ur_[0], ur_[1], ub_, op_[0], nullptr, nullptr, nullptr);
            rc |= launch_check("autotune pass");
            rc |= check(hipEventRecord(e1, st_), "event");
            rc |= check(hipEventSynchronize(e1), "event");
            float ms = 0.f;
            rc |= check(hipEventElapsedTime(&ms, e0, e1), "event");
            if (r > 0 && ms < cand_ms) cand_ms = ms;  // the first launch warms the caches
        }
        if (cand_ms < best) {
            best = cand_ms;
            bq = q;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc || bq < 0) {
        sweep_variant_ = keep_v;
        sweep_tb_ = keep_tb;
        sweep_tk_ = keep_tk;
        sweep_lc_ = keep_lc;
        return rc ? -1 : 0;
    }
    sweep_variant_ = cand[bq][0];
    sweep_tb_ = cand[bq][1];
    sweep_tk_ = cand[bq][2];
    sweep_lc_ = cand[bq][3];
    if (best_ms) *best_ms = best;
    return 0;
}

bool Engine::tiled_order() const { return sweep_map().tiled != 0; }

SweepMap Engine::sweep_map(int elem_bytes) const {
    SweepMap M;
    M.nband = 1;
    M.bpb = 0;
    M.kp = (sweep_variant_ & 16) ? 4 : (sweep_variant_ & 8) ? 2 : 1;
    M.gx = (int)cdiv((long)H_ * (J_ - 2), kBlock / M.kp);
    M.gk0 = (int)cdiv((long)(K_ - 2), M.kp);
    M.tiled = (sweep_variant_ & 2) ? 1 : 0;
    const int nl = L_ - 2;
    int lc = sweep_lc_;
    if (lc <= 0) {
        // automatic order: in the linear order every XCD works through a
        // contiguous range of (chunk, k) and re-reads the other colour's
        // k-1, k, k+1 planes of the chunk's frames and its two neighbours
        // from L2 (and the Infinity Cache behind it); chunks of 2 frames keep
        // that window small (measured at 256x256x128x16, 3.2 MB: red pass
        // 1.259 ms against 1.352 ms for whole marches in the tiled order).
        // Planes whose window exceeds 7 MB are split into bands of runs, each
        // band walked through the k planes in turn (at 512x512x256x8, 12.7 MB:
        // 2 bands 10.01 ms per red + black pass pair, 4 bands 10.11, 8 bands
        // 10.14, whole marches in the tiled order 10.38).
        const long window = 3L * 4 * (long)H_ * J_ * elem_bytes;
        int nband = 1;
        while (nband < 16 && window / nband > (7L << 20)) nband *= 2;
        bool fits = window / nband <= (7L << 20) && M.gx >= 8 * nband;
        if (sweep_nband_ > 0) {  // forced band count
            nband = sweep_nband_ < M.gx ? sweep_nband_ : M.gx;
            fits = true;
        }
        if ((sweep_variant_ & 32) && fits) {
            lc = 2;
            M.tiled = 0;
            M.nband = nband;
            M.bpb = (M.gx + nband - 1) / nband;
        } else {
            // a volume with fewer than ~4096 blocks of whole l-marches splits
            // its frames so that the launch fills the GPU (even chunk lengths
            // keep the two-step load phases whole)
            const long b0 = (long)M.gx * M.gk0;
            const long want = b0 >= 4096 ? 1 : cdiv(4096, b0);
            lc = (int)cdiv(nl, want);
            lc = lc < 2 ? 2 : lc + (lc & 1);
        }
    }
    if (lc > nl) lc = nl;
    M.lc = lc;
    M.gk = M.gk0 * (int)cdiv(nl, lc);
    M.tb = sweep_tb_;
    M.tk = sweep_tk_;
    M.ntb = (M.gx + M.tb - 1) / M.tb;
    M.ntiles = M.ntb * ((M.gk + M.tk - 1) / M.tk);
    return M;
}

// One red-black SOR solve of precision T on the colour layout, pipelined:
// red(z+1) yields the red residual terms of iteration z, reduce(z) applies
// the stopping test on the device, black(z+1) is enqueued before the host has
// read the decision and becomes a no-op if iteration z stopped.
template <typename T>
struct SorJob {
    T c, om;
    double h4;
    const T* opnd[2];   // operand blocks per colour
    T* red[2];          // red ping-pong
    T* black;
    SorCtl* ctl;        // device control block (tol already set)
    SorCtl* hctl;       // pinned host mirror
    SorCtl* hctl_dev = nullptr;  // its device alias (null: copied in the stream instead)
    int cap;
    int kind0;          // profiling kinds kind0 .. kind0+2
};

// The whole solve as one cooperative launch of k_sor_fused: as many blocks
// as the GPU holds at once (a multiple of 8, at most the pass's blocks).
// Returns 1 when the launch does not apply (the caller then runs the solve
// pass by pass, with identical results), 0 on success, -1 on error.
template <typename T>
int sor_run_fused(Engine& E, const SorJob<T>& J, const Geo& G, const SweepMap& M, int nb,
                  int rnc, int rch, int* zout) {
    // the fused kernel is built for the default pass variant (two l-steps per
    // load phase, plain operand loads)
    if ((E.sweep_variant_ & 5) != 1 || E.prof.on) return 1;
    auto kern = k_sor_fused<T, 2>;
    int& fb = E.fused_blocks_[sizeof(T) == 4 ? 1 : 0];
    if (fb == 0) {
        int dev = 0, ncu = 0, per = 0;
        if (E.check(hipGetDevice(&dev), "hipGetDevice") ||
            E.check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev),
                    "CU count") ||
            E.check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)kern, kBlock,
                                                                   0), "occupancy"))
            return -1;
        fb = per * ncu > 0 ? per * ncu : -1;
    }
    if (fb < 0) return 1;
    // a multiple of 8 blocks keeps every work item on its own launch's XCD;
    // grids of fewer than 8 items run as they are
    const int items = sweep_grid(M);
    int nblk = std::min(fb, items);
    if (nblk >= 8) nblk = nblk / 8 * 8;
    if (!E.gbar_) {
        if (E.check(hipMalloc((void**)&E.gbar_, sizeof(GridBar)), "hipMalloc barrier") ||
            E.check(hipHostMalloc((void**)&E.hgbar_, sizeof(GridBar), hipHostMallocDefault),
                    "hipHostMalloc barrier"))
            return -1;
    }
    k_bar_reset<<<1, 64, 0, E.st_>>>(E.gbar_);
    if (E.launch_check("k_bar_reset")) return -1;
    Geo g = G;
    SweepMap m = M;
    T c = J.c, om = J.om;
    double h4 = J.h4;
    T* red0 = J.red[0];
    T* red1 = J.red[1];
    T* black = J.black;
    const T* op0 = J.opnd[0];
    const T* op1 = J.opnd[1];
    double* part_r = E.part_;
    double* part_b = E.part_ + E.npart_;
    double* red2 = E.red2_;
    SorCtl* ctl = J.ctl;
    int cap = J.cap;
    GridBar* gb = E.gbar_;
    void* args[] = {&g, &m, &c, &om, &h4, &red0, &red1, &black, &op0, &op1, &part_r, &part_b,
                    &nb, &rnc, &rch, &red2, &ctl, &cap, &gb};
    const hipError_t el = hipLaunchCooperativeKernel((const void*)kern, dim3((unsigned)nblk),
                                                     dim3(kBlock), args, 0, E.st_);
    if (el != hipSuccess) {
        // refused (too large for the device, or no cooperative launch): pass
        // by pass from now on, with the same results; nothing was enqueued
        (void)hipGetLastError();
        fb = -1;
        return 1;
    }
    if (E.check(hipMemcpyAsync(J.hctl, J.ctl, sizeof(SorCtl), hipMemcpyDeviceToHost, E.st_),
                "ctl copy") ||
        E.check(hipMemcpyAsync(E.hgbar_, E.gbar_, sizeof(GridBar), hipMemcpyDeviceToHost,
                               E.st_), "barrier copy") ||
        E.check(hipStreamSynchronize(E.st_), "fused solve"))
        return -1;
    if (E.hgbar_->abort)
        return E.check(hipErrorLaunchTimeOut, "fused solve: a grid barrier timed out");
    E.fused_last_ = true;
    *zout = J.hctl->zfinal;
    return 0;
}

template <typename T>
int sor_run(Engine& E, const SorJob<T>& J, int* zout) {
    const Geo G = make_geo(E.I_, E.J_, E.K_, E.L_, E.H_);
    const SweepMap M = E.sweep_map((int)sizeof(T));
    const int nb = M.gx * M.gk;
    if ((long)nb > E.npart_) {
        std::fprintf(stderr, "seg4d: SOR grid of %d blocks exceeds the partial buffer\n", nb);
        return -1;
    }
    int rnc = 0, rch = 0;
    resid_chunking(nb, &rnc, &rch);
    E.fused_last_ = false;
    if (E.fused_ == 1) {
        const int r = sor_run_fused<T>(E, J, G, M, nb, rnc, rch, zout);
        if (r <= 0) return r;
    }
    const dim3 grid((unsigned)sweep_grid(M));
    double* part_r = E.part_;
    double* part_b = E.part_ + E.npart_;
    hipStream_t st = E.st_;
    int pe = 0;
    std::vector<int> kinds;
    std::vector<int> iters;  // iteration index of each timed pass
    auto ev_pair = [&](int kind, int iter) -> hipEvent_t* {
        while ((int)E.pev_.size() < pe + 2) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            E.pev_.push_back(e);
        }
        hipEvent_t* r = &E.pev_[pe];
        pe += 2;
        kinds.push_back(kind);
        iters.push_back(iter);
        return r;
    };
    const bool u2 = (E.sweep_variant_ & 1) != 0;
    const bool nto = (E.sweep_variant_ & 4) != 0;
    // one colour pass: template variant by (resid kind, U, nontemporal operands)
    auto pass = [&](int rk, int colour, const T* xin, T* xout, const T* y, const T* ob,
                    double* part, const SorCtl* ctl) {
#define S4D_PASS(RK, UU, NT)                                                              \
    k_sor_sweep<T, RK, UU, NT><<<grid, kBlock, 0, st>>>(G, M, J.c, J.om, J.h4, colour, xin, \
                                                        xout, y, ob, part, ctl, nullptr)
        if (rk == 0) {
            if (u2) { if (nto) S4D_PASS(0, 2, true); else S4D_PASS(0, 2, false); }
            else { if (nto) S4D_PASS(0, 1, true); else S4D_PASS(0, 1, false); }
        } else if (rk == 1) {
            if (u2) { if (nto) S4D_PASS(1, 2, true); else S4D_PASS(1, 2, false); }
            else { if (nto) S4D_PASS(1, 1, true); else S4D_PASS(1, 1, false); }
        } else {
            if (u2) { if (nto) S4D_PASS(2, 2, true); else S4D_PASS(2, 2, false); }
            else { if (nto) S4D_PASS(2, 1, true); else S4D_PASS(2, 1, false); }
        }
#undef S4D_PASS
    };
    auto red = [&](int z, bool resid) -> int {
        hipEvent_t* e = E.prof.on ? ev_pair(J.kind0 + (resid ? 1 : 0), z) : nullptr;
        if (e) (void)hipEventRecord(e[0], st);
        pass(resid ? 1 : 0, 0, J.red[(z - 1) & 1], J.red[z & 1], J.black, J.opnd[0],
             resid ? part_r : nullptr, J.ctl);
        if (e) (void)hipEventRecord(e[1], st);
        return E.launch_check("k_sor_sweep red");
    };
    auto black = [&](int z) -> int {
        hipEvent_t* e = E.prof.on ? ev_pair(J.kind0 + 2, z) : nullptr;
        if (e) (void)hipEventRecord(e[0], st);
        pass(2, 1, J.black, J.black, J.red[z & 1], J.opnd[1], part_b, J.ctl);
        if (e) (void)hipEventRecord(e[1], st);
        return E.launch_check("k_sor_sweep black");
    };
    if (red(1, false) || black(1)) return -1;
    // Iterations are enqueued up to `depth` ahead of the stopping decision the
    // host has seen; after a stop every queued pass and reduce is a no-op (they
    // read the device flag), so a deeper queue only hides the host round trip.
    const int nred_pts = (int)std::min<long>(E.prof.nred > 0 ? E.prof.nred : 1, 1L << 30);
    const int depth = E.pipe_depth_ > 0 ? E.pipe_depth_ : (nred_pts < (4 << 20) ? 4 : 1);
    while ((int)E.evq_.size() < depth) {
        hipEvent_t e;
        if (E.check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event")) return -1;
        E.evq_.push_back(e);
    }
    int z = 1;       // next iteration whose residual gets reduced
    int queued = 0;  // reduced iterations the host has not looked at yet
    for (;;) {
        while (queued < depth) {
            if (red(z + 1, true)) return -1;
            // two fixed-order levels: chunk sums (many blocks), then the test
            k_resid_chunks<<<2 * rnc, 1024, 0, st>>>(part_r, part_b, nb, rnc, rch, E.red2_, J.ctl);
            k_reduce_resid<<<1, 1024, 0, st>>>(E.red2_, rnc, E.red2_ + rnc, rnc, J.ctl, nullptr, z,
                                               J.cap, 1, J.hctl_dev);
            if (E.launch_check("k_reduce_resid")) return -1;
            if (!J.hctl_dev &&
                E.check(hipMemcpyAsync(J.hctl, J.ctl, sizeof(SorCtl), hipMemcpyDeviceToHost, st),
                        "ctl copy"))
                return -1;
            if (E.check(hipEventRecord(E.evq_[z % depth], st), "event")) return -1;
            if (black(z + 1)) return -1;  // a no-op once iteration z stopped
            z++;
            queued++;
        }
        const int zo = z - queued;  // oldest iteration not looked at
        hipError_t q;
        while ((q = hipEventQuery(E.evq_[zo % depth])) == hipErrorNotReady) {
        }
        if (E.check(q, "solve wait")) return -1;
        queued--;
        // the host copy holds the state after reduce(zo) or a later one
        if (J.hctl->stop) break;
    }
    if (E.check(hipStreamSynchronize(st), "solve sync")) return -1;
    z = J.hctl->zfinal;
    if (E.prof.on) {
        for (int q = 0; q < pe; q += 2) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, E.pev_[q], E.pev_[q + 1]);
            // passes enqueued after the stop returned at once: not counted
            // (red(z') yields the residual of z' - 1, black(z') is part of z')
            const int kd = kinds[q / 2] - J.kind0, it = iters[q / 2];
            const bool real = (kd == 2) ? it <= z : it <= z + 1;
            if (!real) continue;
            E.prof.sweep_ms[kinds[q / 2]] += ms;
            E.prof.sweep_launches[kinds[q / 2]]++;
        }
    }
    *zout = z;
    return 0;
}

int Engine::solve(const StepParams& P, SolveStats* out) {
    const long nint = (long)(I_ - 2) * (J_ - 2) * (K_ - 2) * (L_ - 2);
    // interior red / black point counts for the byte model: the corner
    // (1,1,1,1) is red, so red has the extra point when the count is odd
    prof.nred = (nint + 1) / 2;
    prof.nblack = nint - prof.nred;
    if (precision_ == 1) return solve_mixed(P, out);
    SorJob<double> J;
    J.c = P.tau / (P.h * P.h);
    J.om = P.omega;
    J.h4 = P.h * P.h * P.h * P.h;
    J.opnd[0] = op_[0];
    J.opnd[1] = op_[1];
    J.red[0] = ur_[0];
    J.red[1] = ur_[1];
    J.black = ub_;
    J.ctl = ctl_;
    J.hctl = hctl_;
    J.hctl_dev = hctl_dev_;
    J.cap = P.cap > 0x7fffffffL ? 0x7fffffff : (int)P.cap;
    J.kind0 = 0;
    int z = 0;
    if (sor_run<double>(*this, J, &z)) return -1;
    red_final_ = z & 1;
    if (out) {
        out->z = hctl_->zfinal;
        out->R = hctl_->R;
        out->tol = hctl_->tol;
        out->sumsq = hctl_->sumsq;
        out->rounds = 0;
    }
    if (prof.on) {
        prof.iterations += z;
        prof.voxel_updates += (long)z * nint;
    }
    return 0;
}

int Engine::ensure_mixed() {
    if (op32_[0]) return 0;
    auto allocf = [&](float** p, long cnt) -> int {
        if (check(hipMalloc((void**)p, (size_t)cnt * sizeof(float)), "hipMalloc")) return -1;
        return check(hipMemsetAsync(*p, 0, (size_t)cnt * sizeof(float), st_), "hipMemset");
    };
    for (int c = 0; c < 2; c++)
        if (allocf(&op32_[c], nop_) || allocf(&dr_[c], np_)) return -1;
    if (allocf(&db_, np_)) return -1;
    if (check(hipMalloc((void**)&ctl2_, sizeof(SorCtl)), "hipMalloc ctl")) return -1;
    if (check(hipHostMalloc((void**)&hctl2_, sizeof(SorCtl), hipHostMallocDefault), "hipHostMalloc"))
        return -1;
    std::memset(hctl2_, 0, sizeof(SorCtl));
    hctl2_dev_ = host_alias(hctl2_);
    return check(hipStreamSynchronize(st_), "mixed alloc");
}

// FP64 residual of the reference system for the current colour-layout state
// (red in ur_[0], black in ub_); writes the FP32 defect into the FP32 operand
// blocks and the residual into ctl_->R (host copy in hctl_).
// copy32: also write the coefficients' high words into the FP32 blocks (first
// defect of a step whose set-up wrote full FP64 blocks);
// dred != null: first apply the correction (red: dred, black: db_), i.e.
// u += d fused into the defect passes.
int Engine::defect64(const StepParams& P, bool copy32, const float* dred) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    const SweepMap M = sweep_map();
    const int nb = M.gx * M.gk;
    const dim3 grid((unsigned)sweep_grid(M));
    const double c = P.tau / (P.h * P.h);
    const double h4 = P.h * P.h * P.h * P.h;
    double* part_r = part_;
    double* part_b = part_ + npart_;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof.on) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st_);
    }
    // full FP64 operand blocks for the first defect (RESID 4 / 3); the split
    // operands of the fused set-up only come with a correction (RESID 6)
    if ((copy32 || !dred) && split_)
        return check(hipErrorInvalidValue, "defect: split operands without a correction");
    if (copy32) {
        k_sor_sweep<double, 4, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 0, ur_[0], nullptr, ub_, op_[0], part_r, nullptr, op32_[0]);
        k_sor_sweep<double, 4, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 1, ub_, nullptr, ur_[0], op_[1], part_b, nullptr, op32_[1]);
    } else if (dred) {
        // red: own u_r + d_r (stored), black neighbours u_b + d_b; black: own
        // u_b + d_b (stored), red neighbours already corrected
        if (split_) {
            k_sor_sweep<double, 6, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 0, ur_[0], ur_[0], ub_, op_[0], part_r, nullptr, op32_[0], dred, db_);
            k_sor_sweep<double, 6, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 1, ub_, ub_, ur_[0], op_[1], part_b, nullptr, op32_[1], db_, nullptr);
        } else {
            k_sor_sweep<double, 5, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 0, ur_[0], ur_[0], ub_, op_[0], part_r, nullptr, op32_[0], dred, db_);
            k_sor_sweep<double, 5, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 1, ub_, ub_, ur_[0], op_[1], part_b, nullptr, op32_[1], db_, nullptr);
        }
    } else {
        k_sor_sweep<double, 3, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 0, ur_[0], nullptr, ub_, op_[0], part_r, nullptr, op32_[0]);
        k_sor_sweep<double, 3, 2><<<grid, kBlock, 0, st_>>>(G, M, c, P.omega, h4, 1, ub_, nullptr, ur_[0], op_[1], part_b, nullptr, op32_[1]);
    }
    if (launch_check("k_sor_sweep defect")) return -1;
    if (prof.on) (void)hipEventRecord(e1, st_);
    // two fixed-order levels, as in the solve (the chunk sums take no control
    // block: the stop flag of the previous round is still raised); z = 1,
    // cap = 1: the reduction always records R and raises stop, the outer
    // decision is taken on the host from R and tol
    int rnc = 0, rch = 0;
    resid_chunking(nb, &rnc, &rch);
    k_resid_chunks<<<2 * rnc, 1024, 0, st_>>>(part_r, part_b, nb, rnc, rch, red2_, nullptr);
    k_reduce_resid<<<1, 1024, 0, st_>>>(red2_, rnc, red2_ + rnc, rnc, ctl_, nullptr, 1, 1, 0);
    if (launch_check("k_reduce_resid")) return -1;
    if (check(hipMemcpyAsync(hctl_, ctl_, sizeof(SorCtl), hipMemcpyDeviceToHost, st_), "ctl copy"))
        return -1;
    if (check(hipStreamSynchronize(st_), "defect sync")) return -1;
    if (prof.on) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        prof.defect_ms += ms;
        prof.defect_launches += 2;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

// FP32-refined solve: FP64 defect correction around FP32 red-black SOR on the
// correction equation A d = b - A u (same matrix, coefficients rounded to
// FP32).  Stops on the FP64 residual: R <= tol (the reference's test), or
// after max_rounds_ corrections.
int Engine::solve_mixed(const StepParams& P, SolveStats* out) {
    if (ensure_mixed()) return -1;
    const long nint = (long)(I_ - 2) * (J_ - 2) * (K_ - 2) * (L_ - 2);
    SorJob<float> J;
    J.c = (float)(P.tau / (P.h * P.h));
    J.om = (float)(inner_omega_ > 0.0 ? inner_omega_ : P.omega);
    J.h4 = P.h * P.h * P.h * P.h;
    J.opnd[0] = op32_[0];
    J.opnd[1] = op32_[1];
    J.red[0] = dr_[0];
    J.red[1] = dr_[1];
    J.black = db_;
    J.ctl = ctl2_;
    J.hctl = hctl2_;
    J.hctl_dev = hctl2_dev_;
    J.cap = P.cap > 0x7fffffffL ? 0x7fffffff : (int)P.cap;
    J.kind0 = 3;
    long sweeps = 0;
    int rounds = 0;
    // the first defect: formed by the set-up kernel (u = u^n, d = start - u^n
    // already in place), or a defect pass of the start (with the FP32
    // coefficient copies) when the set-up did not form it
    const bool fused = defect_fused_;
    defect_fused_ = false;
    if (fused) {
        if (check(hipMemcpyAsync(hctl_, ctl_, sizeof(SorCtl), hipMemcpyDeviceToHost, st_),
                  "ctl copy") ||
            check(hipStreamSynchronize(st_), "ctl sync"))
            return -1;
    } else if (defect64(P, true)) {
        return -1;
    }
    while (hctl_->R > hctl_->tol && rounds < max_rounds_) {
        // d = 0 (ghosts stay 0: homogeneous boundary of the correction) unless
        // the first round starts from the set-up's d0; the second red buffer
        // needs no clearing (its interior is written before it is read)
        if (rounds > 0 || !fused)
            for (float* d : {dr_[0], db_})
                if (check(hipMemsetAsync(d, 0, (size_t)np_ * sizeof(float), st_), "memset d"))
                    return -1;
        k_inner_ctl<<<1, 64, 0, st_>>>(ctl_, ctl2_, inner_frac_, inner_floor2_);
        if (launch_check("k_inner_ctl")) return -1;
        int z = 0;
        if (sor_run<float>(*this, J, &z)) return -1;
        sweeps += z;
        rounds++;
        // u += d (the accepted red correction is the buffer written by red(z))
        // fused into the next FP64 defect
        if (defect64(P, false, dr_[z & 1])) return -1;
    }
    red_final_ = 0;
    if (out) {
        out->z = (int)sweeps;
        out->R = hctl_->R;
        out->tol = hctl_->tol;
        out->sumsq = hctl_->sumsq;
        out->rounds = rounds;
    }
    if (prof.on) {
        prof.iterations += sweeps;
        prof.voxel_updates += sweeps * nint;
        prof.rounds += rounds;
    }
    return 0;
}

int Engine::learned_preview(float* host_c, double* host_q) {
    if (!net_ || !net_->ready()) return check(hipErrorInvalidValue, "learned preview: no network");
    const int nb = net_->base();
    if ((nb >= 1 && !hist_ok_) || (nb == 3 && !(hist2_ok_ && hist3_ok_)))
        return check(hipErrorInvalidValue, "learned preview: the network's history does not exist");
    LearnedStart ls{nullptr, nullptr, 0.0, 1.f, 1};
    if (net_->predict(*this, nb >= 1 ? full_[A_PREV] : nullptr, nb == 3 ? prev2_ : nullptr,
                      nb == 3 ? prev3_ : nullptr, &ls))
        return check(hipErrorUnknown, net_->err.c_str());
    double sum = 0.0;
    if (check(hipMemcpyAsync(host_c, ls.out, (size_t)n_ * sizeof(float), hipMemcpyDeviceToHost, st_),
              "preview copy") ||
        check(hipMemcpyAsync(&sum, ls.dsum, sizeof(double), hipMemcpyDeviceToHost, st_),
              "preview copy") ||
        check(hipStreamSynchronize(st_), "preview sync"))
        return -1;
    *host_q = std::sqrt(sum * ls.inv_n);
    return 0;
}

int Engine::unpack(const double* dmask) {
    if (wait_side()) return -1;  // e.g. the mask uploaded beside the step
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    k_unpack<<<(unsigned)cdiv(n_, kBlock), kBlock, 0, st_>>>(G, ur_[red_final_], ub_, full_[A_U],
                                                            dmask);
    return launch_check("k_unpack");
}

int Engine::unpack_gated(const double* dmask, const SorCtl* sel, const int* gate) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    k_unpack_gated<<<(unsigned)cdiv(n_, kBlock), kBlock, 0, st_>>>(G, ur_[0], ur_[1], ub_,
                                                                  full_[A_U], dmask, sel, gate);
    return launch_check("k_unpack_gated");
}

double* Engine::scratch(ArrayId a) {
    if (!full_[a] &&
        check(hipMalloc((void**)&full_[a], (size_t)n_ * sizeof(double)), "hipMalloc scratch"))
        return nullptr;
    return full_[a];
}

int Engine::put_balls(const Balls& b, BallList* out) {
    const size_t ni = (size_t)L_ * sizeof(int);
    const size_t nd = (size_t)(b.total > 0 ? b.total : 1) * sizeof(double);
    const size_t bytes = 2 * ni + 3 * nd;
    const bool same = balls_valid_ && b.total == balls_sent_.total &&
                      b.count == balls_sent_.count && b.offset == balls_sent_.offset &&
                      b.cx == balls_sent_.cx && b.cy == balls_sent_.cy && b.cz == balls_sent_.cz;
    if (!same) {
        if ((int)b.count.size() != L_ || (int)b.offset.size() != L_ ||
            (int)b.cx.size() < b.total || (int)b.cy.size() < b.total ||
            (int)b.cz.size() < b.total)
            return check(hipErrorInvalidValue, "ball list sizes");
        for (int l = 0; l < L_; l++)
            if (b.count[l] < 0 || b.offset[l] < 0 || b.offset[l] + b.count[l] > b.total)
                return check(hipErrorInvalidValue, "ball list offsets");
        if (bytes > balls_cap_) {
            if (balls_buf_) {
                (void)hipStreamSynchronize(st_);
                (void)hipFree(balls_buf_);
                balls_buf_ = nullptr;
                balls_cap_ = 0;
            }
            if (check(hipMalloc((void**)&balls_buf_, bytes), "hipMalloc balls")) return -1;
            balls_cap_ = bytes;
        }
        // copies ordered after earlier kernels that read the old list, then
        // waited for: the host vectors need not outlive this call
        int rc = 0;
        auto put = [&](size_t off, const void* src, size_t n) {
            rc |= check(hipMemcpyAsync(balls_buf_ + off, src, n, hipMemcpyHostToDevice, st_),
                        "balls");
        };
        put(0, b.count.data(), ni);
        put(ni, b.offset.data(), ni);
        if (b.total > 0) {
            const size_t nb = (size_t)b.total * sizeof(double);
            put(2 * ni, b.cx.data(), nb);
            put(2 * ni + nd, b.cy.data(), nb);
            put(2 * ni + 2 * nd, b.cz.data(), nb);
        }
        rc |= check(hipStreamSynchronize(st_), "balls sync");
        if (rc) {
            balls_valid_ = false;
            return -1;
        }
        balls_sent_ = b;
        balls_valid_ = true;
    }
    out->count = (const int*)balls_buf_;
    out->offset = (const int*)(balls_buf_ + ni);
    out->cx = (const double*)(balls_buf_ + 2 * ni);
    out->cy = (const double*)(balls_buf_ + 2 * ni + nd);
    out->cz = (const double*)(balls_buf_ + 2 * ni + 2 * nd);
    out->R = b.radius;
    return 0;
}

int Engine::rescale(double* d, const Balls& balls, bool apply_mask) {
    if (wait_side()) return -1;
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    hist_ok_ = hist2_ok_ = hist3_ok_ = false;  // u changed outside a time step
    if (balls.total > 0) {
        BallList B;
        if (put_balls(balls, &B)) return -1;
        k_rescale_balls<<<L_ - 2, kBlock, 0, st_>>>(G, d, B);
        if (launch_check("k_rescale_balls")) return -1;
    }
    if (apply_mask) {
        if (!full_[A_MASK]) return check(hipErrorInvalidValue, "rescale: no mask array");
        dim3 grid(cdiv((long)(I_ - 2) * (J_ - 2), kBlock), K_ - 2, L_ - 2);
        k_apply_mask<<<grid, kBlock, 0, st_>>>(G, d, full_[A_MASK]);
        if (launch_check("k_apply_mask")) return -1;
    }
    return 0;
}

// Interior max into dscal_[1] (device) -- shared by threshold and max_interior.
static int max_interior_dev(Engine& E, const double* d) {
    const Geo G = make_geo(E.I_, E.J_, E.K_, E.L_, E.H_);
    dim3 grid(cdiv((long)(E.I_ - 2) * (E.J_ - 2), kBlock), E.K_ - 2, E.L_ - 2);
    k_max_interior<<<grid, kBlock, 0, E.st_>>>(G, d, E.part_);
    if (E.launch_check("k_max_interior")) return -1;
    k_reduce_max<<<1, 1024, 0, E.st_>>>(E.part_, (int)(grid.x * grid.y * grid.z), E.dscal_ + 1);
    return E.launch_check("k_reduce_max");
}

int Engine::threshold(const double* in, double* thr, const Balls& balls, double alpha,
                      double beta) {
    hist_ok_ = hist2_ok_ = hist3_ok_ = false;
    if (max_interior_dev(*this, in)) return -1;
    if (balls.total <= 0) return 0;
    BallList B;
    if (put_balls(balls, &B)) return -1;
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    k_threshold_balls<<<L_ - 2, kBlock, 0, st_>>>(G, in, thr, B, alpha, beta, dscal_ + 1);
    return launch_check("k_threshold_balls");
}

int Engine::seed(double* d, double* mask, const Balls& balls) {
    hist_ok_ = hist2_ok_ = hist3_ok_ = false;
    if (balls.total <= 0) return 0;
    BallList B;
    if (put_balls(balls, &B)) return -1;
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    k_seed_balls<<<L_ - 2, kBlock, 0, st_>>>(G, d, mask, B);
    return launch_check("k_seed_balls");
}

int Engine::max_interior(const double* d, double* out) {
    if (max_interior_dev(*this, d)) return -1;
    if (check(hipMemcpyAsync(hscal_ + 1, dscal_ + 1, sizeof(double), hipMemcpyDeviceToHost, st_),
              "scalar copy") ||
        check(hipStreamSynchronize(st_), "max sync"))
        return -1;
    *out = hscal_[1];
    return 0;
}

int Engine::ball_minmax(const double* d, int l, double cx, double cy, double cz, double R,
                        double* mn, double* mx) {
    if (l < 0 || l >= L_) return check(hipErrorInvalidValue, "ball_minmax: frame out of range");
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    k_ball_minmax<<<1, kBlock, 0, st_>>>(G, d, l, cx, cy, cz, R, dscal_ + 2);
    if (launch_check("k_ball_minmax")) return -1;
    if (check(hipMemcpyAsync(hscal_ + 2, dscal_ + 2, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             st_),
              "scalar copy") ||
        check(hipStreamSynchronize(st_), "ball sync"))
        return -1;
    *mn = hscal_[2];
    *mx = hscal_[3];
    return 0;
}

namespace {
__global__ __launch_bounds__(kBlock) void k_scaled_sum(const double* __restrict__ a, long n,
                                                      double h4, double* __restrict__ partials) {
    double acc = 0.0;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < n; q += (long)gridDim.x * kBlock)
        acc += (a[q] * h4);
    const double bs = block_sum256(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}
}  // namespace

int Engine::mass(const double* d, double h, double* out) {
    const double h4 = h * h * h * h;
    const int nblk = 2048;
    k_scaled_sum<<<nblk, kBlock, 0, st_>>>(d, n_, h4, part_);
    if (launch_check("k_scaled_sum")) return -1;
    return reduce_to_host(nblk, out);
}

int Engine::seg_step(const StepParams& P, const Balls* balls, SolveStats* st) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof.on) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st_);
    }
    if (setup(P, true, true, nullptr, precision_ == 1)) {
        undo_rotation();  // set-up may fail after it rotated (its reduction)
        return -1;
    }
    if (inject_ == 1) {
        inject_ = 0;
        (void)hipStreamSynchronize(st_);
        undo_rotation();
        return check(hipErrorUnknown, "injected failure after the set-up");
    }
    if (solve(P, st)) {
        undo_rotation();
        return -1;
    }
    const bool has_balls = balls && balls->total > 0;
    if (has_balls ? (unpack(nullptr) || rescale(full_[A_U], *balls, true))
                  : unpack(full_[A_MASK])) {
        undo_rotation();
        return -1;
    }
    if (prof.on) {
        (void)hipEventRecord(e1, st_);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        prof.step_ms += ms;
        prof.steps++;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (check(hipStreamSynchronize(st_), "seg_step")) {
        undo_rotation();
        return -1;
    }
    if (inject_ == 2) {  // as a failed final wait: k_unpack has run
        inject_ = 0;
        undo_rotation();
        return check(hipErrorUnknown, "injected failure at the final wait");
    }
    step_done();
    return 0;
}

int Engine::mcf_step(const StepParams& P, double t, SolveStats* st) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    if (setup(P, false, true)) {
        undo_rotation();  // set-up may fail after it rotated (its reduction)
        return -1;
    }
    // the boundary values at t go to the ghosts of u's new buffer (k_unpack
    // copies them back from the colour arrays with the interior)
    dim3 grid(cdiv((long)I_ * J_, kBlock), K_, L_);
    k_mcf_exact<<<grid, kBlock, 0, st_>>>(G, P.h, t, full_[A_U], 1);
    int rc = launch_check("k_mcf_exact");
    if (!rc) {
        k_pack_ghosts<<<grid, kBlock, 0, st_>>>(G, full_[A_U], ur_[0], ur_[1], ub_);
        rc = launch_check("k_pack_ghosts");
    }
    if (!rc) rc = solve(P, st);
    if (!rc) rc = unpack(nullptr);
    if (!rc) rc = check(hipStreamSynchronize(st_), "mcf_step");
    if (rc) {
        undo_rotation();
        return -1;
    }
    step_done();
    return 0;
}

int Engine::reflect(double* d) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    dim3 grid(cdiv((long)I_ * J_, kBlock), K_, L_);
    k_reflect<<<grid, kBlock, 0, st_>>>(G, d);
    return launch_check("k_reflect");
}

int Engine::reduce_to_host(int np, double* out) {
    k_reduce_sum<false><<<1, 1024, 0, st_>>>(part_, np, dscal_);
    if (launch_check("k_reduce_sum")) return -1;
    if (check(hipMemcpyAsync(hscal_, dscal_, sizeof(double), hipMemcpyDeviceToHost, st_),
              "scalar copy"))
        return -1;
    if (check(hipStreamSynchronize(st_), "reduce sync")) return -1;
    *out = hscal_[0];
    return 0;
}

namespace {
__global__ __launch_bounds__(kBlock) void k_sumsq_interior(Geo G, const double* __restrict__ a,
                                                          double h4,
                                                          double* __restrict__ partials) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int k = blockIdx.y + 1, l = blockIdx.z + 1;
    const int w = G.I - 2;
    double acc = 0.0;
    if (t < w * (G.J - 2)) {
        const int j = t / w + 1, i = t - (j - 1) * w + 1;
        const double v = a[((l * G.K + k) * G.J + j) * G.I + i];
        acc = v * v * h4;
    }
    const double bs = block_sum256(acc);
    if (threadIdx.x == 0)
        partials[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = bs;
}
}  // namespace

int Engine::sumsq_interior(const double* d, double h, double* out) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    const double h4 = h * h * h * h;
    dim3 grid(cdiv((long)(I_ - 2) * (J_ - 2), kBlock), K_ - 2, L_ - 2);
    k_sumsq_interior<<<grid, kBlock, 0, st_>>>(G, d, h4, part_);
    if (launch_check("k_sumsq_interior")) return -1;
    return reduce_to_host((int)(grid.x * grid.y * grid.z), out);
}

int Engine::presmooth(double* d, double sigma, double h, double omega, double gsTol,
                      double rel2, double cap, int* iters, double* Rout) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    const double c = sigma / (h * h);
    const double h4 = h * h * h * h;
    hist_ok_ = hist2_ok_ = hist3_ok_ = false;  // prev is overwritten with the image
    if (check(hipMemcpyAsync(full_[A_PREV], d, (size_t)n_ * sizeof(double),
                             hipMemcpyDeviceToDevice, st_), "presmooth copy"))
        return -1;
    if (reflect(d)) return -1;
    double tol = gsTol;
    if (rel2 > 0.0) {
        double S = 0.0;
        if (sumsq_interior(full_[A_PREV], h, &S)) return -1;
        tol = rel2 * S;
    }
    // right-hand side packed per colour in the interior-plane order
    const long nhp = 2L * (K_ - 2) * (L_ - 2) * G.RP;
    if (!hp_ && check(hipMalloc((void**)&hp_, (size_t)nhp * sizeof(double)), "hipMalloc heat prev"))
        return -1;
    const dim3 gfull(cdiv((long)I_ * J_, kBlock), K_, L_);
    k_heat_pack<<<gfull, kBlock, 0, st_>>>(G, d, ur_[0], ur_[1], ub_, hp_);
    if (launch_check("k_heat_pack")) return -1;
    // red-black passes on the colour layout with the residual fused: red(z+1)
    // yields the red terms of iteration z (its update is discarded when z is
    // the last iteration: red ping-pong), black(z) the black terms of z.  The
    // stopping test is the reference's `while (R > gsTol && z < cap)`
    // (segmentation4D.c:405) on the host, with cap a double as there.
    const SweepMap M = sweep_map();
    const int nb = M.gx * M.gk;
    if ((long)nb > npart_) return check(hipErrorInvalidValue, "presmooth partials");
    int rnc = 0, rch = 0;
    resid_chunking(nb, &rnc, &rch);
    const dim3 grid((unsigned)sweep_grid(M));
    double* part_r = part_;
    double* part_b = part_ + npart_;
    int z = 1;
    double R = 0.0;
    k_heat_sweep<0><<<grid, kBlock, 0, st_>>>(G, M, c, omega, h4, 0, ur_[0], ur_[1], ub_, hp_, nullptr);
    k_heat_sweep<2><<<grid, kBlock, 0, st_>>>(G, M, c, omega, h4, 1, ub_, ub_, ur_[1], hp_, part_b);
    for (;;) {
        k_heat_sweep<1><<<grid, kBlock, 0, st_>>>(G, M, c, omega, h4, 0, ur_[z & 1],
                                                   ur_[(z + 1) & 1], ub_, hp_, part_r);
        k_resid_chunks<<<2 * rnc, 1024, 0, st_>>>(part_r, part_b, nb, rnc, rch, red2_, nullptr);
        k_reduce_sum<false><<<1, 1024, 0, st_>>>(red2_, 2 * rnc, dscal_);
        if (launch_check("presmooth kernels")) return -1;
        if (check(hipMemcpyAsync(hscal_, dscal_, sizeof(double), hipMemcpyDeviceToHost, st_),
                  "scalar copy") ||
            check(hipStreamSynchronize(st_), "presmooth sync"))
            return -1;
        R = hscal_[0];
        if (!(R > tol && z < cap)) break;
        z = z + 1;
        k_heat_sweep<2><<<grid, kBlock, 0, st_>>>(G, M, c, omega, h4, 1, ub_, ub_, ur_[z & 1],
                                                   hp_, part_b);
    }
    // the accepted red values are red(z)'s; every padded point back to d
    k_unpack<<<(unsigned)cdiv(n_, kBlock), kBlock, 0, st_>>>(G, ur_[z & 1], ub_, d, nullptr);
    if (launch_check("k_unpack heat")) return -1;
    if (iters) *iters = z;
    if (Rout) *Rout = R;
    return 0;
}

int Engine::gfun(const double* a, const double* b, double h, double K, double theta,
                 double rho, bool b_finite) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    const Scale S = make_scale(h);
    g_version_++;
    Ptr8 g;
    for (int f = 0; f < 8; f++) {
        g.p[f] = full_[A_G0 + f];
        if (!g.p[f]) return check(hipErrorInvalidValue, "gfun: no g arrays");
    }
    // XCD-resident tiles over the interior planes (runs of kBlock points of
    // the padded (j,i) plane, k - 1, l - 1), shaped as the set-up kernel's
    const PlaneMap PM = plane_map(cdiv((long)I_ * J_, kBlock), K_ - 2, L_ - 2);
    const dim3 grid(plane_grid(PM));
    // the second image's term can be dropped only where it is known to be a
    // signed zero: rho == 0 and every value of b finite
    const bool both = !(rho == 0.0 && b_finite);
    if (is_pow2(h)) {
        if (both) k_gfun<true, true><<<grid, kBlock, 0, st_>>>(G, PM, S, K, theta, rho, a, b, g);
        else k_gfun<true, false><<<grid, kBlock, 0, st_>>>(G, PM, S, K, theta, rho, a, b, g);
    } else {
        if (both) k_gfun<false, true><<<grid, kBlock, 0, st_>>>(G, PM, S, K, theta, rho, a, b, g);
        else k_gfun<false, false><<<grid, kBlock, 0, st_>>>(G, PM, S, K, theta, rho, a, b, g);
    }
    return launch_check("k_gfun");
}

int Engine::l2norm(const double* a, const double* b, double h, double* out) {
    const double h4 = h * h * h * h;
    const int nblk = l2_blocks(n_);  // <= the partial capacity (>= n_ / kBlock)
    k_sqdiff<<<nblk, kBlock, 0, st_>>>(a, b, n_, h4, part_);
    if (launch_check("k_sqdiff")) return -1;
    double S = 0.0;
    if (reduce_to_host(nblk, &S)) return -1;
    *out = std::sqrt(S);
    return 0;
}

// ---- allocation-free _l2norm_ ------------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the block partial sums live in `scratch`, a
// file-scope device array of fixed size, instead of the engine's buffers.
// Every GPU the library runs on holds its own copy until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (kept per device in g_scratch_dev), never through
// the symbol itself.  Nothing in it needs clearing: k_sqdiff writes every
// partial it uses before k_reduce_sum reads it.  Calls that overlap on one
// GPU share the array and race.
namespace {
static __device__ __align__(256) double scratch[kL2MaxPartials];
constexpr int kScratchDevices = 64;
double* g_scratch_dev[kScratchDevices] = {};

int scratch_address(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("seg4dL2normNoAlloc: hipGetDevice: ") + hipGetErrorString(e));
        return -1;
    }
    if (dev < 0 || dev >= kScratchDevices) {
        set_error("seg4dL2normNoAlloc: device index beyond the per-device scratch table");
        return -1;
    }
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("seg4dL2normNoAlloc: hipGetSymbolAddress(scratch): ") +
                      hipGetErrorString(e));
            return -1;
        }
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}
}  // namespace

int l2norm_noalloc(const double* a, const double* b, long n, double h, double* result,
                   hipStream_t st) {
    if (!a || !b || !result) {
        set_error("seg4dL2normNoAlloc: null device pointer");
        return -1;
    }
    if (n < 1 || n > kL2MaxPartials * kL2PerPartial) {
        set_error("seg4dL2normNoAlloc: element count " + std::to_string(n) +
                  " outside 1.." + std::to_string(kL2MaxPartials * kL2PerPartial) +
                  " (the scratch array holds 65536 partial sums of 4096 elements)");
        return -1;
    }
    double* sp = nullptr;
    if (scratch_address(&sp)) return -1;
    const int nblk = l2_blocks(n);
    const double h4 = h * h * h * h;
    k_sqdiff<<<nblk, kBlock, 0, st>>>(a, b, n, h4, sp);
    k_reduce_sum<true><<<1, 1024, 0, st>>>(sp, nblk, result);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("seg4dL2normNoAlloc: launch: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int Engine::mcf_exact(double* d, double h, double t, bool ghosts_only) {
    const Geo G = make_geo(I_, J_, K_, L_, H_);
    dim3 grid(cdiv((long)I_ * J_, kBlock), K_, L_);
    k_mcf_exact<<<grid, kBlock, 0, st_>>>(G, h, t, d, ghosts_only ? 1 : 0);
    return launch_check("k_mcf_exact");
}

int Engine::sync() {
    if (wait_side()) return -1;
    return check(hipStreamSynchronize(st_), "sync");
}

}  // namespace s4d
