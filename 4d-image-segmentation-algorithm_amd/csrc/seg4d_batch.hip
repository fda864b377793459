// seg4d_batch.hip -- the batched form (see seg4d_batch.h).
#include "seg4d_batch.h"

#include <cmath>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>

#include "seg4d_batch_kernels.h"

namespace s4d {

namespace {
inline int cdivi(long a, long b) { return (int)((a + b - 1) / b); }
}  // namespace

Batch::Batch() = default;
Batch::~Batch() { release(); }

int Batch::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    set_error(err);
    return -1;
}

void Batch::release() {
    if (st_) (void)hipStreamSynchronize(st_);
    if (g64_) (void)hipGraphExecDestroy(g64_);
    if (g32_) (void)hipGraphExecDestroy(g32_);
    if (gstep_) (void)hipGraphExecDestroy(gstep_);
    g64_ = g32_ = gstep_ = nullptr;
    g64_cap_ = g32_cap_ = -1;
    graph_iters = graph_rounds = rounds_seen_ = 0;
    graph_steps = graph_redone = 0;
    graph_host_ms = 0.0;
    if (gate_) (void)hipFree(gate_);
    if (hgate_) (void)hipHostFree(hgate_);
    gate_ = hgate_ = nullptr;
    E_.clear();  // engines release their buffers (the stream is not theirs)
    for (auto s : gst_) (void)hipStreamDestroy(s);
    gst_.clear();
    groups_.clear();
    adopted_ = 0;
    if (ctl_) (void)hipFree(ctl_);
    if (ctl2_) (void)hipFree(ctl2_);
    if (hctl_) (void)hipHostFree(hctl_);
    if (hctl2_) (void)hipHostFree(hctl2_);
    if (dj64_) (void)hipFree(dj64_);
    if (dj32_) (void)hipFree(dj32_);
    if (vmap_) (void)hipFree(vmap_);
    vmap_ = nullptr;
    vmap_n_ = 0;
    ctl_ = ctl2_ = hctl_ = hctl2_ = nullptr;
    dj64_ = nullptr;
    dj32_ = nullptr;
    j64_.clear();
    j32_.clear();
    stats.clear();
    for (auto e : evq_) (void)hipEventDestroy(e);
    evq_.clear();
    jobs_ok_ = false;
    if (st_) (void)hipStreamDestroy(st_);
    st_ = nullptr;
}

int Batch::init(int n, const int* dims) {
    release();
    if (n < 1 || !dims) {
        err = "batch: need at least one volume";
        set_error(err);
        return -1;
    }
    if (check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate")) return -1;
    for (int v = 0; v < n; v++) {
        const int* d = dims + 4 * v;
        E_.emplace_back(new Engine());
        Engine& E = *E_.back();
        // no fused (cooperative) solves: the concurrent form runs several
        // volumes' solves at once, and a cooperative grid assumes the GPU's
        // every CU for itself
        E.fused_ = 0;
        if (E.init(d[0] + 2, d[1] + 2, d[2] + 2, d[3] + 2, true, false) || E.adopt_stream(st_)) {
            err = "batch volume " + std::to_string(v) + ": " + E.err;
            set_error(err);
            return -1;
        }
    }
    const size_t cb = (size_t)n * sizeof(SorCtl);
    if (check(hipMalloc((void**)&ctl_, cb), "hipMalloc ctl") ||
        check(hipMalloc((void**)&ctl2_, cb), "hipMalloc ctl") ||
        check(hipHostMalloc((void**)&hctl_, cb, hipHostMallocDefault), "hipHostMalloc") ||
        check(hipHostMalloc((void**)&hctl2_, cb, hipHostMallocDefault), "hipHostMalloc") ||
        check(hipMalloc((void**)&dj64_, (size_t)n * sizeof(BatchJob<double>)), "hipMalloc jobs") ||
        check(hipMalloc((void**)&dj32_, (size_t)n * sizeof(BatchJob<float>)), "hipMalloc jobs") ||
        check(hipMalloc((void**)&gate_, (size_t)(n + 2) * sizeof(int)), "hipMalloc gate") ||
        check(hipHostMalloc((void**)&hgate_, (size_t)(n + 2) * sizeof(int), hipHostMallocDefault),
              "hipHostMalloc"))
        return -1;
    std::memset(hctl_, 0, cb);
    std::memset(hctl2_, 0, cb);
    if (check(hipMemsetAsync(ctl_, 0, cb, st_), "memset ctl") ||
        check(hipMemsetAsync(ctl2_, 0, cb, st_), "memset ctl"))
        return -1;
    stats.assign(n, SolveStats());
    return check(hipStreamSynchronize(st_), "batch init");
}

int Batch::set_volume(int v, const double* img, const double* u0, const double* mask,
                      const EdgeParams& ep) {
    if (v < 0 || v >= size() || !img || !u0 || !mask) {
        err = "batch: bad volume index or null array";
        set_error(err);
        return -1;
    }
    Engine& E = *E_[v];
    if (E.upload(A_I0, img) || E.upload(A_U, u0) || E.upload(A_MASK, mask)) return -1;
    int z = 0;
    double R = 0.0;
    // the edge indicator of the image alone: rho = 0, so the second
    // image's term vanishes exactly and the image serves as both arguments
    if (E.presmooth(E.full(A_I0), ep.sigma, ep.h, ep.omega, ep.gsTol, ep.rel2, ep.cap, &z, &R) ||
        E.gfun(E.full(A_I0), E.full(A_I0), ep.h, ep.K, ep.theta, 0.0, std::isfinite(R)))
        return -1;
    return E.sync();
}

int Batch::build_jobs(const StepParams& P) {
    const int n = size();
    const double c = P.tau / (P.h * P.h);
    const double h4 = P.h * P.h * P.h * P.h;
    const double om32 = inner_omega > 0.0 ? inner_omega : P.omega;
    if (jobs_ok_ && c == jobs_c_ && P.omega == jobs_om_ && om32 == jobs_om32_ && h4 == jobs_h4_ &&
        precision == jobs_prec_)
        return 0;
    j64_.assign(n, BatchJob<double>());
    j32_.assign(n, BatchJob<float>());
    long b0 = 0, b032 = 0, cb0 = 0, r0 = 0, r032 = 0;
    for (int v = 0; v < n; v++) {
        Engine& E = *E_[v];
        if (precision == 1 && E.ensure_mixed()) return -1;
        // the launch maps of the volume's own FP64 and FP32 passes
        const SweepMap M = E.sweep_map(8), M32 = E.sweep_map(4);
        const int nbk = sweep_grid(M);
        const int nbk32 = sweep_grid(M32);
        const int nbp = M.gx * M.gk, nbp32 = M32.gx * M32.gk;
        if ((long)nbp > E.npart_ || (long)nbp32 > E.npart_) {
            err = "batch: SOR grid exceeds a volume's partial buffer";
            set_error(err);
            return -1;
        }
        const int cnb = std::min(cdivi(E.np_, kBlock), 2048);
        BatchJob<double>& A = j64_[v];
        BatchJob<float>& B = j32_[v];
        A.G = B.G = E.geometry();
        A.M = M;
        B.M = M32;
        A.block0 = (int)b0;
        B.block0 = (int)b032;
        A.nblocks = nbk;
        B.nblocks = nbk32;
        A.cblock0 = B.cblock0 = (int)cb0;
        A.cnblocks = B.cnblocks = cnb;
        b0 += ((nbk + 7) / 8) * 8;  // volumes start on XCD 0
        b032 += ((nbk32 + 7) / 8) * 8;
        cb0 += cnb;
        A.c = c;
        A.om = P.omega;
        B.c = (float)c;
        B.om = (float)om32;
        A.h4 = B.h4 = h4;
        A.opnd[0] = E.op_[0];
        A.opnd[1] = E.op_[1];
        A.red[0] = E.ur_[0];
        A.red[1] = E.ur_[1];
        A.black = E.ub_;
        A.part_r = B.part_r = E.part_;
        A.part_b = B.part_b = E.part_ + E.npart_;
        A.nb = nbp;
        B.nb = nbp32;
        A.red2 = B.red2 = E.red2_;
        resid_chunking(nbp, &A.nc, &A.ch);
        resid_chunking(nbp32, &B.nc, &B.ch);
        A.r1block0 = (int)r0;
        B.r1block0 = (int)r032;
        r0 += 2 * A.nc;
        r032 += 2 * B.nc;
        A.np = B.np = E.np_;
        A.ctl = ctl_ + v;
        B.ctl = ctl2_ + v;
        A.defect_ob[0] = E.op32_[0];
        A.defect_ob[1] = E.op32_[1];
        B.defect_ob[0] = B.defect_ob[1] = nullptr;
        B.opnd[0] = E.op32_[0];
        B.opnd[1] = E.op32_[1];
        B.red[0] = E.dr_[0];
        B.red[1] = E.dr_[1];
        B.black = E.db_;
    }
    if (b0 >= (1L << 31) || b032 >= (1L << 31) || cb0 >= (1L << 31)) {
        err = "batch: grid too large";
        set_error(err);
        return -1;
    }
    grid_ = (int)b0;
    grid32_ = (int)b032;
    // volume of each group of 8 blocks: FP64 launches, then FP32 launches
    const long nm = b0 / 8 + b032 / 8;
    std::vector<int> vm((size_t)nm);
    for (int v = 0; v < n; v++) {
        const long g0 = j64_[v].block0 / 8, g1 = (j64_[v].block0 + j64_[v].nblocks + 7) / 8;
        for (long q = g0; q < g1; q++) vm[(size_t)q] = v;
        const long f0 = j32_[v].block0 / 8, f1 = (j32_[v].block0 + j32_[v].nblocks + 7) / 8;
        for (long q = f0; q < f1; q++) vm[(size_t)(b0 / 8 + q)] = v;
    }
    if (nm > vmap_n_) {
        if (vmap_) (void)hipFree(vmap_);
        vmap_ = nullptr;
        vmap_n_ = 0;
        if (check(hipMalloc((void**)&vmap_, (size_t)nm * sizeof(int)), "hipMalloc volume map"))
            return -1;
        vmap_n_ = nm;
    }
    vmap32_ = vmap_ + b0 / 8;
    rgrid_ = (int)r0;
    rgrid32_ = (int)r032;
    cgrid_ = (int)cb0;
    if (check(hipMemcpyAsync(vmap_, vm.data(), (size_t)nm * sizeof(int), hipMemcpyHostToDevice,
                             st_), "volume map") ||
        check(hipMemcpyAsync(dj64_, j64_.data(), (size_t)n * sizeof(BatchJob<double>),
                             hipMemcpyHostToDevice, st_), "jobs") ||
        check(hipMemcpyAsync(dj32_, j32_.data(), (size_t)n * sizeof(BatchJob<float>),
                             hipMemcpyHostToDevice, st_), "jobs") ||
        check(hipStreamSynchronize(st_), "jobs"))
        return -1;
    // captured graphs hold the old grid sizes
    if (g64_) (void)hipGraphExecDestroy(g64_);
    if (g32_) (void)hipGraphExecDestroy(g32_);
    if (gstep_) (void)hipGraphExecDestroy(gstep_);
    g64_ = g32_ = gstep_ = nullptr;
    g64_cap_ = g32_cap_ = -1;
    jobs_ok_ = true;
    jobs_c_ = c;
    jobs_om_ = P.omega;
    jobs_om32_ = om32;
    jobs_h4_ = h4;
    jobs_prec_ = precision;
    return 0;
}

// Batched red-black SOR of every volume of `hj` to its stopping test; the
// loop structure is the single-volume solve's (red(z+1) yields the red
// residual terms of z, reduce(z), black(z+1) a no-op once z stopped), with
// one launch per pass for all volumes and the decision per volume.
template <typename T>
int Batch::sor_run(std::vector<BatchJob<T>>& hj, BatchJob<T>* dj, SorCtl* dctl, SorCtl* hctl,
                   int cap, hipGraphExec_t* gexec, int* gcap) {
    (void)hj;
    const int n = size();
    const dim3 grid((unsigned)(sizeof(T) == 4 ? grid32_ : grid_));
    const int* vm = sizeof(T) == 4 ? vmap32_ : vmap_;
    int rc = 0;
    auto red = [&](int zr, int zw, bool resid) {
        if (resid) k_sor_batch<T, 1, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 0, zr, zw);
        else k_sor_batch<T, 0, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 0, zr, zw);
    };
    auto black = [&](int zr) {
        k_sor_batch<T, 2, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 1, zr, 0);
    };
    // iteration z + 1 and the stopping test of iteration z
    const dim3 rgrid((unsigned)(sizeof(T) == 4 ? rgrid32_ : rgrid_));
    auto iter = [&](int z) {
        red(z & 1, (z + 1) & 1, true);
        k_resid_chunks_batch<T><<<rgrid, 1024, 0, st_>>>(dj, n);
        k_reduce_batch<T><<<n, 1024, 0, st_>>>(dj, cap);
        rc |= check(hipMemcpyAsync(hctl, dctl, (size_t)n * sizeof(SorCtl), hipMemcpyDeviceToHost,
                                   st_), "ctl copy");
        black((z + 1) & 1);
    };
    const bool replay = graph == 1;
    if (replay && (*gexec == nullptr || *gcap != cap)) {
        if (*gexec) (void)hipGraphExecDestroy(*gexec);
        *gexec = nullptr;
        // two iterations from an odd z: every replay starts at an odd
        // iteration, so the ping-pong parities of the capture hold
        hipGraph_t g = nullptr;
        if (check(hipStreamSynchronize(st_), "capture sync") ||
            check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture"))
            return -1;
        iter(1);
        iter(2);
        const hipError_t ec = hipStreamEndCapture(st_, &g);
        if (check(ec, "end capture") || rc) return -1;
        if (check(hipGetLastError(), "captured launches")) return -1;
        const hipError_t ei = hipGraphInstantiateWithFlags(gexec, g, 0);
        (void)hipGraphDestroy(g);
        if (check(ei, "graph instantiate")) return -1;
        *gcap = cap;
    }
    red(0, 1, false);
    black(1);
    if (check(hipGetLastError(), "batched pass")) return -1;
    const int per = replay ? 2 : 1;  // iterations per queued unit
    const int dq = depth > 0 ? depth : 1;
    while ((int)evq_.size() < dq) {
        hipEvent_t e;
        if (check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event")) return -1;
        evq_.push_back(e);
    }
    int z = 1;
    long unit = 0;
    int queued = 0;
    for (;;) {
        while (queued < dq) {
            if (replay) rc |= check(hipGraphLaunch(*gexec, st_), "graph launch");
            else iter(z);
            rc |= check(hipGetLastError(), "batched iteration");
            rc |= check(hipEventRecord(evq_[unit % dq], st_), "event");
            if (rc) return -1;
            z += per;
            unit++;
            queued++;
        }
        hipError_t q;
        while ((q = hipEventQuery(evq_[(unit - queued) % dq])) == hipErrorNotReady) {
        }
        if (check(q, "batch solve wait")) return -1;
        queued--;
        bool all = true;
        for (int v = 0; v < n && all; v++) all = hctl[v].stop != 0;
        if (all) break;
    }
    return check(hipStreamSynchronize(st_), "batch solve sync");
}

int Batch::use_streams(int k) {
    const int n = size();
    if (k > n) k = n;
    if (k == adopted_) return 0;
    if (k == 0) {
        for (auto& e : E_)
            if (e->adopt_stream(st_)) {
                err = e->err;
                set_error(err);
                return -1;
            }
        adopted_ = 0;
        return 0;
    }
    while ((int)gst_.size() < k) {
        hipStream_t s = nullptr;
        if (check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return -1;
        gst_.push_back(s);
    }
    // largest volumes first, each to the group with the fewest voxels so far
    std::vector<int> order(n);
    for (int v = 0; v < n; v++) order[v] = v;
    auto vox = [&](int v) { return (long)E_[v]->I_ * E_[v]->J_ * E_[v]->K_ * E_[v]->L_; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return vox(a) > vox(b); });
    groups_.assign(k, std::vector<int>());
    std::vector<long> load(k, 0);
    for (int v : order) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        groups_[g].push_back(v);
        load[g] += vox(v);
    }
    for (int g = 0; g < k; g++)
        for (int v : groups_[g])
            if (E_[v]->adopt_stream(gst_[g])) {
                err = E_[v]->err;
                set_error(err);
                return -1;
            }
    adopted_ = k;
    return 0;
}

int Batch::step_streams(const StepParams& P) {
    const int n = size();
    if (use_streams(streams)) return -1;
    for (auto& e : E_) {
        e->max_rounds_ = max_rounds;
        e->inner_frac_ = inner_frac;
        e->inner_floor2_ = inner_floor2;
        e->inner_omega_ = inner_omega;
    }
    int dev = 0;
    if (check(hipGetDevice(&dev), "hipGetDevice")) return -1;
    const int k = adopted_;
    std::vector<std::string> errs(k);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    th.reserve(k);
    for (int g = 0; g < k; g++)
        th.emplace_back([&, g] {
            if (hipSetDevice(dev) != hipSuccess) {
                errs[g] = "batch: hipSetDevice in a group thread";
                return;
            }
            for (int v : groups_[g])
                if (E_[v]->seg_step(P, nullptr, &stats[v])) {
                    errs[g] = "batch volume " + std::to_string(v) + ": " + E_[v]->err;
                    return;
                }
        });
    for (auto& t : th) t.join();
    last_step_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    last_iterations = 0;
    for (int v = 0; v < n; v++) last_iterations += stats[v].z;
    for (int g = 0; g < k; g++)
        if (!errs[g].empty()) {
            err = errs[g];
            set_error(err);
            return -1;
        }
    return 0;
}

int Batch::step(const StepParams& P) {
    const int n = size();
    if (n == 0) {
        err = "batch: not initialised";
        set_error(err);
        return -1;
    }
    for (auto& e : E_) {
        if (e->precision_ != precision || e->warm_ != warm) graph_iters = rounds_seen_ = 0;
        e->precision_ = precision;
        e->warm_ = warm;
    }
    if (streams > 0) return step_streams(P);
    if (use_streams(0)) return -1;
    if (build_jobs(P)) return -1;
    const int cap = P.cap > 0x7fffffffL ? 0x7fffffff : (int)P.cap;
    bool profiled = false;
    for (auto& e : E_) profiled = profiled || e->prof.on;
    if (graph == 2 && graph_iters > 0 && !profiled) {
        const int r = step_graph(P, cap);
        if (r <= 0) return r;
        graph_redone++;  // a solve needed more than the graph held: the step again, eagerly
    }
    return step_eager(P, cap);
}

// Iterations per solve and correction rounds of the next captured step: the
// last step's largest sweep count plus a margin of max(2, count / 16), and
// the most rounds any step of this batch has needed (a spare round costs
// about as much as its spare iterations; a second round, once needed, tends
// to recur).  Sweep counts fall or stay from step to step, with rare jumps;
// a spare iteration is four launches whose blocks each load a stop flag and
// return (about 0.3 ms at 32 volumes of 355.7 M voxels), an overrun one step
// redone eagerly, so the margin only covers small rises.
void Batch::plan_graph(int max_sweeps, int max_rounds_used) {
    const int s = max_sweeps > 1 ? max_sweeps : 1;
    graph_iters = s + std::max(2, s / 16);
    rounds_seen_ = std::max(rounds_seen_, max_rounds_used);
    graph_rounds = std::max(1, rounds_seen_);
}

// Iterations z = 1 .. nit of a solve after its first red and black pass,
// enqueued without host copies, so that they can be captured: the stopping
// test of each volume stays on the device and passes after it return at once.
template <typename T>
void Batch::enqueue_iters(BatchJob<T>* dj, SorCtl* dctl, int cap, int nit) {
    const int n = size();
    const dim3 grid((unsigned)(sizeof(T) == 4 ? grid32_ : grid_));
    const dim3 rgrid((unsigned)(sizeof(T) == 4 ? rgrid32_ : rgrid_));
    const int* vm = sizeof(T) == 4 ? vmap32_ : vmap_;
    k_sor_batch<T, 0, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 0, 0, 1);
    k_sor_batch<T, 2, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 1, 1, 0);
    for (int z = 1; z <= nit; z++) {
        k_sor_batch<T, 1, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 0, z & 1, (z + 1) & 1);
        k_resid_chunks_batch<T><<<rgrid, 1024, 0, st_>>>(dj, n);
        k_reduce_batch<T><<<n, 1024, 0, st_>>>(dj, cap);
        k_sor_batch<T, 2, 2><<<grid, kBlock, 0, st_>>>(dj, vm, dctl, 1, (z + 1) & 1, 0);
    }
}

// The whole time step of every volume as one graph: the set-ups, the solve
// (FP64: graph_iters iterations; FP32-refined: the first decision and
// graph_rounds correction rounds of graph_iters FP32 iterations each), the
// gate and the gated epilogue, and the copies of the control blocks to the
// host.  The history buffers rotate every step, so the step is captured
// every step and the instantiated graph updated.  Returns 0 (done), 1 (a
// solve overran: nothing was written to u and the rotation is undone, so the
// step can be redone) or -1 (error).
int Batch::step_graph(const StepParams& P, int cap) {
    const int n = size();
    const int nit = std::min(graph_iters, cap);
    const auto h0 = std::chrono::steady_clock::now();
    if (check(hipStreamSynchronize(st_), "capture sync") ||
        check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture step"))
        return -1;
    int rc = 0;
    k_gate_begin<<<1, 64, 0, st_>>>(gate_, n);
    for (int v = 0; v < n && !rc; v++)
        if (E_[v]->setup(P, true, true, ctl_ + v, precision == 1)) {
            err = E_[v]->err;
            rc = -1;
        }
    if (!rc && precision == 0) {
        enqueue_iters<double>(dj64_, ctl_, cap, nit);
    } else if (!rc) {
        const dim3 grid((unsigned)grid_), cgrid((unsigned)cgrid_);
        const unsigned vb = (unsigned)((n + 63) / 64);
        k_decide_batch<<<vb, 64, 0, st_>>>(dj64_, n);
        for (int r = 0; r < graph_rounds; r++) {
            k_zero_batch<<<cgrid, kBlock, 0, st_>>>(dj64_, dj32_, n);
            k_inner_ctl_batch<<<vb, 64, 0, st_>>>(dj64_, dj32_, n, inner_frac, inner_floor2);
            enqueue_iters<float>(dj32_, ctl2_, cap, nit);
            k_gate_inner<<<vb, 64, 0, st_>>>(dj64_, dj32_, n, gate_);
            k_defect_corr_batch<<<grid, kBlock, 0, st_>>>(dj64_, dj32_, vmap_, ctl_, 0);
            k_defect_corr_batch<<<grid, kBlock, 0, st_>>>(dj64_, dj32_, vmap_, ctl_, 1);
            k_resid_chunks_batch<double><<<(unsigned)rgrid_, 1024, 0, st_>>>(dj64_, n);
            k_defect_decide<<<n, 1024, 0, st_>>>(dj64_, max_rounds);
        }
    }
    if (!rc) {
        k_gate_done<<<1, 256, 0, st_>>>(dj64_, n, gate_);
        for (int v = 0; v < n && !rc; v++)
            if (E_[v]->unpack_gated(E_[v]->full(A_MASK), precision == 0 ? ctl_ + v : nullptr,
                                    gate_)) {
                err = E_[v]->err;
                rc = -1;
            }
    }
    if (!rc)
        rc = check(hipMemcpyAsync(hctl_, ctl_, (size_t)n * sizeof(SorCtl), hipMemcpyDeviceToHost,
                                  st_), "ctl copy") |
             check(hipMemcpyAsync(hgate_, gate_, (size_t)(n + 2) * sizeof(int),
                                  hipMemcpyDeviceToHost, st_), "gate copy") |
             check(hipGetLastError(), "captured step");
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st_, &g);
    if (rc || check(ec, "end capture")) {
        if (g) (void)hipGraphDestroy(g);
        for (auto& e : E_) e->undo_rotation(true);  // nothing ran
        return -1;
    }
    if (gstep_) {
        hipGraphNode_t en = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(gstep_, g, &en, &res) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(gstep_);
            gstep_ = nullptr;
        }
    }
    if (!gstep_) {
        const hipError_t ei = hipGraphInstantiateWithFlags(&gstep_, g, 0);
        if (ei != hipSuccess) gstep_ = nullptr;
        rc = check(ei, "step graph instantiate");
    }
    (void)hipGraphDestroy(g);
    graph_host_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = check(hipEventCreate(&e0), "event") | check(hipEventCreate(&e1), "event");
    if (!rc) rc = check(hipEventRecord(e0, st_), "event");
    if (!rc) rc = check(hipGraphLaunch(gstep_, st_), "step graph launch");
    if (!rc) rc = check(hipEventRecord(e1, st_), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "step graph wait");
    float ms = 0.f;
    if (!rc) rc = check(hipEventElapsedTime(&ms, e0, e1), "event");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        for (auto& e : E_) e->undo_rotation();  // the epilogue may have run
        return -1;
    }
    if (hgate_[1] == 0) {
        for (auto& e : E_) e->undo_rotation(true);  // the gated epilogue wrote nothing
        return 1;
    }
    last_step_ms = ms;
    last_iterations = 0;
    int ms_max = 0, mr_max = 0;
    for (int v = 0; v < n; v++) {
        const int z = precision == 0 ? hctl_[v].zfinal : hgate_[2 + v];
        stats[v].z = z;
        stats[v].R = hctl_[v].R;
        stats[v].tol = hctl_[v].tol;
        stats[v].sumsq = hctl_[v].sumsq;
        stats[v].rounds = precision == 0 ? 0 : hctl_[v].z;
        E_[v]->red_final_ = precision == 0 ? (hctl_[v].zfinal & 1) : 0;
        last_iterations += z;
        ms_max = std::max(ms_max, z);
        mr_max = std::max(mr_max, stats[v].rounds);
    }
    for (auto& e : E_) e->step_done();
    graph_steps++;
    plan_graph(ms_max, mr_max);
    return 0;
}

int Batch::step_eager(const StepParams& P, int cap) {
    const int n = size();
    int max_round_sweeps = 0, max_rounds_used = 0;
    hipEvent_t e0, e1;
    if (check(hipEventCreate(&e0), "event")) return -1;
    if (check(hipEventCreate(&e1), "event")) {
        (void)hipEventDestroy(e0);
        return -1;
    }
    int rc = check(hipEventRecord(e0, st_), "event");
    for (int v = 0; v < n && !rc; v++)
        if (E_[v]->setup(P, true, true, ctl_ + v, precision == 1)) {
            err = E_[v]->err;
            rc = -1;
        }
    last_iterations = 0;
    if (!rc && precision == 0) {
        rc = sor_run<double>(j64_, dj64_, ctl_, hctl_, cap, &g64_, &g64_cap_);
        for (int v = 0; v < n && !rc; v++) {
            stats[v].z = hctl_[v].zfinal;
            stats[v].R = hctl_[v].R;
            stats[v].tol = hctl_[v].tol;
            stats[v].sumsq = hctl_[v].sumsq;
            stats[v].rounds = 0;
            E_[v]->red_final_ = hctl_[v].zfinal & 1;
            last_iterations += hctl_[v].zfinal;
            max_round_sweeps = std::max(max_round_sweeps, hctl_[v].zfinal);
        }
    } else if (!rc) {
        const dim3 grid((unsigned)grid_), cgrid((unsigned)cgrid_);
        std::vector<int> sweeps(n, 0);
        // the first defect was formed by each volume's set-up kernel; later
        // ones follow a correction (u += d fused into the defect)
        auto sync_ctl = [&]() {
            int r = check(hipGetLastError(), "batched defect");
            r |= check(hipMemcpyAsync(hctl_, ctl_, (size_t)n * sizeof(SorCtl),
                                      hipMemcpyDeviceToHost, st_), "ctl copy");
            return r | check(hipStreamSynchronize(st_), "defect sync");
        };
        auto defect = [&]() {
            k_defect_corr_batch<<<grid, kBlock, 0, st_>>>(dj64_, dj32_, vmap_, ctl_, 0);
            k_defect_corr_batch<<<grid, kBlock, 0, st_>>>(dj64_, dj32_, vmap_, ctl_, 1);
            k_resid_chunks_batch<double><<<(unsigned)rgrid_, 1024, 0, st_>>>(dj64_, n);
            k_defect_decide<<<n, 1024, 0, st_>>>(dj64_, max_rounds);
            return sync_ctl();
        };
        k_decide_batch<<<(n + 63) / 64, 64, 0, st_>>>(dj64_, n);
        rc = sync_ctl();
        for (;;) {
            if (rc) break;
            std::vector<char> active(n);
            bool any = false;
            for (int v = 0; v < n; v++) {
                active[v] = hctl_[v].stop == 0;
                any = any || active[v];
            }
            if (!any) break;
            k_zero_batch<<<cgrid, kBlock, 0, st_>>>(dj64_, dj32_, n);
            k_inner_ctl_batch<<<(n + 63) / 64, 64, 0, st_>>>(dj64_, dj32_, n, inner_frac,
                                                            inner_floor2);
            rc = check(hipGetLastError(), "batched round set-up");
            if (!rc) rc = sor_run<float>(j32_, dj32_, ctl2_, hctl2_, cap, &g32_, &g32_cap_);
            if (rc) break;
            for (int v = 0; v < n; v++)
                if (active[v]) {
                    sweeps[v] += hctl2_[v].zfinal;
                    max_round_sweeps = std::max(max_round_sweeps, hctl2_[v].zfinal);
                }
            rc = defect();
        }
        for (int v = 0; v < n && !rc; v++) {
            stats[v].z = sweeps[v];
            stats[v].R = hctl_[v].R;
            stats[v].tol = hctl_[v].tol;
            stats[v].sumsq = hctl_[v].sumsq;
            stats[v].rounds = hctl_[v].z;
            E_[v]->red_final_ = 0;
            last_iterations += sweeps[v];
            max_rounds_used = std::max(max_rounds_used, hctl_[v].z);
        }
    }
    for (int v = 0; v < n && !rc; v++)
        if (E_[v]->unpack(E_[v]->full(A_MASK))) {
            err = E_[v]->err;
            rc = -1;
        }
    if (!rc) rc = check(hipEventRecord(e1, st_), "event");
    if (!rc) rc = check(hipEventSynchronize(e1), "event");
    float ms = 0.f;
    if (!rc) rc = check(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        for (auto& e : E_) e->undo_rotation();
        return -1;
    }
    last_step_ms = ms;
    for (auto& e : E_) e->step_done();
    if (graph == 2) plan_graph(max_round_sweeps, max_rounds_used);
    return 0;
}

int Batch::autotune(int reps, std::vector<float>* best_ms) {
    if (best_ms) best_ms->assign(size(), 0.f);
    for (int v = 0; v < size(); v++) {
        float ms = 0.f;
        if (E_[v]->autotune(reps, &ms)) {
            err = E_[v]->err;
            return -1;
        }
        if (best_ms) (*best_ms)[v] = ms;
    }
    jobs_ok_ = false;
    return 0;
}

int Batch::download_u(int v, double* host) {
    if (v < 0 || v >= size() || !host) {
        err = "batch: bad volume index or null array";
        set_error(err);
        return -1;
    }
    return E_[v]->download(A_U, host);
}

}  // namespace s4d
