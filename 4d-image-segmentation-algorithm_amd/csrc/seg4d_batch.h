// seg4d_batch.h -- the batched form: many volumes of different sizes stepped
// together, each colour pass of all volumes in one launch.
//
// Every volume is an Engine (its own buffers, set-up kernel, warm-start
// history and tuned launch order) on the batch's stream.  The SOR solves,
// FP64 defects, corrections and stopping tests of all volumes run as single
// batched launches over per-volume descriptors; every decision is taken on
// the device per volume, the host only waits for "all volumes stopped".  The
// iterations of a solve can be replayed from a captured HIP graph (two
// iterations per replay), or the whole time step can run as one graph
// (graph mode 2: set-up, the solve's iterations, the correction rounds and
// the epilogue, re-captured every step because the history buffers rotate;
// the iterations it holds follow the previous step's sweep counts, and a
// step whose solve needs more is redone eagerly).  Each volume's result is
// bitwise that of the same step on a single-volume engine with the same
// settings.
//
// The concurrent form (streams > 0) instead splits the volumes into groups
// of about equal voxel count, one stream and one host thread per group, and
// steps each volume on the single-volume path (its own launches and
// pipelined stopping test); the groups' launches overlap on the GPU, which
// hides each solve's launch gaps and pass tails behind the other groups'
// passes.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "seg4d_engine.h"

namespace s4d {

template <typename T>
struct BatchJob;

struct EdgeParams {         // presmoothing + edge indicator of one image
    double sigma = 0.5, h = 1.0, omega = 1.3, gsTol = 0.0, rel2 = 0.0, cap = 100000.0;
    double K = 1000.0, theta = 1.0;
};

class Batch {
public:
    Batch();  // out of line: the job types are complete in seg4d_batch.hip only
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // n volumes of interior sizes dims[4v .. 4v+3] = (nx, ny, nz, nt).
    int init(int n, const int* dims);
    void release();
    int size() const { return (int)E_.size(); }
    Engine& vol(int v) { return *E_[v]; }

    // Host arrays of padded size: the image (its edge indicator g becomes the
    // volume's; presmoothing on the device), u0 and the mask.
    int set_volume(int v, const double* img, const double* u0, const double* mask,
                   const EdgeParams& ep);
    // One time step of every volume (set-up, solve, mask epilogue).
    int step(const StepParams& P);
    // Per-volume tile shapes of the SOR launch order (Engine::autotune).
    int autotune(int reps, std::vector<float>* best_ms);
    int download_u(int v, double* host);

    int precision = 0;      // 0 FP64 (the reference's arithmetic), 1 FP32-refined
    int warm = 0;           // SOR start (Engine::warm_ modes 0, 1, 3, 4)
    int graph = 0;          // 1: solve iterations replayed from a captured graph; 2: the
                            // whole step as one graph
    int depth = 2;          // iterations (graph: replays) queued ahead of the test
    int streams = 0;        // 0: one launch per pass for all volumes; k: k concurrent groups
    int max_rounds = 8;
    double inner_frac = 0.9, inner_floor2 = 1e-10, inner_omega = 1.2;
    std::vector<SolveStats> stats;  // last step, per volume
    double last_step_ms = 0.0;      // device time (concurrent form: host wall time)
    long last_iterations = 0;       // sum over volumes of SOR / FP32 sweeps
    // graph mode 2: iterations per solve and correction rounds the next
    // captured step holds (from the last step: its largest sweep count plus a
    // margin), steps run as one graph and steps redone eagerly after an overrun
    int graph_iters = 0, graph_rounds = 0;
    long graph_steps = 0, graph_redone = 0;
    double graph_host_ms = 0.0;     // host time spent capturing and instantiating, all steps
    std::string err;

private:
    int check(hipError_t e, const char* what);
    int build_jobs(const StepParams& P);
    int step_streams(const StepParams& P);
    int step_eager(const StepParams& P, int cap);
    int step_graph(const StepParams& P, int cap);  // 1: overran, redo eagerly
    template <typename T>
    void enqueue_iters(BatchJob<T>* dj, SorCtl* dctl, int cap, int nit);
    void plan_graph(int max_sweeps, int max_rounds_used);
    int use_streams(int k);  // engines onto k group streams (0: the batch stream)
    template <typename T>
    int sor_run(std::vector<BatchJob<T>>& hj, BatchJob<T>* dj, SorCtl* dctl, SorCtl* hctl,
                int cap, hipGraphExec_t* gexec, int* gcap);

    hipStream_t st_ = nullptr;
    std::vector<std::unique_ptr<Engine>> E_;
    SorCtl* ctl_ = nullptr;   // [n] FP64 control blocks (solve or defect rounds)
    SorCtl* hctl_ = nullptr;  // pinned mirrors
    SorCtl* ctl2_ = nullptr;  // [n] correction-solve control blocks
    SorCtl* hctl2_ = nullptr;
    BatchJob<double>* dj64_ = nullptr;
    BatchJob<float>* dj32_ = nullptr;
    int* vmap_ = nullptr;     // volume of each group of 8 blocks (FP64 launches, then FP32)
    int* vmap32_ = nullptr;   // the FP32 launches' part of vmap_
    long vmap_n_ = 0;
    std::vector<BatchJob<double>> j64_;
    std::vector<BatchJob<float>> j32_;
    int grid_ = 0, grid32_ = 0, cgrid_ = 0;  // FP64 / FP32 SOR launches, element-wise
    int rgrid_ = 0, rgrid32_ = 0;            // chunk-sum launches of the residual reduction
    bool jobs_ok_ = false;
    double jobs_c_ = 0.0, jobs_om_ = 0.0, jobs_om32_ = 0.0, jobs_h4_ = 0.0;
    int jobs_prec_ = -1;
    hipGraphExec_t g64_ = nullptr, g32_ = nullptr;
    hipGraphExec_t gstep_ = nullptr;  // the captured time step (updated every step)
    int* gate_ = nullptr;     // [2 + n] device gate of the captured step (k_gate_*)
    int* hgate_ = nullptr;    // pinned mirror
    int rounds_seen_ = 0;     // most correction rounds of any step so far
    int g64_cap_ = -1, g32_cap_ = -1;
    std::vector<hipEvent_t> evq_;
    std::vector<hipStream_t> gst_;           // group streams of the concurrent form
    std::vector<std::vector<int>> groups_;   // volumes of each group
    int adopted_ = 0;                        // groups the engines run on (0: st_)
};

}  // namespace s4d
