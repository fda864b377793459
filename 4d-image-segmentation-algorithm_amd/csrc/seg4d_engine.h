// seg4d_engine.h -- device state and operations of one padded 4D volume.
//
// The engine owns every device buffer of the time step and launches the
// kernels of seg4d_kernels.h on one HIP stream.  It knows nothing about the
// reference's globals; the C-ABI layers (seg4d_abi.cpp, mcf4d_abi.cpp) map
// those onto an engine.  All entry points return 0 on success and -1 on
// failure with a message in Engine::err.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace s4d {

struct Geo;
struct SorCtl;
struct GridBar;
struct SweepMap;
struct BallList;
class Net;

struct StepParams {
    double h = 1.0, tau = 0.5, eps2 = 1e-6;
    double omega = 1.8;     // SOR relaxation
    double gsTol = 1e-12;   // absolute bound on sum r^2 h^4
    double rel2 = 0.0;      // > 0: bound = rel2 * sum(prev^2 h^4) instead
    long cap = 1000;        // sweep cap (maxNoGSIteration)
};

struct SolveStats {
    int z = 0;              // SOR iterations performed (FP32-refined: FP32 sweeps)
    double R = 0.0;         // residual at exit (FP64 system)
    double tol = 0.0;       // bound used
    double sumsq = 0.0;     // sum(prev^2 h^4) over the interior
    int rounds = 0;         // FP32-refined: defect-correction rounds
};

struct Balls {              // ball centres per frame (host copies)
    std::vector<int> count, offset;
    std::vector<double> cx, cy, cz;
    double radius = 0.0;
    int total = 0;
};

struct Profile {            // HIP-event timing (events on the engine's stream)
    bool on = false;
    // SOR colour passes by kernel variant: 0 = red without residual (first
    // sweep of a solve), 1 = red with the fused residual, 2 = black with the
    // fused residual (the last black launch of a solve is a no-op); 3..5 the
    // same for the FP32 correction solves.
    double sweep_ms[6] = {0, 0, 0, 0, 0, 0};
    long sweep_launches[6] = {0, 0, 0, 0, 0, 0};
    double defect_ms = 0.0; // FP64 defect passes of the FP32-refined form
    long defect_launches = 0;
    long rounds = 0;
    double setup_ms = 0.0;  // coefficient / set-up kernel
    long setup_launches = 0;
    double step_ms = 0.0;   // whole time steps (setup + solve + epilogue)
    long steps = 0;
    long voxel_updates = 0; // interior voxels x reference iterations
    long iterations = 0;
    long nred = 0, nblack = 0;  // interior points per colour
    double copy_ms = 0.0;   // k_copy_f64 (a plain stream of known size)
    long copy_launches = 0;
};

enum ArrayId {
    A_U = 0, A_PREV, A_I0, A_MASK, A_EXACT,
    A_G0, A_G7 = A_G0 + 7,
    A_TMP, A_TMP2,
    A_COUNT
};

class Engine {
public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Allocates device arrays for padded dims I,J,K,L.  `seg` adds i0, Mmask
    // and the 8 g arrays; `mcf` adds exact_u.  Arrays start zeroed.
    int init(int I, int J, int K, int L, bool seg, bool mcf);
    void release();
    bool ready() const { return ready_; }

    double* full(ArrayId a) const { return full_[a]; }
    long elems() const { return n_; }
    int dims(int d) const { return d == 0 ? I_ : d == 1 ? J_ : d == 2 ? K_ : L_; }

    int upload(ArrayId a, const double* host);
    int download(ArrayId a, double* host);
    // Upload on a second stream (default priority) that overlaps the work
    // queued on the engine's stream; the next consumer (the step's epilogue,
    // any other transfer, sync) waits for it.  `host` must stay unchanged
    // until finish_side() has returned.
    int upload_side(ArrayId a, const double* host);
    int finish_side();  // the side copy has completed (host buffer free again)
    int copy(ArrayId src, ArrayId dst);
    // Coefficient face f (0..7, order e w n s t b fw bw) in full layout.
    int download_coef(int f, double* host);

    // --- time step pieces -------------------------------------------------
    // Coefficients from u (and g if with_g) into the colour layout, u packed
    // into the colour arrays, and (if write_prev) prev = u.
    // fuse_defect (FP32-refined form): the set-up kernel also forms the first
    // FP64 defect (b - A u^n) and the start's offset as the initial correction.
    int setup(const StepParams& P, bool with_g, bool write_prev, SorCtl* ctl_target = nullptr,
              bool fuse_defect = false);
    // Red-black SOR from the packed state to the stopping test.
    int solve(const StepParams& P, SolveStats* st);
    // Colour layout -> u (interior), with the Mmask pass fused if mask.
    int unpack(const double* dmask);
    // Test hook of the learned start: the network's correction c of the next
    // solve's start (FP32, full layout, interior points) and its scale q, as
    // the set-up kernel would receive them; needs the history the network's
    // base uses.  Leaves the state unchanged.
    int learned_preview(float* host_c, double* host_q);
    int rescale(double* d, const Balls& balls, bool apply_mask);

    // --- initialisation (segmentation4D.c:876-1148) -----------------------------
    // thr = global interior max of `in` where `in` exceeds the ball's
    // alpha*min + beta*max, inside every ball (thr may alias in).
    int threshold(const double* in, double* thr, const Balls& balls, double alpha,
                  double beta);
    // Seed function max(u, |1/(d+1/2) - 1/12.5|) and mask = 1 around centres.
    int seed(double* d, double* mask, const Balls& balls);
    int max_interior(const double* d, double* out);
    int ball_minmax(const double* d, int l, double cx, double cy, double cz, double R,
                    double* mn, double* mx);
    // sum of d h^4 over the whole padded array
    int mass(const double* d, double h, double* out);
    // Lazily allocated full-layout scratch array (A_TMP or A_TMP2).
    double* scratch(ArrayId a);
    // Whole segmentation step: setup, solve, unpack, rescale epilogue.
    int seg_step(const StepParams& P, const Balls* balls, SolveStats* st);
    // MCF step: setup (g == 1), exact boundary at time t, solve, unpack.
    int mcf_step(const StepParams& P, double t, SolveStats* st);

    // --- edge indicator -----------------------------------------------------
    int reflect(double* d);
    // Implicit heat step in place on device array d (prev receives input).
    int presmooth(double* d, double sigma, double h, double omega, double gsTol,
                  double rel2, double cap, int* iters, double* R);
    // 8 face g's from the presmoothed images a and b.  b_finite: every value
    // of b is known finite (e.g. its presmoothing residual was finite: the
    // ghosts are copies of interior values and a non-finite interior value
    // makes the residual non-finite), which lets rho == 0 skip b.
    int gfun(const double* a, const double* b, double h, double K, double theta,
             double rho, bool b_finite = false);

    // --- helpers --------------------------------------------------------------
    int l2norm(const double* a, const double* b, double h, double* out);
    int mcf_exact(double* d, double h, double t, bool ghosts_only);
    int sync();

    hipStream_t stream() const { return st_; }
    // Runs this engine's work on stream s (not owned) instead of its own.
    int adopt_stream(hipStream_t s);
    // Completed time step: the history levels the setup filled become valid.
    void step_done();
    // puts the buffer pointers back as they were before setup() rotated
    // them (a step that failed after its set-up leaves u = u^n).
    // history_intact: nothing wrote the recycled buffer (no k_unpack ran), so
    // the history levels stay valid; otherwise they are dropped.
    void undo_rotation(bool history_intact = false);
    // k_unpack behind a device gate (gate[1] != 0) with the red buffer chosen
    // on the device from sel->zfinal's parity (sel null: the first buffer):
    // the epilogue of a captured step, whose solve ends on the device.
    int unpack_gated(const double* dmask, const SorCtl* sel, const int* gate);
    struct Geo geometry() const;
    // Times the red pass (FP64, no residual) over tile shapes of the tiled
    // order and the linear order, `reps` launches each, and keeps the
    // fastest; *best_ms receives its time.  Results are unchanged by the
    // choice; the pass writes only the second red buffer, whose interior
    // every solve rewrites first.
    int autotune(int reps, float* best_ms);
    // SOR pass kernel variant (tuning knob, identical results), bit flags:
    // 1 = two l-steps per load phase, 2 = tiled XCD-resident launch order,
    // 4 = nontemporal operand-block loads, 8 = two k planes per block,
    // 16 = four k planes per block, 32 = automatic order: with automatic
    // frame chunks (sweep_lc_ = 0), chunks of 2 frames in the linear order
    // whenever the other colour's sliding window (3 k planes x 4 frames)
    // fits in an XCD's L2, which replaces bit 2 and the chunk rule.
    int sweep_variant_ = 35;
    // tile of the tiled order in blocks (runs x k groups); 16 x 8 measured
    // fastest at 256x256x128x16 (8 x 32: 1.401 ms, 16 x 16: 1.352 ms, 16 x 8:
    // 1.347 ms per red pass)
    int sweep_tb_ = 16, sweep_tk_ = 8;
    int sweep_lc_ = 0;     // l steps per block of a colour pass; 0 = automatic
    int sweep_nband_ = 0;  // bands of the automatic linear order; 0 = as the L2 window needs
    int pipe_depth_ = 0;   // iterations enqueued ahead of the stopping test; 0 = automatic
    std::vector<hipEvent_t> evq_;  // one event per queued iteration
    // Fused sweeps: a whole SOR solve in one cooperative launch (k_sor_fused,
    // grid barriers between the passes).  0 = off (one launch per pass),
    // 1 = whenever the grid fits on the GPU.  Identical results either way.
    // Off by default: measured slower at every size (a grid barrier with its
    // agent-scope release and acquire in every block costs ~40 us, a whole
    // iteration of the pass-per-launch solve of 64x64x32x8 36 us).
    int fused_ = 0;
    bool fused_last_ = false;      // the last SOR solve ran fused
    GridBar* gbar_ = nullptr;      // the fused launch's barrier state
    GridBar* hgbar_ = nullptr;     // pinned mirror (abort flag)
    int fused_blocks_[2] = {0, 0}; // co-resident blocks of k_sor_fused<double / float>
                                   // (0: not yet known, -1: the launch never fits)
    // 0 = FP64 (the reference's arithmetic); 1 = FP32-refined: FP32 SOR on
    // the correction equation inside FP64 defect correction.
    int precision_ = 0;
    int max_rounds_ = 8;           // FP32-refined: correction rounds per solve
    // SOR start: 0 = u^n (the reference); 1 = 2 u^n - u^(n-1) when the
    // previous call was a completed time step on the device (hist_ok_);
    // 2 = the prediction of net_ (learned warm start); 3 = 3 u^n - 3 u^(n-1)
    // + u^(n-2) when two completed steps precede (hist2_ok_), else as 1;
    // 4 = 4 u^n - 6 u^(n-1) + 4 u^(n-2) - u^(n-3) after three, else as 3.
    int warm_ = 0;
    bool hist_ok_ = false;
    bool hist2_ok_ = false, hist3_ok_ = false;
    bool shifted2_ = false, shifted3_ = false;  // after this step prev2_ / prev3_ hold valid levels
    bool rotated_ = false;                       // setup() rotated; step_done() not yet called
    double* rot_[4] = {};                        // u, prev_full, prev2_, prev3_ before it
    double* hp_ = nullptr;         // presmoothing right-hand side, colour-packed (lazy)
    double* prev2_ = nullptr;      // u^(n-2) for the quadratic and cubic starts
    double* prev3_ = nullptr;      // u^(n-3) for the cubic start
    Net* net_ = nullptr;           // not owned
    long g_version_ = 0;           // bumped whenever the g arrays change
    // inner stop: R32 <= max(frac*tol, floor2*R).  frac 0.9 measured at
    // 256x256x128x16: 825 FP32 sweeps for 30 steps against 910 at 0.25, one
    // correction round in every step either way
    double inner_frac_ = 0.9;
    double inner_omega_ = 1.2;     // FP32 correction solve's SOR factor (0: omega_c)
    double inner_floor2_ = 1e-10;  // (1e-5)^2: what an FP32 solve can resolve
    // Test hook of the failure paths: 1 = the next seg_step fails right
    // after its set-up (buffers rotated), 2 = it fails at its final wait,
    // after k_unpack wrote u^(n+1) into the recycled buffer.  Forced from
    // the host: every kernel still runs to completion.  Cleared when used.
    int inject_ = 0;
    Profile prof;
    std::string err;

    // --- internals (used by the solver templates in seg4d_engine.hip) --------
    int check(hipError_t e, const char* what);
    int launch_check(const char* what);
    int sumsq_interior(const double* d, double h, double* out);
    int reduce_to_host(int np, double* out);
    // elem_bytes: 8 for the FP64 passes, 4 for the FP32 correction passes
    struct SweepMap sweep_map(int elem_bytes = 8) const;
    bool tiled_order() const;  // the FP64 passes run in the tiled order
    int ensure_mixed();
    int defect64(const StepParams& P, bool copy32, const float* dred = nullptr);
    int solve_mixed(const StepParams& P, SolveStats* st);

    bool ready_ = false;
    int I_ = 0, J_ = 0, K_ = 0, L_ = 0, H_ = 0;
    long n_ = 0, np_ = 0;        // full / colour-plane element counts
    hipStream_t st_ = nullptr;
    bool own_st_ = true;
    double* full_[A_COUNT] = {};
    double* op_[2] = {};         // operand blocks (8 coefficients + prev) per colour
    long nop_ = 0;               // elements of one operand-block array
    double* ur_[2] = {};         // red ping-pong
    double* ub_ = nullptr;       // black
    double* part_ = nullptr;     // partial sums (two halves)
    long npart_ = 0;
    double* dscal_ = nullptr;    // device scalar scratch
    double* red2_ = nullptr;     // second-level partial sums (4096)
    bool defect_fused_ = false;  // the last set-up formed the first defect
    double* hscal_ = nullptr;    // pinned host scalar
    SorCtl* ctl_ = nullptr;      // device control block
    SorCtl* hctl_ = nullptr;     // pinned host mirror
    SorCtl* hctl_dev_ = nullptr; // its device alias (the reduce writes it directly)
    hipEvent_t ev_ = nullptr;
    hipStream_t side_ = nullptr;   // upload_side's stream
    hipEvent_t side_ev_ = nullptr;
    bool side_pending_ = false;    // the engine's stream has not waited for it yet
    int wait_side();
    std::vector<hipEvent_t> pev_; // profiling events (pairs)
    int red_final_ = 0;          // which red buffer holds the accepted state
    bool split_ = false;         // op_ holds the split operands of the FP32-refined set-up
    // FP32-refined form (allocated on first use)
    float* op32_[2] = {};        // FP32 operand blocks: 8 coefficients + defect
    float* dr_[2] = {};          // correction, red ping-pong
    float* db_ = nullptr;        // correction, black
    SorCtl* ctl2_ = nullptr;     // control block of the correction solves
    SorCtl* hctl2_ = nullptr;
    SorCtl* hctl2_dev_ = nullptr;
    // device ball list (count, offset, cx, cy, cz), re-sent when it changes
    int put_balls(const Balls& b, BallList* out);
    char* balls_buf_ = nullptr;
    size_t balls_cap_ = 0;
    Balls balls_sent_;
    bool balls_valid_ = false;
};

// Allocation-free _l2norm_ on caller-owned device arrays a, b of n elements:
// sqrt(sum (a-b)^2 h^4) into the device double *result, enqueued on stream st
// (nothing waited for).  Block partials go to a fixed device scratch array;
// n is limited to 65536 x 4096 elements.  Returns 0, or -1 with last_error().
int l2norm_noalloc(const double* a, const double* b, long n, double h, double* result,
                   hipStream_t st);

// Chunking of the two-level residual reduction over nb partials per colour.
void resid_chunking(int nb, int* nc, int* ch);

// Last error of the calling library (set by the ABI layers).
void set_error(const std::string& msg);
const char* last_error();

}  // namespace s4d
