// seg4d_net.h -- the learned warm start of the SOR solve.
//
// A separable 4D convolutional network predicts the next time step's
// solution from u^n, its recent increments and g, and the solver starts SOR
// from that prediction instead of u^n; the stopping test is unchanged, so
// the plain method still decides convergence and a better start only shows
// up as fewer sweeps.
//
// Network (parameters FP32, kernels BF16 with FP32 accumulation):
//   d = u^n - u^(n-1), sigma = rms of d over the interior (this step's motion)
//   t = u^n - 3u^(n-1) + 3u^(n-2) - u^(n-3) (third difference, the error the
//       quadratic extrapolation made at the last step), kappa = its rms
//   base = what the network corrects: u^n itself (base order 0, scale q = 1:
//          the inputs are u^n and g only), the linear extrapolation
//          2u^n - u^(n-1) (order 1, scale q = sigma) or the cubic
//          4u^n - 6u^(n-1) + 4u^(n-2) - u^(n-3) (order 3, scale q = kappa)
//   features x0 = [u^n, s_d d / sigma, g_mean, s_e t / kappa, 0 x 4]
//                 (8 channels; channel 3 is zero below order 3, channel 1
//                 at order 0)
//   `layers` x { 3x3x3 spatial conv -> 3-tap temporal conv + bias -> ReLU },
//   16 channels (1 to 4 layers)
//   1x1 projection -> c;  start = base + q c / s_t
// i.e. it learns the correction to the extrapolation in units of the
// step's own scale (the increment, or the last extrapolation error), so that
// one network serves early (fast) and late (slow) steps alike.  Training
// pairs come from the solver itself: record() snapshots the device u after
// each step of a run; train() samples patches of those runs (core 64^3
// points, all frames, a 4-point halo of real data) at the levels whose base
// exists and minimises the mean squared error of c against
// s_t (u^(n+1) - base_n) / q_n with Adam.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace s4d {

class Engine;

// The learned start's inputs to the set-up kernel: the network's output c
// (FP32, full layout) and the device sum over the interior whose root mean
// is the target's scale q (sigma for order 1, kappa for order 3), so that
// start = base + q c inv_st with q = sqrt(dsum inv_n), base the linear
// (order 1) or cubic (order 3) extrapolation.
struct LearnedStart {
    const float* out;
    const double* dsum;
    double inv_n;
    float inv_st;
    int order;
};

namespace net {
struct NGeo;
struct Taps;
}  // namespace net

class Net {
public:
    Net() = default;
    ~Net();
    Net(const Net&) = delete;
    Net& operator=(const Net&) = delete;

    // He-initialised weights from `seed` for `layers` (1..4) layers and a
    // base of order `base` (0, 1 or 3); empties the training pool.
    int init(uint64_t seed, int layers = 4, int base = 1);
    int layers() const { return nlayers_; }
    int base() const { return base_; }
    // Time levels before u^n the prediction needs (the base's order).
    int history_needed() const { return base_; }
    bool ready() const { return wm_ != nullptr; }
    void release();

    // Training pool: snapshot of the device u of `E` (FP32), appended to the
    // current run; `new_run` starts a run (and snapshots E's mean g).
    int record(Engine& E, bool new_run);
    long pool_pairs() const;
    // `iters` Adam steps on batches of `batch` patches; mean loss of the last
    // 10% of steps in *loss.  Prints progress every `log_every` steps (0: no).
    int train(int iters, int batch, float lr, uint64_t seed, int log_every, double* loss);
    int save(const char* path);
    int load(const char* path);
    // Re-derives the BF16 and transposed copies after wm_ was written.
    int refresh();

    // The network's correction c of the next solve's start (FP32, full
    // layout, interior points) from u^n (E's u) and the history levels
    // h1 = u^(n-1), h2 = u^(n-2), h3 = u^(n-3) (full layout; h2 and h3 only
    // for the cubic base), on E's stream.  *ls receives what the set-up
    // kernel needs to form start = base + q c / s_t (c and the device sum
    // of the increment, valid until the next prediction).  The engine asks
    // only when every level the base needs exists.
    int predict(Engine& E, const double* h1, const double* h2, const double* h3,
                LearnedStart* ls);

    // One forward pass on a caller-built batch (tests): feat BF16
    // [points][8], out FP32 [points].
    int forward_batch(int I, int J, int K, int L, int P, const void* feat_dev, float* out_dev);
    // Forward and backward on a caller-built batch (tests): leaves the
    // gradient of the mean squared error over the lmask points in grad_
    // (parameter layout) and the loss in *loss; no update.
    int grad_batch(int I, int J, int K, int L, int P, const void* feat_dev,
                   const float* target_dev, const unsigned char* lmask_dev, double* loss);

    std::string err;
    double predict_ms = 0.0;
    long predicts = 0;
    float s_d = 1.f, s_t = 1.f;  // feature / target scales
    float s_e = 1.f;             // scale of the cubic base's extra feature

    // parameter offsets (floats) into the parameter vector
    static constexpr int kMaxLayers = 4;
    int off_ws[kMaxLayers] = {}, off_wt[kMaxLayers] = {}, off_b[kMaxLayers] = {};
    int off_wp = 0, off_bp = 0, nparam = 0;

    // device parameters
    float* wm_ = nullptr;   // FP32 master weights
    __bf16* wb_ = nullptr;  // BF16 copy (same offsets)
    __bf16* wtr_ = nullptr; // per-tap transposed copies for the backward-data convs
    float* m1_ = nullptr;
    float* m2_ = nullptr;
    float* grad_ = nullptr;
    __bf16* w0s_ = nullptr; // first-layer spatial weights (4 channels) with the feature scales
    long adam_t_ = 0;

private:
    int nlayers_ = 4;  // layers of the current network
    int base_ = 1;     // order of the base it corrects (0: u^n, 1 linear, 3 cubic)
    int check(hipError_t e, const char* what);
    int layout();
    int ensure_acts(int I, int J, int K, int L, bool training);
    // w0: spatial weights of the first layer for 4 input channels, used
    // instead of wb_'s 8-channel ones (inference only: the prediction's
    // scaled copy, with feat_ holding 4 channels per point)
    int forward(const net::NGeo& G, hipStream_t st, bool training, const __bf16* w0 = nullptr);
    int backward(const net::NGeo& G, hipStream_t st, float inv_n);
    int compute_scales(hipStream_t st);

    struct Run {
        int I = 0, J = 0, K = 0, L = 0;
        __bf16* gmean = nullptr;       // BF16 [points]
        std::vector<double*> snaps;    // FP64 [points] per time level
        std::vector<double> sigma;     // rms of u^n - u^(n-1) over the interior (0 for n = 0)
        std::vector<double> kappa;     // rms of the third difference (0 for n < 3)
    };
    std::vector<Run> runs_;

    // activations of the current geometry
    int aI_ = 0, aJ_ = 0, aK_ = 0, aL_ = 0;
    bool a_training_ = false;
    __bf16* feat_ = nullptr;          // [pts][8]
    __bf16* S_[kMaxLayers] = {};      // spatial conv outputs [pts][16]
    __bf16* Y_[kMaxLayers] = {};      // layer outputs [pts][16]
    __bf16* dA_ = nullptr;            // gradient buffers [pts][16]
    __bf16* dB_ = nullptr;
    float* out_ = nullptr;            // [pts]
    float* target_ = nullptr;         // [pts]
    unsigned char* lmask_ = nullptr;  // [pts]
    float* partial_ = nullptr;        // reduction partials
    long npartial_ = 0;
    double* dscal_ = nullptr;
    double* hscal_ = nullptr;
    long g_version_ = -1;             // engine g generation of gmean_
    __bf16* gmean_ = nullptr;         // inference: mean g of the engine [pts]
    long gpts_ = 0;
};

}  // namespace s4d
