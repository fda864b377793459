/*
 * seg4d.h -- C ABI of libseg4d_amd.so: MI355X (gfx950) implementation of the
 * semi-implicit 4D subjective-surface segmentation time step.
 *
 * Drop-in boundary.  The functions and globals below carry the names, types
 * and meaning of the reference's serial segmentation program
 * (88MARK08/4D-image-segmentation-algorithm, segmentation4D.h:8-62 for the
 * prototypes, segmentation4D.c:13-60 for the globals), so a driver written
 * against that program links against this library unchanged.  The caller
 * sets the dimensions and parameters in the globals, calls _allocateMem_,
 * writes the host arrays through the global pointers, and calls the entry
 * points; the library runs them on the GPU.
 *
 * Coherence ("mirror" mode, the default): host arrays are authoritative
 * between calls.  Each entry point uploads the arrays it reads and
 * downloads the arrays it writes, with two exceptions that avoid moving
 * eight full volumes per step:
 *   - g_e..g_bw are uploaded only when the library did not produce them:
 *     after _allocateMem_ (first use), or after seg4dHostWrote(g_x);
 *   - e..bw (coefficients) are downloaded by _gs_coefficients_ and
 *     seg4dDownload, not by _segTimeStep_ (the reference's drivers never
 *     read them between steps).
 * "Resident" mode (seg4dSetResident(1)): device arrays are authoritative;
 * entry points never copy implicitly; seg4dUpload / seg4dDownload move
 * arrays explicitly.
 *
 * Differences from the reference, all on error paths only:
 *   - an entry point returns false on a HIP error, on a non-finite SOR
 *     residual (the reference silently stops, segmentation4D.c:309) or on
 *     arguments it cannot resolve, with the reason in seg4dLastError();
 *   - the per-ball loops of the rescale epilogue are clamped to the array
 *     (the reference indexes them unchecked, segmentation4D.c:1028-1090);
 *   - no_of_centers == NULL is read as zero balls in every frame.
 * Not thread-safe (global state, as the reference); one device, one stream.
 */
#ifndef SEG4D_H
#define SEG4D_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "seg4d_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference globals (segmentation4D.c:13-60) ------------------------ */
extern int kMax, iMax, jMax, lMax;          /* padded dims (interior + 2)   */
extern double *u, *prev_u, *i0, *Mmask;     /* segmentation4D.c:16-19       */
extern double *e, *w, *n, *s, *t, *b, *fw, *bw;              /* :21-28      */
extern double *g_e, *g_w, *g_n, *g_s, *g_t, *g_b, *g_fw, *g_bw; /* :30-37   */
extern int *no_of_centers;                  /* :39, per frame, caller-owned */
extern double **center_x, **center_y, **center_z;            /* :40         */
extern double spacing_x, spacing_y, spacing_z;               /* :41         */
extern double origin_x, origin_y, origin_z;                  /* :42         */
extern int maxNoGSIteration;                /* :44 sweep cap of a time step */
extern double eps2, K;                      /* :45-46                       */
extern int numberOfTimeStep, maxNoOfTimeSteps;               /* :47-48      */
extern double segTol;                       /* :49                          */
extern double tau, h, omega_c, gsTol, omega_sigma;           /* :50         */
extern double theta_img, rho_thr, alpha, beta;               /* :53         */
extern int mod;                             /* :54                          */
extern double sigma;                        /* :55                          */
extern double maxNoGSIteration_sigma;       /* :56                          */
extern int timeStepsNum;                    /* :57                          */
extern double ballRadius;                   /* :58                          */

/* ---- reference entry points on the hot path ---------------------------- */
/* segmentation4D.h:8 / segmentation4D.c:71-155: allocates and zero-fills
 * the 20 padded arrays (host) and their device mirrors, and the per-frame
 * centre arrays sized by no_of_centers[l]. */
bool _allocateMem_(void);
/* segmentation4D.h:10 / segmentation4D.c:156-192 */
bool _deallocateMem_(void);
/* segmentation4D.h:12 / segmentation4D.c:194-213: sqrt(sum (a-b)^2 h^4)
 * over the padded array; a and b must be library arrays (u, prev_u, i0,
 * Mmask, g_*) or, in mirror mode, any host arrays of the padded size. */
double _l2norm_(double *dataArray3DPtr1, double *dataArray3DPtr2);
/* segmentation4D.h:14 / segmentation4D.c:446-580: the 8 face coefficients
 * e..bw from u and g_*. */
bool _gs_coefficients_(void);
/* segmentation4D.h:16 / segmentation4D.c:215-315: prev_u <- u,
 * coefficients, red-black SOR to (R <= gsTol or maxNoGSIteration sweeps),
 * rescale epilogue.  On failure (false) u is as before the call; prev_u
 * too in resident mode, prev_u = u in mirror mode. */
bool _segTimeStep_(void);
/* segmentation4D.h:18 / segmentation4D.c:337-409: implicit heat step in
 * place (prev_u receives the input; ghosts reflected once). */
bool _presmoothStep_(double *u);
/* segmentation4D.h:28 / segmentation4D.c:582-823 */
bool _gf_inputImage_plus_thresholdedImage_(double *i0, double *i0_thr);
/* segmentation4D.h:30 / segmentation4D.c:866-874 */
void _copy4DdataFromSrcToDest_(double *srcDataPtr, double *destDataPtr);
/* segmentation4D.h:32 / segmentation4D.c:61-65 */
int ijkl(int i, int j, int k, int l);
/* segmentation4D.h:34 / segmentation4D.c:67-70 */
double _gFunction_(double value, double coef);
/* segmentation4D.h:24 / segmentation4D.c:1020-1104 */
bool _rescale4DdataToZeroOne_(double *imagePtr);
/* segmentation4D.h:56 / segmentation4D.c:826-864 */
bool _reflection4Ddata_(double *u);

/* ---- initialisation of the segmentation function ----------------------- */
/* The balls are the per-frame centres center_x/y/z[l][c], c < no_of_centers[l],
 * with radius ballRadius; boxes are clamped to the array as in the rescale. */
/* segmentation4D.h:26 / segmentation4D.c:971-1018: inside every ball, points
 * of the input brighter than alpha*min + beta*max of that ball are set to
 * the input's interior maximum (from 0) in the output (which may alias the
 * input). */
bool _locallyThreshold4Ddata_(double *inputImagePtr, double *thresholdedImagePtr);
/* segmentation4D.h:22 / segmentation4D.c:1106-1148: within each ball's box
 * and distance d <= 12: value = max(value, |1/(d + 0.5) - 1/12.5|) and
 * Mmask = 1. */
bool _generate4DsegmFunct_(double *segmFuntionPtr);
/* segmentation4D.h:38 / segmentation4D.c:876-895: interior maximum, starting
 * from 0.  Exits the process on NULL, as the reference. */
double _get4DMax_(double *imagePtr);
/* segmentation4D.h:40 / segmentation4D.c:934-969: maximum (from 0) of frame
 * l over |x - c| <= R.  Exits the process on NULL, as the reference. */
double _get3DMaxWithinR_(double *imagePtr, int l, double center_x, double center_y,
                         double center_z, double R);
/* segmentation4D.h:42 / segmentation4D.c:897-932: minimum (from 100000). */
double _get3DMinWithinR_(double *imagePtr, int l, double center_x, double center_y,
                         double center_z, double R);
/* segmentation4D.h:62 / segmentation4D.c:317-335: sum of value * h^4 over
 * the whole padded array (device reduction: equal to the reference's
 * sequential sum to rounding, ~1e-15 relative). */
double _mass4D_(double *dataArray3DPtr1);

/* ---- file formats (host; csrc/seg4d_io.c) -------------------------------- */
/* They read and write the host arrays.  In resident mode, seg4dUpload after
 * reading and seg4dDownload before writing. */
extern FILE *input_file, *input_center_file, *output_file; /* segmentation4D.c:60 */
/* segmentation4D.h:52 / segmentation4D.c:1149-1216: frames l = 1..lMax-2
 * from <dir>/movingBalls_457_at_time_%03d.vtk (bytes / 255.) and centres
 * from <dir>/movingBalls_457_centers_at_timestep_%03d.vtk, each mapped to
 * (double)(int)(x / spacing + 0.5).  Exits the process on a missing or
 * malformed file, as the reference (with a message). */
void _readData_(double *inputImage, unsigned char *fileDirectory);
/* segmentation4D.h:48 / segmentation4D.c:1282-1313: one frame's bytes after
 * ten header lines into frame l's interior. */
bool _load3dDataArrayVTK_(double *imageDataPtr, int l, unsigned char *pathPtr);
/* segmentation4D.h:50 / segmentation4D.c:1218-1253: <path>_at_time_%03d.vtk
 * per frame, bytes (unsigned char)(255 v + 0.5).  False if a file cannot be
 * created. */
bool _writeData_(double *array4DPtr, unsigned char *pathPtr);
/* segmentation4D.h:44 / segmentation4D.c:1255-1266 */
void _readNumberFromFile_(char *dim, FILE *input_file);
/* segmentation4D.h:46 / segmentation4D.c:1268-1280 */
void _readMetaData_(FILE *input_file);

/* ---- extensions (no counterpart in the reference) ---------------------- */
/* The mask writer of the MPI program (segmentation4D_mpi.c:1429-1480): the
 * _writeData_ files with bytes 1 where value >= 0.5, else 0. */
bool seg4dWriteMaskVTK(const double *array4DPtr, const char *pathPtr);
/* 0 = mirror (default), 1 = resident.  Returns the previous mode. */
int seg4dSetResident(int on);
/* > 0: before every SOR solve (time step and presmoothing) the library
 * replaces gsTol by rel^2 * sum(prev^2 h^4) over the interior, i.e. stops at
 * relative residual ||b - A u|| / ||b|| <= rel.  0 (default): gsTol as set. */
void seg4dSetRelativeTolerance(double rel);
/* Explicit transfers of a library array (any global pointer above).  The
 * coefficient arrays e..bw download in the reference layout. */
bool seg4dUpload(double *hostArray);
bool seg4dDownload(double *hostArray);
/* Tells the library that the caller wrote a library array directly. */
void seg4dHostWrote(double *hostArray);
/* Mirror-mode host-write tracking of u, prev_u and Mmask (default off; also
 * switched on by SEG4D_TRACK_WRITES=1 in the environment at the first
 * _allocateMem_).  After a mirror _segTimeStep_ the three arrays equal their
 * device copies; the library keeps them read-only (mprotect) and catches the
 * caller's first write to each with a SIGSEGV handler that chains to the one
 * installed before it.  _segTimeStep_ and _l2norm_ then upload only arrays
 * written since, which removes four of the five volume transfers of the
 * reference's loop (segmentation4D.c:1398-1411: the step, then
 * _l2norm_(prev_u, u)).  Caveats: a system call that writes into a clean
 * array (read(2), fread into it) fails with EFAULT -- call seg4dHostWrote on
 * it first; a SIGSEGV handler installed after the library's turns the
 * protection off for later steps but must not meet a write to an array
 * still protected -- call seg4dSetWriteTracking(0) before installing one.
 * Only page-aligned arrays of at least 64 pages are tracked (the library's
 * page-locked host arrays are page-aligned).
 * Returns the previous setting. */
int seg4dSetWriteTracking(int on);
/* Bit 0: tracking on and its handler the installed one; bits 1, 2, 3: u,
 * prev_u, Mmask known equal to their device copies (no upload needed). */
int seg4dWriteTrackingState(void);
/* 0 silences the reference's per-step printf lines (default 1). */
void seg4dSetVerbose(int on);
const char *seg4dLastError(void);
/* Statistics of the last solve: SOR iterations, residual, bound used. */
int seg4dLastSweeps(void);
double seg4dLastResidual(void);
double seg4dLastTolerance(void);
/* HIP-event profiling (events on the library's stream).  seg4dProfileRead
 * fills out[0..13] = {ms, launches} for the three SOR pass kernels (red
 * without residual, red with fused residual, black with fused residual),
 * then setup_ms, setup_launches, step_ms, steps, voxel_updates (interior
 * voxels x SOR iterations), iterations, red points, black points;
 * out[14..19] the same three {ms, launches} pairs for the FP32 correction
 * sweeps, out[20..22] = FP64 defect ms, defect launches, correction rounds,
 * out[23..24] = ms and launches of resident-mode copies (a plain stream of
 * known size).  Returns the number of values written (at most 25). */
void seg4dProfile(int on);
int seg4dProfileRead(double *out, int nout);
void seg4dProfileReset(void);
/* SOR pass kernel variant (tuning knob), bit flags: 1 = two l-steps per
 * load phase; 2 = tiled launch order (the blocks that share neighbour planes
 * are co-resident on one XCD); 4 = nontemporal loads of the coefficient
 * stream; 8 = two, 16 = four consecutive k planes per block (8 and 16
 * exclusive); 32 = automatic order: with automatic frame chunks
 * (seg4dSetSweepChunk(0)), chunks of 2 frames in the linear order, the runs
 * of each plane split into bands when the other colour's window of 3 k planes
 * x 4 frames exceeds 7 MB (seg4dSetSweepBands), replacing bit 2.  Default 35.  u and the sweep counts are those
 * of any other variant; bits 8/16 and the frame chunk regroup the residual
 * sums (last-bit differences in R).  Returns the previous value, -1 before
 * _allocateMem_. */
int seg4dSetSweepVariant(int v);
/* Tile of the tiled launch order in blocks: `runs` runs x `planes` k-plane
 * groups per XCD residency round (default 16 x 8).  0 on success. */
int seg4dSetSweepTile(int runs, int planes);
/* Frames per block of a SOR colour pass (tuning knob, identical results):
 * 0 (default) = all frames unless the volume is too small to fill the GPU,
 * then chunks of an even length; n > 0 = chunks of n frames.  Returns the
 * previous value. */
int seg4dSetSweepChunk(int lsteps);
/* Bands of the automatic order (variant bit 32, automatic chunks): 0
 * (default) = as many as keep the other colour's window within 7 MB (planes
 * larger than that are split into bands of runs, each band walked through
 * the k planes in turn); n > 0 forces n bands.  Tuning knob: the
 * same u and sweep counts.  Returns the previous value. */
int seg4dSetSweepBands(int nband);
/* SOR iterations enqueued ahead of the stopping test the host has seen
 * (1..16; 0 = automatic: 4 for volumes under 4M points per colour, else 1).
 * Passes enqueued after the stop are no-ops; results are identical.
 * Returns the previous value. */
int seg4dSetPipelineDepth(int depth);
/* Fused sweeps: 1 runs every SOR solve (and FP32 correction solve) as one
 * cooperative launch that holds all its red passes, black passes and
 * stopping tests, with grid-wide barriers between them, whenever the grid
 * fits on the GPU; 0 (the default) one launch per pass.  Results and sweep
 * counts are identical; on an MI355X the fused solve is the slower one at
 * every size measured (its grid barriers cost more than the launch gaps they
 * remove).  Returns the previous value (-2 without a volume). */
int seg4dSetFusedSweeps(int mode);
/* 1 if the last SOR solve (or correction solve) ran fused. */
int seg4dLastSolveFused(void);
/* Arithmetic of the SOR solve.  0 (default): FP64, the reference's
 * arithmetic and bits.  1: FP32-refined -- FP64 defect correction in which
 * the defect b - A u of the FP64 system is recomputed in FP64 on the GPU and
 * an FP32 red-black SOR solve of A d = b - A u (coefficients to 20 bits)
 * is applied; the stopping test stays the reference's on the FP64 residual
 * (R <= gsTol, or the relative bound), and maxNoGSIteration caps each FP32
 * solve.  Returns the previous mode, -1 before _allocateMem_. */
int seg4dSetPrecision(int mode);
/* FP32-refined controls: at most max_rounds corrections per solve (default
 * 8); each FP32 solve stops once its residual is below max(frac * tol,
 * inner_floor^2 * R) of the current FP64 residual R (inner_floor default
 * 1e-5, frac: seg4dSetInnerFraction).  Non-positive values keep the current
 * setting. */
void seg4dSetRefinement(int max_rounds, double inner_floor);
/* The frac above (0 < frac < 1): fraction of the outer bound at which an
 * FP32 correction solve stops.  A larger fraction saves FP32 sweeps and
 * risks a second correction round (the FP64 stopping test is unchanged).
 * Default 0.9. */
void seg4dSetInnerFraction(double frac);
/* Relaxation factor of the FP32 correction solves (0 < omega < 2; 0 uses
 * omega_c, as the reference's solve does).  The correction solve only has to
 * reduce the defect: any convergent factor gives a result that meets the
 * same FP64 stopping test.  Default 1.2: omega_c = 1.8 over-relaxes this
 * system (the error of an SOR sweep at 1.8 contracts by about 0.8), and 1.2
 * took the fewest sweeps on the 64x64x32x8 and 256x256x128x16 workloads.
 * Other values are ignored, so a negative argument queries the setting.
 * Returns the previous setting. */
double seg4dSetInnerOmega(double omega);
/* Correction rounds of the last FP32-refined solve (0 in FP64 mode);
 * seg4dLastSweeps then counts FP32 sweeps. */
int seg4dLastRounds(void);
/* Start of the SOR solve.  0 (default): u^n, as the reference.  1: the
 * linear extrapolation 2 u^n - u^(n-1), used when the previous call was a
 * completed _segTimeStep_ whose state stayed on the device (resident mode).
 * 2: the prediction of the learned network (seg4dNet*): the network's
 * correction of u^n (base order 0, from the first step on), of the linear
 * extrapolation (order 1, under the condition of 1) or of the cubic one
 * (order 3, under the condition of 4; before that, as 4).  3: the quadratic
 * extrapolation 3 u^n - 3 u^(n-1) + u^(n-2) after two such steps (after
 * one, as 1; one more padded FP64 array on the device).  4: the cubic one
 * 4 u^n - 6 u^(n-1) + 4 u^(n-2) - u^(n-3) after three (else as 3; two more
 * arrays).  In every mode
 * the system solved, its right-hand side and the stopping test are
 * unchanged; only the number of sweeps (and the last bits of u) differ.
 * Returns the previous mode, -1 before _allocateMem_ or (mode 2) without a
 * network. */
int seg4dSetWarmStart(int mode);

/* ---- learned warm start (no counterpart in the reference) ---------------
 * A separable 4D convolutional network (1 to 4 x {3x3x3 spatial conv, 3-tap
 * temporal conv, bias, ReLU}, 16 channels, 1x1 projection; BF16 MFMA
 * kernels, FP32 accumulation) predicts the correction of a base -- u^n
 * itself (base order 0: the network sees only u^n and the mean g and
 * predicts the increment u^(n+1) - u^n), the linear extrapolation
 * 2 u^n - u^(n-1) (order 1) or the cubic 4 u^n - 6 u^(n-1) + 4 u^(n-2) -
 * u^(n-3) (order 3) -- from u^n, the increment u^n - u^(n-1) (orders 1, 3),
 * the cubic's difference from the linear one (order 3) and the mean g.  It is trained on solver runs: seg4dNetRecord
 * after each _segTimeStep_ (and once before the first) fills the pool;
 * seg4dNetTrain fits it with Adam on patches of 64^3 points x all frames.
 * The network outlives _deallocateMem_, so runs of several volumes can be
 * pooled.  Functions return 0 (or the documented value) on success, -1 on
 * failure. */
int seg4dNetInit(uint64_t seed);          /* fresh weights: 4 layers, order 1 */
/* Fresh weights of `layers` (1..4) layers correcting the base of order
 * `base` (0, 1 or 3); empties the pool. */
int seg4dNetInitShape(uint64_t seed, int layers, int base);
int seg4dNetShape(int *layers, int *base);  /* shape of the current network */
int seg4dNetLoad(const char *path);       /* weights file of seg4dNetSave   */
int seg4dNetSave(const char *path);
/* Snapshot of the current u (the device copy; uploaded first in mirror
 * mode); new_run != 0 starts a run and snapshots the mean g. */
int seg4dNetRecord(int new_run);
long seg4dNetPoolPairs(void);             /* training levels pooled (with the
                                             base's history and u^(n+1))  */
/* iters Adam steps (lr) on batches of `batch` patches; progress line every
 * log_every steps (0: silent).  Returns the mean loss of the last 10% of
 * steps, -1 on failure.  Runs need equal frame counts and >= 74 padded
 * points in x, y, z. */
double seg4dNetTrain(int iters, int batch, double lr, uint64_t seed, int log_every);
/* Parameter vector (FP32) out of / into the network; seg4dNetParams(NULL, 0)
 * returns the count.  s_d, s_t: feature and target scales. */
int seg4dNetParams(float *out, int n);
int seg4dNetSetParams(const float *in, int n, float s_d, float s_t);
void seg4dNetScales(float *s_d, float *s_t);
/* Test hooks on a caller-built batch of padded dims I, J, K, L (frame
 * period P): features BF16 bit patterns [points][8]; one forward pass
 * (out: FP32 [points]) or one forward + backward pass of the mean squared
 * error over lmask points (gradient in parameter layout, loss). */
int seg4dNetForward(int I, int J, int K, int L, int P, const uint16_t *feat_bf16, float *out);
int seg4dNetGrad(int I, int J, int K, int L, int P, const uint16_t *feat_bf16,
                 const float *target, const unsigned char *lmask, float *grad_out,
                 double *loss);
/* Test hook: the network's correction c of the next solve's start (FP32,
 * padded layout, interior points; the caller's array holds one float per
 * padded point) and its scale q (1 for base order 0, rms increment for
 * order 1, rms third difference for order 3), so that start = base +
 * q c / s_t, from the
 * device state after completed resident time steps; needs the history the
 * base uses.  The state is left unchanged. */
int seg4dNetPreview(float *c, double *q);
double seg4dNetPredictMs(void);           /* total time of the predictions  */
long seg4dNetPredicts(void);
/* Launch-order autotuner of the current volume (no counterpart in the
 * reference): times the FP64 red pass over tile shapes of the XCD-resident
 * order and the linear order (reps launches each, default 3) on the current
 * device state and keeps the fastest.  Results are unchanged by the choice.
 * *runs / *planes receive the tile (0 for the linear order), *best_ms its
 * time; any output may be NULL.  Returns 0, or -1 without a volume. */
int seg4dAutotune(int reps, int *runs, int *planes, float *best_ms);

/* ---- Batched form (no counterpart in the reference) ------------------------
 * nvol volumes of different sizes stepped together: every SOR colour pass,
 * FP64 defect, correction and stopping test of all volumes is one launch,
 * with the stopping decisions per volume on the device.  Each volume's
 * result is bitwise that of the same step on the single-volume library with
 * the same settings.  The batch is independent of _allocateMem_'s arrays;
 * its parameters are the reference's globals at each call (h, tau, eps2,
 * omega_c, gsTol or the relative tolerance, maxNoGSIteration; for
 * seg4dBatchSetVolume sigma, omega_sigma, maxNoGSIteration_sigma, K,
 * theta_img).  One batch at a time; all calls return 0, or -1 with
 * seg4dLastError set. */
/* dims[4v .. 4v+3] = interior (nx, ny, nz, nt) of volume v. */
int seg4dBatchCreate(int nvol, const int *dims);
void seg4dBatchDestroy(void);
/* Host arrays of volume v's padded size (x fastest): the image, whose edge
 * indicator (presmoothing, then g with rho_thr = 0) becomes the volume's,
 * the initial u and Mmask. */
int seg4dBatchSetVolume(int v, const double *img, const double *u0, const double *mask);
/* One time step of every volume (set-up, SOR solve, mask epilogue). */
int seg4dBatchStep(void);
int seg4dBatchGetU(int v, double *u);
/* Last step of volume v: SOR iterations (FP32-refined: FP32 sweeps),
 * correction rounds, FP64 residual and its bound. */
int seg4dBatchStats(int v, int *sweeps, int *rounds, double *R, double *tol);
double seg4dBatchLastStepMs(void);   /* device time of the last step */
/* precision 0 (FP64) / 1 (FP32-refined, correction solves at the factor of
 * seg4dSetInnerOmega's default 1.2); warm start 0, 1, 3 or 4 as
 * seg4dSetWarmStart; graph 1 replays the solve iterations from a captured
 * HIP graph (two iterations per replay); graph 2 runs every time step after
 * the first as one captured HIP graph (set-ups, solve, correction rounds,
 * epilogue; every decision on the device), holding as many iterations per
 * solve and rounds as the previous step needed plus a margin; a step whose
 * solve needs more writes nothing and is redone eagerly; depth 1..16
 * iterations (replays) queued ahead of the stopping test of eager solves.
 * Defaults 0, 0, 0, 2. */
int seg4dBatchConfigure(int precision, int warm, int graph, int depth);
/* Graph mode 2: iterations per solve and rounds the next captured step
 * holds, steps run as one graph and steps redone eagerly so far, and the
 * host time spent capturing and instantiating them (any output may be
 * NULL). */
int seg4dBatchGraphInfo(int *iters, int *rounds, long *steps, long *redone, double *host_ms);
/* Graph mode 2: the iterations per solve and rounds the next captured step
 * holds, in place of the plan from the last step (both 0: run the next step
 * eagerly).  Later steps plan from their predecessor again. */
int seg4dBatchSetGraphPlan(int iters, int rounds);
/* Concurrent form: nstreams >= 1 splits the volumes into that many groups of
 * about equal voxel count (at most one per volume), each on its own stream
 * driven by its own host thread, and steps every volume on the
 * single-volume path (its own launches and pipelined stopping test), so
 * that the groups' passes overlap on the GPU; results are bitwise those of
 * the one-launch form.  seg4dBatchLastStepMs then reports the host wall time
 * of the step.  A failed step leaves the volumes of other groups stepped
 * (seg4dBatchStats tells which).  0 (the default) restores one launch per
 * pass for all volumes. */
int seg4dBatchSetStreams(int nstreams);
/* Per-volume launch-order tiles (seg4dAutotune on each volume); best_ms[v]
 * (may be NULL) receives each volume's red-pass time. */
int seg4dBatchAutotune(int reps, float *best_ms);
int seg4dBatchTile(int v, int *runs, int *planes);

/* Allocation-free variant of _l2norm_ (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops).  a, b: device arrays of n doubles; result: a device
 * double receiving sqrt(sum over the n elements of (a-b)^2 h^4), the value
 * _l2norm_ returns for the same arrays, bit for bit.  The work is enqueued on
 * `stream` (a hipStream_t; NULL = the default stream) and not waited for:
 * synchronise before reading *result.  It calls no allocator: the partial
 * sums go to a fixed device array of 65536 doubles (512 KB) per GPU, so
 * 1 <= n <= 65536 * 4096 = 268,435,456 (larger n is refused), and calls that
 * overlap on one GPU share that array and race.  Independent of
 * _allocateMem_.  Returns 0, or -1 with seg4dLastError() set. */
int seg4dL2normNoAlloc(const double *a, const double *b, long n, double h, double *result,
                       void *stream);

/* Test hook of the failure paths (no counterpart in the reference): makes
 * the next _segTimeStep_ fail, forced from the host, 1 = right after its
 * set-up kernel, 2 = at its final wait (after the new u was written into
 * the recycled device buffer).  Every kernel still runs to completion; the
 * step then returns false as after a device error, with u = u^n on the
 * device, the warm-start history dropped and seg4dLastError() naming the
 * injection.  0 clears it.  Returns 0, or -1 before _allocateMem_. */
int seg4dInjectStepFailure(int where);

/* Blocks until all queued device work is done. */
bool seg4dSynchronize(void);
/* Device name / CU count / arch into buf; returns 0 on success. */
int seg4dDeviceInfo(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
