/*
 * mcf4d.h -- C ABI of libmcf4d_amd.so: MI355X (gfx950) implementation of the
 * 4D level-set mean-curvature-flow verification step.
 *
 * Drop-in for the reference's MCF programs (88MARK08/4D-image-segmentation-
 * algorithm, mcf_4D.h:8-23 for the prototypes, mcf_4D.c:11-34 for the
 * globals; the OpenMP program mcf_4D_openmp.c has the same interface).  It
 * is a separate library because the reference's two programs define
 * clashing global symbols (u, e, ijkl, ...).  Coherence modes, error
 * behaviour and threading are those of seg4d.h.
 */
#ifndef MCF4D_H
#define MCF4D_H
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference globals (mcf_4D.c:11-34) --------------------------------- */
extern int kMax, iMax, jMax, lMax;
extern double *u, *prev_u, *exact_u;
extern double *e, *w, *n, *s, *t, *b, *fw, *bw;
extern int maxNoGSIteration;
extern double eps2;
extern int numberOfTimeStep, maxNoOfTimeSteps;
extern double mcfTol;
extern double tau, h, omega_c, gsTol;

/* ---- reference entry points ---------------------------------------------- */
/* mcf_4D.h:8 / mcf_4D.c:42-63 */
bool _allocateMem_mcf(void);
/* mcf_4D.h:10 / mcf_4D.c:65-79 */
bool _deallocateMem_mcf(void);
/* mcf_4D.h:12 / mcf_4D.c:81-100 */
double _l2norm_mcf(double *dataArray3DPtr1, double *dataArray3DPtr2);
/* mcf_4D.h:14 / mcf_4D.c:208-342: coefficients with g == 1 */
bool _gs_coefficients_mcf(void);
/* mcf_4D.h:16 / mcf_4D.c:102-205: copy, coefficients, exact boundary at
 * t = tau * numberOfTimeStep, red-black SOR.  On failure (false) u is as
 * before the call; prev_u too in resident mode, prev_u = u in mirror mode. */
bool _mcfTimeStep(void);
/* mcf_4D.h:18 / mcf_4D.c:344-443 */
bool _set4DBoundaryToExactValues_mcf(double t);
/* mcf_4D.h:20 / mcf_4D.c:444-462 */
void _exactSolution_mcf(double t);
/* mcf_4D.h:22 / mcf_4D.c:463-471 */
void _copy4DdataFromSrcToDest_mcf(double *srcDataPtr, double *destDataPtr);
/* mcf_4D.h:24 / mcf_4D.c:36-40 */
int ijkl(int i, int j, int k, int l);

/* ---- extensions (no counterpart in the reference) ------------------------ */
int mcf4dSetResident(int on);
void mcf4dSetRelativeTolerance(double rel);
bool mcf4dUpload(double *hostArray);
bool mcf4dDownload(double *hostArray);
/* Mirror-mode host-write tracking of u, prev_u and exact_u, as
 * seg4dSetWriteTracking (include/seg4d.h; environment MCF4D_TRACK_WRITES=1
 * at the first _allocateMem_mcf): _mcfTimeStep, _exactSolution_mcf and
 * _l2norm_mcf upload only arrays the host wrote since they last matched
 * their device copies, so the reference's loop (mcf_4D.c:500-516: step,
 * exact solution, _l2norm_mcf(exact_u, u)) moves u and exact_u down and
 * nothing up.  Same caveats: a system call writing into a clean array fails
 * with EFAULT; a SIGSEGV handler installed later must come after
 * mcf4dSetWriteTracking(0).  Returns the previous setting. */
int mcf4dSetWriteTracking(int on);
/* Bit 0: tracking active; bits 1, 2, 3: u, prev_u, exact_u clean. */
int mcf4dWriteTrackingState(void);
void mcf4dSetVerbose(int on);
const char *mcf4dLastError(void);
int mcf4dLastSweeps(void);
double mcf4dLastResidual(void);
bool mcf4dSynchronize(void);

#ifdef __cplusplus
}
#endif
#endif
