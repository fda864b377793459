/*
 * seg4d_oracle.h -- CPU oracle for the semi-implicit subjective-surface /
 * mean-curvature-flow time step (TEST INFRASTRUCTURE ONLY).
 *
 * This is a clean-room FP64 restatement of the reference's hot path, written
 * so that every floating-point operation happens in the same order as in the
 * reference and therefore produces bitwise-identical results when compiled
 * with gcc -O2 -ffp-contract=off on x86-64.  Only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * library never links or calls it.
 *
 * State is passed explicitly (no globals).  Arrays are padded, x-fastest:
 *   idx = ((l*kMax + k)*jMax + j)*iMax + i,   interior 1..n, ghosts 0 and n+1.
 * Face order everywhere: 0=e(+x) 1=w(-x) 2=n(+y) 3=s(-y) 4=t(+z) 5=b(-z)
 *                        6=fw(+t) 7=bw(-t).
 */
#ifndef SEG4D_ORACLE_H
#define SEG4D_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int iMax, jMax, kMax, lMax;      /* padded dims (interior + 2)          */
    double h, tau, eps2, K;           /* grid size, time step, eps^2, K     */
    double omega_c, omega_sigma;      /* SOR relaxation (step, presmooth)   */
    double sigma;                     /* implicit heat-step time            */
    double gsTol;                     /* absolute bound on sum r^2 h^4      */
    double theta_img, rho_thr;        /* weights of image / threshold term  */
    int maxNoGSIteration;             /* sweep cap of the time step         */
    double maxNoGSIteration_sigma;    /* sweep cap of presmoothing (double) */
    int nthreads;                     /* 0/1 = serial loops, >1 = OpenMP    */
    int par_reduce;                   /* 1 = OpenMP reduction for residual
                                         sums (the reference OpenMP program's
                                         order); 0 = lexicographic serial sum
                                         (the serial reference's order);
                                         2 = residual terms formed by the
                                         threads, then summed serially in
                                         lexicographic order (bitwise the
                                         serial sum, most of the speed)     */
} orc_params;

typedef struct {                      /* ball list for the rescale epilogue */
    const int *count;                 /* count[l], l = 0..lMax-1            */
    const int *offset;                /* offset[l] into cx/cy/cz            */
    const double *cx, *cy, *cz;
    double radius;
} orc_balls;

size_t orc_size(const orc_params *p);
int orc_ijkl(const orc_params *p, int i, int j, int k, int l);

/* a3 */ double orc_gfunction(double value, double coef);
/* a9 */ void orc_copy(const orc_params *p, const double *src, double *dst);
/* a11 */ double orc_l2norm(const orc_params *p, const double *a, const double *b);
/* a4 */ void orc_reflection(const orc_params *p, double *u);
/* sum over interior of a^2 h^4, lexicographic (l,k,j,i) order */
double orc_sumsq_interior(const orc_params *p, const double *a);

/* a5: in-place implicit heat step; prev is scratch (receives the input). */
int orc_presmooth(const orc_params *p, double *u, double *prev, double *res_out);
/* a6: presmooths i0 and i0_thr in place, then the 8 face g's. */
void orc_gf(const orc_params *p, double *i0, double *i0_thr, double *prev,
            double *g[8], int *iters_out, double *res_out);
/* a7: 8 eps-regularised face coefficients from u; g == NULL is the MCF
 * variant (no edge-indicator factor). */
void orc_coefficients(const orc_params *p, const double *u, double *const g[8],
                      double *coef[8]);
/* Norm of one face gradient (helper, exposed for the face-sharing test). */
double orc_face_norm(const orc_params *p, const double *u, int i, int j, int k,
                     int l, int face);
/* a8 inner solve: red-black SOR to (R <= gsTol || z >= cap).  Returns z.
 * hist (optional, length >= cap) receives R after every iteration. */
int orc_sor(const orc_params *p, double *u, const double *prev,
            double *const coef[8], double *res_out, double *hist);
/* one SOR colour pass (0 = red, 1 = black) and the residual; exposed so that
 * tests can compare single passes. */
void orc_sor_pass(const orc_params *p, double *u, const double *prev,
                  double *const coef[8], int colour);
double orc_residual(const orc_params *p, const double *u, const double *prev,
                    double *const coef[8]);
/* a10 */ void orc_rescale(const orc_params *p, double *u, const double *mask,
                          const orc_balls *balls);

/* Segmentation initialisation (segmentation4D.c:317-335 mass, :876-1018
 * maxima, ball minima / maxima and threshold, :1106-1147 seed function),
 * ball boxes clamped to the array like orc_rescale. */
double orc_get4dmax(const orc_params *p, const double *img);
double orc_ball_minmax(const orc_params *p, const double *img, int l, double cx,
                       double cy, double cz, double R, int want_max);
void orc_threshold(const orc_params *p, const double *in, double *thr,
                   const orc_balls *balls, double alpha, double beta);
void orc_generate(const orc_params *p, double *u, double *mask,
                  const orc_balls *balls);
/* sum of a h^4 over the whole padded array, lexicographic order */
double orc_mass(const orc_params *p, const double *a);
/* a8 whole step: copy, coefficients, SOR, rescale.  Returns z. */
int orc_seg_time_step(const orc_params *p, double *u, double *prev,
                      double *const g[8], double *coef[8], const double *mask,
                      const orc_balls *balls, double *res_out);

/* a12 MCF verification twin */
void orc_mcf_boundary(const orc_params *p, double *u, double t);
void orc_mcf_exact(const orc_params *p, double *exact, double t);
int orc_mcf_time_step(const orc_params *p, double *u, double *prev,
                      double *coef[8], int step, double *res_out);

/* a14 segmentation mask and IoU over the interior */
void orc_mask(const orc_params *p, const double *u, unsigned char *mask);
double orc_iou(const orc_params *p, const unsigned char *a,
               const unsigned char *b);

/* CPU baseline helper: wall time of `passes` SOR iterations plus one
 * coefficient pass, using OpenMP over l like the reference OpenMP path. */
double orc_wtime(void);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
