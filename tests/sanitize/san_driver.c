/*
 * san_driver.c -- host code under AddressSanitizer + UndefinedBehaviorSanitizer
 * (test infrastructure; built by tests/sanitize/Makefile, run by
 * tests/test_sanitized_host.py).
 *
 * Links the CPU oracle (oracle/seg4d_oracle.c), the file formats
 * (csrc/seg4d_io.c) and the synthetic generator (csrc/seg4d_gen.c) into one
 * executable with the sanitizers on, and drives the paths where the
 * reference indexes without bounds checks: the ball loops of thresholding,
 * seed generation and rescaling with centres on and beyond the volume's
 * edges (segmentation4D.c:907-930, 991-1015, 1033-1090, 1119-1145), the
 * VTK readers and writers (segmentation4D.c:1149-1324), and every threaded
 * branch of the oracle.  Any sanitizer report aborts with a non-zero exit.
 *
 *   san_driver io <vtkdir> I J K L <c_0> .. <c_{L-1}> <spacing> <out.raw> <writebase>
 *   san_driver oracle
 *   san_driver gen
 *   san_driver selfcheck     (reads one element past an array: must abort)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "seg4d.h"
#include "seg4d_gen.h"
#include "seg4d_oracle.h"

/* ---- the reference globals seg4d_io.c reads (segmentation4D.c:13-60) ---- */
int kMax, iMax, jMax, lMax;
double *u, *prev_u, *i0, *Mmask;
double *e, *w, *n, *s, *t, *b, *fw, *bw;
double *g_e, *g_w, *g_n, *g_s, *g_t, *g_b, *g_fw, *g_bw;
int *no_of_centers;
double **center_x, **center_y, **center_z;
double spacing_x, spacing_y, spacing_z;
double origin_x, origin_y, origin_z;
int maxNoGSIteration;
double eps2, K;
int numberOfTimeStep, maxNoOfTimeSteps;
double segTol;
double tau, h, omega_c, gsTol, omega_sigma;
double theta_img, rho_thr, alpha, beta;
int mod;
double sigma;
double maxNoGSIteration_sigma;
int timeStepsNum;
double ballRadius;

static double *alloc0(size_t nn)
{
    double *p = (double *)calloc(nn, sizeof(double));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(3);
    }
    return p;
}

static int run_io(int argc, char **argv)
{
    if (argc < 8) return 2;
    const char *dir = argv[2];
    iMax = atoi(argv[3]);
    jMax = atoi(argv[4]);
    kMax = atoi(argv[5]);
    lMax = atoi(argv[6]);
    if (argc != 7 + lMax + 3) return 2;
    int *counts = (int *)calloc((size_t)lMax, sizeof(int));
    center_x = (double **)calloc((size_t)lMax, sizeof(double *));
    center_y = (double **)calloc((size_t)lMax, sizeof(double *));
    center_z = (double **)calloc((size_t)lMax, sizeof(double *));
    for (int l = 0; l < lMax; l++) {
        counts[l] = atoi(argv[7 + l]);
        /* exactly no_of_centers[l] slots: an overrun is a sanitizer report */
        const int c = counts[l] > 0 ? counts[l] : 1;
        center_x[l] = alloc0((size_t)c);
        center_y[l] = alloc0((size_t)c);
        center_z[l] = alloc0((size_t)c);
    }
    no_of_centers = counts;
    spacing_x = spacing_y = spacing_z = atof(argv[7 + lMax]);
    const size_t N = (size_t)iMax * jMax * kMax * lMax;
    double *img = alloc0(N);
    _readData_(img, (unsigned char *)dir);
    FILE *f = fopen(argv[8 + lMax], "wb");
    if (!f || fwrite(img, sizeof(double), N, f) != N) return 4;
    for (int l = 0; l < lMax; l++) {
        for (int c = 0; c < counts[l]; c++) {
            const double p[3] = {center_x[l][c], center_y[l][c], center_z[l][c]};
            if (fwrite(p, sizeof(double), 3, f) != 3) return 4;
        }
    }
    fclose(f);
    /* writers over values inside and outside [0, 1] */
    for (size_t q = 0; q < N; q++) img[q] = (double)(q % 23) / 11.0 - 0.5;
    if (!_writeData_(img, (unsigned char *)argv[9 + lMax])) return 5;
    char mpath[4096];
    snprintf(mpath, sizeof mpath, "%s_mask", argv[9 + lMax]);
    if (!seg4dWriteMaskVTK(img, mpath)) return 5;
    char fpath[4096];
    snprintf(fpath, sizeof fpath, "%s_at_time_001.vtk", argv[9 + lMax]);
    if (!_load3dDataArrayVTK_(img, 1, (unsigned char *)fpath)) return 6;
    for (int l = 0; l < lMax; l++) {
        free(center_x[l]);
        free(center_y[l]);
        free(center_z[l]);
    }
    free(center_x);
    free(center_y);
    free(center_z);
    free(counts);
    free(img);
    printf("io ok\n");
    return 0;
}

static void fill_volume(const orc_params *p, double *img, double *u0)
{
    for (int l = 0; l < p->lMax; l++)
        for (int k = 0; k < p->kMax; k++)
            for (int j = 0; j < p->jMax; j++)
                for (int i = 0; i < p->iMax; i++) {
                    const size_t q = (size_t)orc_ijkl(p, i, j, k, l);
                    const int in = i > 0 && j > 0 && k > 0 && l > 0 && i < p->iMax - 1 &&
                                   j < p->jMax - 1 && k < p->kMax - 1 && l < p->lMax - 1;
                    const double dx = i - p->iMax / 2.0, dy = j - p->jMax / 2.0,
                                 dz = k - p->kMax / 2.0;
                    const double r2 = dx * dx / 16 + dy * dy / 9 + dz * dz / 4;
                    img[q] = in ? ((r2 < 1.0 + 0.1 * l) ? 200.0 : 40.0) / 255.0 +
                                      ((i * 7 + j * 3 + k * 5 + l) % 9) / 255.0
                                : 0.0;
                    u0[q] = in ? 1.0 / (dx * dx + dy * dy + dz * dz + 1.0) : 0.0;
                }
}

static double checksum(const orc_params *p, const double *a)
{
    double acc = 0;
    for (size_t q = 0; q < orc_size(p); q++) acc += a[q] * (double)(q % 97 + 1);
    return acc;
}

static int run_oracle(void)
{
    for (int variant = 0; variant < 4; variant++) {
        orc_params p;
        memset(&p, 0, sizeof p);
        p.iMax = 19; p.jMax = 15; p.kMax = 11; p.lMax = 7;
        p.h = 1.0; p.tau = 0.5; p.eps2 = 1e-6; p.K = 1000.0;
        p.omega_c = 1.8; p.omega_sigma = 1.3; p.sigma = 0.5;
        p.gsTol = 1e-10; p.theta_img = 0.5; p.rho_thr = 0.5;
        p.maxNoGSIteration = 500; p.maxNoGSIteration_sigma = 50.0;
        p.nthreads = variant == 0 ? 1 : 3;
        p.par_reduce = variant == 0 ? 0 : variant - 1;
        const size_t N = orc_size(&p);
        double *img = alloc0(N), *u0 = alloc0(N), *thr = alloc0(N), *prev = alloc0(N);
        double *mask = alloc0(N), *g[8], *coef[8];
        for (int f = 0; f < 8; f++) {
            g[f] = alloc0(N);
            coef[f] = alloc0(N);
        }
        fill_volume(&p, img, u0);
        /* balls: inside, on the edges and beyond every face of the array */
        int count[7] = {0, 3, 2, 3, 1, 2, 0}, offset[7];
        double cx[11] = {9, 0, 18, 25, -6, 3, 9, -30, 14, 2, 17};
        double cy[11] = {7, 0, 14, 2, 20, 7, 7, 40, 1, 13, 0};
        double cz[11] = {5, 0, 10, -4, 12, 5, 5, 5, 9, 1, 10};
        for (int l = 0, o = 0; l < 7; l++) {
            offset[l] = o;
            o += count[l];
        }
        orc_balls balls = {count, offset, cx, cy, cz, 4.0};
        orc_threshold(&p, img, thr, &balls, 0.5, 0.5);
        orc_threshold(&p, thr, thr, &balls, 0.3, 0.6);  /* in place */
        double mx = orc_get4dmax(&p, img);
        double mn = orc_ball_minmax(&p, img, 2, -6.0, 20.0, 12.0, 4.0, 0);
        int it[2];
        double res[2];
        orc_gf(&p, img, thr, prev, g, it, res);
        double *uu = alloc0(N);
        memcpy(uu, u0, N * sizeof(double));
        orc_generate(&p, uu, mask, &balls);
        for (size_t q = 0; q < N; q++) mask[q] = (q % 5) ? 1.0 : 0.0;
        orc_rescale(&p, uu, mask, &balls);
        double R = 0;
        for (int st = 0; st < 2; st++) orc_seg_time_step(&p, uu, prev, g, coef, mask, &balls, &R);
        double l2 = orc_l2norm(&p, uu, prev), ms = orc_mass(&p, uu);
        unsigned char *m1 = (unsigned char *)calloc(N, 1), *m2 = (unsigned char *)calloc(N, 1);
        orc_mask(&p, uu, m1);
        orc_mask(&p, prev, m2);
        double iou = orc_iou(&p, m1, m2);
        /* MCF twin */
        orc_params pm = p;
        pm.iMax = pm.jMax = pm.kMax = pm.lMax = 9;
        pm.h = 2.5 / 7;
        pm.tau = pm.eps2 = pm.h * pm.h;
        pm.gsTol = 1e-12;
        const size_t Nm = orc_size(&pm);
        double *um = alloc0(Nm), *pv = alloc0(Nm), *cm[8], *ex = alloc0(Nm);
        for (int f = 0; f < 8; f++) cm[f] = alloc0(Nm);
        orc_mcf_exact(&pm, um, 0.0);
        for (int st = 1; st <= 2; st++) orc_mcf_time_step(&pm, um, pv, cm, st, &R);
        orc_mcf_exact(&pm, ex, 2 * pm.tau);
        orc_mcf_boundary(&pm, ex, 2 * pm.tau);
        printf("oracle variant %d: %d/%d presmooth sweeps, max %.6f, min %.6f, l2 %.6e, "
               "mass %.6e, IoU %.4f, MCF error %.6e, checksum %.12e\n",
               variant, it[0], it[1], mx, mn, l2, ms, iou, orc_l2norm(&pm, ex, um),
               checksum(&p, uu));
        for (int f = 0; f < 8; f++) {
            free(g[f]);
            free(coef[f]);
            free(cm[f]);
        }
        free(img); free(u0); free(thr); free(prev); free(mask); free(uu);
        free(m1); free(m2); free(um); free(pv); free(ex);
    }
    return 0;
}

static int run_gen(void)
{
    for (int cfg = 0; cfg <= 4; cfg++) {
        seg4d_gen_spec sp;
        if (seg4d_gen_preset(cfg, (uint64_t)(cfg + 11), &sp) != 0) return 7;
        /* the preset's objects and noise on a volume small enough to run here */
        const int full[4] = {sp.nx, sp.ny, sp.nz, sp.nt};
        sp.nx = 40; sp.ny = 36; sp.nz = 20; sp.nt = full[3] > 6 ? 6 : full[3];
        for (int o = 0; o < sp.n_objects; o++)
            for (int a = 0; a < 3; a++)
                if (sp.axes[o][a] > 0) sp.axes[o][a] = 3.0 + o % 4;
        sp.semi_min = 3.0;
        sp.semi_max = 8.0;
        sp.margin = 2.0;
        const size_t N = (size_t)(sp.nx + 2) * (sp.ny + 2) * (sp.nz + 2) * (sp.nt + 2);
        double *img = alloc0(N), *u0 = alloc0(N), c[3];
        if (seg4d_generate(&sp, img, u0, c) != 0) return 8;
        double acc = 0;
        for (size_t q = 0; q < N; q++) acc += img[q] + u0[q];
        printf("gen preset %d (full %dx%dx%dx%d): sum %.6f, seed centre %.1f %.1f %.1f\n", cfg,
               full[0], full[1], full[2], full[3], acc, c[0], c[1], c[2]);
        free(img);
        free(u0);
    }
    return 0;
}

/* Proves the sanitizer is active in this build: a read one past the end of
 * a heap array must be reported (non-zero exit). */
static int run_selfcheck(int argc)
{
    volatile double *a = alloc0(4);
    const double v = a[3 + (argc > 0)];
    free((void *)a);
    printf("selfcheck read %f without a report\n", v);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    if (!strcmp(argv[1], "selfcheck")) return run_selfcheck(argc);
    if (!strcmp(argv[1], "io")) return run_io(argc, argv);
    if (!strcmp(argv[1], "oracle")) return run_oracle();
    if (!strcmp(argv[1], "gen")) return run_gen();
    return 2;
}
