/*
 * seg4d_run.c -- C host driver of the segmentation time step on the GPU.
 *
 * The call sequence is the reference program's (segmentation4D.c:1326-1424):
 * set globals, _allocateMem_, fill the input, edge indicator
 * (_gf_inputImage_plus_thresholdedImage_), initial function, then
 * _segTimeStep_ / _l2norm_ until maxNoOfTimeSteps or ||u - u_prev|| <= segTol.
 * The input is synthetic (include/seg4d_gen.h) instead of VTK files, the
 * ball-based thresholding/rescale are disabled (no centres, Mmask = 1), and
 * the SOR tolerance is relative (seg4dSetRelativeTolerance).  Everything
 * below uses only the C ABI of libseg4d_amd.so.
 *
 *   seg4d_run [preset=0] [steps] [rel_tol=1e-6] [--resident] [--mask out.raw]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "seg4d.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    int preset = 0, steps = -1, resident = 0;
    double rel = 1e-6;
    const char *mask_path = NULL;
    int pos = 0;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--resident")) resident = 1;
        else if (!strcmp(argv[a], "--mask") && a + 1 < argc) mask_path = argv[++a];
        else if (pos == 0) { preset = atoi(argv[a]); pos++; }
        else if (pos == 1) { steps = atoi(argv[a]); pos++; }
        else if (pos == 2) { rel = atof(argv[a]); pos++; }
    }
    seg4d_gen_spec spec;
    if (seg4d_gen_preset(preset, (uint64_t)preset, &spec) != 0) {
        fprintf(stderr, "unknown preset %d\n", preset);
        return 2;
    }
    if (steps < 0) steps = preset == 0 ? 20 : 30;

    /* globals as segmentation4D.c:1332-1371 sets them, with this workload's values */
    iMax = spec.nx + 2; jMax = spec.ny + 2; kMax = spec.nz + 2; lMax = spec.nt + 2;
    no_of_centers = NULL;          /* no balls: thresholding/rescale disabled */
    ballRadius = 0;
    theta_img = 1.0; rho_thr = 0.0;
    maxNoGSIteration = 100000; maxNoOfTimeSteps = steps; segTol = 0.0;
    omega_c = 1.8; omega_sigma = 1.3;
    gsTol = 0.0; h = 1.0; tau = 0.5; eps2 = 1e-6; K = 1000.0;
    sigma = 0.5; maxNoGSIteration_sigma = 100000;
    seg4dSetRelativeTolerance(rel);

    if (!_allocateMem_()) {
        fprintf(stderr, "allocation failed: %s\n", seg4dLastError());
        return 1;
    }
    const size_t N = (size_t)iMax * jMax * kMax * lMax;
    double *u0 = (double *)malloc(N * sizeof(double));
    double centre[3];
    if (!u0 || seg4d_generate(&spec, i0, u0, centre) != 0) {
        fprintf(stderr, "input generation failed\n");
        return 1;
    }
    for (size_t q = 0; q < N; q++) Mmask[q] = 1.0;  /* thresholded image u stays 0 */

    double t0 = now();
    if (!_gf_inputImage_plus_thresholdedImage_(i0, u)) return 1;
    double t1 = now();
    memcpy(u, u0, N * sizeof(double));
    if (resident) {
        seg4dSetResident(1);
        seg4dUpload(u);
        seg4dUpload(Mmask);
    }
    double diff = 0;
    int tCount = 1;
    long sweeps = 0;
    double t2 = now();
    do {
        numberOfTimeStep = tCount;
        if (!_segTimeStep_()) return 1;
        sweeps += seg4dLastSweeps();
        diff = _l2norm_(prev_u, u);
        printf("l2 norm of difference at step %d is %e\n", tCount, diff);
        tCount++;
    } while (tCount <= maxNoOfTimeSteps && diff > segTol);
    seg4dSynchronize();
    double t3 = now();
    if (resident) seg4dDownload(u);
    const double nint = (double)spec.nx * spec.ny * spec.nz * spec.nt;
    printf("edge indicator %.3f s; %d steps, %ld SOR sweeps in %.3f s: %.3e voxel-updates/s\n",
           t1 - t0, tCount - 1, sweeps, t3 - t2, nint * sweeps / (t3 - t2));

    if (mask_path) {  /* u >= (max+min)/2 over the interior, as bytes */
        double mx = -1e300, mn = 1e300;
        for (int l = 1; l <= lMax - 2; l++)
            for (int k = 1; k <= kMax - 2; k++)
                for (int j = 1; j <= jMax - 2; j++)
                    for (int i = 1; i <= iMax - 2; i++) {
                        double v = u[ijkl(i, j, k, l)];
                        if (v > mx) mx = v;
                        if (v < mn) mn = v;
                    }
        FILE *f = fopen(mask_path, "wb");
        if (!f) return 1;
        const double thr = 0.5 * (mx + mn);
        for (int l = 1; l <= lMax - 2; l++)
            for (int k = 1; k <= kMax - 2; k++)
                for (int j = 1; j <= jMax - 2; j++)
                    for (int i = 1; i <= iMax - 2; i++)
                        fputc(u[ijkl(i, j, k, l)] >= thr, f);
        fclose(f);
    }
    free(u0);
    _deallocateMem_();
    return 0;
}
