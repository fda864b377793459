/*
 * mcf4d_run.c -- C host driver of the 4D MCF verification on the GPU.
 *
 * Same call sequence as the reference's MCF programs (mcf_4D.c:473-524,
 * mcf_4D_openmp.c:487-539): exact solution at t = 0 copied to u, then per
 * step _mcfTimeStep, _exactSolution_mcf(t), _l2norm_mcf(exact_u, u), with
 * the reference's parameters (h = 2.5/n, tau = eps2 = h^2, omega 1.8,
 * gsTol 1e-12, at most 1000 sweeps).  Only the C ABI of libmcf4d_amd.so is
 * used.
 *
 *   mcf4d_run [n_interior=80] [steps=64]
 *       one run, the error printed after every step (the reference's output)
 *   mcf4d_run eoc [n0=20] [levels=4] [T=0.0625]
 *       n = n0, 2 n0, 4 n0, ... to the same final time T (T / h^2 steps,
 *       which must be whole), device-resident; per level the L2 error at T
 *       and the experimental order of convergence log2(e(n/2) / e(n)).
 *       The reference's own run is the level n = 160: 256 steps to 0.0625.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mcf4d.h"

static double now(void)
{
    struct timespec a;
    clock_gettime(CLOCK_MONOTONIC, &a);
    return a.tv_sec + 1e-9 * a.tv_nsec;
}

static void set_params(int nint, int steps)
{
    iMax = jMax = kMax = lMax = nint + 2;
    maxNoGSIteration = 1000;
    maxNoOfTimeSteps = steps;
    mcfTol = 1.0e-12;
    omega_c = 1.8;
    gsTol = 1.0e-12;
    h = 2.5 / (iMax - 2);
    tau = h * h;
    eps2 = h * h;
}

/* One run of the reference's loop.  Returns 0 on success; err = error after
 * the last step, sweeps = SOR iterations of all steps. */
static int run(int nint, int steps, int resident, int print_steps, double *err, long *sweeps,
               double *wall)
{
    set_params(nint, steps);
    mcf4dSetVerbose(0);
    if (!_allocateMem_mcf()) {
        fprintf(stderr, "allocation failed: %s\n", mcf4dLastError());
        return 1;
    }
    mcf4dSetResident(resident);
    const double t0 = now();
    _exactSolution_mcf(0.);
    _copy4DdataFromSrcToDest_mcf(exact_u, u);
    int tCount = 1;
    *sweeps = 0;
    do {
        numberOfTimeStep = tCount;
        const double time = numberOfTimeStep * tau;
        if (!_mcfTimeStep()) {
            fprintf(stderr, "step %d failed: %s\n", tCount, mcf4dLastError());
            return 1;
        }
        *sweeps += mcf4dLastSweeps();
        _exactSolution_mcf(time);
        *err = _l2norm_mcf(exact_u, u);
        if (print_steps) {
            printf("l2 norm of error at time %lf is %e (%d sweeps)\n", time, *err,
                   mcf4dLastSweeps());
            fflush(stdout);
        }
        tCount++;
    } while (tCount <= maxNoOfTimeSteps);
    mcf4dSynchronize();
    *wall = now() - t0;
    mcf4dSetResident(0);
    _deallocateMem_mcf();
    return 0;
}

static int eoc(int n0, int levels, double T)
{
    printf("%6s %6s %14s %7s %9s %10s %14s\n", "n", "steps", "L2 error at T", "EOC", "sweeps",
           "wall s", "voxel-upd/s");
    double prev = 0.0;
    for (int q = 0; q < levels; q++) {
        const int n = n0 << q;
        const double hh = 2.5 / n;
        const double sf = T / (hh * hh);
        const int steps = (int)(sf + 0.5);
        if (steps < 1 || fabs(sf - steps) > 1e-9 * sf) {
            fprintf(stderr, "T / h^2 = %.12g is not a whole number of steps at n = %d\n", sf, n);
            return 1;
        }
        double err = 0, wall = 0;
        long sweeps = 0;
        printf("[level] n = %d, %d steps\n", n, steps);
        fflush(stdout);
        if (run(n, steps, 1, n >= 80, &err, &sweeps, &wall)) return 1;
        const double ups = (double)n * n * n * n * (double)sweeps / wall;
        if (q == 0)
            printf("%6d %6d %14.6e %7s %9ld %10.3f %14.4e\n", n, steps, err, "-", sweeps, wall,
                   ups);
        else
            printf("%6d %6d %14.6e %7.3f %9ld %10.3f %14.4e\n", n, steps, err,
                   log2(prev / err), sweeps, wall, ups);
        fflush(stdout);
        prev = err;
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && strcmp(argv[1], "eoc") == 0) {
        const int n0 = argc > 2 ? atoi(argv[2]) : 20;
        const int levels = argc > 3 ? atoi(argv[3]) : 4;
        const double T = argc > 4 ? atof(argv[4]) : 0.0625;
        if (n0 < 1 || levels < 1 || levels > 8 || !(T > 0)) {
            fprintf(stderr, "usage: mcf4d_run eoc [n0>=1] [1<=levels<=8] [T>0]\n");
            return 2;
        }
        return eoc(n0, levels, T);
    }
    const int nint = argc > 1 ? atoi(argv[1]) : 80;
    const int steps = argc > 2 ? atoi(argv[2]) : 64;
    if (nint < 1 || steps < 1) {
        fprintf(stderr, "usage: mcf4d_run [n_interior>=1] [steps>=1]\n");
        return 2;
    }
    double err = 0, wall = 0;
    long sweeps = 0;
    if (run(nint, steps, 0, 1, &err, &sweeps, &wall)) return 1;
    printf("wall time is: %e secs\n", wall);
    return 0;
}
