/*
 * mcf4d_run.c -- C host driver of the 4D MCF verification on the GPU.
 *
 * Same call sequence as the reference's MCF programs (mcf_4D.c:473-524,
 * mcf_4D_openmp.c:487-539): exact solution at t = 0 copied to u, then per
 * step _mcfTimeStep, _exactSolution_mcf(t), _l2norm_mcf(exact_u, u).  Only the
 * C ABI of libmcf4d_amd.so is used.
 *
 *   mcf4d_run [n_interior=80] [steps=64]
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "mcf4d.h"

int main(int argc, char **argv)
{
    const int nint = argc > 1 ? atoi(argv[1]) : 80;
    const int steps = argc > 2 ? atoi(argv[2]) : 64;
    iMax = jMax = kMax = lMax = nint + 2;
    maxNoGSIteration = 1000;
    maxNoOfTimeSteps = steps;
    mcfTol = 1.0e-12;
    omega_c = 1.8;
    gsTol = 1.0e-12;
    h = 2.5 / (iMax - 2);
    tau = h * h;
    eps2 = h * h;
    mcf4dSetVerbose(0);
    if (!_allocateMem_mcf()) {
        fprintf(stderr, "allocation failed: %s\n", mcf4dLastError());
        return 1;
    }
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    _exactSolution_mcf(0.);
    _copy4DdataFromSrcToDest_mcf(exact_u, u);
    int tCount = 1;
    double err = 0;
    do {
        numberOfTimeStep = tCount;
        double time = numberOfTimeStep * tau;
        if (!_mcfTimeStep()) return 1;
        _exactSolution_mcf(time);
        err = _l2norm_mcf(exact_u, u);
        printf("l2 norm of error at time %lf is %e (%d sweeps)\n", time, err, mcf4dLastSweeps());
        tCount++;
    } while (tCount <= maxNoOfTimeSteps);
    clock_gettime(CLOCK_MONOTONIC, &b);
    printf("wall time is: %e secs\n", (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec));
    _deallocateMem_mcf();
    return 0;
}
