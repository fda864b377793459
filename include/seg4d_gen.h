/*
 * seg4d_gen.h -- deterministic synthetic 4D test volumes (x, y, z, t).
 *
 * The reference reads VTK frames from fixed paths (segmentation4D.c:1149-1213);
 * the workloads of this project are synthetic instead: ellipsoids of
 * intensity 200 on a background of 40, with time-periodic radii, azimuthal
 * wedge gaps cut out of their boundary shells and i.i.d. Gaussian noise,
 * quantised to uint8 and divided by 255 as the reference's loader does
 * (segmentation4D.c:1320).  All randomness is counter-based (SplitMix64 +
 * Box-Muller), so the result is independent of thread count.
 *
 * Arrays are padded, x-fastest FP64 (iMax = nx+2, ...), ghosts left at 0.
 */
#ifndef SEG4D_GEN_H
#define SEG4D_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEG4D_GEN_MAX_OBJECTS 64

typedef struct {
    int nx, ny, nz, nt;            /* interior dims                         */
    int n_objects;                 /* number of ellipsoids                  */
    double axes[SEG4D_GEN_MAX_OBJECTS][3]; /* fixed semi-axes (if > 0)      */
    double semi_min, semi_max;     /* otherwise drawn uniformly from this   */
    double margin;                 /* centre margin in voxels; < 0 = auto   */
    double amp;                    /* radius modulation 1 + amp sin(2 pi t/T) */
    int wedges;                    /* wedge gaps per object                 */
    double wedge_deg;              /* full opening angle of one wedge       */
    double shell;                  /* thickness of the cut shell (voxels)   */
    double noise_sigma;            /* Gaussian noise (intensity units)      */
    double background, foreground; /* 40, 200                               */
    uint64_t seed;
} seg4d_gen_spec;

/* Fills `spec` with preset `cfg` (0..4) of the benchmark configurations:
 * 0: 64x64x32x8, 2 objects (14/10/8, 8/8/6), amp .15, 3 wedges, sigma 20
 * 1: 256x256x128x16, 6 objects U[12,40], margin 20, amp .2, 4 wedges, sigma 25
 * 2: same volume as 1
 * 3: one member of the batched set (sizes drawn from the seed)
 * 4: 512x512x256x8, 12 objects, sigma 25.  Returns 0 on success. */
int seg4d_gen_preset(int cfg, uint64_t seed, seg4d_gen_spec *spec);

/* Generates the image (scaled to [0,1]) and, if u0 != NULL, the initial
 * function u0 = 1/(|x - s| + 1) with s the first object's centre (3-D
 * distance per frame).  Writes the first object's centre (padded index
 * coordinates) to centre_out[3] if non-NULL.  Returns 0 on success. */
int seg4d_generate(const seg4d_gen_spec *spec, double *image, double *u0,
                   double centre_out[3]);

/* SplitMix64 primitives (exposed for known-answer tests). */
uint64_t seg4d_splitmix64_next(uint64_t *state);
uint64_t seg4d_rng_u64(uint64_t seed, uint64_t stream, uint64_t n);
double seg4d_rng_normal(uint64_t seed, uint64_t stream, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
