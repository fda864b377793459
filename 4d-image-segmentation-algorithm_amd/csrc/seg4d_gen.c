/*
 * seg4d_gen.c -- deterministic synthetic 4D volumes (see include/seg4d_gen.h).
 *
 * Randomness: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014; constants as in Vigna's public-domain
 * splitmix64.c), used as a counter-based generator: the n-th draw of stream
 * s under seed S is mix(mix(S + (s+1) gamma) + (n+1) gamma).  Normals use
 * the cosine branch of Box-Muller on draws 2n and 2n+1.
 */
#include "seg4d_gen.h"

#include <math.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ULL

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t seg4d_splitmix64_next(uint64_t *state)
{
    *state += GOLDEN;
    return mix64(*state);
}

uint64_t seg4d_rng_u64(uint64_t seed, uint64_t stream, uint64_t n)
{
    uint64_t key = mix64(seed + (stream + 1) * GOLDEN);
    return mix64(key + (n + 1) * GOLDEN);
}

static double u01_open0(uint64_t x) /* (0, 1] */
{
    return (double)((x >> 11) + 1) * 0x1.0p-53;
}

static double u01(uint64_t x) /* [0, 1) */
{
    return (double)(x >> 11) * 0x1.0p-53;
}

double seg4d_rng_normal(uint64_t seed, uint64_t stream, uint64_t n)
{
    double a = u01_open0(seg4d_rng_u64(seed, stream, 2 * n));
    double b = u01(seg4d_rng_u64(seed, stream, 2 * n + 1));
    return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * b);
}

enum { STREAM_OBJECTS = 0, STREAM_NOISE = 1, STREAM_BATCH = 2 };

int seg4d_gen_preset(int cfg, uint64_t seed, seg4d_gen_spec *s)
{
    if (!s) return -1;
    memset(s, 0, sizeof(*s));
    s->seed = seed;
    s->background = 40.0;
    s->foreground = 200.0;
    s->wedge_deg = 30.0;
    s->shell = 2.0;
    s->margin = -1.0;
    switch (cfg) {
    case 0:
        s->nx = 64; s->ny = 64; s->nz = 32; s->nt = 8;
        s->n_objects = 2;
        s->axes[0][0] = 14; s->axes[0][1] = 10; s->axes[0][2] = 8;
        s->axes[1][0] = 8;  s->axes[1][1] = 8;  s->axes[1][2] = 6;
        s->amp = 0.15; s->wedges = 3; s->noise_sigma = 20.0;
        return 0;
    case 1:
    case 2:
        s->nx = 256; s->ny = 256; s->nz = 128; s->nt = 16;
        s->n_objects = 6; s->semi_min = 12; s->semi_max = 40; s->margin = 20;
        s->amp = 0.2; s->wedges = 4; s->noise_sigma = 25.0;
        return 0;
    case 3: {
        /* one member of the uneven batch: x,y in {48..192 step 16},
         * z in {32..128 step 16}, t in {6..16}, 2-6 objects */
        s->nx = 48 + 16 * (int)(seg4d_rng_u64(seed, STREAM_BATCH, 0) % 10);
        s->ny = 48 + 16 * (int)(seg4d_rng_u64(seed, STREAM_BATCH, 1) % 10);
        s->nz = 32 + 16 * (int)(seg4d_rng_u64(seed, STREAM_BATCH, 2) % 7);
        s->nt = 6 + (int)(seg4d_rng_u64(seed, STREAM_BATCH, 3) % 11);
        s->n_objects = 2 + (int)(seg4d_rng_u64(seed, STREAM_BATCH, 4) % 5);
        int mn = s->nx < s->ny ? s->nx : s->ny;
        if (s->nz < mn) mn = s->nz;
        s->semi_min = 6; s->semi_max = 0.22 * mn > 8 ? 0.22 * mn : 8;
        s->amp = 0.15; s->wedges = 3; s->noise_sigma = 20.0;
        return 0;
    }
    case 4:
        s->nx = 512; s->ny = 512; s->nz = 256; s->nt = 8;
        s->n_objects = 12; s->semi_min = 16; s->semi_max = 64; s->margin = 20;
        s->amp = 0.2; s->wedges = 4; s->noise_sigma = 25.0;
        return 0;
    default:
        return -1;
    }
}

typedef struct {
    double a, b, c;          /* base semi-axes */
    double cx, cy, cz;       /* centre, padded index coordinates */
    double wedge[SEG4D_GEN_MAX_OBJECTS];
} object_t;

static double place_centre(int n, double margin, double u)
{
    /* interior occupies padded coordinates 1..n */
    if (2.0 * margin >= (double)(n - 1)) return 0.5 * (double)(n + 1);
    return 1.0 + margin + u * ((double)(n - 1) - 2.0 * margin);
}

static void draw_objects(const seg4d_gen_spec *s, object_t *obj)
{
    uint64_t ctr = 0;
    for (int o = 0; o < s->n_objects; o++) {
        object_t *b = &obj[o];
        if (s->axes[o][0] > 0) {
            b->a = s->axes[o][0]; b->b = s->axes[o][1]; b->c = s->axes[o][2];
        } else {
            double span = s->semi_max - s->semi_min;
            b->a = s->semi_min + span * u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++));
            b->b = s->semi_min + span * u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++));
            b->c = s->semi_min + span * u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++));
        }
        double big = b->a > b->b ? b->a : b->b;
        if (b->c > big) big = b->c;
        double mg = s->margin >= 0 ? s->margin : big * (1.0 + s->amp) + 2.0;
        b->cx = place_centre(s->nx, mg, u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++)));
        b->cy = place_centre(s->ny, mg, u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++)));
        b->cz = place_centre(s->nz, mg, u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++)));
        for (int w = 0; w < s->wedges && w < SEG4D_GEN_MAX_OBJECTS; w++)
            b->wedge[w] = 2.0 * M_PI * u01(seg4d_rng_u64(s->seed, STREAM_OBJECTS, ctr++));
    }
}

static int in_wedge(const seg4d_gen_spec *s, const object_t *b, double phi)
{
    const double half = 0.5 * s->wedge_deg * M_PI / 180.0;
    for (int w = 0; w < s->wedges; w++) {
        double d = fabs(phi - b->wedge[w]);
        d = fmod(d, 2.0 * M_PI);
        if (d > M_PI) d = 2.0 * M_PI - d;
        if (d <= half) return 1;
    }
    return 0;
}

int seg4d_generate(const seg4d_gen_spec *s, double *image, double *u0,
                   double centre_out[3])
{
    if (!s || !image || s->nx < 1 || s->ny < 1 || s->nz < 1 || s->nt < 1 ||
        s->n_objects < 1 || s->n_objects > SEG4D_GEN_MAX_OBJECTS ||
        s->wedges < 0 || s->wedges > SEG4D_GEN_MAX_OBJECTS)
        return -1;
    const long I = s->nx + 2, J = s->ny + 2, K = s->nz + 2, L = s->nt + 2;
    const long N = I * J * K * L;
    object_t obj[SEG4D_GEN_MAX_OBJECTS];
    draw_objects(s, obj);
    if (centre_out) {
        centre_out[0] = obj[0].cx; centre_out[1] = obj[0].cy; centre_out[2] = obj[0].cz;
    }

    #pragma omp parallel for schedule(static)
    for (long q = 0; q < N; q++) {
        image[q] = 0.0;
        if (u0) u0[q] = 0.0;
    }

    #pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (long l = 1; l <= s->nt; l++)
        for (long k = 1; k <= s->nz; k++) {
            const double tt = (double)(l - 1);
            const double scale = 1.0 + s->amp * sin(2.0 * M_PI * tt / (double)s->nt);
            for (long j = 1; j <= s->ny; j++) {
                double *row = image + ((l * K + k) * J + j) * I;
                for (long i = 1; i <= s->nx; i++) row[i] = s->background;
                for (int o = 0; o < s->n_objects; o++) {
                    const object_t *b = &obj[o];
                    const double a = b->a * scale, bb = b->b * scale, c = b->c * scale;
                    const double dz = (double)k - b->cz, dy = (double)j - b->cy;
                    const double ryz = (dy / bb) * (dy / bb) + (dz / c) * (dz / c);
                    if (ryz > 1.0) continue;
                    const double half = a * sqrt(1.0 - ryz);
                    long i0 = (long)ceil(b->cx - half), i1 = (long)floor(b->cx + half);
                    if (i0 < 1) i0 = 1;
                    if (i1 > s->nx) i1 = s->nx;
                    double mn = a < bb ? a : bb;
                    if (c < mn) mn = c;
                    const double inner = 1.0 - s->shell / mn;
                    for (long i = i0; i <= i1; i++) {
                        const double dx = (double)i - b->cx;
                        const double r2 = (dx / a) * (dx / a) + ryz;
                        if (r2 > 1.0) continue;
                        if (s->wedges > 0 && sqrt(r2) > inner &&
                            in_wedge(s, b, atan2(dy, dx)))
                            continue;
                        row[i] = s->foreground;
                    }
                }
                for (long i = 1; i <= s->nx; i++) {
                    const uint64_t n = (uint64_t)((((l - 1) * s->nz + (k - 1)) * s->ny +
                                                   (j - 1)) * s->nx + (i - 1));
                    double v = row[i] + s->noise_sigma * seg4d_rng_normal(s->seed, STREAM_NOISE, n);
                    if (v < 0.0) v = 0.0;
                    if (v > 255.0) v = 255.0;
                    row[i] = floor(v + 0.5) / 255.0;
                }
                if (u0) {
                    /* seed function 1/(|x - s| + 1), 3-D distance per frame */
                    double *urow = u0 + ((l * K + k) * J + j) * I;
                    const double ey = (double)j - obj[0].cy, ez = (double)k - obj[0].cz;
                    for (long i = 1; i <= s->nx; i++) {
                        const double ex = (double)i - obj[0].cx;
                        urow[i] = 1.0 / (sqrt(ex * ex + ey * ey + ez * ez) + 1.0);
                    }
                }
            }
        }
    return 0;
}
