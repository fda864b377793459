/*
 * seg4d_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see the header).
 *
 * Each function names the reference lines it restates.  Files are in
 * 88MARK08/4D-image-segmentation-algorithm:
 *   segmentation4D.c   (serial segmentation program)
 *   mcf_4D.c           (serial MCF verification program)
 *   mcf_4D_openmp.c    (OpenMP MCF program: loop/pragma structure)
 *
 * Bitwise parity with the reference rests on three facts:
 *   - every sum below is evaluated in the reference's left-to-right order
 *     (IEEE addition is commutative but not associative, so only grouping
 *     matters -- swapped operands inside one addition are harmless);
 *   - the build uses -ffp-contract=off (no fused multiply-add);
 *   - pow(x,2) in the reference is folded to x*x by gcc.
 * The OpenMP branches parallelise the outer l loop exactly like the
 * reference OpenMP program; SOR colour passes and coefficient passes are
 * order-independent, only the residual sum changes order (as in the
 * reference OpenMP program).
 */
#include "seg4d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define NT(p) ((p)->nthreads > 1 ? (p)->nthreads : 1)
#define NR(p) ((p)->par_reduce == 1 ? NT(p) : 1)
/* par_reduce == 2 with threads: the terms of a residual sum are formed in
 * parallel into a buffer and added up serially in lexicographic order, so
 * the sum is bitwise the serial one (the threaded oracle then gives the
 * serial oracle's sweep counts at any size). */
#define TERMS(p) ((p)->par_reduce == 2 && NT(p) > 1)

/* interior-sized term buffer of the TERMS mode (test infrastructure: one
 * caller at a time, kept between calls) */
static double *terms_buffer(size_t n)
{
    static double *buf = NULL;
    static size_t cap = 0;
    if (n > cap) {
        free(buf);
        buf = (double *)malloc(n * sizeof(double));
        cap = buf ? n : 0;
    }
    return buf;
}

static size_t interior_index(const orc_params *p, int i, int j, int k, int l)
{
    return (((size_t)(l - 1) * (p->kMax - 2) + (k - 1)) * (p->jMax - 2) + (j - 1)) *
               (p->iMax - 2) + (i - 1);
}

static size_t interior_size(const orc_params *p)
{
    return (size_t)(p->iMax - 2) * (p->jMax - 2) * (p->kMax - 2) * (p->lMax - 2);
}

/* serial left-to-right sum of n terms: acc = ((0 + t0) + t1) + ... */
static double serial_sum(const double *t, size_t n)
{
    double acc = 0.0;
    for (size_t q = 0; q < n; q++) acc += t[q];
    return acc;
}

size_t orc_size(const orc_params *p)
{
    return (size_t)p->iMax * p->jMax * p->kMax * p->lMax;
}

/* segmentation4D.c:61-65 -- padded x-fastest index */
int orc_ijkl(const orc_params *p, int i, int j, int k, int l)
{
    return ((l * p->kMax + k) * p->jMax + j) * p->iMax + i;
}

/* segmentation4D.c:67-70 */
double orc_gfunction(double value, double coef)
{
    return 1.0 / (1 + coef * value);
}

/* segmentation4D.c:866-874 (copy over the whole padded array) */
void orc_copy(const orc_params *p, const double *src, double *dst)
{
    memcpy(dst, src, orc_size(p) * sizeof(double));
}

/* segmentation4D.c:194-213 -- sqrt(sum over padded array of (a-b)^2 h^4),
 * lexicographic order.  Returns 0 for NULL inputs (the reference returns
 * `false`). */
double orc_l2norm(const orc_params *p, const double *a, const double *b)
{
    if (!a || !b) return 0.0;
    const double h4 = p->h * p->h * p->h * p->h;
    const size_t n = orc_size(p);
    double acc = 0;
    if (NR(p) > 1) {
        const size_t slab = (size_t)p->iMax * p->jMax * p->kMax;
        #pragma omp parallel for num_threads(NR(p)) reduction(+:acc) schedule(static)
        for (int l = 0; l < p->lMax; l++)
            for (size_t q = (size_t)l * slab; q < (size_t)(l + 1) * slab; q++) {
                double d = a[q] - b[q];
                acc += d * d * h4;
            }
    } else {
        for (size_t q = 0; q < n; q++) {
            double d = a[q] - b[q];
            acc += d * d * h4;
        }
    }
    return sqrt(acc);
}

double orc_sumsq_interior(const orc_params *p, const double *a)
{
    const double h4 = p->h * p->h * p->h * p->h;
    double acc = 0;
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    double v = a[orc_ijkl(p, i, j, k, l)];
                    acc += v * v * h4;
                }
    return acc;
}

/* segmentation4D.c:826-864 -- ghost shell copies in X, then Y, Z, W order.
 * Each later pass overwrites edges/corners written by an earlier one, so the
 * result equals u[clamp(i),clamp(j),clamp(k),clamp(l)] on every ghost. */
void orc_reflection(const orc_params *p, double *u)
{
    const int I = p->iMax, J = p->jMax, K = p->kMax, L = p->lMax;
    for (int l = 0; l < L; l++)
        for (int k = 0; k < K; k++)
            for (int j = 0; j < J; j++) {
                u[orc_ijkl(p, 0, j, k, l)] = u[orc_ijkl(p, 1, j, k, l)];
                u[orc_ijkl(p, I - 1, j, k, l)] = u[orc_ijkl(p, I - 2, j, k, l)];
            }
    for (int l = 0; l < L; l++)
        for (int k = 0; k < K; k++)
            for (int i = 0; i < I; i++) {
                u[orc_ijkl(p, i, 0, k, l)] = u[orc_ijkl(p, i, 1, k, l)];
                u[orc_ijkl(p, i, J - 1, k, l)] = u[orc_ijkl(p, i, J - 2, k, l)];
            }
    for (int l = 0; l < L; l++)
        for (int j = 0; j < J; j++)
            for (int i = 0; i < I; i++) {
                u[orc_ijkl(p, i, j, 0, l)] = u[orc_ijkl(p, i, j, 1, l)];
                u[orc_ijkl(p, i, j, K - 1, l)] = u[orc_ijkl(p, i, j, K - 2, l)];
            }
    for (int k = 0; k < K; k++)
        for (int i = 0; i < I; i++)
            for (int j = 0; j < J; j++) {
                u[orc_ijkl(p, i, j, k, 0)] = u[orc_ijkl(p, i, j, k, 1)];
                u[orc_ijkl(p, i, j, k, L - 1)] = u[orc_ijkl(p, i, j, k, L - 2)];
            }
}

/* Neighbour offsets in the numerator order of segmentation4D.c:242-249:
 * +x, -x, -y, +y, -z, +z, -t, +t (paired with e, w, s, n, b, t, bw, fw). */
static void nb_offsets(const orc_params *p, long off[8])
{
    const long sy = p->iMax, sz = (long)p->iMax * p->jMax,
               sw = (long)p->iMax * p->jMax * p->kMax;
    off[0] = 1;   off[1] = -1;
    off[2] = -sy; off[3] = sy;
    off[4] = -sz; off[5] = sz;
    off[6] = -sw; off[7] = sw;
}

/* One constant-coefficient SOR colour pass of the implicit heat step,
 * segmentation4D.c:355-377 (red) / 380-401 (black). */
static void heat_pass(const orc_params *p, double *u, const double *prev,
                      double c, int colour)
{
    long off[8];
    nb_offsets(p, off);
    const double om = p->omega_sigma;
    #pragma omp parallel for num_threads(NT(p)) if (NT(p) > 1) schedule(static)
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    if ((i + j + k + l) % 2 != colour) continue;
                    const long q = orc_ijkl(p, i, j, k, l);
                    double nb = u[q + off[0]] + u[q + off[1]] + u[q + off[2]] +
                                u[q + off[3]] + u[q + off[4]] + u[q + off[5]] +
                                u[q + off[6]] + u[q + off[7]];
                    double gs = (prev[q] + c * nb) / (1 + c * 8);
                    u[q] = u[q] + om * (gs - u[q]);
                }
}

/* segmentation4D.c:392-403 */
static double heat_residual(const orc_params *p, const double *u,
                            const double *prev, double c)
{
    long off[8];
    nb_offsets(p, off);
    const double h4 = p->h * p->h * p->h * p->h;
    double *terms = TERMS(p) ? terms_buffer(interior_size(p)) : NULL;
    double acc = 0.0;
    #pragma omp parallel for num_threads(terms ? NT(p) : NR(p)) if (terms || NR(p) > 1) \
        reduction(+:acc) schedule(static)
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    const long q = orc_ijkl(p, i, j, k, l);
                    double nb = u[q + off[0]] + u[q + off[1]] + u[q + off[2]] +
                                u[q + off[3]] + u[q + off[4]] + u[q + off[5]] +
                                u[q + off[6]] + u[q + off[7]];
                    double r = u[q] * (1 + c * 8) - c * nb - prev[q];
                    if (terms) terms[interior_index(p, i, j, k, l)] = r * r * h4;
                    else acc += r * r * h4;
                }
    return terms ? serial_sum(terms, interior_size(p)) : acc;
}

/* segmentation4D.c:337-409 -- copy to prev, reflect once (ghosts then stay
 * frozen), red-black SOR while (R > gsTol && z < cap). */
int orc_presmooth(const orc_params *p, double *u, double *prev, double *res_out)
{
    const double c = p->sigma / (p->h * p->h);
    int z = 0;
    double R;
    orc_copy(p, u, prev);
    orc_reflection(p, u);
    do {
        z = z + 1;
        heat_pass(p, u, prev, c, 0);
        heat_pass(p, u, prev, c, 1);
        R = heat_residual(p, u, prev, c);
    } while (R > p->gsTol && z < p->maxNoGSIteration_sigma);
    if (res_out) *res_out = R;
    return z;
}

/* Gradient components of face `face` at point q (segmentation4D.c:456-560
 * for u; :596-820 for the image).  Axis A = face/2, direction d = +1 for
 * even faces, -1 for odd.  Normal component: (U(q+dA) - U(q)) / h.
 * Transverse component along B: ((U(q+dA+B) + U(q+B)) - (U(q+dA-B) + U(q-B)))
 * / (4h).  The reference writes some pairs as (U(q+B) + U(q+dA+B)); the two
 * operands of one addition commute exactly, so the grouping is identical. */
static void face_components(const orc_params *p, const double *u, long q,
                            int face, double comp[4])
{
    const long st[4] = {1, p->iMax, (long)p->iMax * p->jMax,
                        (long)p->iMax * p->jMax * p->kMax};
    const int A = face >> 1;
    const long dA = (face & 1) ? -st[A] : st[A];
    const double quotient = 4.0 * p->h;
    for (int B = 0; B < 4; B++) {
        if (B == A) {
            comp[B] = (u[q + dA] - u[q]) / p->h;
        } else {
            comp[B] = ((u[q + dA + st[B]] + u[q + st[B]]) -
                       (u[q + dA - st[B]] + u[q - st[B]])) / quotient;
        }
    }
}

double orc_face_norm(const orc_params *p, const double *u, int i, int j, int k,
                     int l, int face)
{
    double c[4];
    face_components(p, u, orc_ijkl(p, i, j, k, l), face, c);
    return sqrt((c[0] * c[0]) + (c[1] * c[1]) + (c[2] * c[2]) + (c[3] * c[3]) +
                p->eps2);
}

/* segmentation4D.c:582-823 -- presmooth both inputs in place, then
 * g_f = 1/(1 + K (theta |grad I0|_f^2 + rho |grad Ithr|_f^2)) on 8 faces. */
void orc_gf(const orc_params *p, double *i0, double *i0_thr, double *prev,
            double *g[8], int *iters_out, double *res_out)
{
    double r0 = 0, r1 = 0;
    int z0 = orc_presmooth(p, i0, prev, &r0);
    int z1 = orc_presmooth(p, i0_thr, prev, &r1);
    if (iters_out) { iters_out[0] = z0; iters_out[1] = z1; }
    if (res_out) { res_out[0] = r0; res_out[1] = r1; }
    #pragma omp parallel for num_threads(NT(p)) if (NT(p) > 1) schedule(static)
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    const long q = orc_ijkl(p, i, j, k, l);
                    for (int f = 0; f < 8; f++) {
                        double a[4], b[4];
                        face_components(p, i0, q, f, a);
                        face_components(p, i0_thr, q, f, b);
                        double va = (a[0] * a[0]) + (a[1] * a[1]) + (a[2] * a[2]) + (a[3] * a[3]);
                        double vb = (b[0] * b[0]) + (b[1] * b[1]) + (b[2] * b[2]) + (b[3] * b[3]);
                        g[f][q] = orc_gfunction(p->theta_img * va + p->rho_thr * vb, p->K);
                    }
                }
}

/* segmentation4D.c:446-580 (g != NULL) and mcf_4D.c:208-342 (g == NULL):
 * N_f = sqrt(|grad u|_f^2 + eps2); avg = (N_e+...+N_bw)/8;
 * vc = sqrt(avg^2 + eps2); coef_f = vc * g_f * (1/N_f). */
void orc_coefficients(const orc_params *p, const double *u, double *const g[8],
                      double *coef[8])
{
    #pragma omp parallel for num_threads(NT(p)) if (NT(p) > 1) schedule(static)
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    const long q = orc_ijkl(p, i, j, k, l);
                    double N[8];
                    for (int f = 0; f < 8; f++) {
                        double c[4];
                        face_components(p, u, q, f, c);
                        N[f] = sqrt((c[0] * c[0]) + (c[1] * c[1]) + (c[2] * c[2]) +
                                    (c[3] * c[3]) + p->eps2);
                    }
                    double avg = (N[0] + N[1] + N[2] + N[3] + N[4] + N[5] + N[6] + N[7]) / 8.0;
                    double vc = sqrt(avg * avg + p->eps2);
                    for (int f = 0; f < 8; f++)
                        coef[f][q] = g ? vc * g[f][q] * (1.0 / N[f]) : vc * (1.0 / N[f]);
                }
}

/* coefficient indices in the numerator order e, w, s, n, b, t, bw, fw */
static const int NUM_ORDER[8] = {0, 1, 3, 2, 5, 4, 7, 6};

/* segmentation4D.c:233-258 (red) / 260-285 (black) */
void orc_sor_pass(const orc_params *p, double *u, const double *prev,
                  double *const coef[8], int colour)
{
    long off[8];
    nb_offsets(p, off);
    const double c = p->tau / (p->h * p->h);
    const double om = p->omega_c;
    #pragma omp parallel for num_threads(NT(p)) if (NT(p) > 1) schedule(static)
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    if ((i + j + k + l) % 2 != colour) continue;
                    const long q = orc_ijkl(p, i, j, k, l);
                    double num = 0;
                    for (int m = 0; m < 8; m++) {
                        double term = coef[NUM_ORDER[m]][q] * u[q + off[m]];
                        num = (m == 0) ? term : num + term;
                    }
                    double diag = coef[0][q] + coef[1][q] + coef[2][q] + coef[3][q] +
                                  coef[4][q] + coef[5][q] + coef[6][q] + coef[7][q];
                    double gs = (prev[q] + c * num) / (1 + c * diag);
                    u[q] = u[q] + om * (gs - u[q]);
                }
}

/* segmentation4D.c:287-308 -- absolute sum r^2 h^4 over the interior */
double orc_residual(const orc_params *p, const double *u, const double *prev,
                    double *const coef[8])
{
    long off[8];
    nb_offsets(p, off);
    const double c = p->tau / (p->h * p->h);
    const double h4 = p->h * p->h * p->h * p->h;
    double *terms = TERMS(p) ? terms_buffer(interior_size(p)) : NULL;
    double acc = 0.0;
    #pragma omp parallel for num_threads(terms ? NT(p) : NR(p)) if (terms || NR(p) > 1) \
        reduction(+:acc) schedule(static)
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    const long q = orc_ijkl(p, i, j, k, l);
                    double num = 0;
                    for (int m = 0; m < 8; m++) {
                        double term = coef[NUM_ORDER[m]][q] * u[q + off[m]];
                        num = (m == 0) ? term : num + term;
                    }
                    double diag = coef[0][q] + coef[1][q] + coef[2][q] + coef[3][q] +
                                  coef[4][q] + coef[5][q] + coef[6][q] + coef[7][q];
                    double r = u[q] * (1 + c * diag) - c * num - prev[q];
                    if (terms) terms[interior_index(p, i, j, k, l)] = r * r * h4;
                    else acc += r * r * h4;
                }
    return terms ? serial_sum(terms, interior_size(p)) : acc;
}

/* segmentation4D.c:229-309 -- do { red; black; R } while (R > gsTol && z < cap) */
int orc_sor(const orc_params *p, double *u, const double *prev,
            double *const coef[8], double *res_out, double *hist)
{
    int z = 0;
    double R;
    do {
        z = z + 1;
        orc_sor_pass(p, u, prev, coef, 0);
        orc_sor_pass(p, u, prev, coef, 1);
        R = orc_residual(p, u, prev, coef);
        if (hist) hist[z - 1] = R;
    } while (R > p->gsTol && z < p->maxNoGSIteration);
    if (res_out) *res_out = R;
    return z;
}

/* segmentation4D.c:1020-1104, with the ball bounding boxes clamped to the
 * padded array (the reference indexes them unchecked; the clamp changes
 * nothing where the reference stays in bounds). */
void orc_rescale(const orc_params *p, double *u, const double *mask,
                 const orc_balls *balls)
{
    const double R = balls ? balls->radius : 0.0;
    for (int l = 1; l <= p->lMax - 2 && balls && balls->count; l++) {
        for (int c = 0; c < balls->count[l]; c++) {
            const double cx = balls->cx[balls->offset[l] + c];
            const double cy = balls->cy[balls->offset[l] + c];
            const double cz = balls->cz[balls->offset[l] + c];
            int k0 = (int)(cz - R - 0.5), k1 = (int)(cz + R + 0.5);
            int j0 = (int)(cy - R - 0.5), j1 = (int)(cy + R + 0.5);
            int i0 = (int)(cx - R - 0.5), i1 = (int)(cx + R + 0.5);
            if (k0 < 0) k0 = 0;
            if (j0 < 0) j0 = 0;
            if (i0 < 0) i0 = 0;
            if (k1 > p->kMax - 1) k1 = p->kMax - 1;
            if (j1 > p->jMax - 1) j1 = p->jMax - 1;
            if (i1 > p->iMax - 1) i1 = p->iMax - 1;
            double mx = 0, mn = 100000;
            for (int k = k0; k <= k1; k++)
                for (int j = j0; j <= j1; j++)
                    for (int i = i0; i <= i1; i++) {
                        double dx = i - cx, dy = j - cy, dz = k - cz;
                        double nd = sqrt((dx * dx) + (dy * dy) + (dz * dz));
                        if (nd <= R) {
                            double v = u[orc_ijkl(p, i, j, k, l)];
                            if (v < mn) mn = v;
                            if (v > mx) mx = v;
                        }
                    }
            double quot = 1. / (mx - mn);
            double offs = mn * quot;
            for (int k = k0; k <= k1; k++)
                for (int j = j0; j <= j1; j++)
                    for (int i = i0; i <= i1; i++) {
                        double dx = i - cx, dy = j - cy, dz = k - cz;
                        double nd = sqrt((dx * dx) + (dy * dy) + (dz * dz));
                        const long q = orc_ijkl(p, i, j, k, l);
                        double nv = (quot * u[q]) - offs;
                        if (nd <= R) {
                            if (c == 0) u[q] = nv;
                            else if (u[q] < nv) u[q] = nv;
                        }
                    }
        }
    }
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++)
                    if (mask[orc_ijkl(p, i, j, k, l)] == 0)
                        u[orc_ijkl(p, i, j, k, l)] = 0;
}

/* Ball bounding box of segmentation4D.c:907-912 etc., clamped to the array. */
static void ball_box(const orc_params *p, double cx, double cy, double cz, double R,
                     int box[6])
{
    box[0] = (int)(cx - R - 0.5); box[1] = (int)(cx + R + 0.5);
    box[2] = (int)(cy - R - 0.5); box[3] = (int)(cy + R + 0.5);
    box[4] = (int)(cz - R - 0.5); box[5] = (int)(cz + R + 0.5);
    const int hi[3] = {p->iMax - 1, p->jMax - 1, p->kMax - 1};
    for (int a = 0; a < 3; a++) {
        if (box[2 * a] < 0) box[2 * a] = 0;
        if (box[2 * a + 1] > hi[a]) box[2 * a + 1] = hi[a];
    }
}

/* segmentation4D.c:876-895 -- max over the interior, starting from 0 */
double orc_get4dmax(const orc_params *p, const double *img)
{
    double mx = 0;
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++)
                    if (img[orc_ijkl(p, i, j, k, l)] > mx) mx = img[orc_ijkl(p, i, j, k, l)];
    return mx;
}

/* segmentation4D.c:897-969 -- min (start 100000) or max (start 0) of frame
 * l inside the ball |x - c| <= R */
double orc_ball_minmax(const orc_params *p, const double *img, int l, double cx,
                       double cy, double cz, double R, int want_max)
{
    int box[6];
    ball_box(p, cx, cy, cz, R, box);
    double v = want_max ? 0 : 100000;
    for (int k = box[4]; k <= box[5]; k++)
        for (int j = box[2]; j <= box[3]; j++)
            for (int i = box[0]; i <= box[1]; i++) {
                double dx = i - cx, dy = j - cy, dz = k - cz;
                if (sqrt((dx * dx) + (dy * dy) + (dz * dz)) <= R) {
                    double x = img[orc_ijkl(p, i, j, k, l)];
                    if (want_max ? (x > v) : (x < v)) v = x;
                }
            }
    return v;
}

/* segmentation4D.c:971-1018 -- inside each ball, voxels brighter than
 * alpha*min + beta*max (of the ball) are set to the image's global max */
void orc_threshold(const orc_params *p, const double *in, double *thr,
                   const orc_balls *balls, double alpha, double beta)
{
    const double gmax = orc_get4dmax(p, in);
    const double R = balls->radius;
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int c = 0; c < balls->count[l]; c++) {
            const double cx = balls->cx[balls->offset[l] + c];
            const double cy = balls->cy[balls->offset[l] + c];
            const double cz = balls->cz[balls->offset[l] + c];
            const double mx = orc_ball_minmax(p, in, l, cx, cy, cz, R, 1);
            const double mn = orc_ball_minmax(p, in, l, cx, cy, cz, R, 0);
            const double tv = alpha * mn + beta * mx;
            int box[6];
            ball_box(p, cx, cy, cz, R, box);
            for (int k = box[4]; k <= box[5]; k++)
                for (int j = box[2]; j <= box[3]; j++)
                    for (int i = box[0]; i <= box[1]; i++) {
                        double dx = i - cx, dy = j - cy, dz = k - cz;
                        if (sqrt((dx * dx) + (dy * dy) + (dz * dz)) <= R &&
                            in[orc_ijkl(p, i, j, k, l)] > tv)
                            thr[orc_ijkl(p, i, j, k, l)] = gmax;
                    }
        }
}

/* segmentation4D.c:1106-1148 -- seed function max(u, |1/(d+v) - 1/(Rs+v)|)
 * inside d <= Rs = 12 (v = 0.5) around every centre, within the
 * ballRadius box; Mmask = 1 there */
void orc_generate(const orc_params *p, double *u, double *mask, const orc_balls *balls)
{
    const double v = 0.5, Rs = 12.;
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int c = 0; c < balls->count[l]; c++) {
            const double cx = balls->cx[balls->offset[l] + c];
            const double cy = balls->cy[balls->offset[l] + c];
            const double cz = balls->cz[balls->offset[l] + c];
            int box[6];
            ball_box(p, cx, cy, cz, balls->radius, box);
            for (int k = box[4]; k <= box[5]; k++)
                for (int j = box[2]; j <= box[3]; j++)
                    for (int i = box[0]; i <= box[1]; i++) {
                        double dx = i - cx, dy = j - cy, dz = k - cz;
                        double nd = sqrt((dx * dx) + (dy * dy) + (dz * dz));
                        if (nd <= Rs) {
                            double nv = fabs((1. / (sqrt((dx * dx) + (dy * dy) + (dz * dz)) + v)) -
                                             (1. / (Rs + v)));
                            const long q = orc_ijkl(p, i, j, k, l);
                            if (u[q] < nv) u[q] = nv;
                            mask[q] = 1.;
                        }
                    }
        }
}

/* segmentation4D.c:317-335 -- ghosts included, h^4 multiplied per term */
double orc_mass(const orc_params *p, const double *a)
{
    const double h4 = p->h * p->h * p->h * p->h;
    const size_t n = orc_size(p);
    double s = 0;
    for (size_t q = 0; q < n; q++) s += (a[q] * h4);
    return s;
}

/* segmentation4D.c:215-315 */
int orc_seg_time_step(const orc_params *p, double *u, double *prev,
                      double *const g[8], double *coef[8], const double *mask,
                      const orc_balls *balls, double *res_out)
{
    orc_copy(p, u, prev);
    orc_coefficients(p, u, g, coef);
    int z = orc_sor(p, u, prev, coef, res_out, NULL);
    orc_rescale(p, u, mask, balls);
    return z;
}

/* mcf_4D.c:344-443 -- exact-value Dirichlet ghosts, passes in the order
 * x-low, x-high, y-low, y-high, z-low, z-high, w-low, w-high. */
void orc_mcf_boundary(const orc_params *p, double *u, double t)
{
    const double h = p->h;
    const int I = p->iMax, J = p->jMax, K = p->kMax, L = p->lMax;
#define CC(n) (((n) - 1) * h + (h / 2.) - 1.25)
#define EX(x, y, z, w) ((((x) * (x) + (y) * (y) + (z) * (z) + (w) * (w) - 1.0) / 6.0) + t)
    for (int side = 0; side < 2; side++)
        for (int l = 0; l < L; l++)
            for (int k = 0; k < K; k++)
                for (int j = 0; j < J; j++) {
                    double x = side ? (h / 2.) + 1.25 : -(h / 2.) - 1.25;
                    double y = CC(j), z = CC(k), w = CC(l);
                    u[orc_ijkl(p, side ? I - 1 : 0, j, k, l)] = EX(x, y, z, w);
                }
    for (int side = 0; side < 2; side++)
        for (int l = 0; l < L; l++)
            for (int k = 0; k < K; k++)
                for (int i = 0; i < I; i++) {
                    double y = side ? (h / 2.) + 1.25 : -(h / 2.) - 1.25;
                    double x = CC(i), z = CC(k), w = CC(l);
                    u[orc_ijkl(p, i, side ? J - 1 : 0, k, l)] = EX(x, y, z, w);
                }
    for (int side = 0; side < 2; side++)
        for (int l = 0; l < L; l++)
            for (int j = 0; j < J; j++)
                for (int i = 0; i < I; i++) {
                    double z = side ? (h / 2.) + 1.25 : -(h / 2.) - 1.25;
                    double x = CC(i), y = CC(j), w = CC(l);
                    u[orc_ijkl(p, i, j, side ? K - 1 : 0, l)] = EX(x, y, z, w);
                }
    for (int side = 0; side < 2; side++)
        for (int k = 0; k < K; k++)
            for (int i = 0; i < I; i++)
                for (int j = 0; j < J; j++) {
                    double w = side ? (h / 2.) + 1.25 : -(h / 2.) - 1.25;
                    double x = CC(i), y = CC(j), z = CC(k);
                    u[orc_ijkl(p, i, j, k, side ? L - 1 : 0)] = EX(x, y, z, w);
                }
}

/* mcf_4D.c:444-462 */
void orc_mcf_exact(const orc_params *p, double *exact, double t)
{
    const double h = p->h;
    for (int l = 0; l < p->lMax; l++)
        for (int k = 0; k < p->kMax; k++)
            for (int j = 0; j < p->jMax; j++)
                for (int i = 0; i < p->iMax; i++) {
                    double x = CC(i), y = CC(j), z = CC(k), w = CC(l);
                    exact[orc_ijkl(p, i, j, k, l)] = EX(x, y, z, w);
                }
#undef CC
#undef EX
}

/* mcf_4D.c:102-205 -- copy, coefficients (g == 1), exact ghosts at
 * t^n = tau*n (set after the coefficients), then red-black SOR. */
int orc_mcf_time_step(const orc_params *p, double *u, double *prev,
                      double *coef[8], int step, double *res_out)
{
    double t = p->tau * step;
    orc_copy(p, u, prev);
    orc_coefficients(p, u, NULL, coef);
    orc_mcf_boundary(p, u, t);
    return orc_sor(p, u, prev, coef, res_out, NULL);
}

/* Mask: u >= 0.5 (max + min) over the interior (the reference quantises to
 * bytes at segmentation4D.c:1255 and thresholds at 0.5 in
 * segmentation4D_mpi.c:1474-1477). */
void orc_mask(const orc_params *p, const double *u, unsigned char *mask)
{
    double mx = -INFINITY, mn = INFINITY;
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    double v = u[orc_ijkl(p, i, j, k, l)];
                    if (v > mx) mx = v;
                    if (v < mn) mn = v;
                }
    double thr = 0.5 * (mx + mn);
    for (size_t q = 0; q < orc_size(p); q++) mask[q] = 0;
    for (int l = 1; l <= p->lMax - 2; l++)
        for (int k = 1; k <= p->kMax - 2; k++)
            for (int j = 1; j <= p->jMax - 2; j++)
                for (int i = 1; i <= p->iMax - 2; i++) {
                    const long q = orc_ijkl(p, i, j, k, l);
                    mask[q] = u[q] >= thr;
                }
}

double orc_iou(const orc_params *p, const unsigned char *a,
               const unsigned char *b)
{
    size_t inter = 0, uni = 0;
    for (size_t q = 0; q < orc_size(p); q++) {
        inter += (a[q] && b[q]);
        uni += (a[q] || b[q]);
    }
    return uni ? (double)inter / (double)uni : 1.0;
}

double orc_wtime(void) { return omp_get_wtime(); }
int orc_max_threads(void) { return omp_get_max_threads(); }
