/*
 * seg4d_io.c -- the segmentation program's file formats (host code): the
 * uint8 VTK STRUCTURED_POINTS frames it reads and writes and the ASCII VTK
 * centre files it parses (segmentation4D.c:1149-1324), plus the binary mask
 * writer of the MPI program (segmentation4D_mpi.c:1429-1480).
 *
 * Same names, file names, header text and value mapping as the reference:
 *   read:  value = byte / 255.
 *   write: byte  = (unsigned char)(255 * value + 0.5), i.e. truncation
 *          toward zero to int, then the low 8 bits (values outside [0, 1]
 *          wrap, as the reference's x86-64 build does)
 *   mask:  byte  = value >= 0.5 ? 1 : 0
 * Differences, all on malformed input: reads stop at end of file instead of
 * spinning on EOF, number tokens are bounded to 199 characters, a centre
 * file with more points than no_of_centers[l] or an unreadable coordinate is
 * an error (the reference writes past its arrays), and _writeData_ returns
 * false when a file cannot be created (the reference writes to NULL).
 * Fatal errors of _readData_ exit(1) like the reference, with a message.
 *
 * Coherence: these functions read and write the host arrays; in resident
 * mode call seg4dUpload after _readData_ and seg4dDownload before a writer.
 */
#include <ctype.h>
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "seg4d.h"

#define API __attribute__((visibility("default")))

/* segmentation4D.c:60 */
API FILE *input_file = NULL, *input_center_file = NULL, *output_file = NULL;

static long frame_points(void) { return (long)(iMax - 2) * (jMax - 2) * (kMax - 2); }

static long idx(int i, int j, int k, int l)
{
    return (((long)l * kMax + k) * jMax + j) * iMax + i;
}

/* segmentation4D.c:1255-1266: characters up to and including the first
 * white-space character. */
API void _readNumberFromFile_(char *dim, FILE *f)
{
    int count = 0, c;
    do {
        c = fgetc(f);
        if (c == EOF) break;
        dim[count++] = (char)c;
    } while (!isspace(c) && count < 199);
    dim[count] = '\0';
}

/* Skips to just past the next newline; 0 at end of file. */
static int skip_line(FILE *f)
{
    int c;
    while ((c = getc(f)) != '\n')
        if (c == EOF) return 0;
    return 1;
}

/* segmentation4D.c:1268-1280: comment lines (first character '#'), then the
 * rest of the line whose first character ended the comment loop, then two
 * more lines (title, ASCII, DATASET in a standard legacy file). */
API void _readMetaData_(FILE *f)
{
    int c;
    while ((c = getc(f)) == '#')
        if (!skip_line(f)) return;
    if (c == EOF) return;
    if (!skip_line(f)) return;  /* even when c was that line's newline */
    if (!skip_line(f)) return;
    skip_line(f);
}

/* segmentation4D.c:1282-1313: ten header lines (fgets of 255 characters
 * each), then (kMax-2)(jMax-2)(iMax-2) bytes, x fastest, into frame l's
 * interior.  Bytes missing at the end of the file read as EOF / 255.,
 * like the reference's getc. */
API bool _load3dDataArrayVTK_(double *img, int l, unsigned char *path)
{
    if (img == NULL || path == NULL) return false;
    FILE *f = fopen((const char *)path, "rb");
    if (f == NULL) return false;
    char line[256];
    for (int q = 0; q < 10; q++)
        if (fgets(line, 256, f) == NULL) break;
    const long n = frame_points();
    unsigned char *buf = (unsigned char *)malloc(n > 0 ? (size_t)n : 1);
    if (buf == NULL) {
        fclose(f);
        return false;
    }
    const long got = (long)fread(buf, 1, (size_t)n, f);
    fclose(f);
    long q = 0;
    for (int k = 1; k <= kMax - 2; k++)
        for (int j = 1; j <= jMax - 2; j++)
            for (int i = 1; i <= iMax - 2; i++, q++)
                img[idx(i, j, k, l)] = (q < got ? (int)buf[q] : EOF) / 255.;
    free(buf);
    return true;
}

static void die(const char *what, const char *path)
{
    fprintf(stderr, "seg4d: _readData_: %s: %s\n", what, path);
    exit(1);
}

/* segmentation4D.c:1149-1216: frames l = 1..lMax-2 from
 * <dir>/movingBalls_457_at_time_%03d.vtk and their ball centres from
 * <dir>/movingBalls_457_centers_at_timestep_%03d.vtk, each coordinate
 * mapped to (double)(int)(x / spacing + 0.5). */
API void _readData_(double *inputImage, unsigned char *fileDirectory)
{
    char path[4096], temp[16], dim[200];
    for (int l = 1; l <= lMax - 2; l++) {
        snprintf(path, sizeof path, "%s/movingBalls_457_at_time_%03d.vtk",
                 (const char *)fileDirectory, l);
        if (!_load3dDataArrayVTK_(inputImage, l, (unsigned char *)path))
            die("cannot read image frame", path);

        snprintf(path, sizeof path, "%s/movingBalls_457_centers_at_timestep_%03d.vtk",
                 (const char *)fileDirectory, l);
        input_center_file = fopen(path, "r");
        if (input_center_file == NULL) die("cannot open centre file", path);
        _readMetaData_(input_center_file);
        if (fgets(temp, 8, input_center_file) == NULL) die("no POINTS line", path);
        _readNumberFromFile_(dim, input_center_file);
        const int points = atoi(dim);
        /* rest of the POINTS line (the next line if the count ended it) */
        if (!skip_line(input_center_file) && points > 0) die("truncated POINTS line", path);
        const int cap = (no_of_centers != NULL) ? no_of_centers[l] : 0;
        if (points > cap) die("more points than no_of_centers[l]", path);
        for (int c = 0; c < points; c++) {
            double x, y, z;
            if (fscanf(input_center_file, "%lf", &x) != 1 ||
                fscanf(input_center_file, "%lf", &y) != 1 ||
                fscanf(input_center_file, "%lf", &z) != 1)
                die("unreadable centre coordinate", path);
            center_x[l][c] = (double)(int)(x / spacing_x + 0.5);
            center_y[l][c] = (double)(int)(y / spacing_y + 0.5);
            center_z[l][c] = (double)(int)(z / spacing_z + 0.5);
        }
        fclose(input_center_file);
        input_center_file = NULL;
    }
}

/* The header of segmentation4D.c:1236-1245 (and segmentation4D_mpi.c:
 * 1451-1460). */
static void vtk_header(FILE *f)
{
    fprintf(f, "# vtk DataFile Version 3.0\n");
    fprintf(f, "file in binary format\n");
    fprintf(f, "BINARY\n");
    fprintf(f, "DATASET STRUCTURED_POINTS\n");
    fprintf(f, "DIMENSIONS %d %d %d\n", iMax - 2, jMax - 2, kMax - 2);
    fprintf(f, "ORIGIN %f %f %f\n", origin_x, origin_y, origin_z);
    fprintf(f, "SPACING %f %f %f\n", spacing_x, spacing_y, spacing_z);
    fprintf(f, "POINT_DATA %ld\n", frame_points());
    fprintf(f, "SCALARS scalars unsigned_char\n");
    fprintf(f, "LOOKUP_TABLE default\n");
}

/* (unsigned char)(255 v + 0.5) as the reference's x86-64 build evaluates it:
 * truncation to a 32-bit int (out of range and NaN give INT_MIN), then the
 * low byte. */
static unsigned char to_byte(double v)
{
    const double x = 255 * v + 0.5;
    if (!(x > (double)INT_MIN - 1.0 && x < (double)INT_MAX + 1.0)) return 0;
    return (unsigned char)((unsigned)(int)x & 0xffu);
}

static bool write_frames(const double *a, const char *path, int mask)
{
    if (a == NULL || path == NULL) return false;
    const long n = frame_points();
    unsigned char *buf = (unsigned char *)malloc(n > 0 ? (size_t)n : 1);
    if (buf == NULL) return false;
    char name[4096];
    bool ok = true;
    for (int l = 1; l <= lMax - 2; l++) {
        snprintf(name, sizeof name, "%s_at_time_%03d.vtk", path, l);
        output_file = fopen(name, "w");
        if (output_file == NULL) {
            puts("File could not be opened.");
            ok = false;
            continue;
        }
        vtk_header(output_file);
        long q = 0;
        for (int k = 1; k <= kMax - 2; k++)
            for (int j = 1; j <= jMax - 2; j++)
                for (int i = 1; i <= iMax - 2; i++, q++) {
                    const double v = a[idx(i, j, k, l)];
                    buf[q] = mask ? (unsigned char)(v >= 0.5 ? 1 : 0) : to_byte(v);
                }
        if (fwrite(buf, 1, (size_t)n, output_file) != (size_t)n) ok = false;
        if (fclose(output_file) != 0) ok = false;
        output_file = NULL;
    }
    free(buf);
    return ok;
}

/* segmentation4D.c:1218-1253: <path>_at_time_%03d.vtk for l = 1..lMax-2. */
API bool _writeData_(double *array4DPtr, unsigned char *pathPtr)
{
    return write_frames(array4DPtr, (const char *)pathPtr, 0);
}

/* segmentation4D_mpi.c:1429-1480: the same files with the segmentation mask
 * (1 where value >= 0.5, else 0) as the data bytes. */
API bool seg4dWriteMaskVTK(const double *array4DPtr, const char *pathPtr)
{
    return write_frames(array4DPtr, pathPtr, 1);
}
