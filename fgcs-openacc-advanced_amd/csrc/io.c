/*
 * io.c — host-side readers and writers of T fields.
 *
 *   fold3d_write_text / fold3d_read_text: the reference's solution dump
 *     (Stencil-OpenACC/AN_ISO_1_OpenACC.c:1244-1260): one "%f\n" line per
 *     node of the padded grid, ghosts included, with i outermost and k
 *     innermost, i.e. T[i + (nx+2)*(j + (ny+2)*k)] for i, then j, then k.
 *     Outputs of reference builds elsewhere can be compared with ours line
 *     for line.
 *   fold3d_write_vtk: a legacy-format VTK file of the interior nodes
 *     (STRUCTURED_POINTS, origin at node (1,1,1) = (0,0,0) as the reference's
 *     x(i) = (i-1)*DX, spacing dx, dy, dz), ASCII or big-endian binary.  The
 *     reference's Write_VTK_Structured_Grid (:1098-1135) writes headers only.
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fold3d.h"

void fold3d_set_error_text(const char *msg);

static int io_fail(int code, const char *what, const char *path) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s: %s (%s)", what, path ? path : "(null)",
           errno ? strerror(errno) : "bad argument");
  fold3d_set_error_text(buf);
  return code;
}

int fold3d_write_text(const char *path, const double *T, int nx, int ny, int nz) {
  errno = 0;
  if (!path || !T || nx < 1 || ny < 1 || nz < 1)
    return io_fail(FOLD3D_E_INVALID, "fold3d_write_text: bad argument", path);
  FILE *f = fopen(path, "w");
  if (!f) return io_fail(FOLD3D_E_IO, "fold3d_write_text: cannot open", path);
  const long long H0 = nx + 2, H1 = ny + 2, H2 = nz + 2;
  /* a 1 MiB buffer that stdio allocates for this stream alone (writers on
     several threads do not share it) and frees at fclose */
  setvbuf(f, NULL, _IOFBF, 1 << 20);
  for (long long i = 0; i < H0; ++i)
    for (long long j = 0; j < H1; ++j)
      for (long long k = 0; k < H2; ++k)
        if (fprintf(f, "%f\n", T[i + H0 * (j + H1 * k)]) < 0) {
          fclose(f);
          return io_fail(FOLD3D_E_IO, "fold3d_write_text: write failed", path);
        }
  if (fclose(f) != 0) return io_fail(FOLD3D_E_IO, "fold3d_write_text: close failed", path);
  return FOLD3D_OK;
}

int fold3d_read_text(const char *path, double *T, int nx, int ny, int nz) {
  errno = 0;
  if (!path || !T || nx < 1 || ny < 1 || nz < 1)
    return io_fail(FOLD3D_E_INVALID, "fold3d_read_text: bad argument", path);
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(FOLD3D_E_IO, "fold3d_read_text: cannot open", path);
  const long long H0 = nx + 2, H1 = ny + 2, H2 = nz + 2;
  for (long long i = 0; i < H0; ++i)
    for (long long j = 0; j < H1; ++j)
      for (long long k = 0; k < H2; ++k) {
        double v;
        if (fscanf(f, "%lf", &v) != 1) {
          fclose(f);
          errno = 0;
          return io_fail(FOLD3D_E_IO, "fold3d_read_text: too few values", path);
        }
        T[i + H0 * (j + H1 * k)] = v;
      }
  double extra;
  const int more = fscanf(f, "%lf", &extra);
  fclose(f);
  if (more == 1) {
    errno = 0;
    return io_fail(FOLD3D_E_IO, "fold3d_read_text: more values than the grid holds", path);
  }
  return FOLD3D_OK;
}

static void put_be64(unsigned char *p, double v) {
  uint64_t u;
  memcpy(&u, &v, sizeof u);
  for (int b = 0; b < 8; ++b) p[b] = (unsigned char)(u >> (56 - 8 * b));
}

int fold3d_write_vtk(const char *path, const double *T, int nx, int ny, int nz, double dx,
                     double dy, double dz, int binary) {
  errno = 0;
  if (!path || !T || nx < 1 || ny < 1 || nz < 1 || !(dx > 0 && dy > 0 && dz > 0))
    return io_fail(FOLD3D_E_INVALID, "fold3d_write_vtk: bad argument", path);
  FILE *f = fopen(path, binary ? "wb" : "w");
  if (!f) return io_fail(FOLD3D_E_IO, "fold3d_write_vtk: cannot open", path);
  const long long H0 = nx + 2, H1 = ny + 2;
  const long long npts = (long long)nx * ny * nz;
  fprintf(f, "# vtk DataFile Version 3.0\n");
  fprintf(f, "fold3d traveltime field\n");
  fprintf(f, "%s\n", binary ? "BINARY" : "ASCII");
  fprintf(f, "DATASET STRUCTURED_POINTS\n");
  fprintf(f, "DIMENSIONS %d %d %d\n", nx, ny, nz);
  fprintf(f, "ORIGIN 0 0 0\n");
  fprintf(f, "SPACING %.17g %.17g %.17g\n", dx, dy, dz);
  fprintf(f, "POINT_DATA %lld\n", npts);
  fprintf(f, "SCALARS T double 1\n");
  fprintf(f, "LOOKUP_TABLE default\n");
  int bad = 0;
  if (binary) {
    unsigned char *row = (unsigned char *)malloc((size_t)nx * 8);
    if (!row) {
      fclose(f);
      return io_fail(FOLD3D_E_NOMEM, "fold3d_write_vtk: out of host memory", path);
    }
    for (long long k = 1; k <= nz && !bad; ++k)
      for (long long j = 1; j <= ny && !bad; ++j) {
        for (long long i = 1; i <= nx; ++i) put_be64(row + 8 * (i - 1), T[i + H0 * (j + H1 * k)]);
        bad = fwrite(row, 8, (size_t)nx, f) != (size_t)nx;
      }
    free(row);
    if (!bad) bad = fputc('\n', f) == EOF;
  } else {
    for (long long k = 1; k <= nz && !bad; ++k)
      for (long long j = 1; j <= ny && !bad; ++j)
        for (long long i = 1; i <= nx && !bad; ++i)
          bad = fprintf(f, "%.17g\n", T[i + H0 * (j + H1 * k)]) < 0;
  }
  if (fclose(f) != 0 || bad) return io_fail(FOLD3D_E_IO, "fold3d_write_vtk: write failed", path);
  return FOLD3D_OK;
}
