/*
 * host_driver.c — runs the library's host C code (the reference initializer
 * restated in csrc/host_init.c, the file formats in csrc/io.c) on a few grids,
 * for a build under AddressSanitizer and UndefinedBehaviorSanitizer.  A
 * finding aborts the process with the sanitizer's report; the checks below
 * only confirm the code ran.
 *
 * usage: host_driver <scratch directory>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/Fold3dPMM.h"
#include "../../include/fold3d.h"

/* io.c reports failures through this hook (libfold3d.so's error text) */
static char g_msg[512];
void fold3d_set_error_text(const char *msg) { snprintf(g_msg, sizeof g_msg, "%s", msg); }

static int run(const char *dir, int nx, int ny, int nz) {
  const size_t count = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  /* R also holds the (nx+1) x ny x nz grid of the mismatched read below */
  const size_t count1 = (size_t)(nx + 3) * (ny + 2) * (nz + 2);
  double *T = malloc(count * sizeof *T), *R = malloc(count1 * sizeof *R);
  if (!T || !R) return 1;
  const double dx = 10.0 / (nx > 1 ? nx - 1 : 1), dy = 10.0 / (ny > 1 ? ny - 1 : 1),
               dz = 10.0 / (nz > 1 ? nz - 1 : 1);
  ImplicitInitialiser(T, nx, ny, nz, dx, dy, dz);
  size_t reached = 0;
  for (size_t q = 0; q < count; ++q) reached += fabs(T[q]) < 1e15;
  char path[1024];
  snprintf(path, sizeof path, "%s/dump_%d_%d_%d.txt", dir, nx, ny, nz);
  int rc = fold3d_write_text(path, T, nx, ny, nz);
  if (rc == FOLD3D_OK) rc = fold3d_read_text(path, R, nx, ny, nz);
  for (size_t q = 0; rc == FOLD3D_OK && q < count; ++q)
    if (fabs(R[q] - T[q]) > 5e-7 * (1.0 + fabs(T[q]))) rc = -100;
  /* a grid that does not match the file is an error code, not an overrun */
  if (rc == FOLD3D_OK && fold3d_read_text(path, R, nx + 1, ny, nz) != FOLD3D_E_IO) rc = -101;
  snprintf(path, sizeof path, "%s/field_%d_%d_%d.vtk", dir, nx, ny, nz);
  if (rc == FOLD3D_OK) rc = fold3d_write_vtk(path, T, nx, ny, nz, dx, dy, dz, 1);
  if (rc == FOLD3D_OK) rc = fold3d_write_vtk(path, T, nx, ny, nz, dx, dy, dz, 0);
  printf("grid %d x %d x %d: %zu reached nodes, rc %d\n", nx, ny, nz, reached, rc);
  free(T);
  free(R);
  return rc != FOLD3D_OK;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  static const int grids[][3] = {{16, 16, 16}, {7, 9, 11}, {3, 3, 3}, {30, 5, 12}, {1, 4, 2}};
  int bad = 0;
  for (size_t g = 0; g < sizeof grids / sizeof grids[0]; ++g)
    bad += run(argv[1], grids[g][0], grids[g][1], grids[g][2]);
  /* argument and file errors */
  bad += fold3d_write_text(NULL, NULL, 4, 4, 4) != FOLD3D_E_INVALID;
  bad += fold3d_read_text("/nonexistent/dir/file.txt", (double[8]){0}, 0, 0, 0) !=
         FOLD3D_E_INVALID;
  bad += fold3d_write_vtk("/nonexistent/dir/f.vtk", (double[1000]){0}, 8, 8, 8, 1, 1, 1, 1) !=
         FOLD3D_E_IO;
  printf("%s\n", bad ? "FAILED" : "ok");
  return bad != 0;
}
