/*
 * A C driver of the reference's shape (Stencil-OpenACC/AN_ISO_1_OpenACC.c:
 * 1150-1260) written against the drop-in header only: the host initializer,
 * device allocation and copies through the fold3d helpers, the six
 * reference-signature plane steps in the reference's sweep order, the copy
 * back and the reference's "%f" dump.  Used by tests/test_gpu_dropin.py.
 *
 *   driver N ITERATIONS DUMP_PATH
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "Fold3dPMM.h"
#include "fold3d.h"

typedef void (*plane_fn)(double *, int, double, double, double, double, double, double,
                         double, double, double, double);

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s N ITERATIONS DUMP_PATH\n", argv[0]);
    return 2;
  }
  const int n = atoi(argv[1]), iters = atoi(argv[2]);
  const double h = (_xmax - _xmin) / (n - 1.0); /* DX = dy = dz on the cube */
  const double hd = sqrt(h * h + h * h);      /* DXYP = DXZP = DYZP */
  const size_t count = (size_t)(n + 2) * (n + 2) * (n + 2);
  double *T = malloc(sizeof(double) * count);
  if (!T) return 1;
  ImplicitInitialiser(T, n, n, n, h, h, h);
  double *dT = NULL;
  if (fold3d_device_alloc((void **)&dT, sizeof(double) * count) ||
      fold3d_copy_to_device(dT, T, sizeof(double) * count)) {
    fprintf(stderr, "%s\n", fold3d_last_error());
    return 1;
  }
  /* the reference's per-direction (F, ay, ax, az) arguments and plane order */
  const struct {
    plane_fn f;
    int forward;
    double ay, ax, az;
  } dirs[6] = {
      {function_x_down, 1, _ayy, _azz, _axx},  {function_x_up, 0, -_ayy, -_azz, -_axx},
      {function_k_up, 1, _axx, _ayy, _azz},    {function_k_down, 0, -_axx, -_ayy, -_azz},
      {function_y_up, 0, -_azz, -_axx, -_ayy}, {function_y_down, 1, _azz, _axx, _ayy},
  };
  for (int it = 0; it < iters; ++it)
    for (int d = 0; d < 6; ++d) {
      if (dirs[d].forward)
        for (int p = 1; p < n; ++p)
          dirs[d].f(dT, p, h, h, h, hd, hd, hd, _Fc, dirs[d].ay, dirs[d].ax, dirs[d].az);
      else
        for (int p = n; p > 1; --p)
          dirs[d].f(dT, p, h, h, h, hd, hd, hd, _Fc, dirs[d].ay, dirs[d].ax, dirs[d].az);
    }
  if (fold3d_copy_to_host(T, dT, sizeof(double) * count) ||
      fold3d_write_text(argv[3], T, n, n, n)) {
    fprintf(stderr, "%s\n", fold3d_last_error());
    return 1;
  }
  fold3d_device_free(dT);
  free(T);
  printf("driver: %d^3, %d iterations, dump written to %s\n", n, iters, argv[3]);
  return 0;
}
