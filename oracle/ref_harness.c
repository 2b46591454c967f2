/*
 * Thin C entry points over the reference's own single-device solver, so the
 * checker library oracle/_ref/libfoldref_<N>.so can be driven from Python
 * (ctypes) to write golden vectors and to cross-check the oracle.
 *
 * Linked against objects compiled from the unmodified reference sources
 *   Stencil-OpenACC/AN_ISO_1_OpenACC.c  (main renamed to fold_ref_main)
 *   Stencil-OpenACC/Fold3dPMM.c
 * with the grid size set by ref_grid.h.  Test infrastructure only.
 */
#include <math.h>
#include <stdlib.h>

/* Reference symbols (signatures from AN_ISO_1_OpenACC.c:59-62,120-124,180-183
 * and Fold3dPMM.h:58-86). */
void FoldEdge2Points(double *tnew, const double st, const double xt,
                     const double F, const double ax, const double ay,
                     const double az, const double dxx, const double dzz);
void FoldSurf3Points(double *tnew, const double st, const double xt,
                     const double yt, const double F, const double ax,
                     const double ay, const double az, const double dxx,
                     const double dyy, const double dzz, const double dxz);
typedef void plane_fn(double *T, int plane, double dzz, double dxx, double dyy,
                      double dxy, double dxz, double dyz, double F, double ay,
                      double ax, double az);
plane_fn function_x_down, function_x_up, function_y_up, function_y_down,
    function_k_up, function_k_down;
int fold_ref_main(void);

int ref_n(void) { return REF_N; }
int ref_height(void) { return HEIGHT; }

double ref_edge(double tnew, double st, double xt, double F, double ax,
                double ay, double az, double dxx, double dzz)
{
  FoldEdge2Points(&tnew, st, xt, F, ax, ay, az, dxx, dzz);
  return tnew;
}

double ref_surf(double tnew, double st, double xt, double yt, double F,
                double ax, double ay, double az, double dxx, double dyy,
                double dzz, double dxz)
{
  FoldSurf3Points(&tnew, st, xt, yt, F, ax, ay, az, dxx, dyy, dzz, dxz);
  return tnew;
}

double ref_init_fold(double F, double ax, double ay, double az, double x,
                     double y, double z)
{
  return InitFoldFromPoint(F, ax, ay, az, x, y, z);
}

double ref_impsurf(double x, double y, double z) { return ImpSurfF(x, y, z); }

/* Grid spacings exactly as the reference driver computes them (AN_ISO_1_OpenACC.c:1154-1161). */
static void spacings(double *DX, double *dy, double *dz, double *DXYP,
                     double *DXZP, double *DYZP)
{
  *DX = (_xmax - _xmin) / (_nx - 1.0);
  *dy = (_ymax - _ymin) / (_ny - 1.0);
  *dz = (_zmax - _zmin) / (_nz - 1.0);
  *DXYP = sqrt(*DX * *DX + *dy * *dy);
  *DXZP = sqrt(*DX * *DX + *dz * *dz);
  *DYZP = sqrt(*dy * *dy + *dz * *dz);
}

void ref_initialiser(double *T)
{
  double DX, dy, dz, a, b, c;
  spacings(&DX, &dy, &dz, &a, &b, &c);
  ImplicitInitialiser(T, _nx, _ny, _nz, DX, dy, dz);
}

/* One plane step with the arguments the reference driver passes
 * (AN_ISO_1_OpenACC.c:1198,1205,1212,1219,1227,1234).
 * dir: 0 x_down, 1 x_up, 2 k_up, 3 k_down, 4 y_up, 5 y_down. */
void ref_plane(int dir, double *T, int plane)
{
  double DX, dy, dz, DXYP, DXZP, DYZP;
  const double F = _Fc;
  spacings(&DX, &dy, &dz, &DXYP, &DXZP, &DYZP);
  switch (dir) {
  case 0: function_x_down(T, plane, DX, dz, dy, DYZP, DXZP, DXYP, F, _ayy, _azz, _axx); break;
  case 1: function_x_up(T, plane, DX, dz, dy, DYZP, DXZP, DXYP, F, -_ayy, -_azz, -_axx); break;
  case 2: function_k_up(T, plane, dz, dy, DX, DXYP, DYZP, DXZP, F, _axx, _ayy, _azz); break;
  case 3: function_k_down(T, plane, dz, dy, DX, DXYP, DYZP, DXZP, F, -_axx, -_ayy, -_azz); break;
  case 4: function_y_up(T, plane, dy, DX, dz, DXZP, DXYP, DYZP, F, -_azz, -_axx, -_ayy); break;
  case 5: function_y_down(T, plane, dy, DX, dz, DXZP, DXYP, DYZP, F, _azz, _axx, _ayy); break;
  default: break;
  }
}

/* The driver's plane order for one sweep direction (AN_ISO_1_OpenACC.c:1195-1236). */
void ref_direction(int dir, double *T)
{
  int p;
  switch (dir) {
  case 0: for (p = 1; p < _nx; p++) ref_plane(0, T, p); break;
  case 1: for (p = _nx; p > 1; --p) ref_plane(1, T, p); break;
  case 2: for (p = 1; p < _nz; p++) ref_plane(2, T, p); break;
  case 3: for (p = _nz; p > 1; p--) ref_plane(3, T, p); break;
  case 4: for (p = _ny; p > 1; p--) ref_plane(4, T, p); break;
  case 5: for (p = 1; p < _ny; p++) ref_plane(5, T, p); break;
  default: break;
  }
}

void ref_sweeps(double *T, int iterations)
{
  int it, d;
  for (it = 0; it < iterations; it++)
    for (d = 0; d < 6; d++) ref_direction(d, T);
}

/* The reference's unmodified main(): writes an_iso_1_openacc.txt in the
 * current directory. */
int ref_main(void) { return fold_ref_main(); }
