/*
 * host_init.c — the reference's host-side initializer, behind the drop-in
 * prototypes of include/Fold3dPMM.h (reference Stencil-OpenACC/Fold3dPMM.c).
 *
 * It writes the reference's input field: T = sign(IF)*MAXT everywhere, then
 * the closed-form fold traveltime around every point where the implicit
 * surface IF = 0 crosses a voxel edge (and blends of those points).  Host C,
 * not timed by the reference (AN_ISO_1_OpenACC.c:1179-1182); built without FP
 * contraction so the field is bit-identical to the reference's.
 *
 * Stride: the reference indexes with the compile-time HEIGHT = n+2
 * (Fold3dPMM.c:143); here it is the run-time (nx+2, ny+2) of the last
 * ImplicitInitialiser()/fold3d_set_grid() call.
 */
#include <math.h>
#include <stdlib.h>

#include "../../include/Fold3dPMM.h"

static int g_nx = _nx, g_ny = _ny, g_nz = _nz;

void fold3d_set_grid(int nx, int ny, int nz)
{
  g_nx = nx;
  g_ny = ny;
  g_nz = nz;
}

void fold3d_get_grid(int *nx, int *ny, int *nz)
{
  *nx = g_nx;
  *ny = g_ny;
  *nz = g_nz;
}

static inline long node_of(int i, int j, int k, int nx, int ny)
{
  return i + (long)(nx + 2) * (j + (long)(ny + 2) * k);
}

static inline double sgn(double v) { return (v > 0) ? 1.0 : -1.0; }

/* Fold3dPMM.c:9-26: traveltime from a point source at offset d, or MAXT when
 * the quadratic has no admissible root. */
double InitFoldFromPoint(double F, double ax, double ay, double az, double x,
                         double y, double z)
{
  const double r2 = x * x + y * y + z * z;
  const double ad = ax * x + ay * y + az * z;
  double q = (F * F - (ax * ax + ay * ay + az * az)) * r2 + ad * ad;
  if (q > 0.0) {
    q = ad + sqrt(q);
    if (q > 0.0) return r2 / q;
  }
  return 1.0 * MAXT;
}

/* Fold3dPMM.c:29-33 */
double ImpSurfF(double x, double y, double z)
{
  return 5.0 + (sin(x) + cos(y + x) + sin(x + y + z)) * 0.4 - z +
         (x - 5.0) * (y - 5.0) * (z - 5) / 55.0;
}

/* Fold3dPMM.c:35-54: (x,y,z) is the surface point relative to node (i,j,k)
 * of a grid with nx x ny interior nodes per plane. */
void fold3d_init_node(double x, double y, double z, int i, int j, int k, double *T, int nx,
                      int ny, int nz)
{
  (void)nz;  /* the layout needs only the two faster strides */
  double *t = T + node_of(i, j, k, nx, ny);
  const double s = (*t < 0.0) ? -1.0 : 1.0;
  const double tn = InitFoldFromPoint(_Fc, s * _axx, s * _ayy, s * _azz, -x, -y, -z);
  if (tn < fabs(*t)) *t = sgn(*t) * tn;
}

void InitNode(double x, double y, double z, int i, int j, int k, double *T)
{
  fold3d_init_node(x, y, z, i, j, k, T, g_nx, g_ny, g_nz);
}

/* Fold3dPMM.c:56-112: the 44 nodes around voxel (i,j,k).  Row = {node offset
 * di,dj,dk ; coordinate shift cx,cy,cz}: the node gets (xp - cx*dxx, ...),
 * where a shift of -1 is written xp + dxx and 2 is xp - 2*dxx.  Rows 18 and
 * 22 keep the reference's pairing of node (i+1,j,k) with the shift of
 * (i+1,j-1,k) (:82) and node (i+2,j,k) with the shift of (i+1,j+2,k) (:86). */
static const signed char kVoxel[44][6] = {
    {0, 0, 0, 0, 0, 0},     {1, 0, 0, 1, 0, 0},     {0, 1, 0, 0, 1, 0},
    {0, 0, 1, 0, 0, 1},     {1, 1, 0, 1, 1, 0},     {1, 0, 1, 1, 0, 1},
    {0, 1, 1, 0, 1, 1},     {1, 1, 1, 1, 1, 1},
    {-1, 0, 0, -1, 0, 0},   {-1, 1, 0, -1, 1, 0},   {-1, 0, 1, -1, 0, 1},
    {-1, 1, 1, -1, 1, 1},   {2, 0, 0, 2, 0, 0},     {2, 1, 0, 2, 1, 0},
    {2, 0, 1, 2, 0, 1},     {2, 1, 1, 2, 1, 1},
    {0, -1, 0, 0, -1, 0},   {1, 0, 0, 1, -1, 0},    {0, -1, 1, 0, -1, 1},
    {1, -1, 1, 1, -1, 1},   {0, 2, 0, 0, 2, 0},     {2, 0, 0, 1, 2, 0},
    {0, 2, 1, 0, 2, 1},     {1, 2, 1, 1, 2, 1},
    {0, 0, -1, 0, 0, -1},   {1, 0, -1, 1, 0, -1},   {0, 1, -1, 0, 1, -1},
    {1, 1, -1, 1, 1, -1},   {0, 0, 2, 0, 0, 2},     {1, 0, 2, 1, 0, 2},
    {0, 1, 2, 0, 1, 2},     {1, 1, 2, 1, 1, 2},
    {0, -1, -1, 0, -1, -1}, {-1, 0, -1, -1, 0, -1}, {-1, -1, 0, -1, -1, 0},
    {0, 2, 2, 0, 2, 2},     {2, 0, 2, 2, 0, 2},     {2, 2, 0, 2, 2, 0},
    {1, -1, -1, 1, -1, -1}, {-1, 1, -1, -1, 1, -1}, {-1, -1, 1, -1, -1, 1},
    {1, 2, 2, 1, 2, 2},     {2, 1, 2, 2, 1, 2},     {2, 2, 1, 2, 2, 1},
};

static inline double shifted(double p, int c, double h)
{
  return c == 0 ? p : c == 1 ? p - h : c == -1 ? p + h : p - 2 * h;
}

void fold3d_update_voxel(double *T, int i, int j, int k, double xp, double yp, double zp,
                         double dxx, double dyy, double dzz, int nx, int ny, int nz)
{
  for (int e = 0; e < 44; ++e) {
    const signed char *r = kVoxel[e];
    fold3d_init_node(shifted(xp, r[3], dxx), shifted(yp, r[4], dyy), shifted(zp, r[5], dzz),
                     i + r[0], j + r[1], k + r[2], T, nx, ny, nz);
  }
}

void updateVoxel(double *T, int i, int j, int k, double xp, double yp,
                 double zp, double dxx, double dyy, double dzz)
{
  fold3d_update_voxel(T, i, j, k, xp, yp, zp, dxx, dyy, dzz, g_nx, g_ny, g_nz);
}

double x(int i, double DX) { return (i - 1) * DX + _xmin; } /* Fold3dPMM.c:115-126 */
double y(int j, double dy) { return (j - 1) * dy + _ymin; }
double z(int k, double dz) { return (k - 1) * dz + _zmin; }

typedef struct { double x[12], y[12], z[12]; int n; } points_t;

static void add_point(points_t *P, double px, double py, double pz)
{
  P->x[P->n] = px;
  P->y[P->n] = py;
  P->z[P->n] = pz;
  P->n++;
}

/* Seed points of the voxel and their blends (Fold3dPMM.c:271-383).  For
 * point l and each earlier point m = l-1..l-5 (as far as they exist): the
 * thirds (2p_l+p_m)/3 and (p_l+2p_m)/3; for m = l-3 also four (l,l-2,l-3)
 * means, for m = l-4 three and for m = l-5 five three-point means.  In those
 * multi-point means the z coordinate adds the middle point's *y* coordinate,
 * exactly as the reference does. */
static void seed_voxel(double *T, int i, int j, int k, const points_t *P,
                       double dxx, double dyy, double dzz)
{
  const double *X = P->x, *Y = P->y, *Z = P->z;
  for (int l = 0; l < P->n; ++l) {
    updateVoxel(T, i, j, k, X[l], Y[l], Z[l], dxx, dyy, dzz);
    for (int back = 1; back <= 5 && back <= l; ++back) {
      const int m = l - back;
      updateVoxel(T, i, j, k, (2 * X[l] + X[m]) / 3, (2 * Y[l] + Y[m]) / 3,
                  (2 * Z[l] + Z[m]) / 3, dxx, dyy, dzz);
      updateVoxel(T, i, j, k, (X[l] + 2 * X[m]) / 3, (Y[l] + 2 * Y[m]) / 3,
                  (Z[l] + 2 * Z[m]) / 3, dxx, dyy, dzz);
      if (back == 3) {
        const int b = l - 2, c = l - 3;
        updateVoxel(T, i, j, k, (X[l] + X[b] + X[c]) / 3, (Y[l] + Y[b] + Y[c]) / 3,
                    (Z[l] + Y[b] + Z[c]) / 3, dxx, dyy, dzz);
        updateVoxel(T, i, j, k, (2 * X[l] + X[b] + X[c]) / 4,
                    (2 * Y[l] + Y[b] + Y[c]) / 4, (2 * Z[l] + Y[b] + Z[c]) / 4,
                    dxx, dyy, dzz);
        updateVoxel(T, i, j, k, (X[l] + 2 * X[b] + X[c]) / 4,
                    (Y[l] + 2 * Y[b] + Y[c]) / 4, (Z[l] + 2 * Y[b] + Z[c]) / 4,
                    dxx, dyy, dzz);
        updateVoxel(T, i, j, k, (X[l] + X[b] + 2 * X[c]) / 4,
                    (Y[l] + Y[b] + 2 * Y[c]) / 4, (Z[l] + Y[b] + 2 * Z[c]) / 4,
                    dxx, dyy, dzz);
      } else if (back >= 4) {
        /* (middle, last) offsets behind l */
        static const int pairs4[3][2] = {{2, 3}, {2, 4}, {3, 4}};
        static const int pairs5[5][2] = {{2, 3}, {2, 4}, {2, 5}, {3, 5}, {3, 4}};
        const int(*pr)[2] = back == 4 ? pairs4 : pairs5;
        const int np = back == 4 ? 3 : 5;
        for (int t = 0; t < np; ++t) {
          const int b = l - pr[t][0], c = l - pr[t][1];
          updateVoxel(T, i, j, k, (X[l] + X[b] + X[c]) / 3, (Y[l] + Y[b] + Y[c]) / 3,
                      (Z[l] + Y[b] + Z[c]) / 3, dxx, dyy, dzz);
        }
      }
    }
  }
}

void ImplicitInitialiser(double *T, int nx, int ny, int nz, double dxx,
                         double dyy, double dzz)
{
  fold3d_set_grid(nx, ny, nz);
  const long count = (long)(nx + 2) * (ny + 2) * (nz + 2);
  double *IF = (double *)malloc(sizeof(double) * count);
  if (!IF) return; /* the reference would crash here; leave T untouched */

  for (int k = 0; k < nz + 2; k++)
    for (int j = 0; j < ny + 2; j++)
      for (int i = 0; i < nx + 2; i++) {
        const long c = node_of(i, j, k, nx, ny);
        IF[c] = ImpSurfF(x(i, dxx), y(j, dyy), z(k, dzz));
        T[c] = sgn(IF[c]) * MAXT;
      }

  for (int i = 1; i < nx; i++)
    for (int j = 1; j < ny; j++)
      for (int k = 1; k < nz; k++) {
        /* corner values c[di + 2*dj + 4*dk] */
        double c[8];
        for (int q = 0; q < 8; ++q) c[q] = IF[node_of(i + (q & 1), j + (q >> 1 & 1), k + (q >> 2), nx, ny)];
        points_t P;
        P.n = 0;
        double t;
        /* x-edges at (y,z) = (0,0), (0,dzz), (dyy,0), (dyy,dzz) (Fold3dPMM.c:173-204) */
        static const int xe[4] = {0, 4, 2, 6};
        for (int e = 0; e < 4; ++e) {
          const double f0 = c[xe[e]], f1 = c[xe[e] + 1];
          t = dxx * f0 / (f0 - f1);
          if (0.0 <= t && t <= dxx) add_point(&P, t, (xe[e] & 2) ? dyy : 0.0, (xe[e] & 4) ? dzz : 0.0);
        }
        /* y-edges at (x,z) = (0,0), (0,dzz), (dxx,0), (dxx,dzz) (:205-236) */
        static const int ye[4] = {0, 4, 1, 5};
        for (int e = 0; e < 4; ++e) {
          const double f0 = c[ye[e]], f1 = c[ye[e] + 2];
          t = dyy * f0 / (f0 - f1);
          if (0.0 <= t && t <= dyy) add_point(&P, (ye[e] & 1) ? dxx : 0., t, (ye[e] & 4) ? dzz : 0.0);
        }
        /* z-edges at (x,y) = (0,0), (dxx,dyy), (dxx,0), (0,dyy) (:238-269) */
        static const int ze[4] = {0, 3, 1, 2};
        for (int e = 0; e < 4; ++e) {
          const double f0 = c[ze[e]], f1 = c[ze[e] + 4];
          t = dzz * f0 / (f0 - f1);
          if (0.0 <= t && t <= dzz) add_point(&P, (ze[e] & 1) ? dxx : 0.0, (ze[e] & 2) ? dyy : 0.0, t);
        }
        seed_voxel(T, i, j, k, &P, dxx, dyy, dzz);
      }
  free(IF);
}
