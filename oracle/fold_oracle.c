/*
 * ORACLE — CPU restatement of the reference's fold-model PMM solve (FP64).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg as the checker; never linked into the shipped library.
 *
 * Every expression keeps the reference's operation order (no contraction:
 * built with -ffp-contract=off), because the HIP path is held to bit parity
 * with this file.  Citations are to the reference tree:
 *   AN_ISO_1_OpenACC.c  local solvers :59-175, plane kernels :180-1092, driver :1150-1241
 *   Fold3dPMM.c         initializer :9-389,  Fold3dPMM.h constants :6-52
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXT (10e15 * 1.0) /* Fold3dPMM.h:6 */

/* The reference's macro semantics: min returns the second operand on ties
 * and NaN (AN_ISO_1_OpenACC.c:32-34); sign(0) is -1 (Fold3dPMM.h:32-34). */
static inline double rmin(double a, double b) { return (a < b) ? a : b; }
static inline double rsign(double a) { return (a > 0) ? 1.0 : -1.0; }

static int g_threads = 0;
void orc_set_threads(int nthreads) { g_threads = nthreads; }

/* Edge candidate + one-point candidate (AN_ISO_1_OpenACC.c:59-114).
 * Source edge st=(0,0)->xt=(dxx,0) in the upwind plane, target (0,0,dzz). */
static void edge2(double *tnew, double st, double xt, double F, double ax,
                  double ay, double az, double dxx, double dzz)
{
  if (*tnew < rmin(xt, st) || (ax > 0.0 && xt > st)) return; /* :71 */

  double ga = F * F - ay * ay;
  double dtx = (xt - st) / dxx;
  double c = 1.0 - dtx * ax;
  double disc = ga * (c * c - dtx * dtx * (ga - az * az));
  if (0.0 < disc) {
    double pz = (-az * c + sqrt(disc)) / (ga - az * az);
    double cand = st + dzz * pz;
    if (cand < *tnew && cand > rmin(st, xt)) {
      /* foot point of the characteristic; :92 is the one that counts */
      double g = sqrt((dtx * dtx + pz * pz) * F * F / ga);
      double xe = -dzz * (dtx + ax * g) / (pz + az * g);
      if (0.0 < xe && xe < dxx) *tnew = cand;
    }
  }
  if (*tnew > xt) { /* :99-113, straight line from xt */
    double c1 = (-dxx * ax + dzz * az);
    double len2 = dxx * dxx + dzz * dzz;
    double d1 = (F * F - ax * ax - ay * ay - az * az) * len2 + c1 * c1;
    if (d1 > 0.0) {
      c1 = c1 + sqrt(d1);
      if (c1 > 0.0) *tnew = rmin(*tnew, xt + len2 / c1);
    }
  }
}

/* Triangle candidate over st=(0,0), xt=(dxx,0), yt=(dxx,dyy) plus the edge
 * xt->yt in a frame rotated in the x-z plane (AN_ISO_1_OpenACC.c:119-175;
 * UNROLLED defined, Fold3dPMM.h:11, so :137-142 is inactive). */
static void surf3(double *tnew, double st, double xt, double yt, double F,
                  double ax, double ay, double az, double dxx, double dyy,
                  double dzz, double dxz)
{
  if (*tnew < rmin(st, rmin(xt, yt)) || (ax > 0.0 && xt > st) ||
      (ay > 0.0 && yt > xt))
    return; /* :132 */

  double dtx = (xt - st) / dxx;
  double dty = (yt - xt) / dyy;
  double p2 = dtx * dtx + dty * dty;
  double ga = F * F - az * az;
  double al = 1.0 - ax * dtx - ay * dty;
  double disc = al * al - p2 * ga;
  if (disc > 0.0) {
    double pz = (-az * al + F * sqrt(disc)) / ga;
    double cand = st + dzz * pz;
    if (cand < *tnew && cand > rmin(rmin(st, xt), yt)) {
      double g = sqrt(p2 + pz * pz);
      double s = -dzz / (F * pz + az * g);
      double xe = s * (F * dtx + ax * g);
      double ye = s * (F * dty + ay * g);
      if (0.0 <= rmin(xe, ye) && ye * dxx <= xe * dyy && xe <= dxx) *tnew = cand;
    }
  }
  if (*tnew > rmin(xt, yt) && (!(ax * dzz + az * dxx > 0.0 && xt > yt)))
    edge2(tnew, xt, yt, F, ay, (ax * dzz + az * dxx) / dxz,
          (-ax * dxx + az * dzz) / dxz, dyy, dxz);
}

double orc_edge(double tnew, double st, double xt, double F, double ax,
                double ay, double az, double dxx, double dzz)
{
  edge2(&tnew, st, xt, F, ax, ay, az, dxx, dzz);
  return tnew;
}

double orc_surf(double tnew, double st, double xt, double yt, double F,
                double ax, double ay, double az, double dxx, double dyy,
                double dzz, double dxz)
{
  surf3(&tnew, st, xt, yt, F, ax, ay, az, dxx, dyy, dzz, dxz);
  return tnew;
}

/* ---------------------------------------------------------------------- */
/* Direction frames.  For each sweep: march axis m and step s, the local x
 * axis (ux, sign sx) and local y axis (uy, sign sy) of the upwind plane.
 * Read off the loads and model constants of the six kernels
 * (AN_ISO_1_OpenACC.c:195-212, 348-364, 500-516, 652-668, 805-821, 960-976)
 * and the reference driver's spacing arguments (:1198-1235).                        */
typedef struct { int m, s, ux, sx, uy, sy; } frame_t;
static const frame_t FRAMES[6] = {
    {0, +1, 2, +1, 1, +1}, /* x_down: i -> i+1, x^=+z, y^=+y */
    {0, -1, 2, -1, 1, -1}, /* x_up:   i -> i-1, x^=-z, y^=-y */
    {2, +1, 1, +1, 0, +1}, /* k_up:   k -> k+1, x^=+y, y^=+x */
    {2, -1, 1, -1, 0, -1}, /* k_down: k -> k-1, x^=-y, y^=-x */
    {1, -1, 0, -1, 2, -1}, /* y_up:   j -> j-1, x^=-x, y^=-z */
    {1, +1, 0, +1, 2, +1}, /* y_down: j -> j+1, x^=+x, y^=+z */
};

typedef struct {
  double F, ax, ay, az;                  /* model in the local frame */
  double dzz, dxx, dyy, dxy, dxz, dyz;   /* spacings in the local frame */
  long sm, su, sv;                       /* signed index strides: march, x^, y^ */
} local_t;

static double diag(const double h[3], int a, int b)
{
  /* DXYP/DXZP/DYZP of the reference driver (:1159-1161): lower axis first. */
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return sqrt(h[lo] * h[lo] + h[hi] * h[hi]);
}

static local_t make_local(int dir, const int n[3], double F, const double a[3],
                          const double h[3])
{
  const frame_t *f = &FRAMES[dir];
  const long H0 = n[0] + 2, H1 = n[1] + 2;
  const long stride[3] = {1, H0, H0 * H1};
  local_t L;
  L.F = F;
  L.ax = f->sx * a[f->ux];
  L.ay = f->sy * a[f->uy];
  L.az = f->s * a[f->m];
  L.dzz = h[f->m];
  L.dxx = h[f->ux];
  L.dyy = h[f->uy];
  L.dxy = diag(h, f->ux, f->uy);
  L.dxz = diag(h, f->ux, f->m);
  L.dyz = diag(h, f->uy, f->m);
  L.sm = f->s * stride[f->m];
  L.su = f->sx * stride[f->ux];
  L.sv = f->sy * stride[f->uy];
  return L;
}

/* Per-cell update shared by the six kernels (AN_ISO_1_OpenACC.c:195-326).
 * c = linear index of the source-plane node, the target is c + sm. */
static void cell_update(double *T, long c, const local_t *L)
{
  const long su = L->su, sv = L->sv;
  const double F = L->F;
  const double dzz = L->dzz, dxx = L->dxx, dyy = L->dyy;
  const double dxy = L->dxy, dxz = L->dxz, dyz = L->dyz;

  double told = T[c + L->sm];
  double tnew = fabs(told);
  double st = fabs(T[c]);
  double xt = fabs(T[c + su]);
  double yt = fabs(T[c + sv]);
  double xnt = fabs(T[c - su]);
  double ynt = fabs(T[c - sv]);
  double txm = fabs(T[c + su - sv]);
  double txy = fabs(T[c + su + sv]);
  double tym = fabs(T[c - su + sv]);
  double txnyn = fabs(T[c - su - sv]);

  double ax = L->ax, ay = L->ay, az = L->az;
  if (told < 0) { /* negative side of the surface: a flips (:207-212) */
    ax *= -1;
    ay *= -1;
    az *= -1;
  }

  if (st < tnew) { /* one-point update from st (:214-228) */
    double d = (F * F - (ax * ax + ay * ay)) * dzz * dzz;
    if (d > 0.0) {
      d = az * dzz + sqrt(d);
      if (d > 0.0) tnew = rmin(tnew, st + dzz * dzz / d);
    }
  }

  /* the reference's gate (:231): xnt and ynt are not part of it */
  if (!(tnew < st && tnew < rmin(xt, yt) && tnew < rmin(txy, txm) &&
        tnew < rmin(txnyn, tym))) {
    surf3(&tnew, st, xt, txy, F, ax, ay, az, dxx, dyy, dzz, dxz);
    surf3(&tnew, st, xt, txm, F, ax, -ay, az, dxx, dyy, dzz, dxz);
    surf3(&tnew, st, xnt, txnyn, F, -ax, -ay, az, dxx, dyy, dzz, dxz);
    surf3(&tnew, st, xnt, tym, F, -ax, ay, az, dxx, dyy, dzz, dxz);
    surf3(&tnew, st, ynt, txnyn, F, -ay, -ax, az, dyy, dxx, dzz, dyz);
    surf3(&tnew, st, ynt, txm, F, -ay, ax, az, dyy, dxx, dzz, dyz);
    surf3(&tnew, st, yt, txy, F, ay, ax, az, dyy, dxx, dzz, dyz);
    surf3(&tnew, st, yt, tym, F, ay, -ax, az, dyy, dxx, dzz, dyz);

    edge2(&tnew, st, xt, F, ax, ay, az, dxx, dzz);
    edge2(&tnew, st, txy, F, (ax * dxx + ay * dyy) / dxy,
          (-ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
    edge2(&tnew, st, yt, F, ay, ax, az, dyy, dzz);
    edge2(&tnew, st, tym, F, -(ax * dxx - ay * dyy) / dxy,
          (ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
    edge2(&tnew, st, xnt, F, -ax, ay, az, dxx, dzz);
    edge2(&tnew, st, txnyn, F, -(ax * dxx + ay * dyy) / dxy,
          (-ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
    edge2(&tnew, st, ynt, F, -ay, ax, az, dyy, dzz);
    edge2(&tnew, st, txm, F, (ax * dxx - ay * dyy) / dxy,
          (ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
  }

  T[c + L->sm] = rsign(told) * tnew; /* :326 */
}

/* Box grids (nx, ny, nz); the reference itself is cubic (one HEIGHT). */
void orc_plane3(int dir, double *T, const int n[3], int p, double F,
                const double a[3], const double h[3])
{
  const frame_t *f = &FRAMES[dir];
  const local_t L = make_local(dir, n, F, a, h);
  const long H0 = n[0] + 2, H1 = n[1] + 2;
  const long stride[3] = {1, H0, H0 * H1};
  const long base = (long)p * stride[f->m];
  const long s_u = stride[f->ux], s_v = stride[f->uy];
  const long nu = n[f->ux], nv = n[f->uy];
  long u;
#pragma omp parallel for schedule(static) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (u = 1; u <= nu; ++u) {
    long v;
    for (v = 1; v <= nv; ++v) cell_update(T, base + u * s_u + v * s_v, &L);
  }
}

void orc_direction3(int dir, double *T, const int n[3], double F,
                    const double a[3], const double h[3])
{
  const int nm = n[FRAMES[dir].m];
  int p;
  if (FRAMES[dir].s > 0)
    for (p = 1; p < nm; ++p) orc_plane3(dir, T, n, p, F, a, h);
  else
    for (p = nm; p > 1; --p) orc_plane3(dir, T, n, p, F, a, h);
}

void orc_sweeps3(double *T, const int n[3], int iterations, double F,
                 const double a[3], const double h[3])
{
  int it, d;
  for (it = 0; it < iterations; ++it)
    for (d = 0; d < 6; ++d) orc_direction3(d, T, n, F, a, h);
}

void orc_plane(int dir, double *T, int n, int p, double F, const double a[3],
               const double h[3])
{
  const int n3[3] = {n, n, n};
  orc_plane3(dir, T, n3, p, F, a, h);
}

void orc_direction(int dir, double *T, int n, double F, const double a[3],
                   const double h[3])
{
  const int n3[3] = {n, n, n};
  orc_direction3(dir, T, n3, F, a, h);
}

void orc_sweeps(double *T, int n, int iterations, double F, const double a[3],
                const double h[3])
{
  const int n3[3] = {n, n, n};
  orc_sweeps3(T, n3, iterations, F, a, h);
}

/* ---------------------------------------------------------------------- */
/* Initializer (Fold3dPMM.c).  Model constants of Fold3dPMM.h:49-52.       */

#define ORC_AX (-0.35)
#define ORC_AY 0.4
#define ORC_AZ 0.7
#define ORC_F 1.1

/* Closed-form traveltime from a point source for the fold model
 * (Fold3dPMM.c:9-26); d = node - source. */
double orc_fold_point(double F, double ax, double ay, double az, double x,
                      double y, double z)
{
  double r2 = x * x + y * y + z * z;
  double ad = ax * x + ay * y + az * z;
  double disc = (F * F - (ax * ax + ay * ay + az * az)) * r2 + ad * ad;
  if (disc > 0.0) {
    disc = ad + sqrt(disc);
    if (disc > 0.0) return r2 / disc;
  }
  return 1.0 * ORC_MAXT;
}

/* Implicit surface whose zero set seeds the front (Fold3dPMM.c:29-33). */
double orc_impsurf(double x, double y, double z)
{
  return 5.0 + (sin(x) + cos(y + x) + sin(x + y + z)) * 0.4 - z +
         (x - 5.0) * (y - 5.0) * (z - 5) / 55.0;
}

typedef struct { double *T; long H; } grid_t;

/* Node update from a surface point at offset (x,y,z) from node (i,j,k),
 * i.e. the node sits at -(x,y,z) from the point (Fold3dPMM.c:35-54). */
static void init_node(const grid_t *g, double x, double y, double z, int i,
                      int j, int k)
{
  double *t = &g->T[i + g->H * (j + g->H * (long)k)];
  double ax = ORC_AX, ay = ORC_AY, az = ORC_AZ;
  if (*t < 0.0) {
    ax *= -1;
    ay *= -1;
    az *= -1;
  }
  double tn = orc_fold_point(ORC_F, ax, ay, az, -x, -y, -z);
  if (tn < fabs(*t)) *t = rsign(*t) * tn;
}

/* The 44 nodes around voxel (i,j,k) touched per surface point
 * (Fold3dPMM.c:56-112).  Entries {di,dj,dk, ox,oy,oz}: node (i+di, j+dj,
 * k+dk) receives offset (xp - ox*dxx, yp - oy*dyy, zp - oz*dzz).  Two rows
 * keep the reference's index/offset mismatch (:82 and :86 write node
 * (i+1,j,k) resp. (i+2,j,k) with the offsets of (i+1,j-1,k) resp.
 * (i+1,j+2,k)). */
static const signed char VOXEL_STENCIL[44][6] = {
    {0, 0, 0, 0, 0, 0},    {1, 0, 0, 1, 0, 0},    {0, 1, 0, 0, 1, 0},
    {0, 0, 1, 0, 0, 1},    {1, 1, 0, 1, 1, 0},    {1, 0, 1, 1, 0, 1},
    {0, 1, 1, 0, 1, 1},    {1, 1, 1, 1, 1, 1},    {-1, 0, 0, -1, 0, 0},
    {-1, 1, 0, -1, 1, 0},  {-1, 0, 1, -1, 0, 1},  {-1, 1, 1, -1, 1, 1},
    {2, 0, 0, 2, 0, 0},    {2, 1, 0, 2, 1, 0},    {2, 0, 1, 2, 0, 1},
    {2, 1, 1, 2, 1, 1},    {0, -1, 0, 0, -1, 0},  {1, 0, 0, 1, -1, 0},
    {0, -1, 1, 0, -1, 1},  {1, -1, 1, 1, -1, 1},  {0, 2, 0, 0, 2, 0},
    {2, 0, 0, 1, 2, 0},    {0, 2, 1, 0, 2, 1},    {1, 2, 1, 1, 2, 1},
    {0, 0, -1, 0, 0, -1},  {1, 0, -1, 1, 0, -1},  {0, 1, -1, 0, 1, -1},
    {1, 1, -1, 1, 1, -1},  {0, 0, 2, 0, 0, 2},    {1, 0, 2, 1, 0, 2},
    {0, 1, 2, 0, 1, 2},    {1, 1, 2, 1, 1, 2},    {0, -1, -1, 0, -1, -1},
    {-1, 0, -1, -1, 0, -1}, {-1, -1, 0, -1, -1, 0}, {0, 2, 2, 0, 2, 2},
    {2, 0, 2, 2, 0, 2},    {2, 2, 0, 2, 2, 0},    {1, -1, -1, 1, -1, -1},
    {-1, 1, -1, -1, 1, -1}, {-1, -1, 1, -1, -1, 1}, {1, 2, 2, 1, 2, 2},
    {2, 1, 2, 2, 1, 2},    {2, 2, 1, 2, 2, 1},
};

static double offs(double p, int o, double h)
{
  /* xp, xp-dxx, xp+dxx, xp-2*dxx exactly as written in the reference */
  switch (o) {
  case 0: return p;
  case 1: return p - h;
  case -1: return p + h;
  default: return p - 2 * h;
  }
}

static void update_voxel(const grid_t *g, int i, int j, int k, double xp,
                         double yp, double zp, double dxx, double dyy,
                         double dzz)
{
  int e;
  for (e = 0; e < 44; ++e) {
    const signed char *s = VOXEL_STENCIL[e];
    init_node(g, offs(xp, s[3], dxx), offs(yp, s[4], dyy), offs(zp, s[5], dzz),
              i + s[0], j + s[1], k + s[2]);
  }
}

/* Blended extra seed points between the l-th crossing and earlier ones
 * (Fold3dPMM.c:271-383).  Weighted means with weights summing to 3 or 4;
 * the z coordinate of several 3-point blends uses the y coordinate of the
 * middle point, as the reference does (:309-379). */
static void seed_points(const grid_t *g, int i, int j, int k, const double *xs,
                        const double *ys, const double *zs, int np, double dxx,
                        double dyy, double dzz)
{
  int l;
  for (l = 0; l < np; ++l) {
    update_voxel(g, i, j, k, xs[l], ys[l], zs[l], dxx, dyy, dzz);
    /* two-point blends (2p_l + p_m)/3 and (p_l + 2p_m)/3 for m = l-1..l-5,
     * in the order the reference visits them */
    int back;
    for (back = 1; back <= 5 && back <= l; ++back) {
      const int m = l - back;
      double xp = (2 * xs[l] + xs[m]) / 3;
      double yp = (2 * ys[l] + ys[m]) / 3;
      double zp = (2 * zs[l] + zs[m]) / 3;
      update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
      xp = (xs[l] + 2 * xs[m]) / 3;
      yp = (ys[l] + 2 * ys[m]) / 3;
      zp = (zs[l] + 2 * zs[m]) / 3;
      update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
      if (back == 3) {
        /* (l, l-2, l-3) blends (:305-322) */
        const int a2 = l - 2, a3 = l - 3;
        double xp = (xs[l] + xs[a2] + xs[a3]) / 3;
        double yp = (ys[l] + ys[a2] + ys[a3]) / 3;
        double zp = (zs[l] + ys[a2] + zs[a3]) / 3;
        update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
        xp = (2 * xs[l] + xs[a2] + xs[a3]) / 4;
        yp = (2 * ys[l] + ys[a2] + ys[a3]) / 4;
        zp = (2 * zs[l] + ys[a2] + zs[a3]) / 4;
        update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
        xp = (xs[l] + 2 * xs[a2] + xs[a3]) / 4;
        yp = (ys[l] + 2 * ys[a2] + ys[a3]) / 4;
        zp = (zs[l] + 2 * ys[a2] + zs[a3]) / 4;
        update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
        xp = (xs[l] + xs[a2] + 2 * xs[a3]) / 4;
        yp = (ys[l] + ys[a2] + 2 * ys[a3]) / 4;
        zp = (zs[l] + ys[a2] + 2 * zs[a3]) / 4;
        update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
      }
      if (back == 4 || back == 5) {
        /* three-point means (p_l + p_a + p_b)/3 (:334-346, :359-375) */
        static const signed char tri4[3][2] = {{2, 3}, {2, 4}, {3, 4}};
        static const signed char tri5[5][2] = {{2, 3}, {2, 4}, {2, 5}, {3, 5}, {3, 4}};
        const signed char(*tri)[2] = back == 4 ? tri4 : tri5;
        int nt = back == 4 ? 3 : 5, t;
        for (t = 0; t < nt; ++t) {
          const int pa = l - tri[t][0], pb = l - tri[t][1];
          double xp = (xs[l] + xs[pa] + xs[pb]) / 3;
          double yp = (ys[l] + ys[pa] + ys[pb]) / 3;
          double zp = (zs[l] + ys[pa] + zs[pb]) / 3;
          update_voxel(g, i, j, k, xp, yp, zp, dxx, dyy, dzz);
        }
      }
    }
  }
}

static double node_coord(int i, double h) { return (i - 1) * h + 0.; } /* :115-126 */

void orc_initialiser(double *T, int n)
{
  const double h = (10. - 0.) / (n - 1.0); /* AN_ISO_1_OpenACC.c:1154 */
  const double dxx = h, dyy = h, dzz = h;
  const long H = n + 2;
  grid_t g = {T, H};
  double *IF = (double *)malloc(sizeof(double) * H * H * H);
  int i, j, k;

  for (k = 0; k < n + 2; k++)
    for (j = 0; j < n + 2; j++)
      for (i = 0; i < n + 2; i++) {
        long c = i + H * (j + H * (long)k);
        IF[c] = orc_impsurf(node_coord(i, dxx), node_coord(j, dyy), node_coord(k, dzz));
        T[c] = rsign(IF[c]) * ORC_MAXT;
      }

  /* Surface crossings on the 12 voxel edges by linear interpolation of IF
   * (Fold3dPMM.c:163-269), visited in the reference's order: x-edges, y-edges,
   * z-edges.  Edge table {corner a, corner b, axis, fixed offsets}. */
  for (i = 1; i < n; i++)
    for (j = 1; j < n; j++)
      for (k = 1; k < n; k++) {
        double f[2][2][2]; /* f[di][dj][dk] */
        int a, b, cc;
        for (a = 0; a < 2; a++)
          for (b = 0; b < 2; b++)
            for (cc = 0; cc < 2; cc++)
              f[a][b][cc] = IF[(i + a) + H * ((j + b) + H * (long)(k + cc))];
        double xs[12], ys[12], zs[12];
        int np = 0;
        double t;
        /* x-edges at (y,z) = (0,0), (0,dzz), (dyy,0), (dyy,dzz) */
        static const signed char xe[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
        static const signed char ye[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}}; /* (x,z) */
        static const signed char ze[4][2] = {{0, 0}, {1, 1}, {1, 0}, {0, 1}}; /* (x,y) */
        int e;
        for (e = 0; e < 4; e++) {
          double f0 = f[0][xe[e][0]][xe[e][1]], f1 = f[1][xe[e][0]][xe[e][1]];
          t = dxx * f0 / (f0 - f1);
          if (0.0 <= t && t <= dxx) {
            xs[np] = t; ys[np] = xe[e][0] ? dyy : 0.0; zs[np] = xe[e][1] ? dzz : 0.0; np++;
          }
        }
        for (e = 0; e < 4; e++) {
          double f0 = f[ye[e][0]][0][ye[e][1]], f1 = f[ye[e][0]][1][ye[e][1]];
          t = dyy * f0 / (f0 - f1);
          if (0.0 <= t && t <= dyy) {
            xs[np] = ye[e][0] ? dxx : 0.0; ys[np] = t; zs[np] = ye[e][1] ? dzz : 0.0; np++;
          }
        }
        for (e = 0; e < 4; e++) {
          double f0 = f[ze[e][0]][ze[e][1]][0], f1 = f[ze[e][0]][ze[e][1]][1];
          t = dzz * f0 / (f0 - f1);
          if (0.0 <= t && t <= dzz) {
            xs[np] = ze[e][0] ? dxx : 0.0; ys[np] = ze[e][1] ? dyy : 0.0; zs[np] = t; np++;
          }
        }
        seed_points(&g, i, j, k, xs, ys, zs, np, dxx, dyy, dzz);
      }
  free(IF);
}
