/*
 * ORACLE — FP64 CPU restatement of the tilted elliptic PMM model
 *     sqrt(grad T^T M grad T) = s(x) = 1/v(x),   M symmetric positive definite.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no such model;
 * this file restates the project's own local solver (the sweep structure is
 * the reference's: AN_ISO_1_OpenACC.c:1192-1237 plane order, :195-326 the
 * 3x3 upwind patch of one target node).
 *
 * Local solver.  With N = M^-1 the traveltime of a straight segment d is
 * s*|d|_N.  The new value of target X is the minimum over the current value
 * and, for every simplex S of the upwind patch (9 vertices, 16 edges, 8
 * triangles fanned around the centre node), of
 *     min_{Y in S}  T_lin(Y) + s |X - Y|_N .
 * For a simplex with base Y0, edge matrix E (k columns) and D = X - Y0:
 *     G = E^T N E, b = E^T N D, c = D^T N D, P = G^-1, l0 = P b, kappa = c - b.l0
 * and with the value differences delta_i = T_i - T_0:
 *     mu = P delta, q = delta.mu, r = s^2 - q,  root = sqrt(kappa r)
 *     barycentrics  lambda = l0 - (root/r) mu   (must lie in the simplex)
 *     candidate     T_0 + delta.l0 + root                      (r > 0)
 * (stationary point of the convex function above).  Here every candidate of
 * every simplex is evaluated; the GPU kernel prunes, and must agree.
 * Candidates are evaluated with fused multiply-adds at fixed places (fma()
 * below: one rounding for a*b + c), the operations the GPU's kernels run, so
 * that a constant-metric solve agrees with them bit for bit in FP64.  The
 * per-node constants are correctly rounded divisions here; the GPU's
 * per-node-metric kernels use hardware reciprocals there and agree to a few
 * ulps.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BIG 1e15 /* a vertex at or above this is unreached (MAXT = 1e16) */

/* sweep frames of the reference's six kernels (m, s, ux, sx, uy, sy) */
static const int FR[6][6] = {
    {0, +1, 2, +1, 1, +1}, {0, -1, 2, -1, 1, -1}, {2, +1, 1, +1, 0, +1},
    {2, -1, 1, -1, 0, -1}, {1, -1, 0, -1, 2, -1}, {1, +1, 0, +1, 2, +1},
};

/* ring order of the 8 neighbours around the centre (a, b) in units of
 * (dxx, dyy): +x, +x+y, +y, -x+y, -x, -x-y, -y, +x-y */
static const int RING[8][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1},
                               {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};

void orc_tilted_metric(double theta, double phi, double a1, double a2,
                       double a3, double M[6])
{
  /* M = R diag(a1^2, a2^2, a3^2) R^T, R = Rz(phi) Ry(theta) */
  const double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
  const double R[3][3] = {{cp * ct, -sp, cp * st},
                          {sp * ct, cp, sp * st},
                          {-st, 0.0, ct}};
  const double d[3] = {a1 * a1, a2 * a2, a3 * a3};
  double F[3][3];
  int i, j, k;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      double acc = 0.0;
      for (k = 0; k < 3; k++) acc += R[i][k] * d[k] * R[j][k];
      F[i][j] = acc;
    }
  M[0] = F[0][0]; M[1] = F[1][1]; M[2] = F[2][2];
  M[3] = F[0][1]; M[4] = F[0][2]; M[5] = F[1][2];
}

/* inverse of a symmetric 3x3 given as (xx, yy, zz, xy, xz, yz) */
void orc_sym3_inverse(const double M[6], double N[6])
{
  const double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5];
  const double A = b * c - f * f, B = a * c - e * e, C = a * b - d * d;
  const double D = e * f - c * d, E = d * f - b * e, F = d * e - a * f;
  const double det = a * A + d * D + e * E;
  N[0] = A / det; N[1] = B / det; N[2] = C / det;
  N[3] = D / det; N[4] = E / det; N[5] = F / det;
}

/* Local-frame constants of one direction for inverse metric N (global axes). */
void orc_ellip_constants(int dir, const double N[6], const double h[3],
                         orc_ellip_consts *K)
{
  const int *f = FR[dir];
  const int u[3] = {f[2], f[4], f[0]};
  const double sg[3] = {f[3], f[5], f[1]};
  /* full symmetric N in global axes */
  const double Ng[3][3] = {{N[0], N[3], N[4]}, {N[3], N[1], N[5]}, {N[4], N[5], N[2]}};
  double NL[3][3];
  int i, j, k;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) NL[i][j] = sg[i] * sg[j] * Ng[u[i]][u[j]];
  const double dxx = h[u[0]], dyy = h[u[1]], dzz = h[u[2]];
  double pos[9][3];
  pos[0][0] = 0.0; pos[0][1] = 0.0; pos[0][2] = 0.0;
  for (k = 0; k < 8; k++) {
    pos[k + 1][0] = RING[k][0] * dxx;
    pos[k + 1][1] = RING[k][1] * dyy;
    pos[k + 1][2] = 0.0;
  }
  const double X[3] = {0.0, 0.0, dzz};
#define NDOT(p, q) \
  ((p)[0] * (NL[0][0] * (q)[0] + NL[0][1] * (q)[1] + NL[0][2] * (q)[2]) + \
   (p)[1] * (NL[1][0] * (q)[0] + NL[1][1] * (q)[1] + NL[1][2] * (q)[2]) + \
   (p)[2] * (NL[2][0] * (q)[0] + NL[2][1] * (q)[1] + NL[2][2] * (q)[2]))
  for (k = 0; k < 9; k++) {
    double D[3] = {X[0] - pos[k][0], X[1] - pos[k][1], X[2] - pos[k][2]};
    K->vert[k] = sqrt(NDOT(D, D));
  }
  /* edges: 0..7 spokes centre -> ring k, 8..15 rims ring k -> ring k+1 */
  for (k = 0; k < 16; k++) {
    const int v0 = k < 8 ? 0 : 1 + (k - 8);
    const int v1 = k < 8 ? 1 + k : 1 + ((k - 8 + 1) % 8);
    double E[3], D[3];
    for (i = 0; i < 3; i++) {
      E[i] = pos[v1][i] - pos[v0][i];
      D[i] = X[i] - pos[v0][i];
    }
    const double g = NDOT(E, E), b = NDOT(E, D), c = NDOT(D, D);
    const double P = 1.0 / g;
    const double l0 = P * b;
    K->edge[k][0] = P;
    K->edge[k][1] = l0;
    K->edge[k][2] = c - b * l0;
  }
  /* triangles: centre, ring k, ring k+1 */
  for (k = 0; k < 8; k++) {
    const int v1 = 1 + k, v2 = 1 + (k + 1) % 8;
    double E1[3], E2[3], D[3];
    for (i = 0; i < 3; i++) {
      E1[i] = pos[v1][i];
      E2[i] = pos[v2][i];
      D[i] = X[i];
    }
    const double g11 = NDOT(E1, E1), g12 = NDOT(E1, E2), g22 = NDOT(E2, E2);
    const double b1 = NDOT(E1, D), b2 = NDOT(E2, D), c = NDOT(D, D);
    const double det = g11 * g22 - g12 * g12;
    const double P11 = g22 / det, P12 = -g12 / det, P22 = g11 / det;
    const double l1 = P11 * b1 + P12 * b2, l2 = P12 * b1 + P22 * b2;
    K->tri[k][0] = P11;
    K->tri[k][1] = P12;
    K->tri[k][2] = P22;
    K->tri[k][3] = l1;
    K->tri[k][4] = l2;
    K->tri[k][5] = c - (b1 * l1 + b2 * l2);
  }
#undef NDOT
}

/* Edge candidate; t0 at the base, t1 at the far end. Returns +inf if invalid. */
double orc_ellip_edge(double t0, double t1, double s2, const double E[3])
{
  const double P = E[0], l0 = E[1], kap = E[2];
  if (!(t0 < BIG && t1 < BIG)) return INFINITY;
  const double d = t1 - t0;
  const double mu = P * d;
  const double r = fma(-d, mu, s2);
  if (!(r > 0.0)) return INFINITY;
  const double root = sqrt(kap * r);
  const double rm = -root * mu;
  if (!(fma(l0, r, rm) >= 0.0)) return INFINITY;
  if (!(fma(l0 - 1.0, r, rm) <= 0.0)) return INFINITY;
  return fma(d, l0, t0) + root;
}

/* Triangle candidate over (centre t0, ring t1, ring t2). */
double orc_ellip_tri(double t0, double t1, double t2, double s2, const double Tr[6])
{
  if (!(t0 < BIG && t1 < BIG && t2 < BIG)) return INFINITY;
  const double d1 = t1 - t0, d2 = t2 - t0;
  const double m1 = fma(Tr[0], d1, Tr[1] * d2);
  const double m2 = fma(Tr[1], d1, Tr[2] * d2);
  const double r = fma(-d2, m2, fma(-d1, m1, s2));
  if (!(r > 0.0)) return INFINITY;
  const double root = sqrt(Tr[5] * r);
  const double rm1 = -root * m1, rm2 = -root * m2;
  if (!(fma(Tr[3], r, rm1) >= 0.0)) return INFINITY;
  if (!(fma(Tr[4], r, rm2) >= 0.0)) return INFINITY;
  if (!(fma(Tr[3] + Tr[4] - 1.0, r, rm1 + rm2) <= 0.0)) return INFINITY;
  return fma(d2, Tr[4], fma(d1, Tr[3], t0)) + root;
}

/* New value of one node from its patch values p[9] (centre + ring order). */
double orc_ellip_node(double told, const double p[9], double s,
                      const orc_ellip_consts *K)
{
  double best = told;
  const double s2 = s * s;
  int k;
  for (k = 0; k < 9; k++) {
    if (p[k] < BIG) {
      const double c = fma(s, K->vert[k], p[k]);
      if (c < best) best = c;
    }
  }
  for (k = 0; k < 16; k++) {
    const double t0 = k < 8 ? p[0] : p[1 + (k - 8)];
    const double t1 = k < 8 ? p[1 + k] : p[1 + ((k - 8 + 1) % 8)];
    const double c = orc_ellip_edge(t0, t1, s2, K->edge[k]);
    if (c < best) best = c;
  }
  for (k = 0; k < 8; k++) {
    const double c = orc_ellip_tri(p[0], p[1 + k], p[1 + (k + 1) % 8], s2, K->tri[k]);
    if (c < best) best = c;
  }
  return best;
}

static int g_ell_threads = 0;
void orc_ellip_set_threads(int t) { g_ell_threads = t; }

/* One plane of one direction.  T, speed, theta.. on the padded grid
 * (nx+2)(ny+2)(nz+2).  Constant model when the field pointers are NULL:
 * M6 = metric, speed = v.  Returns max |dT| over the plane. */
double orc_ellip_plane(int dir, double *T, const int n[3], int p,
                       const double h[3], const double M6[6], double speed,
                       const double *speed_field, const double *theta,
                       const double *phi, const double *ra, const double *rb)
{
  const int *f = FR[dir];
  const long H0 = n[0] + 2, H1 = n[1] + 2;
  const long stride[3] = {1, H0, H0 * H1};
  const long sm = f[1] * stride[f[0]], su = f[3] * stride[f[2]], sv = f[5] * stride[f[4]];
  const int nu = n[f[2]], nv = n[f[4]];
  const long base = (long)p * stride[f[0]];
  orc_ellip_consts Kc;
  const int per_node = theta != NULL;
  if (!per_node) {
    double N[6];
    orc_sym3_inverse(M6, N);
    orc_ellip_constants(dir, N, h, &Kc);
  }
  double maxd = 0.0;
  long a;
#pragma omp parallel for schedule(static) reduction(max : maxd) num_threads(g_ell_threads > 0 ? g_ell_threads : omp_get_max_threads())
  for (a = 1; a <= nu; a++) {
    long b;
    for (b = 1; b <= nv; b++) {
      const long c = base + a * stride[f[2]] + b * stride[f[4]];
      const long t = c + sm;
      double pv[9];
      pv[0] = T[c];
      int k;
      for (k = 0; k < 8; k++) pv[1 + k] = T[c + RING[k][0] * su + RING[k][1] * sv];
      const double v = speed_field ? speed_field[t] : speed;
      const double s = 1.0 / v;
      orc_ellip_consts Kn;
      const orc_ellip_consts *K = &Kc;
      if (per_node) {
        double M[6], N[6];
        orc_tilted_metric(theta[t], phi[t], 1.0, ra[t], rb[t], M);
        orc_sym3_inverse(M, N);
        orc_ellip_constants(dir, N, h, &Kn);
        K = &Kn;
      }
      const double told = T[t];
      const double tn = orc_ellip_node(told, pv, s, K);
      T[t] = tn;
      const double d = fabs(tn - told);
      if (d > maxd) maxd = d;
    }
  }
  return maxd;
}

int orc_ellip_sweeps(double *T, const int n[3], int max_iters, double tol,
                     const double h[3], const double M6[6], double speed,
                     const double *speed_field, const double *theta,
                     const double *phi, const double *ra, const double *rb,
                     double *max_dt)
{
  int it;
  double last = -1.0;
  for (it = 0; it < max_iters;) {
    double m = 0.0;
    int d;
    for (d = 0; d < 6; d++) {
      const int ax = FR[d][0];
      int p;
      if (FR[d][1] > 0) {
        for (p = 1; p < n[ax]; p++) {
          double x = orc_ellip_plane(d, T, n, p, h, M6, speed, speed_field, theta, phi, ra, rb);
          if (x > m) m = x;
        }
      } else {
        for (p = n[ax]; p > 1; p--) {
          double x = orc_ellip_plane(d, T, n, p, h, M6, speed, speed_field, theta, phi, ra, rb);
          if (x > m) m = x;
        }
      }
    }
    it++;
    last = m;
    if (tol > 0 && m < tol) break;
  }
  if (max_dt) *max_dt = last;
  return it;
}

/* ---------------------------------------------------------------------- */
/* Adjoint (FP64, independent of the GPU path): winning stencil per node by
 * exhaustive search over the six directions' candidates, local derivatives by
 * central differences of that candidate, the transposed upwind system solved
 * by a push in order of decreasing T followed by Jacobi passes to a fixed
 * point. */

static const double *g_sort_key;
static int cmp_desc(const void *a, const void *b)
{
  const double x = g_sort_key[*(const long *)a], y = g_sort_key[*(const long *)b];
  return x < y ? 1 : x > y ? -1 : 0;
}
static void sort_by_value_desc(long *idx, long m, const double *key)
{
  g_sort_key = key;
  qsort(idx, m, sizeof(long), cmp_desc);
}

typedef struct {
  int dir, type, id;  /* type 1 vertex, 2 edge, 3 triangle; id as in the kernels */
  long v[3];
  double w[3];
  double val;
} orc_stencil;

static void node_metric(const double *M6, const double *theta, const double *phi,
                        const double *ra, const double *rb, long t, double dth,
                        double N[6])
{
  if (theta) {
    double M[6];
    orc_tilted_metric(theta[t] + dth, phi[t], 1.0, ra[t], rb[t], M);
    orc_sym3_inverse(M, N);
  } else {
    orc_sym3_inverse(M6, N);
  }
}

/* Value of one candidate (type/id of direction d) at node X with inverse
 * metric N and slowness s; weights of its vertices in w (if non-NULL).
 * Returns +inf when the candidate is not admissible. */
static double eval_candidate_k(const double *T, const int n[3], long X, int d, int type,
                               int id, const orc_ellip_consts *Kp, double s, long v[3],
                               double w[3]);

static double eval_candidate(const double *T, const int n[3], const double h[3], long X,
                             int d, int type, int id, const double N[6], double s,
                             long v[3], double w[3])
{
  orc_ellip_consts K;
  orc_ellip_constants(d, N, h, &K);
  return eval_candidate_k(T, n, X, d, type, id, &K, s, v, w);
}

/* The same with the direction's simplex constants K (orc_ellip_constants of
 * direction d and metric N) given. */
static double eval_candidate_k(const double *T, const int n[3], long X, int d, int type,
                               int id, const orc_ellip_consts *Kp, double s, long v[3],
                               double w[3])
{
  const int *f = FR[d];
  const long H0 = n[0] + 2, H1 = n[1] + 2;
  const long stride[3] = {1, H0, H0 * H1};
  const long sm = f[1] * stride[f[0]], su = f[3] * stride[f[2]], sv = f[5] * stride[f[4]];
  const long c = X - sm;
  const orc_ellip_consts *K = Kp;
  long pidx[9];
  int k;
  pidx[0] = c;
  for (k = 0; k < 8; k++) pidx[1 + k] = c + RING[k][0] * su + RING[k][1] * sv;
  const double s2 = s * s;
  if (type == 1) {
    const double t0 = T[pidx[id]];
    if (!(t0 < BIG)) return INFINITY;
    if (v) { v[0] = pidx[id]; w[0] = 1.0; }
    return fma(s, K->vert[id], t0);
  }
  if (type == 2) {
    const int q0 = id < 8 ? 0 : 1 + (id - 8), q1 = id < 8 ? 1 + id : 1 + ((id - 8 + 1) % 8);
    const double t0 = T[pidx[q0]], t1 = T[pidx[q1]];
    const double val = orc_ellip_edge(t0, t1, s2, K->edge[id]);
    if (v && isfinite(val)) {
      const double P = K->edge[id][0], l0 = K->edge[id][1], kap = K->edge[id][2];
      const double dd = t1 - t0, mu = P * dd, r = s2 - dd * mu;
      const double lam = l0 - sqrt(kap * r) / r * mu;
      v[0] = pidx[q0]; v[1] = pidx[q1];
      w[0] = 1.0 - lam; w[1] = lam;
    }
    return val;
  }
  {
    const int q1 = 1 + id, q2 = 1 + (id + 1) % 8;
    const double t0 = T[pidx[0]], t1 = T[pidx[q1]], t2 = T[pidx[q2]];
    const double *Tr = K->tri[id];
    const double val = orc_ellip_tri(t0, t1, t2, s2, Tr);
    if (v && isfinite(val)) {
      const double d1 = t1 - t0, d2 = t2 - t0;
      const double m1 = Tr[0] * d1 + Tr[1] * d2, m2 = Tr[1] * d1 + Tr[2] * d2;
      const double r = s2 - (d1 * m1 + d2 * m2);
      const double q = sqrt(Tr[5] * r) / r;
      const double x1 = Tr[3] - q * m1, x2 = Tr[4] - q * m2;
      v[0] = pidx[0]; v[1] = pidx[q1]; v[2] = pidx[q2];
      w[0] = 1.0 - x1 - x2; w[1] = x1; w[2] = x2;
    }
    return val;
  }
}

/* Minimum over every candidate of the six patches of node X (the node's local
 * solver without its current value), with the winning candidate. */
static double node_min(const double *T, const int n[3], const double h[3], long X,
                       const double N[6], double s, int *dir, int *type, int *id)
{
  const long H0 = n[0] + 2, H1 = n[1] + 2;
  const long ijk[3] = {X % H0, (X / H0) % H1, X / (H0 * H1)};
  double best = INFINITY;
  int d, t, k;
  *type = 0;
  for (d = 0; d < 6; d++) {
    const long cm = ijk[FR[d][0]] - FR[d][1];
    if (cm < 1 || cm > n[FR[d][0]]) continue;
    orc_ellip_consts K;
    orc_ellip_constants(d, N, h, &K);
    for (t = 1; t <= 3; t++) {
      const int nid = t == 1 ? 9 : t == 2 ? 16 : 8;
      for (k = 0; k < nid; k++) {
        const double val = eval_candidate_k(T, n, X, d, t, k, &K, s, NULL, NULL);
        if (val < best) {
          best = val;
          *dir = d; *type = t; *id = k;
        }
      }
    }
  }
  return best;
}

/* Minimum over the candidates of one simplex of direction d and of its faces
 * (a triangle, its three edges and three vertices; an edge and its two
 * vertices; a vertex): where the simplex's candidate wins, this is the node's
 * local solver restricted to that simplex, continuously differentiable across
 * the simplex's boundary, and blind to near-ties with other simplices. */
static double closure_min(const double *T, const int n[3], const double h[3], long X, int d,
                          int type, int id, const double N[6], double s)
{
  orc_ellip_consts K;
  orc_ellip_constants(d, N, h, &K);
  double best = eval_candidate_k(T, n, X, d, type, id, &K, s, NULL, NULL);
  int qv[3], nq = 0, ev[3], ne = 0, k;
  if (type == 2) {
    qv[nq++] = id < 8 ? 0 : 1 + (id - 8);
    qv[nq++] = id < 8 ? 1 + id : 1 + ((id - 8 + 1) % 8);
  } else if (type == 3) {
    qv[nq++] = 0;
    qv[nq++] = 1 + id;
    qv[nq++] = 1 + (id + 1) % 8;
    ev[ne++] = id;             /* spoke centre -> ring id */
    ev[ne++] = (id + 1) % 8;   /* spoke centre -> ring id+1 */
    ev[ne++] = 8 + id;         /* rim ring id -> ring id+1 */
  }
  for (k = 0; k < ne; k++) {
    const double v = eval_candidate_k(T, n, X, d, 2, ev[k], &K, s, NULL, NULL);
    if (v < best) best = v;
  }
  for (k = 0; k < nq; k++) {
    const double v = eval_candidate_k(T, n, X, d, 1, qv[k], &K, s, NULL, NULL);
    if (v < best) best = v;
  }
  return best;
}

/* Local derivatives of the adjoint: 1 (default) = central differences of the
 * winning simplex's closure (closure_min), the derivative of the candidate the
 * node follows; 0 = of the node's whole local solver (every candidate of the
 * six patches), which at a node where two patches' candidates tie within the
 * step averages the two slopes across the kink. */
static int g_adj_local = 1;
void orc_ellip_adjoint_local(int mode) { g_adj_local = mode; }

int orc_ellip_adjoint(const double *T, const int n[3], const double h[3], const double M6[6],
                      double speed, const double *speed_field, const double *theta,
                      const double *phi, const double *ra, const double *rb, int nrec,
                      const long *rec_idx, const double *rec_res, double *grad_s,
                      double *grad_th, double rel_tol, int max_passes, double *mu_out,
                      unsigned char *fixed_out)
{
  const long H0 = n[0] + 2, H1 = n[1] + 2, H2 = n[2] + 2;
  const long count = H0 * H1 * H2;
  const long stride[3] = {1, H0, H0 * H1};
  orc_stencil *S = (orc_stencil *)calloc(count, sizeof(orc_stencil));
  double *dTds = (double *)calloc(count, sizeof(double));
  double *dTdth = (double *)calloc(count, sizeof(double));
  double *mu = (double *)calloc(count, sizeof(double));
  double *nxt = (double *)calloc(count, sizeof(double));
  double *res = (double *)calloc(count, sizeof(double));
  long *order = (long *)malloc(sizeof(long) * count);
  long i, nord = 0;
  int r;
  for (r = 0; r < nrec; r++) res[rec_idx[r]] += rec_res[r];

#pragma omp parallel for schedule(dynamic, 64)
  for (i = 0; i < count; i++) {
    const long x = i % H0, y = (i / H0) % H1, z = i / (H0 * H1);
    S[i].type = 0;
    if (x < 1 || x > n[0] || y < 1 || y > n[1] || z < 1 || z > n[2]) continue;
    if (!(T[i] < BIG)) continue;
    const double s = 1.0 / (speed_field ? speed_field[i] : speed);
    double N[6];
    node_metric(M6, theta, phi, ra, rb, i, 0.0, N);
    const double best = node_min(T, n, h, i, N, s, &S[i].dir, &S[i].type, &S[i].id);
    if (!(best <= T[i] * (1.0 + rel_tol))) { S[i].type = 0; continue; }
    S[i].val = best;
    eval_candidate(T, n, h, i, S[i].dir, S[i].type, S[i].id, N, s, S[i].v, S[i].w);
    /* local derivatives: central differences of the node's whole local
     * solver (the minimum over every candidate of the six patches), which is
     * smooth across the boundary between a simplex and its faces */
    int dd, tt, ii;
    const int wd = S[i].dir, wt = S[i].type, wi = S[i].id;
    const double es = 1e-6 * s;
    if (g_adj_local)
      dTds[i] = (closure_min(T, n, h, i, wd, wt, wi, N, s + es) -
                 closure_min(T, n, h, i, wd, wt, wi, N, s - es)) / (2 * es);
    else
      dTds[i] = (node_min(T, n, h, i, N, s + es, &dd, &tt, &ii) -
                 node_min(T, n, h, i, N, s - es, &dd, &tt, &ii)) / (2 * es);
    if (theta) {
      const double et = 1e-6;
      double Np[6], Nm[6];
      node_metric(M6, theta, phi, ra, rb, i, et, Np);
      node_metric(M6, theta, phi, ra, rb, i, -et, Nm);
      if (g_adj_local)
        dTdth[i] = (closure_min(T, n, h, i, wd, wt, wi, Np, s) -
                    closure_min(T, n, h, i, wd, wt, wi, Nm, s)) / (2 * et);
      else
        dTdth[i] = (node_min(T, n, h, i, Np, s, &dd, &tt, &ii) -
                    node_min(T, n, h, i, Nm, s, &dd, &tt, &ii)) / (2 * et);
    }
  }
  (void)stride;
  /* push in order of decreasing T */
  for (i = 0; i < count; i++)
    if (S[i].type) order[nord++] = i;
  sort_by_value_desc(order, nord, T);
  for (i = 0; i < count; i++) mu[i] = res[i];
  for (i = 0; i < nord; i++) {
    const orc_stencil *st = &S[order[i]];
    const int nv = st->type == 1 ? 1 : st->type == 2 ? 2 : 3;
    int q;
    for (q = 0; q < nv; q++) mu[st->v[q]] += st->w[q] * mu[order[i]];
  }
  /* Jacobi passes: mu = res + A^T mu, until a pass changes nothing */
  int pass;
  for (pass = 0; pass < max_passes; pass++) {
    for (i = 0; i < count; i++) nxt[i] = res[i];
    for (i = 0; i < count; i++) {
      if (!S[i].type) continue;
      const int nv = S[i].type == 1 ? 1 : S[i].type == 2 ? 2 : 3;
      int q;
      for (q = 0; q < nv; q++) nxt[S[i].v[q]] += S[i].w[q] * mu[i];
    }
    double dmax = 0.0;
    for (i = 0; i < count; i++) {
      const double dd = fabs(nxt[i] - mu[i]);
      if (dd > dmax) dmax = dd;
      mu[i] = nxt[i];
    }
    if (dmax == 0.0) break;
  }
  if (mu_out) memcpy(mu_out, mu, sizeof(double) * count);
  if (fixed_out)
    for (i = 0; i < count; i++) fixed_out[i] = !S[i].type && T[i] < BIG;
  for (i = 0; i < count; i++) {
    if (!S[i].type) continue;
    grad_s[i] += mu[i] * dTds[i];
    if (grad_th && theta) grad_th[i] += mu[i] * dTdth[i];
  }
  free(S); free(dTds); free(dTdth); free(mu); free(nxt); free(res); free(order);
  return pass;
}

