/*
 * ORACLE — CPU restatement of the reference's single-device Fold3dPMM solve.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline).  The shipped library never links it.
 *
 * Grid layout (reference Fold3dPMM.h:27-30, AN_ISO_1_OpenACC.c:195):
 *   T[i + H*(j + H*k)], H = n+2, nodes 1..n per axis, ghost layer 0 and n+1.
 * Sweep direction numbering (reference driver order, AN_ISO_1_OpenACC.c:1195-1236):
 *   0 x_down (+x)  1 x_up (-x)  2 k_up (+z)  3 k_down (-z)  4 y_up (-y)  5 y_down (+y)
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   orc_edge, orc_surf           bitwise vs the compiled reference (oracle/_ref) on
 *                                seeded tuples and committed KAT fixtures; a=0 case vs the
 *                                textbook isotropic update; brute-force simplex minimum.
 *   orc_fold_point               bitwise vs reference; closed form |d|/... properties.
 *   orc_plane / orc_sweeps       bitwise vs reference plane kernels and full solves
 *                                (committed fixtures at 16^3/24^3/32^3; _ref up to 128^3).
 *   orc_initialiser              bitwise vs reference ImplicitInitialiser fixtures.
 *   orc_ellip_*                  (elliptic model, no reference code) closed-form point
 *                                source, isotropic reduction, brute-force simplex minimum.
 */
#ifndef FOLD3D_ORACLE_H
#define FOLD3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fold model (F|grad T| + a.grad T = 1), reference AN_ISO_1_OpenACC.c ---- */
double orc_edge(double tnew, double st, double xt, double F, double ax,
                double ay, double az, double dxx, double dzz);
double orc_surf(double tnew, double st, double xt, double yt, double F,
                double ax, double ay, double az, double dxx, double dyy,
                double dzz, double dxz);
double orc_fold_point(double F, double ax, double ay, double az, double x,
                      double y, double z);
double orc_impsurf(double x, double y, double z);

/* Reference input on an n^3 grid over [0,10]^3 (Fold3dPMM.c:129-389). */
void orc_initialiser(double *T, int n);

/* One plane step: reads plane p along the march axis, writes plane p+step.
 * F, a = (ax, ay, az) in global axes; h = spacings (dx, dy, dz). */
void orc_plane(int dir, double *T, int n, int p, double F, const double a[3],
               const double h[3]);
/* Whole sweep along one direction in the reference plane order. */
void orc_direction(int dir, double *T, int n, double F, const double a[3],
                   const double h[3]);
/* `iterations` x the six directions in reference order. */
void orc_sweeps(double *T, int n, int iterations, double F, const double a[3],
                const double h[3]);

/* The same on a box grid n = (nx, ny, nz) (the reference's grid is cubic). */
void orc_plane3(int dir, double *T, const int n[3], int p, double F,
                const double a[3], const double h[3]);
void orc_direction3(int dir, double *T, const int n[3], double F,
                    const double a[3], const double h[3]);
void orc_sweeps3(double *T, const int n[3], int iterations, double F,
                 const double a[3], const double h[3]);

/* OpenMP thread count used by the plane loops (0 = OpenMP default). */
void orc_set_threads(int nthreads);

/* ---- tilted elliptic model sqrt(grad T^T M grad T) = 1/v (no reference code) ---- */
typedef struct {
  double vert[9];     /* |X - Y|_N for the centre and the 8 ring nodes */
  double edge[16][3]; /* P, l0, kappa: 8 spokes centre->ring k, 8 rims ring k->k+1 */
  double tri[8][6];   /* P11, P12, P22, l1, l2, kappa: centre, ring k, ring k+1 */
} orc_ellip_consts;

void orc_tilted_metric(double theta, double phi, double a1, double a2,
                       double a3, double M[6]);
void orc_sym3_inverse(const double M[6], double N[6]);
void orc_ellip_constants(int dir, const double N[6], const double h[3],
                         orc_ellip_consts *K);
double orc_ellip_edge(double t0, double t1, double s2, const double E[3]);
double orc_ellip_tri(double t0, double t1, double t2, double s2, const double Tr[6]);
double orc_ellip_node(double told, const double p[9], double s,
                      const orc_ellip_consts *K);
void orc_ellip_set_threads(int t);
double orc_ellip_plane(int dir, double *T, const int n[3], int p,
                       const double h[3], const double M6[6], double speed,
                       const double *speed_field, const double *theta,
                       const double *phi, const double *ra, const double *rb);
int orc_ellip_sweeps(double *T, const int n[3], int max_iters, double tol,
                     const double h[3], const double M6[6], double speed,
                     const double *speed_field, const double *theta,
                     const double *phi, const double *ra, const double *rb,
                     double *max_dt);

/* Adjoint of a converged elliptic field T (FP64, independent of the GPU path).
 * Adds dJ/d(slowness) to grad_s and, with theta..rb given, dJ/d(theta) to
 * grad_th for J = 1/2 sum_r res_r^2 with res_r the residual at node rec_idx[r].
 * A node's stencil is its best candidate over all six directions if that
 * reproduces T within rel_tol, else the node is held fixed.  Optional outputs:
 * the adjoint state mu and a flag per node that is reached but held fixed.
 * Returns the Jacobi passes used after the ordered push. */
/* 1 (default): local derivatives of the winning simplex's closure; 0: of the
 * node's whole local solver. */
void orc_ellip_adjoint_local(int mode);
int orc_ellip_adjoint(const double *T, const int n[3], const double h[3], const double M6[6],
                      double speed, const double *speed_field, const double *theta,
                      const double *phi, const double *ra, const double *rb, int nrec,
                      const long *rec_idx, const double *rec_res, double *grad_s,
                      double *grad_th, double rel_tol, int max_passes, double *mu_out,
                      unsigned char *fixed_out);

#ifdef __cplusplus
}
#endif
#endif
