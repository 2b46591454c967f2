/*
 * Fold3dPMM.h — drop-in for the reference's Stencil-OpenACC/Fold3dPMM.h.
 *
 * Same macros (MAXT, _DOUBLE_, domain, model constants; Fold3dPMM.h:6-52) and
 * the same eight host prototypes with the same names and argument order
 * (Fold3dPMM.h:58-86), served by libfold3d.so.  Differences, all additive:
 *   - the grid stride is a run-time value (nx+2), not the compile-time HEIGHT.
 *     ImplicitInitialiser() sets it; InitNode()/updateVoxel(), which take no
 *     size, use the last value set (or fold3d_set_grid()).  _nx/_ny/_nz/HEIGHT
 *     remain only as defaults for drivers that still use them.
 *   - the six plane-step kernels of AN_ISO_1_OpenACC.c:180-1092 are provided
 *     with their reference signatures (device pointer T, plane index, spacings,
 *     F, ay, ax, az), run on the MI355X and return when the plane is done, like
 *     the reference's synchronous OpenACC regions.  They use F and a as passed
 *     (as the CUDA twin Stencil-CUDA/AN_ISO_2_GPU.cu:151 does; the OpenACC
 *     kernels overwrite them with the same macro values).
 * Whole solves, FP32, the elliptic model and error codes live in fold3d.h.
 */
#ifndef FOLD3DPMM_H
#define FOLD3DPMM_H

#include <stdint.h>

#define MAXT 10e15*1.0
#define EDGE_GODUNOV_2D 1.0
#define CBD
#define FOLD

#define UNROLLED

/* SINGLE is not honoured by the reference's kernels (they take double *T,
 * AN_ISO_1_OpenACC.c:180); this header keeps the macro for source
 * compatibility.  FP32 solves go through fold3d_solve(). */
#ifdef SINGLE
#define _DOUBLE_ float
#else
#define _DOUBLE_ double
#endif

/* default grid of the reference build (Fold3dPMM.h:27-30) */
#ifndef _nx
#define _nx 1024
#define _ny 1024
#define _nz 1024
#define HEIGHT 1026
#endif

#ifndef sign
#define sign(a) (a > 0) ? 1 : -1
#endif

#define _xmin 0.
#define _xmax 10.

#define _ymin 0.
#define _ymax 10.

#define _zmin 0.
#define _zmax 10.

#define _axx -0.35
#define _ayy  0.4
#define _azz  0.7
#define _Fc 1.1

#ifdef __cplusplus
extern "C" {
#endif

/* Host initializer (reference Fold3dPMM.c:9-389). */
_DOUBLE_ InitFoldFromPoint(_DOUBLE_ F, _DOUBLE_ ax, _DOUBLE_ ay, _DOUBLE_ az,
                           _DOUBLE_ x, _DOUBLE_ y, _DOUBLE_ z);
_DOUBLE_ ImpSurfF(_DOUBLE_ x, _DOUBLE_ y, _DOUBLE_ z);
void InitNode(_DOUBLE_ x, _DOUBLE_ y, _DOUBLE_ z, int i, int j, int k, _DOUBLE_ *T);
void updateVoxel(_DOUBLE_ *T, int i, int j, int k, _DOUBLE_ xp, _DOUBLE_ yp,
                 _DOUBLE_ zp, _DOUBLE_ dxx, _DOUBLE_ dyy, _DOUBLE_ dzz);
_DOUBLE_ x(int i, _DOUBLE_ DX);
_DOUBLE_ y(int j, _DOUBLE_ dy);
_DOUBLE_ z(int k, _DOUBLE_ dz);
void ImplicitInitialiser(_DOUBLE_ *T, int nx, int ny, int nz, _DOUBLE_ dxx,
                         _DOUBLE_ dyy, _DOUBLE_ dzz);

/* Grid used by InitNode/updateVoxel and the plane steps below. */
void fold3d_set_grid(int nx, int ny, int nz);
void fold3d_get_grid(int *nx, int *ny, int *nz);

/* InitNode / updateVoxel on an explicit grid of nx x ny x nz interior nodes
 * (same results; they ignore the grid recorded by ImplicitInitialiser or
 * fold3d_set_grid).  As in the reference, the node indices, with updateVoxel's
 * offsets of -1..2, must lie in 0..n+1 per axis. */
void fold3d_init_node(double x, double y, double z, int i, int j, int k, double *T, int nx,
                      int ny, int nz);
void fold3d_update_voxel(double *T, int i, int j, int k, double xp, double yp, double zp,
                         double dxx, double dyy, double dzz, int nx, int ny, int nz);

/* Plane steps on a device field (reference AN_ISO_1_OpenACC.c:180, 334, 485,
 * 636, 790, 944).  Synchronous. */
void function_x_down(double *T, int i, double dzz, double dxx, double dyy,
                     double dxy, double dxz, double dyz, double F, double ay,
                     double ax, double az);
void function_x_up(double *T, int i, double dzz, double dxx, double dyy,
                   double dxy, double dxz, double dyz, double F, double ay,
                   double ax, double az);
void function_y_up(double *T, int j, double dzz, double dxx, double dyy,
                   double dxy, double dxz, double dyz, double F, double ay,
                   double ax, double az);
void function_y_down(double *T, int j, double dzz, double dxx, double dyy,
                     double dxy, double dxz, double dyz, double F, double ay,
                     double ax, double az);
void function_k_up(double *T, int k, double dzz, double dxx, double dyy,
                   double dxy, double dxz, double dyz, double F, double ay,
                   double ax, double az);
void function_k_down(double *T, int k, double dzz, double dxx, double dyy,
                     double dxy, double dxz, double dyz, double F, double ay,
                     double ax, double az);

#ifdef __cplusplus
}
#endif
#endif
