/*
 * Pre-included (gcc -include) when the reference's own sources are compiled
 * into oracle/_ref/ as a checker.  It pulls in the reference's unmodified
 * Stencil-OpenACC/Fold3dPMM.h and then resizes the compile-time grid that
 * header fixes at 1024^3 (Fold3dPMM.h:27-30); REF_N comes from the command
 * line.  The include guard of Fold3dPMM.h keeps the sources' own
 * #include "Fold3dPMM.h" from undoing the resize.
 *
 * Test infrastructure only: nothing in the shipped library includes this.
 */
#include "Fold3dPMM.h"

#ifndef REF_N
#error "REF_N (nodes per axis) must be defined"
#endif

#undef _nx
#undef _ny
#undef _nz
#undef HEIGHT
#define _nx REF_N
#define _ny REF_N
#define _nz REF_N
#define HEIGHT (REF_N + 2)
