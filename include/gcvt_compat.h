/*
 * gcvt_compat.h -- the reference's two hot-path entry points, kept with their
 * exact signatures and C++ linkage so that the reference's host headers
 * (gcvt.h, discretization.h) link against libcvt_hip.so unchanged.
 *
 *   gCVT              replaces source/gcvt.cu:1087 (declared source/gcvt.h:29);
 *                     runs cvt_compat_gcvt on the caller's host buffers.
 *   discretization_d  replaces source/discretization.cu:87-120 (declared
 *                     source/discretization.h:66-67); runs cvt_rasterize.
 *
 * Like the reference, both return nothing and end the process (exit(1),
 * message on stderr) when the engine reports an error.
 */
#ifndef GCVT_COMPAT_H
#define GCVT_COMPAT_H

#ifdef __cplusplus
void gCVT(short *Voronoi, float *density_d, bool *mask, int size, int depth, int maxIter);
void discretization_d(double *points, double *weight, int num_point, int *triangle, int num_tri,
                      float *density, double scale, int n);
#endif

#endif /* GCVT_COMPAT_H */
