// grad27.h -- the 27-point gradient and Laplacian of one site
// (gradient_3d_27pt_fluid.c:219-368), shared by the direct kernels
// (fe_ch.hip) and the LDS-tiled ones (fe_tile.h).
//
// p(i, j, k), i, j, k in {0, 1, 2}, is phi at offset (i-1, j-1, k-1).  Each
// gradient component sums the 9 differences across the cube in the
// reference's order, left to right, and is scaled by 0.5/9; the Laplacian
// sums the 26 neighbours in x, y, z order, subtracts 26 phi and is scaled
// by 1/9.  With -ffp-contract=off the FP64 results equal the reference's.

#ifndef LDW_GRAD27_H
#define LDW_GRAD27_H

#include <hip/hip_runtime.h>

namespace ldw {

template <typename T, class P>
__device__ __forceinline__ void grad27(const P & p, T (&g)[3], T & delsq) {
  const T r9 = T(1.0/9.0);
  const T h9 = T(0.5)*r9;
  T sx = + p(2, 0, 0) - p(0, 0, 0);
  sx = sx + p(2, 0, 1) - p(0, 0, 1);
  sx = sx + p(2, 0, 2) - p(0, 0, 2);
  sx = sx + p(2, 1, 0) - p(0, 1, 0);
  sx = sx + p(2, 1, 1) - p(0, 1, 1);
  sx = sx + p(2, 1, 2) - p(0, 1, 2);
  sx = sx + p(2, 2, 0) - p(0, 2, 0);
  sx = sx + p(2, 2, 1) - p(0, 2, 1);
  sx = sx + p(2, 2, 2) - p(0, 2, 2);
  T sy = + p(0, 2, 0) - p(0, 0, 0);
  sy = sy + p(0, 2, 1) - p(0, 0, 1);
  sy = sy + p(0, 2, 2) - p(0, 0, 2);
  sy = sy + p(1, 2, 0) - p(1, 0, 0);
  sy = sy + p(1, 2, 1) - p(1, 0, 1);
  sy = sy + p(1, 2, 2) - p(1, 0, 2);
  sy = sy + p(2, 2, 0) - p(2, 0, 0);
  sy = sy + p(2, 2, 1) - p(2, 0, 1);
  sy = sy + p(2, 2, 2) - p(2, 0, 2);
  T sz = + p(0, 0, 2) - p(0, 0, 0);
  sz = sz + p(0, 1, 2) - p(0, 1, 0);
  sz = sz + p(0, 2, 2) - p(0, 2, 0);
  sz = sz + p(1, 0, 2) - p(1, 0, 0);
  sz = sz + p(1, 1, 2) - p(1, 1, 0);
  sz = sz + p(1, 2, 2) - p(1, 2, 0);
  sz = sz + p(2, 0, 2) - p(2, 0, 0);
  sz = sz + p(2, 1, 2) - p(2, 1, 0);
  sz = sz + p(2, 2, 2) - p(2, 2, 0);
  T sd = + p(0, 0, 0);
  sd = sd + p(0, 0, 1); sd = sd + p(0, 0, 2);
  sd = sd + p(0, 1, 0); sd = sd + p(0, 1, 1); sd = sd + p(0, 1, 2);
  sd = sd + p(0, 2, 0); sd = sd + p(0, 2, 1); sd = sd + p(0, 2, 2);
  sd = sd + p(1, 0, 0); sd = sd + p(1, 0, 1); sd = sd + p(1, 0, 2);
  sd = sd + p(1, 1, 0); sd = sd + p(1, 1, 2);
  sd = sd + p(1, 2, 0); sd = sd + p(1, 2, 1); sd = sd + p(1, 2, 2);
  sd = sd + p(2, 0, 0); sd = sd + p(2, 0, 1); sd = sd + p(2, 0, 2);
  sd = sd + p(2, 1, 0); sd = sd + p(2, 1, 1); sd = sd + p(2, 1, 2);
  sd = sd + p(2, 2, 0); sd = sd + p(2, 2, 1); sd = sd + p(2, 2, 2);
  sd = sd - T(26.0)*p(1, 1, 1);
  g[0] = h9*sx;
  g[1] = h9*sy;
  g[2] = h9*sz;
  delsq = r9*sd;
}

}  // namespace ldw

#endif
