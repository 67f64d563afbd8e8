// le_lb.h -- the Lees-Edwards plane condition on the distributions, per
// plane site: the reprojection of the five distributions that cross the
// plane (le_reproject, model_le.c:245-330).  Used by the staged plane
// kernel (lees_edwards.hip) and by the single-pass step's edge launch
// (le_step_tile.h), which forms it from the post-collision values it holds
// in registers; the arithmetic is one function, so both give the same bits.

#ifndef LDW_LE_LB_H
#define LDW_LE_LB_H

#include <hip/hip_runtime.h>

#include "d3q19.h"

namespace ldw {

struct LeProps {
  int p[2][5];            // cx = +1 (side 0) and cx = -1 (side 1), in p order
};

constexpr LeProps le_make_props() {
  LeProps t{};
  int n0 = 0, n1 = 0;
  for (int p = 1; p < NVEL; p++) {
    if (CV.v[p][0] == +1) t.p[0][n0++] = p;
    if (CV.v[p][0] == -1) t.p[1][n1++] = p;
  }
  return t;
}

constexpr LeProps LE_PROPS = le_make_props();

// moments of the post-collision f at a plane site, then for each crossing
// p: f_p + w_p (rho u.c / cs2 + ds:Q / 2 cs2^2), with the jump
// du = (0, duy, 0); side 0 is x = L (cx = +1 crosses), side 1 x = L + 1
__device__ __forceinline__ void le_reproject_site(const double (&f)[NVEL], int side, double duy,
						  double (&out)[5]) {
  double rho = 0.0;
#pragma unroll
  for (int q = 0; q < NVEL; q++) rho += f[q];
  double gm[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < NVEL; q++)
#pragma unroll
    for (int n = 0; n < 3; n++) gm[n] += CV.v[q][n]*f[q];
  const double du[3] = {0.0, duy, 0.0};
  double ds[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) ds[a][b] = (gm[a]*du[b] + du[a]*gm[b] + rho*du[a]*du[b]);
  const double cs2 = (1.0/3.0);
  const double rcs2 = 1.0/cs2;
#pragma unroll
  for (int ip = 0; ip < 5; ip++) {
    const int q = side ? LE_PROPS.p[1][ip] : LE_PROPS.p[0][ip];
    const double udotc = du[1]*CV.v[q][1];
    double sdotq = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
	const double dab = cs2*(a == b);
	const double qq = (CV.v[q][a]*CV.v[q][b] - dab);
	sdotq += ds[a][b]*qq;
      }
    double fq = f[q];
    fq += weight(q)*(rho*udotc*rcs2 + 0.5*sdotq*rcs2*rcs2);
    out[ip] = fq;
  }
}

}  // namespace ldw

#endif
