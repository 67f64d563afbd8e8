// advect.h -- the face fluxes of the higher-order advection schemes and
// of the external chemical-potential gradient, shared by every
// Cahn-Hilliard kernel (direct, tiled, with and without Lees-Edwards
// planes).
//
// Each face lies between a lower cell L and the upper cell U = L + 1 along
// one axis; the flux is the face velocity uf times an interpolation of phi.
// The operands are summed left to right in the reference's order, so FP64
// results are bit-identical:
//   order 4 (advection.c:1153-1262): uf (-a1 phi(L-1) + a2 phi(L)
//            + a2 phi(U) - a1 phi(U+1)), a1 = 1/16, a2 = 9/16;
//   order 5 (advection.c:1264-1430, Li 1997): five points biased upwind,
//            from below when `lower` holds, else from above.
// The west face of cell x takes `lower` = (uf > 0), every other face
// `lower` = !(uf < 0), as the reference's branches do.

#ifndef LDW_ADVECT_H
#define LDW_ADVECT_H

#include <hip/hip_runtime.h>

namespace ldw {
namespace adv {

template <typename T>
__device__ __forceinline__ T order4(T uf, T lm1, T l0, T u0, T up1) {
  const T a1 = T(1.0/16.0), a2 = T(9.0/16.0);
  return uf*(- a1*lm1 + a2*l0 + a2*u0 - a1*up1);
}

template <typename T>
__device__ __forceinline__ T order5(T uf, bool lower, T lm2, T lm1, T l0, T u0, T up1, T up2) {
  const T a1 = T(0.055453), a2 = T(-0.305147), a3 = T(0.916054), a4 = T(0.361520), a5 = T(-0.027880);
  if (lower) return uf*(a1*lm2 + a2*lm1 + a3*l0 + a4*u0 + a5*up1);
  return uf*(a1*up2 + a2*up1 + a3*u0 + a4*l0 + a5*lm1);
}

// phi_ch_flux_mu_ext_kernel (phi_cahn_hilliard.c:1373-1398): every face
// flux loses M grad mu_ext along its axis, after the diffusive fluxes
template <typename T>
__device__ __forceinline__ void mu_ext(T mob, const double (&gm)[3], T & fw, T & fe, T & fyp,
				       T & fym, T & fzp, T & fzm) {
  const T gx = mob*T(gm[0]), gy = mob*T(gm[1]), gz = mob*T(gm[2]);
  fw -= gx; fe -= gx;
  fyp -= gy; fym -= gy;
  fzp -= gz; fzm -= gz;
}

}  // namespace adv
}  // namespace ldw

#endif
