// ch_conserve.h -- the compensated Cahn-Hilliard update of
// cahn_hilliard_options_conserve 1, shared by the Cahn-Hilliard kernels.
//
// Built without value-changing optimisations (no fast-math, no FMA
// contraction), so every operation keeps its order and rounding.

#ifndef LDW_CH_CONSERVE_H
#define LDW_CH_CONSERVE_H

#include <hip/hip_runtime.h>

namespace ldw {

// kahan_add_double (util_sum.c:30-40)
template <typename T>
__device__ __forceinline__ void kahan_add(T & sum, T & cs, T val) {
  const T y = val + cs;
  const T t = sum + y;
  cs = y - (t - sum);
  sum = t;
}

// phi_ch_csum_kernel (phi_cahn_hilliard.c:1181-1218) at one site, wz = 1:
// phi and its compensation cs updated by the six face fluxes in the
// reference's order (fe, fw, fy, fy(y-1), fz, fz(z-1))
template <typename T>
__device__ __forceinline__ void ch_update_compensated(T & phi, T & cs, T fe, T fw, T fyp,
						      T fym, T fzp, T fzm) {
  kahan_add(phi, cs, -fe);
  kahan_add(phi, cs, fw);
  kahan_add(phi, cs, -fyp);
  kahan_add(phi, cs, fym);
  kahan_add(phi, cs, -fzp);
  kahan_add(phi, cs, fzm);
}

}  // namespace ldw

#endif
