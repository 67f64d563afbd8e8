// lb_site.h -- per-site D3Q19 MRT collision and pull-index helpers shared
// by the collide kernel (lb_collide.hip) and the single-pass step kernel
// (binary_step.hip).  See lb_collide.hip for the arithmetic contract.

#ifndef LDW_LB_SITE_H
#define LDW_LB_SITE_H

#include <hip/hip_runtime.h>

#include "d3q19.h"
#include "kernels.h"

namespace ldw {
namespace site {

constexpr bool pow2_coeff(double c) {
  if (c < 0) c = -c;
  return c == 2.0 || c == 0.5 || c == 0.25 || c == 0.125 || c == 4.0;
}

template <typename T> struct Prec;
template <> struct Prec<double> {
  static __device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
};
template <> struct Prec<float> {
  static __device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};

// acc + x*c with c a compile-time table entry (c != 0)
template <typename T, int R, int C, bool MODE2F>
__device__ __forceinline__ T addterm(T acc, T x) {
  constexpr double c = MODE2F ? MI.v[R][C] : MA.v[R][C];
  if constexpr (c == 1.0) return acc + x;
  else if constexpr (c == -1.0) return acc - x;
  else if constexpr (pow2_coeff(c)) return Prec<T>::fma_(x, T(c), acc);
  else return acc + x*T(c);
}

template <typename T, int R, int C, bool MODE2F>
__device__ __forceinline__ T firstterm(T x) {
  constexpr double c = MODE2F ? MI.v[R][C] : MA.v[R][C];
  if constexpr (c == 1.0) return x;
  else if constexpr (c == -1.0) return -x;
  else return x*T(c);
}

// Row R of the transform: sum over C < NC of table[R][C]*v[C], summed in
// increasing C, zero entries skipped.
template <typename T, int R, int C, int NC, bool MODE2F, bool FIRST>
struct Row {
  static __device__ __forceinline__ T run(T acc, const T * v) {
    if constexpr (C == NC) {
      return acc;
    }
    else {
      constexpr double c = MODE2F ? MI.v[R][C] : MA.v[R][C];
      if constexpr (c == 0.0) {
	return Row<T, R, C + 1, NC, MODE2F, FIRST>::run(acc, v);
      }
      else if constexpr (FIRST) {
	return Row<T, R, C + 1, NC, MODE2F, false>::run(firstterm<T, R, C, MODE2F>(v[C]), v);
      }
      else {
	return Row<T, R, C + 1, NC, MODE2F, false>::run(addterm<T, R, C, MODE2F>(acc, v[C]), v);
      }
    }
  }
};

template <typename T, int NROW, int NC, bool MODE2F, int R = 0>
__device__ __forceinline__ void transform(const T * in, T * out) {
  if constexpr (R < NROW) {
    out[R] = Row<T, R, 0, NC, MODE2F, true>::run(T(0), in);
    transform<T, NROW, NC, MODE2F, R + 1>(in, out);
  }
}

// Neighbour offsets: for direction p the pull source is x - c_p.
template <int P>
__device__ __forceinline__ int64_t src_index(const int64_t xs[3], const int ys[3],
					     const int zs[3]) {
  constexpr int cx = CV.v[P][0], cy = CV.v[P][1], cz = CV.v[P][2];
  return xs[1 - cx] + ys[1 - cy] + zs[1 - cz];
}

template <typename T, int P = 0>
__device__ __forceinline__ void load_pull(const T * __restrict__ g, int64_t nsite,
					  const int64_t xs[3], const int ys[3],
					  const int zs[3], T * f) {
  if constexpr (P < NVEL) {
    f[P] = g[P*nsite + src_index<P>(xs, ys, zs)];
    load_pull<T, P + 1>(g, nsite, xs, ys, zs, f);
  }
}

template <typename T, bool GHOSTS_ZERO>
__device__ __forceinline__ void collide_site(const LbParams & prm, T force[3],
					     T f[NVEL], T & rho_out, T u[3]) {

  constexpr int NMODE = GHOSTS_ZERO ? 10 : NVEL;
  T mode[NVEL];
  transform<T, NMODE, NVEL, false>(f, mode);

  const T rdim = T(1.0)/T(3.0);
  T rho = mode[0];
  if constexpr (sizeof(T) == 4) rho = T(prm.rho0) + mode[0];  // shifted storage

  const T rrho = T(1.0)/rho;
  u[0] = rrho*(mode[1] + T(0.5)*force[0]);
  u[1] = rrho*(mode[2] + T(0.5)*force[1]);
  u[2] = rrho*(mode[3] + T(0.5)*force[2]);

  const T rtau = T(prm.rtau_shear);
  const T rtau_b = T(prm.rtau_bulk);

  // upper-triangle stress s_ab and equilibrium rho u_a u_b
  T s00 = mode[4], s01 = mode[5], s02 = mode[6];
  T s11 = mode[7], s12 = mode[8], s22 = mode[9];
  T q00 = rho*u[0]*u[0], q01 = rho*u[0]*u[1], q02 = rho*u[0]*u[2];
  T q11 = rho*u[1]*u[1], q12 = rho*u[1]*u[2], q22 = rho*u[2]*u[2];

  T tr_s = s00 + s11 + s22;
  T tr_q = q00 + q11 + q22;
  s00 -= rdim*tr_s; s11 -= rdim*tr_s; s22 -= rdim*tr_s;
  q00 -= rdim*tr_q; q11 -= rdim*tr_q; q22 -= rdim*tr_q;
  tr_s = tr_s - rtau_b*(tr_s - tr_q);

  const T w2 = T(2.0) - rtau;
#define RELAX_DIAG(s, q, a)						\
  s -= rtau*(s - q);							\
  s += rdim*tr_s;							\
  s += w2*(u[a]*force[a] + force[a]*u[a]);
#define RELAX_OFF(s, q, a, b)						\
  s -= rtau*(s - q);							\
  s += w2*(u[a]*force[b] + force[a]*u[b]);
  RELAX_DIAG(s00, q00, 0)
  RELAX_OFF(s01, q01, 0, 1)
  RELAX_OFF(s02, q02, 0, 2)
  RELAX_DIAG(s11, q11, 1)
  RELAX_OFF(s12, q12, 1, 2)
  RELAX_DIAG(s22, q22, 2)
#undef RELAX_DIAG
#undef RELAX_OFF

  mode[1] += force[0];
  mode[2] += force[1];
  mode[3] += force[2];
  mode[4] = s00; mode[5] = s01; mode[6] = s02;
  mode[7] = s11; mode[8] = s12; mode[9] = s22;

  if constexpr (!GHOSTS_ZERO) {
#pragma unroll
    for (int m = 10; m < NVEL; m++) {
      mode[m] = mode[m] - T(prm.rtau_ghost[m])*mode[m];
    }
  }

  transform<T, NVEL, NMODE, true>(mode, f);
  rho_out = rho;
}

template <typename T, int P = 0>
__device__ __forceinline__ void store_planes(T * __restrict__ g, int64_t nsite,
					     int64_t site, const T * f) {
  if constexpr (P < NVEL) {
    g[P*nsite + site] = f[P];
    store_planes<T, P + 1>(g, nsite, site, f);
  }
}

}  // namespace site
}  // namespace ldw

#endif
