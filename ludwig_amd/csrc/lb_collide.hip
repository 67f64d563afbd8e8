// lb_collide.hip -- K1: D3Q19 MRT collision with body force, fused with
// pull-scheme propagation, on an SoA distribution layout.
//
// Per site (collision.c:253-593, lb_collision_mrt1_site):
//   m = M f (unrolled in the reference's order, zero terms elided),
//   rho = m0, u = (j + F/2)/rho, stress relaxed with the shear and bulk
//   rates, force term (2 - omega)(uF + Fu), j += F, ghosts relaxed to
//   zero, f = M^-1 m.
// Pull propagation (propagation.c:155-199) is folded into the load: the
// buffer holds post-collision values g, and the pre-collision value at x
// in direction p is g(x - c_p).  One kernel reads 19 planes and writes 19
// planes, rho and u: 360 B/site in FP64, 180 B/site in FP32.
//
// FP64 parity build: compiled with -ffp-contract=off, so every product and
// sum rounds as in the reference's gcc -O2 x86-64 build (which has no
// FMA).  fma() is used only where the product is exact (coefficients
// +-2^k), which gives the same rounding as a separate multiply and add.
//
// FP32 storage holds f - w_p rho0.  In that representation mode 0 is the
// density deviation and every other mode is unchanged, because
// sum_p w_p ma_mp = delta_m0; mode -> f then yields the deviation directly
// since mi_p0 = w_p.

#include <hip/hip_runtime.h>

#include "d3q19.h"
#include "kernels.h"

namespace ldw {

namespace {

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

template <typename T, bool GHOSTS_ZERO, bool PULL, bool HAVE_F>
__global__ void __launch_bounds__(256)
collide_kernel(Geom geo, LbParams prm, const T * __restrict__ g_in,
	       T * __restrict__ g_out, const T * __restrict__ force,
	       T * __restrict__ rho, T * __restrict__ u) {

  const int64_t site = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (site >= geo.nsite) return;

  const int nyz = geo.ny*geo.nz;
  const int x = (int) (site / nyz);
  const int rem = (int) (site - (int64_t) x*nyz);
  const int y = rem / geo.nz;
  const int z = rem - y*geo.nz;

  T f[NVEL];
  if constexpr (PULL) {
    // index of the pull source x - c for c = +1, 0, -1 (array slot 1 - c)
    const int xm = (x == 0) ? geo.nx - 1 : x - 1;
    const int xp = (x == geo.nx - 1) ? 0 : x + 1;
    const int ym = (y == 0) ? geo.ny - 1 : y - 1;
    const int yp = (y == geo.ny - 1) ? 0 : y + 1;
    const int zm = (z == 0) ? geo.nz - 1 : z - 1;
    const int zp = (z == geo.nz - 1) ? 0 : z + 1;
    const int64_t xs[3] = {(int64_t) xm*nyz, (int64_t) x*nyz, (int64_t) xp*nyz};
    const int ys[3] = {ym*geo.nz, y*geo.nz, yp*geo.nz};
    const int zs[3] = {zm, z, zp};
    load_pull<T>(g_in, geo.nsite, xs, ys, zs, f);
  }
  else {
#pragma unroll
    for (int p = 0; p < NVEL; p++) f[p] = g_in[p*geo.nsite + site];
  }

  T fc[3];
  if constexpr (HAVE_F) {
    fc[0] = T(prm.fglobal[0]) + force[site];
    fc[1] = T(prm.fglobal[1]) + force[geo.nsite + site];
    fc[2] = T(prm.fglobal[2]) + force[2*geo.nsite + site];
  }
  else {
    fc[0] = T(prm.fglobal[0]);
    fc[1] = T(prm.fglobal[1]);
    fc[2] = T(prm.fglobal[2]);
  }

  T r, uu[3];
  collide_site<T, GHOSTS_ZERO>(prm, fc, f, r, uu);

  store_planes<T>(g_out, geo.nsite, site, f);
  if (rho) rho[site] = r;
  if (u) {
    u[site] = uu[0];
    u[geo.nsite + site] = uu[1];
    u[2*geo.nsite + site] = uu[2];
  }
}

template <typename T>
void launch_collide(const Geom & geo, const LbParams & prm, const T * g_in,
		    T * g_out, const T * force, T * rho, T * u, int pull,
		    hipStream_t s) {
  const int nthr = 256;
  const int64_t nblk = (geo.nsite + nthr - 1)/nthr;
  const bool gz = (prm.relax == 0);
  const bool hf = (force != nullptr);
#define LAUNCH(GZ, PL, HF)						\
  hipLaunchKernelGGL((collide_kernel<T, GZ, PL, HF>), dim3((unsigned) nblk), \
		     dim3(nthr), 0, s, geo, prm, g_in, g_out, force, rho, u)
  if (gz) {
    if (pull) { if (hf) LAUNCH(true, true, true); else LAUNCH(true, true, false); }
    else      { if (hf) LAUNCH(true, false, true); else LAUNCH(true, false, false); }
  }
  else {
    if (pull) { if (hf) LAUNCH(false, true, true); else LAUNCH(false, true, false); }
    else      { if (hf) LAUNCH(false, false, true); else LAUNCH(false, false, false); }
  }
#undef LAUNCH
}

template <typename T>
__global__ void __launch_bounds__(256)
propagate_kernel(Geom geo, const T * __restrict__ in, T * __restrict__ out) {
  const int64_t site = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (site >= geo.nsite) return;
  const int nyz = geo.ny*geo.nz;
  const int x = (int) (site / nyz);
  const int rem = (int) (site - (int64_t) x*nyz);
  const int y = rem / geo.nz;
  const int z = rem - y*geo.nz;
  const int xm = (x == 0) ? geo.nx - 1 : x - 1;
  const int xp = (x == geo.nx - 1) ? 0 : x + 1;
  const int ym = (y == 0) ? geo.ny - 1 : y - 1;
  const int yp = (y == geo.ny - 1) ? 0 : y + 1;
  const int zm = (z == 0) ? geo.nz - 1 : z - 1;
  const int zp = (z == geo.nz - 1) ? 0 : z + 1;
  const int64_t xs[3] = {(int64_t) xm*nyz, (int64_t) x*nyz, (int64_t) xp*nyz};
  const int ys[3] = {ym*geo.nz, y*geo.nz, yp*geo.nz};
  const int zs[3] = {zm, z, zp};
  T f[NVEL];
  load_pull<T>(in, geo.nsite, xs, ys, zs, f);
  store_planes<T>(out, geo.nsite, site, f);
}

}  // namespace

void launch_collide_f64(const Geom & g, const LbParams & p, const double * g_in,
			double * g_out, const double * force, double * rho,
			double * u, int pull, hipStream_t s) {
  launch_collide<double>(g, p, g_in, g_out, force, rho, u, pull, s);
}

void launch_collide_f32(const Geom & g, const LbParams & p, const float * g_in,
			float * g_out, const float * force, float * rho,
			float * u, int pull, hipStream_t s) {
  launch_collide<float>(g, p, g_in, g_out, force, rho, u, pull, s);
}

void launch_propagate_f64(const Geom & g, const double * in, double * out,
			  hipStream_t s) {
  const int64_t nblk = (g.nsite + 255)/256;
  hipLaunchKernelGGL(propagate_kernel<double>, dim3((unsigned) nblk), dim3(256),
		     0, s, g, in, out);
}

void launch_propagate_f32(const Geom & g, const float * in, float * out,
			  hipStream_t s) {
  const int64_t nblk = (g.nsite + 255)/256;
  hipLaunchKernelGGL(propagate_kernel<float>, dim3((unsigned) nblk), dim3(256),
		     0, s, g, in, out);
}

}  // namespace ldw
