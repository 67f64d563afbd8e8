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
#include "lb_site.h"

namespace ldw {

namespace {

using namespace site;

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
