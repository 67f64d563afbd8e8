// step_tile.h -- one tile of the single-pass binary-fluid step: the TY x TZ
// (y, z) columns [y0, y0+TY) x [z0, z0+TZ) of planes [xb, xe) of one
// periodic lattice.  Used by the single-lattice step kernel (binary_step.hip)
// and by the replica-ensemble kernel (ensemble.hip), so a replica advances
// with exactly the arithmetic of a single-lattice run.
//
// For each site the step forms the thermodynamic force and the new phi from
// the LDS tiles (fe_tile.h), then collides the site's 19 pulled
// distributions with that force (lb_site.h).  The force never goes to
// memory.  The distributions and velocities of the current plane are
// requested before the plane's Q values are computed, so their latency
// overlaps that work.

#ifndef LDW_STEP_TILE_H
#define LDW_STEP_TILE_H

#include <hip/hip_runtime.h>

#include "fe_tile.h"
#include "kernels.h"
#include "lb_site.h"

namespace ldw {
namespace tile {

template <typename T>
__device__ __forceinline__ void load_f(T (&f)[NVEL], const T * __restrict__ g_in,
				       const Geom & g, int x, int y, int z) {
  const int nyz = g.ny*g.nz;
  const int xm = (x == 0) ? g.nx - 1 : x - 1;
  const int xp = (x == g.nx - 1) ? 0 : x + 1;
  const int ym = (y == 0) ? g.ny - 1 : y - 1;
  const int yp = (y == g.ny - 1) ? 0 : y + 1;
  const int zm = (z == 0) ? g.nz - 1 : z - 1;
  const int zp = (z == g.nz - 1) ? 0 : z + 1;
  const int64_t xs[3] = {(int64_t) xm*nyz, (int64_t) x*nyz, (int64_t) xp*nyz};
  const int ys[3] = {ym*g.nz, y*g.nz, yp*g.nz};
  const int zs[3] = {zm, z, zp};
  site::load_pull<T>(g_in, g.nsite, xs, ys, zs, f);
}

// phi -> phi_out, u_prev -> u_next (the Cahn-Hilliard part reads u_prev),
// g_in -> g_out with pull indexing, rho written.  fz: the uniform force
// recorded by hydro_f_zero, added to the phi force.  CONS: the compensated
// update of cahn_hilliard_options_conserve 1, with the per-site
// compensation csum read and written.  GMU: the external chemical
// potential gradient is compiled in (fe.gmu_on then decides).  Ends with a
// block barrier, so the caller may reuse the LDS tiles at once.
template <typename T, class S, int ORDER, int METHOD, bool GZ, bool CONS = false, bool GMU = false>
__device__ __forceinline__ void step_tile(Smem<T, S> & s, const Geom & g, const FeParams & fe,
					  const LbParams & prm, int y0, int z0, int xb, int xe,
					  const T * __restrict__ phi, T * __restrict__ phi_out,
					  const T * __restrict__ u_prev, T * __restrict__ u_next,
					  T * __restrict__ rho, const T * __restrict__ g_in,
					  T * __restrict__ g_out, T fz0, T fz1, T fz2,
					  T * __restrict__ csum = nullptr) {
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int y = y0 + ty, z = z0 + tz;
  const bool active = (y < g.ny) && (z < g.nz);

  load_phi_planes<T, S, 4>(s, phi, g, xb - 2, y0, z0);
  T pre[S::NPRE];
  prefetch_phi_plane<T, S>(pre, phi, g, xb + 2, y0, z0);
  __syncthreads();
  compute_q_plane(s, fe, xb - 1);
  compute_q_plane(s, fe, xb);

  for (int x = xb; x < xe; x++) {
    store_phi_plane(s, pre, x + 2);
    if (x + 1 < xe) prefetch_phi_plane<T, S>(pre, phi, g, x + 3, y0, z0);
    __syncthreads();
    T uc[9];
    T fc[NVEL];
    if (active) {
      load_u(uc, u_prev, g, x, y, z);
      load_f(fc, g_in, g, x, y, z);
    }
    compute_q_plane(s, fe, x + 1);
    __syncthreads();

    if (active) {
      const int64_t site = ((int64_t) x*g.ny + y)*g.nz + z;
      T F[3], pnew;
      const T fzv[3] = {fz0, fz1, fz2};
      force_ch_site<T, S, ORDER, METHOD, true, GMU>(s, fe, x, ty, tz, uc, fzv, F, pnew,
					       CONS ? csum + site : nullptr);
      phi_out[site] = pnew;
      // site force: force_global + (hydro_f_zero value + phi force)
      // (collision.c:323-327, phi_force_colloid.c:457-462)
      T fs[3];
      fs[0] = T(prm.fglobal[0]) + F[0];
      fs[1] = T(prm.fglobal[1]) + F[1];
      fs[2] = T(prm.fglobal[2]) + F[2];
      T r, uu[3];
      site::collide_site<T, GZ>(prm, fs, fc, r, uu);
      site::store_planes<T>(g_out, g.nsite, site, fc);
      rho[site] = r;
      u_next[site] = uu[0];
      u_next[g.nsite + site] = uu[1];
      u_next[2*g.nsite + site] = uu[2];
    }
    __syncthreads();
  }
}

}  // namespace tile
}  // namespace ldw

#endif
