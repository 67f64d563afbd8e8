// binary_step.hip -- the whole binary-fluid time step in one launch.
//
// The reference performs, per step (ludwig.c:528-860): the thermodynamic
// force from phi_t, the Cahn-Hilliard update phi_t -> phi_t+1 with the
// velocity of the previous collision, the collision of f_t with that force
// (giving rho_t, u_t), and the propagation.  Here one block owns a TY x TZ
// tile of (y, z) columns and marches along x (the LDS tiling of fe_tile.h);
// for each site of the current plane it forms F and the new phi from the
// tile, then collides the site's 19 pulled distributions with that F
// (lb_site.h).  The force never goes to memory.  The distributions and the
// velocities of the current plane are requested before the current
// plane's Q values are computed.
//
// Traffic per site (FP64): f 19 read + 19 write, phi read + write, u of the
// previous step read, u and rho written: 376 B (against 448 B for three
// separate passes).  Buffers: phi and u are double-buffered so that no block
// reads a value another block has already replaced.  Arithmetic is that of
// fe_ch.hip and lb_collide.hip, so FP64 results are bit-identical to the
// reference.

#include <hip/hip_runtime.h>

#include "fe_tile.h"
#include "kernels.h"
#include "lb_site.h"

namespace ldw {

namespace {

using namespace tile;

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

template <typename T, int ORDER, int METHOD, bool GZ>
__global__ void __launch_bounds__(NT)
binary_step_kernel(Geom g, FeParams fe, LbParams prm, int xchunk, int nbz, int nby, int nbx, int remap,
		   const T * __restrict__ phi, T * __restrict__ phi_out,
		   const T * __restrict__ u_prev, T * __restrict__ u_next,
		   T * __restrict__ rho, const T * __restrict__ g_in,
		   T * __restrict__ g_out, T fz0, T fz1, T fz2) {

  __shared__ Smem<T> s;

  int bzi, byi, bxi;
  if (!tile_of_block(nbz, nby, nbx, remap, bzi, byi, bxi)) return;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int z0 = bzi*TZ, y0 = byi*TY;
  const int xb = bxi*xchunk;
  const int xe = min(xb + xchunk, g.nx);
  const int y = y0 + ty, z = z0 + tz;
  const bool active = (y < g.ny) && (z < g.nz);

  for (int x = xb - 2; x <= xb + 1; x++) load_phi_plane(s, phi, g, x, y0, z0);
  T pre[NPRE];
  prefetch_phi_plane(pre, phi, g, xb + 2, y0, z0);
  __syncthreads();
  compute_q_plane(s, fe, xb - 1);
  compute_q_plane(s, fe, xb);

  for (int x = xb; x < xe; x++) {
    store_phi_plane(s, pre, x + 2);
    if (x + 1 < xe) prefetch_phi_plane(pre, phi, g, x + 3, y0, z0);
    __syncthreads();
    // the site's velocities and pulled distributions are requested before
    // the Q plane is computed, so their latency overlaps that work
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
      force_ch_site<T, ORDER, METHOD, true>(s, fe, x, ty, tz, uc, F, pnew);
      phi_out[site] = pnew;
      // site force: force_global + (hydro_f_zero value + phi force)
      // (collision.c:323-327, phi_force_colloid.c:457-462)
      T fs[3];
      fs[0] = T(prm.fglobal[0]) + (fz0 + F[0]);
      fs[1] = T(prm.fglobal[1]) + (fz1 + F[1]);
      fs[2] = T(prm.fglobal[2]) + (fz2 + F[2]);
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

template <typename T>
void launch(const Geom & g, const FeParams & fe, const LbParams & prm, const T * phi,
	    T * phi_out, const T * u_prev, T * u_next, T * rho, const T * g_in,
	    T * g_out, const double fzero[3], hipStream_t st) {
  const int bz = (g.nz + TZ - 1)/TZ;
  const int by = (g.ny + TY - 1)/TY;
  int xchunk = g.nx;
  const int want = tune_min_blocks();
  while (xchunk > 8 && (int64_t) bz*by*((g.nx + xchunk - 1)/xchunk) < want) xchunk = (xchunk + 1)/2;
  const int bx = (g.nx + xchunk - 1)/xchunk;
  const int ntile = bz*by*bx;
  dim3 grid(((ntile + NXCD - 1)/NXCD)*NXCD), block(TZ, TY, 1);
  const T f0 = T(fzero[0]), f1 = T(fzero[1]), f2 = T(fzero[2]);
  const bool gz = (prm.relax == 0);
#define L_(O, M, G) hipLaunchKernelGGL((binary_step_kernel<T, O, M, G>), grid, block, 0, st, \
				       g, fe, prm, xchunk, bz, by, bx, tune_remap(), phi, phi_out, u_prev, u_next, rho, \
				       g_in, g_out, f0, f1, f2)
#define L_G(O, M) do { if (gz) L_(O, M, true); else L_(O, M, false); } while (0)
#define L_M(O) do { if (fe.force_method == 0) L_G(O, 0); else L_G(O, 1); } while (0)
  if (fe.order == 1) L_M(1);
  else if (fe.order == 2) L_M(2);
  else L_M(3);
#undef L_M
#undef L_G
#undef L_
}

}  // namespace

void launch_binary_step_f64(const Geom & g, const FeParams & fe, const LbParams & prm,
			    const double * phi, double * phi_out, const double * u_prev,
			    double * u_next, double * rho, const double * g_in,
			    double * g_out, const double fzero[3], hipStream_t s) {
  launch<double>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, s);
}

void launch_binary_step_f32(const Geom & g, const FeParams & fe, const LbParams & prm,
			    const float * phi, float * phi_out, const float * u_prev,
			    float * u_next, float * rho, const float * g_in,
			    float * g_out, const double fzero[3], hipStream_t s) {
  launch<float>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, s);
}

}  // namespace ldw
