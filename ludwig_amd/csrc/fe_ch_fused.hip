// fe_ch_fused.hip -- K23: thermodynamic force and Cahn-Hilliard update of
// the symmetric free energy in one pass over phi.
//
// One block owns a TY x TZ tile of (y, z) columns and marches along x.  LDS
// holds a ring of 5 phi planes with a halo of 2 in y and z, and a ring of 3
// planes of Q = (d_x phi, d_y phi, d_z phi, p0, mu) with a halo of 1, so
// every phi value is fetched from memory about once and every stencil
// quantity is evaluated once per point.  For output plane x the block
//   A. loads phi plane x+2,
//   B. evaluates Q on plane x+1 (7-point gradient and Laplacian,
//      gradient_3d_7pt_fluid.c:270-305; p0 and mu, symmetric.c:308-418),
//   C. forms at each site of plane x
//      - the force from the face-averaged stress divergence
//        (phi_force_colloid.c:352-466) or -phi grad mu (phi_grad_mu.c:301-333),
//      - the advective and diffusive face fluxes and the forward-step
//        update of phi (advection.c, phi_cahn_hilliard.c:350-413,1018-1052).
// Reads phi and u (previous collision), writes F and the new phi into the
// other phi buffer: 64 B/site in FP64 instead of 40 + 48 for K2 + K3.  The
// arithmetic is the same expression-for-expression as fe_ch.hip, so FP64
// results stay bit-identical to the reference.

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "fe_tile.h"

namespace ldw {

namespace {

using namespace tile;
using S = Shape64x4;

template <typename T, int ORDER, int METHOD, bool HAVE_U>
__global__ void __launch_bounds__(S::NT)
force_ch_kernel(Geom g, FeParams fe, int xchunk, int nbz, int nby, int nbx, int remap, const T * __restrict__ phi,
		const T * __restrict__ u, T * __restrict__ force, T * __restrict__ phi_out,
		T * __restrict__ csum, int force_mode, T fz0, T fz1, T fz2) {

  __shared__ Smem<T, S> s;

  int bzi, byi, bxi;
  if (!tile_of_block(nbz, nby, nbx, remap, bzi, byi, bxi)) return;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int z0 = bzi*S::TZ, y0 = byi*S::TY;
  const int xb = bxi*xchunk;
  const int xe = min(xb + xchunk, g.nx);
  const int y = y0 + ty, z = z0 + tz;
  const bool active = (y < g.ny) && (z < g.nz);

  // prologue: phi planes xb-2 .. xb+1, Q planes xb-1, xb; registers hold
  // phi plane xb+2 and the velocities of plane xb
  load_phi_planes<T, S, 4>(s, phi, g, xb - 2, y0, z0);
  T pre[S::NPRE];
  prefetch_phi_plane<T, S>(pre, phi, g, xb + 2, y0, z0);
  T uc[9];
  if constexpr (HAVE_U) {
    if (active) load_u(uc, u, g, xb, y, z);
  }
  __syncthreads();
  compute_q_plane(s, fe, xb - 1);
  compute_q_plane(s, fe, xb);

  for (int x = xb; x < xe; x++) {
    // A: phi plane x+2 from registers; fetch plane x+3 for the next round
    store_phi_plane(s, pre, x + 2);
    if (x + 1 < xe) prefetch_phi_plane<T, S>(pre, phi, g, x + 3, y0, z0);
    __syncthreads();
    // B: Q plane x+1
    compute_q_plane(s, fe, x + 1);
    __syncthreads();
    T un[9];
    if constexpr (HAVE_U) {
      if (active && x + 1 < xe) load_u(un, u, g, x + 1, y, z);
    }

    // C: outputs on plane x
    if (active) {
      const int64_t site = ((int64_t) x*g.ny + y)*g.nz + z;
      T F[3], pnew;
      // the force before this calculation: the hydro_f_zero value, or the
      // stored force (read-modify-write)
      T fzv[3] = {fz0, fz1, fz2};
      if (force_mode != 0) {
	fzv[0] = force[site]; fzv[1] = force[g.nsite + site]; fzv[2] = force[2*g.nsite + site];
      }
      force_ch_site<T, S, ORDER, METHOD, HAVE_U>(s, fe, x, ty, tz, uc, fzv, F, pnew,
						 csum ? csum + site : nullptr);
      force[site] = F[0];
      force[g.nsite + site] = F[1];
      force[2*g.nsite + site] = F[2];
      phi_out[site] = pnew;
    }
    __syncthreads();
    if constexpr (HAVE_U) {
#pragma unroll
      for (int k = 0; k < 9; k++) uc[k] = un[k];
    }
  }
}

template <typename T>
void launch(const Geom & g, const FeParams & fe, const T * phi, const T * u, T * force,
	    T * phi_out, T * csum, int force_mode, const double fzero[3], hipStream_t st) {
  const int bz = (g.nz + S::TZ - 1)/S::TZ;
  const int by = (g.ny + S::TY - 1)/S::TY;
  // enough blocks to fill 256 CUs several times; chunk of x planes per block
  int xchunk = g.nx;
  const int want = tune_min_blocks();
  while (xchunk > 8 && (int64_t) bz*by*((g.nx + xchunk - 1)/xchunk) < want) xchunk = (xchunk + 1)/2;
  const int bx = (g.nx + xchunk - 1)/xchunk;
  const int ntile = bz*by*bx;
  dim3 grid(((ntile + NXCD - 1)/NXCD)*NXCD), block(S::TZ, S::TY, 1);
  const T f0 = T(fzero[0]), f1 = T(fzero[1]), f2 = T(fzero[2]);
#define L_(O, M, U) hipLaunchKernelGGL((force_ch_kernel<T, O, M, U>), grid, block, 0, st, \
				       g, fe, xchunk, bz, by, bx, tune_remap(), phi, u, force, phi_out, csum, force_mode, \
				       f0, f1, f2)
#define L_U(O, M) do { if (u) L_(O, M, true); else L_(O, M, false); } while (0)
#define L_M(O) do { if (fe.force_method == 0) L_U(O, 0); else L_U(O, 1); } while (0)
  if (fe.order == 1) L_M(1);
  else if (fe.order == 2) L_M(2);
  else if (fe.order == 3) L_M(3);
  else L_M(4);
#undef L_M
#undef L_U
#undef L_
}

}  // namespace

void launch_force_ch_f64(const Geom & g, const FeParams & fe, const double * phi,
			 const double * u, double * force, double * phi_out, double * csum,
			 int force_mode, const double fzero[3], hipStream_t s) {
  launch<double>(g, fe, phi, u, force, phi_out, csum, force_mode, fzero, s);
}

void launch_force_ch_f32(const Geom & g, const FeParams & fe, const float * phi,
			 const float * u, float * force, float * phi_out, float * csum,
			 int force_mode, const double fzero[3], hipStream_t s) {
  launch<float>(g, fe, phi, u, force, phi_out, csum, force_mode, fzero, s);
}

}  // namespace ldw
