// fe_tile.h -- LDS tiling of phi and of the stencil quantities Q for the
// force + Cahn-Hilliard passes (fe_ch_fused.hip, binary_step.hip).
//
// A block owns TY x TZ (y, z) columns and marches along x.  LDS holds a ring
// of 5 phi planes with a halo of 2 in y and z and a ring of 3 planes of
// Q = (d_x phi, d_y phi, d_z phi, p0, mu) with a halo of 1.  Arithmetic is
// expression-for-expression that of fe_ch.hip (the reference's order).

#ifndef LDW_FE_TILE_H
#define LDW_FE_TILE_H

#include <hip/hip_runtime.h>

#include "advect.h"
#include "ch_conserve.h"
#include "grad27.h"
#include "kernels.h"

namespace ldw {
namespace tile {

// Tile shape: TZ x TY sites of a plane, one thread per site (block
// dimensions (TZ, TY)).  The single-lattice step chooses 64 x 4, 64 x 8 or
// 128 x 4 by precision and lattice size (tune_step_tile); the staged
// force/Cahn-Hilliard kernel uses 64 x 4; the replica ensemble chooses among
// several shapes at run time (ensemble.hip).
template <int TZ_, int TY_>
struct Shape {
  static constexpr int TZ = TZ_, TY = TY_;
  static constexpr int NT = TZ*TY;
  static constexpr int PZ = TZ + 4, PY = TY + 4;      // phi tile with halo 2
  static constexpr int QZ = TZ + 2, QY = TY + 2;      // Q tile with halo 1
  static constexpr int PPLANE = PZ*PY;
  static constexpr int QPLANE = QZ*QY;
  static constexpr int NPRE = (PPLANE + NT - 1)/NT;   // phi values prefetched per thread
};

using Shape64x4 = Shape<64, 4>;

// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs (each
// with its own L2), so physical block k runs on XCD k % 8.  Tiles are
// numbered z-fastest, then y, then x-chunk.  With remap = 1 XCD g is given
// the contiguous range of tiles [g*C, (g+1)*C), so tiles that share halo rows
// and edge cache lines run close together on one L2: that removes about half
// of the excess L2 fetches, but on MI355X it measured 7% slower at 256^3 than
// the linear order (remap = 0, the default), whose duplicate edge lines are
// served by the memory-side Infinity Cache.  Returns false for the padding
// blocks of the last range.
constexpr int NXCD = 8;

__device__ __forceinline__ bool tile_of_block(int nbz, int nby, int nbx, int remap,
					      int & bzi, int & byi, int & bxi) {
  const int ntile = nbz*nby*nbx;
  const int per = (ntile + NXCD - 1)/NXCD;
  const int k = blockIdx.x;
  const int t = remap ? (k % NXCD)*per + k/NXCD : k;
  if (t >= ntile) return false;
  bzi = t % nbz;
  byi = (t/nbz) % nby;
  bxi = t/(nbz*nby);
  return true;
}

__device__ __forceinline__ int modn(int i, int n) {
  int r = i % n;
  return (r < 0) ? r + n : r;
}

template <typename T, class S>
struct Smem {
  T phi[5][S::PPLANE];
  T q[3][5][S::QPLANE];       // g0, g1, g2, p0, mu
};

// Software pipelining: the phi plane needed by the next iteration is
// fetched into registers one iteration ahead, so its memory latency overlaps
// the LDS compute of the current plane.
template <typename T, class S>
__device__ __forceinline__ void prefetch_phi_plane(T (&pre)[S::NPRE], const T * __restrict__ phi,
						   const Geom & g, int x, int y0, int z0) {
  const int xw = modn(x, g.nx);
#pragma unroll
  for (int k = 0; k < S::NPRE; k++) {
    const int i = threadIdx.x + threadIdx.y*S::TZ + k*S::NT;
    if (i < S::PPLANE) {
      const int yy = i / S::PZ, zz = i - yy*S::PZ;
      const int y = modn(y0 - 2 + yy, g.ny);
      const int z = modn(z0 - 2 + zz, g.nz);
      pre[k] = phi[((int64_t) xw*g.ny + y)*g.nz + z];
    }
  }
}

template <typename T, class S>
__device__ __forceinline__ void store_phi_plane(Smem<T, S> & s, const T (&pre)[S::NPRE], int x) {
  const int slot = modn(x, 5);
#pragma unroll
  for (int k = 0; k < S::NPRE; k++) {
    const int i = threadIdx.x + threadIdx.y*S::TZ + k*S::NT;
    if (i < S::PPLANE) s.phi[slot][i] = pre[k];
  }
}

// The prologue's phi planes x0 .. x0+N-1 into the ring.  Every load is
// issued before the first LDS store, so the N planes cost one memory
// latency (a load-then-store loop per plane waits for each load in turn).
template <typename T, class S, int N>
__device__ __forceinline__ void load_phi_planes(Smem<T, S> & s, const T * __restrict__ phi,
						const Geom & g, int x0, int y0, int z0) {
  T p[N][S::NPRE];
#pragma unroll
  for (int k = 0; k < N; k++) prefetch_phi_plane<T, S>(p[k], phi, g, x0 + k, y0, z0);
#pragma unroll
  for (int k = 0; k < N; k++) store_phi_plane(s, p[k], x0 + k);
}

// velocity components the fluxes of site (x, y, z) need:
// ux(x), ux(x-1), ux(x+1), uy(y), uy(y+1), uy(y-1), uz(z), uz(z+1), uz(z-1)
template <typename T>
__device__ __forceinline__ void load_u(T (&v)[9], const T * __restrict__ u, const Geom & g,
				       int x, int y, int z) {
  const int xm = (x == 0) ? g.nx - 1 : x - 1;
  const int xp = (x == g.nx - 1) ? 0 : x + 1;
  const int ym = (y == 0) ? g.ny - 1 : y - 1;
  const int yp = (y == g.ny - 1) ? 0 : y + 1;
  const int zm = (z == 0) ? g.nz - 1 : z - 1;
  const int zp = (z == g.nz - 1) ? 0 : z + 1;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t rowx = (int64_t) x*nyz;
  const int64_t site = rowx + (int64_t) y*g.nz + z;
  const T * __restrict__ ux = u;
  const T * __restrict__ uy = u + g.nsite;
  const T * __restrict__ uz = u + 2*g.nsite;
  v[0] = ux[site];
  v[1] = ux[(int64_t) xm*nyz + (int64_t) y*g.nz + z];
  v[2] = ux[(int64_t) xp*nyz + (int64_t) y*g.nz + z];
  v[3] = uy[site];
  v[4] = uy[rowx + (int64_t) yp*g.nz + z];
  v[5] = uy[rowx + (int64_t) ym*g.nz + z];
  v[6] = uz[site];
  v[7] = uz[rowx + (int64_t) y*g.nz + zp];
  v[8] = uz[rowx + (int64_t) y*g.nz + zm];
}

// Q on plane x at every point of the halo-1 tile
template <typename T, class S>
__device__ __forceinline__ void compute_q_plane(Smem<T, S> & s, const FeParams & fe, int x) {
  constexpr int PZ = S::PZ, QZ = S::QZ;
  const int sm = modn(x - 1, 5), s0 = modn(x, 5), sp = modn(x + 1, 5);
  const int qs = modn(x, 3);
  const T a = T(fe.a), b = T(fe.b), kappa = T(fe.kappa);
  for (int i = threadIdx.x + threadIdx.y*S::TZ; i < S::QPLANE; i += S::NT) {
    const int yy = i / QZ, zz = i - yy*QZ;
    const int c = (yy + 1)*PZ + (zz + 1);         // same point in the phi tile
    const T phi = s.phi[s0][c];
    T g0, g1, g2, delsq;
    if (fe.stencil == 27) {
      const int sl[3] = {sm, s0, sp};
      auto p = [&](int i, int j, int k) { return s.phi[sl[i]][c + (j - 1)*PZ + (k - 1)]; };
      T gg[3];
      grad27<T>(p, gg, delsq);
      g0 = gg[0]; g1 = gg[1]; g2 = gg[2];
    }
    else {
      const T xp = s.phi[sp][c], xm = s.phi[sm][c];
      const T yp = s.phi[s0][c + PZ], ym = s.phi[s0][c - PZ];
      const T zp = s.phi[s0][c + 1], zm = s.phi[s0][c - 1];
      g0 = T(0.5)*(xp - xm);
      g1 = T(0.5)*(yp - ym);
      g2 = T(0.5)*(zp - zm);
      delsq = xp + xm + yp + ym + zp + zm - T(6.0)*phi;
    }
    const T p0 = T(0.5)*a*phi*phi + T(0.75)*b*phi*phi*phi*phi - kappa*phi*delsq
      - T(0.5)*kappa*(g0*g0 + g1*g1 + g2*g2);
    const T mu = a*phi + b*phi*phi*phi - kappa*delsq;
    s.q[qs][0][i] = g0;
    s.q[qs][1][i] = g1;
    s.q[qs][2][i] = g2;
    s.q[qs][3][i] = p0;
    s.q[qs][4][i] = mu;
  }
}

// P_ab at a Q point (symmetric.c:410-413)
template <typename T, int QPLANE>
__device__ __forceinline__ T pab_q(const T (&q)[5][QPLANE], int i, T kappa, int a, int b) {
  const T s = kappa*q[a][i]*q[b][i];
  return (a == b) ? q[3][i] + s : s;
}

template <typename T>
__device__ __forceinline__ T adv1(T uf, T phi_self, T phi_other, bool west) {
  if (west) return uf*((uf > T(0.0)) ? phi_other : phi_self);
  return uf*((uf < T(0.0)) ? phi_other : phi_self);
}


// Force and Cahn-Hilliard update at site (x, y, z) of the output plane x,
// from the LDS rings (Q planes x-1, x, x+1; phi planes x-2 .. x+2) and the
// velocities uc of load_u().  Returns the site's hydrodynamic force
// F = fz + F_phi (fz: the force before this calculation, the
// hydro_f_zero value or the stored force), plus -phi grad mu_ext as a
// second addition under phi grad mu with an external gradient, and the
// updated phi; with csum_site (the site's compensation) the update is the
// compensated one.  Advection orders 1-4 (order 4 reads phi to x +- 2).
// GMU: the external gradient terms are compiled in (they act where
// fe.gmu_on); without them the kernels that never see one keep their
// register count.
template <typename T, class S, int ORDER, int METHOD, bool HAVE_U, bool GMU = true>
__device__ __forceinline__ void force_ch_site(const Smem<T, S> & s, const FeParams & fe,
					      int x, int ty, int tz, const T (&uc)[9], const T (&fz)[3],
					      T (&F)[3], T & pnew, T * csum_site = nullptr) {
  constexpr int PZ = S::PZ, QZ = S::QZ, PPLANE = S::PPLANE, QPLANE = S::QPLANE;
  const T kappa = T(fe.kappa);
  const T mob = T(fe.mobility);
  const T a1 = T(-0.213933), a2 = T(0.927865), a3 = T(0.286067);
  const T (&qm)[5][QPLANE] = s.q[modn(x - 1, 3)];
  const T (&q0)[5][QPLANE] = s.q[modn(x, 3)];
  const T (&qp)[5][QPLANE] = s.q[modn(x + 1, 3)];
  const int c = (ty + 1)*QZ + (tz + 1);      // site in the Q tile

  if constexpr (METHOD == 0) {
#pragma unroll
	for (int a = 0; a < 3; a++) {
	  T s0 = pab_q(q0, c, kappa, a, 0);
	  T fa = T(-0.5)*(pab_q(qp, c, kappa, a, 0) + s0);
	  fa += T(0.5)*(pab_q(qm, c, kappa, a, 0) + s0);
	  s0 = pab_q(q0, c, kappa, a, 1);
	  fa -= T(0.5)*(pab_q(q0, c + QZ, kappa, a, 1) + s0);
	  fa += T(0.5)*(pab_q(q0, c - QZ, kappa, a, 1) + s0);
	  s0 = pab_q(q0, c, kappa, a, 2);
	  fa -= T(0.5)*(pab_q(q0, c + 1, kappa, a, 2) + s0);
	  fa += T(0.5)*(pab_q(q0, c - 1, kappa, a, 2) + s0);
	  F[a] = fz[a] + fa;
	}
  }
  else {
	const T phi0 = s.phi[modn(x, 5)][(ty + 2)*PZ + (tz + 2)];
	F[0] = fz[0] + (T(0.0) + -phi0*T(0.5)*(qp[4][c] - qm[4][c]));
	F[1] = fz[1] + (T(0.0) + -phi0*T(0.5)*(q0[4][c + QZ] - q0[4][c - QZ]));
	F[2] = fz[2] + (T(0.0) + -phi0*T(0.5)*(q0[4][c + 1] - q0[4][c - 1]));
	if (GMU && fe.gmu_on) {
	  F[0] += -phi0*T(fe.gmu[0]);
	  F[1] += -phi0*T(fe.gmu[1]);
	  F[2] += -phi0*T(fe.gmu[2]);
	}
  }
  // Cahn-Hilliard fluxes: phi from the ring, mu from Q
  const T (&P0)[PPLANE] = s.phi[modn(x, 5)];
  const int cp = (ty + 2)*PZ + (tz + 2);     // site in the phi tile
  const T p0 = P0[cp];
  T fw = T(0), fe_ = T(0), fyp = T(0), fym = T(0), fzp = T(0), fzm = T(0);
  if constexpr (HAVE_U) {
	const T ux0 = uc[0], uxm = uc[1], uxp = uc[2];
	const T uy0 = uc[3], uyp = uc[4], uym = uc[5];
	const T uz0 = uc[6], uzp = uc[7], uzm = uc[8];
	const T uw = T(0.5)*(ux0 + uxm);
	const T ue = T(0.5)*(ux0 + uxp);
	const T uyf = T(0.5)*(uy0 + uyp);
	const T uyb = T(0.5)*(uym + uy0);
	const T uzf = T(0.5)*(uz0 + uzp);
	const T uzb = T(0.5)*(uzm + uz0);
	const T (&PM)[PPLANE] = s.phi[modn(x - 1, 5)];
	const T (&PP)[PPLANE] = s.phi[modn(x + 1, 5)];
	if constexpr (ORDER == 1) {
	  fw = adv1(uw, p0, PM[cp], true);
	  fe_ = adv1(ue, p0, PP[cp], false);
	  fyp = adv1(uyf, p0, P0[cp + PZ], false);
	  fym = adv1(uyb, P0[cp - PZ], p0, false);
	  fzp = adv1(uzf, p0, P0[cp + 1], false);
	  fzm = adv1(uzb, P0[cp - 1], p0, false);
	}
	else if constexpr (ORDER == 2) {
	  fw = uw*T(0.5)*(PM[cp] + p0);
	  fe_ = ue*T(0.5)*(p0 + PP[cp]);
	  fyp = uyf*T(0.5)*(p0 + P0[cp + PZ]);
	  fym = uyb*T(0.5)*(P0[cp - PZ] + p0);
	  fzp = uzf*T(0.5)*(p0 + P0[cp + 1]);
	  fzm = uzb*T(0.5)*(P0[cp - 1] + p0);
	}
	else if constexpr (ORDER == 4) {
	  const T (&PM2)[PPLANE] = s.phi[modn(x - 2, 5)];
	  const T (&PP2)[PPLANE] = s.phi[modn(x + 2, 5)];
	  fw = adv::order4(uw, PM2[cp], PM[cp], p0, PP[cp]);
	  fe_ = adv::order4(ue, PM[cp], p0, PP[cp], PP2[cp]);
	  fyp = adv::order4(uyf, P0[cp - PZ], p0, P0[cp + PZ], P0[cp + 2*PZ]);
	  fym = adv::order4(uyb, P0[cp - 2*PZ], P0[cp - PZ], p0, P0[cp + PZ]);
	  fzp = adv::order4(uzf, P0[cp - 1], p0, P0[cp + 1], P0[cp + 2]);
	  fzm = adv::order4(uzb, P0[cp - 2], P0[cp - 1], p0, P0[cp + 1]);
	}
	else {
	  static_assert(ORDER == 3, "the tiles hold phi to +-2: advection orders 1-4");
	  const T (&PM2)[PPLANE] = s.phi[modn(x - 2, 5)];
	  const T (&PP2)[PPLANE] = s.phi[modn(x + 2, 5)];
	  T d1, d2, d3;
	  if (uw > T(0.0)) { d1 = PM2[cp]; d2 = PM[cp]; d3 = p0; }
	  else             { d1 = PP[cp]; d2 = p0; d3 = PM[cp]; }
	  fw = uw*(a1*d1 + a2*d2 + a3*d3);
	  if (ue < T(0.0)) { d1 = PP2[cp]; d2 = PP[cp]; d3 = p0; }
	  else             { d1 = PM[cp]; d2 = p0; d3 = PP[cp]; }
	  fe_ = ue*(a1*d1 + a2*d2 + a3*d3);
	  if (uyf < T(0.0)) { d1 = P0[cp + 2*PZ]; d2 = P0[cp + PZ]; d3 = p0; }
	  else              { d1 = P0[cp - PZ]; d2 = p0; d3 = P0[cp + PZ]; }
	  fyp = uyf*(a1*d1 + a2*d2 + a3*d3);
	  if (uyb < T(0.0)) { d1 = P0[cp + PZ]; d2 = p0; d3 = P0[cp - PZ]; }
	  else              { d1 = P0[cp - 2*PZ]; d2 = P0[cp - PZ]; d3 = p0; }
	  fym = uyb*(a1*d1 + a2*d2 + a3*d3);
	  if (uzf < T(0.0)) { d1 = P0[cp + 2]; d2 = P0[cp + 1]; d3 = p0; }
	  else              { d1 = P0[cp - 1]; d2 = p0; d3 = P0[cp + 1]; }
	  fzp = uzf*(a1*d1 + a2*d2 + a3*d3);
	  if (uzb < T(0.0)) { d1 = P0[cp + 1]; d2 = p0; d3 = P0[cp - 1]; }
	  else              { d1 = P0[cp - 2]; d2 = P0[cp - 1]; d3 = p0; }
	  fzm = uzb*(a1*d1 + a2*d2 + a3*d3);
	}
  }
  {
	const T m0 = q0[4][c];
	fw -= mob*(m0 - qm[4][c]);
	fe_ -= mob*(qp[4][c] - m0);
	fyp -= mob*(q0[4][c + QZ] - m0);
	fym -= mob*(m0 - q0[4][c - QZ]);
	fzp -= mob*(q0[4][c + 1] - m0);
	fzm -= mob*(m0 - q0[4][c - 1]);
  }
  if (GMU && fe.gmu_on) adv::mu_ext(mob, fe.gmu, fw, fe_, fyp, fym, fzp, fzm);
  pnew = p0;
  if (csum_site != nullptr) {
    T cs = *csum_site;
    ch_update_compensated(pnew, cs, fe_, fw, fyp, fym, fzp, fzm);
    *csum_site = cs;
  }
  else {
    pnew -= (+ fe_ - fw + fyp - fym + fzp - fzm);
  }
}

}  // namespace tile
}  // namespace ldw

#endif
