// le_step_tile.h -- the single-pass binary-fluid step (step_tile.h) for a
// lattice with Lees-Edwards planes, FP64.
//
// The lattice is cut into slabs between consecutive planes.  A block
// marches through x-planes of one slab only; its LDS phi ring holds, for
// planes beyond the slab's edges, the buffer planes the reference's
// stencils read across a plane (lees_edw_ic_to_buff): the planes the
// reference displaces and interpolates for that plane (field.c:459-500),
// prepared by le_phi_buffer_kernel before the step.  So every gradient,
// chemical potential, stress and flux a block forms from its rings is the
// one the reference forms with its buffer index mapping, at real sites and
// at the buffer sites next to a plane (gradient_3d_7pt_fluid.c:330-430).
// Velocities across a plane come from the interpolated velocity buffer
// (hydro.c:350-437).
//
// With planes the reference takes the flux form of the force at every site
// (phi_force.c:89-97, 289-520), which rounds differently from the
// stress-divergence form of the plane-free kernel, so the force here is
// the flux form.  The two x-slices at a slab's edges need values formed
// across the whole plane first: the per-plane force correction
// (phi_force_flux_fix_local) and the averaged Cahn-Hilliard face fluxes
// (phi_ch_le_fix_fluxes).  They run in a second launch, after those are
// known; everything else runs in the first.

#ifndef LDW_LE_STEP_TILE_H
#define LDW_LE_STEP_TILE_H

#include "le_lb.h"
#include "step_tile.h"

namespace ldw {
namespace tile {

// One slab: planes lo (below) and hi (above); x is unwrapped, real planes
// slab_lo .. slab_hi (0-based, slab_hi may exceed nx for the slab that
// wraps round x = 0).  Buffer planes of plane p start at 2 p nh.
struct LeSlab {
  const double * phib;        // buffer phi planes
  const double * ub;          // buffer velocity planes, component stride nbuf
  int64_t nbuf;
  int nh;
  int lo, hi;
  int slab_lo, slab_hi;
  // the edge slices' plane values (second launch only; null otherwise)
  const double * fcor;        // [plane][component] force correction sums
  double ra;                  // 1/2 (L_y L_z)
  const double * chface;      // raw Cahn-Hilliard x faces [plane][side][nyz]
  int jw, je;                 // integer and fractional split of +dy and -dy
  double frw, fre;
  double * sbuf;              // LB plane reprojections [plane][side][5][nyz]
  double duy0, duy1;          // their velocity jump, side 0 and side 1
};

// where phi plane x (unwrapped) of the slab's ring comes from
__device__ __forceinline__ const double * le_plane_src(const LeSlab & L, const double * phi,
						       const Geom & g, int x) {
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (x < L.slab_lo) return L.phib + (int64_t) (2*L.lo*L.nh + L.nh - (L.slab_lo - x))*nyz;
  if (x > L.slab_hi) return L.phib + (int64_t) (2*L.hi*L.nh + L.nh + (x - L.slab_hi - 1))*nyz;
  return phi + (int64_t) modn(x, g.nx)*nyz;
}

template <class S>
__device__ __forceinline__ void le_prefetch_phi_plane(double (&pre)[S::NPRE], const double * __restrict__ pl,
						      const Geom & g, int y0, int z0) {
#pragma unroll
  for (int k = 0; k < S::NPRE; k++) {
    const int i = threadIdx.x + threadIdx.y*S::TZ + k*S::NT;
    if (i < S::PPLANE) {
      const int yy = i / S::PZ, zz = i - yy*S::PZ;
      const int y = modn(y0 - 2 + yy, g.ny);
      const int z = modn(z0 - 2 + zz, g.nz);
      pre[k] = pl[(int64_t) y*g.nz + z];
    }
  }
}

// the prologue's four phi planes x0 .. x0+3 (real or buffer planes), every
// load issued before the first LDS store (load_phi_planes)
template <class S>
__device__ __forceinline__ void le_load_phi_planes4(Smem<double, S> & s, const LeSlab & L,
						    const double * __restrict__ phi, const Geom & g,
						    int x0, int y0, int z0) {
  double p[4][S::NPRE];
#pragma unroll
  for (int k = 0; k < 4; k++) le_prefetch_phi_plane<S>(p[k], le_plane_src(L, phi, g, x0 + k), g, y0, z0);
#pragma unroll
  for (int k = 0; k < 4; k++) store_phi_plane(s, p[k], x0 + k);
}

// velocities of load_u(); at the slab's edges u_x of the neighbour across
// the plane is the buffer's
__device__ __forceinline__ void le_load_u(double (&v)[9], const double * __restrict__ u,
					  const LeSlab & L, const Geom & g, int x, int y, int z) {
  load_u<double>(v, u, g, modn(x, g.nx), y, z);
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t s = (int64_t) y*g.nz + z;
  if (x == L.slab_lo) v[1] = L.ub[(int64_t) (2*L.lo*L.nh + L.nh - 1)*nyz + s];
  if (x == L.slab_hi) v[2] = L.ub[(int64_t) (2*L.hi*L.nh + L.nh)*nyz + s];
}

// P_ab at a Q point as fe_symm_str writes it: p0 d_ab + kappa g_a g_b
template <int QPLANE>
__device__ __forceinline__ double pab_str(const double (&q)[5][QPLANE], int i, double kappa, int a, int b) {
  const double dab = (a == b) ? 1.0 : 0.0;
  return q[3][i]*dab + kappa*q[a][i]*q[b][i];
}

// Flux-form force and the Cahn-Hilliard update at site (x, y, z) (x
// unwrapped; y, z lattice coordinates for the face buffers), from the
// rings as force_ch_site() reads them.  GMU: the external chemical
// potential gradient's flux term is compiled in (fe.gmu_on then decides).
template <class S, int ORDER, bool GMU>
__device__ __forceinline__ void le_force_ch_site(const Smem<double, S> & s, const FeParams & fe,
						 const LeSlab & L, const Geom & g, int x, int y, int z,
						 int ty, int tz, const double (&uc)[9], double (&F)[3],
						 double & pnew, double * csum_site) {
  constexpr int PZ = S::PZ, QZ = S::QZ, PPLANE = S::PPLANE, QPLANE = S::QPLANE;
  const double kappa = fe.kappa;
  const double mob = fe.mobility;
  const double a1 = -0.213933, a2 = 0.927865, a3 = 0.286067;
  const double (&qm)[5][QPLANE] = s.q[modn(x - 1, 3)];
  const double (&q0)[5][QPLANE] = s.q[modn(x, 3)];
  const double (&qp)[5][QPLANE] = s.q[modn(x + 1, 3)];
  const int c = (ty + 1)*QZ + (tz + 1);

  // phi_force_compute_fluxes / _fix_local / _divergence (phi_force.c:340-520)
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double s0x = pab_str(q0, c, kappa, a, 0);
    double fe_ = 0.5*(pab_str(qp, c, kappa, a, 0) + s0x);
    double fw = 0.5*(pab_str(qm, c, kappa, a, 0) + s0x);
    const double s0y = pab_str(q0, c, kappa, a, 1);
    const double fy = 0.5*(pab_str(q0, c + QZ, kappa, a, 1) + s0y);
    const double fym = 0.5*(s0y + pab_str(q0, c - QZ, kappa, a, 1));
    const double s0z = pab_str(q0, c, kappa, a, 2);
    const double fz = 0.5*(pab_str(q0, c + 1, kappa, a, 2) + s0z);
    const double fzm = 0.5*(s0z + pab_str(q0, c - 1, kappa, a, 2));
    if (L.fcor != nullptr) {
      if (x == L.slab_hi) fe_ += L.ra*L.fcor[3*L.hi + a];
      if (x == L.slab_lo) fw -= L.ra*L.fcor[3*L.lo + a];
    }
    F[a] = -(+ fe_ - fw + fy - fym + fz - fzm);
  }

  // Cahn-Hilliard fluxes as force_ch_site() forms them
  const double (&P0)[PPLANE] = s.phi[modn(x, 5)];
  const int cp = (ty + 2)*PZ + (tz + 2);
  const double p0 = P0[cp];
  double fw, fe_, fyp, fym, fzp, fzm;
  {
    const double ux0 = uc[0], uxm = uc[1], uxp = uc[2];
    const double uy0 = uc[3], uyp = uc[4], uym = uc[5];
    const double uz0 = uc[6], uzp = uc[7], uzm = uc[8];
    const double uw = 0.5*(ux0 + uxm);
    const double ue = 0.5*(ux0 + uxp);
    const double uyf = 0.5*(uy0 + uyp);
    const double uyb = 0.5*(uym + uy0);
    const double uzf = 0.5*(uz0 + uzp);
    const double uzb = 0.5*(uzm + uz0);
    const double (&PM)[PPLANE] = s.phi[modn(x - 1, 5)];
    const double (&PP)[PPLANE] = s.phi[modn(x + 1, 5)];
    if constexpr (ORDER == 1) {
      fw = adv1(uw, p0, PM[cp], true);
      fe_ = adv1(ue, p0, PP[cp], false);
      fyp = adv1(uyf, p0, P0[cp + PZ], false);
      fym = adv1(uyb, P0[cp - PZ], p0, false);
      fzp = adv1(uzf, p0, P0[cp + 1], false);
      fzm = adv1(uzb, P0[cp - 1], p0, false);
    }
    else if constexpr (ORDER == 2) {
      fw = uw*0.5*(PM[cp] + p0);
      fe_ = ue*0.5*(p0 + PP[cp]);
      fyp = uyf*0.5*(p0 + P0[cp + PZ]);
      fym = uyb*0.5*(P0[cp - PZ] + p0);
      fzp = uzf*0.5*(p0 + P0[cp + 1]);
      fzm = uzb*0.5*(P0[cp - 1] + p0);
    }
    else if constexpr (ORDER == 4) {
      const double (&PM2)[PPLANE] = s.phi[modn(x - 2, 5)];
      const double (&PP2)[PPLANE] = s.phi[modn(x + 2, 5)];
      fw = adv::order4(uw, PM2[cp], PM[cp], p0, PP[cp]);
      fe_ = adv::order4(ue, PM[cp], p0, PP[cp], PP2[cp]);
      fyp = adv::order4(uyf, P0[cp - PZ], p0, P0[cp + PZ], P0[cp + 2*PZ]);
      fym = adv::order4(uyb, P0[cp - 2*PZ], P0[cp - PZ], p0, P0[cp + PZ]);
      fzp = adv::order4(uzf, P0[cp - 1], p0, P0[cp + 1], P0[cp + 2]);
      fzm = adv::order4(uzb, P0[cp - 2], P0[cp - 1], p0, P0[cp + 1]);
    }
    else {
      static_assert(ORDER == 3, "the tiles hold phi to +-2: advection orders 1-4");
      const double (&PM2)[PPLANE] = s.phi[modn(x - 2, 5)];
      const double (&PP2)[PPLANE] = s.phi[modn(x + 2, 5)];
      double d1, d2, d3;
      if (uw > 0.0) { d1 = PM2[cp]; d2 = PM[cp]; d3 = p0; }
      else          { d1 = PP[cp]; d2 = p0; d3 = PM[cp]; }
      fw = uw*(a1*d1 + a2*d2 + a3*d3);
      if (ue < 0.0) { d1 = PP2[cp]; d2 = PP[cp]; d3 = p0; }
      else          { d1 = PM[cp]; d2 = p0; d3 = PP[cp]; }
      fe_ = ue*(a1*d1 + a2*d2 + a3*d3);
      if (uyf < 0.0) { d1 = P0[cp + 2*PZ]; d2 = P0[cp + PZ]; d3 = p0; }
      else           { d1 = P0[cp - PZ]; d2 = p0; d3 = P0[cp + PZ]; }
      fyp = uyf*(a1*d1 + a2*d2 + a3*d3);
      if (uyb < 0.0) { d1 = P0[cp + PZ]; d2 = p0; d3 = P0[cp - PZ]; }
      else           { d1 = P0[cp - 2*PZ]; d2 = P0[cp - PZ]; d3 = p0; }
      fym = uyb*(a1*d1 + a2*d2 + a3*d3);
      if (uzf < 0.0) { d1 = P0[cp + 2]; d2 = P0[cp + 1]; d3 = p0; }
      else           { d1 = P0[cp - 1]; d2 = p0; d3 = P0[cp + 1]; }
      fzp = uzf*(a1*d1 + a2*d2 + a3*d3);
      if (uzb < 0.0) { d1 = P0[cp + 1]; d2 = p0; d3 = P0[cp - 1]; }
      else           { d1 = P0[cp - 2]; d2 = P0[cp - 1]; d3 = p0; }
      fzm = uzb*(a1*d1 + a2*d2 + a3*d3);
    }
  }
  {
    const double m0 = q0[4][c];
    fw -= mob*(m0 - qm[4][c]);
    fe_ -= mob*(qp[4][c] - m0);
    fyp -= mob*(q0[4][c + QZ] - m0);
    fym -= mob*(m0 - q0[4][c - QZ]);
    fzp -= mob*(q0[4][c + 1] - m0);
    fzm -= mob*(m0 - q0[4][c - 1]);
  }
  if (GMU && fe.gmu_on) adv::mu_ext(mob, fe.gmu, fw, fe_, fyp, fym, fzp, fzm);
  // phi_ch_le_fix_fluxes (phi_cahn_hilliard.c:655-720) at the slab's edges
  if (L.chface != nullptr) {
    const int N = g.ny;
    const int64_t nyz = (int64_t) g.ny*g.nz;
    if (x == L.slab_hi) {
      const int j1 = (y - L.jw - 1 + 2*N) % N, j2 = (j1 + 1) % N;
      const double * fwf = L.chface + (int64_t) (2*L.hi + 1)*nyz;
      const double bw = fwf[(int64_t) j1*g.nz + z]*L.frw + fwf[(int64_t) j2*g.nz + z]*(1.0 - L.frw);
      fe_ = 0.5*(fe_ + bw);
    }
    if (x == L.slab_lo) {
      const int j1 = (y - L.je - 1 + 2*N) % N, j2 = (j1 + 1) % N;
      const double * fef = L.chface + (int64_t) (2*L.lo + 0)*nyz;
      const double be = fef[(int64_t) j1*g.nz + z]*L.fre + fef[(int64_t) j2*g.nz + z]*(1.0 - L.fre);
      fw = 0.5*(fw + be);
    }
  }
  pnew = p0;
  if (csum_site != nullptr) {
    double cs = *csum_site;
    ch_update_compensated(pnew, cs, fe_, fw, fyp, fym, fzp, fzm);
    *csum_site = cs;
  }
  else {
    pnew -= (+ fe_ - fw + fyp - fym + fzp - fzm);
  }
}

// step_tile() over the x-planes [xb, xe) of one slab (unwrapped x); the
// force is also stored (force_mode 0: fzero + F).  EDGE (the edge launch):
// every site is a slab edge, and the LB plane reprojection of its
// post-collision distributions goes to L.sbuf (x = slab_lo: side 1 of
// plane lo; x = slab_hi: side 0 of plane hi).
template <class S, int ORDER, bool GZ, bool GMU, bool EDGE = false>
__device__ __forceinline__ void le_step_tile(Smem<double, S> & s, const Geom & g, const FeParams & fe,
					     const LbParams & prm, const LeSlab & L, int y0, int z0,
					     int xb, int xe, const double * __restrict__ phi,
					     double * __restrict__ phi_out, const double * __restrict__ u_prev,
					     double * __restrict__ u_next, double * __restrict__ rho,
					     const double * __restrict__ g_in, double * __restrict__ g_out,
					     double * __restrict__ force, double * __restrict__ csum,
					     double fz0, double fz1, double fz2) {
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int y = y0 + ty, z = z0 + tz;
  const bool active = (y < g.ny) && (z < g.nz);

  le_load_phi_planes4<S>(s, L, phi, g, xb - 2, y0, z0);
  double pre[S::NPRE];
  le_prefetch_phi_plane<S>(pre, le_plane_src(L, phi, g, xb + 2), g, y0, z0);
  __syncthreads();
  compute_q_plane(s, fe, xb - 1);
  compute_q_plane(s, fe, xb);

  for (int x = xb; x < xe; x++) {
    store_phi_plane(s, pre, x + 2);
    if (x + 1 < xe) le_prefetch_phi_plane<S>(pre, le_plane_src(L, phi, g, x + 3), g, y0, z0);
    __syncthreads();
    const int xw = modn(x, g.nx);
    double uc[9];
    double fc[NVEL];
    if (active) {
      le_load_u(uc, u_prev, L, g, x, y, z);
      load_f(fc, g_in, g, xw, y, z);
    }
    compute_q_plane(s, fe, x + 1);
    __syncthreads();

    if (active) {
      const int64_t site = ((int64_t) xw*g.ny + y)*g.nz + z;
      double F[3], pnew;
      le_force_ch_site<S, ORDER, GMU>(s, fe, L, g, x, y, z, ty, tz, uc, F, pnew,
				 csum ? csum + site : nullptr);
      phi_out[site] = pnew;
      double fs[3];
      fs[0] = prm.fglobal[0] + (fz0 + F[0]);
      fs[1] = prm.fglobal[1] + (fz1 + F[1]);
      fs[2] = prm.fglobal[2] + (fz2 + F[2]);
      force[site] = fz0 + F[0];
      force[g.nsite + site] = fz1 + F[1];
      force[2*g.nsite + site] = fz2 + F[2];
      double r, uu[3];
      site::collide_site<double, GZ>(prm, fs, fc, r, uu);
      if constexpr (EDGE) {
	const int side = (x == L.slab_lo) ? 1 : 0;
	const int p = side ? L.lo : L.hi;
	double out[5];
	le_reproject_site(fc, side, side ? L.duy1 : L.duy0, out);
	const int64_t nyz = (int64_t) g.ny*g.nz;
	const int64_t rem = (int64_t) y*g.nz + z;
#pragma unroll
	for (int ip = 0; ip < 5; ip++) L.sbuf[((int64_t) (2*p + side)*5 + ip)*nyz + rem] = out[ip];
      }
      site::store_planes<double>(g_out, g.nsite, site, fc);
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
