// lees_edwards.hip -- Lees-Edwards sliding periodic planes for the binary
// fluid (FP64).
//
// The reference keeps, next to the nx real x planes, nxbuf = 2 nh nplane
// "buffer" planes (leesedwards.c:430-455): copies of the nh real planes on
// either side of each plane, displaced in y by the plane's sliding offset
// and interpolated.  A stencil that crosses a plane reads the buffer
// (lees_edw_ic_to_buff, leesedwards.c:1030-1065).  Here the buffer planes
// are separate device arrays of nxbuf*ny*nz sites, periodic in y and z
// like the real lattice, and a "virtual" x index v < nx names a real
// plane, v >= nx buffer plane v - nx.
//
// Per time step (ludwig.c:528-860 order):
//   phi buffer   4-point Lagrange interpolation in y (field.c:459-500)
//   gradient     real sites with mapped x neighbours, and the buffer sites
//                next to each plane (gradient_3d_7pt_fluid.c:231-310,
//                330-430; gradient_3d_27pt_fluid.c:219-368, 400-560)
//   force        flux form, the only one the reference allows with planes
//                (phi_force.c:74-97, 289-520): x, y, z face fluxes of the
//                chemical stress, the per-plane correction that makes the
//                two sides of a plane carry the same total momentum flux,
//                then the divergence
//   u buffer     linear interpolation plus the velocity jump
//                (hydro.c:350-437)
//   Cahn-Hilliard advective and diffusive face fluxes with mapped x
//                neighbours (advection.c:538-632, 680-760, 946-1145;
//                phi_cahn_hilliard.c:350-413), the plane-face averaging of
//                phi_ch_le_fix_fluxes (phi_cahn_hilliard.c:613-730), and
//                the update (phi_cahn_hilliard.c:1018-1052, 1181-1218)
//   LB planes    after the collision: reprojection of the plane-crossing
//                distributions, integer displacement and linear
//                interpolation (model_le.c:78-625)
// Every sum keeps the reference's operand order; compiled without
// contraction, the results are the reference's bits.  The per-plane
// correction of the force is a sum in site order, which one thread of a
// block per plane and component forms serially, as the reference's loop
// does; those blocks run inside the launch that forms the force off the
// planes, and a small launch then forms the force at the plane sites.

#include <hip/hip_runtime.h>

#include "advect.h"
#include "ch_conserve.h"
#include "d3q19.h"
#include "grad27.h"
#include "kernels.h"
#include "le_lb.h"
#include "le_plane_sum.h"

namespace ldw {

namespace {

constexpr int LB = 256;

__device__ __forceinline__ int wrapi(int i, int n) {
  return (i < 0) ? i + n : ((i >= n) ? i - n : i);
}

// lees_edw_ic_to_buff (leesedwards.c:1030-1065): the 1-based x index of
// the neighbour di of 1-based real x ic; beyond nx + nh it is a buffer
// plane.  The plane tested is ic/(nx/nplane), clamped, as there.
__device__ __forceinline__ int ic_to_buff(const LeGeom & le, int nx, int ic, int di) {
  if (le.nplane > 0) {
    int p = ic/le.sep;
    p = (p < 0) ? 0 : ((p > le.nplane - 1) ? le.nplane - 1 : p);
    const int nh = le.nh;
    int ip = le.loc[p] - (nh - 1);
    if (di > 0 && (ic >= ip && ic < ip + nh) && (ic + di >= ip + nh)) {
      return nx + (1 + 2*p)*nh + (ic - ip + 1) + di;
    }
    ip = le.loc[p] + 1;
    if (di < 0 && (ic >= ip && ic < ip + nh) && (ic + di < ip)) {
      return nx + (2 + 2*p)*nh + (ic - ip + 1) + di;
    }
  }
  return ic + di;
}

// 1-based x (real, halo or buffer) -> virtual plane index
__device__ __forceinline__ int virt(const LeGeom & le, int nx, int xr) {
  if (xr > nx + le.nh) return nx + (xr - (nx + le.nh + 1));
  return wrapi(xr - 1, nx);
}

// mapped x neighbour of real plane x (0-based)
__device__ __forceinline__ int xnb(const LeGeom & le, int nx, int x, int di) {
  return virt(le, nx, ic_to_buff(le, nx, x + 1, di));
}

// one component over the real planes and the buffer planes
struct VF {
  const double * __restrict__ r;
  const double * __restrict__ b;
  int nx, ny, nz;
  __device__ __forceinline__ double at(int v, int y, int z) const {
    y = wrapi(y, ny); z = wrapi(z, nz);
    const int64_t pl = (int64_t) ny*nz;
    const int64_t s = (int64_t) y*nz + z;
    return (v < nx) ? r[(int64_t) v*pl + s] : b[(int64_t) (v - nx)*pl + s];
  }
};

struct Coord { int v, y, z; };

__device__ __forceinline__ Coord coord_of(int64_t i, int ny, int nz) {
  const int64_t nyz = (int64_t) ny*nz;
  Coord c;
  c.v = (int) (i/nyz);
  const int rem = (int) (i - (int64_t) c.v*nyz);
  c.y = rem/nz;
  c.z = rem - c.y*nz;
  return c;
}

/* ---------------------------------------------------------------------- */
/* phi buffer, gradient                                                   */
/* ---------------------------------------------------------------------- */

__global__ void __launch_bounds__(LB)
le_phi_buffer_kernel(Geom g, LeGeom le, LeShift sh, const double * __restrict__ phi,
		     double * __restrict__ phib) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nxbuf*nyz) return;
  const Coord c = coord_of(i, g.ny, g.nz);
  const int b = c.v;
  // lees_edw_ibuff_to_real (leesedwards.c:1010-1025); planes are kept
  // more than nh sites from the x boundaries, so this is a real plane
  const int x = le.loc[b/(2*le.nh)] - (le.nh - 1) + b%(2*le.nh) - 1;
  const int N = g.ny;
  const int jdy = sh.jdy[b];
  const double fr = sh.fr[b];
  // j0 = 1 + (jc - jdy - 3 + 2N) % N, jc = y + 1, and its successors
  const int j0 = (c.y - jdy - 2 + 2*N) % N;
  const int j1 = (j0 + 1) % N, j2 = (j1 + 1) % N, j3 = (j2 + 1) % N;
  const double * pl = phi + (int64_t) x*nyz + c.z;
  const double r6 = (1.0/6.0);
  phib[i] = - r6*fr*(fr - 1.0)*(fr - 2.0)*pl[(int64_t) j0*g.nz]
    + 0.5*(fr*fr - 1.0)*(fr - 2.0)*pl[(int64_t) j1*g.nz]
    - 0.5*fr*(fr + 1.0)*(fr - 2.0)*pl[(int64_t) j2*g.nz]
    + r6*fr*(fr*fr - 1.0)*pl[(int64_t) j3*g.nz];
}

__device__ __forceinline__ void grad_site(const Geom & g, const LeGeom & le, int stencil,
					  const double * __restrict__ phi, const double * __restrict__ phib,
					  double * __restrict__ gr, double * __restrict__ dr,
					  double * __restrict__ gb, double * __restrict__ db,
					  int v, int y, int z) {
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t i = (int64_t) v*nyz + (int64_t) y*g.nz + z;
  int vm, vp;
  if (v < g.nx) {
    vm = xnb(le, g.nx, v, -1);
    vp = xnb(le, g.nx, v, +1);
  }
  else {
    // the buffer sites of grad_3d_7pt_fluid_le (gradient_3d_7pt_fluid.c:
    // 355-425): nh' = 1 .. nh - 1 planes into the buffer on either side
    const int b = v - g.nx, nh = le.nh, s = b%(2*nh);
    const int L = le.loc[b/(2*nh)];
    if (s >= nh) {
      const int k = s - nh + 1;
      if (k > nh - 1) return;
      vm = virt(le, g.nx, ic_to_buff(le, g.nx, L, k - 1));
      vp = virt(le, g.nx, ic_to_buff(le, g.nx, L, k + 1));
    }
    else {
      const int k = nh - s;
      if (k > nh - 1) return;
      vp = virt(le, g.nx, ic_to_buff(le, g.nx, L + 1, -k + 1));
      vm = virt(le, g.nx, ic_to_buff(le, g.nx, L + 1, -k - 1));
    }
  }
  const VF P = {phi, phib, g.nx, g.ny, g.nz};
  double gv[3], d;
  if (stencil == 27) {
    auto p = [&](int a, int bb, int cc) {
      const int vv = (a == 0) ? vm : ((a == 1) ? v : vp);
      return P.at(vv, y + bb - 1, z + cc - 1);
    };
    grad27<double>(p, gv, d);
  }
  else {
    const double xp = P.at(vp, y, z), xm = P.at(vm, y, z);
    const double yp = P.at(v, y + 1, z), ym = P.at(v, y - 1, z);
    const double zp = P.at(v, y, z + 1), zm = P.at(v, y, z - 1);
    gv[0] = 0.5*(xp - xm);
    gv[1] = 0.5*(yp - ym);
    gv[2] = 0.5*(zp - zm);
    d = xp + xm + yp + ym + zp + zm - 6.0*P.at(v, y, z);
  }
  if (v < g.nx) {
    for (int a = 0; a < 3; a++) gr[a*g.nsite + i] = gv[a];
    dr[i] = d;
  }
  else {
    const int64_t nb = (int64_t) le.nxbuf*nyz;
    const int64_t j = i - g.nsite;
    for (int a = 0; a < 3; a++) gb[a*nb + j] = gv[a];
    db[j] = d;
  }
}

__global__ void __launch_bounds__(LB)
le_grad_kernel(Geom g, LeGeom le, int stencil, const double * __restrict__ phi,
	       const double * __restrict__ phib, double * __restrict__ gr,
	       double * __restrict__ dr, double * __restrict__ gb, double * __restrict__ db) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) (g.nx + le.nxbuf)*nyz) return;
  const Coord c = coord_of(i, g.ny, g.nz);
  grad_site(g, le, stencil, phi, phib, gr, dr, gb, db, c.v, c.y, c.z);
}

// the gradient only where the plane pre-pass reads it: for each plane the
// real planes either side (x = L - 1, L, 0-based) and buffer planes nh - 1
// and nh
__global__ void __launch_bounds__(LB)
le_grad_planes_kernel(Geom g, LeGeom le, int stencil, const double * __restrict__ phi,
		      const double * __restrict__ phib, double * __restrict__ gr,
		      double * __restrict__ dr, double * __restrict__ gb, double * __restrict__ db) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nplane*4*nyz) return;
  const Coord c = coord_of(i, g.ny, g.nz);
  const int p = c.v/4, k = c.v%4;
  const int v = (k < 2) ? le.loc[p] - 1 + k : g.nx + 2*p*le.nh + le.nh - 1 + (k - 2);
  grad_site(g, le, stencil, phi, phib, gr, dr, gb, db, v, c.y, c.z);
}

/* ---------------------------------------------------------------------- */
/* Flux-form force                                                        */
/* ---------------------------------------------------------------------- */

struct GradFields {
  VF phi, g[3], d;
};

static GradFields grad_fields(const Geom & g, const LeGeom & le,
						  const double * phi, const double * phib,
						  const double * gr, const double * dr,
						  const double * gb, const double * db) {
  const int64_t nb = (int64_t) le.nxbuf*g.ny*g.nz;
  GradFields f;
  f.phi = {phi, phib, g.nx, g.ny, g.nz};
  for (int a = 0; a < 3; a++) f.g[a] = {gr + a*g.nsite, gb + a*nb, g.nx, g.ny, g.nz};
  f.d = {dr, db, g.nx, g.ny, g.nz};
  return f;
}

// column b of fe_symm_str (symmetric.c:334-360) at a virtual site:
// s[a] = P_ab, with p0 d_ab + kappa d_a phi d_b phi
__device__ __forceinline__ void stress_col(const FeParams & fe, const GradFields & F,
					   int v, int y, int z, int b, double s[3]) {
  const double phi = F.phi.at(v, y, z);
  const double delsq = F.d.at(v, y, z);
  double gp[3];
  for (int a = 0; a < 3; a++) gp[a] = F.g[a].at(v, y, z);
  const double kappa = fe.kappa;
  const double p0 = 0.5*fe.a*phi*phi + 0.75*fe.b*phi*phi*phi*phi
    - kappa*phi*delsq - 0.5*kappa*(gp[0]*gp[0] + gp[1]*gp[1] + gp[2]*gp[2]);
  for (int a = 0; a < 3; a++) {
    const double dab = (a == b) ? 1.0 : 0.0;
    s[a] = p0*dab + kappa*gp[a]*gp[b];
  }
}

// the raw x-face fluxes at the plane sites (phi_force.c:351-378):
// side 0 fluxe at x = L, side 1 fluxw at x = L + 1 (1-based)
__global__ void __launch_bounds__(LB)
le_force_faces_kernel(Geom g, LeGeom le, FeParams fe, GradFields F, double * __restrict__ xface) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nplane*2*nyz) return;
  const int ps = (int) (i/nyz);
  const int rem = (int) (i - (int64_t) ps*nyz);
  const int p = ps/2, side = ps%2;
  const int y = rem/g.nz, z = rem - (rem/g.nz)*g.nz;
  const int x = le.loc[p] - 1 + side;
  const int vn = xnb(le, g.nx, x, side ? -1 : +1);
  double s0[3], s1[3];
  stress_col(fe, F, x, y, z, 0, s0);
  stress_col(fe, F, vn, y, z, 0, s1);
  for (int a = 0; a < 3; a++) xface[((int64_t) ps*3 + a)*nyz + rem] = 0.5*(s1[a] + s0[a]);
}

// the plane sums alone, one block per plane and component.  The lean sum
// takes 48 VGPRs and 16 KB of LDS, so that a block fits on a CU beside a
// block of the single-pass step (2 waves per SIMD at 232 of the 512 VGPRs,
// 112 KB of the 160 KB LDS)
__global__ void __launch_bounds__(LB)
le_plane_sums_kernel(const double * __restrict__ xface, int64_t nyz, double * __restrict__ fcor) {
  plane_sum_block<true>(xface, nyz, fcor, blockIdx.x);
}

// force = -(fluxe - fluxw + fluxy - fluxy(y-1) + fluxz - fluxz(z-1)) at
// real site (x, y, z); at the plane sites fluxe (x = L) or fluxw
// (x = L + 1) carry the correction +/- ra fcor (phi_force.c:412-448, 502-520)
__device__ __forceinline__ void force_at(const Geom & g, const LeGeom & le, const FeParams & fe,
					 const GradFields & F, const double * __restrict__ fcor,
					 double ra, int x, int y, int z, double fo[3]) {
  const int vm = xnb(le, g.nx, x, -1), vp = xnb(le, g.nx, x, +1);
  double s0x[3], sxm[3], sxp[3], s0y[3], syp[3], sym[3], s0z[3], szp[3], szm[3];
  stress_col(fe, F, x, y, z, 0, s0x);
  stress_col(fe, F, vm, y, z, 0, sxm);
  stress_col(fe, F, vp, y, z, 0, sxp);
  stress_col(fe, F, x, y, z, 1, s0y);
  stress_col(fe, F, x, y + 1, z, 1, syp);
  stress_col(fe, F, x, y - 1, z, 1, sym);
  stress_col(fe, F, x, y, z, 2, s0z);
  stress_col(fe, F, x, y, z + 1, 2, szp);
  stress_col(fe, F, x, y, z - 1, 2, szm);
  const int ic = x + 1;
  for (int a = 0; a < 3; a++) {
    double fe_ = 0.5*(sxp[a] + s0x[a]);
    double fw = 0.5*(sxm[a] + s0x[a]);
    const double fy = 0.5*(syp[a] + s0y[a]);
    const double fym = 0.5*(s0y[a] + sym[a]);
    const double fz = 0.5*(szp[a] + s0z[a]);
    const double fzm = 0.5*(s0z[a] + szm[a]);
    if (fcor != nullptr) {
      for (int p = 0; p < le.nplane; p++) {
	if (ic == le.loc[p]) fe_ += ra*fcor[3*p + a];
	if (ic == le.loc[p] + 1) fw -= ra*fcor[3*p + a];
      }
    }
    fo[a] = -(+ fe_ - fw + fy - fym + fz - fzm);
  }
}

__device__ __forceinline__ void force_store(double * __restrict__ force, int64_t nsite, int64_t i,
					    int force_mode, const double fo[3], double fz0, double fz1,
					    double fz2) {
  if (force_mode == 0) {
    force[i] = fz0 + fo[0];
    force[nsite + i] = fz1 + fo[1];
    force[2*nsite + i] = fz2 + fo[2];
  }
  else {
    force[i] += fo[0];
    force[nsite + i] += fo[1];
    force[2*nsite + i] += fo[2];
  }
}

__device__ __forceinline__ bool on_plane(const LeGeom & le, int x) {
  for (int p = 0; p < le.nplane; p++) {
    if (x == le.loc[p] - 1 || x == le.loc[p]) return true;
  }
  return false;
}

// The first 3 nplane blocks form the plane sums; the others the force at
// every site off the planes, which needs no correction, concurrently.
__global__ void __launch_bounds__(LB)
le_force_kernel(Geom g, LeGeom le, FeParams fe, GradFields F, const double * __restrict__ xface,
		double * __restrict__ fcor, double * __restrict__ force, int force_mode,
		double fz0, double fz1, double fz2) {
  const int nsum = 3*le.nplane;
  if ((int) blockIdx.x < nsum) {
    plane_sum_block<false>(xface, (int64_t) g.ny*g.nz, fcor, blockIdx.x);
    return;
  }
  const int64_t i = (int64_t) (blockIdx.x - nsum)*blockDim.x + threadIdx.x;
  if (i >= g.nsite) return;
  const Coord c = coord_of(i, g.ny, g.nz);
  if (on_plane(le, c.v)) return;
  double fo[3];
  force_at(g, le, fe, F, nullptr, 0.0, c.v, c.y, c.z, fo);
  force_store(force, g.nsite, i, force_mode, fo, fz0, fz1, fz2);
}

// the sites either side of each plane, once the sums are known
__global__ void __launch_bounds__(LB)
le_plane_force_kernel(Geom g, LeGeom le, FeParams fe, GradFields F, const double * __restrict__ fcor,
		      double ra, double * __restrict__ force, int force_mode, double fz0, double fz1,
		      double fz2) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nplane*2*nyz) return;
  const int ps = (int) (i/nyz);
  const int rem = (int) (i - (int64_t) ps*nyz);
  const int x = le.loc[ps/2] - 1 + ps%2;
  const int y = rem/g.nz, z = rem - (rem/g.nz)*g.nz;
  double fo[3];
  force_at(g, le, fe, F, fcor, ra, x, y, z, fo);
  force_store(force, g.nsite, (int64_t) x*nyz + rem, force_mode, fo, fz0, fz1, fz2);
}

/* ---------------------------------------------------------------------- */
/* Velocity buffer, Cahn-Hilliard                                         */
/* ---------------------------------------------------------------------- */

__global__ void __launch_bounds__(LB)
le_u_buffer_kernel(Geom g, LeGeom le, LeShift sh, double uy, const double * __restrict__ u,
		   double * __restrict__ ub) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t nb = (int64_t) le.nxbuf*nyz;
  if (i >= nb) return;
  const Coord c = coord_of(i, g.ny, g.nz);
  const int b = c.v;
  const int s = b%(2*le.nh);
  const int x = le.loc[b/(2*le.nh)] - (le.nh - 1) + s - 1;
  const int N = g.ny;
  const int jdy = sh.jdy[b];
  const double fr = sh.fr[b];
  // j1 = 1 + (jc - jdy - 2 + 2N) % N, j2 = 1 + j1 % N
  const int j1 = (c.y - jdy - 1 + 2*N) % N;
  const int j2 = (j1 + 1) % N;
  // lees_edw_buffer_du: (0, uy sign, 0), sign -1 below the plane
  const double ule[3] = {0.0, uy*((s < le.nh) ? -1.0 : 1.0), 0.0};
  const int64_t s1 = (int64_t) x*nyz + (int64_t) j1*g.nz + c.z;
  const int64_t s2 = (int64_t) x*nyz + (int64_t) j2*g.nz + c.z;
  for (int a = 0; a < 3; a++) {
    ub[a*nb + i] = ule[a] + u[a*g.nsite + s1]*fr + u[a*g.nsite + s2]*(1.0 - fr);
  }
}

struct ChFlux {
  double fw, fe, fyp, fym, fzp, fzm;
};

__device__ __forceinline__ double mu_at(const FeParams & fe, const VF & P, const VF & D,
					int v, int y, int z) {
  const double phi = P.at(v, y, z);
  return fe.a*phi + fe.b*phi*phi*phi - fe.kappa*D.at(v, y, z);
}

// the west, east, y, y-1, z and z-1 face fluxes of real site (x, y, z),
// each as the reference forms it at its own site (advection.c:538-632,
// 680-760, 946-1145; phi_cahn_hilliard.c:350-413); x neighbours mapped
__device__ __forceinline__ ChFlux ch_fluxes(const Geom & g, const LeGeom & le, const FeParams & fe,
					    const VF & P, const VF & D, bool adv, const VF & UX,
					    const VF & UY, const VF & UZ, int x, int y, int z) {
  ChFlux f = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int m1 = xnb(le, g.nx, x, -1), p1 = xnb(le, g.nx, x, +1);
  const double a1 = -0.213933, a2 = 0.927865, a3 = 0.286067;
  if (adv) {
    const double ux0 = UX.at(x, y, z), uy0 = UY.at(x, y, z), uz0 = UZ.at(x, y, z);
    const double uxm = UX.at(m1, y, z), uxp = UX.at(p1, y, z);
    const double uyp = UY.at(x, y + 1, z), uym = UY.at(x, y - 1, z);
    const double uzp = UZ.at(x, y, z + 1), uzm = UZ.at(x, y, z - 1);
    const double uw = 0.5*(ux0 + uxm);
    const double ue = 0.5*(ux0 + uxp);
    const double uyf = 0.5*(uy0 + uyp);
    const double uyb = 0.5*(uym + uy0);
    const double uzf = 0.5*(uz0 + uzp);
    const double uzb = 0.5*(uzm + uz0);
    const double p0 = P.at(x, y, z);
    if (fe.order == 1) {
      f.fw = uw*((uw > 0.0) ? P.at(m1, y, z) : p0);
      f.fe = ue*((ue < 0.0) ? P.at(p1, y, z) : p0);
      f.fyp = uyf*((uyf < 0.0) ? P.at(x, y + 1, z) : p0);
      f.fym = uyb*((uyb < 0.0) ? p0 : P.at(x, y - 1, z));
      f.fzp = uzf*((uzf < 0.0) ? P.at(x, y, z + 1) : p0);
      f.fzm = uzb*((uzb < 0.0) ? p0 : P.at(x, y, z - 1));
    }
    else if (fe.order == 2) {
      f.fw = uw*0.5*(P.at(m1, y, z) + p0);
      f.fe = ue*0.5*(p0 + P.at(p1, y, z));
      f.fyp = uyf*0.5*(p0 + P.at(x, y + 1, z));
      f.fym = uyb*0.5*(P.at(x, y - 1, z) + p0);
      f.fzp = uzf*0.5*(p0 + P.at(x, y, z + 1));
      f.fzm = uzb*0.5*(P.at(x, y, z - 1) + p0);
    }
    else if (fe.order == 4) {
      // advection_le_4th (advection.c:1153-1262), x neighbours mapped
      const int m2 = xnb(le, g.nx, x, -2), p2 = xnb(le, g.nx, x, +2);
      f.fw = adv::order4(uw, P.at(m2, y, z), P.at(m1, y, z), p0, P.at(p1, y, z));
      f.fe = adv::order4(ue, P.at(m1, y, z), p0, P.at(p1, y, z), P.at(p2, y, z));
      f.fyp = adv::order4(uyf, P.at(x, y - 1, z), p0, P.at(x, y + 1, z), P.at(x, y + 2, z));
      f.fym = adv::order4(uyb, P.at(x, y - 2, z), P.at(x, y - 1, z), p0, P.at(x, y + 1, z));
      f.fzp = adv::order4(uzf, P.at(x, y, z - 1), p0, P.at(x, y, z + 1), P.at(x, y, z + 2));
      f.fzm = adv::order4(uzb, P.at(x, y, z - 2), P.at(x, y, z - 1), p0, P.at(x, y, z + 1));
    }
    else {
      const int m2 = xnb(le, g.nx, x, -2), p2 = xnb(le, g.nx, x, +2);
      double d1, d2, d3;
      if (uw > 0.0) { d1 = P.at(m2, y, z); d2 = P.at(m1, y, z); d3 = p0; }
      else          { d1 = P.at(p1, y, z); d2 = p0; d3 = P.at(m1, y, z); }
      f.fw = uw*(a1*d1 + a2*d2 + a3*d3);
      if (ue < 0.0) { d1 = P.at(p2, y, z); d2 = P.at(p1, y, z); d3 = p0; }
      else          { d1 = P.at(m1, y, z); d2 = p0; d3 = P.at(p1, y, z); }
      f.fe = ue*(a1*d1 + a2*d2 + a3*d3);
      if (uyf < 0.0) { d1 = P.at(x, y + 2, z); d2 = P.at(x, y + 1, z); d3 = p0; }
      else           { d1 = P.at(x, y - 1, z); d2 = p0; d3 = P.at(x, y + 1, z); }
      f.fyp = uyf*(a1*d1 + a2*d2 + a3*d3);
      if (uyb < 0.0) { d1 = P.at(x, y + 1, z); d2 = p0; d3 = P.at(x, y - 1, z); }
      else           { d1 = P.at(x, y - 2, z); d2 = P.at(x, y - 1, z); d3 = p0; }
      f.fym = uyb*(a1*d1 + a2*d2 + a3*d3);
      if (uzf < 0.0) { d1 = P.at(x, y, z + 2); d2 = P.at(x, y, z + 1); d3 = p0; }
      else           { d1 = P.at(x, y, z - 1); d2 = p0; d3 = P.at(x, y, z + 1); }
      f.fzp = uzf*(a1*d1 + a2*d2 + a3*d3);
      if (uzb < 0.0) { d1 = P.at(x, y, z + 1); d2 = p0; d3 = P.at(x, y, z - 1); }
      else           { d1 = P.at(x, y, z - 2); d2 = P.at(x, y, z - 1); d3 = p0; }
      f.fzm = uzb*(a1*d1 + a2*d2 + a3*d3);
    }
  }
  const double mob = fe.mobility;
  const double m0 = mu_at(fe, P, D, x, y, z);
  f.fw -= mob*(m0 - mu_at(fe, P, D, m1, y, z));
  f.fe -= mob*(mu_at(fe, P, D, p1, y, z) - m0);
  f.fyp -= mob*(mu_at(fe, P, D, x, y + 1, z) - m0);
  f.fym -= mob*(m0 - mu_at(fe, P, D, x, y - 1, z));
  f.fzp -= mob*(mu_at(fe, P, D, x, y, z + 1) - m0);
  f.fzm -= mob*(m0 - mu_at(fe, P, D, x, y, z - 1));
  // phi_ch_flux_mu_ext (phi_cahn_hilliard.c:1373-1398), before the plane
  // faces are averaged
  if (fe.gmu_on) adv::mu_ext(mob, fe.gmu, f.fw, f.fe, f.fyp, f.fym, f.fzp, f.fzm);
  return f;
}

struct ChFields {
  VF phi, d, ux, uy, uz;
  bool adv;
};

// raw total x-face fluxes at the plane sites: side 0 fe at L, side 1 fw at
// L + 1, before phi_ch_le_fix_fluxes averages them
__global__ void __launch_bounds__(LB)
le_ch_faces_kernel(Geom g, LeGeom le, FeParams fe, ChFields C, double * __restrict__ xface) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nplane*2*nyz) return;
  const int ps = (int) (i/nyz);
  const int rem = (int) (i - (int64_t) ps*nyz);
  const int p = ps/2, side = ps%2;
  const int y = rem/g.nz, z = rem - (rem/g.nz)*g.nz;
  const ChFlux f = ch_fluxes(g, le, fe, C.phi, C.d, C.adv, C.ux, C.uy, C.uz,
			     le.loc[p] - 1 + side, y, z);
  xface[i] = side ? f.fw : f.fe;
}

__global__ void __launch_bounds__(LB)
le_ch_update_kernel(Geom g, LeGeom le, FeParams fe, ChFields C, const double * __restrict__ xface,
		    int jdy_w, double fr_w, int jdy_e, double fr_e, double * __restrict__ phi_out,
		    double * __restrict__ csum) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= g.nsite) return;
  const Coord c = coord_of(i, g.ny, g.nz);
  const int x = c.v, y = c.y, z = c.z;
  ChFlux f = ch_fluxes(g, le, fe, C.phi, C.d, C.adv, C.ux, C.uy, C.uz, x, y, z);
  const int ic = x + 1, N = g.ny;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  for (int p = 0; p < le.nplane; p++) {
    if (ic == le.loc[p]) {
      // looking up: fw of plane L + 1 at y - jdy - 1 (+1), weights fr, 1 - fr
      const int j1 = (y - jdy_w - 1 + 2*N) % N, j2 = (j1 + 1) % N;
      const double * fw = xface + (int64_t) (2*p + 1)*nyz;
      const double bw = fw[(int64_t) j1*g.nz + z]*fr_w + fw[(int64_t) j2*g.nz + z]*(1.0 - fr_w);
      f.fe = 0.5*(f.fe + bw);
    }
    if (ic == le.loc[p] + 1) {
      const int j1 = (y - jdy_e - 1 + 2*N) % N, j2 = (j1 + 1) % N;
      const double * fe_ = xface + (int64_t) (2*p + 0)*nyz;
      const double be = fe_[(int64_t) j1*g.nz + z]*fr_e + fe_[(int64_t) j2*g.nz + z]*(1.0 - fr_e);
      f.fw = 0.5*(f.fw + be);
    }
  }
  double phi = C.phi.r[i];
  if (csum != nullptr) {
    double cs = csum[i];
    ch_update_compensated(phi, cs, f.fe, f.fw, f.fyp, f.fym, f.fzp, f.fzm);
    csum[i] = cs;
  }
  else {
    phi -= (+ f.fe - f.fw + f.fyp - f.fym + f.fzp - f.fzm);
  }
  phi_out[i] = phi;
}

/* ---------------------------------------------------------------------- */
/* Lattice Boltzmann plane conditions                                     */
/* ---------------------------------------------------------------------- */

// le_reproject (model_le.c:245-330) at every plane site (le_lb.h)
__global__ void __launch_bounds__(LB)
le_lb_reproject_kernel(Geom g, LeGeom le, double duy0, double duy1,
		       const double * __restrict__ gd, double * __restrict__ sbuf) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nplane*2*nyz) return;
  const int ps = (int) (i/nyz);
  const int rem = (int) (i - (int64_t) ps*nyz);
  const int p = ps/2, side = ps%2;
  const int64_t site = (int64_t) (le.loc[p] - 1 + side)*nyz + rem;
  double f[NVEL];
#pragma unroll
  for (int q = 0; q < NVEL; q++) f[q] = gd[q*g.nsite + site];
  double out[5];
  le_reproject_site(f, side, side ? duy1 : duy0, out);
#pragma unroll
  for (int ip = 0; ip < 5; ip++) sbuf[((int64_t) ps*5 + ip)*nyz + rem] = out[ip];
}

// lb_data_displace_kernel + lb_data_interpolate_kernel (model_le.c:
// 345-395, 585-625): f(y) = (1 - fr) s(y + dj) + fr s(y + dj + 1), periodic
__global__ void __launch_bounds__(LB)
le_lb_interp_kernel(Geom g, LeGeom le, int dj0, int dj1, double fr0, double fr1,
		    const double * __restrict__ sbuf, double * __restrict__ gd) {
  const int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  const int64_t nyz = (int64_t) g.ny*g.nz;
  if (i >= (int64_t) le.nplane*2*nyz) return;
  const int ps = (int) (i/nyz);
  const int rem = (int) (i - (int64_t) ps*nyz);
  const int p = ps/2, side = ps%2;
  const int y = rem/g.nz, z = rem - (rem/g.nz)*g.nz;
  const int N = g.ny;
  const int dj = side ? dj1 : dj0;
  const double fr = side ? fr1 : fr0;
  // js = 1 + (jc + dj - 1 + 2N) % N for jc = y + 1 and y + 2
  const int js0 = (y + dj + 2*N) % N;
  const int js1 = (y + 1 + dj + 2*N) % N;
  const int64_t site = (int64_t) (le.loc[p] - 1 + side)*nyz + rem;
#pragma unroll
  for (int ip = 0; ip < 5; ip++) {
    const double * s = sbuf + ((int64_t) ps*5 + ip)*nyz;
    const double fq = (1.0 - fr)*s[(int64_t) js0*g.nz + z] + fr*s[(int64_t) js1*g.nz + z];
    const int q = side ? LE_PROPS.p[1][ip] : LE_PROPS.p[0][ip];
    gd[q*g.nsite + site] = fq;
  }
}

inline unsigned nblk(int64_t n) { return (unsigned) ((n + LB - 1)/LB); }

}  // namespace

void launch_le_phi_buffer(const Geom & g, const LeGeom & le, const LeShift & sh,
			  const double * phi, double * phib, hipStream_t s) {
  const int64_t n = (int64_t) le.nxbuf*g.ny*g.nz;
  hipLaunchKernelGGL(le_phi_buffer_kernel, dim3(nblk(n)), dim3(LB), 0, s, g, le, sh, phi, phib);
}

void launch_le_grad(const Geom & g, const LeGeom & le, int stencil, const double * phi,
		    const double * phib, double * gr, double * dr, double * gb, double * db,
		    hipStream_t s) {
  const int64_t n = (int64_t) (g.nx + le.nxbuf)*g.ny*g.nz;
  hipLaunchKernelGGL(le_grad_kernel, dim3(nblk(n)), dim3(LB), 0, s, g, le, stencil, phi, phib,
		     gr, dr, gb, db);
}

void launch_le_grad_planes(const Geom & g, const LeGeom & le, int stencil, const double * phi,
			   const double * phib, double * gr, double * dr, double * gb, double * db,
			   hipStream_t s) {
  const int64_t n = (int64_t) le.nplane*4*g.ny*g.nz;
  hipLaunchKernelGGL(le_grad_planes_kernel, dim3(nblk(n)), dim3(LB), 0, s, g, le, stencil, phi, phib,
		     gr, dr, gb, db);
}

void launch_le_plane_sums(const Geom & g, const LeGeom & le, const double * xface, double * fcor,
			  hipStream_t s) {
  hipLaunchKernelGGL(le_plane_sums_kernel, dim3(3*le.nplane), dim3(LB), 0, s, xface,
		     (int64_t) g.ny*g.nz, fcor);
}

void launch_le_force_faces(const Geom & g, const LeGeom & le, const FeParams & fe,
			   const double * phi, const double * phib, const double * gr,
			   const double * dr, const double * gb, const double * db,
			   double * xface, hipStream_t s) {
  const GradFields F = grad_fields(g, le, phi, phib, gr, dr, gb, db);
  const int64_t nf = (int64_t) le.nplane*2*g.ny*g.nz;
  hipLaunchKernelGGL(le_force_faces_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, fe, F, xface);
}

void launch_le_ch_faces(const Geom & g, const LeGeom & le, const FeParams & fe,
			const double * phi, const double * phib, const double * dr,
			const double * db, const double * u, const double * ub,
			double * xface, hipStream_t s) {
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t nb = (int64_t) le.nxbuf*nyz;
  ChFields C;
  C.phi = {phi, phib, g.nx, g.ny, g.nz};
  C.d = {dr, db, g.nx, g.ny, g.nz};
  C.adv = (u != nullptr);
  const double * u0 = u ? u : phi;
  const double * b0 = ub ? ub : phib;
  C.ux = {u0, b0, g.nx, g.ny, g.nz};
  C.uy = {u ? u + g.nsite : u0, ub ? ub + nb : b0, g.nx, g.ny, g.nz};
  C.uz = {u ? u + 2*g.nsite : u0, ub ? ub + 2*nb : b0, g.nx, g.ny, g.nz};
  const int64_t nf = (int64_t) le.nplane*2*nyz;
  hipLaunchKernelGGL(le_ch_faces_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, fe, C, xface);
}

void launch_le_u_buffer(const Geom & g, const LeGeom & le, const LeShift & sh, double uy,
			const double * u, double * ub, hipStream_t s) {
  const int64_t n = (int64_t) le.nxbuf*g.ny*g.nz;
  hipLaunchKernelGGL(le_u_buffer_kernel, dim3(nblk(n)), dim3(LB), 0, s, g, le, sh, uy, u, ub);
}

void launch_le_force(const Geom & g, const LeGeom & le, const FeParams & fe,
		     const double * phi, const double * phib, const double * gr,
		     const double * dr, const double * gb, const double * db,
		     double * xface, double * fcor, double ra, double * force,
		     int force_mode, const double fzero[3], hipStream_t s) {
  const GradFields F = grad_fields(g, le, phi, phib, gr, dr, gb, db);
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t nf = (int64_t) le.nplane*2*nyz;
  hipLaunchKernelGGL(le_force_faces_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, fe, F, xface);
  hipLaunchKernelGGL(le_force_kernel, dim3(3*le.nplane + nblk(g.nsite)), dim3(LB), 0, s, g, le, fe,
		     F, xface, fcor, force, force_mode, fzero[0], fzero[1], fzero[2]);
  hipLaunchKernelGGL(le_plane_force_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, fe, F, fcor, ra,
		     force, force_mode, fzero[0], fzero[1], fzero[2]);
}

void launch_le_cahn_hilliard(const Geom & g, const LeGeom & le, const FeParams & fe,
			     const double * phi, const double * phib, const double * dr,
			     const double * db, const double * u, const double * ub,
			     double * xface, int jdy_w, double fr_w, int jdy_e, double fr_e,
			     double * phi_out, double * csum, hipStream_t s) {
  const int64_t nyz = (int64_t) g.ny*g.nz;
  const int64_t nb = (int64_t) le.nxbuf*nyz;
  ChFields C;
  C.phi = {phi, phib, g.nx, g.ny, g.nz};
  C.d = {dr, db, g.nx, g.ny, g.nz};
  C.adv = (u != nullptr);
  const double * u0 = u ? u : phi;      // not read without advection
  const double * b0 = ub ? ub : phib;
  C.ux = {u0, b0, g.nx, g.ny, g.nz};
  C.uy = {u ? u + g.nsite : u0, ub ? ub + nb : b0, g.nx, g.ny, g.nz};
  C.uz = {u ? u + 2*g.nsite : u0, ub ? ub + 2*nb : b0, g.nx, g.ny, g.nz};
  const int64_t nf = (int64_t) le.nplane*2*nyz;
  hipLaunchKernelGGL(le_ch_faces_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, fe, C, xface);
  hipLaunchKernelGGL(le_ch_update_kernel, dim3(nblk(g.nsite)), dim3(LB), 0, s, g, le, fe, C,
		     xface, jdy_w, fr_w, jdy_e, fr_e, phi_out, csum);
}

void launch_le_lb_bc(const Geom & g, const LeGeom & le, const int dj[2], const double fr[2],
		     const double duy[2], double * gdist, double * sbuf, hipStream_t s) {
  const int64_t nf = (int64_t) le.nplane*2*g.ny*g.nz;
  hipLaunchKernelGGL(le_lb_reproject_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, duy[0], duy[1],
		     gdist, sbuf);
  hipLaunchKernelGGL(le_lb_interp_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, dj[0], dj[1],
		     fr[0], fr[1], sbuf, gdist);
}

void launch_le_lb_interp(const Geom & g, const LeGeom & le, const int dj[2], const double fr[2],
			 double * gdist, const double * sbuf, hipStream_t s) {
  const int64_t nf = (int64_t) le.nplane*2*g.ny*g.nz;
  hipLaunchKernelGGL(le_lb_interp_kernel, dim3(nblk(nf)), dim3(LB), 0, s, g, le, dj[0], dj[1],
		     fr[0], fr[1], sbuf, gdist);
}

}  // namespace ldw
