// fe_ch.hip -- K2 (gradient + mu + thermodynamic force) and K3 (fused
// Cahn-Hilliard fluxes and phi update) for the symmetric free energy.
//
// K2 per site, from phi alone (one pass, periodic):
//   grad, delsq by the 7-point stencil (gradient_3d_7pt_fluid.c:231-310) at
//   the site and its 6 neighbours;
//   P_ab = p0 delta_ab + kappa d_a phi d_b phi (symmetric.c:372-418);
//   F_a = -sum_b 1/2 [P_ab(x+e_b) - P_ab(x-e_b)] written as the
//   reference's face averages in x, y, z order
//   (phi_force_colloid.c:352-466), or F = -phi grad mu
//   (phi_grad_mu.c:284-350);
//   mu = A phi + B phi^3 - kappa delsq (symmetric.c:308-318).
// Writes F (3 planes) and mu: 40 B/site of compulsory traffic in FP64.
//
// K3 per site: the west, east, y and z face fluxes of the reference
// (advection.c:538-632, 770-900, 946-1145; phi_cahn_hilliard.c:350-413)
// at the site, plus the y and z fluxes of the lower neighbours computed by
// the same formula, then phi -= (fe - fw + fy - fy(y-1) + fz - fz(z-1))
// (phi_cahn_hilliard.c:1018-1052).  Reads phi, mu, u; writes phi into the
// other buffer: 48 B/site in FP64.
//
// The stencil values are gathered through the L1/L2 caches.  All sums keep
// the reference's order; FP64 is compiled without contraction.

#include <hip/hip_runtime.h>

#include "advect.h"
#include "ch_conserve.h"
#include "grad27.h"
#include "kernels.h"

namespace ldw {

namespace {

struct Idx {
  int x, y, z;
};

__device__ __forceinline__ int wrapi(int i, int n) {
  return (i < 0) ? i + n : ((i >= n) ? i - n : i);
}

template <typename T>
struct Lat {
  const T * __restrict__ a;
  int nx, ny, nz;
  __device__ __forceinline__ T at(int x, int y, int z) const {
    x = wrapi(x, nx); y = wrapi(y, ny); z = wrapi(z, nz);
    return a[((int64_t) x*ny + y)*nz + z];
  }
};

template <typename T>
struct GradPt {
  T phi, g[3], delsq;
};

// gradient and Laplacian at (x,y,z): the 7-point stencil
// (gradient_3d_7pt_fluid.c:270-305), or the 27-point one (grad27.h) when
// stencil is 27
template <typename T>
__device__ __forceinline__ GradPt<T> grad_at(const Lat<T> & L, int x, int y, int z, int stencil) {
  GradPt<T> r;
  if (stencil == 27) {
    auto p = [&](int i, int j, int k) { return L.at(x + i - 1, y + j - 1, z + k - 1); };
    r.phi = L.at(x, y, z);
    grad27<T>(p, r.g, r.delsq);
    return r;
  }
  const T xp = L.at(x + 1, y, z), xm = L.at(x - 1, y, z);
  const T yp = L.at(x, y + 1, z), ym = L.at(x, y - 1, z);
  const T zp = L.at(x, y, z + 1), zm = L.at(x, y, z - 1);
  r.phi = L.at(x, y, z);
  r.g[0] = T(0.5)*(xp - xm);
  r.g[1] = T(0.5)*(yp - ym);
  r.g[2] = T(0.5)*(zp - zm);
  r.delsq = xp + xm + yp + ym + zp + zm - T(6.0)*r.phi;
  return r;
}

// p0 of fe_symm_str_v (symmetric.c:404-406)
template <typename T>
__device__ __forceinline__ T p0_of(const FeParams & fe, const GradPt<T> & q) {
  const T a = T(fe.a), b = T(fe.b), kappa = T(fe.kappa);
  const T phi = q.phi;
  return T(0.5)*a*phi*phi + T(0.75)*b*phi*phi*phi*phi - kappa*phi*q.delsq
    - T(0.5)*kappa*(q.g[0]*q.g[0] + q.g[1]*q.g[1] + q.g[2]*q.g[2]);
}

// P_ab = p0 d_ab + kappa g_a g_b (symmetric.c:410-413)
template <typename T>
__device__ __forceinline__ T pab(T kappa, T p0, const GradPt<T> & q, int a, int b) {
  T s = kappa*q.g[a]*q.g[b];
  return (a == b) ? p0 + s : s;
}

template <typename T>
__device__ __forceinline__ T mu_of(const FeParams & fe, T phi, T delsq) {
  return T(fe.a)*phi + T(fe.b)*phi*phi*phi - T(fe.kappa)*delsq;
}

template <typename T>
__global__ void __launch_bounds__(256)
phi_force_mu_kernel(Geom geo, FeParams fe, const T * __restrict__ phi,
		    T * __restrict__ force, T * __restrict__ mu,
		    int force_mode, T fz0, T fz1, T fz2) {

  const int64_t site = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (site >= geo.nsite) return;
  const int nyz = geo.ny*geo.nz;
  const int x = (int) (site / nyz);
  const int rem = (int) (site - (int64_t) x*nyz);
  const int y = rem / geo.nz;
  const int z = rem - y*geo.nz;

  const Lat<T> L = {phi, geo.nx, geo.ny, geo.nz};
  const GradPt<T> c0 = grad_at(L, x, y, z, fe.stencil);

  T F[3];

  if (force != nullptr) {
    if (fe.force_method == 0) {
      const T kappa = T(fe.kappa);
      const GradPt<T> cxp = grad_at(L, x + 1, y, z, fe.stencil), cxm = grad_at(L, x - 1, y, z, fe.stencil);
      const GradPt<T> cyp = grad_at(L, x, y + 1, z, fe.stencil), cym = grad_at(L, x, y - 1, z, fe.stencil);
      const GradPt<T> czp = grad_at(L, x, y, z + 1, fe.stencil), czm = grad_at(L, x, y, z - 1, fe.stencil);
      const T p00 = p0_of(fe, c0);
      const T pxp = p0_of(fe, cxp), pxm = p0_of(fe, cxm);
      const T pyp = p0_of(fe, cyp), pym = p0_of(fe, cym);
      const T pzp = p0_of(fe, czp), pzm = p0_of(fe, czm);
#pragma unroll
      for (int a = 0; a < 3; a++) {
	T s0;
	s0 = pab(kappa, p00, c0, a, 0);
	T fa = T(-0.5)*(pab(kappa, pxp, cxp, a, 0) + s0);
	fa += T(0.5)*(pab(kappa, pxm, cxm, a, 0) + s0);
	s0 = pab(kappa, p00, c0, a, 1);
	fa -= T(0.5)*(pab(kappa, pyp, cyp, a, 1) + s0);
	fa += T(0.5)*(pab(kappa, pym, cym, a, 1) + s0);
	s0 = pab(kappa, p00, c0, a, 2);
	fa -= T(0.5)*(pab(kappa, pzp, czp, a, 2) + s0);
	fa += T(0.5)*(pab(kappa, pzm, czm, a, 2) + s0);
	F[a] = fa;
      }
    }
    else {
      // phi grad mu: F_a = 0 + (-phi 1/2)(mu(x+e_a) - mu(x-e_a))
      const T phi0 = c0.phi;
      GradPt<T> m, p;
      m = grad_at(L, x - 1, y, z, fe.stencil); p = grad_at(L, x + 1, y, z, fe.stencil);
      F[0] = T(0.0) + -phi0*T(0.5)*(mu_of(fe, p.phi, p.delsq) - mu_of(fe, m.phi, m.delsq));
      m = grad_at(L, x, y - 1, z, fe.stencil); p = grad_at(L, x, y + 1, z, fe.stencil);
      F[1] = T(0.0) + -phi0*T(0.5)*(mu_of(fe, p.phi, p.delsq) - mu_of(fe, m.phi, m.delsq));
      m = grad_at(L, x, y, z - 1, fe.stencil); p = grad_at(L, x, y, z + 1, fe.stencil);
      F[2] = T(0.0) + -phi0*T(0.5)*(mu_of(fe, p.phi, p.delsq) - mu_of(fe, m.phi, m.delsq));
    }

    // hydro->force += F (phi_force_colloid.c:457-462) on a force that
    // hydro_f_zero() set to fzero, or read-modify-write otherwise; then,
    // with phi grad mu, the external gradient's -phi grad mu_ext as a
    // second addition (phi_grad_mu.c:123-160, 509-534)
    T fx, fy, fz;
    if (force_mode == 0) {
      fx = fz0 + F[0]; fy = fz1 + F[1]; fz = fz2 + F[2];
    }
    else {
      fx = force[site] + F[0];
      fy = force[geo.nsite + site] + F[1];
      fz = force[2*geo.nsite + site] + F[2];
    }
    if (fe.force_method == 1 && fe.gmu_on) {
      const T phi0 = c0.phi;
      fx += -phi0*T(fe.gmu[0]);
      fy += -phi0*T(fe.gmu[1]);
      fz += -phi0*T(fe.gmu[2]);
    }
    force[site] = fx;
    force[geo.nsite + site] = fy;
    force[2*geo.nsite + site] = fz;
  }

  mu[site] = mu_of(fe, c0.phi, c0.delsq);
}

template <typename T>
__global__ void __launch_bounds__(256)
mu_kernel(Geom geo, FeParams fe, const T * __restrict__ phi, T * __restrict__ mu) {
  const int64_t site = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (site >= geo.nsite) return;
  const int nyz = geo.ny*geo.nz;
  const int x = (int) (site / nyz);
  const int rem = (int) (site - (int64_t) x*nyz);
  const int y = rem / geo.nz;
  const int z = rem - y*geo.nz;
  const Lat<T> L = {phi, geo.nx, geo.ny, geo.nz};
  const GradPt<T> c0 = grad_at(L, x, y, z, fe.stencil);
  mu[site] = mu_of(fe, c0.phi, c0.delsq);
}

template <typename T>
__global__ void __launch_bounds__(256)
grad_kernel(Geom geo, int stencil, const T * __restrict__ phi, T * __restrict__ grad,
	     T * __restrict__ delsq) {
  const int64_t site = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (site >= geo.nsite) return;
  const int nyz = geo.ny*geo.nz;
  const int x = (int) (site / nyz);
  const int rem = (int) (site - (int64_t) x*nyz);
  const int y = rem / geo.nz;
  const int z = rem - y*geo.nz;
  const Lat<T> L = {phi, geo.nx, geo.ny, geo.nz};
  const GradPt<T> c0 = grad_at(L, x, y, z, stencil);
  grad[site] = c0.g[0];
  grad[geo.nsite + site] = c0.g[1];
  grad[2*geo.nsite + site] = c0.g[2];
  delsq[site] = c0.delsq;
}

// Advective face flux between `lo` and `hi` along one axis, for face
// velocity uf = 1/2 (u_self + u_other) in the reference's operand order.
// Order 1 (advection.c:560-625), 2 (:820-895), 3 (:990-1140).

template <typename T>
__device__ __forceinline__ T adv_first(T uf, T phi_self, T phi_other, bool west) {
  // west face: upwind cell is the lower one when uf > 0;
  // other faces: the upper one when uf < 0
  if (west) return uf*((uf > T(0.0)) ? phi_other : phi_self);
  return uf*((uf < T(0.0)) ? phi_other : phi_self);
}

template <typename T>
__global__ void __launch_bounds__(256)
cahn_hilliard_kernel(Geom geo, FeParams fe, const T * __restrict__ phi,
		     const T * __restrict__ mu, const T * __restrict__ u,
		     T * __restrict__ phi_out, T * __restrict__ csum) {

  const int64_t site = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  if (site >= geo.nsite) return;
  const int nyz = geo.ny*geo.nz;
  const int x = (int) (site / nyz);
  const int rem = (int) (site - (int64_t) x*nyz);
  const int y = rem / geo.nz;
  const int z = rem - y*geo.nz;

  const Lat<T> P = {phi, geo.nx, geo.ny, geo.nz};
  const Lat<T> M = {mu, geo.nx, geo.ny, geo.nz};
  const T mob = T(fe.mobility);
  const T a1 = T(-0.213933), a2 = T(0.927865), a3 = T(0.286067);

  // advective parts of fw(x), fe(x), fy(y), fy(y-1), fz(z), fz(z-1)
  T fw = T(0), fe_ = T(0), fyp = T(0), fym = T(0), fzp = T(0), fzm = T(0);

  if (u != nullptr) {
    const Lat<T> UX = {u, geo.nx, geo.ny, geo.nz};
    const Lat<T> UY = {u + geo.nsite, geo.nx, geo.ny, geo.nz};
    const Lat<T> UZ = {u + 2*geo.nsite, geo.nx, geo.ny, geo.nz};
    const T ux0 = UX.at(x, y, z), uy0 = UY.at(x, y, z), uz0 = UZ.at(x, y, z);
    const T uxm = UX.at(x - 1, y, z), uxp = UX.at(x + 1, y, z);
    const T uyp = UY.at(x, y + 1, z), uym = UY.at(x, y - 1, z);
    const T uzp = UZ.at(x, y, z + 1), uzm = UZ.at(x, y, z - 1);

    // face velocities in the reference's operand order (u of the cell
    // computing the flux first)
    const T uw = T(0.5)*(ux0 + uxm);
    const T ue = T(0.5)*(ux0 + uxp);
    const T uyf = T(0.5)*(uy0 + uyp);
    const T uyb = T(0.5)*(uym + uy0);     // fy at y-1: u(y-1) + u(y)
    const T uzf = T(0.5)*(uz0 + uzp);
    const T uzb = T(0.5)*(uzm + uz0);     // fz at z-1

    const T p0 = P.at(x, y, z);
    if (fe.order == 1) {
      fw = adv_first(uw, p0, P.at(x - 1, y, z), true);
      fe_ = adv_first(ue, p0, P.at(x + 1, y, z), false);
      fyp = adv_first(uyf, p0, P.at(x, y + 1, z), false);
      fym = adv_first(uyb, P.at(x, y - 1, z), p0, false);
      fzp = adv_first(uzf, p0, P.at(x, y, z + 1), false);
      fzm = adv_first(uzb, P.at(x, y, z - 1), p0, false);
    }
    else if (fe.order == 2) {
      // flux = (u_f 1/2)(phi_a + phi_b), operand order as advection.c:837-895
      fw = uw*T(0.5)*(P.at(x - 1, y, z) + p0);
      fe_ = ue*T(0.5)*(p0 + P.at(x + 1, y, z));
      fyp = uyf*T(0.5)*(p0 + P.at(x, y + 1, z));
      fym = uyb*T(0.5)*(P.at(x, y - 1, z) + p0);
      fzp = uzf*T(0.5)*(p0 + P.at(x, y, z + 1));
      fzm = uzb*T(0.5)*(P.at(x, y, z - 1) + p0);
    }
    else if (fe.order == 4) {
      fw = adv::order4(uw, P.at(x - 2, y, z), P.at(x - 1, y, z), p0, P.at(x + 1, y, z));
      fe_ = adv::order4(ue, P.at(x - 1, y, z), p0, P.at(x + 1, y, z), P.at(x + 2, y, z));
      fyp = adv::order4(uyf, P.at(x, y - 1, z), p0, P.at(x, y + 1, z), P.at(x, y + 2, z));
      fym = adv::order4(uyb, P.at(x, y - 2, z), P.at(x, y - 1, z), p0, P.at(x, y + 1, z));
      fzp = adv::order4(uzf, P.at(x, y, z - 1), p0, P.at(x, y, z + 1), P.at(x, y, z + 2));
      fzm = adv::order4(uzb, P.at(x, y, z - 2), P.at(x, y, z - 1), p0, P.at(x, y, z + 1));
    }
    else if (fe.order == 5) {
      fw = adv::order5(uw, uw > T(0.0), P.at(x - 3, y, z), P.at(x - 2, y, z), P.at(x - 1, y, z), p0,
		       P.at(x + 1, y, z), P.at(x + 2, y, z));
      fe_ = adv::order5(ue, !(ue < T(0.0)), P.at(x - 2, y, z), P.at(x - 1, y, z), p0, P.at(x + 1, y, z),
			P.at(x + 2, y, z), P.at(x + 3, y, z));
      fyp = adv::order5(uyf, !(uyf < T(0.0)), P.at(x, y - 2, z), P.at(x, y - 1, z), p0, P.at(x, y + 1, z),
			P.at(x, y + 2, z), P.at(x, y + 3, z));
      fym = adv::order5(uyb, !(uyb < T(0.0)), P.at(x, y - 3, z), P.at(x, y - 2, z), P.at(x, y - 1, z), p0,
			P.at(x, y + 1, z), P.at(x, y + 2, z));
      fzp = adv::order5(uzf, !(uzf < T(0.0)), P.at(x, y, z - 2), P.at(x, y, z - 1), p0, P.at(x, y, z + 1),
			P.at(x, y, z + 2), P.at(x, y, z + 3));
      fzm = adv::order5(uzb, !(uzb < T(0.0)), P.at(x, y, z - 3), P.at(x, y, z - 2), P.at(x, y, z - 1), p0,
			P.at(x, y, z + 1), P.at(x, y, z + 2));
    }
    else {
      T d1, d2, d3;
      // west face of x
      if (uw > T(0.0)) { d1 = P.at(x - 2, y, z); d2 = P.at(x - 1, y, z); d3 = p0; }
      else             { d1 = P.at(x + 1, y, z); d2 = p0; d3 = P.at(x - 1, y, z); }
      fw = uw*(a1*d1 + a2*d2 + a3*d3);
      // east face of x
      if (ue < T(0.0)) { d1 = P.at(x + 2, y, z); d2 = P.at(x + 1, y, z); d3 = p0; }
      else             { d1 = P.at(x - 1, y, z); d2 = p0; d3 = P.at(x + 1, y, z); }
      fe_ = ue*(a1*d1 + a2*d2 + a3*d3);
      // y face of y (between y, y+1)
      if (uyf < T(0.0)) { d1 = P.at(x, y + 2, z); d2 = P.at(x, y + 1, z); d3 = p0; }
      else              { d1 = P.at(x, y - 1, z); d2 = p0; d3 = P.at(x, y + 1, z); }
      fyp = uyf*(a1*d1 + a2*d2 + a3*d3);
      // y face of y-1 (between y-1, y), computed as at cell y-1
      if (uyb < T(0.0)) { d1 = P.at(x, y + 1, z); d2 = p0; d3 = P.at(x, y - 1, z); }
      else              { d1 = P.at(x, y - 2, z); d2 = P.at(x, y - 1, z); d3 = p0; }
      fym = uyb*(a1*d1 + a2*d2 + a3*d3);
      // z faces
      if (uzf < T(0.0)) { d1 = P.at(x, y, z + 2); d2 = P.at(x, y, z + 1); d3 = p0; }
      else              { d1 = P.at(x, y, z - 1); d2 = p0; d3 = P.at(x, y, z + 1); }
      fzp = uzf*(a1*d1 + a2*d2 + a3*d3);
      if (uzb < T(0.0)) { d1 = P.at(x, y, z + 1); d2 = p0; d3 = P.at(x, y, z - 1); }
      else              { d1 = P.at(x, y, z - 2); d2 = P.at(x, y, z - 1); d3 = p0; }
      fzm = uzb*(a1*d1 + a2*d2 + a3*d3);
    }
  }

  // diffusive parts: flux -= M (mu_upper - mu_lower) (phi_cahn_hilliard.c:383-407)
  {
    const T m0 = M.at(x, y, z);
    fw -= mob*(m0 - M.at(x - 1, y, z));
    fe_ -= mob*(M.at(x + 1, y, z) - m0);
    fyp -= mob*(M.at(x, y + 1, z) - m0);
    fym -= mob*(m0 - M.at(x, y - 1, z));
    fzp -= mob*(M.at(x, y, z + 1) - m0);
    fzm -= mob*(m0 - M.at(x, y, z - 1));
  }
  if (fe.gmu_on) adv::mu_ext(mob, fe.gmu, fw, fe_, fyp, fym, fzp, fzm);

  // phi_ch_ufs_kernel (phi_cahn_hilliard.c:1039-1044), wz = 1, or the
  // compensated update of conserve 1 (phi_cahn_hilliard.c:1181-1218)
  T p = P.at(x, y, z);
  if (csum != nullptr) {
    T cs = csum[site];
    ch_update_compensated(p, cs, fe_, fw, fyp, fym, fzp, fzm);
    csum[site] = cs;
  }
  else {
    p -= (+ fe_ - fw + fyp - fym + fzp - fzm);
  }
  phi_out[site] = p;
}

template <typename T>
__global__ void fill_kernel(T * a, int64_t n, T v) {
  int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x;
  int64_t stride = (int64_t) gridDim.x*blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

inline unsigned nblocks(int64_t n) { return (unsigned) ((n + 255)/256); }

}  // namespace

void launch_phi_force_mu_f64(const Geom & g, const FeParams & fe,
			     const double * phi, double * force, double * mu,
			     int force_mode, const double fzero[3],
			     hipStream_t s) {
  hipLaunchKernelGGL(phi_force_mu_kernel<double>, dim3(nblocks(g.nsite)), dim3(256),
		     0, s, g, fe, phi, force, mu, force_mode, fzero[0], fzero[1], fzero[2]);
}

void launch_phi_force_mu_f32(const Geom & g, const FeParams & fe,
			     const float * phi, float * force, float * mu,
			     int force_mode, const double fzero[3],
			     hipStream_t s) {
  hipLaunchKernelGGL(phi_force_mu_kernel<float>, dim3(nblocks(g.nsite)), dim3(256),
		     0, s, g, fe, phi, force, mu, force_mode, (float) fzero[0],
		     (float) fzero[1], (float) fzero[2]);
}

void launch_mu_f64(const Geom & g, const FeParams & fe, const double * phi,
		   double * mu, hipStream_t s) {
  hipLaunchKernelGGL(mu_kernel<double>, dim3(nblocks(g.nsite)), dim3(256), 0, s,
		     g, fe, phi, mu);
}

void launch_mu_f32(const Geom & g, const FeParams & fe, const float * phi,
		   float * mu, hipStream_t s) {
  hipLaunchKernelGGL(mu_kernel<float>, dim3(nblocks(g.nsite)), dim3(256), 0, s,
		     g, fe, phi, mu);
}

void launch_grad_f64(const Geom & g, int stencil, const double * phi, double * grad,
		     double * delsq, hipStream_t s) {
  hipLaunchKernelGGL(grad_kernel<double>, dim3(nblocks(g.nsite)), dim3(256), 0, s,
		     g, stencil, phi, grad, delsq);
}

void launch_cahn_hilliard_f64(const Geom & g, const FeParams & fe,
			      const double * phi, const double * mu,
			      const double * u, double * phi_out, double * csum,
			      hipStream_t s) {
  hipLaunchKernelGGL(cahn_hilliard_kernel<double>, dim3(nblocks(g.nsite)), dim3(256),
		     0, s, g, fe, phi, mu, u, phi_out, csum);
}

void launch_cahn_hilliard_f32(const Geom & g, const FeParams & fe,
			      const float * phi, const float * mu,
			      const float * u, float * phi_out, float * csum,
			      hipStream_t s) {
  hipLaunchKernelGGL(cahn_hilliard_kernel<float>, dim3(nblocks(g.nsite)), dim3(256),
		     0, s, g, fe, phi, mu, u, phi_out, csum);
}

void launch_fill_f64(double * a, int64_t n, double v, hipStream_t s) {
  unsigned nb = nblocks(n); if (nb > 65536) nb = 65536;
  if (nb == 0) return;
  hipLaunchKernelGGL(fill_kernel<double>, dim3(nb), dim3(256), 0, s, a, n, v);
}

void launch_fill_f32(float * a, int64_t n, float v, hipStream_t s) {
  unsigned nb = nblocks(n); if (nb > 65536) nb = 65536;
  if (nb == 0) return;
  hipLaunchKernelGGL(fill_kernel<float>, dim3(nb), dim3(256), 0, s, a, n, v);
}

}  // namespace ldw
