/*****************************************************************************
 *
 *  le_oracle.c -- CPU restatement of the Lees-Edwards sliding planes
 *
 *  TEST INFRASTRUCTURE ONLY (see ldw_oracle.h).  It is the checker for the
 *  GPU library's plane code; the product never calls it.
 *
 *  The reference extends its x index past the real planes (and their
 *  halo) with 2 nhalo buffer planes per Lees-Edwards plane
 *  (leesedwards.c:430-455).  Here every field that has buffer values is
 *  one array of nx + nxbuf planes: x = 0 .. nx-1 real, x = nx + ib the
 *  buffer plane ib.  y and z are periodic by index wrap, which gives the
 *  values the reference's halos hold on one rank.  The functions follow
 *  the reference's loops and operation order:
 *
 *    phi_buffer     field_leesedwards            field.c:418-510
 *    gradient       grad_3d_7pt_fluid_operator   gradient_3d_7pt_fluid.c:231-310
 *                   grad_3d_7pt_fluid_le         gradient_3d_7pt_fluid.c:330-430
 *                   (27-point: gradient_3d_27pt_fluid.c:219-368, 400-560)
 *    force          phi_force_flux               phi_force.c:289-520
 *    u_buffer       hydro_lees_edwards           hydro.c:350-437
 *    cahn_hilliard  advection_le_1st/2nd/3rd     advection.c:538-1145
 *                   phi_ch_flux_mu1_kernel       phi_cahn_hilliard.c:350-413
 *                   phi_ch_le_fix_fluxes         phi_cahn_hilliard.c:613-730
 *                   forward step / conserve      phi_cahn_hilliard.c:1018-1218
 *    lb_bc          lb_data_apply_le_boundary_conditions  model_le.c:78-625
 *    shear_profile  lb_le_init_shear_profile     model_le.c:649-715
 *
 *  Pinned by: the golden vectors of the reference's own build with planes
 *  (tests/golden/le_*.npz) and closed-form checks in
 *  tests/test_oracle_lees_edwards.py.
 *
 *****************************************************************************/

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ldw_oracle.h"

#define NVEL 19
#define LE_MAXP 64

static const int cv_le[NVEL][3] = {
  { 0,  0,  0},
  { 1,  1,  0}, { 1,  0,  1}, { 1,  0,  0},
  { 1,  0, -1}, { 1, -1,  0}, { 0,  1,  1},
  { 0,  1,  0}, { 0,  1, -1}, { 0,  0,  1},
  { 0,  0, -1}, { 0, -1,  1}, { 0, -1,  0},
  { 0, -1, -1}, {-1,  1,  0}, {-1,  0,  1},
  {-1,  0,  0}, {-1,  0, -1}, {-1, -1,  0}};

/* LB_WEIGHTS_D3Q19 (lb_d3q19.h:35) */
static double wv_le(int p) {
  static const int w36[NVEL] = {12, 1, 1, 2, 1, 1, 1, 2, 1, 2, 2, 1, 2, 1, 1, 1, 2, 1, 1};
  return w36[p]/36.0;
}

typedef struct {
  int nx, ny, nz;
  int nh;               /* nhalo */
  int nplane;
  int nxb;              /* buffer planes */
  int loc[LE_MAXP];     /* lees_edw_plane_location, 1-based */
  double dx_min, dx_sep;
  double uy, time0;
} le_t;

static int le_init(le_t * g, int nx, int ny, int nz, const ora_le_param_t * p) {
  if (p->nplane < 1 || p->nplane > LE_MAXP || nx % p->nplane) return -1;
  g->nx = nx; g->ny = ny; g->nz = nz;
  g->nh = p->nhalo;
  g->nplane = p->nplane;
  g->nxb = 2*p->nhalo*p->nplane;
  g->dx_sep = 1.0*nx/p->nplane;
  g->dx_min = 0.5*g->dx_sep;
  g->uy = p->uy;
  g->time0 = 1.0*p->nt0;
  for (int np = 0; np < p->nplane; np++) {
    g->loc[np] = g->dx_min + np*g->dx_sep;
    if (g->loc[np] <= g->nh || g->loc[np] > nx - g->nh) return -1;
  }
  return 0;
}

static inline int wrp(int i, int n) { i %= n; return (i < 0) ? i + n : i; }

static inline long idx(const le_t * g, int xe, int y, int z) {
  return ((long) xe*g->ny + wrp(y, g->ny))*g->nz + wrp(z, g->nz);
}

/* lees_edw_ic_to_buff, 1-based in and out */
static int ic_to_buff(const le_t * g, int ic, int di) {
  int ib = ic + di;
  int p = ic/(g->nx/g->nplane);
  if (p < 0) p = 0;
  if (p > g->nplane - 1) p = g->nplane - 1;
  int ip = g->loc[p] - (g->nh - 1);
  if (di > 0 && (ic >= ip && ic < ip + g->nh) && (ic + di >= ip + g->nh)) {
    return g->nx + (1 + 2*p)*g->nh + (ic - ip + 1) + di;
  }
  ip = g->loc[p] + 1;
  if (di < 0 && (ic >= ip && ic < ip + g->nh) && (ic + di < ip)) {
    return g->nx + (2 + 2*p)*g->nh + (ic - ip + 1) + di;
  }
  return ib;
}

/* the extended-array x of a 1-based x: buffer plane ib sits at
 * lees_edw_index x = nx + nhalo + 1 + ib */
static int xe_of(const le_t * g, int xr) {
  if (xr >= g->nx + g->nh + 1) return g->nx + (xr - (g->nx + g->nh + 1));
  return wrp(xr - 1, g->nx);
}

/* lees_edw_ibuff_to_real */
static int ibuff_to_real(const le_t * g, int ib) {
  return g->loc[ib/(2*g->nh)] - (g->nh - 1) + ib % (2*g->nh);
}

/* lees_edw_buffer_duy */
static int buffer_duy(const le_t * g, int ib) {
  return (ib % (2*g->nh) < g->nh) ? -1 : +1;
}

/* lees_edw_buffer_displacement, steady shear */
static double buffer_displacement(const le_t * g, int ib, double t) {
  if (t < 0.0) t = 0.0;
  double tle = t - g->time0;
  return tle*g->uy*buffer_duy(g, ib);
}

/* field_leesedwards (field.c:459-500): phi of the buffer planes */
static void phi_buffer(const le_t * g, double t, double * pe) {
  const double r6 = (1.0/6.0);
  const int N = g->ny;
  for (int ib = 0; ib < g->nxb; ib++) {
    int ic = ibuff_to_real(g, ib);
    double dy = buffer_displacement(g, ib, t + 0.0);
    dy = fmod(dy, (double) N);
    int jdy = floor(dy);
    double fr = 1.0 - (dy - jdy);
    for (int jc = 1; jc <= N; jc++) {
      int j0 = 1 + (jc - jdy - 3 + 2*N) % N;
      int j1 = 1 + j0 % N;
      int j2 = 1 + j1 % N;
      int j3 = 1 + j2 % N;
      for (int kc = 1; kc <= g->nz; kc++) {
	pe[idx(g, g->nx + ib, jc - 1, kc - 1)] =
	  -  r6*fr*(fr-1.0)*(fr-2.0)*pe[idx(g, ic - 1, j0 - 1, kc - 1)]
	  + 0.5*(fr*fr-1.0)*(fr-2.0)*pe[idx(g, ic - 1, j1 - 1, kc - 1)]
	  - 0.5*fr*(fr+1.0)*(fr-2.0)*pe[idx(g, ic - 1, j2 - 1, kc - 1)]
	  +        r6*fr*(fr*fr-1.0)*pe[idx(g, ic - 1, j3 - 1, kc - 1)];
      }
    }
  }
}

/* gradient and Laplacian at (x0, y, z) with x neighbour planes xm, xp */
static void grad_at(const le_t * g, int stencil, const double * pe, int xm, int x0, int xp,
		    int y, int z, double gr[3], double * d) {
  if (stencil == 27) {
    /* grad_3d_27pt_kernel (gradient_3d_27pt_fluid.c:266-365) */
    const double r9 = (1.0/9.0);
    int xs[3] = {xm, x0, xp};
    double p[3][3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
	for (int k = 0; k < 3; k++) p[i][j][k] = pe[idx(g, xs[i], y + j - 1, z + k - 1)];
    gr[0] = 0.5*r9*
      (+ p[2][0][0] - p[0][0][0] + p[2][0][1] - p[0][0][1] + p[2][0][2] - p[0][0][2]
       + p[2][1][0] - p[0][1][0] + p[2][1][1] - p[0][1][1] + p[2][1][2] - p[0][1][2]
       + p[2][2][0] - p[0][2][0] + p[2][2][1] - p[0][2][1] + p[2][2][2] - p[0][2][2]);
    gr[1] = 0.5*r9*
      (+ p[0][2][0] - p[0][0][0] + p[0][2][1] - p[0][0][1] + p[0][2][2] - p[0][0][2]
       + p[1][2][0] - p[1][0][0] + p[1][2][1] - p[1][0][1] + p[1][2][2] - p[1][0][2]
       + p[2][2][0] - p[2][0][0] + p[2][2][1] - p[2][0][1] + p[2][2][2] - p[2][0][2]);
    gr[2] = 0.5*r9*
      (+ p[0][0][2] - p[0][0][0] + p[0][1][2] - p[0][1][0] + p[0][2][2] - p[0][2][0]
       + p[1][0][2] - p[1][0][0] + p[1][1][2] - p[1][1][0] + p[1][2][2] - p[1][2][0]
       + p[2][0][2] - p[2][0][0] + p[2][1][2] - p[2][1][0] + p[2][2][2] - p[2][2][0]);
    *d = r9*(+ p[0][0][0] + p[0][0][1] + p[0][0][2] + p[0][1][0] + p[0][1][1] + p[0][1][2]
	     + p[0][2][0] + p[0][2][1] + p[0][2][2] + p[1][0][0] + p[1][0][1] + p[1][0][2]
	     + p[1][1][0] + p[1][1][2] + p[1][2][0] + p[1][2][1] + p[1][2][2]
	     + p[2][0][0] + p[2][0][1] + p[2][0][2] + p[2][1][0] + p[2][1][1] + p[2][1][2]
	     + p[2][2][0] + p[2][2][1] + p[2][2][2]
	     - 26.0*p[1][1][1]);
    return;
  }
  const double vp = pe[idx(g, xp, y, z)], vm = pe[idx(g, xm, y, z)];
  const double yp = pe[idx(g, x0, y + 1, z)], ym = pe[idx(g, x0, y - 1, z)];
  const double zp = pe[idx(g, x0, y, z + 1)], zm = pe[idx(g, x0, y, z - 1)];
  gr[0] = 0.5*(vp - vm);
  gr[1] = 0.5*(yp - ym);
  gr[2] = 0.5*(zp - zm);
  *d = vp + vm + yp + ym + zp + zm - 6.0*pe[idx(g, x0, y, z)];
}

/* field_grad_compute with planes: buffer phi at time t, then the real
 * sites, then the buffer sites next to each plane */
static void gradient(const le_t * g, int stencil, double t, double * pe, double * ge,
		     double * de) {
  const long ne = (long) (g->nx + g->nxb)*g->ny*g->nz;
  phi_buffer(g, t, pe);
  for (int ic = 1; ic <= g->nx; ic++) {
    int xm = xe_of(g, ic_to_buff(g, ic, -1));
    int xp = xe_of(g, ic_to_buff(g, ic, +1));
    for (int y = 0; y < g->ny; y++)
      for (int z = 0; z < g->nz; z++) {
	double gr[3], d;
	grad_at(g, stencil, pe, xm, ic - 1, xp, y, z, gr, &d);
	long s = idx(g, ic - 1, y, z);
	for (int a = 0; a < 3; a++) ge[a*ne + s] = gr[a];
	de[s] = d;
      }
  }
  const int nextra = g->nh - 1;
  for (int np = 0; np < g->nplane; np++) {
    int ic = g->loc[np];
    for (int nh = 1; nh <= nextra; nh++) {
      int x0 = xe_of(g, ic_to_buff(g, ic, nh - 1));
      int x1 = xe_of(g, ic_to_buff(g, ic, nh));
      int x2 = xe_of(g, ic_to_buff(g, ic, nh + 1));
      for (int y = 0; y < g->ny; y++)
	for (int z = 0; z < g->nz; z++) {
	  double gr[3], d;
	  grad_at(g, stencil, pe, x0, x1, x2, y, z, gr, &d);
	  for (int a = 0; a < 3; a++) ge[a*ne + idx(g, x1, y, z)] = gr[a];
	  de[idx(g, x1, y, z)] = d;
	}
    }
    ic += 1;
    for (int nh = 1; nh <= nextra; nh++) {
      int x2 = xe_of(g, ic_to_buff(g, ic, -nh + 1));
      int x1 = xe_of(g, ic_to_buff(g, ic, -nh));
      int x0 = xe_of(g, ic_to_buff(g, ic, -nh - 1));
      for (int y = 0; y < g->ny; y++)
	for (int z = 0; z < g->nz; z++) {
	  double gr[3], d;
	  grad_at(g, stencil, pe, x0, x1, x2, y, z, gr, &d);
	  for (int a = 0; a < 3; a++) ge[a*ne + idx(g, x1, y, z)] = gr[a];
	  de[idx(g, x1, y, z)] = d;
	}
    }
  }
}

/* fe_symm_str (symmetric.c:334-360) at extended site s */
static void stress(const ora_fe_param_t * fe, const double * pe, const double * ge,
		   const double * de, long ne, long s, double p[3][3]) {
  double phi = pe[s], delsq = de[s];
  double gp[3] = {ge[s], ge[ne + s], ge[2*ne + s]};
  double kappa = fe->kappa;
  double p0 = 0.5*fe->a*phi*phi + 0.75*fe->b*phi*phi*phi*phi
    - kappa*phi*delsq - 0.5*kappa*(gp[0]*gp[0] + gp[1]*gp[1] + gp[2]*gp[2]);
  for (int ia = 0; ia < 3; ia++)
    for (int ib = 0; ib < 3; ib++) {
      double d_ab = (ia == ib);
      p[ia][ib] = p0*d_ab + kappa*gp[ia]*gp[ib];
    }
}

static void alloc_ext(const le_t * g, double ** pe, double ** ge, double ** de) {
  const long ne = (long) (g->nx + g->nxb)*g->ny*g->nz;
  *pe = (double *) calloc(ne, sizeof(double));
  *ge = (double *) calloc(3*ne, sizeof(double));
  *de = (double *) calloc(ne, sizeof(double));
}

/* phi_force_flux (phi_force.c:289-520): F = 0 + divergence of the face
 * fluxes of the chemical stress, plane fluxes corrected per plane */
static void force(const le_t * g, const ora_fe_param_t * fe, const double * pe, const double * ge,
		  const double * de, double * F) {
  const long ne = (long) (g->nx + g->nxb)*g->ny*g->nz;
  const long ns = (long) g->nx*g->ny*g->nz;
  double * fl = (double *) malloc(sizeof(double)*12*ns);
  double * fxe = fl, * fxw = fl + 3*ns, * fxy = fl + 6*ns, * fxz = fl + 9*ns;
  for (int ic = 1; ic <= g->nx; ic++) {
    int icm1 = xe_of(g, ic_to_buff(g, ic, -1));
    int icp1 = xe_of(g, ic_to_buff(g, ic, +1));
    for (int y = 0; y < g->ny; y++)
      for (int z = 0; z < g->nz; z++) {
	long s = idx(g, ic - 1, y, z);
	double p0[3][3], p1[3][3];
	stress(fe, pe, ge, de, ne, s, p0);
	stress(fe, pe, ge, de, ne, idx(g, icm1, y, z), p1);
	for (int a = 0; a < 3; a++) fxw[a*ns + s] = 0.5*(p1[a][0] + p0[a][0]);
	stress(fe, pe, ge, de, ne, idx(g, icp1, y, z), p1);
	for (int a = 0; a < 3; a++) fxe[a*ns + s] = 0.5*(p1[a][0] + p0[a][0]);
	stress(fe, pe, ge, de, ne, idx(g, ic - 1, y + 1, z), p1);
	for (int a = 0; a < 3; a++) fxy[a*ns + s] = 0.5*(p1[a][1] + p0[a][1]);
	stress(fe, pe, ge, de, ne, idx(g, ic - 1, y, z + 1), p1);
	for (int a = 0; a < 3; a++) fxz[a*ns + s] = 0.5*(p1[a][2] + p0[a][2]);
      }
  }
  /* phi_force_flux_fix_local */
  {
    double fbar[3*LE_MAXP] = {0};
    const double ra = 0.5/((double) g->ny*(double) g->nz);
    for (int np = 0; np < g->nplane; np++) {
      int ic = g->loc[np];
      for (int y = 0; y < g->ny; y++)
	for (int z = 0; z < g->nz; z++)
	  for (int a = 0; a < 3; a++) {
	    fbar[3*np + a] += - fxe[a*ns + idx(g, ic - 1, y, z)] + fxw[a*ns + idx(g, ic, y, z)];
	  }
    }
    for (int np = 0; np < g->nplane; np++) {
      int ic = g->loc[np];
      for (int y = 0; y < g->ny; y++)
	for (int z = 0; z < g->nz; z++)
	  for (int a = 0; a < 3; a++) {
	    fxe[a*ns + idx(g, ic - 1, y, z)] += ra*fbar[3*np + a];
	    fxw[a*ns + idx(g, ic, y, z)] -= ra*fbar[3*np + a];
	  }
    }
  }
  /* phi_force_flux_divergence; hydro_f_local_add onto a zeroed force */
  for (int x = 0; x < g->nx; x++)
    for (int y = 0; y < g->ny; y++)
      for (int z = 0; z < g->nz; z++) {
	long s = idx(g, x, y, z), sj = idx(g, x, y - 1, z), sk = idx(g, x, y, z - 1);
	for (int a = 0; a < 3; a++) {
	  double f = -(+ fxe[a*ns + s] - fxw[a*ns + s] + fxy[a*ns + s] - fxy[a*ns + sj]
		       + fxz[a*ns + s] - fxz[a*ns + sk]);
	  F[a*ns + s] = 0.0 + f;
	}
      }
  free(fl);
}

/* hydro_lees_edwards (hydro.c:390-430): u of the buffer planes, time t + 1 */
static void u_buffer(const le_t * g, double t, double * ue) {
  const long ne = (long) (g->nx + g->nxb)*g->ny*g->nz;
  const int N = g->ny;
  for (int ib = 0; ib < g->nxb; ib++) {
    int ic = ibuff_to_real(g, ib);
    double ule[3] = {0.0, g->uy*buffer_duy(g, ib), 0.0};
    double dy = buffer_displacement(g, ib, t + 1.0);
    dy = fmod(dy, (double) N);
    int jdy = floor(dy);
    double fr = dy - jdy;
    for (int jc = 1; jc <= N; jc++) {
      int j1 = 1 + (jc - jdy - 2 + 2*N) % N;
      int j2 = 1 + j1 % N;
      for (int kc = 1; kc <= g->nz; kc++)
	for (int a = 0; a < 3; a++) {
	  ue[a*ne + idx(g, g->nx + ib, jc - 1, kc - 1)] = ule[a]
	    + ue[a*ne + idx(g, ic - 1, j1 - 1, kc - 1)]*fr
	    + ue[a*ne + idx(g, ic - 1, j2 - 1, kc - 1)]*(1.0 - fr);
	}
    }
  }
}

static double mu_at(const ora_fe_param_t * fe, const double * pe, const double * de, long s) {
  return fe->a*pe[s] + fe->b*pe[s]*pe[s]*pe[s] - fe->kappa*de[s];
}

static void kahan_add_le(double * sum, double * cs, double val) {
  double y = val + *cs;
  double t = *sum + y;
  *cs = y - (t - *sum);
  *sum = t;
}

/* phi_cahn_hilliard with planes: face fluxes at the real sites, the plane
 * faces averaged, then the update.  u may be NULL (no hydrodynamics). */
static void cahn_hilliard(const le_t * g, const ora_fe_param_t * fe, double t, double * pe,
			  const double * de, double * ue, double * phi_out, double * csum,
			  double phi_init_sum) {
  const long ne = (long) (g->nx + g->nxb)*g->ny*g->nz;
  const long ns = (long) g->nx*g->ny*g->nz;
  const double a1 = -0.213933, a2 = 0.927865, a3 = 0.286067;
  const double mob = fe->mobility;
  double * fl = (double *) calloc(4*ns, sizeof(double));
  double * fw = fl, * fe_ = fl + ns, * fy = fl + 2*ns, * fz = fl + 3*ns;
  if (ue) u_buffer(g, t, ue);
  for (int ic = 1; ic <= g->nx; ic++) {
    int m1 = xe_of(g, ic_to_buff(g, ic, -1)), p1 = xe_of(g, ic_to_buff(g, ic, +1));
    int m2 = xe_of(g, ic_to_buff(g, ic, -2)), p2 = xe_of(g, ic_to_buff(g, ic, +2));
    int x = ic - 1;
    for (int y = 0; y < g->ny; y++)
      for (int z = 0; z < g->nz; z++) {
	long s = idx(g, x, y, z);
	long sxm1 = idx(g, m1, y, z), sxp1 = idx(g, p1, y, z);
	long sxm2 = idx(g, m2, y, z), sxp2 = idx(g, p2, y, z);
	long sym1 = idx(g, x, y - 1, z), syp1 = idx(g, x, y + 1, z), syp2 = idx(g, x, y + 2, z);
	long szm1 = idx(g, x, y, z - 1), szp1 = idx(g, x, y, z + 1), szp2 = idx(g, x, y, z + 2);
	double w = 0.0, e = 0.0, fyv = 0.0, fzv = 0.0;
	if (ue) {
	  const double * ux = ue, * uy = ue + ne, * uz = ue + 2*ne;
	  double uf;
	  if (fe->order == 1) {
	    uf = 0.5*(ux[s] + ux[sxm1]);
	    w = uf*pe[(uf > 0.0) ? sxm1 : s];
	    uf = 0.5*(ux[s] + ux[sxp1]);
	    e = uf*pe[(uf < 0.0) ? sxp1 : s];
	    uf = 0.5*(uy[s] + uy[syp1]);
	    fyv = uf*pe[(uf < 0.0) ? syp1 : s];
	    uf = 0.5*(uz[s] + uz[szp1]);
	    fzv = uf*pe[(uf < 0.0) ? szp1 : s];
	  }
	  else if (fe->order == 2) {
	    w = 0.5*(ux[s] + ux[sxm1])*1*0.5*(pe[sxm1] + pe[s]);
	    e = 0.5*(ux[s] + ux[sxp1])*1*0.5*(pe[s] + pe[sxp1]);
	    fyv = 0.5*(uy[s] + uy[syp1])*1*0.5*(pe[s] + pe[syp1]);
	    fzv = 0.5*(uz[s] + uz[szp1])*1*0.5*(pe[s] + pe[szp1]);
	  }
	  else if (fe->order == 4) {
	    /* advection_le_4th (advection.c:1153-1262) */
	    const double b1 = 1.0/16.0, b2 = 9.0/16.0;
	    long sym1_ = idx(g, x, y - 1, z), szm1_ = idx(g, x, y, z - 1);
	    uf = 0.5*(ux[s] + ux[sxm1]);
	    w = uf*(- b1*pe[sxm2] + b2*pe[sxm1] + b2*pe[s] - b1*pe[sxp1]);
	    uf = 0.5*(ux[s] + ux[sxp1]);
	    e = uf*(- b1*pe[sxm1] + b2*pe[s] + b2*pe[sxp1] - b1*pe[sxp2]);
	    uf = 0.5*(uy[s] + uy[syp1]);
	    fyv = uf*(- b1*pe[sym1_] + b2*pe[s] + b2*pe[syp1] - b1*pe[syp2]);
	    uf = 0.5*(uz[s] + uz[szp1]);
	    fzv = uf*(- b1*pe[szm1_] + b2*pe[s] + b2*pe[szp1] - b1*pe[szp2]);
	  }
	  else {
	    double d1, d2, d3;
	    uf = 0.5*1*(ux[s] + ux[sxm1]);
	    if (uf > 0.0) { d1 = pe[sxm2]; d2 = pe[sxm1]; d3 = pe[s]; }
	    else          { d1 = pe[sxp1]; d2 = pe[s];    d3 = pe[sxm1]; }
	    w = uf*(a1*d1 + a2*d2 + a3*d3);
	    uf = 0.5*1*(ux[s] + ux[sxp1]);
	    if (uf < 0.0) { d1 = pe[sxp2]; d2 = pe[sxp1]; d3 = pe[s]; }
	    else          { d1 = pe[sxm1]; d2 = pe[s];    d3 = pe[sxp1]; }
	    e = uf*(a1*d1 + a2*d2 + a3*d3);
	    uf = 0.5*1*(uy[s] + uy[syp1]);
	    if (uf < 0.0) { d1 = pe[syp2]; d2 = pe[syp1]; d3 = pe[s]; }
	    else          { d1 = pe[sym1]; d2 = pe[s];    d3 = pe[syp1]; }
	    fyv = uf*(a1*d1 + a2*d2 + a3*d3);
	    uf = 0.5*1*(uz[s] + uz[szp1]);
	    if (uf < 0.0) { d1 = pe[szp2]; d2 = pe[szp1]; d3 = pe[s]; }
	    else          { d1 = pe[szm1]; d2 = pe[s];    d3 = pe[szp1]; }
	    fzv = uf*(a1*d1 + a2*d2 + a3*d3);
	  }
	}
	double mu0 = mu_at(fe, pe, de, s);
	w -= mob*(mu0 - mu_at(fe, pe, de, sxm1));
	e -= mob*(mu_at(fe, pe, de, sxp1) - mu0);
	fyv -= mob*(mu_at(fe, pe, de, syp1) - mu0);
	fzv -= mob*(mu_at(fe, pe, de, szp1) - mu0);
	/* phi_ch_flux_mu_ext, before the plane faces are fixed */
	w -= mob*fe->grad_mu[0];
	e -= mob*fe->grad_mu[0];
	fyv -= mob*fe->grad_mu[1];
	fzv -= mob*fe->grad_mu[2];
	long r = ((long) x*g->ny + y)*g->nz + z;
	fw[r] = w; fe_[r] = e; fy[r] = fyv; fz[r] = fzv;
      }
  }
  /* phi_ch_le_fix_fluxes, serial branch */
  {
    const int N = g->ny;
    const long nyz = (long) g->ny*g->nz;
    double * bw = (double *) malloc(sizeof(double)*nyz);
    double * be = (double *) malloc(sizeof(double)*nyz);
    double dy = t*g->uy;                           /* lees_edw_plane_dy */
    for (int np = 0; np < g->nplane; np++) {
      int ic = g->loc[np];
      double d = fmod(+dy, (double) N);
      int jdy = floor(d);
      double fr = d - jdy;
      for (int jc = 1; jc <= N; jc++) {
	int j1 = 1 + (jc - jdy - 2 + 2*N) % N;
	int j2 = 1 + j1 % N;
	for (int kc = 1; kc <= g->nz; kc++) {
	  bw[(jc - 1)*g->nz + kc - 1] = fw[((long) ic*g->ny + j1 - 1)*g->nz + kc - 1]*fr
	    + fw[((long) ic*g->ny + j2 - 1)*g->nz + kc - 1]*(1.0 - fr);
	}
      }
      d = fmod(-dy, (double) N);
      jdy = floor(d);
      fr = d - jdy;
      for (int jc = 1; jc <= N; jc++) {
	int j1 = 1 + (jc - jdy - 2 + 2*N) % N;
	int j2 = 1 + j1 % N;
	for (int kc = 1; kc <= g->nz; kc++) {
	  be[(jc - 1)*g->nz + kc - 1] = fe_[((long) (ic - 1)*g->ny + j1 - 1)*g->nz + kc - 1]*fr
	    + fe_[((long) (ic - 1)*g->ny + j2 - 1)*g->nz + kc - 1]*(1.0 - fr);
	}
      }
      for (long k = 0; k < nyz; k++) {
	long i0 = (long) (ic - 1)*nyz + k, i1 = (long) ic*nyz + k;
	fe_[i0] = 0.5*(fe_[i0] + bw[k]);
	fw[i1] = 0.5*(fw[i1] + be[k]);
      }
    }
    free(be);
    free(bw);
  }
  /* phi_ch_ufs_kernel / phi_ch_csum_kernel */
  for (int x = 0; x < g->nx; x++)
    for (int y = 0; y < g->ny; y++)
      for (int z = 0; z < g->nz; z++) {
	long r = ((long) x*g->ny + y)*g->nz + z;
	long rj = ((long) x*g->ny + wrp(y - 1, g->ny))*g->nz + z;
	long rk = ((long) x*g->ny + y)*g->nz + wrp(z - 1, g->nz);
	const double wz = 1.0;
	if (fe->conserve == 1) {
	  double sum = pe[idx(g, x, y, z)], cs = csum[r];
	  kahan_add_le(&sum, &cs, -fe_[r]);
	  kahan_add_le(&sum, &cs,  fw[r]);
	  kahan_add_le(&sum, &cs, -fy[r]);
	  kahan_add_le(&sum, &cs,  fy[rj]);
	  kahan_add_le(&sum, &cs, -wz*fz[r]);
	  kahan_add_le(&sum, &cs,  wz*fz[rk]);
	  csum[r] = cs;
	  phi_out[r] = sum;
	}
	else {
	  phi_out[r] = pe[idx(g, x, y, z)]
	    - (+ fe_[r] - fw[r] + fy[r] - fy[rj] + wz*fz[r] - wz*fz[rk]);
	}
      }
  if (fe->conserve == 2) {
    double s = 0.0;
    for (long i = 0; i < ns; i++) s += phi_out[i];
    for (long i = 0; i < ns; i++) phi_out[i] -= (s - phi_init_sum)/(int) ns;
  }
  free(fl);
}

/* lb_data_apply_le_boundary_conditions (model_le.c:78-625) at time step ts:
 * reprojection into a send buffer, displacement into a receive buffer of
 * ny + 1 rows, linear interpolation back into f */
static void lb_bc(const le_t * g, double ts, double * f) {
  const long ns = (long) g->nx*g->ny*g->nz;
  const int N = g->ny, nz = g->nz;
  /* the crossing velocities, cx = +1 and cx = -1, in p order (5 each) */
  const int nprop = 5;
  int prop[2][5];
  for (int p = 1, i0 = 0, i1 = 0; p < NVEL; p++) {
    if (cv_le[p][0] == +1) prop[0][i0++] = p;
    if (cv_le[p][0] == -1) prop[1][i1++] = p;
  }
  const long nrow = (long) nz*5;
  double * sbuf = (double *) malloc(sizeof(double)*2*g->nplane*N*nrow);
  double * rbuf = (double *) malloc(sizeof(double)*2*g->nplane*(N + 1)*nrow);
  const double cs2 = (1.0/3.0), rcs2 = 1.0/cs2;
  for (int np = 0; np < g->nplane; np++)
    for (int side = 0; side < 2; side++) {
      int cx = 1 - 2*side;
      int ic = side + g->loc[np];
      double du[3] = {0.0, g->uy, 0.0};                 /* lees_edw_plane_uy_now */
      du[1] = -1.0*cx*du[1];
      double *sb = sbuf + (long) (2*np + side)*N*nrow;
      for (int jc = 1; jc <= N; jc++)
	for (int kc = 1; kc <= nz; kc++) {
	  long s = ((long) (ic - 1)*N + jc - 1)*nz + kc - 1;
	  double rho = 0.0, gm[3] = {0.0, 0.0, 0.0}, ds[3][3];
	  for (int p = 0; p < NVEL; p++) rho += f[p*ns + s];
	  for (int p = 0; p < NVEL; p++)
	    for (int n = 0; n < 3; n++) gm[n] += cv_le[p][n]*f[p*ns + s];
	  for (int ia = 0; ia < 3; ia++)
	    for (int ib = 0; ib < 3; ib++) ds[ia][ib] = (gm[ia]*du[ib] + du[ia]*gm[ib] + rho*du[ia]*du[ib]);
	  for (int ip = 0; ip < nprop; ip++) {
	    int p = prop[side][ip];
	    double udotc = du[1]*cv_le[p][1];
	    double sdotq = 0.0;
	    for (int ia = 0; ia < 3; ia++)
	      for (int ib = 0; ib < 3; ib++) {
		double dab = cs2*(ia == ib);
		double q = (cv_le[p][ia]*cv_le[p][ib] - dab);
		sdotq += ds[ia][ib]*q;
	      }
	    double fv = f[p*ns + s];
	    fv += wv_le(p)*(rho*udotc*rcs2 + 0.5*sdotq*rcs2*rcs2);
	    sb[((long) (jc - 1)*nz + kc - 1)*5 + ip] = fv;
	  }
	}
      /* displacement (lb_data_displace_kernel), 1 <= jc <= ny + 1 */
      double dy = buffer_displacement(g, g->nh, ts);
      int dj = floor(fmod(dy*cx, (double) N));
      double *rb = rbuf + (long) (2*np + side)*(N + 1)*nrow;
      for (int jc = 1; jc <= N + 1; jc++) {
	int js = 1 + (jc + dj - 1 + 2*N) % N;
	memcpy(rb + (long) (jc - 1)*nrow, sb + (long) (js - 1)*nrow, sizeof(double)*nrow);
      }
      /* interpolation (lb_data_interpolate_kernel) */
      double d = fmod(dy*cx, (double) N);
      int jdy = floor(d);
      double fr = d - jdy;
      for (int jc = 1; jc <= N; jc++)
	for (int kc = 1; kc <= nz; kc++)
	  for (int ip = 0; ip < nprop; ip++) {
	    long s = ((long) (ic - 1)*N + jc - 1)*nz + kc - 1;
	    f[prop[side][ip]*ns + s] = (1.0 - fr)*rb[((long) (jc - 1)*nz + kc - 1)*5 + ip]
	      + fr*rb[((long) jc*nz + kc - 1)*5 + ip];
	  }
    }
  free(rbuf);
  free(sbuf);
}

/* ------------------------------------------------------------------------ */
/* Entry points                                                             */
/* ------------------------------------------------------------------------ */

int ora_le_force(int nx, int ny, int nz, const ora_fe_param_t * fe, const ora_le_param_t * lp,
		 double t, const double * phi, double * F) {
  le_t g;
  if (le_init(&g, nx, ny, nz, lp)) return -1;
  double * pe, * ge, * de;
  alloc_ext(&g, &pe, &ge, &de);
  memcpy(pe, phi, sizeof(double)*nx*ny*nz);
  gradient(&g, fe->stencil, t, pe, ge, de);
  force(&g, fe, pe, ge, de, F);
  free(pe); free(ge); free(de);
  return 0;
}

int ora_le_cahn_hilliard(int nx, int ny, int nz, const ora_fe_param_t * fe,
			 const ora_le_param_t * lp, double t, double * phi, const double * u,
			 double * csum, double phi_init_sum) {
  le_t g;
  if (fe->order > 4) return -1;          /* the buffer planes are nhalo = 2 deep */
  if (le_init(&g, nx, ny, nz, lp)) return -1;
  const long ns = (long) nx*ny*nz, ne = (long) (nx + g.nxb)*ny*nz;
  double * pe, * ge, * de, * ue = NULL;
  alloc_ext(&g, &pe, &ge, &de);
  memcpy(pe, phi, sizeof(double)*ns);
  gradient(&g, fe->stencil, t, pe, ge, de);
  if (u) {
    ue = (double *) calloc(3*ne, sizeof(double));
    for (int a = 0; a < 3; a++) memcpy(ue + a*ne, u + a*ns, sizeof(double)*ns);
  }
  cahn_hilliard(&g, fe, t, pe, de, ue, phi, csum, phi_init_sum);
  free(ue); free(pe); free(ge); free(de);
  return 0;
}

int ora_le_lb_bc(int nx, int ny, int nz, const ora_le_param_t * lp, double timestep, double * f) {
  le_t g;
  if (le_init(&g, nx, ny, nz, lp)) return -1;
  lb_bc(&g, timestep, f);
  return 0;
}

int ora_le_shear_profile(int nx, int ny, int nz, const ora_le_param_t * lp, double rho0,
			 double eta, double * f) {
  le_t g;
  if (le_init(&g, nx, ny, nz, lp)) return -1;
  const long ns = (long) nx*ny*nz;
  double u[3] = {0.0, 0.0, 0.0}, gradu[3][3] = {{0.0}};
  const double cs2 = (1.0/3.0), rcs2 = 1.0/cs2;
  gradu[0][1] = g.uy*g.nplane/(double) nx;          /* lees_edw_shear_rate */
  for (int ic = 1; ic <= nx; ic++) {
    /* lees_edw_steady_uy */
    double xglobal = 0 + (double) ic - 0.5;
    int nplane = (int) ((g.dx_min + xglobal)/g.dx_sep);
    u[1] = xglobal*gradu[0][1] - g.uy*nplane;
    for (int y = 0; y < ny; y++)
      for (int z = 0; z < nz; z++)
	for (int p = 0; p < NVEL; p++) {
	  double cdotu = 0.0, sdotq = 0.0;
	  for (int i = 0; i < 3; i++) {
	    cdotu += cv_le[p][i]*u[i];
	    for (int j = 0; j < 3; j++) {
	      double dij = (i == j);
	      double qij = cv_le[p][i]*cv_le[p][j] - cs2*dij;
	      sdotq += (rho0*u[i]*u[j] - eta*gradu[i][j])*qij;
	    }
	  }
	  f[p*ns + ((long) (ic - 1)*ny + y)*nz + z]
	    = wv_le(p)*(rho0 + rcs2*rho0*cdotu + 0.5*rcs2*rcs2*sdotq);
	}
  }
  return 0;
}

/* the ludwig.c:528-860 loop with planes: steps t_current + 1 .. +nsteps,
 * time t = t_start + t_current - 1 (physics.c:622-630) */
int ora_le_binary_steps(int nx, int ny, int nz, int nsteps, int t_start, int t_current,
			const ora_lb_param_t * lb, const ora_fe_param_t * fe,
			const ora_le_param_t * lp, double * f, double * phi, double * rho,
			double * u, double * csum, double phi_init_sum) {
  le_t g;
  if (fe->order > 4) return -1;          /* the buffer planes are nhalo = 2 deep */
  if (le_init(&g, nx, ny, nz, lp)) return -1;
  const long ns = (long) nx*ny*nz, ne = (long) (nx + g.nxb)*ny*nz;
  double * pe, * ge, * de;
  alloc_ext(&g, &pe, &ge, &de);
  double * ue = (double *) calloc(3*ne, sizeof(double));
  double * F = (double *) malloc(sizeof(double)*3*ns);
  double * ftmp = (double *) malloc(sizeof(double)*NVEL*ns);
  for (int step = 0; step < nsteps; step++) {
    t_current += 1;
    const double t = 1.0*(t_start + t_current - 1.0);
    memcpy(pe, phi, sizeof(double)*ns);
    gradient(&g, fe->stencil, t, pe, ge, de);
    force(&g, fe, pe, ge, de, F);
    for (int a = 0; a < 3; a++) memcpy(ue + a*ne, u + a*ns, sizeof(double)*ns);
    cahn_hilliard(&g, fe, t, pe, de, ue, phi, csum, phi_init_sum);
    ora_collide(nx, ny, nz, lb, f, F, rho, u);
    lb_bc(&g, 1.0*t_current, f);
    ora_propagate(nx, ny, nz, f, ftmp);
    memcpy(f, ftmp, sizeof(double)*NVEL*ns);
  }
  free(ftmp); free(F); free(ue); free(pe); free(ge); free(de);
  return 0;
}
