/*****************************************************************************
 *
 *  ldw_oracle.c -- CPU restatement of the D3Q19 binary-fluid step
 *
 *  TEST INFRASTRUCTURE ONLY (see ldw_oracle.h).  It is the checker for the
 *  GPU library and the timed CPU baseline; the product never calls it.
 *
 *  Each function restates the reference's per-site arithmetic in the
 *  reference's operation order, with periodic index wrap in place of halo
 *  copies.  Terms whose coefficient is exactly zero are skipped: adding
 *  x*0 = +-0 to a sum leaves its value unchanged, so the result is the
 *  same number the reference's unrolled transforms produce.
 *
 *  Pinned by: golden vectors written by the reference's own build
 *  (tests/golden/), test_noise.c known answers, the serial-spin-fd1.log
 *  [phi] line at t = 0, and closed-form checks in tests/test_oracle_*.py.
 *
 *****************************************************************************/

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ldw_oracle.h"

#define NVEL 19

/* D3Q19 velocity set and weights (lb_d3q19.h:24-36); the opposite of
 * p >= 1 is 19 - p. Weights are stored as numerators over 36. */

static const int cv_[NVEL][3] = {
  { 0,  0,  0},
  { 1,  1,  0}, { 1,  0,  1}, { 1,  0,  0},
  { 1,  0, -1}, { 1, -1,  0}, { 0,  1,  1},
  { 0,  1,  0}, { 0,  1, -1}, { 0,  0,  1},
  { 0,  0, -1}, { 0, -1,  1}, { 0, -1,  0},
  { 0, -1, -1}, {-1,  1,  0}, {-1,  0,  1},
  {-1,  0,  0}, {-1,  0, -1}, {-1, -1,  0}};

static const int w36_[NVEL] = {12, 1, 1, 2, 1, 1, 1, 2, 1, 2, 2, 1, 2, 1, 1,
			       1, 2, 1, 1};

/* Rational arithmetic for the mode matrices */

typedef struct { long long n, d; } rat_t;

static long long gcdll(long long a, long long b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) { long long t = a % b; a = b; b = t; }
  return a ? a : 1;
}

static rat_t rat(long long n, long long d) {
  long long g = gcdll(n, d);
  rat_t r = {n/g, d/g};
  if (r.d < 0) { r.n = -r.n; r.d = -r.d; }
  return r;
}

static rat_t rat_mul(rat_t a, rat_t b) { return rat(a.n*b.n, a.d*b.d); }
static rat_t rat_add(rat_t a, rat_t b) { return rat(a.n*b.d + b.n*a.d, a.d*b.d); }
static double rat_double(rat_t a) { return (double) a.n / (double) a.d; }

/* Mode matrix rows (lb_d3q19.c:107-153): 1, c, cc - delta/3, chi1, chi1 c,
 * chi2, chi2 c, chi3, with chi1 = (2c^2 - 3)(3cz^2 - c^2),
 * chi2 = (2c^2 - 3)(cy^2 - cx^2), chi3 = 3c^4 - 6c^2 + 1. */

static rat_t ma_rat(int m, int p) {
  int cx = cv_[p][0], cy = cv_[p][1], cz = cv_[p][2];
  int c2 = cx*cx + cy*cy + cz*cz;
  int chi1 = (2*c2 - 3)*(3*cz*cz - c2);
  int chi2 = (2*c2 - 3)*(cy*cy - cx*cx);
  int chi3 = 3*c2*c2 - 6*c2 + 1;
  switch (m) {
  case 0: return rat(1, 1);
  case 1: return rat(cx, 1);
  case 2: return rat(cy, 1);
  case 3: return rat(cz, 1);
  case 4: return rat(3*cx*cx - 1, 3);
  case 5: return rat(cx*cy, 1);
  case 6: return rat(cx*cz, 1);
  case 7: return rat(3*cy*cy - 1, 3);
  case 8: return rat(cy*cz, 1);
  case 9: return rat(3*cz*cz - 1, 3);
  case 10: return rat(chi1, 1);
  case 11: return rat(chi1*cx, 1);
  case 12: return rat(chi1*cy, 1);
  case 13: return rat(chi1*cz, 1);
  case 14: return rat(chi2, 1);
  case 15: return rat(chi2*cx, 1);
  case 16: return rat(chi2*cy, 1);
  case 17: return rat(chi2*cz, 1);
  default: return rat(chi3, 1);
  }
}

void ora_d3q19_tables(int cv[19][3], double wv[19], double ma[19][19],
		      double mi[19][19]) {
  rat_t na[NVEL];
  for (int m = 0; m < NVEL; m++) {
    /* na = 1 / sum_p w_p ma_mp^2 (lb_d3q19.c:72-80) */
    rat_t s = rat(0, 1);
    for (int p = 0; p < NVEL; p++) {
      rat_t e = ma_rat(m, p);
      s = rat_add(s, rat_mul(rat(w36_[p], 36), rat_mul(e, e)));
    }
    na[m] = rat(s.d, s.n);
  }
  for (int p = 0; p < NVEL; p++) {
    if (cv) for (int a = 0; a < 3; a++) cv[p][a] = cv_[p][a];
    if (wv) wv[p] = rat_double(rat(w36_[p], 36));
    for (int m = 0; m < NVEL; m++) {
      if (ma) ma[m][p] = rat_double(ma_rat(m, p));
      /* mi = w na ma^T (lb_data.c:644) */
      if (mi) mi[p][m] = rat_double(rat_mul(rat(w36_[p], 36),
					    rat_mul(na[m], ma_rat(m, p))));
    }
  }
}

/* Tables initialised once */

static int tables_ready_ = 0;
static double ma_[NVEL][NVEL];
static double mi_[NVEL][NVEL];
static double wv_[NVEL];

static void tables_init(void) {
  if (tables_ready_) return;
  #pragma omp critical (ora_tables)
  {
    if (!tables_ready_) {
      ora_d3q19_tables(NULL, wv_, ma_, mi_);
      /* The reference's unrolled f -> mode transform omits the f[4] term
       * of ghost mode 13 (collision.c:2278 multiplies by c0 where the mode
       * matrix entry chi1*cz is -1).  It is reproduced as executed.  Under
       * M10 the ghost modes relax to zero, so it only shows with BGK/TRT. */
      ma_[13][4] = 0.0;
      tables_ready_ = 1;
    }
  }
}

/* Relaxation rates (collision.c:1163-1246, 1260-1420) */

void ora_relaxation_rates(const ora_lb_param_t * p, double rtau[19]) {
  const double cs2 = (1.0/3.0);
  double rtau_shear = 1.0/(0.5 + p->eta_shear/(p->rho0*cs2));
  double rtau_bulk = 1.0/(0.5 + p->eta_bulk/(p->rho0*cs2));
  for (int m = 0; m < NVEL; m++) rtau[m] = 0.0;
  rtau[4] = rtau_shear;            /* slot LB_TAU_SHEAR */
  rtau[5] = rtau_bulk;             /* slot LB_TAU_BULK */
  if (p->relax == ORA_M10) {
    for (int m = 10; m < NVEL; m++) rtau[m] = 1.0;
  }
  else if (p->relax == ORA_BGK) {
    rtau[5] = rtau_shear;
    for (int m = 10; m < NVEL; m++) rtau[m] = rtau_shear;
  }
  else {
    double tau = p->eta_shear/(p->rho0*cs2);
    double rg = 0.5 + 2.0*tau/(tau + 3.0/8.0);
    if (rg > 2.0) rg = 2.0;
    rtau[10] = rtau_shear; rtau[14] = rtau_shear; rtau[18] = rtau_shear;
    rtau[11] = rg; rtau[12] = rg; rtau[13] = rg;
    rtau[15] = rg; rtau[16] = rg; rtau[17] = rg;
  }
}

/* One site of lb_collision_mrt1_site (collision.c:253-593), no noise,
 * fixed viscosity, no stress relaxation.  rtau[4], rtau[5] hold the shear
 * and bulk rates; rtau[10..18] the ghost rates. */

static void collide_site(const double rtau[19], const double force[3],
			 double fc[19], double * prho, double uout[3]) {

  double mode[NVEL];
  double u[3], s[3][3], seq[3][3];
  const double rdim = (1.0/3.0);

  for (int m = 0; m < NVEL; m++) {
    double acc = 0.0;
    for (int p = 0; p < NVEL; p++) {
      if (ma_[m][p] != 0.0) acc += fc[p]*ma_[m][p];
    }
    mode[m] = acc;
  }

  double rho = mode[0];
  for (int a = 0; a < 3; a++) u[a] = mode[1 + a];
  {
    int m = 0;
    for (int a = 0; a < 3; a++) {
      for (int b = a; b < 3; b++) s[a][b] = mode[4 + m++];
    }
    for (int a = 1; a < 3; a++) {
      for (int b = 0; b < a; b++) s[a][b] = s[b][a];
    }
  }

  double rrho = 1.0/rho;
  for (int a = 0; a < 3; a++) u[a] = rrho*(u[a] + 0.5*force[a]);

  double rtau_s = rtau[4];
  double rtau_b = rtau[5];
  double tr_s = 0.0, tr_seq = 0.0;

  for (int a = 0; a < 3; a++) {
    for (int b = 0; b < 3; b++) seq[a][b] = rho*u[a]*u[b];
    tr_s += s[a][a];
    tr_seq += seq[a][a];
  }
  for (int a = 0; a < 3; a++) {
    s[a][a] -= rdim*tr_s;
    seq[a][a] -= rdim*tr_seq;
  }
  tr_s = tr_s - rtau_b*(tr_s - tr_seq);

  for (int a = 0; a < 3; a++) {
    for (int b = 0; b < 3; b++) {
      s[a][b] -= rtau_s*(s[a][b] - seq[a][b]);
      if (a == b) s[a][b] += rdim*tr_s;
      s[a][b] += (2.0 - rtau_s)*(u[a]*force[b] + force[a]*u[b]);
    }
  }

  for (int a = 0; a < 3; a++) mode[1 + a] += force[a];
  {
    int m = 0;
    for (int a = 0; a < 3; a++) {
      for (int b = a; b < 3; b++) mode[4 + m++] = s[a][b];
    }
  }
  for (int m = 10; m < NVEL; m++) {
    mode[m] = mode[m] - rtau[m]*(mode[m] - 0.0);
  }

  for (int p = 0; p < NVEL; p++) {
    double acc = 0.0;
    for (int m = 0; m < NVEL; m++) {
      if (mi_[p][m] != 0.0) acc += mi_[p][m]*mode[m];
    }
    fc[p] = acc;
  }

  *prho = rho;
  for (int a = 0; a < 3; a++) uout[a] = u[a];
}

void ora_collide(int nx, int ny, int nz, const ora_lb_param_t * p,
		 double * f, const double * F, double * rho, double * u) {

  const long nsite = (long) nx*ny*nz;
  double rtau[19];

  tables_init();
  ora_relaxation_rates(p, rtau);

  #pragma omp parallel for schedule(static)
  for (long is = 0; is < nsite; is++) {
    double fc[NVEL], force[3], uu[3], r;
    for (int q = 0; q < NVEL; q++) fc[q] = f[q*nsite + is];
    for (int a = 0; a < 3; a++) {
      force[a] = p->fglobal[a] + (F ? F[a*nsite + is] : 0.0);
    }
    collide_site(rtau, force, fc, &r, uu);
    for (int q = 0; q < NVEL; q++) f[q*nsite + is] = fc[q];
    if (rho) rho[is] = r;
    if (u) for (int a = 0; a < 3; a++) u[a*nsite + is] = uu[a];
  }
}

static inline int wrap(int i, int n) {
  return (i < 0) ? i + n : ((i >= n) ? i - n : i);
}

static inline long site3(int x, int y, int z, int ny, int nz) {
  return ((long) x*ny + y)*nz + z;
}

/* propagation.c:155-199 */

void ora_propagate(int nx, int ny, int nz, const double * fin, double * fout) {
  const long nsite = (long) nx*ny*nz;
  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	long is = site3(x, y, z, ny, nz);
	for (int p = 0; p < NVEL; p++) {
	  long js = site3(wrap(x - cv_[p][0], nx), wrap(y - cv_[p][1], ny),
			  wrap(z - cv_[p][2], nz), ny, nz);
	  fout[p*nsite + is] = fin[p*nsite + js];
	}
      }
    }
  }
}

/* lb_1st_moment_equilib_set (lb_data.c:809-840) */

void ora_equilibrium(int nx, int ny, int nz, const double * rho,
		     const double * u, double rho_uniform, double * f) {
  const long nsite = (long) nx*ny*nz;
  tables_init();
  #pragma omp parallel for schedule(static)
  for (long is = 0; is < nsite; is++) {
    double uu[3] = {0.0, 0.0, 0.0};
    double r = rho ? rho[is] : rho_uniform;
    if (u) for (int a = 0; a < 3; a++) uu[a] = u[a*nsite + is];
    for (int p = 0; p < NVEL; p++) {
      double cs2 = (1.0/3.0);
      double rcs2 = 1.0/cs2;
      double udotc = 0.0;
      double sdotq = 0.0;
      for (int a = 0; a < 3; a++) {
	udotc += uu[a]*cv_[p][a];
	for (int b = 0; b < 3; b++) {
	  double dab = (a == b);
	  sdotq += (cv_[p][a]*cv_[p][b] - cs2*dab)*uu[a]*uu[b];
	}
      }
      f[p*nsite + is] = r*wv_[p]*(1.0 + rcs2*udotc + 0.5*rcs2*rcs2*sdotq);
    }
  }
}

/* grad_3d_7pt_fluid_kernel_v (gradient_3d_7pt_fluid.c:231-310) */

void ora_grad7(int nx, int ny, int nz, const double * phi, double * grad,
	       double * delsq) {
  const long nsite = (long) nx*ny*nz;
  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	long is = site3(x, y, z, ny, nz);
	double xp = phi[site3(wrap(x + 1, nx), y, z, ny, nz)];
	double xm = phi[site3(wrap(x - 1, nx), y, z, ny, nz)];
	double yp = phi[site3(x, wrap(y + 1, ny), z, ny, nz)];
	double ym = phi[site3(x, wrap(y - 1, ny), z, ny, nz)];
	double zp = phi[site3(x, y, wrap(z + 1, nz), ny, nz)];
	double zm = phi[site3(x, y, wrap(z - 1, nz), ny, nz)];
	grad[0*nsite + is] = 0.5*(xp - xm);
	grad[1*nsite + is] = 0.5*(yp - ym);
	grad[2*nsite + is] = 0.5*(zp - zm);
	delsq[is] = xp + xm + yp + ym + zp + zm - 6.0*phi[is];
      }
    }
  }
}

/* grad_3d_27pt_kernel (gradient_3d_27pt_fluid.c:219-368): each
 * component sums the 9 differences across the cube in the kernel's order
 * (y slowest, then z, for d_x; x slowest, then z, for d_y; x, then y,
 * for d_z), times 0.5/9; the Laplacian sums the 26 neighbours in x, y, z
 * order minus 26 phi, times 1/9. */

void ora_grad27(int nx, int ny, int nz, const double * phi, double * grad,
		double * delsq) {
  const long nsite = (long) nx*ny*nz;
  const double r9 = (1.0/9.0);
  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	double p[3][3][3];
	for (int i = 0; i < 3; i++)
	  for (int j = 0; j < 3; j++)
	    for (int k = 0; k < 3; k++)
	      p[i][j][k] = phi[site3(wrap(x + i - 1, nx), wrap(y + j - 1, ny), wrap(z + k - 1, nz), ny, nz)];
	long is = site3(x, y, z, ny, nz);
	double sx = 0.0, sy = 0.0, sz = 0.0;
	int first = 1;
	for (int j = 0; j < 3; j++)
	  for (int k = 0; k < 3; k++) {
	    if (first) sx = + p[2][j][k] - p[0][j][k];
	    else sx = sx + p[2][j][k] - p[0][j][k];
	    first = 0;
	  }
	first = 1;
	for (int i = 0; i < 3; i++)
	  for (int k = 0; k < 3; k++) {
	    if (first) sy = + p[i][2][k] - p[i][0][k];
	    else sy = sy + p[i][2][k] - p[i][0][k];
	    first = 0;
	  }
	first = 1;
	for (int i = 0; i < 3; i++)
	  for (int j = 0; j < 3; j++) {
	    if (first) sz = + p[i][j][2] - p[i][j][0];
	    else sz = sz + p[i][j][2] - p[i][j][0];
	    first = 0;
	  }
	double sd = 0.0;
	first = 1;
	for (int i = 0; i < 3; i++)
	  for (int j = 0; j < 3; j++)
	    for (int k = 0; k < 3; k++) {
	      if (i == 1 && j == 1 && k == 1) continue;
	      if (first) sd = + p[i][j][k];
	      else sd = sd + p[i][j][k];
	      first = 0;
	    }
	sd = sd - 26.0*p[1][1][1];
	grad[0*nsite + is] = 0.5*r9*sx;
	grad[1*nsite + is] = 0.5*r9*sy;
	grad[2*nsite + is] = 0.5*r9*sz;
	delsq[is] = r9*sd;
      }
    }
  }
}

/* fe_symm_mu (symmetric.c:308-318) */

static inline double mu_site(const ora_fe_param_t * fe, double phi,
			     double delsq) {
  return fe->a*phi + fe->b*phi*phi*phi - fe->kappa*delsq;
}

void ora_mu(int nsite, const ora_fe_param_t * fe, const double * phi,
	    const double * delsq, double * mu) {
  #pragma omp parallel for schedule(static)
  for (long is = 0; is < nsite; is++) mu[is] = mu_site(fe, phi[is], delsq[is]);
}

/* fe_symm_str_v (symmetric.c:372-418) */

static inline void stress_site(const ora_fe_param_t * fe, double phi,
			       const double g[3], double delsq,
			       double s[3][3]) {
  const double a = fe->a, b = fe->b, kappa = fe->kappa;
  double p0 = 0.5*a*phi*phi + 0.75*b*phi*phi*phi*phi - kappa*phi*delsq
    - 0.5*kappa*(g[0]*g[0] + g[1]*g[1] + g[2]*g[2]);
  for (int ia = 0; ia < 3; ia++) {
    for (int ib = 0; ib < 3; ib++) {
      double d = (ia == ib);
      s[ia][ib] = p0*d + kappa*g[ia]*g[ib];
    }
  }
}

void ora_stress(int nsite, const ora_fe_param_t * fe, const double * phi,
		const double * grad, const double * delsq, double * pth) {
  #pragma omp parallel for schedule(static)
  for (long is = 0; is < nsite; is++) {
    double g[3] = {grad[is], grad[nsite + is], grad[2*(long) nsite + is]};
    double s[3][3];
    stress_site(fe, phi[is], g, delsq[is], s);
    for (int ia = 0; ia < 3; ia++)
      for (int ib = 0; ib < 3; ib++) pth[(3*ia + ib)*(long) nsite + is] = s[ia][ib];
  }
}

void ora_force(int nx, int ny, int nz, const ora_fe_param_t * fe,
	       const double * phi, const double * grad, const double * delsq,
	       double * F) {

  const long nsite = (long) nx*ny*nz;

  if (fe->force_method == 0) {
    /* pth_force_fluid_kernel_v (phi_force_colloid.c:315-466) */
    double * pth = (double *) malloc(sizeof(double)*9*nsite);
    ora_stress(nsite, fe, phi, grad, delsq, pth);
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; x++) {
      for (int y = 0; y < ny; y++) {
	for (int z = 0; z < nz; z++) {
	  long is = site3(x, y, z, ny, nz);
	  long nb[6] = {
	    site3(wrap(x + 1, nx), y, z, ny, nz), site3(wrap(x - 1, nx), y, z, ny, nz),
	    site3(x, wrap(y + 1, ny), z, ny, nz), site3(x, wrap(y - 1, ny), z, ny, nz),
	    site3(x, y, wrap(z + 1, nz), ny, nz), site3(x, y, wrap(z - 1, nz), ny, nz)};
	  double force[3];
	  for (int ia = 0; ia < 3; ia++) {
#define P(site, a, b) pth[(3*(a) + (b))*nsite + (site)]
	    force[ia] = -0.5*(P(nb[0], ia, 0) + P(is, ia, 0));
	    force[ia] += 0.5*(P(nb[1], ia, 0) + P(is, ia, 0));
	    force[ia] -= 0.5*(P(nb[2], ia, 1) + P(is, ia, 1));
	    force[ia] += 0.5*(P(nb[3], ia, 1) + P(is, ia, 1));
	    force[ia] -= 0.5*(P(nb[4], ia, 2) + P(is, ia, 2));
	    force[ia] += 0.5*(P(nb[5], ia, 2) + P(is, ia, 2));
#undef P
	  }
	  for (int ia = 0; ia < 3; ia++) F[ia*nsite + is] = 0.0 + force[ia];
	}
      }
    }
    free(pth);
  }
  else {
    /* phi_grad_mu_fluid_kernel (phi_grad_mu.c:284-350) */
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; x++) {
      for (int y = 0; y < ny; y++) {
	for (int z = 0; z < nz; z++) {
	  long is = site3(x, y, z, ny, nz);
	  long nb[6] = {
	    site3(wrap(x - 1, nx), y, z, ny, nz), site3(wrap(x + 1, nx), y, z, ny, nz),
	    site3(x, wrap(y - 1, ny), z, ny, nz), site3(x, wrap(y + 1, ny), z, ny, nz),
	    site3(x, y, wrap(z - 1, nz), ny, nz), site3(x, y, wrap(z + 1, nz), ny, nz)};
	  double phi0 = phi[is];
	  for (int ia = 0; ia < 3; ia++) {
	    double mum1 = mu_site(fe, phi[nb[2*ia]], delsq[nb[2*ia]]);
	    double mup1 = mu_site(fe, phi[nb[2*ia + 1]], delsq[nb[2*ia + 1]]);
	    double force = 0.0;
	    force += -phi0*0.5*(mup1 - mum1);
	    F[ia*nsite + is] = 0.0 + force;
	  }
	  /* phi_grad_mu_external_kernel: a second hydro_f_local_add */
	  if (fe->grad_mu[0] != 0.0 || fe->grad_mu[1] != 0.0 || fe->grad_mu[2] != 0.0) {
	    for (int ia = 0; ia < 3; ia++) F[ia*nsite + is] += -phi0*fe->grad_mu[ia];
	  }
	}
      }
    }
  }
}

/* Cahn-Hilliard: advective fluxes (advection.c:538-632, 770-900,
 * 946-1145), diffusive fluxes (phi_cahn_hilliard.c:350-413), no-flux
 * masks (identity for an all-fluid lattice, advection_bcs.c:80-134) and
 * the forward-step divergence (phi_cahn_hilliard.c:1018-1052). */

/* kahan_add_double (util_sum.c:30-40); built without value-changing
 * optimisations, so the four operations keep their order and rounding */
static void kahan_add(double * sum, double * cs, double val) {
  double y = val + *cs;
  double t = *sum + y;
  *cs = y - (t - *sum);
  *sum = t;
}

void ora_cahn_hilliard(int nx, int ny, int nz, const ora_fe_param_t * fe,
		       double * phi, const double * delsq, const double * u) {
  ora_fe_param_t plain = *fe;
  plain.conserve = 0;
  ora_cahn_hilliard_ext(nx, ny, nz, &plain, phi, delsq, u, NULL, 0.0);
}

void ora_cahn_hilliard_ext(int nx, int ny, int nz, const ora_fe_param_t * fe,
			   double * phi, const double * delsq, const double * u,
			   double * csum, double phi_init_sum) {

  const long nsite = (long) nx*ny*nz;
  const double a1 = -0.213933;
  const double a2 =  0.927865;
  const double a3 =  0.286067;
  const double mob = fe->mobility;

  double * fw = (double *) malloc(sizeof(double)*4*nsite);
  double * fe_ = fw + nsite;
  double * fy = fw + 2*nsite;
  double * fz = fw + 3*nsite;

  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	long i0 = site3(x, y, z, ny, nz);
	long xm1 = site3(wrap(x - 1, nx), y, z, ny, nz);
	long xm2 = site3(wrap(x - 2, nx), y, z, ny, nz);
	long xp1 = site3(wrap(x + 1, nx), y, z, ny, nz);
	long xp2 = site3(wrap(x + 2, nx), y, z, ny, nz);
	long ym1 = site3(x, wrap(y - 1, ny), z, ny, nz);
	long yp1 = site3(x, wrap(y + 1, ny), z, ny, nz);
	long yp2 = site3(x, wrap(y + 2, ny), z, ny, nz);
	long zm1 = site3(x, y, wrap(z - 1, nz), ny, nz);
	long zp1 = site3(x, y, wrap(z + 1, nz), ny, nz);
	long zp2 = site3(x, y, wrap(z + 2, nz), ny, nz);
	const double * ux = u;
	const double * uy = u + nsite;
	const double * uz = u + 2*nsite;
	double w, e, fyv, fzv, uf;

	if (fe->order == 1) {
	  long idx;
	  uf = 0.5*(ux[i0] + ux[xm1]);
	  idx = (uf > 0.0) ? xm1 : i0;
	  w = uf*phi[idx];
	  uf = 0.5*(ux[i0] + ux[xp1]);
	  idx = (uf < 0.0) ? xp1 : i0;
	  e = uf*phi[idx];
	  uf = 0.5*(uy[i0] + uy[yp1]);
	  idx = (uf < 0.0) ? yp1 : i0;
	  fyv = uf*phi[idx];
	  uf = 0.5*(uz[i0] + uz[zp1]);
	  idx = (uf < 0.0) ? zp1 : i0;
	  fzv = uf*phi[idx];
	}
	else if (fe->order == 2) {
	  w = 0.5*(ux[i0] + ux[xm1])*0.5*(phi[xm1] + phi[i0]);
	  e = 0.5*(ux[i0] + ux[xp1])*0.5*(phi[i0] + phi[xp1]);
	  fyv = 0.5*(uy[i0] + uy[yp1])*0.5*(phi[i0] + phi[yp1]);
	  fzv = 0.5*(uz[i0] + uz[zp1])*0.5*(phi[i0] + phi[zp1]);
	}
	else if (fe->order == 4) {
	  /* advection_le_4th (advection.c:1153-1262) */
	  const double b1 = 1.0/16.0, b2 = 9.0/16.0;
	  uf = 0.5*(ux[i0] + ux[xm1]);
	  w = uf*(- b1*phi[xm2] + b2*phi[xm1] + b2*phi[i0] - b1*phi[xp1]);
	  uf = 0.5*(ux[i0] + ux[xp1]);
	  e = uf*(- b1*phi[xm1] + b2*phi[i0] + b2*phi[xp1] - b1*phi[xp2]);
	  uf = 0.5*(uy[i0] + uy[yp1]);
	  fyv = uf*(- b1*phi[ym1] + b2*phi[i0] + b2*phi[yp1] - b1*phi[yp2]);
	  uf = 0.5*(uz[i0] + uz[zp1]);
	  fzv = uf*(- b1*phi[zm1] + b2*phi[i0] + b2*phi[zp1] - b1*phi[zp2]);
	}
	else if (fe->order == 5) {
	  /* advection_le_5th (advection.c:1264-1430): five points, upwind
	   * biased; the points of each face in order from the upwind side */
	  const double c[5] = {0.055453, -0.305147, 0.916054, 0.361520, -0.027880};
	  long xm3 = site3(wrap(x - 3, nx), y, z, ny, nz);
	  long xp3 = site3(wrap(x + 3, nx), y, z, ny, nz);
	  long ym2 = site3(x, wrap(y - 2, ny), z, ny, nz);
	  long yp3 = site3(x, wrap(y + 3, ny), z, ny, nz);
	  long zm2 = site3(x, y, wrap(z - 2, nz), ny, nz);
	  long zp3 = site3(x, y, wrap(z + 3, nz), ny, nz);
	  long s5[5];
	  uf = 0.5*(ux[i0] + ux[xm1]);
	  if (uf > 0.0) { long t[5] = {xm3, xm2, xm1, i0, xp1}; memcpy(s5, t, sizeof(t)); }
	  else          { long t[5] = {xp2, xp1, i0, xm1, xm2}; memcpy(s5, t, sizeof(t)); }
	  w = uf*(c[0]*phi[s5[0]] + c[1]*phi[s5[1]] + c[2]*phi[s5[2]] + c[3]*phi[s5[3]] + c[4]*phi[s5[4]]);
	  uf = 0.5*(ux[i0] + ux[xp1]);
	  if (uf < 0.0) { long t[5] = {xp3, xp2, xp1, i0, xm1}; memcpy(s5, t, sizeof(t)); }
	  else          { long t[5] = {xm2, xm1, i0, xp1, xp2}; memcpy(s5, t, sizeof(t)); }
	  e = uf*(c[0]*phi[s5[0]] + c[1]*phi[s5[1]] + c[2]*phi[s5[2]] + c[3]*phi[s5[3]] + c[4]*phi[s5[4]]);
	  uf = 0.5*(uy[i0] + uy[yp1]);
	  if (uf < 0.0) { long t[5] = {yp3, yp2, yp1, i0, ym1}; memcpy(s5, t, sizeof(t)); }
	  else          { long t[5] = {ym2, ym1, i0, yp1, yp2}; memcpy(s5, t, sizeof(t)); }
	  fyv = uf*(c[0]*phi[s5[0]] + c[1]*phi[s5[1]] + c[2]*phi[s5[2]] + c[3]*phi[s5[3]] + c[4]*phi[s5[4]]);
	  uf = 0.5*(uz[i0] + uz[zp1]);
	  if (uf < 0.0) { long t[5] = {zp3, zp2, zp1, i0, zm1}; memcpy(s5, t, sizeof(t)); }
	  else          { long t[5] = {zm2, zm1, i0, zp1, zp2}; memcpy(s5, t, sizeof(t)); }
	  fzv = uf*(c[0]*phi[s5[0]] + c[1]*phi[s5[1]] + c[2]*phi[s5[2]] + c[3]*phi[s5[3]] + c[4]*phi[s5[4]]);
	}
	else {
	  double d1, d2, d3;
	  uf = 0.5*(ux[i0] + ux[xm1]);
	  if (uf > 0.0) { d1 = phi[xm2]; d2 = phi[xm1]; d3 = phi[i0]; }
	  else          { d1 = phi[xp1]; d2 = phi[i0];  d3 = phi[xm1]; }
	  w = uf*(a1*d1 + a2*d2 + a3*d3);
	  uf = 0.5*(ux[i0] + ux[xp1]);
	  if (uf < 0.0) { d1 = phi[xp2]; d2 = phi[xp1]; d3 = phi[i0]; }
	  else          { d1 = phi[xm1]; d2 = phi[i0];  d3 = phi[xp1]; }
	  e = uf*(a1*d1 + a2*d2 + a3*d3);
	  uf = 0.5*(uy[i0] + uy[yp1]);
	  if (uf < 0.0) { d1 = phi[yp2]; d2 = phi[yp1]; d3 = phi[i0]; }
	  else          { d1 = phi[ym1]; d2 = phi[i0];  d3 = phi[yp1]; }
	  fyv = uf*(a1*d1 + a2*d2 + a3*d3);
	  uf = 0.5*(uz[i0] + uz[zp1]);
	  if (uf < 0.0) { d1 = phi[zp2]; d2 = phi[zp1]; d3 = phi[i0]; }
	  else          { d1 = phi[zm1]; d2 = phi[i0];  d3 = phi[zp1]; }
	  fzv = uf*(a1*d1 + a2*d2 + a3*d3);
	}

	{
	  double mu0 = mu_site(fe, phi[i0], delsq[i0]);
	  double mu1;
	  mu1 = mu_site(fe, phi[xm1], delsq[xm1]);
	  w -= mob*(mu0 - mu1);
	  mu1 = mu_site(fe, phi[xp1], delsq[xp1]);
	  e -= mob*(mu1 - mu0);
	  mu1 = mu_site(fe, phi[yp1], delsq[yp1]);
	  fyv -= mob*(mu1 - mu0);
	  mu1 = mu_site(fe, phi[zp1], delsq[zp1]);
	  fzv -= mob*(mu1 - mu0);
	}
	/* phi_ch_flux_mu_ext_kernel: every flux, unconditionally */
	w -= mob*fe->grad_mu[0];
	e -= mob*fe->grad_mu[0];
	fyv -= mob*fe->grad_mu[1];
	fzv -= mob*fe->grad_mu[2];
	fw[i0] = w; fe_[i0] = e; fy[i0] = fyv; fz[i0] = fzv;
      }
    }
  }

  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	long i0 = site3(x, y, z, ny, nz);
	long ym1 = site3(x, wrap(y - 1, ny), z, ny, nz);
	long zm1 = site3(x, y, wrap(z - 1, nz), ny, nz);
	const double wz = 1.0;
	if (fe->conserve == 1) {
	  /* phi_ch_csum_kernel (phi_cahn_hilliard.c:1181-1218) */
	  double sum = phi[i0], cs = csum[i0];
	  kahan_add(&sum, &cs, -fe_[i0]);
	  kahan_add(&sum, &cs,  fw[i0]);
	  kahan_add(&sum, &cs, -fy[i0]);
	  kahan_add(&sum, &cs,  fy[ym1]);
	  kahan_add(&sum, &cs, -wz*fz[i0]);
	  kahan_add(&sum, &cs,  wz*fz[zm1]);
	  csum[i0] = cs;
	  phi[i0] = sum;
	}
	else {
	  phi[i0] -= (+ fe_[i0] - fw[i0] + fy[i0] - fy[ym1]
		      + wz*fz[i0] - wz*fz[zm1]);
	}
      }
    }
  }

  if (fe->conserve == 2) {
    /* phi_ch_subtract_kernel1/2: the sum in site order from 0.0, then the
     * same correction at every (fluid) site */
    double s = 0.0;
    for (long i = 0; i < nsite; i++) s += phi[i];
    for (long i = 0; i < nsite; i++) phi[i] -= (s - phi_init_sum)/(int) nsite;
  }

  free(fw);
}

double ora_phi_init_sum(int64_t n, const double * phi) {
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0, s3 = 0.0, c3 = 0.0;
  for (int64_t i = 0; i < n; i++) kahan_add(&s1, &c1, phi[i]);
  kahan_add(&s2, &c2, s1);       /* kahan_add(&sum, phib[0]) */
  kahan_add(&s2, &c2, c1);
  kahan_add(&s3, &c3, s2);       /* kahan_add(&stats->sum1, sum) */
  kahan_add(&s3, &c3, c2);
  return s3 + c3;                /* kahan_sum() */
}

void ora_binary_steps(int nx, int ny, int nz, int nsteps,
		      const ora_lb_param_t * lb, const ora_fe_param_t * fe,
		      double * f, double * phi, double * rho, double * u) {
  ora_fe_param_t plain = *fe;
  plain.conserve = 0;
  ora_binary_steps_ext(nx, ny, nz, nsteps, lb, &plain, f, phi, rho, u, NULL, 0.0);
}

void ora_binary_steps_ext(int nx, int ny, int nz, int nsteps,
			  const ora_lb_param_t * lb, const ora_fe_param_t * fe,
			  double * f, double * phi, double * rho, double * u,
			  double * csum, double phi_init_sum) {

  const long nsite = (long) nx*ny*nz;
  double * grad = (double *) malloc(sizeof(double)*3*nsite);
  double * delsq = (double *) malloc(sizeof(double)*nsite);
  double * F = (double *) malloc(sizeof(double)*3*nsite);
  double * ftmp = (double *) malloc(sizeof(double)*NVEL*nsite);

  for (int step = 0; step < nsteps; step++) {
    if (fe->stencil == 27) ora_grad27(nx, ny, nz, phi, grad, delsq);
    else ora_grad7(nx, ny, nz, phi, grad, delsq);
    ora_force(nx, ny, nz, fe, phi, grad, delsq, F);
    ora_cahn_hilliard_ext(nx, ny, nz, fe, phi, delsq, u, csum, phi_init_sum);
    ora_collide(nx, ny, nz, lb, f, F, rho, u);
    ora_propagate(nx, ny, nz, f, ftmp);
    memcpy(f, ftmp, sizeof(double)*NVEL*nsite);
  }

  free(ftmp); free(F); free(delsq); free(grad);
}

void ora_single_steps(int nx, int ny, int nz, int nsteps,
		      const ora_lb_param_t * lb, double * f, double * rho,
		      double * u) {
  const long nsite = (long) nx*ny*nz;
  double * ftmp = (double *) malloc(sizeof(double)*NVEL*nsite);
  for (int step = 0; step < nsteps; step++) {
    ora_collide(nx, ny, nz, lb, f, NULL, rho, u);
    ora_propagate(nx, ny, nz, f, ftmp);
    memcpy(f, ftmp, sizeof(double)*NVEL*nsite);
  }
  free(ftmp);
}

/* ns_uniform (noise.h:96-109): L'Ecuyer-Simard combination as used by the
 * reference */

uint32_t ora_ns_uniform(uint32_t s[4]) {
  uint32_t b;
  s[0] = 69069u*s[0] + 1234567u;
  b = s[1] ^ (s[1] << 17);
  b ^= (b >> 13);
  s[1] = b ^ (b << 5);
  s[2] = 36969u*(s[2] & 0xffffu) + (s[2] >> 16);
  s[3] = 18000u*(s[3] & 0xffffu) + (s[3] >> 16);
  b = (s[2] << 16) + s[3];
  return (s[1] + (s[0] ^ b));
}

/* noise_initialise_state (noise.c:278-341) for one site; the generator
 * is warmed through four draws which become the site state. */

void ora_noise_state(uint32_t seed, int ig, int jg, int kg, uint32_t s[4]) {
  uint32_t local[4] = {seed + (uint32_t) ig, 12953u + (uint32_t) jg,
		       712357u + (uint32_t) kg, 22383979u};
  s[0] = ora_ns_uniform(local);
  s[1] = ora_ns_uniform(local);
  s[2] = ora_ns_uniform(local);
  s[3] = ora_ns_uniform(local);
}

static uint32_t seed_of(int seed) {
  /* noise_create_seed (noise.c:87-94): seeds <= 1 fall back to 13 */
  return (seed > 1) ? (uint32_t) seed : 13u;
}

void ora_spinodal(int nx, int ny, int nz, int seed, double phi0, double amp,
		  double * phi) {
  uint32_t sd = seed_of(seed);
  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	uint32_t s[4];
	ora_noise_state(sd, x + 1, y + 1, z + 1, s);
	double ran = (1.0/4294967295u)*ora_ns_uniform(s);
	phi[site3(x, y, z, ny, nz)] = phi0 + amp*(ran - 0.5);
      }
    }
  }
}

void ora_random_velocity(int nx, int ny, int nz, int seed, double amp,
			 double * u) {
  const long nsite = (long) nx*ny*nz;
  uint32_t sd = seed_of(seed);
  #pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; x++) {
    for (int y = 0; y < ny; y++) {
      for (int z = 0; z < nz; z++) {
	uint32_t s[4];
	long is = site3(x, y, z, ny, nz);
	ora_noise_state(sd, x + 1, y + 1, z + 1, s);
	for (int a = 0; a < 3; a++) {
	  double r = (1.0/4294967295u)*ora_ns_uniform(s);
	  u[a*nsite + is] = amp*(2.0*r - 1.0);
	}
      }
    }
  }
  for (int a = 0; a < 3; a++) {
    double mean = 0.0;
    for (long is = 0; is < nsite; is++) mean += u[a*nsite + is];
    mean = mean/nsite;
    for (long is = 0; is < nsite; is++) u[a*nsite + is] -= mean;
  }
}

/* util_sum.c:30-110 */

double ora_kahan_sum(int64_t n, const double * a) {
  volatile double sum = 0.0, cs = 0.0;
  for (int64_t i = 0; i < n; i++) {
    volatile double y = a[i] + cs;
    volatile double t = sum + y;
    cs = y - (t - sum);
    sum = t;
  }
  return sum + cs;
}

double ora_klein_sum(int64_t n, const double * a) {
  volatile double sum = 0.0, cs = 0.0, ccs = 0.0;
  for (int64_t i = 0; i < n; i++) {
    volatile double c, cc;
    volatile double t = sum + a[i];
    if (fabs(sum) >= fabs(a[i])) c = (sum - t) + a[i];
    else c = (a[i] - t) + sum;
    sum = t;
    t = cs + c;
    if (fabs(cs) >= fabs(c)) cc = (cs - t) + c;
    else cc = (c - t) + cs;
    cs = t;
    ccs = ccs + cc;
  }
  return sum + cs + ccs;
}
