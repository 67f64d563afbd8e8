/*****************************************************************************
 *
 *  ldw_oracle.h -- CPU restatement of the D3Q19 binary-fluid step
 *
 *  TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 *  the cpu_baseline leg of bench.py may load this library, and only as
 *  the checker or the timed CPU baseline.  The product (libludwig_hip.so)
 *  never links or calls it.
 *
 *  Layout (same as the GPU library's device layout): a periodic lattice of
 *  nx*ny*nz sites with no halo, site = (x*ny + y)*nz + z, z fastest.
 *  Multi-component fields are component major: f[p*nsite + site],
 *  u[a*nsite + site].  Periodicity is applied by index wrap, which gives
 *  the same values as the reference's halo copies on one rank.
 *
 *  Every function follows the reference's operation order so that results
 *  are bit-identical with the reference built by gcc -O2 on x86-64 (no
 *  FMA).  Citations are to files in the reference tree.
 *
 *****************************************************************************/

#ifndef LDW_ORACLE_H
#define LDW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Relaxation schemes, numbered as lb_relaxation_enum_t (lb_data_options.h:20) */
enum {ORA_M10 = 0, ORA_BGK = 1, ORA_TRT = 2};

typedef struct ora_lb_param_s {
  int    relax;            /* ORA_M10 / ORA_BGK / ORA_TRT */
  double eta_shear;        /* shear viscosity */
  double eta_bulk;         /* bulk viscosity */
  double rho0;             /* reference density */
  double fglobal[3];       /* uniform body force */
} ora_lb_param_t;

typedef struct ora_fe_param_s {
  double a, b, kappa;      /* symmetric free energy A, B, kappa */
  double mobility;         /* Cahn-Hilliard mobility M */
  int    order;            /* advection order 1 to 5 */
  int    force_method;     /* 0 stress divergence, 1 phi grad mu */
  int    stencil;          /* gradient: 7 (3d_7pt_fluid, also 0) or 27 (3d_27pt_fluid) */
  int    conserve;         /* cahn_hilliard_options_conserve 0, 1 or 2 */
  double grad_mu[3];       /* external chemical potential gradient (physics grad_mu) */
} ora_fe_param_t;

/* D3Q19 tables: velocities, weights, ma (modes x velocities),
 * mi (velocities x modes), as exact rationals rounded once to double. */
void ora_d3q19_tables(int cv[19][3], double wv[19], double ma[19][19],
		      double mi[19][19]);

/* Relaxation rates as the reference computes them (collision.c:1163-1246) */
void ora_relaxation_rates(const ora_lb_param_t * p, double rtau[19]);

/* Collision at every site: f updated in place; rho, u written.
 * F is the site force (may be NULL for zero). */
void ora_collide(int nx, int ny, int nz, const ora_lb_param_t * p,
		 double * f, const double * F, double * rho, double * u);

/* Pull propagation fout[p][x] = fin[p][x - c_p] (periodic) */
void ora_propagate(int nx, int ny, int nz, const double * fin, double * fout);

/* Equilibrium f from rho, u (lb_data.c:809-840); u may be NULL (rest) */
void ora_equilibrium(int nx, int ny, int nz, const double * rho,
		     const double * u, double rho_uniform, double * f);

/* 7-point gradient and Laplacian (gradient_3d_7pt_fluid.c:231-310) */
/* grad_3d_27pt_kernel (gradient_3d_27pt_fluid.c:219-368) */
void ora_grad27(int nx, int ny, int nz, const double * phi, double * grad,
		double * delsq);
void ora_grad7(int nx, int ny, int nz, const double * phi, double * grad,
	       double * delsq);

/* mu = A phi + B phi^3 - kappa del^2 phi (symmetric.c:308-318) */
void ora_mu(int nsite, const ora_fe_param_t * fe, const double * phi,
	    const double * delsq, double * mu);

/* Chemical stress P_ab, 9 components ab-major (symmetric.c:372-418) */
void ora_stress(int nsite, const ora_fe_param_t * fe, const double * phi,
		const double * grad, const double * delsq, double * pth);

/* Thermodynamic force: method 0 stress divergence
 * (phi_force_colloid.c:315-466), 1 -phi grad mu (phi_grad_mu.c:284-350),
 * then under method 1 -phi grad_mu added when grad_mu is non-zero
 * (phi_grad_mu.c:123-160, 509-534).  F is overwritten (the reference adds
 * to a zeroed force). */
void ora_force(int nx, int ny, int nz, const ora_fe_param_t * fe,
	       const double * phi, const double * grad, const double * delsq,
	       double * F);

/* One Cahn-Hilliard update of phi in place (phi_cahn_hilliard.c:213-290,
 * advection.c, advection_bcs.c), using delsq of the same phi and u:
 * advective fluxes of order 1-5, then -M grad mu, then -M grad_mu
 * (phi_ch_flux_mu_ext, phi_cahn_hilliard.c:1373-1398), then the update. */
void ora_cahn_hilliard(int nx, int ny, int nz, const ora_fe_param_t * fe,
		       double * phi, const double * delsq, const double * u);
/* The same with the conservation option fe->conserve:
 *   1: the update as per-site compensated sums, with the compensation
 *      carried between steps in csum (phi_cahn_hilliard.c:1181-1218,
 *      util_sum.c:30-40);
 *   2: the forward step, then phi -= (S - phi_init_sum)/nsite with S the
 *      plain sum of phi in site order, as the reference computes it on one
 *      thread (phi_cahn_hilliard.c:1102-1175, 1225-1327).
 * csum may be NULL unless conserve is 1. */
void ora_cahn_hilliard_ext(int nx, int ny, int nz, const ora_fe_param_t * fe,
			   double * phi, const double * delsq, const double * u,
			   double * csum, double phi_init_sum);
/* field_init_sum as cahn_hilliard_stats_time0 forms it on one thread
 * (cahn_hilliard_stats.c:59-76, 233-276): a Kahan sum over sites, added
 * as a Kahan pair into a zero sum twice, then sum + cs. */
double ora_phi_init_sum(int64_t n, const double * phi);

/* Full binary-fluid steps in the reference's loop order
 * (ludwig.c:528-860): force, Cahn-Hilliard, collision, propagation.
 * f, phi are in/out; rho, u are in/out (u of the previous step is used by
 * the Cahn-Hilliard advection); scratch needs 8*nsite doubles. */
void ora_binary_steps(int nx, int ny, int nz, int nsteps,
		      const ora_lb_param_t * lb, const ora_fe_param_t * fe,
		      double * f, double * phi, double * rho, double * u);
/* The same with the conservation state of ora_cahn_hilliard_ext */
void ora_binary_steps_ext(int nx, int ny, int nz, int nsteps,
			  const ora_lb_param_t * lb, const ora_fe_param_t * fe,
			  double * f, double * phi, double * rho, double * u,
			  double * csum, double phi_init_sum);

/* Plain fluid steps (free_energy none): collision with uniform force
 * fglobal only, then propagation. */
void ora_single_steps(int nx, int ny, int nz, int nsteps,
		      const ora_lb_param_t * lb, double * f, double * rho,
		      double * u);

/* The reference lattice generator ns_uniform (noise.h:96-109) */
uint32_t ora_ns_uniform(uint32_t s[4]);

/* Per-site generator state for global site (ig, jg, kg), 1-based
 * (noise.c:278-341) */
void ora_noise_state(uint32_t seed, int ig, int jg, int kg, uint32_t s[4]);

/* Spinodal initial phi = phi0 + amp (r - 1/2) (field_phi_init.c:484-518) */
void ora_spinodal(int nx, int ny, int nz, int seed, double phi0, double amp,
		  double * phi);

/* Random velocity u_a = 1e-3 (2 r - 1), three draws per site, mean
 * removed per component; used for the plain-fluid start. */
void ora_random_velocity(int nx, int ny, int nz, int seed, double amp,
			 double * u);

/* Kahan / Klein compensated sums (util_sum.c:30-110) */
double ora_kahan_sum(int64_t n, const double * a);
double ora_klein_sum(int64_t n, const double * a);

/* ------------------------------------------------------------------------ */
/* Lees-Edwards planes (le_oracle.c), steady shear, one rank                */
/* ------------------------------------------------------------------------ */

typedef struct ora_le_param_s {
  int    nplane;           /* N_LE_plane */
  int    nhalo;            /* coordinate halo (2 for the symmetric free energy) */
  double uy;               /* LE_plane_vel */
  int    nt0;              /* LE_time_offset */
} ora_le_param_t;

/* field_grad_compute + phi_force_calculation at physics time t (the
 * flux form, phi_force.c:289-520): F = 0 + force.  Returns -1 for a plane
 * geometry the reference refuses. */
int ora_le_force(int nx, int ny, int nz, const ora_fe_param_t * fe, const ora_le_param_t * le,
		 double t, const double * phi, double * F);

/* field_grad_compute + phi_cahn_hilliard at time t; phi updated in place;
 * u may be NULL (no hydrodynamics) */
int ora_le_cahn_hilliard(int nx, int ny, int nz, const ora_fe_param_t * fe,
			 const ora_le_param_t * le, double t, double * phi, const double * u,
			 double * csum, double phi_init_sum);

/* lb_data_apply_le_boundary_conditions at time step timestep, in place */
int ora_le_lb_bc(int nx, int ny, int nz, const ora_le_param_t * le, double timestep, double * f);

/* lb_le_init_shear_profile: f for rho0 and the shear viscosity eta */
int ora_le_shear_profile(int nx, int ny, int nz, const ora_le_param_t * le, double rho0,
			 double eta, double * f);

/* nsteps of the ludwig.c loop with planes, from the time counter state
 * (t_start, t_current) of physics_control */
int ora_le_binary_steps(int nx, int ny, int nz, int nsteps, int t_start, int t_current,
			const ora_lb_param_t * lb, const ora_fe_param_t * fe,
			const ora_le_param_t * le, double * f, double * phi, double * rho,
			double * u, double * csum, double phi_init_sum);

#ifdef __cplusplus
}
#endif

#endif
