/*****************************************************************************
 *
 *  ref_harness.c  --  golden-vector generator (test infrastructure only)
 *
 *  Drives the reference implementation's own public C API (objects built
 *  from the reference sources by oracle/Makefile into oracle/_ref/) on
 *  seeded inputs and writes the results as raw little-endian float64
 *  arrays.  Nothing here is part of the product; the GPU library never
 *  links or calls it.
 *
 *  Output layout (chosen to match the repository's device layout, not the
 *  reference's): interior sites only, site = (x*ny + y)*nz + z with
 *  x, y, z zero-based, z fastest; multi-component fields are component
 *  major, e.g. f[p*nsites + site], u[a*nsites + site].
 *
 *  Cases (argv[1]):
 *    collide  N relax seed tau           one lb_collide() on random f, F
 *    prop     N seed                     lb_halo() + lb_propagation()
 *    force    N seed method A B K        field_grad_compute +
 *                                        phi_force_calculation
 *    ch       N seed order A B K M       one phi_cahn_hilliard() with
 *                                        random phi, u
 *    binary   N steps seed order A B K M tau phi0 amp
 *                                        the full symmetric FD step loop
 *                                        (ludwig.c:528-860 order)
 *    single   N steps seed tau           plain fluid, random-velocity
 *                                        equilibrium start
 *    lebc     N seed                     lb_data_apply_le_boundary_conditions
 *                                        on random post-collision f
 *
 *  Lees-Edwards planes: LDW_HARNESS_LE="nplanes uy [nt0]" creates the
 *  planes (steady shear, lees_edw_create leesedwards.c:81; nt0 the
 *  LE_time_offset, default 0); LDW_HARNESS_T0=k
 *  advances the physics time by k steps before a single-call case runs
 *  (the plane displacement depends on it); LDW_HARNESS_LE_PROFILE=1 starts
 *  the binary case from lb_le_init_shear_profile (model_le.c:649).  The
 *  binary case applies lb_data_apply_le_boundary_conditions (model_le.c:78)
 *  after each collision, as ludwig.c:810-812 does.
 *
 *  LDW_HARNESS_GRADMU="gx gy gz" sets the external chemical potential
 *  gradient (physics_grad_mu_set, physics.c:587); LDW_HARNESS_METHOD=1
 *  runs the binary case with fe_force_method phi_gradmu;
 *  LDW_HARNESS_NHALO=n gives the order-parameter cases a coordinate halo
 *  of n (3 for advection order 5, advection.c:1264-1430).
 *
 *  Reference functions used (file:line in the reference tree):
 *    lb_collide          collision.c:143
 *    lb_halo             lb_data.c:754
 *    lb_propagation      propagation.c:43
 *    field_grad_compute  field_grad.c:319 (grad_3d_7pt_fluid_d2, or
 *                        grad_3d_27pt_fluid_d2 with LDW_HARNESS_STENCIL=27)
 *    phi_ch_create       phi_cahn_hilliard.c:128, with
 *                        LDW_HARNESS_CONSERVE=1|2 as the conserve option
 *    cahn_hilliard_stats_time0 cahn_hilliard_stats.c:59 (conserve 2)
 *    phi_force_calculation phi_force.c:74
 *    phi_cahn_hilliard   phi_cahn_hilliard.c:213
 *    field_phi_init_spinodal field_phi_init.c:484
 *    lb_1st_moment_equilib_set lb_data.c:809
 *
 *****************************************************************************/

#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pe.h"
#include "coords.h"
#include "physics.h"
#include "lb_data.h"
#include "collision.h"
#include "propagation.h"
#include "hydro.h"
#include "map.h"
#include "field.h"
#include "field_grad.h"
#include "gradient_3d_7pt_fluid.h"
#include "gradient_3d_27pt_fluid.h"
#include "symmetric.h"
#include "phi_force.h"
#include "phi_force_stress.h"
#include "phi_cahn_hilliard.h"
#include "cahn_hilliard_stats.h"
#include "advection.h"
#include "noise.h"
#include "leesedwards.h"
#include "wall.h"
#include "fe_force_method.h"
#include "field_phi_init.h"
#include "model_le.h"

/* splitmix64: input generator for the harness only */

static uint64_t sm_state = 0;

static uint64_t sm_next(void) {
  uint64_t z = (sm_state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static double sm_uniform(void) {            /* [0, 1) with 53 bits */
  return (sm_next() >> 11) * (1.0/9007199254740992.0);
}

static const char * outdir = ".";

static void usage(void);

/* LDW_HARNESS_T0=k: the physics time of a single-call case */
static void advance_time(physics_t * phys) {
  const char * t0 = getenv("LDW_HARNESS_T0");
  int k = t0 ? atoi(t0) : 0;
  for (int i = 0; i < k; i++) physics_control_next_step(phys);
}

static void dump(const char * name, const double * a, size_t n) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s.f64", outdir, name);
  FILE * fp = fopen(path, "wb");
  if (fp == NULL) { perror(path); exit(2); }
  if (fwrite(a, sizeof(double), n, fp) != n) { perror(path); exit(2); }
  fclose(fp);
}

typedef struct {
  pe_t * pe;
  cs_t * cs;
  physics_t * phys;
  lees_edw_t * le;
  lb_t * lb;
  hydro_t * hydro;
  map_t * map;
  wall_t * wall;
  field_t * phi;
  field_grad_t * dphi;
  fe_symm_t * fe;
  pth_t * pth;
  phi_ch_t * pch;
  int n;
  int nsites;                  /* interior sites n^3 */
} sys_t;

static int site_of(const sys_t * s, int ic, int jc, int kc) {
  return ((ic - 1)*s->n + (jc - 1))*s->n + (kc - 1);
}

static void sys_create(sys_t * s, int n, int nhalo, lb_relaxation_enum_t relax,
		       double eta, int with_phi, int force_method) {

  memset(s, 0, sizeof(*s));
  s->n = n;
  s->nsites = n*n*n;
  {
    /* LDW_HARNESS_NHALO=3: a deeper coordinate halo for the order
     * parameter cases (advection order 5 reads phi three sites away) */
    const char * nh = getenv("LDW_HARNESS_NHALO");
    if (nh && with_phi) nhalo = atoi(nh);
  }

  pe_create(MPI_COMM_WORLD, PE_QUIET, &s->pe);
  physics_create(s->pe, &s->phys);
  physics_eta_shear_set(s->phys, eta);
  physics_eta_bulk_set(s->phys, eta);
  physics_rho0_set(s->phys, 1.0);
  {
    /* LDW_HARNESS_GRADMU="gx gy gz": the external chemical potential
     * gradient (input key grad_mu) */
    const char * gm = getenv("LDW_HARNESS_GRADMU");
    if (gm) {
      double v[3] = {0.0, 0.0, 0.0};
      if (sscanf(gm, "%lf %lf %lf", v, v + 1, v + 2) != 3) usage();
      physics_grad_mu_set(s->phys, v);
    }
  }

  cs_create(s->pe, &s->cs);
  {
    int ntotal[3] = {n, n, n};
    cs_nhalo_set(s->cs, nhalo);
    cs_ntotal_set(s->cs, ntotal);
  }
  cs_init(s->cs);

  {
    lees_edw_options_t opts = {0};
    const char * le = getenv("LDW_HARNESS_LE");
    if (le) {
      opts.type = LE_SHEAR_TYPE_STEADY;
      const int nread = sscanf(le, "%d %lf %d", &opts.nplanes, &opts.uy, &opts.nt0);
      if (nread < 2) usage();
    }
    lees_edw_create(s->pe, s->cs, &opts, &s->le);
  }
  {
    lb_data_options_t opts = lb_data_options_default();
    opts.ndim = 3; opts.nvel = 19; opts.ndist = 1;
    opts.halo = LB_HALO_FULL;
    lb_data_create(s->pe, s->cs, &opts, &s->lb);
    lb_collision_relaxation_set(s->lb, relax);
  }
  {
    hydro_options_t opts = hydro_options_default();
    hydro_create(s->pe, s->cs, s->le, &opts, &s->hydro);
  }
  {
    map_options_t opts = map_options_default();
    map_create(s->pe, s->cs, &opts, &s->map);
  }
  wall_create(s->pe, s->cs, s->map, s->lb, &s->wall);

  if (with_phi) {
    field_options_t opts = field_options_ndata_nhalo(1, nhalo);
    field_create(s->pe, s->cs, s->le, "phi", &opts, &s->phi);
    field_grad_create(s->pe, s->phi, 2, &s->dphi);
    {
      /* LDW_HARNESS_STENCIL=27 selects the 27-point gradient */
      const char * st = getenv("LDW_HARNESS_STENCIL");
      if (st && atoi(st) == 27) field_grad_set(s->dphi, grad_3d_27pt_fluid_d2, NULL);
      else field_grad_set(s->dphi, grad_3d_7pt_fluid_d2, NULL);
    }
    fe_symm_create(s->pe, s->cs, s->phi, s->dphi, &s->fe);
    pth_create(s->pe, s->cs, force_method, &s->pth);
    {
      /* LDW_HARNESS_CONSERVE=1|2 selects cahn_hilliard_options_conserve */
      const char * cv = getenv("LDW_HARNESS_CONSERVE");
      phi_ch_info_t info = {0};
      if (cv) info.conserve = atoi(cv);
      phi_ch_create(s->pe, s->cs, s->le, &info, &s->pch);
    }
  }
}

static void fe_params(sys_t * s, double a, double b, double kappa) {
  fe_symm_param_t p = {.a = a, .b = b, .kappa = kappa, .c = 0.0, .h = 0.0};
  fe_symm_param_set(s->fe, p);
}

/* Extraction of interior values into the output layout */

static double * get_f(sys_t * s) {
  double * a = (double *) malloc(sizeof(double)*19*s->nsites);
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	int index = cs_index(s->cs, ic, jc, kc);
	for (int p = 0; p < 19; p++) {
	  double v;
	  lb_f(s->lb, index, p, 0, &v);
	  a[(size_t) p*s->nsites + site_of(s, ic, jc, kc)] = v;
	}
      }
  return a;
}

static void set_f(sys_t * s, const double * a) {
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	int index = cs_index(s->cs, ic, jc, kc);
	for (int p = 0; p < 19; p++) {
	  lb_f_set(s->lb, index, p, 0, a[(size_t) p*s->nsites + site_of(s, ic, jc, kc)]);
	}
      }
}

static double * get_field(sys_t * s, field_t * fld, int nf) {
  double * a = (double *) malloc(sizeof(double)*nf*s->nsites);
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	int index = cs_index(s->cs, ic, jc, kc);
	for (int m = 0; m < nf; m++) {
	  a[(size_t) m*s->nsites + site_of(s, ic, jc, kc)]
	    = fld->data[addr_rank1(fld->nsites, nf, index, m)];
	}
      }
  return a;
}

static void set_field(sys_t * s, field_t * fld, int nf, const double * a) {
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	int index = cs_index(s->cs, ic, jc, kc);
	for (int m = 0; m < nf; m++) {
	  fld->data[addr_rank1(fld->nsites, nf, index, m)]
	    = a[(size_t) m*s->nsites + site_of(s, ic, jc, kc)];
	}
      }
}

static double * get_grad(sys_t * s, int which) {
  /* which = 0: grad (3 comps), 1: delsq */
  int nc = (which == 0) ? 3 : 1;
  double * a = (double *) malloc(sizeof(double)*nc*s->nsites);
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	int index = cs_index(s->cs, ic, jc, kc);
	int site = site_of(s, ic, jc, kc);
	if (which == 0) {
	  double g[3];
	  field_grad_scalar_grad(s->dphi, index, g);
	  for (int m = 0; m < 3; m++) a[(size_t) m*s->nsites + site] = g[m];
	}
	else {
	  double d;
	  field_grad_scalar_delsq(s->dphi, index, &d);
	  a[site] = d;
	}
      }
  return a;
}

static double * get_mu(sys_t * s) {
  double * a = (double *) malloc(sizeof(double)*s->nsites);
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	double mu;
	fe_symm_mu(s->fe, cs_index(s->cs, ic, jc, kc), &mu);
	a[site_of(s, ic, jc, kc)] = mu;
      }
  return a;
}

static void dump_hydro(sys_t * s, const char * prefix) {
  char name[256];
  double * rho = get_field(s, s->hydro->rho, 1);
  double * u = get_field(s, s->hydro->u, 3);
  snprintf(name, sizeof(name), "%srho", prefix); dump(name, rho, s->nsites);
  snprintf(name, sizeof(name), "%su", prefix); dump(name, u, 3*(size_t) s->nsites);
  free(rho); free(u);
}

/* Random f near equilibrium: rho in [0.9,1.1], |u| < 0.05, plus 1e-3
 * non-equilibrium noise on each f_p (so ghost modes are non-zero). */

static double * random_f(sys_t * s) {
  double * a = (double *) malloc(sizeof(double)*19*s->nsites);
  for (int ic = 1; ic <= s->n; ic++)
    for (int jc = 1; jc <= s->n; jc++)
      for (int kc = 1; kc <= s->n; kc++) {
	int index = cs_index(s->cs, ic, jc, kc);
	double rho = 0.9 + 0.2*sm_uniform();
	double u[3];
	for (int a3 = 0; a3 < 3; a3++) u[a3] = 0.1*(sm_uniform() - 0.5);
	lb_1st_moment_equilib_set(s->lb, index, rho, u);
	for (int p = 0; p < 19; p++) {
	  double v;
	  lb_f(s->lb, index, p, 0, &v);
	  v += 1.0e-3*(sm_uniform() - 0.5);
	  a[(size_t) p*s->nsites + site_of(s, ic, jc, kc)] = v;
	}
      }
  set_f(s, a);
  return a;
}

static double * random_field(sys_t * s, int nf, double centre, double amp) {
  double * a = (double *) malloc(sizeof(double)*nf*s->nsites);
  for (size_t i = 0; i < (size_t) nf*s->nsites; i++) {
    a[i] = centre + amp*(sm_uniform() - 0.5);
  }
  return a;
}

static void usage(void) {
  fprintf(stderr, "usage: ref_harness outdir case args...\n");
  exit(1);
}

int main(int argc, char ** argv) {

  int provided = 0;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);

  if (argc < 3) usage();
  outdir = argv[1];
  const char * which = argv[2];
  sys_t s;

  if (strcmp(which, "collide") == 0) {
    /* N relax seed tau */
    if (argc < 7) usage();
    int n = atoi(argv[3]);
    int relax = atoi(argv[4]);
    sm_state = strtoull(argv[5], NULL, 10);
    double tau = atof(argv[6]);
    double eta = (tau - 0.5)/3.0;
    sys_create(&s, n, 1, (lb_relaxation_enum_t) relax, eta, 0, 0);
    double * f0 = random_f(&s);
    double * F = random_field(&s, 3, 0.0, 2.0e-4);
    set_field(&s, s.hydro->force, 3, F);
    {
      double uzero[3] = {0};
      hydro_u_zero(s.hydro, uzero);
    }
    lb_collide(s.lb, s.hydro, s.map, NULL, NULL, NULL);
    double * f1 = get_f(&s);
    dump("f_in", f0, 19*(size_t) s.nsites);
    dump("force", F, 3*(size_t) s.nsites);
    dump("f_out", f1, 19*(size_t) s.nsites);
    dump_hydro(&s, "");
    {
      double p[2] = {eta, eta};
      dump("eta", p, 2);
    }
  }
  else if (strcmp(which, "prop") == 0) {
    if (argc < 5) usage();
    int n = atoi(argv[3]);
    sm_state = strtoull(argv[4], NULL, 10);
    sys_create(&s, n, 1, LB_RELAXATION_M10, 0.1, 0, 0);
    double * f0 = random_f(&s);
    lb_halo(s.lb);
    lb_propagation(s.lb);
    double * f1 = get_f(&s);
    dump("f_in", f0, 19*(size_t) s.nsites);
    dump("f_out", f1, 19*(size_t) s.nsites);
  }
  else if (strcmp(which, "force") == 0) {
    /* N seed method A B K */
    if (argc < 9) usage();
    int n = atoi(argv[3]);
    sm_state = strtoull(argv[4], NULL, 10);
    int method = atoi(argv[5]);
    double a = atof(argv[6]), b = atof(argv[7]), k = atof(argv[8]);
    int fm = (method == 0) ? FE_FORCE_METHOD_STRESS_DIVERGENCE
                           : FE_FORCE_METHOD_PHI_GRADMU;
    sys_create(&s, n, 2, LB_RELAXATION_M10, 0.1, 1, fm);
    fe_params(&s, a, b, k);
    double * phi = random_field(&s, 1, 0.0, 1.6);
    set_field(&s, s.phi, 1, phi);
    {
      double fzero[3] = {0};
      hydro_f_zero(s.hydro, fzero);
    }
    advance_time(s.phys);
    field_halo(s.phi);
    field_grad_compute(s.dphi);
    phi_force_calculation(s.pe, s.cs, s.le, s.wall, s.pth, (fe_t *) s.fe,
			  s.map, s.phi, s.hydro);
    double * g = get_grad(&s, 0);
    double * d = get_grad(&s, 1);
    double * mu = get_mu(&s);
    double * F = get_field(&s, s.hydro->force, 3);
    dump("phi", phi, s.nsites);
    dump("grad", g, 3*(size_t) s.nsites);
    dump("delsq", d, s.nsites);
    dump("mu", mu, s.nsites);
    dump("force", F, 3*(size_t) s.nsites);
  }
  else if (strcmp(which, "ch") == 0) {
    /* N seed order A B K M */
    if (argc < 10) usage();
    int n = atoi(argv[3]);
    sm_state = strtoull(argv[4], NULL, 10);
    int order = atoi(argv[5]);
    double a = atof(argv[6]), b = atof(argv[7]), k = atof(argv[8]);
    double mob = atof(argv[9]);
    sys_create(&s, n, 2, LB_RELAXATION_M10, 0.1, 1, FE_FORCE_METHOD_STRESS_DIVERGENCE);
    physics_mobility_set(s.phys, mob);
    advection_order_set(order);
    fe_params(&s, a, b, k);
    double * phi = random_field(&s, 1, 0.0, 1.6);
    double * u = random_field(&s, 3, 0.0, 0.02);
    set_field(&s, s.phi, 1, phi);
    set_field(&s, s.hydro->u, 3, u);
    advance_time(s.phys);
    field_halo(s.phi);
    field_grad_compute(s.dphi);
    phi_cahn_hilliard(s.pch, (fe_t *) s.fe, s.phi, s.hydro, s.map, NULL);
    double * phi1 = get_field(&s, s.phi, 1);
    dump("phi_in", phi, s.nsites);
    dump("u", u, 3*(size_t) s.nsites);
    dump("phi_out", phi1, s.nsites);
  }
  else if (strcmp(which, "binary") == 0) {
    /* N steps seed order A B K M tau phi0 amp [dumpevery] */
    if (argc < 14) usage();
    int n = atoi(argv[3]);
    int nsteps = atoi(argv[4]);
    int seed = atoi(argv[5]);
    int order = atoi(argv[6]);
    double a = atof(argv[7]), b = atof(argv[8]), k = atof(argv[9]);
    double mob = atof(argv[10]);
    double tau = atof(argv[11]);
    double phi0 = atof(argv[12]), amp = atof(argv[13]);
    int dumpevery = (argc > 14) ? atoi(argv[14]) : 0;
    double eta = (tau - 0.5)/3.0;
    /* LDW_HARNESS_METHOD=1 selects fe_force_method phi_gradmu */
    const char * fmv = getenv("LDW_HARNESS_METHOD");
    int fm = (fmv && atoi(fmv) == 1) ? FE_FORCE_METHOD_PHI_GRADMU : FE_FORCE_METHOD_STRESS_DIVERGENCE;
    sys_create(&s, n, 2, LB_RELAXATION_M10, eta, 1, fm);
    physics_mobility_set(s.phys, mob);
    advection_order_set(order);
    fe_params(&s, a, b, k);
    field_phi_init_spinodal(s.phi, seed, phi0, amp);
    lb_init_rest_f(s.lb, 1.0);
    {
      const char * pr = getenv("LDW_HARNESS_LE_PROFILE");
      if (pr && atoi(pr) == 1) lb_le_init_shear_profile(s.lb, s.le);
    }
    {
      double * p0 = get_field(&s, s.phi, 1);
      dump("phi_init", p0, s.nsites);
      free(p0);
    }
    /* the reference sum of conserve 2 (ludwig.c:420-425) */
    if (s.pch->info.conserve == 2) {
      cahn_hilliard_stats_time0(s.pch, s.phi, s.map);
      dump("phi_init_sum", &s.phi->field_init_sum, 1);
    }
    for (int step = 1; step <= nsteps; step++) {
      double fzero[3] = {0}, uzero[3] = {0};
      physics_control_next_step(s.phys);
      hydro_f_zero(s.hydro, fzero);
      field_halo(s.phi);
      field_grad_compute(s.dphi);
      phi_force_calculation(s.pe, s.cs, s.le, s.wall, s.pth, (fe_t *) s.fe,
			    s.map, s.phi, s.hydro);
      phi_cahn_hilliard(s.pch, (fe_t *) s.fe, s.phi, s.hydro, s.map, NULL);
      hydro_u_zero(s.hydro, uzero);
      lb_collide(s.lb, s.hydro, s.map, NULL, (fe_t *) s.fe, NULL);
      lb_data_apply_le_boundary_conditions(s.lb, s.le);
      lb_halo(s.lb);
      lb_propagation(s.lb);
      if (dumpevery > 0 && (step % dumpevery) == 0 && step != nsteps) {
	char name[64];
	double * p = get_field(&s, s.phi, 1);
	snprintf(name, sizeof(name), "phi_%d", step);
	dump(name, p, s.nsites);
	free(p);
      }
    }
    double * p = get_field(&s, s.phi, 1);
    double * f = get_f(&s);
    dump("phi", p, s.nsites);
    dump("f", f, 19*(size_t) s.nsites);
    dump_hydro(&s, "");
    if (s.pch->info.conserve == 1) {
      double * cs = get_field(&s, s.pch->csum, 1);
      dump("csum", cs, s.nsites);
      free(cs);
    }
  }
  else if (strcmp(which, "single") == 0) {
    /* N steps seed tau */
    if (argc < 7) usage();
    int n = atoi(argv[3]);
    int nsteps = atoi(argv[4]);
    int seed = atoi(argv[5]);
    double tau = atof(argv[6]);
    double eta = (tau - 0.5)/3.0;
    sys_create(&s, n, 1, LB_RELAXATION_M10, eta, 0, 0);
    {
      /* u_a = 1e-3 (2r - 1), three successive draws of the site's
       * generator, per-component mean subtracted, then equilibrium f. */
      noise_t * rng = NULL;
      double * u = (double *) malloc(sizeof(double)*3*s.nsites);
      double mean[3] = {0};
      noise_create_seed(s.pe, s.cs, seed, &rng);
      for (int ic = 1; ic <= n; ic++)
	for (int jc = 1; jc <= n; jc++)
	  for (int kc = 1; kc <= n; kc++) {
	    int index = cs_index(s.cs, ic, jc, kc);
	    for (int ia = 0; ia < 3; ia++) {
	      double r;
	      noise_uniform_double_reap(rng, index, &r);
	      u[(size_t) ia*s.nsites + site_of(&s, ic, jc, kc)] = 1.0e-3*(2.0*r - 1.0);
	    }
	  }
      for (int ia = 0; ia < 3; ia++) {
	for (int i = 0; i < s.nsites; i++) mean[ia] += u[(size_t) ia*s.nsites + i];
	mean[ia] = mean[ia]/s.nsites;
	for (int i = 0; i < s.nsites; i++) u[(size_t) ia*s.nsites + i] -= mean[ia];
      }
      for (int ic = 1; ic <= n; ic++)
	for (int jc = 1; jc <= n; jc++)
	  for (int kc = 1; kc <= n; kc++) {
	    int site = site_of(&s, ic, jc, kc);
	    double u0[3] = {u[site], u[s.nsites + site], u[2*(size_t) s.nsites + site]};
	    lb_1st_moment_equilib_set(s.lb, cs_index(s.cs, ic, jc, kc), 1.0, u0);
	  }
      dump("u_init", u, 3*(size_t) s.nsites);
      {
	double * f0 = get_f(&s);
	dump("f_init", f0, 19*(size_t) s.nsites);
	free(f0);
      }
      noise_free(&rng);
      free(u);
    }
    for (int step = 1; step <= nsteps; step++) {
      double fzero[3] = {0}, uzero[3] = {0};
      physics_control_next_step(s.phys);
      hydro_f_zero(s.hydro, fzero);
      hydro_u_zero(s.hydro, uzero);
      lb_collide(s.lb, s.hydro, s.map, NULL, NULL, NULL);
      lb_halo(s.lb);
      lb_propagation(s.lb);
    }
    double * f = get_f(&s);
    dump("f", f, 19*(size_t) s.nsites);
    dump_hydro(&s, "");
  }
  else if (strcmp(which, "lebc") == 0) {
    /* N seed */
    if (argc < 5) usage();
    int n = atoi(argv[3]);
    sm_state = strtoull(argv[4], NULL, 10);
    sys_create(&s, n, 2, LB_RELAXATION_M10, 0.1, 0, 0);
    double * f0 = random_f(&s);
    advance_time(s.phys);
    lb_data_apply_le_boundary_conditions(s.lb, s.le);
    double * f1 = get_f(&s);
    dump("f_in", f0, 19*(size_t) s.nsites);
    dump("f_out", f1, 19*(size_t) s.nsites);
  }
  else {
    usage();
  }

  MPI_Finalize();
  return 0;
}
