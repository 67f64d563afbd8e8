/*****************************************************************************
 *
 *  ludwig_hip_main.c -- host driver: an unchanged Ludwig input file in,
 *  Ludwig's statistics lines out, the time steps on the GPU.
 *
 *  Reads the run-time input file ("input" in the working directory, as
 *  main.c:38-49 of the reference, or the path given as the only argument)
 *  with the reference's key/value rules (runtime.c:150-190, 611-650):
 *  one "key value" per line, '#' comments, duplicate keys fatal.  Builds
 *  the object graph as ludwig_rt() does for the supported branch
 *  (ludwig.c:202-429, 1175-1262), runs the time loop through the C ABI of
 *  libludwig_hip.so and prints the statistics blocks of
 *  ludwig_report_statistics / ludwig_report_momentum / stats_velocity_minmax
 *  (ludwig.c:978-992, 2391-2477, 1073-1124; stats_distribution.c:55-120;
 *  cahn_hilliard_stats.c:84-115; stats_free_energy.c:38-135;
 *  stats_velocity.c:54-122) in the same format, so logs compare with the
 *  reference's regression logs token by token.
 *
 *  Supported keys (others are reported as unused at the end, as
 *  ludwig.c:1056 does):
 *    N_cycles N_start size grid periodicity boundary_walls
 *    viscosity viscosity_bulk fluid_rho0
 *    free_energy (none | symmetric)  A B K | symmetric_a symmetric_b
 *    symmetric_kappa  mobility  phi0  phi_initialisation (spinodal, drop)
 *    phi_init_drop_radius phi_init_drop_amplitude phi_init_drop_centred
 *    noise  random_seed  fd_gradient_calculation (3d_7pt_fluid, 3d_27pt_fluid)
 *    fd_advection_scheme_order  fe_force_method  lb_relaxation_scheme
 *    ghost_modes  lb_ghost_modes  cahn_hilliard_options_conserve
 *    distribution_initialisation (3d_uniform_u) distribution_uniform_u
 *    freq_statistics  freq_config  config_at_end  colloid_init  hydrodynamics
 *    isothermal_fluctuations lb_fluctuations temperature
 *    {default, default_input, default_output, lb, lb_input, lb_output, phi,
 *     rho, vel, ...}_io_{mode, format, report}, {default, lb, phi, rho,
 *     vel}_io_{grid, freq}   (io_info_args_rt.c:33-140, io_options_rt.c)
 *
 *  Restart (N_start > 0) reads the distribution, phi, rho and vel files of
 *  step N_start from the working directory (ludwig.c:313-364); a
 *  configuration step (freq_config, config_at_end; control.c:138-147)
 *  writes them (ludwig.c:868-962), in the reference's file format.
 *
 *****************************************************************************/

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ludwig_hip.h"

/* ------------------------------------------------------------------------ */
/* Input file                                                               */
/* ------------------------------------------------------------------------ */

#define MAXKEYS 512

typedef struct {
  char key[128];
  char val[512];
  int used;
  int line;
} kv_t;

static kv_t keys[MAXKEYS];
static int nkeys = 0;

static void fatal(const char * msg, const char * arg) {
  fprintf(stdout, "%s%s\n", msg, arg ? arg : "");
  fflush(stdout);
  exit(1);
}

static void read_input(const char * path) {
  FILE * fp = fopen(path, "r");
  char line[1024];
  int nline = 0;
  if (fp == NULL) fatal("Input file not found: ", path);
  while (fgets(line, sizeof(line), fp)) {
    char k[128], v[512];
    nline++;
    char * hash = strchr(line, '#');
    if (hash) *hash = '\0';
    v[0] = '\0';
    int n = sscanf(line, "%127s %511[^\n]", k, v);
    if (n < 1) continue;
    /* trim trailing blanks of the value */
    size_t lv = strlen(v);
    while (lv > 0 && (v[lv - 1] == ' ' || v[lv - 1] == '\t' || v[lv - 1] == '\r')) v[--lv] = '\0';
    for (int i = 0; i < nkeys; i++) {
      if (strcmp(keys[i].key, k) == 0) {
	fprintf(stdout, "Duplicate key '%s' at lines %d and %d\n", k, keys[i].line, nline);
	fatal("Please check the input file and try again.", NULL);
      }
    }
    if (nkeys == MAXKEYS) fatal("Too many keys in input", NULL);
    snprintf(keys[nkeys].key, sizeof(keys[nkeys].key), "%s", k);
    snprintf(keys[nkeys].val, sizeof(keys[nkeys].val), "%s", v);
    keys[nkeys].used = 0;
    keys[nkeys].line = nline;
    nkeys++;
  }
  fclose(fp);
}

static const char * get(const char * key) {
  for (int i = 0; i < nkeys; i++) {
    if (strcmp(keys[i].key, key) == 0) {
      keys[i].used = 1;
      return keys[i].val;
    }
  }
  return NULL;
}

static int get_int(const char * key, int * v) {
  const char * s = get(key);
  if (s == NULL) return 0;
  *v = atoi(s);
  return 1;
}

static int get_double(const char * key, double * v) {
  const char * s = get(key);
  if (s == NULL) return 0;
  *v = atof(s);
  return 1;
}

static int get_string(const char * key, char * out, size_t n) {
  const char * s = get(key);
  if (s == NULL) return 0;
  snprintf(out, n, "%s", s);
  return 1;
}

/* "a_b_c" vectors (runtime.c) */
static int get_ivec(const char * key, int v[3]) {
  const char * s = get(key);
  if (s == NULL) return 0;
  if (sscanf(s, "%d_%d_%d", v, v + 1, v + 2) != 3) fatal("Bad integer vector for key ", key);
  return 1;
}

static int get_dvec(const char * key, double v[3]) {
  const char * s = get(key);
  if (s == NULL) return 0;
  if (sscanf(s, "%lf_%lf_%lf", v, v + 1, v + 2) != 3) fatal("Bad vector for key ", key);
  return 1;
}

/* runtime.c:535-554 */
static int switch_on(const char * key) {
  char s[64] = "";
  if (!get_string(key, s, sizeof(s))) return 0;
  return (strcmp(s, "Yes") == 0 || strcmp(s, "yes") == 0 || strcmp(s, "1") == 0
	  || strcmp(s, "on") == 0 || strcmp(s, "true") == 0);
}

static void lower(char * s) {
  for (; *s; s++) if (*s >= 'A' && *s <= 'Z') *s = (char) (*s - 'A' + 'a');
}

/* io_options_rt.c:38-180: "<stub>_io_mode", "<stub>_io_format",
 * "<stub>_io_report", where present */
static void io_options_keys(const char * stub, io_options_t * o) {
  char key[192], s[64];
  snprintf(key, sizeof(key), "%s_io_mode", stub);
  if (get_string(key, s, sizeof(s))) {
    lower(s);
    if (strcmp(s, "mpiio") != 0) fatal("I/O mode not recognised (mpiio is) for key ", key);
    o->mode = IO_MODE_MPIIO;
  }
  snprintf(key, sizeof(key), "%s_io_format", stub);
  if (get_string(key, s, sizeof(s))) {
    lower(s);
    if (strcmp(s, "ascii") == 0) o->iorformat = IO_RECORD_ASCII;
    else if (strcmp(s, "binary") == 0) o->iorformat = IO_RECORD_BINARY;
    else fatal("I/O record format should be either 'ascii' or 'binary' for key ", key);
  }
  snprintf(key, sizeof(key), "%s_io_report", stub);
  if (get(key)) o->report = switch_on(key);
}

/* io_info_args_rt.c:33-140: input and output options in increasing order
 * of precedence, then the shared I/O grid and the output frequency */
static io_info_args_t io_args_keys(const char * stub) {
  io_info_args_t a = io_info_args_default();
  char k[160];
  io_options_keys("default", &a.input);
  io_options_keys("default_input", &a.input);
  io_options_keys(stub, &a.input);
  snprintf(k, sizeof(k), "%s_input", stub);
  io_options_keys(k, &a.input);
  io_options_keys("default", &a.output);
  io_options_keys("default_output", &a.output);
  io_options_keys(stub, &a.output);
  snprintf(k, sizeof(k), "%s_output", stub);
  io_options_keys(k, &a.output);
  get_ivec("default_io_grid", a.grid);
  snprintf(k, sizeof(k), "%s_io_grid", stub);
  get_ivec(k, a.grid);
  if (a.grid[0] != 1 || a.grid[1] != 1 || a.grid[2] != 1) {
    fatal("One GPU writes one file: the I/O grid must be 1_1_1 for ", stub);
  }
  for (int d = 0; d < 3; d++) a.input.iogrid[d] = a.output.iogrid[d] = a.grid[d];
  get_int("default_io_freq", &a.iofreq);
  snprintf(k, sizeof(k), "%s_io_freq", stub);
  get_int(k, &a.iofreq);
  if (a.iofreq < 0) fatal("The I/O frequency must be a non-negative integer for ", stub);
  return a;
}

/* util.c:1082-1085 */
static int util_mod(int a, int b) { return (b == 0) ? a : a % b; }

/* ------------------------------------------------------------------------ */
/* Statistics in the reference's formats                                    */
/* ------------------------------------------------------------------------ */

static int nx, ny, nz;

static double dmin(double a, double b) { return (a < b) ? a : b; }
static double dmax(double a, double b) { return (a > b) ? a : b; }

/* stats_distribution.c:55-120 */
static void stats_rho(lb_t * lb, cs_t * cs) {
  double s[5] = {0.0, 0.0, 0.0, +DBL_MAX, -DBL_MAX};
  for (int ic = 1; ic <= nx; ic++)
    for (int jc = 1; jc <= ny; jc++)
      for (int kc = 1; kc <= nz; kc++) {
	double rho;
	lb_0th_moment(lb, cs_index(cs, ic, jc, kc), LB_RHO, &rho);
	s[0] += 1.0;
	s[1] += rho;
	s[2] += rho*rho;
	s[3] = dmin(rho, s[3]);
	s[4] = dmax(rho, s[4]);
      }
  double mean = s[1]/s[0];
  double var = (s[2]/s[0]) - mean*mean;
  printf("\nScalars - total mean variance min max\n");
  printf("[rho] %14.2f %14.11f %14.7e %14.11f %14.11f\n", s[1], mean, fabs(var), s[3], s[4]);
}

/* cahn_hilliard_stats.c:84-115 with the Kahan sum of util_sum.c:30-40 */
static void stats_phi(field_t * phi, cs_t * cs) {
  volatile double ksum = 0.0, kcs = 0.0;
  double var = 0.0, vmin = +FLT_MAX, vmax = -FLT_MAX, vol = 0.0;
  for (int ic = 1; ic <= nx; ic++)
    for (int jc = 1; jc <= ny; jc++)
      for (int kc = 1; kc <= nz; kc++) {
	double p;
	field_scalar(phi, cs_index(cs, ic, jc, kc), &p);
	volatile double y = p + kcs;
	volatile double t = ksum + y;
	kcs = y - (t - ksum);
	ksum = t;
	var += p*p;
	vmin = dmin(vmin, p);
	vmax = dmax(vmax, p);
	vol += 1.0;
      }
  double sum = ksum + kcs;
  double rvol = 1.0/vol;
  double fbar = rvol*sum;
  double fvar = rvol*var - fbar*fbar;
  printf("[phi] %14.7e %14.7e%14.7e %14.7e%14.7e\n", sum, fbar, fvar, vmin, vmax);
}

/* stats_free_energy.c:38-135, fluid only */
static void stats_fed(fe_symm_t * fe, cs_t * cs, int step) {
  double total = 0.0, fluid = 0.0, vol = 0.0;
  for (int ic = 1; ic <= nx; ic++)
    for (int jc = 1; jc <= ny; jc++)
      for (int kc = 1; kc <= nz; kc++) {
	double fed;
	fe_symm_fed(fe, cs_index(cs, ic, jc, kc), &fed);
	total += fed;
	fluid += fed;
	vol += 1.0;
      }
  double rv = 1.0/((double) nx*(double) ny*(double) nz);
  printf("\nFree energy density - timestep total fluid\n");
  printf("[fed] %14d %17.10e %17.10e\n", step, rv*total, fluid/vol);
}

/* ludwig.c:1073-1124 with the momentum of stats_distribution.c:281-330:
 * every term f_p c_p (p >= 1) is added to a Kahan sum per component */
static void stats_momentum(lb_t * lb, cs_t * cs) {
  static const int cv[19][3] = {
    { 0,  0,  0}, { 1,  1,  0}, { 1,  0,  1}, { 1,  0,  0}, { 1,  0, -1},
    { 1, -1,  0}, { 0,  1,  1}, { 0,  1,  0}, { 0,  1, -1}, { 0,  0,  1},
    { 0,  0, -1}, { 0, -1,  1}, { 0, -1,  0}, { 0, -1, -1}, {-1,  1,  0},
    {-1,  0,  1}, {-1,  0,  0}, {-1,  0, -1}, {-1, -1,  0}};
  volatile double g[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  for (int ic = 1; ic <= nx; ic++)
    for (int jc = 1; jc <= ny; jc++)
      for (int kc = 1; kc <= nz; kc++) {
	int index = cs_index(cs, ic, jc, kc);
	for (int p = 1; p < 19; p++) {
	  double f;
	  lb_f(lb, index, p, 0, &f);
	  for (int a = 0; a < 3; a++) {
	    double term = f*cv[p][a];
	    volatile double y = term + c[a];
	    volatile double t = g[a] + y;
	    c[a] = y - (t - g[a]);
	    g[a] = t;
	  }
	}
      }
  double gt[3];
  for (int a = 0; a < 3; a++) gt[a] = g[a] + c[a];
  printf("\nMomentum - x y z\n");
  printf("[total   ] %14.7e %14.7e %14.7e\n", gt[0], gt[1], gt[2]);
  printf("[fluid   ] %14.7e %14.7e %14.7e\n", gt[0], gt[1], gt[2]);
}

/* stats_velocity.c:54-122; note umax starts at FLT_MIN as there */
static void stats_velocity(hydro_t * hydro, cs_t * cs) {
  double umin[3], umax[3];
  for (int a = 0; a < 3; a++) { umin[a] = FLT_MAX; umax[a] = FLT_MIN; }
  for (int ic = 1; ic <= nx; ic++)
    for (int jc = 1; jc <= ny; jc++)
      for (int kc = 1; kc <= nz; kc++) {
	double u[3];
	hydro_u(hydro, cs_index(cs, ic, jc, kc), u);
	for (int a = 0; a < 3; a++) {
	  umin[a] = dmin(umin[a], u[a]);
	  umax[a] = dmax(umax[a], u[a]);
	}
      }
  printf("\nVelocity - x y z\n");
  printf("[minimum ] %14.7e %14.7e %14.7e\n", umin[0], umin[1], umin[2]);
  printf("[maximum ] %14.7e %14.7e %14.7e\n", umax[0], umax[1], umax[2]);
}

/* ------------------------------------------------------------------------ */

static double wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1.0e-9*ts.tv_nsec;
}

int main(int argc, char ** argv) {

  const char * path = (argc > 1) ? argv[1] : "input";
  read_input(path);

  int ncycles = 0, nstart = 0;
  if (!get_int("N_cycles", &ncycles)) fatal("Please set N_cycles in the input file", NULL);
  get_int("N_start", &nstart);
  if (nstart < 0) fatal("N_start must not be negative", NULL);

  int size[3] = {64, 64, 64}, grid[3] = {1, 1, 1}, per[3] = {1, 1, 1}, walls[3] = {0, 0, 0};
  get_ivec("size", size);
  get_ivec("grid", grid);              /* one GPU: the decomposition is ignored */
  get_ivec("periodicity", per);
  get_ivec("boundary_walls", walls);
  if (!per[0] || !per[1] || !per[2]) fatal("Only fully periodic systems are supported", NULL);
  if (walls[0] || walls[1] || walls[2]) fatal("Walls are not supported by this build", NULL);
  nx = size[0]; ny = size[1]; nz = size[2];

  double eta = 1.0/6.0, eta_bulk, rho0 = 1.0;
  get_double("viscosity", &eta);
  eta_bulk = eta;                      /* physics_rt.c:99-106 */
  get_double("viscosity_bulk", &eta_bulk);
  get_double("fluid_rho0", &rho0);

  char fe_name[64] = "none";
  get_string("free_energy", fe_name, sizeof(fe_name));
  int symmetric = (strcmp(fe_name, "symmetric") == 0);
  if (!symmetric && strcmp(fe_name, "none") != 0) fatal("free_energy not supported by this build: ", fe_name);

  char s[256];
  if (get_string("hydrodynamics", s, sizeof(s))) {
    if (strcmp(s, "off") == 0 || strcmp(s, "0") == 0 || strcmp(s, "no") == 0) {
      fatal("hydrodynamics off is not supported by this build", NULL);
    }
  }
  if (switch_on("lb_fluctuations")) fatal("lb_fluctuations are not supported by this build", NULL);
  if (get_string("isothermal_fluctuations", s, sizeof(s)) && strcmp(s, "on") == 0) {
    fatal("isothermal_fluctuations are not supported by this build", NULL);
  }
  if (get_string("colloid_init", s, sizeof(s)) && strcmp(s, "none") != 0
      && strcmp(s, "no_colloids") != 0) {
    fatal("colloids are not supported by this build", NULL);
  }
  get_double("temperature", &(double) {0.0});
  /* physics_rt.c:144-146: an external chemical potential gradient adds
   * -M grad_mu to every Cahn-Hilliard face flux (phi_cahn_hilliard.c:
   * 1373-1398) and, with fe_force_method phi_gradmu, -phi grad_mu to the
   * force (phi_grad_mu.c:123-160) */
  double grad_mu[3] = {0.0, 0.0, 0.0};
  get_dvec("grad_mu", grad_mu);

  pe_t * pe = NULL;
  cs_t * cs = NULL;
  physics_t * phys = NULL;
  lees_edw_t * le = NULL;
  lb_t * lb = NULL;
  hydro_t * hydro = NULL;
  map_t * map = NULL;
  wall_t * wall = NULL;
  field_t * phi = NULL;
  field_grad_t * dphi = NULL;
  fe_symm_t * fe = NULL;
  pth_t * pth = NULL;
  phi_ch_t * pch = NULL;

  pe_create(0, PE_VERBOSE, &pe);
  physics_create(pe, &phys);
  physics_eta_shear_set(phys, eta);
  physics_eta_bulk_set(phys, eta_bulk);
  physics_rho0_set(phys, rho0);
  physics_grad_mu_set(phys, grad_mu);

  cs_create(pe, &cs);
  cs_nhalo_set(cs, symmetric ? 2 : 1);     /* ludwig.c:1183,1198 */
  cs_ntotal_set(cs, size);
  cs_init(cs);
  /* leesedwards_rt.c:18-37, leesedwards.c:330-360 */
  lees_edw_options_t leo = {0};
  leo.type = LE_SHEAR_TYPE_STEADY;
  get_int("N_LE_plane", &leo.nplanes);
  get_double("LE_plane_vel", &leo.uy);
  if (get_int("LE_oscillation_period", &leo.period)) {
    fatal("oscillatory Lees-Edwards shear is not supported by this build", NULL);
  }
  get_int("LE_time_offset", &leo.nt0);
  lees_edw_create(pe, cs, &leo, &le);
  if (leo.nplanes > 0) {
    double gammadot = 0.0;
    lees_edw_shear_rate(le, &gammadot);
    printf("\nLees-Edwards boundary conditions are active:\n");
    for (int np = 0; np < leo.nplanes; np++) {
      printf("LE plane %d is at x = %d with speed %f\n", np + 1, lees_edw_plane_location(le, np),
	     leo.uy);
    }
    printf("Overall shear rate = %f\n", gammadot);
    printf("\nLees-Edwards time offset (time steps): %8d\n", leo.nt0);
  }
  physics_control_init_time(phys, nstart, ncycles);

  lb_data_options_t lbo = lb_data_options_default();
  lbo.iodata = io_args_keys("lb");
  if (get_string("lb_relaxation_scheme", s, sizeof(s))) {
    if (strcmp(s, "m10") == 0 || strcmp(s, "M10") == 0) lbo.nrelax = LB_RELAXATION_M10;
    else if (strcmp(s, "bgk") == 0 || strcmp(s, "BGK") == 0) lbo.nrelax = LB_RELAXATION_BGK;
    else if (strcmp(s, "trt") == 0 || strcmp(s, "TRT") == 0) lbo.nrelax = LB_RELAXATION_TRT;
    else fatal("Unrecognised relaxation time key ", s);
  }
  get_string("lb_halo_scheme", s, sizeof(s));
  lb_data_create(pe, cs, &lbo, &lb);
  lb_collision_relaxation_set(lb, lbo.nrelax);
  if (get_string("ghost_modes", s, sizeof(s)) && strcmp(s, "off") == 0) lb_collision_ghost_modes_off(lb);
  if (get_string("lb_ghost_modes", s, sizeof(s)) && strcmp(s, "off") == 0) lb_collision_ghost_modes_off(lb);

  hydro_options_t ho = hydro_options_default();
  ho.rho.iodata = io_args_keys("rho");
  ho.u.iodata = io_args_keys("vel");
  hydro_create(pe, cs, le, &ho, &hydro);
  map_options_t mo = map_options_default();
  map_create(pe, cs, &mo, &map);
  wall_create(pe, cs, map, lb, &wall);

  int seed = 13;                           /* field_phi_init_rt.c:23 */
  int phi_iofreq = 0;
  int conserve = 0;
  get_int("random_seed", &seed);
  printf("[User   ] Random number seed: %d\n", seed);

  if (symmetric) {
    field_options_t fo = field_options_ndata_nhalo(1, 2);
    fo.iodata = io_args_keys("phi");
    phi_iofreq = fo.iodata.iofreq;
    field_create(pe, cs, le, "phi", &fo, &phi);
    field_grad_create(pe, phi, 2, &dphi);
    grad_ft d2 = grad_3d_7pt_fluid_d2;
    if (get_string("fd_gradient_calculation", s, sizeof(s))) {
      if (strcmp(s, "3d_27pt_fluid") == 0) d2 = grad_3d_27pt_fluid_d2;
      else if (strcmp(s, "3d_7pt_fluid") != 0)
	fatal("fd_gradient_calculation not supported by this build (3d_7pt_fluid, 3d_27pt_fluid are): ", s);
    }
    field_grad_set(dphi, d2, NULL);
    fe_symm_create(pe, cs, phi, dphi, &fe);
    fe_symm_param_t fp = {-0.003125, 0.003125, 0.002, 0.0, 0.0};   /* symmetric.c:40-42 */
    get_double("A", &fp.a); get_double("symmetric_a", &fp.a);
    get_double("B", &fp.b); get_double("symmetric_b", &fp.b);
    get_double("K", &fp.kappa); get_double("symmetric_kappa", &fp.kappa);
    fe_symm_param_set(fe, fp);
    double mob = 0.0;
    if (!get_double("mobility", &mob)) fatal("Please set mobility in the input file", NULL);
    physics_mobility_set(phys, mob);
    get_int("cahn_hilliard_options_conserve", &conserve);
    phi_ch_info_t chi = {conserve, 0};
    phi_ch_create(pe, cs, le, &chi, &pch);
    int order = 1;
    get_int("fd_advection_scheme_order", &order);
    advection_order_set(order);
    int method = FE_FORCE_METHOD_STRESS_DIVERGENCE;
    if (get_string("fe_force_method", s, sizeof(s))) {
      if (strcmp(s, "stress_divergence") == 0) method = FE_FORCE_METHOD_STRESS_DIVERGENCE;
      else if (strcmp(s, "phi_gradmu") == 0) method = FE_FORCE_METHOD_PHI_GRADMU;
      else fatal("fe_force_method not supported by this build: ", s);
    }
    pth_create(pe, cs, method, &pth);

    double phi0 = 0.0, noise = 0.1;
    get_double("phi0", &phi0);
    physics_phi0_set(phys, phi0);
    /* field_phi_init_rt.c:40-160 */
    char init[64] = "spinodal";
    get_string("phi_initialisation", init, sizeof(init));
    if (strcmp(init, "spinodal") == 0) {
      printf("Initialising phi for spinodal\n");
      get_double("noise", &noise);
      field_phi_init_spinodal(phi, seed, phi0, noise);
    }
    else if (strcmp(init, "drop") == 0) {
      double radius = 8.0, phistar = 1.0, xi = 0.0;
      get_double("phi_init_drop_radius", &radius);
      get_double("phi_init_drop_amplitude", &phistar);
      int centred = switch_on("phi_init_drop_centred");
      printf("Initialising droplet radius:     %14.7e\n", radius);
      printf("Initialising droplet amplitude:  %14.7e\n", phistar);
      fe_symm_interfacial_width(fe, &xi);
      field_phi_init_drop(phi, xi, radius, phistar, centred);
    }
    else {
      fatal("phi_initialisation not supported by this build (spinodal, drop are): ", init);
    }
  }

  /* LB initial state (distribution_rt.c): rest at rho0, or uniform u */
  if (get_string("distribution_initialisation", s, sizeof(s)) && strcmp(s, "3d_uniform_u") == 0) {
    double u0[3] = {0.0, 0.0, 0.0};
    get_dvec("distribution_uniform_u", u0);
    printf("Velocity:             %14.7e %14.7e %14.7e\n", u0[0], u0[1], u0[2]);
    for (int ic = 1; ic <= nx; ic++)
      for (int jc = 1; jc <= ny; jc++)
	for (int kc = 1; kc <= nz; kc++) {
	  lb_1st_moment_equilib_set(lb, cs_index(cs, ic, jc, kc), rho0, u0);
	}
    lb_memcpy(lb, tdpMemcpyHostToDevice);
  }
  else {
    lb_init_rest_f(lb, rho0);
  }
  /* ludwig.c:305-312 */
  if (nstart == 0 && switch_on("LE_init_profile")) {
    if (leo.nplanes == 0) fatal("Cannot use LE_init_profile with no planes", NULL);
    printf("Initialising shear profile\n");
    lb_le_init_shear_profile(lb, le);
  }

  /* Restart (ludwig.c:313-364): distributions, phi, then rho and vel */
  if (nstart > 0) {
    io_event_t event = {0};
    printf("Re-starting simulation at step %d with data read from file\n", nstart);
    printf("Reading distribution files for step %d\n", nstart);
    lb_io_read(lb, nstart, &event);
    if (symmetric) {
      printf("Reading phi files for step %d\n", nstart);
      field_io_read(phi, nstart, &event);
    }
    printf("Reading rho/vel files for step %d\n", nstart);
    hydro_io_read(hydro, nstart, &event);
  }

  /* the reference sum of conserve 2 (ludwig.c:420-425) */
  if (symmetric && conserve == 2) cahn_hilliard_stats_time0(pch, phi, map);

  /* control.c:42-86, 138-147 */
  int freq_stats = 0, freq_config = 100000000, config_at_end = 1;
  get_int("freq_statistics", &freq_stats);
  get_int("freq_config", &freq_config);
  if (freq_config < 1) freq_config = nstart + ncycles + 1;
  if (get_string("config_at_end", s, sizeof(s)) && strcmp(s, "no") == 0) config_at_end = 0;
  get_string("colloid_io_freq", s, sizeof(s));
  const int nsteps_end = nstart + ncycles;

  /* step 0 statistics (ludwig.c:517-518) */
  lb_memcpy(lb, tdpMemcpyDeviceToHost);
  stats_rho(lb, cs);
  if (symmetric) {
    field_grad_compute(dphi);
    field_memcpy(phi, tdpMemcpyDeviceToHost);
    field_grad_memcpy(dphi, tdpMemcpyDeviceToHost);
    stats_phi(phi, cs);
    stats_fed(fe, cs, nstart);
  }
  stats_momentum(lb, cs);
  printf("\nStarting time step loop.\n");
  fflush(stdout);

  double tloop = 0.0;
  for (int step = nstart + 1; step <= nstart + ncycles; step++) {
    double t0 = wtime();
    if (symmetric) {
      ldw_step_binary(lb, hydro, phi, dphi, (fe_t *) fe, pth, pch, 1);
    }
    else {
      ldw_step_single(lb, hydro, 1);
    }
    ldw_device_synchronize();
    tloop += wtime() - t0;

    /* Configuration and field output (ludwig.c:868-962) */
    const int config = (step % freq_config == 0) || (config_at_end && step == nsteps_end);
    if (config) {
      io_event_t event = {0};
      printf("Writing distribution output at step %d!\n", step);
      lb_io_write(lb, step, &event);
    }
    if (symmetric && (config || util_mod(step, phi_iofreq) == 0)) {
      io_event_t event = {0};
      printf("Writing phi file at step %d!\n", step);
      field_io_write(phi, step, &event);
    }
    if (config) {
      io_event_t event = {0};
      printf("Writing rho/velocity output at step %d!\n", step);
      hydro_io_write(hydro, step, &event);
    }
    else {
      /* individual requests (ludwig.c:951-960) */
      if (util_mod(step, ho.rho.iodata.iofreq) == 0) {
	io_event_t event = {0};
	printf("Writing rho output at step %d!\n", step);
	hydro_io_write_rho(hydro, step, &event);
      }
      if (util_mod(step, ho.u.iodata.iofreq) == 0) {
	io_event_t event = {0};
	printf("Writing velocity output at step %d!\n", step);
	hydro_io_write_vel(hydro, step, &event);
      }
    }

    if (freq_stats > 0 && step % freq_stats == 0) {
      lb_memcpy(lb, tdpMemcpyDeviceToHost);
      stats_rho(lb, cs);
      if (symmetric) {
	field_memcpy(phi, tdpMemcpyDeviceToHost);
	field_grad_memcpy(dphi, tdpMemcpyDeviceToHost);
	stats_phi(phi, cs);
	stats_fed(fe, cs, step);
      }
      stats_momentum(lb, cs);
      hydro_memcpy(hydro, tdpMemcpyDeviceToHost);
      stats_velocity(hydro, cs);
      printf("\nCompleted cycle %d\n", step);
      fflush(stdout);
    }
  }

  for (int i = 0; i < nkeys; i++) {
    if (!keys[i].used) printf("Warning: key/value present in input but not used:\n (Line %d) %s %s\n",
			      keys[i].line, keys[i].key, keys[i].val);
  }
  printf("\nTimer statistics\n");
  printf("      Time step loop: %10.3f %10.3f %10.3f %10.6f (%d calls)\n",
	 ncycles ? tloop/ncycles : 0.0, ncycles ? tloop/ncycles : 0.0, tloop,
	 ncycles ? tloop/ncycles : 0.0, ncycles);
  printf("MLUPS: %.1f\n", ncycles ? (double) nx*ny*nz*ncycles/tloop/1.0e6 : 0.0);
  printf("Ludwig (libludwig_hip) finished normally.\n");

  if (symmetric) {
    phi_ch_free(pch); pth_free(pth); fe_symm_free(fe);
    field_grad_free(dphi); field_free(phi);
  }
  wall_free(wall); map_free(&map); hydro_free(hydro); lb_free(lb);
  lees_edw_free(le); cs_free(cs); physics_free(phys); pe_free(pe);
  return 0;
}
