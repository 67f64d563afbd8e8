/*****************************************************************************
 *
 *  ludwig_hip.h -- C ABI of libludwig_hip.so
 *
 *  The D3Q19 binary-fluid time step of Ludwig (lattice Boltzmann MRT
 *  collision + propagation, symmetric free energy mu and force, finite
 *  difference Cahn-Hilliard update) on one MI355X GPU, behind the
 *  reference's own C object API.  Each declaration cites the reference
 *  interface it replaces (file:line in the reference's src/ tree).
 *
 *  Conventions kept from the reference:
 *    - every function returns int, 0 on success;
 *    - errors are fatal by default (message with file:line, then abort),
 *      as pe_fatal()/tdpAssert() are (pe.c:240, target/target.h:136-138);
 *      ldw_error_mode_set(0) makes them return -1 instead, with the message
 *      in ldw_last_error();
 *    - host arrays keep the reference's SoA layout with halos
 *      (memory.h:184-230 "reverse" order, -DADDR_SOA), indexed by
 *      cs_index() (coords.c:617-631);
 *    - state lives on the device between *_memcpy() calls
 *      (ludwig.c:501-509).
 *
 *  Device side (not visible through this header): a periodic lattice
 *  without halos, SoA, site = (x*ny + y)*nz + z; *_halo() calls are
 *  therefore no-ops and propagation is folded into the next collision.
 *
 *****************************************************************************/

#ifndef LUDWIG_HIP_H
#define LUDWIG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Library control (no counterpart in the reference)                       */
/* ------------------------------------------------------------------------ */

/* 1: errors abort (default, as the reference); 0: functions return -1 */
int ldw_error_mode_set(int fatal);
const char * ldw_last_error(void);

/* Storage and arithmetic precision of objects created afterwards:
 * 64 = FP64 (default; bit-compatible with the reference),
 * 32 = FP32 with distributions stored as f_i - w_i rho0.
 * The environment variable LDW_PRECISION gives the initial value. */
int ldw_precision_set(int bits);
int ldw_precision(void);

/* Wait for all device work of the library */
int ldw_device_synchronize(void);

/* Device memory through the library's HIP runtime, for callers that keep
 * data on the GPU.  ldw_memcpy takes a tdpMemcpyKind and is complete on
 * return. */
int ldw_device_alloc(long long bytes, void ** ptr);
int ldw_device_free(void * ptr);
int ldw_memcpy(void * dst, const void * src, long long bytes, int kind);

/* ------------------------------------------------------------------------ */
/* Parallel environment and coordinates                                     */
/* ------------------------------------------------------------------------ */

typedef struct pe_s pe_t;
typedef enum {PE_QUIET = 0, PE_VERBOSE, PE_OPTION_MAX} pe_enum_t;

/* pe.h:26 pe_create(MPI_Comm, pe_enum_t, pe_t **); one rank, so the
 * communicator argument is an integer placeholder and is ignored */
int pe_create(int comm, pe_enum_t flag, pe_t ** ppe);
int pe_free(pe_t * pe);                                       /* pe.h:27 */

typedef struct cs_s cs_t;

int cs_create(pe_t * pe, cs_t ** pcs);                        /* coords.h:73 */
int cs_free(cs_t * cs);                                       /* coords.h:74 */
int cs_init(cs_t * cs);                                       /* coords.h:76 */
int cs_ntotal_set(cs_t * cs, const int ntotal[3]);            /* coords.h:82 */
int cs_nhalo_set(cs_t * cs, int nhalo);                       /* coords.h:83 */
int cs_nhalo(cs_t * cs, int * nhalo);                         /* coords.h:88 */
int cs_ntotal(cs_t * cs, int ntotal[3]);                      /* coords.h:90 */
int cs_nlocal(cs_t * cs, int n[3]);                           /* coords.h:95 */
int cs_nsites(cs_t * cs, int * nsites);                       /* coords.h:104 */
int cs_index(cs_t * cs, int ic, int jc, int kc);              /* coords.h:102 */
int cs_lmin(cs_t * cs, double lmin[3]);                      /* coords.h:96 */
int cs_ltot(cs_t * cs, double ltot[3]);                      /* coords.h:97 */

/* ------------------------------------------------------------------------ */
/* Physical constants (physics.h)                                           */
/* ------------------------------------------------------------------------ */

typedef struct physics_s physics_t;

int physics_create(pe_t * pe, physics_t ** phys);             /* physics.h:20 */
int physics_free(physics_t * phys);                           /* physics.h:21 */
int physics_ref(physics_t ** ref);                            /* physics.h:39 */
int physics_rho0_set(physics_t * phys, double rho0);          /* physics.h:23 */
int physics_phi0_set(physics_t * phys, double phi0);          /* physics.h:24 */
int physics_eta_shear_set(physics_t * phys, double eta);      /* physics.h:25 */
int physics_eta_bulk_set(physics_t * phys, double zeta);      /* physics.h:26 */
int physics_fbody_set(physics_t * phys, double f[3]);         /* physics.h:31 */
int physics_mobility_set(physics_t * phys, double mobility);  /* physics.h:32 */
int physics_rho0(physics_t * phys, double * rho0);
int physics_phi0(physics_t * phys, double * phi0);
int physics_eta_shear(physics_t * phys, double * eta);
int physics_eta_bulk(physics_t * phys, double * eta);
int physics_mobility(physics_t * phys, double * mobility);
int physics_fbody(physics_t * phys, double f[3]);
/* External chemical potential gradient ("grad_mu"): every Cahn-Hilliard
 * face flux loses M grad_mu (phi_cahn_hilliard.c:1373-1398); with the
 * phi_gradmu force the fluid also feels -phi grad_mu (phi_grad_mu.c:123-160) */
int physics_grad_mu_set(physics_t * phys, double gm[3]);      /* physics.h:37 */
int physics_grad_mu(physics_t * phys, double gm[3]);          /* physics.h:56 */
/* the time step counter (ludwig.c:528-531); the Lees-Edwards planes move
 * with physics_control_time() = t_start + t_current - 1 */
int physics_control_init_time(physics_t * phys, int nstart, int nsteps); /* physics.h:34 */
int physics_control_next_step(physics_t * phys);              /* physics.h:33 */
int physics_control_timestep(physics_t * phys);               /* physics.h:51 */
int physics_control_time(physics_t * phys, double * t);       /* physics.h:52 */

/* ------------------------------------------------------------------------ */
/* Host/device copies (target/target.h tdpMemcpyKind)                       */
/* ------------------------------------------------------------------------ */

typedef enum {
  tdpMemcpyHostToHost = 0,
  tdpMemcpyHostToDevice = 1,
  tdpMemcpyDeviceToHost = 2,
  tdpMemcpyDeviceToDevice = 3
} tdpMemcpyKind;

/* ------------------------------------------------------------------------ */
/* I/O options (io_options.h, io_info_args.h, io_event.h)                   */
/* ------------------------------------------------------------------------ */

typedef enum io_mode_enum {IO_MODE_INVALID, IO_MODE_MPIIO} io_mode_enum_t;   /* io_options.h:30 */
typedef enum io_record_format_enum {IO_RECORD_INVALID,
				    IO_RECORD_ASCII,
				    IO_RECORD_BINARY} io_record_format_enum_t; /* io_options.h:34 */
typedef enum io_metadata_version_enum {IO_METADATA_INVALID,
				       IO_METADATA_SINGLE_V1,
				       IO_METADATA_MULTI_V1,
				       IO_METADATA_V2} io_metadata_version_enum_t; /* io_options.h:40 */

typedef struct io_options_s {                                 /* io_options.h:51 */
  io_mode_enum_t mode;
  io_record_format_enum_t iorformat;
  io_metadata_version_enum_t metadata_version;
  int report;
  int asynchronous;
  int compression_levl;
  int iogrid[3];
} io_options_t;

typedef struct io_info_args_s {                               /* io_info_args.h:30 */
  io_options_t input;
  io_options_t output;
  int grid[3];
  int iofreq;
} io_info_args_t;

/* mode mpiio, binary records, metadata version 2, I/O grid 1 1 1
 * (io_options.c:26-38, io_info_args.c:28-35) */
io_options_t io_options_default(void);
io_info_args_t io_info_args_default(void);

typedef struct io_event_s {                                   /* io_event.h:35 */
  const char * extra_name;
  void * extra_json;                  /* not used: no extra metadata here */
  double time[6];
} io_event_t;

/* ------------------------------------------------------------------------ */
/* Lattice Boltzmann distributions (lb_data.h)                              */
/* ------------------------------------------------------------------------ */

typedef enum lb_relaxation_enum {LB_RELAXATION_M10,
				 LB_RELAXATION_BGK,
				 LB_RELAXATION_TRT} lb_relaxation_enum_t;
typedef enum lb_halo_enum {LB_HALO_FULL = 1, LB_HALO_REDUCED = 2} lb_halo_enum_t;
typedef enum lb_dist_enum_type {LB_RHO = 0, LB_PHI = 1} lb_dist_enum_t;

/* lb_data_options.h:32-42 */
typedef struct lb_data_options_s {
  int ndim;
  int nvel;
  int ndist;
  lb_relaxation_enum_t nrelax;
  lb_halo_enum_t halo;
  int reportimbalance;
  int usefirsttouch;
  io_info_args_t iodata;
} lb_data_options_t;

typedef struct lb_data_s lb_t;
typedef struct hydro_s hydro_t;
typedef struct map_s map_t;
typedef struct noise_s noise_t;
typedef struct fe_s fe_t;
typedef struct visc_s visc_t;

lb_data_options_t lb_data_options_default(void);              /* lb_data_options.h:44 */
int lb_data_create(pe_t * pe, cs_t * cs, const lb_data_options_t * opts,
		   lb_t ** lb);                               /* lb_data.h:162 */
int lb_free(lb_t * lb);                                       /* lb_data.h:164 */
int lb_memcpy(lb_t * lb, tdpMemcpyKind flag);                 /* lb_data.c:478 */
int lb_halo(lb_t * lb);                                       /* lb_data.c:754 */
int lb_init_rest_f(lb_t * lb, double rho0);                   /* lb_data.c:659 */
int lb_1st_moment_equilib_set(lb_t * lb, int index, double rho,
			      double u[3]);                   /* lb_data.c:809 */
int lb_ndist(lb_t * lb, int * ndist);
int lb_nvel(lb_t * lb, int * nvel);

/* Host accessors, static inline in the reference (lb_data.h:198-290) */
int lb_f(lb_t * lb, int index, int p, int n, double * f);
int lb_f_set(lb_t * lb, int index, int p, int n, double f);
int lb_0th_moment(lb_t * lb, int index, lb_dist_enum_t nd, double * rho);
int lb_1st_moment(lb_t * lb, int index, lb_dist_enum_t nd, double g[3]);

/* collision.h:27-35 */
int lb_collision_relaxation_set(lb_t * lb, lb_relaxation_enum_t nrelax);
int lb_collision_ghost_modes_on(lb_t * lb);
int lb_collision_ghost_modes_off(lb_t * lb);
int lb_collide(lb_t * lb, hydro_t * hydro, map_t * map, noise_t * noise,
	       fe_t * fe, visc_t * visc);                     /* collision.c:143 */

int lb_propagation(lb_t * lb);                                /* propagation.c:43 */

/* Distribution files "dist-ttttttttt.001-001" (io_subfile.c:186-204) in
 * the working directory, with "dist-metadata.001-001" (io_metadata.c:
 * 246-281) written before the first data file.  Records are the 19
 * distributions of each site as doubles in host byte order, sites in
 * x-slowest, z-fastest order (binary), or " %22.15e" per value and one line
 * per velocity (ascii), as lb_write_buf / lb_write_buf_ascii
 * (lb_data.c:1533-1640).  The record format comes from the options'
 * iodata at lb_data_create(); the I/O grid must be 1 1 1 (one GPU).  A
 * missing or short file is an error. */
int lb_io_write(lb_t * lb, int timestep, io_event_t * event); /* lb_data.c:1716 */
int lb_io_read(lb_t * lb, int timestep, io_event_t * event);  /* lb_data.c:1779 */

/* ------------------------------------------------------------------------ */
/* Lees-Edwards planes (steady shear, one rank); map and wall: the          */
/* all-fluid, wall-free case only                                           */
/* ------------------------------------------------------------------------ */

typedef struct lees_edw_s lees_edw_t;
typedef enum lees_edw_enum {                                  /* lees_edwards_options.h:26 */
  LE_SHEAR_TYPE_INVALID,
  LE_SHEAR_TYPE_STEADY,
  LE_SHEAR_TYPE_OSCILLATORY
} lees_edw_enum_t;
typedef struct lees_edw_options_s {                           /* lees_edwards_options.h:32 */
  int nplanes;
  int type;                   /* lees_edw_enum_t; planes need LE_SHEAR_TYPE_STEADY */
  int period;
  int nt0;
  double uy;
} lees_edw_options_t;

/* nplanes equally spaced planes, plane p between x = loc and loc + 1 with
 * loc = (int) (L/2n + p L/n); the FP64 storage only; fails for planes
 * within nhalo of the x boundaries, as the reference does */
int lees_edw_create(pe_t * pe, cs_t * cs, const lees_edw_options_t * opts,
		    lees_edw_t ** le);                        /* leesedwards.h:26 */
int lees_edw_free(lees_edw_t * le);                           /* leesedwards.h:29 */
int lees_edw_nplane_total(lees_edw_t * le);                   /* leesedwards.h:57 */
int lees_edw_shear_rate(lees_edw_t * le, double * gammadot);  /* leesedwards.h:63 */
int lees_edw_plane_location(lees_edw_t * le, int np);         /* leesedwards.h:65 */

/* after lb_collide(), before lb_propagation(): reprojection, displacement
 * and interpolation of the plane-crossing distributions; no-op without
 * planes */
int lb_data_apply_le_boundary_conditions(lb_t * lb, lees_edw_t * le); /* model_le.h:19 */
/* steady linear shear profile f for rho0 and the plane speed */
int lb_le_init_shear_profile(lb_t * lb, lees_edw_t * le);     /* model_le.h:20 */

typedef struct map_options_s {int ndata; int is_porous_media;} map_options_t;
map_options_t map_options_default(void);                      /* map_options.h:29 */
int map_create(pe_t * pe, cs_t * cs, const map_options_t * options,
	       map_t ** map);                                 /* map.h:51 */
int map_free(map_t ** map);                                   /* map.h:53 */

typedef struct wall_s wall_t;
int wall_create(pe_t * pe, cs_t * cs, map_t * map, lb_t * lb,
		wall_t ** p);                                 /* wall.h:88 */
int wall_free(wall_t * wall);

/* ------------------------------------------------------------------------ */
/* Fields (field.h, field_grad.h)                                           */
/* ------------------------------------------------------------------------ */

typedef struct field_options_s {                              /* field_options.h:28 */
  int ndata;
  int nhcomm;
  int haloscheme;
  int haloverbose;
  int usefirsttouch;
  io_info_args_t iodata;
} field_options_t;

typedef struct field_s field_t;
typedef struct field_grad_s field_grad_t;
typedef int (* grad_ft)(field_grad_t * fgrad);

field_options_t field_options_ndata_nhalo(int ndata, int nhalo); /* field_options.h:40 */
int field_create(pe_t * pe, cs_t * cs, lees_edw_t * le, const char * name,
		 const field_options_t * opts, field_t ** pobj); /* field.h:97 */
int field_free(field_t * obj);                                /* field.h:101 */
int field_memcpy(field_t * obj, tdpMemcpyKind flag);          /* field.h:102 */
int field_halo(field_t * obj);                                /* field.h:104 */
int field_nf(field_t * obj, int * nf);
int field_scalar(field_t * obj, int index, double * phi);     /* field.h:109 */
int field_scalar_set(field_t * obj, int index, double phi);   /* field.h:110 */

/* "<name>-ttttttttt.001-001" with "<name>-metadata.001-001": nf doubles
 * per site (binary) or nf " %22.15e" values and a newline (ascii), as
 * field_write_buf / field_write_buf_ascii (field.c:896-996); otherwise as
 * lb_io_write. */
int field_io_write(field_t * field, int timestep, io_event_t * event); /* field.c:1633 */
int field_io_read(field_t * field, int timestep, io_event_t * event);  /* field.c:1690 */

int field_grad_create(pe_t * pe, field_t * f, int level,
		      field_grad_t ** pobj);                  /* field_grad.h:44 */
void field_grad_free(field_grad_t * obj);                     /* field_grad.h:46 */
int field_grad_set(field_grad_t * obj, grad_ft d2, grad_ft d4); /* field_grad.h:47 */
int field_grad_compute(field_grad_t * obj);                   /* field_grad.c:319 */
int field_grad_memcpy(field_grad_t * obj, tdpMemcpyKind flag);
int field_grad_scalar_grad(field_grad_t * obj, int index, double grad[3]);
int field_grad_scalar_delsq(field_grad_t * obj, int index, double * delsq);

/* The 7-point fluid stencil (gradient_3d_7pt_fluid.c:73) */
int grad_3d_7pt_fluid_d2(field_grad_t * fgrad);
int grad_3d_27pt_fluid_d2(field_grad_t * fgrad);   /* gradient_3d_27pt_fluid.h */

/* ------------------------------------------------------------------------ */
/* Hydrodynamic fields (hydro.h)                                            */
/* ------------------------------------------------------------------------ */

typedef struct hydro_options_s {                              /* hydro_options.h:22 */
  int nhcomm;
  field_options_t rho;
  field_options_t u;
  field_options_t force;
  field_options_t eta;
} hydro_options_t;

hydro_options_t hydro_options_default(void);                  /* hydro_options.h:32 */
int hydro_create(pe_t * pe, cs_t * cs, lees_edw_t * le,
		 const hydro_options_t * opts, hydro_t ** pobj); /* hydro.h:49 */
int hydro_free(hydro_t * obj);                                /* hydro.h:51 */
int hydro_memcpy(hydro_t * obj, tdpMemcpyKind flag);          /* hydro.h:52 */
int hydro_u_halo(hydro_t * obj);                              /* hydro.h:53 */
int hydro_f_zero(hydro_t * obj, const double fzero[3]);       /* hydro.h:59 */
int hydro_u_zero(hydro_t * obj, const double uzero[3]);       /* hydro.h:60 */
/* accessors, static inline in the reference (hydro_impl.h:46-140) */
int hydro_u(hydro_t * obj, int index, double u[3]);
int hydro_u_set(hydro_t * obj, int index, const double u[3]);
int hydro_rho(hydro_t * obj, int index, double * rho);
int hydro_f_local(hydro_t * obj, int index, double force[3]);
int hydro_f_local_set(hydro_t * obj, int index, const double force[3]);

/* Configuration output and input of the fields "rho" (1 double per site)
 * and "vel" (3), as field_io_write/read with the options' rho.iodata and
 * u.iodata (hydro.c:941-967). */
int hydro_io_write(hydro_t * obj, int timestep, io_event_t * event); /* hydro.c:941 */
int hydro_io_read(hydro_t * obj, int timestep, io_event_t * event);  /* hydro.c:959 */
/* rho alone and velocity alone, the individual requests between
 * configuration steps: the reference calls field_io_write() on hydro->rho
 * and hydro->u (ludwig.c:951-960); hydro_t is opaque here */
int hydro_io_write_rho(hydro_t * obj, int timestep, io_event_t * event);
int hydro_io_write_vel(hydro_t * obj, int timestep, io_event_t * event);

/* ------------------------------------------------------------------------ */
/* Symmetric free energy (symmetric.h)                                      */
/* ------------------------------------------------------------------------ */

typedef struct fe_symm_s fe_symm_t;
typedef struct fe_symm_param_s {                              /* symmetric.h:40 */
  double a;
  double b;
  double kappa;
  double c;
  double h;
} fe_symm_param_t;

int fe_symm_create(pe_t * pe, cs_t * cs, field_t * f, field_grad_t * grd,
		   fe_symm_t ** p);                           /* symmetric.h:48 */
int fe_symm_free(fe_symm_t * fe);                             /* symmetric.h:50 */
int fe_symm_param_set(fe_symm_t * fe, fe_symm_param_t values); /* symmetric.h:51 */
int fe_symm_param(fe_symm_t * fe, fe_symm_param_t * values);
int fe_symm_interfacial_width(fe_symm_t * fe, double * xi);   /* symmetric.c:267 */
/* host evaluation on the host copies of phi and its gradients */
int fe_symm_mu(fe_symm_t * fe, int index, double * mu);       /* symmetric.c:308 */
int fe_symm_fed(fe_symm_t * fe, int index, double * fed);     /* symmetric.c:285 */
int fe_symm_str(fe_symm_t * fe, int index, double s[3][3]);   /* symmetric.c:334 */

/* ------------------------------------------------------------------------ */
/* Thermodynamic force (fe_force_method.h, phi_force_stress.h, phi_force.h) */
/* ------------------------------------------------------------------------ */

typedef enum {
  FE_FORCE_METHOD_INVALID,
  FE_FORCE_METHOD_NO_FORCE,
  FE_FORCE_METHOD_STRESS_DIVERGENCE,
  FE_FORCE_METHOD_PHI_GRADMU,
  FE_FORCE_METHOD_PHI_GRADMU_CORRECTION,
  FE_FORCE_METHOD_RELAXATION_SYMM,
  FE_FORCE_METHOD_RELAXATION_ANTI,
  FE_FORCE_METHOD_MAX
} fe_force_method_enum_t;                                     /* fe_force_method.h:20 */

typedef struct pth_s pth_t;
int pth_create(pe_t * pe, cs_t * cs, int method, pth_t ** pth); /* phi_force_stress.h:36 */
int pth_free(pth_t * pth);                                    /* phi_force_stress.h:37 */
int phi_force_calculation(pe_t * pe, cs_t * cs, lees_edw_t * le,
			  wall_t * wall, pth_t * pth, fe_t * fe,
			  map_t * map, field_t * phi,
			  hydro_t * hydro);                   /* phi_force.c:74 */

/* ------------------------------------------------------------------------ */
/* Cahn-Hilliard (phi_cahn_hilliard.h, advection.h)                         */
/* ------------------------------------------------------------------------ */

typedef struct phi_ch_s phi_ch_t;
typedef struct phi_ch_info_s {                                /* phi_cahn_hilliard.h:36 */
  int conserve;
  int noise;
} phi_ch_info_t;

int phi_ch_create(pe_t * pe, cs_t * cs, lees_edw_t * le,
		  phi_ch_info_t * info, phi_ch_t ** pch);     /* phi_cahn_hilliard.c:131 */
int phi_ch_free(phi_ch_t * pch);                              /* phi_cahn_hilliard.c:171 */
int phi_cahn_hilliard(phi_ch_t * pch, fe_t * fe, field_t * phi,
		      hydro_t * hydro, map_t * map,
		      noise_t * noise);                       /* phi_cahn_hilliard.c:213 */
/* cahn_hilliard_options_conserve (phi_ch_info_t.conserve):
 *   0  the forward step;
 *   1  the update as per-site compensated sums, the compensation kept
 *      between steps (phi_cahn_hilliard.c:1059-1097, 1181-1218);
 *   2  the forward step, then phi -= (S - field_init_sum)/N with S the sum
 *      of phi (phi_cahn_hilliard.c:1102-1175, 1225-1327).  Needs
 *      cahn_hilliard_stats_time0() first.  The reference's S is a plain
 *      sum whose order it does not fix (atomics on the GPU, per-thread
 *      partials under OpenMP); S here is, by default, a compensated tree
 *      reduction (deterministic, within a few ulp of the exact sum).
 *      ldw_conserve2_site_order(1) (or LDW_CONSERVE2_SITE_ORDER=1 in the
 *      environment) takes the plain sum in site order on one thread
 *      instead, the reference's one-thread order, bit-identical with it;
 *      that sum is one serial chain of additions (~10 ms at 128^3). */
int ldw_conserve2_site_order(int on);
/* field_init_sum of phi, as the reference computes it on one thread
 * (cahn_hilliard_stats.c:59-76) */
int cahn_hilliard_stats_time0(phi_ch_t * pch, field_t * phi, map_t * map);
int advection_order_set(const int order);                     /* advection.h:40 */
int advection_order(int * order);                             /* advection.h:41 */

/* ------------------------------------------------------------------------ */
/* Initial conditions                                                       */
/* ------------------------------------------------------------------------ */

/* phi = phi0 + amp (r - 1/2), r from the per-site lattice generator of the
 * reference (field_phi_init.c:484-518, noise.c:278-341, noise.h:96-109);
 * writes the host copy and the device copy */
int field_phi_init_spinodal(field_t * phi, int seed, double phi0, double amp);

/* phi = phistar tanh((r - radius)/xi), r the distance of the site from the
 * box centre (0.5 L, or Lmin + 0.5 L when is_centred), as
 * field_phi_init.c:38-82; writes the host copy and the device copy */
int field_phi_init_drop(field_t * phi, double xi, double radius, double phistar,
			int is_centred);

/* ------------------------------------------------------------------------ */
/* Additions (no counterpart in the reference)                              */
/* ------------------------------------------------------------------------ */

/* Plain-fluid start: u_a = amp (2 r - 1) from three draws of the same
 * per-site generator, per-component mean removed, f = equilibrium(1, u) */
int ldw_lb_init_random_velocity(lb_t * lb, int seed, double amp);

/* Bulk copies of interior values in the device layout
 * (site = (x*ny + y)*nz + z, component major), as FP64 on the host side.
 * ldw_lb_f_read returns the distributions the reference's lb->f would
 * hold at this point of the step. */
int ldw_lb_f_read(lb_t * lb, double * f);
int ldw_lb_f_write(lb_t * lb, const double * f);
int ldw_field_read(field_t * obj, double * data);
int ldw_field_write(field_t * obj, const double * data);
int ldw_hydro_read(hydro_t * obj, double * rho, double * u, double * force);
int ldw_hydro_write(hydro_t * obj, const double * rho, const double * u,
		    const double * force);

/* nsteps of the binary-fluid time step in the reference's loop order
 * (ludwig.c:528-860): hydro_f_zero, field_halo, field_grad_compute,
 * phi_force_calculation, phi_cahn_hilliard, hydro_u_zero, lb_collide,
 * lb_halo, lb_propagation; the same result as calling those in turn. */
int ldw_step_binary(lb_t * lb, hydro_t * hydro, field_t * phi,
		    field_grad_t * dphi, fe_t * fe, pth_t * pth,
		    phi_ch_t * pch, int nsteps);

/* nsteps of the single-fluid step (free_energy none) */
int ldw_step_single(lb_t * lb, hydro_t * hydro, int nsteps);

/* Compensated FP64 sums on the device: out[0] = sum_x sum_p f (total
 * density), out[1] = sum_x phi (phi may be NULL -> 0) */
int ldw_conserved_sums(lb_t * lb, field_t * phi, double out[2]);

/* Per-kernel device timing with HIP events on the library's stream.
 * ldw_timer_enable(1) starts recording; ldw_timer_read(name, &ms, &n)
 * returns the accumulated time and launch count of a kernel family:
 * "collide", "phi_force_mu", "mu", "cahn_hilliard", "propagate". */
int ldw_timer_enable(int on);
int ldw_timer_reset(void);
int ldw_timer_read(const char * name, double * ms, int * count);

/* Region timing on the library's stream: ldw_stream_mark(slot) records an
 * event (slots 0..7); ldw_stream_elapsed(a, b, &ms) waits for slot b and
 * returns the time from slot a to slot b. */
int ldw_stream_mark(int slot);
int ldw_stream_elapsed(int a, int b, double * ms);

/* Bandwidth calibration: mean time per launch of a plain device copy of
 * nbytes (read nbytes + write nbytes) over reps launches */
int ldw_copy_calibration(long long nbytes, int reps, double * ms);

/* Raw device arrays (layout as described at the top of this file) for
 * callers that keep data on the GPU; bytes per element in *elsize. */
int ldw_lb_device_f(lb_t * lb, void ** ptr, int * elsize, int * shifted);
int ldw_field_device(field_t * obj, void ** ptr, int * elsize);
int ldw_hydro_device(hydro_t * obj, void ** rho, void ** u, void ** force,
		     int * elsize);

/* ------------------------------------------------------------------------ */
/* Replica ensemble (additive; no counterpart in the reference)             */
/* ------------------------------------------------------------------------ */

/* Many independent periodic binary-fluid lattices of different sizes and
 * parameters advanced together, one persistent launch per time step
 * (captured in a HIP graph).  Each replica follows exactly the arithmetic
 * of a single lattice driven by ldw_step_binary() with the same size,
 * parameters and seed:
 *   phi = phi0 + noise (U - 1/2), the field_phi_init_spinodal() generator
 *   (field_phi_init.c:484-518) with the replica's seed;
 *   f = rest equilibrium at rho0 (lb_init_rest_f), u = 0;
 *   M10 collision with eta_shear / eta_bulk; symmetric free energy with
 *   a, b, kappa; Cahn-Hilliard with the mobility; 7-point gradient. */
typedef struct {
  int nx, ny, nz;
  int seed;
  double phi0, noise;
  double a, b, kappa, mobility;
  double eta_shear, eta_bulk;
  double rho0;
} ldw_replica_t;

typedef struct ldw_ensemble_s ldw_ensemble_t;

/* precision 64 or 32 (FP32 stores f - w rho0); adv_order 1..3;
 * force_method FE_FORCE_METHOD_STRESS_DIVERGENCE or _PHI_GRADMU */
int ldw_ensemble_create(int nreplica, const ldw_replica_t * spec, int precision,
			int adv_order, int force_method, ldw_ensemble_t ** pens);
int ldw_ensemble_free(ldw_ensemble_t * ens);

/* Autotuner: times every tile shape (64x4, 32x8, 16x16, 64x8, 128x4 in z x y) and
 * several x planes per work item on the current state, `reps` launches
 * each, and keeps the fastest.  The state is not advanced.  Runs by itself
 * before the first step unless a configuration was set. */
int ldw_ensemble_tune(ldw_ensemble_t * ens, int reps);
/* The autotuner's trials: up to maxrows rows of {tile z, tile y, x planes,
 * blocks, ms per step} into rows (may be NULL); returns the number of trials */
int ldw_ensemble_tune_results(ldw_ensemble_t * ens, int maxrows, double * rows);
/* cfg = {tile z, tile y, x planes per item, blocks}; ms = the tuned step
 * time (0 if not tuned) */
int ldw_ensemble_config(ldw_ensemble_t * ens, int cfg[4], double * ms);
/* Fix the configuration (tz x ty one of the shapes above; xplanes >= 1;
 * blocks 0 -> as many as are resident at once) */
int ldw_ensemble_config_set(ldw_ensemble_t * ens, int tz, int ty, int xplanes,
			    int blocks);
/* Use a captured HIP graph per step (default 1) or direct launches (0) */
int ldw_ensemble_graph_set(ldw_ensemble_t * ens, int on);

int ldw_ensemble_step(ldw_ensemble_t * ens, int nsteps);
/* number of steps taken */
long long ldw_ensemble_time(ldw_ensemble_t * ens);

/* Replica r's current state as FP64 in the device layout
 * (site = (x*ny + y)*nz + z): phi (nsite), rho (nsite), u (3 nsite), each
 * may be NULL.  f: 19 nsite values, the distributions the reference's
 * lb->f would hold after lb_propagation(). */
int ldw_ensemble_snapshot(ldw_ensemble_t * ens, int r, double * phi, double * rho,
			  double * u);
int ldw_ensemble_f_read(ldw_ensemble_t * ens, int r, double * f);
/* phi, ux, uy, uz of replica r as 4 nsite floats into device memory dst,
 * ordered on the library's stream (the library's later calls see it; the
 * host or another stream waits with ldw_device_synchronize) */
int ldw_ensemble_snapshot_device(ldw_ensemble_t * ens, int r, float * dst);
/* Compensated FP64 sums of replica r: out[0] = total density, out[1] = sum phi */
int ldw_ensemble_conserved_sums(ldw_ensemble_t * ens, int r, double out[2]);
/* Replace replica r's phi and u by src (device memory, 4 nsite floats:
 * phi, ux, uy, uz) and its distributions by the equilibrium at the current
 * density and the new velocity (a surrogate jump) */
int ldw_ensemble_load_device(ldw_ensemble_t * ens, int r, const float * src);

/* ------------------------------------------------------------------------ */
/* Surrogate network kernels (additive; no counterpart in the reference)    */
/* ------------------------------------------------------------------------ */

/* 3 x 3 x 3 convolutions with periodic padding on the MFMA units: BF16
 * operands (raw 16-bit arrays), FP32 accumulation and results.  Device
 * pointers; activations channels-last, x[p][c], p = ((b nx + x) ny + y) nz + z;
 * weights w[27][cout][cin] (tap t = 9 (dx+1) + 3 (dy+1) + (dz+1)); channel
 * counts multiples of 16.  accumulate = 1 adds to the output.
 *   y  = bias + sum_t w_t x(p + o_t)                      (ldw_conv3d_fwd)
 *   dx = sum_t w_t^T dy(p - o_t), wd from weights_prep      (ldw_conv3d_bwd_data)
 *   dw_t = sum_p dy(p) x(p + o_t)^T, cin, cout <= 160        (ldw_conv3d_bwd_weight) */
int ldw_conv3d_fwd(const void * x, const void * w, const float * bias, float * y,
		   int batch, int nx, int ny, int nz, int cin, int cout, int accumulate);
/* the forward pass with SiLU fused: z (may be NULL) = bias + conv,
 * a = bf16(SiLU(z)) */
int ldw_conv3d_fwd_silu(const void * x, const void * w, const float * bias, float * z, void * a,
			int batch, int nx, int ny, int nz, int cin, int cout);
int ldw_conv3d_bwd_data(const void * dy, const void * wd, float * dx, int batch, int nx,
			int ny, int nz, int cin, int cout, int accumulate);
int ldw_conv3d_bwd_weight(const void * x, const void * dy, float * dw, int batch, int nx,
			  int ny, int nz, int cin, int cout, int accumulate);
/* FP32 master weights w -> BF16 wf (same layout) and wd (flipped and
 * transposed for ldw_conv3d_bwd_data); either output may be NULL */
int ldw_conv3d_weights_prep(const float * w, void * wf, void * wd, int cin, int cout);

/* Elementwise and reshaping kernels of the network (device pointers) */
int ldw_nn_silu_fwd(const float * z, void * a, long long n);
int ldw_nn_silu_bwd(const float * z, const float * da, float * dz, void * dzb, long long n);
int ldw_nn_bias_grad(const float * dz, float * db, long long npos, int c, int accumulate);
int ldw_nn_pool2_fwd(const void * in, void * out, int batch, int nx, int ny, int nz, int c);
int ldw_nn_pool2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c,
		     int accumulate);
int ldw_nn_up2_fwd(const void * in, void * out, int batch, int nx, int ny, int nz, int c);
int ldw_nn_up2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c);
int ldw_nn_concat(const void * a, int ca, const void * b, int cb, void * out, long long npos);
/* out (fine lattice fx x fy x fz) = concat(nearest 2x upsampling of coarse
 * (cu channels), skip (cs channels)); cu, cs multiples of 8 */
int ldw_nn_up2_concat(const void * coarse, int cu, const void * skip, int cs, void * out,
		      int batch, int fx, int fy, int fz);
int ldw_nn_split(const float * d, int ca, int cb, float * da, float * db, long long npos,
		 int acc_a, int acc_b);
int ldw_nn_to_bf16(const float * a, void * b, long long n);
/* input channels of one sample from 4 SoA planes (phi, ux, uy, uz):
 * norm = {phi shift, phi scale, u scale, p0, p1, p2} (p: scaled parameters) */
int ldw_nn_encode(const float * state, long long nsite, const float norm[6], void * out, int c);
int ldw_nn_target(const float * now, const float * next, long long nsite, float * tgt,
		  float dphi_scale, float du_scale);
/* now + y / scale back to 4 SoA planes, clamped to |phi| <= phi_max and
 * |u_a| <= u_max */
int ldw_nn_decode(const float * now, const float * y, int cy, long long nsite, float * out,
		  float dphi_scale, float du_scale, float phi_max, float u_max);
/* mean of w_c (y_c - tgt_c)^2 over y's first 4 channels (weights may be
 * NULL: all 1); gradient into dy (BF16) and dyf (FP32, may be NULL);
 * *loss (may be NULL) = the mean, compensated sum */
int ldw_nn_mse(const float * y, int cy, const float * tgt, long long npos, void * dy,
	       float * dyf, const float weights[4], double * loss);
/* Adam step `step` (1-based) on FP32 weights; the gradient is multiplied
 * by gscale first (global-norm clipping) */
int ldw_nn_adam(float * w, const float * g, float * m, float * v, long long n, float lr,
		float b1, float b2, float eps, float wd, float gscale, int step);
/* *out = sum a[i]^2 in FP64 (fixed order), complete on return */
int ldw_nn_sumsq(const float * a, long long n, double * out);
/* a[i] *= f */
int ldw_nn_scale(float * a, long long n, float f);
/* *out = sum a[i] in FP64 (fixed order), complete on return */
int ldw_nn_sum(const float * a, long long n, double * out);
/* a[i] += v */
int ldw_nn_shift(float * a, long long n, float v);
/* out = an element of the cubic group applied to the state in (4 planes
 * of n^3 floats: phi, ux, uy, uz, sites z fastest): output axis a is input
 * axis perm[a], reflected (t -> n-1-t) where flip[a] != 0, and the velocity
 * turns with it.  The lattice dynamics commute with these maps, so they
 * make new training pairs from old ones. */
int ldw_nn_cubic(const float * in, float * out, int n, const int perm[3], const int flip[3]);
/* For each of nplane planes of ns floats: out[p] += (sum now[p] - sum
 * out[p])/ns, sums in FP64 in a fixed order, all on the stream (no host
 * read-back) */
int ldw_nn_restore_sums(const float * now, float * out, long long ns, int nplane);
/* ldw_nn_restore_sums without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid hipMalloc).  The partial
 * sums live in a fixed device array of 64 KiB per GPU (nplane <= 64; more
 * is refused with an error) that every call on that GPU shares: calls that
 * overlap on one GPU race.  Results equal ldw_nn_restore_sums's bit for
 * bit. */
int ldw_nn_restore_sums_noalloc(const float * now, float * out, long long ns, int nplane);

#ifdef __cplusplus
}
#endif

#endif
