// ludwig_api.cpp -- the C ABI of libludwig_hip.so (include/ludwig_hip.h).
//
// Host objects mirror the reference's lb_t / field_t / hydro_t / fe_symm_t
// / pth_t / phi_ch_t (lb_data.h:107-138, field.h:68-88, hydro.h:32-47,
// symmetric.h:28-42, phi_force_stress.h, phi_cahn_hilliard.h:36-46): a host
// copy in the reference's SoA layout with halos, and a device copy in the
// device layout (periodic, no halo, SoA).  Device work is ordered on one
// HIP stream of default priority.
//
// The time-step calls keep the reference's semantics while letting the
// kernels fuse:
//   - lb_propagation() marks the distributions as "streamed"; the next
//     lb_collide() reads them through the pull indexing (K1 fuses both);
//   - hydro_f_zero()/hydro_u_zero() record the uniform value; the next
//     kernel that needs the array takes it as a constant;
//   - field_grad_compute() is deferred into phi_force_calculation() (K2),
//     which writes the force and mu; phi_cahn_hilliard() (K3) consumes mu;
//   - field_halo(), lb_halo(), hydro_u_halo() are no-ops: the kernels wrap
//     indices periodically.
// Any *_memcpy(DeviceToHost) materialises the deferred state first, so the
// host sees exactly what the reference would hold.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ludwig_hip.h"
#include "api_internal.h"
#include "d3q19.h"
#include "io_format.h"
#include "kernels.h"

using ldw::Geom;
using ldw::LbParams;
using ldw::FeParams;

/* ------------------------------------------------------------------------ */
/* Errors                                                                   */
/* ------------------------------------------------------------------------ */

static int g_fatal = 1;
static char g_errmsg[1024] = "";

static int ldw_fail(const char * file, int line, const char * fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  snprintf(g_errmsg, sizeof(g_errmsg), "%s (%s:%d)", buf, file, line);
  if (g_fatal) {
    fprintf(stderr, "libludwig_hip: fatal: %s\n", g_errmsg);
    fflush(stderr);
    abort();
  }
  return -1;
}

#define FAIL(...) return ldw_fail(__FILE__, __LINE__, __VA_ARGS__)

#define HIP_CHECK(call)							\
  do {									\
    hipError_t err_ = (call);						\
    if (err_ != hipSuccess) {						\
      return ldw_fail(__FILE__, __LINE__, "%s: %s", #call,		\
		      hipGetErrorString(err_));				\
    }									\
  } while (0)

#define CHECK_LAUNCH() HIP_CHECK(hipGetLastError())

extern "C" int ldw_error_mode_set(int fatal) { g_fatal = fatal ? 1 : 0; return 0; }
extern "C" const char * ldw_last_error(void) { return g_errmsg; }

int ldw::api::fail(const char * file, int line, const char * fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return ldw_fail(file, line, "%s", buf);
}

/* ------------------------------------------------------------------------ */
/* Device, stream, precision, timers                                        */
/* ------------------------------------------------------------------------ */

static hipStream_t g_stream = nullptr;
// a second stream, default priority like the first, for work that runs
// beside a launch on g_stream (the Lees-Edwards plane sums), and the
// events that order it
static hipStream_t g_stream2 = nullptr;
static hipEvent_t g_ev_fork = nullptr, g_ev_join = nullptr;
static int g_ready = 0;
static int flush_force(void);
static int materialise_derived(void);
static int g_precision = 0;

static int device_init(void) {
  if (g_ready) return 0;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    FAIL("no HIP device available: libludwig_hip runs on an MI355X (gfx950)");
  }
  HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&g_stream2, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreateWithFlags(&g_ev_fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&g_ev_join, hipEventDisableTiming));
  if (ldw::api::noalloc_scratch_init()) return -1;
  g_ready = 1;
  return 0;
}

int ldw::api::device_init(void) { return ::device_init(); }
hipStream_t ldw::api::stream(void) { return g_stream; }

static int precision_now(void) {
  if (g_precision == 0) {
    const char * e = getenv("LDW_PRECISION");
    g_precision = (e && atoi(e) == 32) ? 32 : 64;
  }
  return g_precision;
}

extern "C" int ldw_precision_set(int bits) {
  if (bits != 32 && bits != 64) FAIL("precision must be 32 or 64, got %d", bits);
  g_precision = bits;
  return 0;
}

extern "C" int ldw_precision(void) { return precision_now(); }

extern "C" int ldw_device_synchronize(void) {
  if (!g_ready) return 0;
  if (flush_force()) return -1;
  HIP_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

/* Device memory for callers that keep their own data on the GPU (the
 * library's HIP runtime; a second runtime in the same process, such as the
 * one a framework bundles, cannot share the device with it) */
extern "C" int ldw_device_alloc(long long bytes, void ** ptr) {
  if (ptr == nullptr || bytes < 0) FAIL("ldw_device_alloc: bad argument");
  *ptr = nullptr;
  if (device_init()) return -1;
  if (bytes == 0) return 0;
  HIP_CHECK(hipMalloc(ptr, (size_t) bytes));
  return 0;
}

extern "C" int ldw_device_free(void * ptr) {
  if (ptr == nullptr) return 0;
  HIP_CHECK(hipStreamSynchronize(g_stream));
  HIP_CHECK(hipFree(ptr));
  return 0;
}

/* Copy on the library's stream, complete on return */
extern "C" int ldw_memcpy(void * dst, const void * src, long long bytes, int kind) {
  if (bytes < 0) FAIL("ldw_memcpy: negative size");
  if (bytes == 0) return 0;
  if (device_init()) return -1;
  hipMemcpyKind k;
  switch (kind) {
  case tdpMemcpyHostToDevice: k = hipMemcpyHostToDevice; break;
  case tdpMemcpyDeviceToHost: k = hipMemcpyDeviceToHost; break;
  case tdpMemcpyDeviceToDevice: k = hipMemcpyDeviceToDevice; break;
  default: FAIL("ldw_memcpy: kind %d", kind);
  }
  if (flush_force()) return -1;
  HIP_CHECK(hipMemcpyAsync(dst, src, (size_t) bytes, k, g_stream));
  HIP_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

struct TimerAcc {
  double ms = 0.0;
  int count = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

static int g_timing = 0;
static std::map<std::string, TimerAcc> g_timers;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t event_get(void) {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static void timers_drain(void) {
  for (auto & kv : g_timers) {
    for (auto & pr : kv.second.pending) {
      float ms = 0.0f;
      (void) hipEventSynchronize(pr.second);
      (void) hipEventElapsedTime(&ms, pr.first, pr.second);
      kv.second.ms += ms;
      kv.second.count += 1;
      g_event_pool.push_back(pr.first);
      g_event_pool.push_back(pr.second);
    }
    kv.second.pending.clear();
  }
}

struct KTimer {
  const char * name;
  hipEvent_t a = nullptr, b = nullptr;
  explicit KTimer(const char * n) : name(n) {
    if (g_timing) {
      a = event_get(); b = event_get();
      if (a) (void) hipEventRecord(a, g_stream);
    }
  }
  ~KTimer() {
    if (g_timing && a && b) {
      (void) hipEventRecord(b, g_stream);
      g_timers[name].pending.push_back({a, b});
      if (g_timers[name].pending.size() > 256) timers_drain();
    }
  }
};

extern "C" int ldw_timer_enable(int on) { g_timing = on ? 1 : 0; return 0; }

extern "C" int ldw_timer_reset(void) {
  timers_drain();
  g_timers.clear();
  return 0;
}

extern "C" int ldw_timer_read(const char * name, double * ms, int * count) {
  timers_drain();
  auto it = g_timers.find(name);
  if (it == g_timers.end()) { *ms = 0.0; *count = 0; return 0; }
  *ms = it->second.ms;
  *count = it->second.count;
  return 0;
}

/* Events on the library's stream around a whole region of launches: mark
 * slot a before, slot b after, then the elapsed time between them.  Every
 * launch the region makes is on that stream, so this is the device time of
 * the region, gaps between launches included. */
static hipEvent_t g_marks[8] = {nullptr};

extern "C" int ldw_stream_mark(int slot) {
  if (slot < 0 || slot >= 8) FAIL("ldw_stream_mark: slot %d outside 0..7", slot);
  if (device_init()) return -1;
  if (!g_marks[slot]) HIP_CHECK(hipEventCreate(&g_marks[slot]));
  HIP_CHECK(hipEventRecord(g_marks[slot], g_stream));
  return 0;
}

extern "C" int ldw_stream_elapsed(int a, int b, double * ms) {
  if (a < 0 || a >= 8 || b < 0 || b >= 8 || !g_marks[a] || !g_marks[b]) {
    FAIL("ldw_stream_elapsed: slots %d, %d not both marked", a, b);
  }
  float t = 0.0f;
  HIP_CHECK(hipEventSynchronize(g_marks[b]));
  HIP_CHECK(hipEventElapsedTime(&t, g_marks[a], g_marks[b]));
  *ms = t;
  return 0;
}

static int dmalloc(void ** p, size_t bytes) {
  if (device_init()) return -1;
  HIP_CHECK(hipMalloc(p, bytes));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* pe, cs, physics                                                          */
/* ------------------------------------------------------------------------ */

struct pe_s { int verbose; };

extern "C" int pe_create(int, pe_enum_t flag, pe_t ** ppe) {
  if (!ppe) FAIL("pe_create: null argument");
  pe_t * pe = (pe_t *) calloc(1, sizeof(pe_t));
  pe->verbose = (flag == PE_VERBOSE);
  *ppe = pe;
  return 0;
}

extern "C" int pe_free(pe_t * pe) { free(pe); return 0; }

struct cs_s {
  pe_t * pe;
  int ntotal[3];
  int nhalo;
  int init;
  int str[3];
  int nsites;
  lees_edw_options_t leopts;   // copy for the I/O metadata (leesedwards.c:99)
};

extern "C" int cs_create(pe_t * pe, cs_t ** pcs) {
  if (!pcs) FAIL("cs_create: null argument");
  cs_t * cs = (cs_t *) calloc(1, sizeof(cs_t));
  cs->pe = pe;
  cs->ntotal[0] = cs->ntotal[1] = cs->ntotal[2] = 64;   /* coords.c default */
  cs->nhalo = 1;
  *pcs = cs;
  return 0;
}

extern "C" int cs_free(cs_t * cs) { free(cs); return 0; }

extern "C" int cs_ntotal_set(cs_t * cs, const int n[3]) {
  if (cs->init) FAIL("cs_ntotal_set after cs_init");
  for (int a = 0; a < 3; a++) {
    if (n[a] < 1) FAIL("cs_ntotal_set: bad size %d", n[a]);
    cs->ntotal[a] = n[a];
  }
  return 0;
}

extern "C" int cs_nhalo_set(cs_t * cs, int nhalo) {
  if (cs->init) FAIL("cs_nhalo_set after cs_init");
  if (nhalo < 1) FAIL("cs_nhalo_set: nhalo %d < 1", nhalo);
  cs->nhalo = nhalo;
  return 0;
}

/* strides as coords.c:211-215 */
extern "C" int cs_init(cs_t * cs) {
  cs->str[2] = 1;
  cs->str[1] = cs->str[2]*(cs->ntotal[2] + 2*cs->nhalo);
  cs->str[0] = cs->str[1]*(cs->ntotal[1] + 2*cs->nhalo);
  long long ns = (long long) cs->str[0]*(cs->ntotal[0] + 2*cs->nhalo);
  if (ns > 2147483647LL) FAIL("cs_init: lattice too large for int host indexing");
  cs->nsites = (int) ns;
  cs->init = 1;
  return 0;
}

extern "C" int cs_nhalo(cs_t * cs, int * nhalo) { *nhalo = cs->nhalo; return 0; }
extern "C" int cs_ntotal(cs_t * cs, int n[3]) { for (int a = 0; a < 3; a++) n[a] = cs->ntotal[a]; return 0; }
extern "C" int cs_nlocal(cs_t * cs, int n[3]) { for (int a = 0; a < 3; a++) n[a] = cs->ntotal[a]; return 0; }
extern "C" int cs_nsites(cs_t * cs, int * nsites) { *nsites = cs->nsites; return 0; }

/* coords.c:419-445: the box is [Lmin, Lmin + L) with Lmin = 0.5 */
extern "C" int cs_ltot(cs_t * cs, double ltot[3]) {
  for (int a = 0; a < 3; a++) ltot[a] = (double) cs->ntotal[a];
  return 0;
}

extern "C" int cs_lmin(cs_t *, double lmin[3]) {
  for (int a = 0; a < 3; a++) lmin[a] = 0.5;
  return 0;
}

/* coords.c:617-631 */
extern "C" int cs_index(cs_t * cs, int ic, int jc, int kc) {
  return cs->str[0]*(cs->nhalo + ic - 1) + cs->str[1]*(cs->nhalo + jc - 1)
    + cs->str[2]*(cs->nhalo + kc - 1);
}

static Geom geom_of(cs_t * cs) {
  Geom g;
  g.nx = cs->ntotal[0]; g.ny = cs->ntotal[1]; g.nz = cs->ntotal[2];
  g.nsite = (int64_t) g.nx*g.ny*g.nz;
  return g;
}

struct physics_s {
  double rho0, phi0, eta_shear, eta_bulk, mobility, fbody[3];
  double grad_mu[3];                // external chemical potential gradient
  int t_start, t_current, nsteps;   // time step control (physics.c:600-670)
};

static physics_t * g_phys = nullptr;

extern "C" int physics_create(pe_t *, physics_t ** phys) {
  physics_t * p = (physics_t *) calloc(1, sizeof(physics_t));
  p->rho0 = 1.0;            /* physics.c RHO_DEFAULT */
  p->eta_shear = 1.0/6.0;   /* physics.c:36 ETA_DEFAULT */
  p->eta_bulk = 1.0/6.0;
  p->mobility = 0.0;
  *phys = p;
  g_phys = p;
  return 0;
}

extern "C" int physics_free(physics_t * phys) {
  if (g_phys == phys) g_phys = nullptr;
  free(phys);
  return 0;
}

extern "C" int physics_ref(physics_t ** ref) {
  if (g_phys == nullptr) physics_create(nullptr, &g_phys);
  *ref = g_phys;
  return 0;
}

extern "C" int physics_rho0_set(physics_t * p, double v) { p->rho0 = v; return 0; }
extern "C" int physics_phi0_set(physics_t * p, double v) { p->phi0 = v; return 0; }
extern "C" int physics_eta_shear_set(physics_t * p, double v) { p->eta_shear = v; return 0; }
extern "C" int physics_eta_bulk_set(physics_t * p, double v) { p->eta_bulk = v; return 0; }
extern "C" int physics_mobility_set(physics_t * p, double v) { p->mobility = v; return 0; }
extern "C" int physics_fbody_set(physics_t * p, double f[3]) { for (int a = 0; a < 3; a++) p->fbody[a] = f[a]; return 0; }
extern "C" int physics_rho0(physics_t * p, double * v) { *v = p->rho0; return 0; }
extern "C" int physics_phi0(physics_t * p, double * v) { *v = p->phi0; return 0; }
extern "C" int physics_eta_shear(physics_t * p, double * v) { *v = p->eta_shear; return 0; }
extern "C" int physics_eta_bulk(physics_t * p, double * v) { *v = p->eta_bulk; return 0; }
extern "C" int physics_mobility(physics_t * p, double * v) { *v = p->mobility; return 0; }
extern "C" int physics_fbody(physics_t * p, double f[3]) { for (int a = 0; a < 3; a++) f[a] = p->fbody[a]; return 0; }
extern "C" int physics_grad_mu_set(physics_t * p, double gm[3]) { for (int a = 0; a < 3; a++) p->grad_mu[a] = gm[a]; return 0; }
extern "C" int physics_grad_mu(physics_t * p, double gm[3]) { for (int a = 0; a < 3; a++) gm[a] = p->grad_mu[a]; return 0; }

/* physics.c:600-670: the step counter runs nstart + 1 .. nstart + nsteps;
 * the "time" of step t_current is t_start + t_current - 1 */
extern "C" int physics_control_init_time(physics_t * p, int nstart, int nsteps) {
  p->t_start = nstart;
  p->nsteps = nsteps;
  p->t_current = nstart;
  return 0;
}

extern "C" int physics_control_next_step(physics_t * p) {
  p->t_current += 1;
  return (p->t_start + p->nsteps - p->t_current + 1);
}

extern "C" int physics_control_timestep(physics_t * p) { return p->t_current; }

extern "C" int physics_control_time(physics_t * p, double * t) {
  *t = 1.0*(p->t_start + p->t_current - 1.0);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Host <-> device layout                                                   */
/* ------------------------------------------------------------------------ */

// device site of interior point (ic, jc, kc), 1-based
static inline int64_t dsite(const Geom & g, int ic, int jc, int kc) {
  return ((int64_t) (ic - 1)*g.ny + (jc - 1))*g.nz + (kc - 1);
}

static inline int wrap1(int i, int n) {   // 1-based periodic image
  while (i < 1) i += n;
  while (i > n) i -= n;
  return i;
}

// Copy component planes between a host SoA array with halo (nhost sites,
// component stride nhost) and a device-layout array (component stride
// nsite).  Host halos receive the periodic images.
template <typename T>
static void pack_host(cs_t * cs, const double * host, int ncomp, T * dev) {
  Geom g = geom_of(cs);
  for (int c = 0; c < ncomp; c++)
    for (int ic = 1; ic <= g.nx; ic++)
      for (int jc = 1; jc <= g.ny; jc++)
	for (int kc = 1; kc <= g.nz; kc++)
	  dev[c*g.nsite + dsite(g, ic, jc, kc)]
	    = (T) host[(int64_t) c*cs->nsites + cs_index(cs, ic, jc, kc)];
}

template <typename T>
static void unpack_host(cs_t * cs, const T * dev, int ncomp, double * host) {
  Geom g = geom_of(cs);
  const int h = cs->nhalo;
  for (int c = 0; c < ncomp; c++)
    for (int ic = 1 - h; ic <= g.nx + h; ic++)
      for (int jc = 1 - h; jc <= g.ny + h; jc++)
	for (int kc = 1 - h; kc <= g.nz + h; kc++)
	  host[(int64_t) c*cs->nsites + cs_index(cs, ic, jc, kc)]
	    = (double) dev[c*g.nsite + dsite(g, wrap1(ic, g.nx), wrap1(jc, g.ny), wrap1(kc, g.nz))];
}

/* ------------------------------------------------------------------------ */
/* lb_t                                                                     */
/* ------------------------------------------------------------------------ */

struct lb_data_s {
  pe_t * pe;
  cs_t * cs;
  int ndim, nvel, ndist;
  lb_relaxation_enum_t nrelax;
  int isghost;
  int nsite;               // host sites incl. halo
  double * f;              // host copy, SoA f[p*nsite + index]
  int prec;                // 64 or 32
  void * d_g[2];           // device buffers (post-collision values)
  int cur;                 // current buffer
  int streamed;            // 1: logical f(x)[p] = g(x - c_p)[p]
  Geom geo;
  double * d_work;         // reduction scratch
  double * d_out;
  io_info_args_t iodata;   // distribution file options
  int meta_written;        // dist-metadata written
};

extern "C" io_options_t io_options_default(void) {
  io_options_t o = {IO_MODE_MPIIO, IO_RECORD_BINARY, IO_METADATA_V2, 0, 0, 0, {1, 1, 1}};
  return o;
}

extern "C" io_info_args_t io_info_args_default(void) {
  io_info_args_t a;
  a.input = io_options_default();
  a.output = io_options_default();
  a.grid[0] = a.grid[1] = a.grid[2] = 1;
  a.iofreq = 0;
  return a;
}

// Both directions of an object's I/O options are ones this library handles
static int io_args_check(const char * who, const io_info_args_t & a) {
  std::string err;
  if (ldw::io::options_supported(a.input, err) || ldw::io::options_supported(a.output, err)) {
    FAIL("%s: %s", who, err.c_str());
  }
  return 0;
}

extern "C" lb_data_options_t lb_data_options_default(void) {
  lb_data_options_t o;
  o.ndim = 3; o.nvel = 19; o.ndist = 1;
  o.nrelax = LB_RELAXATION_M10; o.halo = LB_HALO_FULL;
  o.reportimbalance = 0; o.usefirsttouch = 0;
  o.iodata = io_info_args_default();
  return o;
}

extern "C" int lb_data_create(pe_t * pe, cs_t * cs, const lb_data_options_t * opts,
			      lb_t ** plb) {
  if (!cs || !cs->init || !opts || !plb) FAIL("lb_data_create: bad arguments");
  if (opts->ndim != 3 || opts->nvel != 19) FAIL("lb_data_create: only D3Q19 is built (ndim %d nvel %d)", opts->ndim, opts->nvel);
  if (opts->ndist != 1) FAIL("lb_data_create: ndist %d not supported (the two-distribution binary LB is out of scope)", opts->ndist);
  if (io_args_check("lb_data_create", opts->iodata)) return -1;
  if (device_init()) return -1;

  lb_t * lb = (lb_t *) calloc(1, sizeof(lb_t));
  lb->pe = pe; lb->cs = cs;
  lb->ndim = 3; lb->nvel = 19; lb->ndist = 1;
  lb->nrelax = opts->nrelax;
  lb->isghost = 1;
  lb->nsite = cs->nsites;
  lb->geo = geom_of(cs);
  lb->prec = precision_now();
  lb->f = (double *) calloc((size_t) 19*lb->nsite, sizeof(double));
  size_t es = (lb->prec == 64) ? 8 : 4;
  for (int b = 0; b < 2; b++) {
    if (dmalloc(&lb->d_g[b], es*19*(size_t) lb->geo.nsite)) return -1;
  }
  if (dmalloc((void **) &lb->d_work, sizeof(double)*2*1024)) return -1;
  if (dmalloc((void **) &lb->d_out, sizeof(double)*4)) return -1;
  lb->cur = 0;
  lb->streamed = 0;
  lb->iodata = opts->iodata;
  *plb = lb;
  return 0;
}

extern "C" int lb_free(lb_t * lb) {
  if (!lb) return 0;
  if (g_ready) (void) hipStreamSynchronize(g_stream);
  for (int b = 0; b < 2; b++) (void) hipFree(lb->d_g[b]);
  (void) hipFree(lb->d_work);
  (void) hipFree(lb->d_out);
  free(lb->f);
  free(lb);
  return 0;
}

extern "C" int lb_ndist(lb_t * lb, int * ndist) { *ndist = lb->ndist; return 0; }
extern "C" int lb_nvel(lb_t * lb, int * nvel) { *nvel = lb->nvel; return 0; }

extern "C" int lb_f(lb_t * lb, int index, int p, int n, double * f) {
  *f = lb->f[((int64_t) n*19 + p)*lb->nsite + index];
  return 0;
}

extern "C" int lb_f_set(lb_t * lb, int index, int p, int n, double f) {
  lb->f[((int64_t) n*19 + p)*lb->nsite + index] = f;
  return 0;
}

/* lb_data.h:140-190 */
extern "C" int lb_0th_moment(lb_t * lb, int index, lb_dist_enum_t nd, double * rho) {
  *rho = 0.0;
  for (int p = 0; p < 19; p++) *rho += lb->f[((int64_t) nd*19 + p)*lb->nsite + index];
  return 0;
}

extern "C" int lb_1st_moment(lb_t * lb, int index, lb_dist_enum_t nd, double g[3]) {
  for (int a = 0; a < 3; a++) g[a] = 0.0;
  for (int p = 0; p < 19; p++)
    for (int a = 0; a < 3; a++)
      g[a] += ldw::CV.v[p][a]*lb->f[((int64_t) nd*19 + p)*lb->nsite + index];
  return 0;
}

/* lb_data.c:809-840, evaluated on the host copy */
double ldw::api::equilibrium_f(int p, double rho, const double u[3]) {
  double cs2 = (1.0/3.0);
  double rcs2 = 1.0/cs2;
  double udotc = 0.0;
  double sdotq = 0.0;
  for (int ia = 0; ia < 3; ia++) {
    udotc += u[ia]*ldw::CV.v[p][ia];
    for (int ib = 0; ib < 3; ib++) {
      double dab = (ia == ib);
      sdotq += (ldw::CV.v[p][ia]*ldw::CV.v[p][ib] - cs2*dab)*u[ia]*u[ib];
    }
  }
  return rho*ldw::weight(p)*(1.0 + rcs2*udotc + 0.5*rcs2*rcs2*sdotq);
}

extern "C" int lb_1st_moment_equilib_set(lb_t * lb, int index, double rho, double u[3]) {
  for (int p = 0; p < 19; p++)
    lb->f[(int64_t) p*lb->nsite + index] = ldw::api::equilibrium_f(p, rho, u);
  return 0;
}

/* lb_data.c:659-685 */
extern "C" int lb_init_rest_f(lb_t * lb, double rho0) {
  Geom g = lb->geo;
  for (int ic = 1; ic <= g.nx; ic++)
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	double u0[3] = {0.0, 0.0, 0.0};
	lb_1st_moment_equilib_set(lb, cs_index(lb->cs, ic, jc, kc), rho0, u0);
      }
  return lb_memcpy(lb, tdpMemcpyHostToDevice);
}

static double lb_rho0(void) {
  physics_t * p = nullptr;
  physics_ref(&p);
  return p->rho0;
}

extern "C" int lb_memcpy(lb_t * lb, tdpMemcpyKind flag) {
  const Geom g = lb->geo;
  const size_t n = (size_t) 19*g.nsite;
  const double rho0 = lb_rho0();

  if (flag == tdpMemcpyHostToDevice) {
    // device holds g with streamed = 0: logical f = g
    if (lb->prec == 64) {
      std::vector<double> buf(n);
      pack_host<double>(lb->cs, lb->f, 19, buf.data());
      HIP_CHECK(hipMemcpyAsync(lb->d_g[lb->cur], buf.data(), n*8, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    else {
      std::vector<float> buf(n);
      Geom gg = g;
      for (int p = 0; p < 19; p++)
	for (int ic = 1; ic <= gg.nx; ic++)
	  for (int jc = 1; jc <= gg.ny; jc++)
	    for (int kc = 1; kc <= gg.nz; kc++)
	      buf[p*gg.nsite + dsite(gg, ic, jc, kc)]
		= (float) (lb->f[(int64_t) p*lb->nsite + cs_index(lb->cs, ic, jc, kc)]
			   - ldw::weight(p)*rho0);
      HIP_CHECK(hipMemcpyAsync(lb->d_g[lb->cur], buf.data(), n*4, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    lb->streamed = 0;
    return 0;
  }

  if (flag == tdpMemcpyDeviceToHost) {
    std::vector<double> gd(n);
    if (lb->prec == 64) {
      HIP_CHECK(hipMemcpyAsync(gd.data(), lb->d_g[lb->cur], n*8, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    else {
      std::vector<float> gf(n);
      HIP_CHECK(hipMemcpyAsync(gf.data(), lb->d_g[lb->cur], n*4, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
      for (int p = 0; p < 19; p++)
	for (int64_t s = 0; s < g.nsite; s++)
	  gd[p*g.nsite + s] = (double) gf[p*g.nsite + s] + ldw::weight(p)*rho0;
    }
    // logical f: g itself, or g(x - c_p) once streamed; halos periodic
    const int h = lb->cs->nhalo;
    for (int p = 0; p < 19; p++) {
      const int sh = lb->streamed ? 1 : 0;
      const int cx = ldw::CV.v[p][0]*sh, cy = ldw::CV.v[p][1]*sh, cz = ldw::CV.v[p][2]*sh;
      for (int ic = 1 - h; ic <= g.nx + h; ic++)
	for (int jc = 1 - h; jc <= g.ny + h; jc++)
	  for (int kc = 1 - h; kc <= g.nz + h; kc++) {
	    int64_t s = dsite(g, wrap1(ic - cx, g.nx), wrap1(jc - cy, g.ny), wrap1(kc - cz, g.nz));
	    lb->f[(int64_t) p*lb->nsite + cs_index(lb->cs, ic, jc, kc)] = gd[p*g.nsite + s];
	  }
    }
    return 0;
  }
  return 0;
}

extern "C" int lb_halo(lb_t *) { return 0; }

extern "C" int lb_collision_relaxation_set(lb_t * lb, lb_relaxation_enum_t nrelax) {
  if (nrelax != LB_RELAXATION_M10 && nrelax != LB_RELAXATION_BGK && nrelax != LB_RELAXATION_TRT) {
    FAIL("lb_collision_relaxation_set: bad scheme %d", (int) nrelax);
  }
  lb->nrelax = nrelax;
  return 0;
}

extern "C" int lb_collision_ghost_modes_on(lb_t * lb) { lb->isghost = 1; return 0; }
extern "C" int lb_collision_ghost_modes_off(lb_t * lb) { lb->isghost = 0; return 0; }

/* ------------------------------------------------------------------------ */
/* le, map, wall, noise, visc                                               */
/* ------------------------------------------------------------------------ */

/* Lees-Edwards planes (leesedwards.c), one rank, steady shear.  Besides
 * the plane geometry the object owns the device scratch of the plane
 * kernels: x-face fluxes at the planes, the per-plane force corrections
 * and the reprojected distributions. */
struct lees_edw_s {
  int nplanes;
  lees_edw_options_t opts;
  cs_t * cs;
  int nh;
  int n[3];
  double dx_min, dx_sep, time0;
  ldw::LeGeom geo;
  double * d_xface;        // nplane*2*3*ny*nz
  double * d_fcor;         // nplane*3
  double * d_sbuf;         // nplane*2*5*ny*nz
  double * d_chface;       // nplane*2*ny*nz (Cahn-Hilliard faces of the single-pass step)
};

static int le_active(const lees_edw_t * le) { return le != nullptr && le->nplanes > 0; }

/* leesedwards.c:81-135, 233-275, 417-455, 464-510 */
extern "C" int lees_edw_create(pe_t *, cs_t * cs, const lees_edw_options_t * opts, lees_edw_t ** ple) {
  if (!ple) FAIL("lees_edw_create: null argument");
  lees_edw_t * le = (lees_edw_t *) calloc(1, sizeof(lees_edw_t));
  *ple = le;
  le->cs = cs;
  if (opts && cs) cs->leopts = *opts;
  if (opts == nullptr || opts->nplanes == 0) return 0;
  if (!cs || !cs->init) FAIL("lees_edw_create: the coordinate system must be initialised first");
  le->opts = *opts;
  le->nplanes = opts->nplanes;
  le->nh = cs->nhalo;
  for (int a = 0; a < 3; a++) le->n[a] = cs->ntotal[a];
  if (opts->nplanes < 0) FAIL("lees_edw_create: %d planes", opts->nplanes);
  if (le->n[0] % opts->nplanes) {
    FAIL("lees_edw_create: Number of planes must divide system size (%d planes, x size %d)",
	 opts->nplanes, le->n[0]);
  }
  if (opts->type != LE_SHEAR_TYPE_STEADY) {
    FAIL("lees_edw_create: only steady shear (LE_SHEAR_TYPE_STEADY) is built, type %d", opts->type);
  }
  if (opts->nplanes > ldw::LE_MAX_PLANES) FAIL("lees_edw_create: at most %d planes", ldw::LE_MAX_PLANES);
  if (!std::isfinite(opts->uy)) FAIL("lees_edw_create: plane speed %g", opts->uy);
  if (le->nh > ldw::LE_MAX_NH) FAIL("lees_edw_create: halo %d above %d", le->nh, ldw::LE_MAX_NH);
  le->dx_sep = 1.0*le->n[0]/opts->nplanes;
  le->dx_min = 0.5*le->dx_sep;
  le->time0 = 1.0*opts->nt0;
  ldw::LeGeom & g = le->geo;
  g.nplane = opts->nplanes;
  g.nh = le->nh;
  g.nxbuf = 2*le->nh*opts->nplanes;
  g.sep = le->n[0]/opts->nplanes;
  for (int p = 0; p < opts->nplanes; p++) {
    g.loc[p] = (int) (le->dx_min + p*le->dx_sep);     // lees_edw_plane_location
    // lees_edw_checks: no plane within nhalo of the x boundaries
    if (g.loc[p] <= le->nh || g.loc[p] > le->n[0] - le->nh) {
      FAIL("lees_edw_create: Wall at domain boundary (plane %d at x = %d, halo %d)", p + 1,
	   g.loc[p], le->nh);
    }
  }
  const size_t nyz = (size_t) le->n[1]*le->n[2];
  if (dmalloc((void **) &le->d_xface, 8*(size_t) opts->nplanes*2*3*nyz)) return -1;
  if (dmalloc((void **) &le->d_fcor, 8*(size_t) opts->nplanes*3)) return -1;
  if (dmalloc((void **) &le->d_sbuf, 8*(size_t) opts->nplanes*2*5*nyz)) return -1;
  if (dmalloc((void **) &le->d_chface, 8*(size_t) opts->nplanes*2*nyz)) return -1;
  return 0;
}

extern "C" int lees_edw_free(lees_edw_t * le) {
  if (!le) return 0;
  if (g_ready) (void) hipStreamSynchronize(g_stream);
  if (le->d_xface) (void) hipFree(le->d_xface);
  if (le->d_fcor) (void) hipFree(le->d_fcor);
  if (le->d_sbuf) (void) hipFree(le->d_sbuf);
  if (le->d_chface) (void) hipFree(le->d_chface);
  free(le);
  return 0;
}

extern "C" int lees_edw_nplane_total(lees_edw_t * le) { return le ? le->nplanes : 0; }

/* lees_edw_shear_rate (leesedwards.c:700-712) */
extern "C" int lees_edw_shear_rate(lees_edw_t * le, double * gammadot) {
  *gammadot = le->nplanes ? le->opts.uy*le->nplanes/(double) le->n[0] : 0.0;
  return 0;
}

/* lees_edw_plane_location (leesedwards.c:617-636), 1-based local x */
extern "C" int lees_edw_plane_location(lees_edw_t * le, int np) {
  return (le && np >= 0 && np < le->nplanes) ? le->geo.loc[np] : -1;
}

static double phys_time(void) {
  physics_t * p = nullptr;
  physics_ref(&p);
  double t;
  physics_control_time(p, &t);
  return t;
}

static int phys_timestep(void) {
  physics_t * p = nullptr;
  physics_ref(&p);
  return physics_control_timestep(p);
}

/* lees_edw_buffer_displacement (leesedwards.c:645-670), steady shear */
static double le_buffer_displacement(const lees_edw_t * le, int ib, double t) {
  if (t < 0.0) t = 0.0;
  const double tle = t - le->time0;
  const int duy = (ib % (2*le->nh) < le->nh) ? -1 : +1;   // lees_edw_buffer_duy
  return tle*le->opts.uy*duy;
}

/* the split of each buffer plane's displacement for field_leesedwards
 * (field.c:470-475: t0 = 0, fr = 1 - fraction) or hydro_lees_edwards
 * (hydro.c:398-403: t0 = 1, fr = fraction) */
/* integer part of a displacement already reduced modulo ly; it lies in
 * [-ly, ly] for any finite input, which the kernels' periodic indexing
 * relies on, so anything else (a non-finite time or speed) is refused */
static int le_floor(double dy, double ly, int * jdy) {
  if (!(fabs(dy) <= ly)) FAIL("Lees-Edwards displacement %g outside [-%g, %g]", dy, ly, ly);
  *jdy = (int) floor(dy);
  return 0;
}

static int le_buffer_shift(const lees_edw_t * le, double t0, int one_minus, ldw::LeShift * sh) {
  memset(sh, 0, sizeof(*sh));
  const double ly = (double) le->n[1];
  const double t = phys_time();
  for (int ib = 0; ib < le->geo.nxbuf; ib++) {
    double dy = le_buffer_displacement(le, ib, t + t0);
    dy = fmod(dy, ly);
    int jdy = 0;
    if (le_floor(dy, ly, &jdy)) return -1;
    sh->jdy[ib] = jdy;
    sh->fr[ib] = one_minus ? 1.0 - (dy - jdy) : dy - jdy;
  }
  return 0;
}

struct map_s { int ndata; };

extern "C" map_options_t map_options_default(void) { map_options_t o = {0, 0}; return o; }

extern "C" int map_create(pe_t *, cs_t *, const map_options_t * options, map_t ** pmap) {
  if (options && options->is_porous_media) FAIL("map_create: porous media not supported");
  map_t * m = (map_t *) calloc(1, sizeof(map_t));
  m->ndata = options ? options->ndata : 0;
  *pmap = m;
  return 0;
}

extern "C" int map_free(map_t ** map) { if (map) { free(*map); *map = nullptr; } return 0; }

struct wall_s { int present; };

extern "C" int wall_create(pe_t *, cs_t *, map_t *, lb_t *, wall_t ** p) {
  *p = (wall_t *) calloc(1, sizeof(wall_t));
  return 0;
}

extern "C" int wall_free(wall_t * wall) { free(wall); return 0; }

/* ------------------------------------------------------------------------ */
/* field_t, field_grad_t                                                    */
/* ------------------------------------------------------------------------ */

static void pending_drop_field(field_t * f);

struct field_s {
  pe_t * pe;
  cs_t * cs;
  char name[64];
  int nf;
  int nhcomm;
  int nsites;
  double * data;           // host, SoA data[n*nsites + index]
  int prec;
  void * d_data[2];        // device ping-pong
  int cur;
  long version;            // bumped whenever the device values change
  Geom geo;
  io_info_args_t iodata;   // data file options
  int meta_written;
  double init_sum;         // field_init_sum (cahn_hilliard_stats_time0)
  lees_edw_t * le;         // planes: phi of the buffer planes in d_phib
  double * d_phib;
};

struct field_grad_s {
  field_t * field;
  int level;
  double * grad;           // host grad[a*nsite + index]
  double * delsq;          // host
  void * d_mu;             // device mu for the version mu_version
  long mu_version;
  void * d_grad;           // materialised only for reads
  void * d_delsq;
  grad_ft d2;
  long grad_version;       // phi version at the last field_grad_compute()
  int grad_buf;            // phi buffer holding that version
  // with Lees-Edwards planes the gradient is formed by field_grad_compute()
  // itself: real sites (d_gr, d_dr) and buffer sites (d_gb, d_db)
  double * d_gr, * d_dr, * d_gb, * d_db;
  long le_version;
  int le_partial;          // after a single-pass step: gradient only next to the planes
};

extern "C" field_options_t field_options_ndata_nhalo(int ndata, int nhalo) {
  field_options_t o = {ndata, nhalo, 0, 0, 0, io_info_args_default()};
  return o;
}

extern "C" int field_create(pe_t * pe, cs_t * cs, lees_edw_t * le, const char * name,
			    const field_options_t * opts, field_t ** pobj) {
  if (!cs || !cs->init || !opts || !pobj) FAIL("field_create: bad arguments");
  if (opts->ndata != 1) FAIL("field_create: only scalar fields (ndata 1) are on the path, got %d", opts->ndata);
  if (opts->nhcomm > cs->nhalo) FAIL("field_create: halo %d exceeds coordinate halo %d", opts->nhcomm, cs->nhalo);
  if (io_args_check("field_create", opts->iodata)) return -1;
  field_t * f = (field_t *) calloc(1, sizeof(field_t));
  f->pe = pe; f->cs = cs;
  snprintf(f->name, sizeof(f->name), "%s", name ? name : "");
  f->nf = opts->ndata;
  f->nhcomm = opts->nhcomm;
  f->nsites = cs->nsites;
  f->geo = geom_of(cs);
  f->prec = precision_now();
  f->data = (double *) calloc((size_t) f->nf*f->nsites, sizeof(double));
  size_t es = (f->prec == 64) ? 8 : 4;
  for (int b = 0; b < 2; b++) {
    if (dmalloc(&f->d_data[b], es*f->geo.nsite)) return -1;
    HIP_CHECK(hipMemsetAsync(f->d_data[b], 0, es*f->geo.nsite, g_stream));
  }
  f->version = 1;
  f->iodata = opts->iodata;
  *pobj = f;
  if (le_active(le)) {
    if (f->prec != 64) FAIL("field_create: Lees-Edwards planes are built for the FP64 storage");
    // stencils reach 2 sites across a plane (third-order advection) and the
    // gradient is needed one buffer plane deep (lees_edw_ic_to_buff asserts
    // |di| <= nhalo)
    if (cs->nhalo < 2) FAIL("field_create: a field with Lees-Edwards planes needs nhalo >= 2, got %d", cs->nhalo);
    f->le = le;
    const size_t nb = (size_t) le->geo.nxbuf*f->geo.ny*f->geo.nz;
    if (dmalloc((void **) &f->d_phib, 8*nb)) return -1;
    HIP_CHECK(hipMemsetAsync(f->d_phib, 0, 8*nb, g_stream));
  }
  return 0;
}

extern "C" int field_free(field_t * f) {
  if (!f) return 0;
  pending_drop_field(f);
  if (g_ready) (void) hipStreamSynchronize(g_stream);
  for (int b = 0; b < 2; b++) (void) hipFree(f->d_data[b]);
  if (f->d_phib) (void) hipFree(f->d_phib);
  free(f->data);
  free(f);
  return 0;
}

extern "C" int field_nf(field_t * f, int * nf) { *nf = f->nf; return 0; }

extern "C" int field_scalar(field_t * f, int index, double * phi) {
  *phi = f->data[index];
  return 0;
}

extern "C" int field_scalar_set(field_t * f, int index, double phi) {
  f->data[index] = phi;
  return 0;
}

extern "C" int field_memcpy(field_t * f, tdpMemcpyKind flag) {
  if (flag == tdpMemcpyHostToDevice && (flush_force() || materialise_derived())) return -1;
  const Geom g = f->geo;
  if (flag == tdpMemcpyHostToDevice) {
    if (f->prec == 64) {
      std::vector<double> buf(g.nsite);
      pack_host<double>(f->cs, f->data, 1, buf.data());
      HIP_CHECK(hipMemcpyAsync(f->d_data[f->cur], buf.data(), 8*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    else {
      std::vector<float> buf(g.nsite);
      pack_host<float>(f->cs, f->data, 1, buf.data());
      HIP_CHECK(hipMemcpyAsync(f->d_data[f->cur], buf.data(), 4*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    f->version++;
  }
  else if (flag == tdpMemcpyDeviceToHost) {
    if (f->prec == 64) {
      std::vector<double> buf(g.nsite);
      HIP_CHECK(hipMemcpyAsync(buf.data(), f->d_data[f->cur], 8*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
      unpack_host<double>(f->cs, buf.data(), 1, f->data);
    }
    else {
      std::vector<float> buf(g.nsite);
      HIP_CHECK(hipMemcpyAsync(buf.data(), f->d_data[f->cur], 4*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
      unpack_host<float>(f->cs, buf.data(), 1, f->data);
    }
  }
  return 0;
}

extern "C" int field_halo(field_t *) { return 0; }

/* The stencils are evaluated inside the kernels; these entry points only
 * name the choice for field_grad_set() (gradient_3d_7pt_fluid.h,
 * gradient_3d_27pt_fluid.h) */
extern "C" int grad_3d_7pt_fluid_d2(field_grad_t *) { return 0; }
extern "C" int grad_3d_27pt_fluid_d2(field_grad_t *) { return 0; }

extern "C" int field_grad_create(pe_t *, field_t * f, int level, field_grad_t ** pobj) {
  if (!f || !pobj) FAIL("field_grad_create: bad arguments");
  field_grad_t * g = (field_grad_t *) calloc(1, sizeof(field_grad_t));
  g->field = f;
  g->level = level;
  g->grad = (double *) calloc((size_t) 3*f->nsites, sizeof(double));
  g->delsq = (double *) calloc((size_t) f->nsites, sizeof(double));
  size_t es = (f->prec == 64) ? 8 : 4;
  if (dmalloc(&g->d_mu, es*f->geo.nsite)) return -1;
  g->mu_version = 0;
  g->d2 = grad_3d_7pt_fluid_d2;
  *pobj = g;
  if (le_active(f->le)) {
    const size_t n = (size_t) f->geo.nsite;
    const size_t nb = (size_t) f->le->geo.nxbuf*f->geo.ny*f->geo.nz;
    if (dmalloc((void **) &g->d_gr, 8*3*n) || dmalloc((void **) &g->d_dr, 8*n)) return -1;
    if (dmalloc((void **) &g->d_gb, 8*3*nb) || dmalloc((void **) &g->d_db, 8*nb)) return -1;
    HIP_CHECK(hipMemsetAsync(g->d_gb, 0, 8*3*nb, g_stream));
    HIP_CHECK(hipMemsetAsync(g->d_db, 0, 8*nb, g_stream));
  }
  return 0;
}

extern "C" void field_grad_free(field_grad_t * g) {
  if (!g) return;
  if (g_ready) (void) hipStreamSynchronize(g_stream);
  (void) hipFree(g->d_mu);
  if (g->d_grad) (void) hipFree(g->d_grad);
  if (g->d_delsq) (void) hipFree(g->d_delsq);
  if (g->d_gr) { (void) hipFree(g->d_gr); (void) hipFree(g->d_dr); }
  if (g->d_gb) { (void) hipFree(g->d_gb); (void) hipFree(g->d_db); }
  free(g->grad);
  free(g->delsq);
  free(g);
}

extern "C" int field_grad_set(field_grad_t * g, grad_ft d2, grad_ft) {
  if (d2 != grad_3d_7pt_fluid_d2 && d2 != grad_3d_27pt_fluid_d2)
    FAIL("field_grad_set: the built stencils are grad_3d_7pt_fluid_d2 and grad_3d_27pt_fluid_d2");
  g->d2 = d2;
  return 0;
}

static int grad_stencil(const field_grad_t * g) {
  return (g && g->d2 == grad_3d_27pt_fluid_d2) ? 27 : 7;
}

/* Deferred: the stencil is evaluated inside the kernels that consume it.
 * The phi version is recorded, so that a later field_grad_memcpy() returns
 * the gradient of that phi, as the reference's stored gradient would be
 * (its statistics after a step use the gradient from the step's start). */
extern "C" int field_grad_compute(field_grad_t * g) {
  g->grad_version = g->field->version;
  g->grad_buf = g->field->cur;
  field_t * f = g->field;
  if (le_active(f->le)) {
    // field_leesedwards() then the gradient (field_grad.c:319-325): the
    // buffer planes at the current time, then real and buffer sites
    if (flush_force() || materialise_derived()) return -1;
    ldw::LeShift sh;
    if (le_buffer_shift(f->le, 0.0, 1, &sh)) return -1;
    KTimer t("le_gradient");
    ldw::launch_le_phi_buffer(f->geo, f->le->geo, sh, (const double *) f->d_data[f->cur],
			      f->d_phib, g_stream);
    ldw::launch_le_grad(f->geo, f->le->geo, grad_stencil(g), (const double *) f->d_data[f->cur],
			f->d_phib, g->d_gr, g->d_dr, g->d_gb, g->d_db, g_stream);
    CHECK_LAUNCH();
    g->le_version = f->version;
    g->le_partial = 0;
  }
  return 0;
}

extern "C" int field_grad_memcpy(field_grad_t * g, tdpMemcpyKind flag) {
  if (flag != tdpMemcpyDeviceToHost) return 0;
  field_t * f = g->field;
  const Geom geo = f->geo;
  if (f->prec != 64) FAIL("field_grad_memcpy: gradient read-back is built for FP64 only");
  if (le_active(f->le)) {
    // the gradient the last field_grad_compute() formed; after a
    // single-pass step only the planes' neighbourhood was formed, so the
    // rest is formed now from the phi and buffer planes of that step's start
    if (g->le_partial) {
      ldw::launch_le_grad(geo, f->le->geo, grad_stencil(g), (const double *) f->d_data[g->grad_buf],
			  f->d_phib, g->d_gr, g->d_dr, g->d_gb, g->d_db, g_stream);
      CHECK_LAUNCH();
      g->le_partial = 0;
    }
    std::vector<double> b1(3*geo.nsite), b2(geo.nsite);
    HIP_CHECK(hipMemcpyAsync(b1.data(), g->d_gr, 8*3*geo.nsite, hipMemcpyDeviceToHost, g_stream));
    HIP_CHECK(hipMemcpyAsync(b2.data(), g->d_dr, 8*geo.nsite, hipMemcpyDeviceToHost, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
    unpack_host<double>(f->cs, b1.data(), 3, g->grad);
    unpack_host<double>(f->cs, b2.data(), 1, g->delsq);
    return 0;
  }
  if (!g->d_grad) {
    if (dmalloc(&g->d_grad, 8*3*(size_t) geo.nsite)) return -1;
    if (dmalloc(&g->d_delsq, 8*(size_t) geo.nsite)) return -1;
  }
  if (materialise_derived()) return -1;
  // the phi of the last field_grad_compute(): the current one, or the one
  // a single update ago (still in the other buffer)
  int buf = f->cur;
  if (f->version == g->grad_version + 1 && g->grad_buf == 1 - f->cur) buf = g->grad_buf;
  ldw::launch_grad_f64(geo, grad_stencil(g), (const double *) f->d_data[buf], (double *) g->d_grad,
		       (double *) g->d_delsq, g_stream);
  CHECK_LAUNCH();
  std::vector<double> b1(3*geo.nsite), b2(geo.nsite);
  HIP_CHECK(hipMemcpyAsync(b1.data(), g->d_grad, 8*3*geo.nsite, hipMemcpyDeviceToHost, g_stream));
  HIP_CHECK(hipMemcpyAsync(b2.data(), g->d_delsq, 8*geo.nsite, hipMemcpyDeviceToHost, g_stream));
  HIP_CHECK(hipStreamSynchronize(g_stream));
  unpack_host<double>(f->cs, b1.data(), 3, g->grad);
  unpack_host<double>(f->cs, b2.data(), 1, g->delsq);
  return 0;
}

extern "C" int field_grad_scalar_grad(field_grad_t * g, int index, double grad[3]) {
  for (int a = 0; a < 3; a++) grad[a] = g->grad[(int64_t) a*g->field->nsites + index];
  return 0;
}

extern "C" int field_grad_scalar_delsq(field_grad_t * g, int index, double * delsq) {
  *delsq = g->delsq[index];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* hydro_t                                                                  */
/* ------------------------------------------------------------------------ */

struct hydro_s {
  pe_t * pe;
  cs_t * cs;
  int nhcomm;
  int nsite;
  double * rho, * u, * force;      // host SoA with halo
  int prec;
  void * d_rho, * d_u, * d_force;
  int force_uniform;               // 1: force is uniform fzero (not materialised)
  double fzero[3];
  int u_uniform;                   // 1: u is uniform uzero
  double uzero[3];
  Geom geo;
  // deferred phi_force_calculation (launched fused with the next
  // phi_cahn_hilliard on the same phi, or alone when the force is needed)
  int fp_pending;
  field_t * fp_phi;
  long fp_version;
  FeParams fp_params;
  field_grad_t * fp_dphi;
  // after a single-pass step the force is not stored: it is re-formed from
  // the phi the step started from (still in the other phi buffer) on demand
  void * d_u2;                     // second velocity buffer for the step kernel
  int fd_derived;
  field_t * fd_phi;
  int fd_buf;
  FeParams fd_params;
  field_grad_t * fd_dphi;
  io_info_args_t rho_io, u_io;     // "rho" and "vel" file options
  int rho_meta_written, u_meta_written;
  lees_edw_t * le;                 // planes: velocity of the buffer planes
  double * d_ub;
};

static hydro_t * g_pending = nullptr;   // at most one deferred force
static hydro_t * g_derived = nullptr;   // at most one force to re-form

static void pending_drop_field(field_t * f) {
  if (g_pending && g_pending->fp_phi == f) {
    g_pending->fp_pending = 0;
    g_pending = nullptr;
  }
  if (g_derived && g_derived->fd_phi == f) {
    g_derived->fd_derived = 0;
    g_derived = nullptr;
  }
}

/* hydro_options.c:32-46 */
extern "C" hydro_options_t hydro_options_default(void) {
  hydro_options_t o;
  o.nhcomm = 1;
  o.rho = field_options_ndata_nhalo(1, 1);
  o.u = field_options_ndata_nhalo(3, 1);
  o.force = field_options_ndata_nhalo(3, 1);
  o.eta = field_options_ndata_nhalo(1, 1);
  return o;
}

extern "C" int hydro_create(pe_t * pe, cs_t * cs, lees_edw_t * le, const hydro_options_t * opts,
			    hydro_t ** pobj) {
  if (!cs || !cs->init || !pobj) FAIL("hydro_create: bad arguments");
  const hydro_options_t dflt = hydro_options_default();
  if (!opts) opts = &dflt;
  if (io_args_check("hydro_create (rho)", opts->rho.iodata)) return -1;
  if (io_args_check("hydro_create (vel)", opts->u.iodata)) return -1;
  hydro_t * h = (hydro_t *) calloc(1, sizeof(hydro_t));
  h->pe = pe; h->cs = cs;
  h->nhcomm = opts->nhcomm;
  h->rho_io = opts->rho.iodata;
  h->u_io = opts->u.iodata;
  h->nsite = cs->nsites;
  h->geo = geom_of(cs);
  h->prec = precision_now();
  h->rho = (double *) calloc((size_t) h->nsite, sizeof(double));
  h->u = (double *) calloc((size_t) 3*h->nsite, sizeof(double));
  h->force = (double *) calloc((size_t) 3*h->nsite, sizeof(double));
  size_t es = (h->prec == 64) ? 8 : 4;
  if (dmalloc(&h->d_rho, es*h->geo.nsite)) return -1;
  if (dmalloc(&h->d_u, es*3*h->geo.nsite)) return -1;
  if (dmalloc(&h->d_force, es*3*h->geo.nsite)) return -1;
  // hydro_create zeroes rho? The reference initialises rho to rho0 at
  // start (hydro_rho0); u and force are zero.
  HIP_CHECK(hipMemsetAsync(h->d_rho, 0, es*h->geo.nsite, g_stream));
  h->u_uniform = 1;
  h->force_uniform = 1;
  *pobj = h;
  if (le_active(le)) {
    if (h->prec != 64) FAIL("hydro_create: Lees-Edwards planes are built for the FP64 storage");
    h->le = le;
    if (dmalloc((void **) &h->d_ub, 8*3*(size_t) le->geo.nxbuf*h->geo.ny*h->geo.nz)) return -1;
  }
  return 0;
}

extern "C" int hydro_free(hydro_t * h) {
  if (!h) return 0;
  if (g_pending == h) g_pending = nullptr;
  if (g_derived == h) g_derived = nullptr;
  if (g_ready) (void) hipStreamSynchronize(g_stream);
  if (h->d_u2) (void) hipFree(h->d_u2);
  if (h->d_ub) (void) hipFree(h->d_ub);
  (void) hipFree(h->d_rho); (void) hipFree(h->d_u); (void) hipFree(h->d_force);
  free(h->rho); free(h->u); free(h->force);
  free(h);
  return 0;
}

static int hydro_materialise(hydro_t * h) {
  if (flush_force()) return -1;
  if (materialise_derived()) return -1;
  const Geom g = h->geo;
  if (h->force_uniform) {
    for (int a = 0; a < 3; a++) {
      if (h->prec == 64) ldw::launch_fill_f64((double *) h->d_force + a*g.nsite, g.nsite, h->fzero[a], g_stream);
      else ldw::launch_fill_f32((float *) h->d_force + a*g.nsite, g.nsite, (float) h->fzero[a], g_stream);
    }
    CHECK_LAUNCH();
    h->force_uniform = 0;
  }
  if (h->u_uniform) {
    for (int a = 0; a < 3; a++) {
      if (h->prec == 64) ldw::launch_fill_f64((double *) h->d_u + a*g.nsite, g.nsite, h->uzero[a], g_stream);
      else ldw::launch_fill_f32((float *) h->d_u + a*g.nsite, g.nsite, (float) h->uzero[a], g_stream);
    }
    CHECK_LAUNCH();
    h->u_uniform = 0;
  }
  return 0;
}

extern "C" int hydro_memcpy(hydro_t * h, tdpMemcpyKind flag) {
  if (flush_force()) return -1;
  if (materialise_derived()) return -1;
  const Geom g = h->geo;
  if (flag == tdpMemcpyHostToDevice) {
    if (h->prec == 64) {
      std::vector<double> b(g.nsite), b3(3*g.nsite), f3(3*g.nsite);
      pack_host<double>(h->cs, h->rho, 1, b.data());
      pack_host<double>(h->cs, h->u, 3, b3.data());
      pack_host<double>(h->cs, h->force, 3, f3.data());
      HIP_CHECK(hipMemcpyAsync(h->d_rho, b.data(), 8*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipMemcpyAsync(h->d_u, b3.data(), 24*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipMemcpyAsync(h->d_force, f3.data(), 24*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    else {
      std::vector<float> b(g.nsite), b3(3*g.nsite), f3(3*g.nsite);
      pack_host<float>(h->cs, h->rho, 1, b.data());
      pack_host<float>(h->cs, h->u, 3, b3.data());
      pack_host<float>(h->cs, h->force, 3, f3.data());
      HIP_CHECK(hipMemcpyAsync(h->d_rho, b.data(), 4*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipMemcpyAsync(h->d_u, b3.data(), 12*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipMemcpyAsync(h->d_force, f3.data(), 12*g.nsite, hipMemcpyHostToDevice, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
    }
    h->force_uniform = 0;
    h->u_uniform = 0;
    return 0;
  }
  if (flag == tdpMemcpyDeviceToHost) {
    if (hydro_materialise(h)) return -1;
    if (h->prec == 64) {
      std::vector<double> b(g.nsite), b3(3*g.nsite), f3(3*g.nsite);
      HIP_CHECK(hipMemcpyAsync(b.data(), h->d_rho, 8*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipMemcpyAsync(b3.data(), h->d_u, 24*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipMemcpyAsync(f3.data(), h->d_force, 24*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
      unpack_host<double>(h->cs, b.data(), 1, h->rho);
      unpack_host<double>(h->cs, b3.data(), 3, h->u);
      unpack_host<double>(h->cs, f3.data(), 3, h->force);
    }
    else {
      std::vector<float> b(g.nsite), b3(3*g.nsite), f3(3*g.nsite);
      HIP_CHECK(hipMemcpyAsync(b.data(), h->d_rho, 4*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipMemcpyAsync(b3.data(), h->d_u, 12*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipMemcpyAsync(f3.data(), h->d_force, 12*g.nsite, hipMemcpyDeviceToHost, g_stream));
      HIP_CHECK(hipStreamSynchronize(g_stream));
      unpack_host<float>(h->cs, b.data(), 1, h->rho);
      unpack_host<float>(h->cs, b3.data(), 3, h->u);
      unpack_host<float>(h->cs, f3.data(), 3, h->force);
    }
  }
  return 0;
}

extern "C" int hydro_u_halo(hydro_t *) { return 0; }

extern "C" int hydro_f_zero(hydro_t * h, const double fzero[3]) {
  if (g_derived == h) { g_derived = nullptr; h->fd_derived = 0; }
  h->force_uniform = 1;
  for (int a = 0; a < 3; a++) h->fzero[a] = fzero[a];
  return 0;
}

extern "C" int hydro_u_zero(hydro_t * h, const double uzero[3]) {
  h->u_uniform = 1;
  for (int a = 0; a < 3; a++) h->uzero[a] = uzero[a];
  return 0;
}

extern "C" int hydro_u(hydro_t * h, int index, double u[3]) {
  for (int a = 0; a < 3; a++) u[a] = h->u[(int64_t) a*h->nsite + index];
  return 0;
}

extern "C" int hydro_u_set(hydro_t * h, int index, const double u[3]) {
  for (int a = 0; a < 3; a++) h->u[(int64_t) a*h->nsite + index] = u[a];
  return 0;
}

extern "C" int hydro_rho(hydro_t * h, int index, double * rho) { *rho = h->rho[index]; return 0; }

extern "C" int hydro_f_local(hydro_t * h, int index, double force[3]) {
  for (int a = 0; a < 3; a++) force[a] = h->force[(int64_t) a*h->nsite + index];
  return 0;
}

extern "C" int hydro_f_local_set(hydro_t * h, int index, const double force[3]) {
  for (int a = 0; a < 3; a++) h->force[(int64_t) a*h->nsite + index] = force[a];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Symmetric free energy                                                    */
/* ------------------------------------------------------------------------ */

enum {LDW_FE_SYMMETRIC = 1};

struct fe_s {
  int id;
  int use_stress_relaxation;
};

struct fe_symm_s {
  fe_t super;
  pe_t * pe;
  cs_t * cs;
  field_t * phi;
  field_grad_t * dphi;
  fe_symm_param_t param;
};

extern "C" int fe_symm_create(pe_t * pe, cs_t * cs, field_t * f, field_grad_t * grd,
			      fe_symm_t ** p) {
  fe_symm_t * fe = (fe_symm_t *) calloc(1, sizeof(fe_symm_t));
  fe->super.id = LDW_FE_SYMMETRIC;
  fe->pe = pe; fe->cs = cs; fe->phi = f; fe->dphi = grd;
  fe->param.a = -0.003125;      /* symmetric.c:40-42 defaults */
  fe->param.b = +0.003125;
  fe->param.kappa = +0.002;
  *p = fe;
  return 0;
}

extern "C" int fe_symm_free(fe_symm_t * fe) { free(fe); return 0; }

extern "C" int fe_symm_param_set(fe_symm_t * fe, fe_symm_param_t values) {
  fe->param = values;
  fe->dphi->mu_version = 0;     // mu depends on the parameters
  return 0;
}

extern "C" int fe_symm_param(fe_symm_t * fe, fe_symm_param_t * values) {
  *values = fe->param;
  return 0;
}

/* symmetric.c:267-274 */
extern "C" int fe_symm_interfacial_width(fe_symm_t * fe, double * xi) {
  *xi = sqrt(-2.0*fe->param.kappa/fe->param.a);
  return 0;
}

extern "C" int fe_symm_mu(fe_symm_t * fe, int index, double * mu) {
  double phi = fe->phi->data[index];
  double delsq = fe->dphi->delsq[index];
  *mu = fe->param.a*phi + fe->param.b*phi*phi*phi - fe->param.kappa*delsq;
  return 0;
}

extern "C" int fe_symm_fed(fe_symm_t * fe, int index, double * fed) {
  double phi = fe->phi->data[index];
  double g[3];
  field_grad_scalar_grad(fe->dphi, index, g);
  *fed = (0.5*fe->param.a + 0.25*fe->param.b*phi*phi)*phi*phi
    + 0.5*fe->param.kappa*(g[0]*g[0] + g[1]*g[1] + g[2]*g[2]);
  return 0;
}

extern "C" int fe_symm_str(fe_symm_t * fe, int index, double s[3][3]) {
  double phi = fe->phi->data[index];
  double delsq = fe->dphi->delsq[index];
  double g[3];
  field_grad_scalar_grad(fe->dphi, index, g);
  const double kappa = fe->param.kappa;
  double p0 = 0.5*fe->param.a*phi*phi + 0.75*fe->param.b*phi*phi*phi*phi
    - kappa*phi*delsq - 0.5*kappa*(g[0]*g[0] + g[1]*g[1] + g[2]*g[2]);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) s[a][b] = p0*(a == b) + kappa*g[a]*g[b];
  return 0;
}

static FeParams fe_params_of(fe_symm_t * fe, int method) {
  FeParams p;
  physics_t * phys = nullptr;
  physics_ref(&phys);
  p.a = fe->param.a; p.b = fe->param.b; p.kappa = fe->param.kappa;
  p.mobility = phys->mobility;
  int order = 1;
  advection_order(&order);
  p.order = order;
  p.force_method = (method == FE_FORCE_METHOD_PHI_GRADMU) ? 1 : 0;
  p.stencil = grad_stencil(fe->dphi);
  // the external gradient acts on every Cahn-Hilliard face flux
  // (phi_cahn_hilliard.c:1344) and, under phi grad mu, on the force when
  // it is non-zero (phi_grad_mu.c:138-145)
  for (int a = 0; a < 3; a++) p.gmu[a] = phys->grad_mu[a];
  p.gmu_on = (p.gmu[0] != 0.0 || p.gmu[1] != 0.0 || p.gmu[2] != 0.0);
  return p;
}

/* ------------------------------------------------------------------------ */
/* Force                                                                    */
/* ------------------------------------------------------------------------ */

struct pth_s { int method; };

extern "C" int pth_create(pe_t *, cs_t *, int method, pth_t ** pth) {
  if (method != FE_FORCE_METHOD_NO_FORCE && method != FE_FORCE_METHOD_STRESS_DIVERGENCE
      && method != FE_FORCE_METHOD_PHI_GRADMU) {
    FAIL("pth_create: force method %d is not built (no_force, stress_divergence, phi_gradmu are)", method);
  }
  pth_t * p = (pth_t *) calloc(1, sizeof(pth_t));
  p->method = method;
  *pth = p;
  return 0;
}

extern "C" int pth_free(pth_t * pth) { free(pth); return 0; }

static fe_symm_t * as_symm(fe_t * fe) {
  if (fe == nullptr || fe->id != LDW_FE_SYMMETRIC) return nullptr;
  return (fe_symm_t *) fe;
}

/* phi_force.c:74-148, fluid-only branch */
extern "C" int phi_force_calculation(pe_t *, cs_t *, lees_edw_t * le, wall_t *, pth_t * pth,
				     fe_t * fe, map_t *, field_t * phi, hydro_t * hydro) {
  if (hydro == nullptr) return 0;
  if (pth == nullptr || pth->method == FE_FORCE_METHOD_NO_FORCE) return 0;
  fe_symm_t * fs = as_symm(fe);
  if (!fs) FAIL("phi_force_calculation: only the symmetric free energy is built");
  if (phi->prec != hydro->prec) FAIL("phi_force_calculation: precision mismatch");
  if (flush_force()) return -1;
  if (materialise_derived()) return -1;
  if (le_active(le)) {
    // with planes the reference always takes the flux form (phi_force.c:89-97)
    if (!le_active(phi->le) || phi->le != le) FAIL("phi_force_calculation: phi was not created with these Lees-Edwards planes");
    field_grad_t * dphi = fs->dphi;
    if (dphi->le_version != phi->version) {
      FAIL("phi_force_calculation: with Lees-Edwards planes field_grad_compute() of the current phi must come first");
    }
    const Geom g = phi->geo;
    const FeParams p = fe_params_of(fs, pth->method);
    const double ra = 0.5/((double) g.ny*(double) g.nz);
    const int mode = hydro->force_uniform ? 0 : 1;
    {
      KTimer t("le_force");
      ldw::launch_le_force(g, le->geo, p, (const double *) phi->d_data[phi->cur], phi->d_phib,
			   dphi->d_gr, dphi->d_dr, dphi->d_gb, dphi->d_db, le->d_xface, le->d_fcor,
			   ra, (double *) hydro->d_force, mode, hydro->fzero, g_stream);
    }
    CHECK_LAUNCH();
    hydro->force_uniform = 0;
    return 0;
  }
  hydro->fp_pending = 1;
  hydro->fp_phi = phi;
  hydro->fp_version = phi->version;
  hydro->fp_params = fe_params_of(fs, pth->method);
  hydro->fp_dphi = fs->dphi;
  g_pending = hydro;
  return 0;
}

/* Launch a deferred force calculation on its own (K2), writing F and mu */
static int flush_force(void) {
  hydro_t * hydro = g_pending;
  if (hydro == nullptr) return 0;
  g_pending = nullptr;
  hydro->fp_pending = 0;
  field_t * phi = hydro->fp_phi;
  field_grad_t * dphi = hydro->fp_dphi;
  if (phi->version != hydro->fp_version) FAIL("deferred force: phi changed before the force was formed");
  const Geom g = phi->geo;
  const int mode = hydro->force_uniform ? 0 : 1;
  {
    KTimer t("phi_force_mu");
    if (phi->prec == 64) {
      ldw::launch_phi_force_mu_f64(g, hydro->fp_params, (const double *) phi->d_data[phi->cur],
				   (double *) hydro->d_force, (double *) dphi->d_mu,
				   mode, hydro->fzero, g_stream);
    }
    else {
      ldw::launch_phi_force_mu_f32(g, hydro->fp_params, (const float *) phi->d_data[phi->cur],
				   (float *) hydro->d_force, (float *) dphi->d_mu,
				   mode, hydro->fzero, g_stream);
    }
  }
  CHECK_LAUNCH();
  hydro->force_uniform = 0;
  dphi->mu_version = phi->version;
  return 0;
}

/* Re-form the force of the last single-pass step from the phi it started
 * from (K2 on the other phi buffer); mu is not kept */
static int materialise_derived(void) {
  hydro_t * hydro = g_derived;
  if (hydro == nullptr) return 0;
  g_derived = nullptr;
  hydro->fd_derived = 0;
  field_t * phi = hydro->fd_phi;
  const Geom g = phi->geo;
  const double fz[3] = {0.0, 0.0, 0.0};
  {
    KTimer t("phi_force_mu");
    if (phi->prec == 64) {
      ldw::launch_phi_force_mu_f64(g, hydro->fd_params, (const double *) phi->d_data[hydro->fd_buf],
				   (double *) hydro->d_force, (double *) hydro->fd_dphi->d_mu,
				   0, fz, g_stream);
    }
    else {
      ldw::launch_phi_force_mu_f32(g, hydro->fd_params, (const float *) phi->d_data[hydro->fd_buf],
				   (float *) hydro->d_force, (float *) hydro->fd_dphi->d_mu,
				   0, fz, g_stream);
    }
  }
  CHECK_LAUNCH();
  hydro->fd_dphi->mu_version = 0;   // mu of the previous phi, not the current one
  hydro->force_uniform = 0;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Cahn-Hilliard                                                            */
/* ------------------------------------------------------------------------ */

static int g_adv_order = 1;    /* advection.c:74 default */

extern "C" int advection_order_set(const int order) {
  // advection.c:463-472 dispatches orders 1-5; order 5 reads phi three
  // sites away, which the periodic kernels here do by index wrap (the
  // reference needs a coordinate halo of 3 for it)
  if (order < 1 || order > 5) FAIL("advection_order_set: order %d (1 to 5)", order);
  g_adv_order = order;
  return 0;
}

extern "C" int advection_order(int * order) { *order = g_adv_order; return 0; }

struct phi_ch_s {
  phi_ch_info_t info;
  pe_t * pe;
  cs_t * cs;
  lees_edw_t * le;
  int prec;
  void * d_csum;           // conserve 1: per-site compensation, kept between steps
  double * d_sum;          // conserve 2: the sum of phi (3 doubles: rounded sum, pair)
  double * d_work;         // conserve 2: partial sums of the tree reduction
};

/* Conserve 2 subtracts the drift of the plain sum of phi from the initial
 * sum (phi_cahn_hilliard.c:1102-1175).  That sum's rounding, not the true
 * drift, dominates the difference, so the correction depends on the order
 * of the additions, which the reference itself does not fix (atomics on
 * the GPU, per-thread partial sums under OpenMP).  By default the sum here
 * is the compensated tree reduction (Neumaier pairs merged with TwoSum,
 * deterministic, within a few ulp of the exact sum, a few microseconds).
 * ldw_conserve2_site_order(1), or LDW_CONSERVE2_SITE_ORDER=1 in the
 * environment, selects the plain sum in site order instead: one thread
 * adds every value in turn, as the reference's one-thread build does, so
 * results are bit-identical with it (about 10 ms per step at 128^3). */
static int g_conserve2_site_order = -1;

static int conserve2_site_order(void) {
  if (g_conserve2_site_order < 0) {
    const char * e = getenv("LDW_CONSERVE2_SITE_ORDER");
    g_conserve2_site_order = (e && atoi(e) != 0) ? 1 : 0;
  }
  return g_conserve2_site_order;
}

extern "C" int ldw_conserve2_site_order(int on) {
  g_conserve2_site_order = on ? 1 : 0;
  return 0;
}

/* phi_cahn_hilliard.c:128-163: conserve 1 keeps a compensation field
 * ("compensated sum"), zero at the start */
extern "C" int phi_ch_create(pe_t * pe, cs_t * cs, lees_edw_t * le, phi_ch_info_t * info,
			     phi_ch_t ** pch) {
  if (!cs || !cs->init || !pch) FAIL("phi_ch_create: bad arguments");
  if (info && info->noise) FAIL("phi_ch_create: order parameter noise not supported");
  if (info && (info->conserve < 0 || info->conserve > 2)) {
    FAIL("phi_ch_create: cahn_hilliard_options_conserve %d (0, 1 or 2)", info->conserve);
  }
  phi_ch_t * p = (phi_ch_t *) calloc(1, sizeof(phi_ch_t));
  if (info) p->info = *info;
  p->pe = pe; p->cs = cs; p->le = le;
  p->prec = precision_now();
  if (p->info.conserve == 1) {
    const Geom g = geom_of(cs);
    const size_t bytes = ((p->prec == 64) ? 8 : 4)*(size_t) g.nsite;
    if (dmalloc(&p->d_csum, bytes)) return -1;
    HIP_CHECK(hipMemsetAsync(p->d_csum, 0, bytes, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
  }
  if (p->info.conserve == 2) {
    if (dmalloc((void **) &p->d_sum, 3*sizeof(double))) return -1;
    if (dmalloc((void **) &p->d_work, 2*sizeof(double)*(size_t) ldw::sum_blocks(geom_of(cs).nsite))) return -1;
  }
  *pch = p;
  return 0;
}

extern "C" int phi_ch_free(phi_ch_t * pch) {
  if (!pch) return 0;
  if (g_ready) (void) hipStreamSynchronize(g_stream);
  if (pch->d_csum) (void) hipFree(pch->d_csum);
  if (pch->d_sum) (void) hipFree(pch->d_sum);
  if (pch->d_work) (void) hipFree(pch->d_work);
  free(pch);
  return 0;
}

/* kahan_add_double (util_sum.c:30-40), host side; no value-changing
 * optimisations are enabled, so the operations keep their order */
static void kahan_add_host(double & sum, double & cs, double val) {
  const double y = val + cs;
  const double t = sum + y;
  cs = y - (t - sum);
  sum = t;
}

/* cahn_hilliard_stats.c:59-76: the reference sum of conserve 2, formed as
 * the reference forms it on one thread (cahn_hilliard_stats.c:233-276:
 * a Kahan sum over the sites in order, added as a Kahan pair into a zero
 * sum twice, then sum + cs) */
extern "C" int cahn_hilliard_stats_time0(phi_ch_t * pch, field_t * phi, map_t *) {
  if (!pch || !phi) FAIL("cahn_hilliard_stats_time0: null argument");
  if (field_memcpy(phi, tdpMemcpyDeviceToHost)) return -1;
  double s[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  const Geom g = phi->geo;
  for (int ic = 1; ic <= g.nx; ic++)
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	kahan_add_host(s[0], c[0], phi->data[cs_index(phi->cs, ic, jc, kc)]);
      }
  kahan_add_host(s[1], c[1], s[0]);
  kahan_add_host(s[1], c[1], c[0]);
  kahan_add_host(s[2], c[2], s[1]);
  kahan_add_host(s[2], c[2], c[1]);
  phi->init_sum = s[2] + c[2];
  return 0;
}

/* conserve 2 after the forward step (phi_cahn_hilliard.c:1102-1175):
 * phi -= (S - field_init_sum)/nfluid, S the site-order sum; both kernels
 * stay on the stream (the reference copies S to the host and back) */
static int ch_subtract_drift(phi_ch_t * pch, field_t * phi) {
  if (pch->info.conserve != 2) return 0;
  const Geom g = phi->geo;
  const bool site_order = conserve2_site_order();
  if (phi->prec == 64) {
    const double * a = (const double *) phi->d_data[phi->cur];
    if (site_order) ldw::launch_site_order_sum_f64(a, g.nsite, pch->d_sum, g_stream);
    else ldw::launch_klein_sum_f64(a, g.nsite, pch->d_work, pch->d_sum, g_stream);
    ldw::launch_subtract_drift_f64((double *) phi->d_data[phi->cur], g.nsite, pch->d_sum,
				   phi->init_sum, g_stream);
  }
  else {
    const float * a = (const float *) phi->d_data[phi->cur];
    if (site_order) ldw::launch_site_order_sum_f32(a, g.nsite, pch->d_sum, g_stream);
    else ldw::launch_klein_sum_f32(a, g.nsite, pch->d_work, pch->d_sum, g_stream);
    ldw::launch_subtract_drift_f32((float *) phi->d_data[phi->cur], g.nsite, pch->d_sum,
				   phi->init_sum, g_stream);
  }
  CHECK_LAUNCH();
  return 0;
}

/* phi_cahn_hilliard.c:213-290 */
extern "C" int phi_cahn_hilliard(phi_ch_t * pch, fe_t * fe, field_t * phi, hydro_t * hydro,
				 map_t *, noise_t * noise) {
  if (noise && pch->info.noise) FAIL("phi_cahn_hilliard: noise not supported");
  if (pch->info.conserve && pch->prec != phi->prec) {
    FAIL("phi_cahn_hilliard: phi_ch_t (%d-bit) and phi (%d-bit) differ in precision", pch->prec, phi->prec);
  }
  fe_symm_t * fs = as_symm(fe);
  if (!fs) FAIL("phi_cahn_hilliard: only the symmetric free energy is built");
  field_grad_t * dphi = fs->dphi;
  FeParams p = fe_params_of(fs, FE_FORCE_METHOD_STRESS_DIVERGENCE);
  const Geom g = phi->geo;
  if (materialise_derived()) return -1;

  if (le_active(phi->le)) {
    // hydro_lees_edwards(), advection_x() and the diffusive fluxes with
    // the planes, phi_ch_le_fix_fluxes(), the update (phi_cahn_hilliard.c:
    // 213-290)
    lees_edw_t * le = phi->le;
    if (dphi->le_version != phi->version) {
      FAIL("phi_cahn_hilliard: with Lees-Edwards planes field_grad_compute() of the current phi must come first");
    }
    if (p.order > 4) {
      FAIL("phi_cahn_hilliard: advection order %d with Lees-Edwards planes: the buffer planes hold the "
	   "two sites next to each plane (orders 1-4)", p.order);
    }
    if (flush_force()) return -1;
    const double * u = nullptr;
    double * ub = nullptr;
    if (hydro) {
      if (hydro->le != le) FAIL("phi_cahn_hilliard: hydro was not created with the planes of phi");
      if (hydro_materialise(hydro)) return -1;
      u = (const double *) hydro->d_u;
      ub = hydro->d_ub;
      ldw::LeShift sh;
      if (le_buffer_shift(le, 1.0, 0, &sh)) return -1;
      ldw::launch_le_u_buffer(g, le->geo, sh, le->opts.uy, u, ub, g_stream);
      CHECK_LAUNCH();
    }
    // lees_edw_plane_dy (leesedwards.c:880-893): t uy, no time offset
    const double dy = phys_time()*le->opts.uy;
    const double ly = (double) g.ny;
    const double dyw = fmod(+dy, ly), dye = fmod(-dy, ly);
    int jw = 0, je = 0;
    if (le_floor(dyw, ly, &jw) || le_floor(dye, ly, &je)) return -1;
    {
      KTimer t("le_cahn_hilliard");
      ldw::launch_le_cahn_hilliard(g, le->geo, p, (const double *) phi->d_data[phi->cur], phi->d_phib,
				   dphi->d_dr, dphi->d_db, u, ub, le->d_xface, jw, dyw - jw, je, dye - je,
				   (double *) phi->d_data[1 - phi->cur], (double *) pch->d_csum, g_stream);
    }
    CHECK_LAUNCH();
    phi->cur = 1 - phi->cur;
    phi->version++;
    return ch_subtract_drift(pch, phi);
  }

  if (hydro && hydro->fp_pending && g_pending == hydro && hydro->fp_phi == phi
      && hydro->fp_version == phi->version && p.order <= 4) {
    // force and Cahn-Hilliard update in one pass (K23); a uniform zero
    // velocity is passed as no velocity (advective flux exactly zero)
    const void * uf = hydro->d_u;
    if (hydro->u_uniform) {
      if (hydro->uzero[0] == 0.0 && hydro->uzero[1] == 0.0 && hydro->uzero[2] == 0.0) {
	uf = nullptr;
      }
      else {
	// materialise u only (the pending force is launched fused below)
	const Geom gg = hydro->geo;
	for (int a = 0; a < 3; a++) {
	  if (hydro->prec == 64) ldw::launch_fill_f64((double *) hydro->d_u + a*gg.nsite, gg.nsite, hydro->uzero[a], g_stream);
	  else ldw::launch_fill_f32((float *) hydro->d_u + a*gg.nsite, gg.nsite, (float) hydro->uzero[a], g_stream);
	}
	CHECK_LAUNCH();
	hydro->u_uniform = 0;
      }
    }
    FeParams q = hydro->fp_params;
    q.order = p.order;
    q.mobility = p.mobility;
    const int mode = hydro->force_uniform ? 0 : 1;
    g_pending = nullptr;
    hydro->fp_pending = 0;
    {
      KTimer t("force_ch");
      if (phi->prec == 64) {
	ldw::launch_force_ch_f64(g, q, (const double *) phi->d_data[phi->cur], (const double *) uf,
				 (double *) hydro->d_force, (double *) phi->d_data[1 - phi->cur],
				 (double *) pch->d_csum, mode, hydro->fzero, g_stream);
      }
      else {
	ldw::launch_force_ch_f32(g, q, (const float *) phi->d_data[phi->cur], (const float *) uf,
				 (float *) hydro->d_force, (float *) phi->d_data[1 - phi->cur],
				 (float *) pch->d_csum, mode, hydro->fzero, g_stream);
      }
    }
    CHECK_LAUNCH();
    hydro->force_uniform = 0;
    phi->cur = 1 - phi->cur;
    phi->version++;
    return ch_subtract_drift(pch, phi);
  }
  if (flush_force()) return -1;

  if (dphi->mu_version != phi->version) {
    KTimer t("mu");
    if (phi->prec == 64) ldw::launch_mu_f64(g, p, (const double *) phi->d_data[phi->cur], (double *) dphi->d_mu, g_stream);
    else ldw::launch_mu_f32(g, p, (const float *) phi->d_data[phi->cur], (float *) dphi->d_mu, g_stream);
    CHECK_LAUNCH();
    dphi->mu_version = phi->version;
  }

  const void * u = nullptr;
  if (hydro) {
    if (hydro->u_uniform) {
      if (hydro->uzero[0] != 0.0 || hydro->uzero[1] != 0.0 || hydro->uzero[2] != 0.0) {
	if (hydro_materialise(hydro)) return -1;
	u = hydro->d_u;
      }
      // uniform zero velocity: the advective flux is u_f phi = 0 exactly
      // (sign of zero aside), the same as no hydrodynamics
    }
    else {
      u = hydro->d_u;
    }
  }

  {
    KTimer t("cahn_hilliard");
    if (phi->prec == 64) {
      ldw::launch_cahn_hilliard_f64(g, p, (const double *) phi->d_data[phi->cur],
				    (const double *) dphi->d_mu, (const double *) u,
				    (double *) phi->d_data[1 - phi->cur], (double *) pch->d_csum,
				    g_stream);
    }
    else {
      ldw::launch_cahn_hilliard_f32(g, p, (const float *) phi->d_data[phi->cur],
				    (const float *) dphi->d_mu, (const float *) u,
				    (float *) phi->d_data[1 - phi->cur], (float *) pch->d_csum,
				    g_stream);
    }
  }
  CHECK_LAUNCH();
  phi->cur = 1 - phi->cur;
  phi->version++;
  return ch_subtract_drift(pch, phi);
}

/* ------------------------------------------------------------------------ */
/* Collision and propagation                                                */
/* ------------------------------------------------------------------------ */

/* Relaxation rates as collision.c:1163-1246 and the per-site helpers
 * (collision.c:1255-1420) with fixed viscosities */
LbParams ldw::api::lb_params(double rho0, double eta, double eta_nu, int relax,
			     const double fglobal[3]) {
  LbParams p;
  memset(&p, 0, sizeof(p));
  const double cs2 = (1.0/3.0);
  const double rtau_shear = 1.0/(0.5 + eta/(rho0*cs2));
  p.rtau_shear = rtau_shear;
  p.relax = relax;
  if (relax == LB_RELAXATION_M10) {
    p.rtau_bulk = 1.0/(0.5 + eta_nu/(rho0*cs2));
    for (int m = 10; m < 19; m++) p.rtau_ghost[m] = 1.0;
  }
  else if (relax == LB_RELAXATION_BGK) {
    p.rtau_bulk = 1.0/(0.5 + eta/(rho0*cs2));
    for (int m = 10; m < 19; m++) p.rtau_ghost[m] = rtau_shear;
  }
  else {
    p.rtau_bulk = 1.0/(0.5 + eta_nu/(rho0*cs2));
    double tau = eta/(rho0*cs2);
    double rg = 0.5 + 2.0*tau/(tau + 3.0/8.0);
    if (rg > 2.0) rg = 2.0;
    p.rtau_ghost[10] = rtau_shear; p.rtau_ghost[14] = rtau_shear; p.rtau_ghost[18] = rtau_shear;
    p.rtau_ghost[11] = rg; p.rtau_ghost[12] = rg; p.rtau_ghost[13] = rg;
    p.rtau_ghost[15] = rg; p.rtau_ghost[16] = rg; p.rtau_ghost[17] = rg;
  }
  for (int a = 0; a < 3; a++) p.fglobal[a] = fglobal ? fglobal[a] : 0.0;
  p.rho0 = rho0;
  return p;
}

/* The lattice's collision parameters from the physics object */
static LbParams lb_params_of(lb_t * lb, const double fextra[3]) {
  physics_t * phys = nullptr;
  physics_ref(&phys);
  // collision.c:1929-1940: force_global = fbody + pulsatile (zero); the
  // site force is force_global + hydro force
  double fg[3];
  for (int a = 0; a < 3; a++) {
    fg[a] = phys->fbody[a];
    if (fextra) fg[a] = fg[a] + fextra[a];
  }
  return ldw::api::lb_params(phys->rho0, phys->eta_shear, phys->eta_bulk, (int) lb->nrelax, fg);
}

/* collision.c:143-162 */
extern "C" int lb_collide(lb_t * lb, hydro_t * hydro, map_t *, noise_t * noise, fe_t * fe,
			  visc_t * visc) {
  if (hydro == nullptr) return 0;
  if (noise) FAIL("lb_collide: lattice fluctuations are not supported");
  if (visc) FAIL("lb_collide: viscosity models are not supported");
  if (fe && fe->use_stress_relaxation) FAIL("lb_collide: stress relaxation force method is not supported");
  if (hydro->prec != lb->prec) FAIL("lb_collide: precision mismatch between lb and hydro");
  if (flush_force()) return -1;
  if (materialise_derived()) return -1;

  const Geom g = lb->geo;
  const void * force = nullptr;
  LbParams p;
  if (hydro->force_uniform) {
    p = lb_params_of(lb, hydro->fzero);
  }
  else {
    p = lb_params_of(lb, nullptr);
    force = hydro->d_force;
  }
  {
    KTimer t("collide");
    if (lb->prec == 64) {
      ldw::launch_collide_f64(g, p, (const double *) lb->d_g[lb->cur], (double *) lb->d_g[1 - lb->cur],
			      (const double *) force, (double *) hydro->d_rho,
			      (double *) hydro->d_u, lb->streamed, g_stream);
    }
    else {
      ldw::launch_collide_f32(g, p, (const float *) lb->d_g[lb->cur], (float *) lb->d_g[1 - lb->cur],
			      (const float *) force, (float *) hydro->d_rho,
			      (float *) hydro->d_u, lb->streamed, g_stream);
    }
  }
  CHECK_LAUNCH();
  lb->cur = 1 - lb->cur;
  lb->streamed = 0;
  hydro->u_uniform = 0;
  return 0;
}

/* model_le.c:78-170: on the post-collision distributions of lb_collide();
 * the displacement is that of time step t = physics_control_timestep() */
/* the integer displacement, fraction and velocity jump of each plane side
 * at time step t = physics_control_timestep() (model_le.c:100-105, 245-395,
 * 585-625) */
static int le_lb_shifts(const lees_edw_t * le, double ly, int dj[2], double fr[2], double duy[2]) {
  const double t = 1.0*phys_timestep();
  // lees_edw_buffer_displacement at ib = nhalo (sign +1); plane_uy_now
  const double dy = le_buffer_displacement(le, le->nh, t);
  for (int side = 0; side < 2; side++) {
    const int cx = 1 - 2*side;
    const double d = fmod(dy*cx, ly);
    int jdy = 0;
    if (le_floor(d, ly, &jdy)) return -1;
    dj[side] = jdy;                // model_le.c:378 floor(fmod(dy cx, L_y))
    fr[side] = d - jdy;
    duy[side] = -1.0*cx*le->opts.uy;
  }
  return 0;
}

extern "C" int lb_data_apply_le_boundary_conditions(lb_t * lb, lees_edw_t * le) {
  if (!lb || !le) FAIL("lb_data_apply_le_boundary_conditions: null argument");
  if (!le_active(le)) return 0;
  if (lb->prec != 64) FAIL("lb_data_apply_le_boundary_conditions: Lees-Edwards planes are built for FP64");
  if (lb->streamed) FAIL("lb_data_apply_le_boundary_conditions: needs the post-collision distributions (call it between lb_collide and lb_propagation)");
  int dj[2];
  double fr[2], duy[2];
  if (le_lb_shifts(le, (double) lb->geo.ny, dj, fr, duy)) return -1;
  {
    KTimer tm("le_lb_bc");
    ldw::launch_le_lb_bc(lb->geo, le->geo, dj, fr, duy, (double *) lb->d_g[lb->cur], le->d_sbuf, g_stream);
  }
  CHECK_LAUNCH();
  return 0;
}

/* model_le.c:649-715: f = w (rho0 + rho0 c.u/cs2 + (rho0 u u - eta grad u):Q
 * / 2 cs2^2) with u_y = lees_edw_steady_uy(x), on the host copy */
extern "C" int lb_le_init_shear_profile(lb_t * lb, lees_edw_t * le) {
  if (!lb || !le) FAIL("lb_le_init_shear_profile: null argument");
  if (!le_active(le)) FAIL("lb_le_init_shear_profile: Cannot use LE_init_profile with no planes");
  physics_t * phys = nullptr;
  physics_ref(&phys);
  const double rho0 = phys->rho0, eta = phys->eta_shear;
  double u[3] = {0.0, 0.0, 0.0}, gradu[3][3] = {{0.0}};
  lees_edw_shear_rate(le, &gradu[0][1]);
  const Geom g = lb->geo;
  const double cs2 = (1.0/3.0);
  const double rcs2 = 1.0/cs2;
  for (int ic = 1; ic <= g.nx; ic++) {
    // lees_edw_steady_uy (leesedwards.c:520-545)
    const double xglobal = 0 + (double) ic - 0.5;
    const int nplane = (int) ((le->dx_min + xglobal)/le->dx_sep);
    u[1] = xglobal*gradu[0][1] - le->opts.uy*nplane;
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	const int index = cs_index(lb->cs, ic, jc, kc);
	for (int p = 0; p < 19; p++) {
	  double cdotu = 0.0, sdotq = 0.0;
	  for (int i = 0; i < 3; i++) {
	    cdotu += ldw::CV.v[p][i]*u[i];
	    for (int j = 0; j < 3; j++) {
	      const double dij = (i == j);
	      const double qij = ldw::CV.v[p][i]*ldw::CV.v[p][j] - cs2*dij;
	      sdotq += (rho0*u[i]*u[j] - eta*gradu[i][j])*qij;
	    }
	  }
	  lb->f[(int64_t) p*lb->nsite + index]
	    = ldw::weight(p)*(rho0 + rcs2*rho0*cdotu + 0.5*rcs2*rcs2*sdotq);
	}
      }
  }
  return lb_memcpy(lb, tdpMemcpyHostToDevice);
}

/* propagation.c:43-105: the streaming is deferred into the next collision
 * (pull); two propagations in a row materialise the first. */
extern "C" int lb_propagation(lb_t * lb) {
  if (lb->streamed) {
    KTimer t("propagate");
    if (lb->prec == 64) ldw::launch_propagate_f64(lb->geo, (const double *) lb->d_g[lb->cur], (double *) lb->d_g[1 - lb->cur], g_stream);
    else ldw::launch_propagate_f32(lb->geo, (const float *) lb->d_g[lb->cur], (float *) lb->d_g[1 - lb->cur], g_stream);
    CHECK_LAUNCH();
    lb->cur = 1 - lb->cur;
  }
  lb->streamed = 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Additions                                                                */
/* ------------------------------------------------------------------------ */

static int step_staged(lb_t * lb, hydro_t * hydro, field_t * phi, field_grad_t * dphi,
		       fe_t * fe, pth_t * pth, phi_ch_t * pch) {
  const double fzero[3] = {0.0, 0.0, 0.0};
  const double uzero[3] = {0.0, 0.0, 0.0};
  lees_edw_t * le = phi->le;
  if (hydro_f_zero(hydro, fzero)) return -1;
  if (field_halo(phi)) return -1;
  if (field_grad_compute(dphi)) return -1;
  if (phi_force_calculation(nullptr, nullptr, le, nullptr, pth, fe, nullptr, phi, hydro)) return -1;
  if (phi_cahn_hilliard(pch, fe, phi, hydro, nullptr, nullptr)) return -1;
  if (hydro_u_zero(hydro, uzero)) return -1;
  if (lb_collide(lb, hydro, nullptr, nullptr, fe, nullptr)) return -1;
  if (le_active(le) && lb_data_apply_le_boundary_conditions(lb, le)) return -1;
  if (lb_halo(lb)) return -1;
  if (lb_propagation(lb)) return -1;
  return 0;
}

/* The same step as step_staged() in one kernel (binary_step.hip).  Needs
 * streamed distributions, a stored velocity field and one precision. */
// the single-pass step; conserve 1 takes the compensated update (pch's
// per-site compensation), conserve 2 the global subtraction after it, as
// phi_cahn_hilliard does
static int step_fused(lb_t * lb, hydro_t * hydro, field_t * phi, fe_symm_t * fs, pth_t * pth,
		      phi_ch_t * pch) {
  const Geom g = lb->geo;
  if (hydro->d_u2 == nullptr) {
    size_t es = (hydro->prec == 64) ? 8 : 4;
    if (dmalloc(&hydro->d_u2, es*3*(size_t) g.nsite)) return -1;
  }
  if (flush_force()) return -1;
  // the force of the previous fused step is superseded, not re-formed
  if (g_derived == hydro) { g_derived = nullptr; hydro->fd_derived = 0; }
  if (materialise_derived()) return -1;
  FeParams p = fe_params_of(fs, pth->method);
  LbParams q = lb_params_of(lb, nullptr);
  const double fz[3] = {0.0, 0.0, 0.0};
  if (field_grad_compute(fs->dphi)) return -1;
  {
    KTimer t("binary_step");
    if (lb->prec == 64) {
      ldw::launch_binary_step_f64(g, p, q, (const double *) phi->d_data[phi->cur],
				  (double *) phi->d_data[1 - phi->cur], (const double *) hydro->d_u,
				  (double *) hydro->d_u2, (double *) hydro->d_rho,
				  (const double *) lb->d_g[lb->cur], (double *) lb->d_g[1 - lb->cur],
				  fz, (double *) pch->d_csum, g_stream);
    }
    else {
      ldw::launch_binary_step_f32(g, p, q, (const float *) phi->d_data[phi->cur],
				  (float *) phi->d_data[1 - phi->cur], (const float *) hydro->d_u,
				  (float *) hydro->d_u2, (float *) hydro->d_rho,
				  (const float *) lb->d_g[lb->cur], (float *) lb->d_g[1 - lb->cur],
				  fz, (float *) pch->d_csum, g_stream);
    }
  }
  CHECK_LAUNCH();
  phi->cur = 1 - phi->cur;
  phi->version++;
  void * t = hydro->d_u; hydro->d_u = hydro->d_u2; hydro->d_u2 = t;
  lb->cur = 1 - lb->cur;
  lb->streamed = 1;
  hydro->u_uniform = 0;
  hydro->force_uniform = 0;
  hydro->fd_derived = 1;
  hydro->fd_phi = phi;
  hydro->fd_buf = 1 - phi->cur;
  hydro->fd_params = p;
  hydro->fd_dphi = fs->dphi;
  g_derived = hydro;
  fs->dphi->mu_version = 0;
  return ch_subtract_drift(pch, phi);
}

/* The step with Lees-Edwards planes through the tiled single-pass kernel
 * (le_step.hip): the plane kernels fill the buffer planes and the raw
 * plane faces, two launches form the step (slab interiors with the force
 * correction sums, then the slab edges), and the LB plane conditions act
 * on the new post-collision distributions.  The same bits as
 * step_staged().  The force is stored; the gradient only next to the
 * planes (field_grad_memcpy forms the rest on demand). */
static int step_le_fused(lb_t * lb, hydro_t * hydro, field_t * phi, fe_symm_t * fs, pth_t * pth,
			 phi_ch_t * pch) {
  lees_edw_t * le = phi->le;
  field_grad_t * dphi = fs->dphi;
  const Geom g = lb->geo;
  const double ly = (double) g.ny;
  if (flush_force() || materialise_derived()) return -1;
  if (g_derived == hydro) { g_derived = nullptr; hydro->fd_derived = 0; }
  if (hydro->d_u2 == nullptr && dmalloc(&hydro->d_u2, 8*3*(size_t) g.nsite)) return -1;
  // the step starts from hydro_f_zero(0) and the velocity of the last
  // collision; it writes the force (zero plus the thermodynamic force) at
  // every site, so the zero is not written first
  const double fz[3] = {0.0, 0.0, 0.0};
  if (hydro_f_zero(hydro, fz)) return -1;
  hydro->force_uniform = 0;
  if (hydro_materialise(hydro)) return -1;
  ldw::LeShift shp, shu;
  if (le_buffer_shift(le, 0.0, 1, &shp) || le_buffer_shift(le, 1.0, 0, &shu)) return -1;
  const double dy = phys_time()*le->opts.uy;
  const double dyw = fmod(+dy, ly), dye = fmod(-dy, ly);
  int jw = 0, je = 0;
  if (le_floor(dyw, ly, &jw) || le_floor(dye, ly, &je)) return -1;
  int dj[2];
  double fr[2], duy[2];
  if (le_lb_shifts(le, ly, dj, fr, duy)) return -1;
  const FeParams p = fe_params_of(fs, pth->method);
  const LbParams q = lb_params_of(lb, nullptr);
  const double ra = 0.5/((double) g.ny*(double) g.nz);
  const double * phi0 = (const double *) phi->d_data[phi->cur];
  dphi->grad_version = phi->version;
  dphi->grad_buf = phi->cur;
  {
    KTimer t("le_planes");
    ldw::launch_le_phi_buffer(g, le->geo, shp, phi0, phi->d_phib, g_stream);
    ldw::launch_le_grad_planes(g, le->geo, grad_stencil(dphi), phi0, phi->d_phib, dphi->d_gr, dphi->d_dr,
			       dphi->d_gb, dphi->d_db, g_stream);
    ldw::launch_le_force_faces(g, le->geo, p, phi0, phi->d_phib, dphi->d_gr, dphi->d_dr, dphi->d_gb,
			       dphi->d_db, le->d_xface, g_stream);
    CHECK_LAUNCH();
    // the serial plane sums, the step's critical path, start as soon as
    // the x-faces are there and run on the second stream beside the rest
    HIP_CHECK(hipEventRecord(g_ev_fork, g_stream));
    HIP_CHECK(hipStreamWaitEvent(g_stream2, g_ev_fork, 0));
    ldw::launch_le_plane_sums(g, le->geo, le->d_xface, le->d_fcor, g_stream2);
    CHECK_LAUNCH();
    HIP_CHECK(hipEventRecord(g_ev_join, g_stream2));
    ldw::launch_le_u_buffer(g, le->geo, shu, le->opts.uy, (const double *) hydro->d_u, hydro->d_ub, g_stream);
    ldw::launch_le_ch_faces(g, le->geo, p, phi0, phi->d_phib, dphi->d_dr, dphi->d_db,
			    (const double *) hydro->d_u, hydro->d_ub, le->d_chface, g_stream);
  }
  CHECK_LAUNCH();
  {
    KTimer t("le_step");
    ldw::launch_le_step(g, p, q, le->geo, phi->d_phib, hydro->d_ub, le->d_fcor, le->d_chface,
			ra, jw, dyw - jw, je, dye - je, duy, le->d_sbuf, phi0, (double *) phi->d_data[1 - phi->cur],
			(const double *) hydro->d_u, (double *) hydro->d_u2, (double *) hydro->d_rho,
			(const double *) lb->d_g[lb->cur], (double *) lb->d_g[1 - lb->cur],
			(double *) hydro->d_force, (double *) pch->d_csum, g_stream, g_ev_join);
  }
  CHECK_LAUNCH();
  {
    // the edge launch formed the reprojections; displace and interpolate
    KTimer t("le_lb_bc");
    ldw::launch_le_lb_interp(g, le->geo, dj, fr, (double *) lb->d_g[1 - lb->cur], le->d_sbuf, g_stream);
  }
  CHECK_LAUNCH();
  phi->cur = 1 - phi->cur;
  phi->version++;
  void * tu = hydro->d_u; hydro->d_u = hydro->d_u2; hydro->d_u2 = tu;
  lb->cur = 1 - lb->cur;
  lb->streamed = 1;
  hydro->u_uniform = 0;
  hydro->force_uniform = 0;
  dphi->le_version = phi->version - 1;
  dphi->le_partial = 1;
  dphi->mu_version = 0;
  return ch_subtract_drift(pch, phi);
}

static int g_le_fused = -1;       // LDW_LE_FUSED=0 keeps the staged plane path

static int g_fused_step = -1;     // LDW_FUSED_STEP=0 selects the staged calls

extern "C" int ldw_step_binary(lb_t * lb, hydro_t * hydro, field_t * phi, field_grad_t * dphi,
			       fe_t * fe, pth_t * pth, phi_ch_t * pch, int nsteps) {
  if (g_fused_step < 0) {
    const char * e = getenv("LDW_FUSED_STEP");
    g_fused_step = (e && atoi(e) == 0) ? 0 : 1;
  }
  if (g_le_fused < 0) {
    const char * e = getenv("LDW_LE_FUSED");
    g_le_fused = (e && atoi(e) == 0) ? 0 : 1;
  }
  fe_symm_t * fs = as_symm(fe);
  // refused on the host before any launch: the conserve-1 compensation
  // buffer is sized for the precision phi_ch_t was created with, and every
  // path reads it with the precision of phi
  if (pch && pch->info.conserve && pch->prec != phi->prec) {
    FAIL("ldw_step_binary: phi_ch_t (%d-bit) and phi (%d-bit) differ in precision", pch->prec, phi->prec);
  }
  int order = 1;
  advection_order(&order);
  if (order > 4 && le_active(phi->le)) {
    FAIL("ldw_step_binary: advection order %d with Lees-Edwards planes (orders 1-4 are built)", order);
  }
  physics_t * phys = nullptr;
  physics_ref(&phys);
  for (int step = 0; step < nsteps; step++) {
    physics_control_next_step(phys);
    const bool fusable = g_fused_step && fs && pth && hydro
      && (pth->method == FE_FORCE_METHOD_STRESS_DIVERGENCE || pth->method == FE_FORCE_METHOD_PHI_GRADMU)
      && lb->streamed && !hydro->u_uniform && !hydro->fp_pending
      && lb->prec == hydro->prec && phi->prec == lb->prec && fs->phi == phi
      && pch && (pch->info.conserve == 0 || pch->prec == phi->prec) && order <= 4
      && !le_active(phi->le) && !le_active(hydro->le);
    // the planes' single-pass step: FP64, a force to form, the velocity of
    // a collision, and slabs at least 3 planes deep
    const bool le_fusable = g_fused_step && g_le_fused && fs && pth && hydro && pch
      && le_active(phi->le) && hydro->le == phi->le && fs->phi == phi
      && pth->method != FE_FORCE_METHOD_NO_FORCE && lb->streamed
      && lb->prec == 64 && phi->prec == 64 && hydro->prec == 64
      && (pch->info.conserve == 0 || pch->prec == 64)
      && phi->geo.nx/phi->le->nplanes >= 3;
    if (fusable) {
      if (step_fused(lb, hydro, phi, fs, pth, pch)) return -1;
    }
    else if (le_fusable) {
      if (step_le_fused(lb, hydro, phi, fs, pth, pch)) return -1;
    }
    else {
      if (step_staged(lb, hydro, phi, dphi, fe, pth, pch)) return -1;
    }
  }
  return 0;
}

extern "C" int ldw_step_single(lb_t * lb, hydro_t * hydro, int nsteps) {
  const double fzero[3] = {0.0, 0.0, 0.0};
  const double uzero[3] = {0.0, 0.0, 0.0};
  physics_t * phys = nullptr;
  physics_ref(&phys);
  for (int step = 0; step < nsteps; step++) {
    physics_control_next_step(phys);
    if (hydro_f_zero(hydro, fzero)) return -1;
    if (hydro_u_zero(hydro, uzero)) return -1;
    if (lb_collide(lb, hydro, nullptr, nullptr, nullptr, nullptr)) return -1;
    if (le_active(hydro->le) && lb_data_apply_le_boundary_conditions(lb, hydro->le)) return -1;
    if (lb_propagation(lb)) return -1;
  }
  return 0;
}

extern "C" int ldw_conserved_sums(lb_t * lb, field_t * phi, double out[2]) {
  double h[3];
  const Geom g = lb->geo;
  if (lb->prec == 64) {
    ldw::launch_klein_sum_f64((const double *) lb->d_g[lb->cur], 19*g.nsite, lb->d_work, lb->d_out, g_stream);
  }
  else {
    ldw::launch_klein_sum_planes_f32((const float *) lb->d_g[lb->cur], g.nsite, 19, lb->d_work, lb->d_out, g_stream);
  }
  CHECK_LAUNCH();
  HIP_CHECK(hipMemcpyAsync(h, lb->d_out, 3*sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HIP_CHECK(hipStreamSynchronize(g_stream));
  out[0] = h[0];
  if (lb->prec == 32) out[0] = (h[1] + lb_rho0()*(double) g.nsite) + h[2];
  out[1] = 0.0;
  if (phi) {
    if (phi->prec == 64) ldw::launch_klein_sum_f64((const double *) phi->d_data[phi->cur], g.nsite, lb->d_work, lb->d_out, g_stream);
    else ldw::launch_klein_sum_f32((const float *) phi->d_data[phi->cur], g.nsite, lb->d_work, lb->d_out, g_stream);
    CHECK_LAUNCH();
    HIP_CHECK(hipMemcpyAsync(h, lb->d_out, 3*sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
    out[1] = h[0];
  }
  return 0;
}

extern "C" int ldw_lb_device_f(lb_t * lb, void ** ptr, int * elsize, int * shifted) {
  *ptr = lb->d_g[lb->cur];
  *elsize = (lb->prec == 64) ? 8 : 4;
  *shifted = lb->streamed;
  return 0;
}

extern "C" int ldw_field_device(field_t * f, void ** ptr, int * elsize) {
  *ptr = f->d_data[f->cur];
  *elsize = (f->prec == 64) ? 8 : 4;
  return 0;
}

extern "C" int ldw_hydro_device(hydro_t * h, void ** rho, void ** u, void ** force, int * elsize) {
  if (hydro_materialise(h)) return -1;
  *rho = h->d_rho; *u = h->d_u; *force = h->d_force;
  *elsize = (h->prec == 64) ? 8 : 4;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Initial conditions and bulk copies                                       */
/* ------------------------------------------------------------------------ */

/* ns_uniform (noise.h:96-109) */
static inline unsigned int ns_uniform(unsigned int s[4]) {
  unsigned int b;
  s[0] = 69069u*s[0] + 1234567u;
  b = s[1] ^ (s[1] << 17);
  b ^= (b >> 13);
  s[1] = b ^ (b << 5);
  s[2] = 36969u*(s[2] & 0xffffu) + (s[2] >> 16);
  s[3] = 18000u*(s[3] & 0xffffu) + (s[3] >> 16);
  b = (s[2] << 16) + s[3];
  return (s[1] + (s[0] ^ b));
}

/* per-site state for global site (ig, jg, kg), 1-based (noise.c:278-341);
 * seeds <= 1 fall back to the default 13 (noise.c:87-94) */
static void site_state(int seed, int ig, int jg, int kg, unsigned int s[4]) {
  unsigned int sd = (seed > 1) ? (unsigned int) seed : 13u;
  unsigned int local[4] = {sd + (unsigned int) ig, 12953u + (unsigned int) jg,
			   712357u + (unsigned int) kg, 22383979u};
  s[0] = ns_uniform(local);
  s[1] = ns_uniform(local);
  s[2] = ns_uniform(local);
  s[3] = ns_uniform(local);
}

double ldw::api::spinodal_value(int seed, int ic, int jc, int kc, double phi0, double amp) {
  unsigned int s[4];
  site_state(seed, ic, jc, kc, s);
  double ran = (1.0/4294967295u)*ns_uniform(s);
  return phi0 + amp*(ran - 0.5);
}

extern "C" int field_phi_init_spinodal(field_t * phi, int seed, double phi0, double amp) {
  const Geom g = phi->geo;
  for (int ic = 1; ic <= g.nx; ic++)
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	unsigned int s[4];
	site_state(seed, ic, jc, kc, s);
	double ran = (1.0/4294967295u)*ns_uniform(s);
	phi->data[cs_index(phi->cs, ic, jc, kc)] = phi0 + amp*(ran - 0.5);
      }
  return field_memcpy(phi, tdpMemcpyHostToDevice);
}

/* field_phi_init.c:38-82 (one rank: no local offset) */
extern "C" int field_phi_init_drop(field_t * phi, double xi, double radius, double phistar,
				   int is_centred) {
  double ltot[3], lmin[3], centre[3];
  cs_ltot(phi->cs, ltot);
  cs_lmin(phi->cs, lmin);
  const double rxi0 = 1.0/xi;
  for (int a = 0; a < 3; a++) centre[a] = is_centred*lmin[a] + 0.5*ltot[a];
  const Geom g = phi->geo;
  for (int ic = 1; ic <= g.nx; ic++)
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	const double pos[3] = {1.0*ic - centre[0], 1.0*jc - centre[1], 1.0*kc - centre[2]};
	const double r = sqrt(pos[0]*pos[0] + pos[1]*pos[1] + pos[2]*pos[2]);
	phi->data[cs_index(phi->cs, ic, jc, kc)] = phistar*tanh(rxi0*(r - radius));
      }
  return field_memcpy(phi, tdpMemcpyHostToDevice);
}

/* ------------------------------------------------------------------------ */
/* Configuration files (lb_data.c:1716-1830, field.c:1633-1738,             */
/* hydro.c:941-967); the format is in io_format.h                           */
/* ------------------------------------------------------------------------ */

// count components of a host array (data[c*stride + cs_index]) as records
// in file order, rec[site*count + c]
static void gather_records(cs_t * cs, const double * data, int64_t stride, int count,
			   double * rec) {
  int64_t s = 0;
  for (int ic = 1; ic <= cs->ntotal[0]; ic++)
    for (int jc = 1; jc <= cs->ntotal[1]; jc++)
      for (int kc = 1; kc <= cs->ntotal[2]; kc++, s++) {
	const int index = cs_index(cs, ic, jc, kc);
	for (int c = 0; c < count; c++) rec[s*count + c] = data[c*stride + index];
      }
}

static void scatter_records(cs_t * cs, const double * rec, int64_t stride, int count,
			    double * data) {
  int64_t s = 0;
  for (int ic = 1; ic <= cs->ntotal[0]; ic++)
    for (int jc = 1; jc <= cs->ntotal[1]; jc++)
      for (int kc = 1; kc <= cs->ntotal[2]; kc++, s++) {
	const int index = cs_index(cs, ic, jc, kc);
	for (int c = 0; c < count; c++) data[c*stride + index] = rec[s*count + c];
      }
}

static ldw::io::Layout io_layout(cs_t * cs, int count, int per_line) {
  ldw::io::Layout l;
  for (int a = 0; a < 3; a++) l.ntotal[a] = cs->ntotal[a];
  l.count = count;
  l.per_line = per_line;
  return l;
}

// Metadata before the first data file of the object, then the data
static int io_write(const char * who, const char * stub, int timestep, const io_options_t & o,
		    int & meta_written, cs_t * cs, const double * data, int64_t stride,
		    int count, int per_line) {
  if (timestep < 0 || timestep >= 1000*1000*1000) FAIL("%s: time step %d out of range", who, timestep);
  const ldw::io::Layout l = io_layout(cs, count, per_line);
  std::string err;
  if (!meta_written) {
    double lmin[3];
    cs_lmin(cs, lmin);
    if (ldw::io::write_text(ldw::io::metadata_name(stub), ldw::io::metadata_json(l, o, lmin, cs->leopts), err)) {
      FAIL("%s: %s", who, err.c_str());
    }
    meta_written = 1;
  }
  std::vector<double> rec((size_t) cs->ntotal[0]*cs->ntotal[1]*cs->ntotal[2]*count);
  gather_records(cs, data, stride, count, rec.data());
  const std::string name = ldw::io::data_name(stub, timestep);
  if (ldw::io::write_data(name, l, o.iorformat, rec.data(), err)) FAIL("%s: %s", who, err.c_str());
  if (o.report) printf("MPIIO wrote to %s\n", name.c_str());
  return 0;
}

static int io_read(const char * who, const char * stub, int timestep, const io_options_t & o,
		   cs_t * cs, double * data, int64_t stride, int count, int per_line) {
  if (timestep < 0 || timestep >= 1000*1000*1000) FAIL("%s: time step %d out of range", who, timestep);
  const ldw::io::Layout l = io_layout(cs, count, per_line);
  std::vector<double> rec((size_t) cs->ntotal[0]*cs->ntotal[1]*cs->ntotal[2]*count);
  const std::string name = ldw::io::data_name(stub, timestep);
  std::string err;
  if (ldw::io::read_data(name, l, o.iorformat, rec.data(), err)) FAIL("%s: %s", who, err.c_str());
  scatter_records(cs, rec.data(), stride, count, data);
  if (o.report) printf("MPIIO read from %s\n", name.c_str());
  return 0;
}

extern "C" int lb_io_write(lb_t * lb, int timestep, io_event_t *) {
  if (!lb) FAIL("lb_io_write: null lb");
  if (lb_memcpy(lb, tdpMemcpyDeviceToHost)) return -1;
  return io_write("lb_io_write", "dist", timestep, lb->iodata.output, lb->meta_written, lb->cs,
		  lb->f, lb->nsite, 19*lb->ndist, lb->ndist);
}

extern "C" int lb_io_read(lb_t * lb, int timestep, io_event_t *) {
  if (!lb) FAIL("lb_io_read: null lb");
  if (io_read("lb_io_read", "dist", timestep, lb->iodata.input, lb->cs, lb->f, lb->nsite,
	      19*lb->ndist, lb->ndist)) return -1;
  return lb_memcpy(lb, tdpMemcpyHostToDevice);
}

extern "C" int field_io_write(field_t * f, int timestep, io_event_t *) {
  if (!f) FAIL("field_io_write: null field");
  if (field_memcpy(f, tdpMemcpyDeviceToHost)) return -1;
  return io_write("field_io_write", f->name, timestep, f->iodata.output, f->meta_written, f->cs,
		  f->data, f->nsites, f->nf, f->nf);
}

extern "C" int field_io_read(field_t * f, int timestep, io_event_t *) {
  if (!f) FAIL("field_io_read: null field");
  if (io_read("field_io_read", f->name, timestep, f->iodata.input, f->cs, f->data, f->nsites,
	      f->nf, f->nf)) return -1;
  return field_memcpy(f, tdpMemcpyHostToDevice);
}

extern "C" int hydro_io_write(hydro_t * h, int timestep, io_event_t *) {
  if (!h) FAIL("hydro_io_write: null hydro");
  if (hydro_memcpy(h, tdpMemcpyDeviceToHost)) return -1;
  if (io_write("hydro_io_write", "rho", timestep, h->rho_io.output, h->rho_meta_written, h->cs,
	       h->rho, h->nsite, 1, 1)) return -1;
  return io_write("hydro_io_write", "vel", timestep, h->u_io.output, h->u_meta_written, h->cs,
		  h->u, h->nsite, 3, 3);
}

// The individual rho and velocity outputs between configuration steps
// (ludwig.c:951-960, field_io_write() of hydro->rho and hydro->u)
extern "C" int hydro_io_write_rho(hydro_t * h, int timestep, io_event_t *) {
  if (!h) FAIL("hydro_io_write_rho: null hydro");
  if (hydro_memcpy(h, tdpMemcpyDeviceToHost)) return -1;
  return io_write("hydro_io_write_rho", "rho", timestep, h->rho_io.output, h->rho_meta_written, h->cs,
		  h->rho, h->nsite, 1, 1);
}

extern "C" int hydro_io_write_vel(hydro_t * h, int timestep, io_event_t *) {
  if (!h) FAIL("hydro_io_write_vel: null hydro");
  if (hydro_memcpy(h, tdpMemcpyDeviceToHost)) return -1;
  return io_write("hydro_io_write_vel", "vel", timestep, h->u_io.output, h->u_meta_written, h->cs,
		  h->u, h->nsite, 3, 3);
}

// rho and u from the host copy to the device; the force is left alone
extern "C" int hydro_io_read(hydro_t * h, int timestep, io_event_t *) {
  if (!h) FAIL("hydro_io_read: null hydro");
  if (flush_force() || materialise_derived()) return -1;
  if (io_read("hydro_io_read", "rho", timestep, h->rho_io.input, h->cs, h->rho, h->nsite, 1, 1)) return -1;
  if (io_read("hydro_io_read", "vel", timestep, h->u_io.input, h->cs, h->u, h->nsite, 3, 3)) return -1;
  const Geom g = h->geo;
  if (h->prec == 64) {
    std::vector<double> b(g.nsite), b3(3*g.nsite);
    pack_host<double>(h->cs, h->rho, 1, b.data());
    pack_host<double>(h->cs, h->u, 3, b3.data());
    HIP_CHECK(hipMemcpyAsync(h->d_rho, b.data(), 8*g.nsite, hipMemcpyHostToDevice, g_stream));
    HIP_CHECK(hipMemcpyAsync(h->d_u, b3.data(), 24*g.nsite, hipMemcpyHostToDevice, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
  }
  else {
    std::vector<float> b(g.nsite), b3(3*g.nsite);
    pack_host<float>(h->cs, h->rho, 1, b.data());
    pack_host<float>(h->cs, h->u, 3, b3.data());
    HIP_CHECK(hipMemcpyAsync(h->d_rho, b.data(), 4*g.nsite, hipMemcpyHostToDevice, g_stream));
    HIP_CHECK(hipMemcpyAsync(h->d_u, b3.data(), 12*g.nsite, hipMemcpyHostToDevice, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
  }
  h->u_uniform = 0;
  return 0;
}

extern "C" int ldw_lb_init_random_velocity(lb_t * lb, int seed, double amp) {
  const Geom g = lb->geo;
  std::vector<double> u(3*g.nsite);
  for (int ic = 1; ic <= g.nx; ic++)
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	unsigned int s[4];
	site_state(seed, ic, jc, kc, s);
	int64_t is = dsite(g, ic, jc, kc);
	for (int a = 0; a < 3; a++) {
	  double r = (1.0/4294967295u)*ns_uniform(s);
	  u[a*g.nsite + is] = amp*(2.0*r - 1.0);
	}
      }
  for (int a = 0; a < 3; a++) {
    double mean = 0.0;
    for (int64_t is = 0; is < g.nsite; is++) mean += u[a*g.nsite + is];
    mean = mean/g.nsite;
    for (int64_t is = 0; is < g.nsite; is++) u[a*g.nsite + is] -= mean;
  }
  for (int ic = 1; ic <= g.nx; ic++)
    for (int jc = 1; jc <= g.ny; jc++)
      for (int kc = 1; kc <= g.nz; kc++) {
	int64_t is = dsite(g, ic, jc, kc);
	double u0[3] = {u[is], u[g.nsite + is], u[2*g.nsite + is]};
	lb_1st_moment_equilib_set(lb, cs_index(lb->cs, ic, jc, kc), 1.0, u0);
      }
  return lb_memcpy(lb, tdpMemcpyHostToDevice);
}

extern "C" int ldw_lb_f_read(lb_t * lb, double * f) {
  const Geom g = lb->geo;
  const size_t n = (size_t) 19*g.nsite;
  const double rho0 = lb_rho0();
  std::vector<double> gd(n);
  if (lb->prec == 64) {
    HIP_CHECK(hipMemcpyAsync(gd.data(), lb->d_g[lb->cur], n*8, hipMemcpyDeviceToHost, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
  }
  else {
    std::vector<float> gf(n);
    HIP_CHECK(hipMemcpyAsync(gf.data(), lb->d_g[lb->cur], n*4, hipMemcpyDeviceToHost, g_stream));
    HIP_CHECK(hipStreamSynchronize(g_stream));
    for (int p = 0; p < 19; p++)
      for (int64_t s = 0; s < g.nsite; s++) gd[p*g.nsite + s] = (double) gf[p*g.nsite + s] + ldw::weight(p)*rho0;
  }
  if (!lb->streamed) {
    memcpy(f, gd.data(), n*sizeof(double));
    return 0;
  }
  for (int p = 0; p < 19; p++) {
    const int cx = ldw::CV.v[p][0], cy = ldw::CV.v[p][1], cz = ldw::CV.v[p][2];
    for (int ic = 1; ic <= g.nx; ic++)
      for (int jc = 1; jc <= g.ny; jc++)
	for (int kc = 1; kc <= g.nz; kc++)
	  f[p*g.nsite + dsite(g, ic, jc, kc)]
	    = gd[p*g.nsite + dsite(g, wrap1(ic - cx, g.nx), wrap1(jc - cy, g.ny), wrap1(kc - cz, g.nz))];
  }
  return 0;
}

template <typename T>
static int upload(void * dst, const double * src, int64_t n, double shift_per_plane_rho0,
		  int64_t plane) {
  std::vector<T> buf(n);
  for (int64_t i = 0; i < n; i++) {
    double v = src[i];
    if (shift_per_plane_rho0 != 0.0) v -= ldw::weight((int) (i/plane))*shift_per_plane_rho0;
    buf[i] = (T) v;
  }
  HIP_CHECK(hipMemcpyAsync(dst, buf.data(), n*sizeof(T), hipMemcpyHostToDevice, g_stream));
  HIP_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

template <typename T>
static int download(double * dst, const void * src, int64_t n) {
  std::vector<T> buf(n);
  HIP_CHECK(hipMemcpyAsync(buf.data(), src, n*sizeof(T), hipMemcpyDeviceToHost, g_stream));
  HIP_CHECK(hipStreamSynchronize(g_stream));
  for (int64_t i = 0; i < n; i++) dst[i] = (double) buf[i];
  return 0;
}

extern "C" int ldw_lb_f_write(lb_t * lb, const double * f) {
  const Geom g = lb->geo;
  int rc = (lb->prec == 64)
    ? upload<double>(lb->d_g[lb->cur], f, 19*g.nsite, 0.0, g.nsite)
    : upload<float>(lb->d_g[lb->cur], f, 19*g.nsite, lb_rho0(), g.nsite);
  lb->streamed = 0;
  return rc;
}

extern "C" int ldw_field_read(field_t * obj, double * data) {
  return (obj->prec == 64) ? download<double>(data, obj->d_data[obj->cur], obj->geo.nsite)
    : download<float>(data, obj->d_data[obj->cur], obj->geo.nsite);
}

extern "C" int ldw_field_write(field_t * obj, const double * data) {
  if (flush_force()) return -1;
  if (materialise_derived()) return -1;
  int rc = (obj->prec == 64) ? upload<double>(obj->d_data[obj->cur], data, obj->geo.nsite, 0.0, 1)
    : upload<float>(obj->d_data[obj->cur], data, obj->geo.nsite, 0.0, 1);
  obj->version++;
  return rc;
}

extern "C" int ldw_hydro_read(hydro_t * h, double * rho, double * u, double * force) {
  if (hydro_materialise(h)) return -1;
  const int64_t n = h->geo.nsite;
  int rc = 0;
  if (h->prec == 64) {
    if (rho) rc |= download<double>(rho, h->d_rho, n);
    if (u) rc |= download<double>(u, h->d_u, 3*n);
    if (force) rc |= download<double>(force, h->d_force, 3*n);
  }
  else {
    if (rho) rc |= download<float>(rho, h->d_rho, n);
    if (u) rc |= download<float>(u, h->d_u, 3*n);
    if (force) rc |= download<float>(force, h->d_force, 3*n);
  }
  return rc;
}

extern "C" int ldw_hydro_write(hydro_t * h, const double * rho, const double * u,
			       const double * force) {
  if (hydro_materialise(h)) return -1;
  const int64_t n = h->geo.nsite;
  int rc = 0;
  if (h->prec == 64) {
    if (rho) rc |= upload<double>(h->d_rho, rho, n, 0.0, 1);
    if (u) rc |= upload<double>(h->d_u, u, 3*n, 0.0, 1);
    if (force) rc |= upload<double>(h->d_force, force, 3*n, 0.0, 1);
  }
  else {
    if (rho) rc |= upload<float>(h->d_rho, rho, n, 0.0, 1);
    if (u) rc |= upload<float>(h->d_u, u, 3*n, 0.0, 1);
    if (force) rc |= upload<float>(h->d_force, force, 3*n, 0.0, 1);
  }
  return rc;
}

/* Bandwidth calibration: a plain copy of nbytes between two device buffers,
 * timed with HIP events over reps launches; *ms is the mean per launch. */
extern "C" int ldw_copy_calibration(long long nbytes, int reps, double * ms) {
  if (device_init()) return -1;
  if (nbytes <= 0 || nbytes % 32) FAIL("ldw_copy_calibration: nbytes must be a positive multiple of 32");
  void * a = nullptr, * b = nullptr;
  HIP_CHECK(hipMalloc(&a, (size_t) nbytes));
  HIP_CHECK(hipMalloc(&b, (size_t) nbytes));
  HIP_CHECK(hipMemsetAsync(a, 0, (size_t) nbytes, g_stream));
  ldw::launch_copy(a, b, nbytes, g_stream);           // warm-up
  CHECK_LAUNCH();
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, g_stream));
  for (int r = 0; r < reps; r++) ldw::launch_copy(a, b, nbytes, g_stream);
  HIP_CHECK(hipEventRecord(e1, g_stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
  *ms = (double) t/reps;
  (void) hipEventDestroy(e0);
  (void) hipEventDestroy(e1);
  (void) hipFree(a);
  (void) hipFree(b);
  return 0;
}
