// ensemble_api.cpp -- host side of the replica ensemble (ldw_ensemble_*,
// include/ludwig_hip.h): shared device pools, work-item lists, the
// autotuner and the per-step HIP graphs.
//
// Pools: all replicas' 19 distribution planes back to back in each of two
// f pools, and likewise phi (two pools), u (two pools, 3 planes per
// replica) and rho (one pool).  Step t reads pool t % 2 and writes the
// other, for every field at once, so one graph per parity replays a step.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ludwig_hip.h"
#include "api_internal.h"
#include "d3q19.h"
#include "kernels.h"

using ldw::EnsItem;
using ldw::EnsRep;
using ldw::FeParams;
using ldw::Geom;

struct ldw_ensemble_s {
  int nrep = 0;
  int prec = 64;
  int order = 1;
  int method = 0;             // 0 stress divergence, 1 phi grad mu
  size_t es = 8;
  std::vector<ldw_replica_t> spec;
  std::vector<EnsRep> reps;
  int64_t nsite_all = 0;
  void * d_f[2] = {nullptr, nullptr};
  void * d_phi[2] = {nullptr, nullptr};
  void * d_u[2] = {nullptr, nullptr};
  void * d_rho = nullptr;
  EnsRep * d_reps = nullptr;
  EnsItem * d_items = nullptr;
  int nitems = 0;
  int items_cap = 0;
  unsigned int * d_ctr = nullptr;
  double * d_work = nullptr;
  double * d_out = nullptr;
  int shape = -1;             // configuration; shape < 0: not chosen yet
  int xplanes = 0;
  int nblocks = 0;
  double tuned_ms = 0.0;
  std::vector<double> tune_rows;   // tz, ty, xplanes, blocks, ms per trial
  int use_graph = 1;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int cur = 0;
  long long t = 0;
};

namespace {

int shape_of(int tz, int ty) {
  for (int s = 0; s < ldw::ENS_NSHAPE; s++)
    if (ldw::ens_shape_tz(s) == tz && ldw::ens_shape_ty(s) == ty) return s;
  return -1;
}

int resident_blocks(ldw_ensemble_t * e, int shape) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const int per = ldw::ens_blocks_per_cu(shape, e->prec, e->order, e->method);
  return per*ncu;
}

// Work items of every replica for tile shape `shape` and x chunks of
// `xplanes` planes, longest first (a persistent grid that takes them in
// this order ends with short items).
void build_items(const ldw_ensemble_t * e, int shape, int xplanes, std::vector<EnsItem> & items) {
  const int tz = ldw::ens_shape_tz(shape), ty = ldw::ens_shape_ty(shape);
  items.clear();
  for (int r = 0; r < e->nrep; r++) {
    const Geom & g = e->reps[r].g;
    for (int xb = 0; xb < g.nx; xb += xplanes)
      for (int y0 = 0; y0 < g.ny; y0 += ty)
	for (int z0 = 0; z0 < g.nz; z0 += tz) {
	  EnsItem w;
	  memset(&w, 0, sizeof(w));
	  w.rep = r; w.y0 = y0; w.z0 = z0; w.xb = xb;
	  w.xe = std::min(xb + xplanes, g.nx);
	  items.push_back(w);
	}
  }
  std::stable_sort(items.begin(), items.end(), [](const EnsItem & a, const EnsItem & b) {
    return (a.xe - a.xb) > (b.xe - b.xb);
  });
}

void drop_graphs(ldw_ensemble_t * e) {
  for (int p = 0; p < 2; p++) {
    if (e->gexec[p]) (void) hipGraphExecDestroy(e->gexec[p]);
    e->gexec[p] = nullptr;
  }
}

int upload_items(ldw_ensemble_t * e, const std::vector<EnsItem> & items) {
  hipStream_t st = ldw::api::stream();
  if ((int) items.size() > e->items_cap) {
    LDW_HIP_CHECK(hipStreamSynchronize(st));
    if (e->d_items) LDW_HIP_CHECK(hipFree(e->d_items));
    e->d_items = nullptr;
    e->items_cap = 0;
    LDW_HIP_CHECK(hipMalloc((void **) &e->d_items, items.size()*sizeof(EnsItem)));
    e->items_cap = (int) items.size();
  }
  LDW_HIP_CHECK(hipMemcpyAsync(e->d_items, items.data(), items.size()*sizeof(EnsItem),
			       hipMemcpyHostToDevice, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  e->nitems = (int) items.size();
  drop_graphs(e);
  return 0;
}

// One step from pools `par` to 1 - par with the given launch shape
hipError_t launch_step(ldw_ensemble_t * e, int par, int shape, int nblocks) {
  hipStream_t st = ldw::api::stream();
  const int q = 1 - par;
  if (e->prec == 64) {
    return ldw::launch_ensemble_step_f64(shape, e->order, e->method, nblocks, e->d_reps, e->d_items,
					 e->nitems, e->d_ctr, (const double *) e->d_phi[par],
					 (double *) e->d_phi[q], (const double *) e->d_u[par],
					 (double *) e->d_u[q], (double *) e->d_rho,
					 (const double *) e->d_f[par], (double *) e->d_f[q], st);
  }
  return ldw::launch_ensemble_step_f32(shape, e->order, e->method, nblocks, e->d_reps, e->d_items,
				       e->nitems, e->d_ctr, (const float *) e->d_phi[par],
				       (float *) e->d_phi[q], (const float *) e->d_u[par],
				       (float *) e->d_u[q], (float *) e->d_rho,
				       (const float *) e->d_f[par], (float *) e->d_f[q], st);
}

int capture_graphs(ldw_ensemble_t * e) {
  hipStream_t st = ldw::api::stream();
  drop_graphs(e);
  for (int p = 0; p < 2; p++) {
    hipGraph_t graph = nullptr;
    LDW_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const hipError_t le = launch_step(e, p, e->shape, e->nblocks);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (le != hipSuccess) LDW_FAIL("ensemble step capture: %s", hipGetErrorString(le));
    if (ce != hipSuccess) LDW_FAIL("hipStreamEndCapture: %s", hipGetErrorString(ce));
    ce = hipGraphInstantiate(&e->gexec[p], graph, nullptr, nullptr, 0);
    (void) hipGraphDestroy(graph);
    if (ce != hipSuccess) LDW_FAIL("hipGraphInstantiate: %s", hipGetErrorString(ce));
  }
  return 0;
}

int apply_config(ldw_ensemble_t * e, int shape, int xplanes, int nblocks) {
  std::vector<EnsItem> items;
  build_items(e, shape, xplanes, items);
  if (upload_items(e, items)) return -1;
  e->shape = shape;
  e->xplanes = xplanes;
  e->nblocks = nblocks;
  return 0;
}

template <typename T>
int upload_scalar(void * pool, int64_t off, const std::vector<double> & v) {
  std::vector<T> buf(v.size());
  for (size_t i = 0; i < v.size(); i++) buf[i] = (T) v[i];
  hipStream_t st = ldw::api::stream();
  LDW_HIP_CHECK(hipMemcpyAsync((T *) pool + off, buf.data(), buf.size()*sizeof(T),
			       hipMemcpyHostToDevice, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  return 0;
}

template <typename T>
int download(const void * pool, int64_t off, int64_t n, double * out) {
  std::vector<T> buf(n);
  hipStream_t st = ldw::api::stream();
  LDW_HIP_CHECK(hipMemcpyAsync(buf.data(), (const T *) pool + off, n*sizeof(T),
			       hipMemcpyDeviceToHost, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  for (int64_t i = 0; i < n; i++) out[i] = (double) buf[i];
  return 0;
}

int check_rep(ldw_ensemble_t * e, int r) {
  if (e == nullptr) LDW_FAIL("ensemble: null handle");
  if (r < 0 || r >= e->nrep) LDW_FAIL("ensemble: replica %d out of range [0, %d)", r, e->nrep);
  return 0;
}

}  // namespace

extern "C" int ldw_ensemble_create(int nrep, const ldw_replica_t * spec, int precision,
				   int adv_order, int force_method, ldw_ensemble_t ** pens) {
  if (pens == nullptr || spec == nullptr) LDW_FAIL("ldw_ensemble_create: null argument");
  *pens = nullptr;
  if (nrep < 1) LDW_FAIL("ldw_ensemble_create: need at least one replica, got %d", nrep);
  if (precision != 32 && precision != 64) LDW_FAIL("ldw_ensemble_create: precision %d (32 or 64)", precision);
  if (adv_order < 1 || adv_order > 3) LDW_FAIL("ldw_ensemble_create: advection order %d (1..3)", adv_order);
  if (force_method != FE_FORCE_METHOD_STRESS_DIVERGENCE && force_method != FE_FORCE_METHOD_PHI_GRADMU)
    LDW_FAIL("ldw_ensemble_create: force method %d (stress_divergence or phi_gradmu)", force_method);
  for (int r = 0; r < nrep; r++) {
    const ldw_replica_t & s = spec[r];
    if (s.nx < 1 || s.ny < 1 || s.nz < 1) LDW_FAIL("replica %d: bad size %d x %d x %d", r, s.nx, s.ny, s.nz);
    if (!(s.rho0 > 0.0) || !(s.eta_shear > 0.0) || !(s.eta_bulk > 0.0))
      LDW_FAIL("replica %d: rho0, eta_shear and eta_bulk must be positive", r);
  }
  if (ldw::api::device_init()) return -1;

  ldw_ensemble_t * e = new ldw_ensemble_t;
  e->nrep = nrep;
  e->prec = precision;
  e->es = (precision == 64) ? 8 : 4;
  e->order = adv_order;
  e->method = (force_method == FE_FORCE_METHOD_PHI_GRADMU) ? 1 : 0;
  e->spec.assign(spec, spec + nrep);
  e->reps.resize(nrep);
  int64_t off = 0;
  for (int r = 0; r < nrep; r++) {
    const ldw_replica_t & s = spec[r];
    EnsRep & R = e->reps[r];
    memset(&R, 0, sizeof(R));
    R.g.nx = s.nx; R.g.ny = s.ny; R.g.nz = s.nz;
    R.g.nsite = (int64_t) s.nx*s.ny*s.nz;
    R.off_s = off;
    R.off_f = 19*off;
    R.fe.a = s.a; R.fe.b = s.b; R.fe.kappa = s.kappa; R.fe.mobility = s.mobility;
    R.fe.order = adv_order;
    R.fe.force_method = e->method;
    R.fe.stencil = 7;
    R.prm = ldw::api::lb_params(s.rho0, s.eta_shear, s.eta_bulk, LB_RELAXATION_M10, nullptr);
    off += R.g.nsite;
  }
  e->nsite_all = off;

  hipStream_t st = ldw::api::stream();
  const int64_t n = e->nsite_all;
  const int nb = ldw::sum_blocks(19*n);
  hipError_t err = hipSuccess;
  for (int p = 0; p < 2 && err == hipSuccess; p++) {
    err = hipMalloc(&e->d_f[p], 19*n*e->es);
    if (err == hipSuccess) err = hipMalloc(&e->d_phi[p], n*e->es);
    if (err == hipSuccess) err = hipMalloc(&e->d_u[p], 3*n*e->es);
  }
  if (err == hipSuccess) err = hipMalloc(&e->d_rho, n*e->es);
  if (err == hipSuccess) err = hipMalloc((void **) &e->d_reps, nrep*sizeof(EnsRep));
  if (err == hipSuccess) err = hipMalloc((void **) &e->d_ctr, 2*sizeof(unsigned int));
  if (err == hipSuccess) err = hipMalloc((void **) &e->d_work, 3*(size_t) nb*sizeof(double));
  if (err == hipSuccess) err = hipMalloc((void **) &e->d_out, 3*sizeof(double));
  if (err != hipSuccess) {
    ldw_ensemble_free(e);
    LDW_FAIL("ldw_ensemble_create: device allocation of %lld sites failed: %s",
	     (long long) n, hipGetErrorString(err));
  }
  LDW_HIP_CHECK(hipMemcpyAsync(e->d_reps, e->reps.data(), nrep*sizeof(EnsRep), hipMemcpyHostToDevice, st));
  LDW_HIP_CHECK(hipMemsetAsync(e->d_ctr, 0, 2*sizeof(unsigned int), st));

  // initial state: spinodal phi, rest distributions, u = 0, rho = rho0
  for (int r = 0; r < nrep; r++) {
    const ldw_replica_t & s = spec[r];
    const EnsRep & R = e->reps[r];
    const Geom & g = R.g;
    std::vector<double> phi(g.nsite);
    for (int ic = 1; ic <= g.nx; ic++)
      for (int jc = 1; jc <= g.ny; jc++)
	for (int kc = 1; kc <= g.nz; kc++)
	  phi[((int64_t) (ic - 1)*g.ny + (jc - 1))*g.nz + (kc - 1)]
	    = ldw::api::spinodal_value(s.seed, ic, jc, kc, s.phi0, s.noise);
    int rc = (e->prec == 64) ? upload_scalar<double>(e->d_phi[0], R.off_s, phi)
			     : upload_scalar<float>(e->d_phi[0], R.off_s, phi);
    if (rc) return -1;
    const double u0[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < 19; p++) {
      const double fp = ldw::api::equilibrium_f(p, s.rho0, u0);
      if (e->prec == 64) ldw::launch_fill_f64((double *) e->d_f[0] + R.off_f + p*g.nsite, g.nsite, fp, st);
      else ldw::launch_fill_f32((float *) e->d_f[0] + R.off_f + p*g.nsite, g.nsite,
				(double) (float) (fp - ldw::weight(p)*s.rho0), st);
    }
    if (e->prec == 64) ldw::launch_fill_f64((double *) e->d_rho + R.off_s, g.nsite, s.rho0, st);
    else ldw::launch_fill_f32((float *) e->d_rho + R.off_s, g.nsite, s.rho0, st);
  }
  if (e->prec == 64) ldw::launch_fill_f64((double *) e->d_u[0], 3*n, 0.0, st);
  else ldw::launch_fill_f32((float *) e->d_u[0], 3*n, 0.0, st);
  LDW_HIP_CHECK(hipGetLastError());
  LDW_HIP_CHECK(hipStreamSynchronize(st));

  const char * cfg = getenv("LDW_ENS_GRAPH");
  if (cfg && atoi(cfg) == 0) e->use_graph = 0;
  *pens = e;
  return 0;
}

extern "C" int ldw_ensemble_free(ldw_ensemble_t * e) {
  if (e == nullptr) return 0;
  hipStream_t st = ldw::api::stream();
  if (st) (void) hipStreamSynchronize(st);
  drop_graphs(e);
  for (int p = 0; p < 2; p++) {
    if (e->d_f[p]) (void) hipFree(e->d_f[p]);
    if (e->d_phi[p]) (void) hipFree(e->d_phi[p]);
    if (e->d_u[p]) (void) hipFree(e->d_u[p]);
  }
  if (e->d_rho) (void) hipFree(e->d_rho);
  if (e->d_reps) (void) hipFree(e->d_reps);
  if (e->d_items) (void) hipFree(e->d_items);
  if (e->d_ctr) (void) hipFree(e->d_ctr);
  if (e->d_work) (void) hipFree(e->d_work);
  if (e->d_out) (void) hipFree(e->d_out);
  delete e;
  return 0;
}

extern "C" int ldw_ensemble_config_set(ldw_ensemble_t * e, int tz, int ty, int xplanes, int blocks) {
  if (e == nullptr) LDW_FAIL("ldw_ensemble_config_set: null handle");
  const int shape = shape_of(tz, ty);
  if (shape < 0) LDW_FAIL("ldw_ensemble_config_set: tile %d x %d is not built (64x4, 32x8, 16x16, 64x8, 128x4)", tz, ty);
  if (xplanes < 1) LDW_FAIL("ldw_ensemble_config_set: xplanes %d < 1", xplanes);
  const int resident = resident_blocks(e, shape);
  if (resident < 1) LDW_FAIL("ldw_ensemble_config_set: occupancy query failed");
  if (blocks < 0) LDW_FAIL("ldw_ensemble_config_set: blocks %d < 0", blocks);
  if (blocks == 0) blocks = resident;
  if (apply_config(e, shape, xplanes, blocks)) return -1;
  e->tuned_ms = 0.0;
  return 0;
}

extern "C" int ldw_ensemble_tune(ldw_ensemble_t * e, int reps) {
  if (e == nullptr) LDW_FAIL("ldw_ensemble_tune: null handle");
  if (reps < 1) reps = 1;
  hipStream_t st = ldw::api::stream();
  int maxn = 1;
  for (const EnsRep & R : e->reps) maxn = std::max(maxn, R.g.nx);
  std::vector<int> xp = {4, 8, 16, 32};
  xp.push_back(maxn);
  std::sort(xp.begin(), xp.end());
  xp.erase(std::unique(xp.begin(), xp.end()), xp.end());

  hipEvent_t t0, t1;
  LDW_HIP_CHECK(hipEventCreate(&t0));
  LDW_HIP_CHECK(hipEventCreate(&t1));
  double best = -1.0;
  int bs = 0, bx = 0, bb = 0;
  e->tune_rows.clear();
  for (int shape = 0; shape < ldw::ENS_NSHAPE; shape++) {
    const int nb = resident_blocks(e, shape);
    if (nb < 1) continue;
    for (int x : xp) {
      if (apply_config(e, shape, x, nb)) return -1;
      // the step reads pools `cur` and writes the others: repeating it
      // leaves the state unchanged
      LDW_HIP_CHECK(launch_step(e, e->cur, shape, nb));
      LDW_HIP_CHECK(hipEventRecord(t0, st));
      for (int k = 0; k < reps; k++) LDW_HIP_CHECK(launch_step(e, e->cur, shape, nb));
      LDW_HIP_CHECK(hipEventRecord(t1, st));
      LDW_HIP_CHECK(hipEventSynchronize(t1));
      float ms = 0.0f;
      LDW_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
      const double per = (double) ms/reps;
      const double row[5] = {(double) ldw::ens_shape_tz(shape), (double) ldw::ens_shape_ty(shape),
			     (double) x, (double) nb, per};
      e->tune_rows.insert(e->tune_rows.end(), row, row + 5);
      if (best < 0.0 || per < best) { best = per; bs = shape; bx = x; bb = nb; }
    }
  }
  (void) hipEventDestroy(t0);
  (void) hipEventDestroy(t1);
  if (best < 0.0) LDW_FAIL("ldw_ensemble_tune: no configuration could be launched");
  if (apply_config(e, bs, bx, bb)) return -1;
  e->tuned_ms = best;
  return 0;
}

extern "C" int ldw_ensemble_tune_results(ldw_ensemble_t * e, int maxrows, double * rows) {
  if (e == nullptr) LDW_FAIL("ldw_ensemble_tune_results: null handle");
  const int n = (int) (e->tune_rows.size()/5);
  if (rows) {
    for (int i = 0; i < std::min(n, maxrows); i++)
      for (int k = 0; k < 5; k++) rows[5*i + k] = e->tune_rows[5*i + k];
  }
  return n;
}

extern "C" int ldw_ensemble_config(ldw_ensemble_t * e, int cfg[4], double * ms) {
  if (e == nullptr) LDW_FAIL("ldw_ensemble_config: null handle");
  cfg[0] = e->shape >= 0 ? ldw::ens_shape_tz(e->shape) : 0;
  cfg[1] = e->shape >= 0 ? ldw::ens_shape_ty(e->shape) : 0;
  cfg[2] = e->xplanes;
  cfg[3] = e->nblocks;
  if (ms) *ms = e->tuned_ms;
  return 0;
}

extern "C" int ldw_ensemble_graph_set(ldw_ensemble_t * e, int on) {
  if (e == nullptr) LDW_FAIL("ldw_ensemble_graph_set: null handle");
  e->use_graph = on ? 1 : 0;
  if (!e->use_graph) drop_graphs(e);
  return 0;
}

extern "C" int ldw_ensemble_step(ldw_ensemble_t * e, int nsteps) {
  if (e == nullptr) LDW_FAIL("ldw_ensemble_step: null handle");
  if (nsteps < 0) LDW_FAIL("ldw_ensemble_step: nsteps %d < 0", nsteps);
  if (e->shape < 0 && ldw_ensemble_tune(e, 3)) return -1;
  hipStream_t st = ldw::api::stream();
  if (e->use_graph && e->gexec[0] == nullptr && capture_graphs(e)) return -1;
  for (int k = 0; k < nsteps; k++) {
    if (e->use_graph) LDW_HIP_CHECK(hipGraphLaunch(e->gexec[e->cur], st));
    else LDW_HIP_CHECK(launch_step(e, e->cur, e->shape, e->nblocks));
    e->cur = 1 - e->cur;
    e->t += 1;
  }
  return 0;
}

extern "C" long long ldw_ensemble_time(ldw_ensemble_t * e) { return e ? e->t : -1; }

extern "C" int ldw_ensemble_snapshot(ldw_ensemble_t * e, int r, double * phi, double * rho, double * u) {
  if (check_rep(e, r)) return -1;
  const EnsRep & R = e->reps[r];
  const int64_t n = R.g.nsite;
  int rc = 0;
  if (e->prec == 64) {
    if (phi) rc |= download<double>(e->d_phi[e->cur], R.off_s, n, phi);
    if (rho) rc |= download<double>(e->d_rho, R.off_s, n, rho);
    if (u) rc |= download<double>(e->d_u[e->cur], 3*R.off_s, 3*n, u);
  }
  else {
    if (phi) rc |= download<float>(e->d_phi[e->cur], R.off_s, n, phi);
    if (rho) rc |= download<float>(e->d_rho, R.off_s, n, rho);
    if (u) rc |= download<float>(e->d_u[e->cur], 3*R.off_s, 3*n, u);
  }
  return rc ? -1 : 0;
}

extern "C" int ldw_ensemble_f_read(ldw_ensemble_t * e, int r, double * f) {
  if (check_rep(e, r)) return -1;
  const EnsRep & R = e->reps[r];
  const Geom & g = R.g;
  std::vector<double> gd(19*g.nsite);
  int rc = (e->prec == 64) ? download<double>(e->d_f[e->cur], R.off_f, 19*g.nsite, gd.data())
			   : download<float>(e->d_f[e->cur], R.off_f, 19*g.nsite, gd.data());
  if (rc) return -1;
  const double rho0 = e->spec[r].rho0;
  // the pools hold post-collision values: after propagation f(x)[p] = g(x - c_p)[p]
  for (int p = 0; p < 19; p++) {
    const int cx = ldw::CV.v[p][0], cy = ldw::CV.v[p][1], cz = ldw::CV.v[p][2];
    const double shift = (e->prec == 32) ? ldw::weight(p)*rho0 : 0.0;
    for (int x = 0; x < g.nx; x++)
      for (int y = 0; y < g.ny; y++)
	for (int z = 0; z < g.nz; z++) {
	  const int xs = (x - cx + g.nx) % g.nx, ys = (y - cy + g.ny) % g.ny, zs = (z - cz + g.nz) % g.nz;
	  const double v = gd[p*g.nsite + ((int64_t) xs*g.ny + ys)*g.nz + zs];
	  f[p*g.nsite + ((int64_t) x*g.ny + y)*g.nz + z] = (e->prec == 32) ? v + shift : v;
	}
  }
  return 0;
}

extern "C" int ldw_ensemble_snapshot_device(ldw_ensemble_t * e, int r, float * dst) {
  if (check_rep(e, r)) return -1;
  if (dst == nullptr) LDW_FAIL("ldw_ensemble_snapshot_device: null destination");
  const EnsRep & R = e->reps[r];
  const int64_t n = R.g.nsite;
  hipStream_t st = ldw::api::stream();
  const char * phi = (const char *) e->d_phi[e->cur] + R.off_s*e->es;
  const char * u = (const char *) e->d_u[e->cur] + 3*R.off_s*e->es;
  if (e->prec == 32) {
    LDW_HIP_CHECK(hipMemcpyAsync(dst, phi, n*4, hipMemcpyDeviceToDevice, st));
    LDW_HIP_CHECK(hipMemcpyAsync(dst + n, u, 3*n*4, hipMemcpyDeviceToDevice, st));
  }
  else {
    ldw::launch_to_f32(phi, 64, dst, n, st);
    ldw::launch_to_f32(u, 64, dst + n, 3*n, st);
    LDW_HIP_CHECK(hipGetLastError());
  }
  // ordered on the library stream like every other call: no host wait
  // (a jump snapshots every replica in turn)
  return 0;
}

extern "C" int ldw_ensemble_load_device(ldw_ensemble_t * e, int r, const float * src) {
  if (check_rep(e, r)) return -1;
  if (src == nullptr) LDW_FAIL("ldw_ensemble_load_device: null source");
  const EnsRep & R = e->reps[r];
  const int64_t n = R.g.nsite;
  hipStream_t st = ldw::api::stream();
  const int c = e->cur;
  ldw::launch_ensemble_set_state(src, n, e->prec, (char *) e->d_phi[c] + R.off_s*e->es,
				 (char *) e->d_u[c] + 3*R.off_s*e->es, (char *) e->d_f[c] + R.off_f*e->es,
				 (const char *) e->d_rho + R.off_s*e->es, e->spec[r].rho0, st);
  LDW_HIP_CHECK(hipGetLastError());
  return 0;
}

extern "C" int ldw_ensemble_conserved_sums(ldw_ensemble_t * e, int r, double out[2]) {
  if (check_rep(e, r)) return -1;
  const EnsRep & R = e->reps[r];
  const int64_t n = R.g.nsite;
  hipStream_t st = ldw::api::stream();
  double h[3];
  if (e->prec == 64) ldw::launch_klein_sum_f64((const double *) e->d_f[e->cur] + R.off_f, 19*n, e->d_work, e->d_out, st);
  else ldw::launch_klein_sum_planes_f32((const float *) e->d_f[e->cur] + R.off_f, n, 19, e->d_work, e->d_out, st);
  LDW_HIP_CHECK(hipGetLastError());
  LDW_HIP_CHECK(hipMemcpyAsync(h, e->d_out, 3*sizeof(double), hipMemcpyDeviceToHost, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  out[0] = (e->prec == 64) ? h[0] : (h[1] + e->spec[r].rho0*(double) n) + h[2];
  if (e->prec == 64) ldw::launch_klein_sum_f64((const double *) e->d_phi[e->cur] + R.off_s, n, e->d_work, e->d_out, st);
  else ldw::launch_klein_sum_f32((const float *) e->d_phi[e->cur] + R.off_s, n, e->d_work, e->d_out, st);
  LDW_HIP_CHECK(hipGetLastError());
  LDW_HIP_CHECK(hipMemcpyAsync(h, e->d_out, 3*sizeof(double), hipMemcpyDeviceToHost, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  out[1] = h[0];
  return 0;
}
