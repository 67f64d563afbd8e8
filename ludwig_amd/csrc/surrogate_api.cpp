// surrogate_api.cpp -- C ABI of the surrogate network's device kernels
// (ldw_conv3d_*, ldw_nn_*; include/ludwig_hip.h).  All arrays are device
// pointers (from ldw_device_alloc); BF16 arrays are raw 16-bit patterns.
// Work is ordered on the library's stream; a call that returns a host
// value (ldw_nn_mse with a loss pointer) completes before it returns.
// Scratch memory for partial sums is kept by the library and grown on
// demand.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/ludwig_hip.h"
#include "api_internal.h"
#include "kernels.h"

namespace {

struct Scratch {
  void * p = nullptr;
  size_t bytes = 0;
};

Scratch g_wgrad, g_bias, g_err, g_sum, g_sumsq, g_restore;

int scratch_buffer(Scratch & s, size_t bytes, void ** out) {
  if (s.bytes < bytes) {
    hipStream_t st = ldw::api::stream();
    LDW_HIP_CHECK(hipStreamSynchronize(st));
    if (s.p) LDW_HIP_CHECK(hipFree(s.p));
    s.p = nullptr;
    s.bytes = 0;
    LDW_HIP_CHECK(hipMalloc(&s.p, bytes));
    s.bytes = bytes;
  }
  *out = s.p;
  return 0;
}

int check_conv(int batch, int nx, int ny, int nz, int cin, int cout) {
  if (batch < 1 || nx < 1 || ny < 1 || nz < 1) LDW_FAIL("conv3d: bad sizes %d x %d x %d x %d", batch, nx, ny, nz);
  if (cin < 16 || cout < 16 || cin % 16 || cout % 16)
    LDW_FAIL("conv3d: channel counts must be multiples of 16 (cin %d, cout %d)", cin, cout);
  return 0;
}

// The scratch of ldw_nn_restore_sums_noalloc: the block partial sums of
// up to NOALLOC_MAX_PLANES planes in a fixed device array (64 KiB), which
// each GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel: it is a host-side stand-in,
// not a device address.  The device address is looked up once per device
// at library set-up (noalloc_scratch_init) and kept in g_noalloc_ptr.
constexpr int NOALLOC_MAX_PLANES = 64;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[2*NOALLOC_MAX_PLANES*64];
double * g_noalloc_ptr[NOALLOC_MAX_DEVICES];

int current_device(int * dev) {
  LDW_HIP_CHECK(hipGetDevice(dev));
  if (*dev < 0 || *dev >= NOALLOC_MAX_DEVICES) LDW_FAIL("device %d: at most %d devices", *dev, NOALLOC_MAX_DEVICES);
  return 0;
}

}  // namespace

// Called by the library's init (device_init) with its device current.
int ldw::api::noalloc_scratch_init(void) {
  int dev;
  if (current_device(&dev)) return -1;
  if (g_noalloc_ptr[dev] == nullptr) {
    void * p = nullptr;
    LDW_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) LDW_FAIL("hipGetSymbolAddress: no device address for the restore-sums scratch");
    g_noalloc_ptr[dev] = (double *) p;
  }
  return 0;
}

#define READY() do { if (ldw::api::device_init()) return -1; } while (0)
#define KCHECK(call) LDW_HIP_CHECK(call)

extern "C" int ldw_conv3d_fwd(const void * x, const void * w, const float * bias, float * y,
			      int batch, int nx, int ny, int nz, int cin, int cout, int accumulate) {
  READY();
  if (check_conv(batch, nx, ny, nz, cin, cout)) return -1;
  if (y == nullptr) LDW_FAIL("ldw_conv3d_fwd: null output");
  KCHECK(ldw::launch_conv3_fwd((const uint16_t *) x, (const uint16_t *) w, bias, y, nullptr, batch, nx,
			       ny, nz, cin, cout, accumulate, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_conv3d_fwd_silu(const void * x, const void * w, const float * bias, float * z, void * a,
				   int batch, int nx, int ny, int nz, int cin, int cout) {
  READY();
  if (check_conv(batch, nx, ny, nz, cin, cout)) return -1;
  if (a == nullptr) LDW_FAIL("ldw_conv3d_fwd_silu: null activation output");
  KCHECK(ldw::launch_conv3_fwd((const uint16_t *) x, (const uint16_t *) w, bias, z, (uint16_t *) a, batch,
			       nx, ny, nz, cin, cout, 0, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_conv3d_bwd_data(const void * dy, const void * wd, float * dx, int batch, int nx,
				   int ny, int nz, int cin, int cout, int accumulate) {
  READY();
  if (check_conv(batch, nx, ny, nz, cin, cout)) return -1;
  // dx (cin channels) = convolution of dy (cout channels) with the flipped,
  // transposed weights
  KCHECK(ldw::launch_conv3_fwd((const uint16_t *) dy, (const uint16_t *) wd, nullptr, dx, nullptr, batch,
			       nx, ny, nz, cout, cin, accumulate, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_conv3d_bwd_weight(const void * x, const void * dy, float * dw, int batch, int nx,
				     int ny, int nz, int cin, int cout, int accumulate) {
  READY();
  if (check_conv(batch, nx, ny, nz, cin, cout)) return -1;
  if (cin > 160 || cout > 160) LDW_FAIL("conv3d weight gradient: at most 160 channels (cin %d, cout %d)", cin, cout);
  int nchunk;
  int64_t chunk;
  const int64_t n = ldw::conv3_wgrad_workspace(batch, nx, ny, nz, cin, cout, &nchunk, &chunk);
  void * work = nullptr;
  if (scratch_buffer(g_wgrad, (size_t) n*sizeof(float), &work)) return -1;
  KCHECK(ldw::launch_conv3_wgrad((const uint16_t *) x, (const uint16_t *) dy, dw, (float *) work, batch,
				 nx, ny, nz, cin, cout, accumulate, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_conv3d_weights_prep(const float * w, void * wf, void * wd, int cin, int cout) {
  READY();
  if (cin < 1 || cout < 1) LDW_FAIL("ldw_conv3d_weights_prep: bad channel counts");
  KCHECK(ldw::launch_conv3_weights_prep(w, (uint16_t *) wf, (uint16_t *) wd, cin, cout, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_silu_fwd(const float * z, void * a, long long n) {
  READY();
  KCHECK(ldw::nn_silu_fwd(z, (uint16_t *) a, n, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_silu_bwd(const float * z, const float * da, float * dz, void * dzb, long long n) {
  READY();
  KCHECK(ldw::nn_silu_bwd(z, da, dz, (uint16_t *) dzb, n, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_bias_grad(const float * dz, float * db, long long npos, int c, int accumulate) {
  READY();
  void * work = nullptr;
  if (scratch_buffer(g_bias, (size_t) ldw::nn_bias_grad_workspace(npos, c)*sizeof(float), &work)) return -1;
  KCHECK(ldw::nn_bias_grad(dz, db, (float *) work, npos, c, accumulate, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_pool2_fwd(const void * in, void * out, int batch, int nx, int ny, int nz, int c) {
  READY();
  if (nx % 2 || ny % 2 || nz % 2) LDW_FAIL("ldw_nn_pool2_fwd: odd lattice %d x %d x %d", nx, ny, nz);
  KCHECK(ldw::nn_pool2_fwd((const uint16_t *) in, (uint16_t *) out, batch, nx, ny, nz, c, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_pool2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c,
				int accumulate) {
  READY();
  if (nx % 2 || ny % 2 || nz % 2) LDW_FAIL("ldw_nn_pool2_bwd: odd lattice %d x %d x %d", nx, ny, nz);
  KCHECK(ldw::nn_pool2_bwd(dout, din, batch, nx, ny, nz, c, accumulate, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_up2_fwd(const void * in, void * out, int batch, int nx, int ny, int nz, int c) {
  READY();
  KCHECK(ldw::nn_up2_fwd((const uint16_t *) in, (uint16_t *) out, batch, nx, ny, nz, c, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_up2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c) {
  READY();
  KCHECK(ldw::nn_up2_bwd(dout, din, batch, nx, ny, nz, c, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_concat(const void * a, int ca, const void * b, int cb, void * out, long long npos) {
  READY();
  KCHECK(ldw::nn_concat((const uint16_t *) a, ca, (const uint16_t *) b, cb, (uint16_t *) out, npos,
			ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_up2_concat(const void * coarse, int cu, const void * skip, int cs, void * out,
				 int batch, int fx, int fy, int fz) {
  READY();
  if (cu % 8 || cs % 8) LDW_FAIL("ldw_nn_up2_concat: channel counts must be multiples of 8 (%d, %d)", cu, cs);
  if (fx % 2 || fy % 2 || fz % 2) LDW_FAIL("ldw_nn_up2_concat: odd fine lattice %d x %d x %d", fx, fy, fz);
  KCHECK(ldw::nn_up2_concat((const uint16_t *) coarse, cu, (const uint16_t *) skip, cs, (uint16_t *) out,
			    batch, fx, fy, fz, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_split(const float * d, int ca, int cb, float * da, float * db, long long npos,
			    int acc_a, int acc_b) {
  READY();
  KCHECK(ldw::nn_split(d, ca, cb, da, db, npos, acc_a, acc_b, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_to_bf16(const float * a, void * b, long long n) {
  READY();
  KCHECK(ldw::nn_to_bf16(a, (uint16_t *) b, n, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_encode(const float * state, long long nsite, const float norm[6], void * out, int c) {
  READY();
  if (c < 7) LDW_FAIL("ldw_nn_encode: need at least 7 channels, got %d", c);
  ldw::EncodeNorm e;
  e.phi_shift = norm[0]; e.phi_scale = norm[1]; e.u_scale = norm[2];
  e.prm[0] = norm[3]; e.prm[1] = norm[4]; e.prm[2] = norm[5];
  KCHECK(ldw::nn_encode(state, nsite, e, (uint16_t *) out, c, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_target(const float * now, const float * next, long long nsite, float * tgt,
			     float dphi_scale, float du_scale) {
  READY();
  KCHECK(ldw::nn_target(now, next, nsite, tgt, dphi_scale, du_scale, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_decode(const float * now, const float * y, int cy, long long nsite, float * out,
			     float dphi_scale, float du_scale, float phi_max, float u_max) {
  READY();
  if (cy < 4) LDW_FAIL("ldw_nn_decode: need at least 4 output channels, got %d", cy);
  if (!(phi_max > 0.0f) || !(u_max > 0.0f)) LDW_FAIL("ldw_nn_decode: clamps must be positive");
  KCHECK(ldw::nn_decode(now, y, cy, nsite, out, dphi_scale, du_scale, phi_max, u_max, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_mse(const float * y, int cy, const float * tgt, long long npos, void * dy,
			  float * dyf, const float weights[4], double * loss) {
  READY();
  if (cy < 4) LDW_FAIL("ldw_nn_mse: need at least 4 output channels, got %d", cy);
  ldw::ChannelWeights cw;
  for (int c = 0; c < 4; c++) cw.w[c] = weights ? weights[c] : 1.0f;
  hipStream_t st = ldw::api::stream();
  void * err = nullptr;
  if (scratch_buffer(g_err, (size_t) 4*npos*sizeof(float), &err)) return -1;
  KCHECK(ldw::nn_mse(y, cy, tgt, npos, (float *) err, (uint16_t *) dy, dyf, cw, st));
  if (loss) {
    const int nb = ldw::sum_blocks(4*npos);
    void * work = nullptr;
    if (scratch_buffer(g_sum, (size_t) (3*nb + 3)*sizeof(double), &work)) return -1;
    double * w = (double *) work;
    ldw::launch_klein_sum_f32((const float *) err, 4*npos, w + 3, w, st);
    LDW_HIP_CHECK(hipGetLastError());
    double h[3];
    LDW_HIP_CHECK(hipMemcpyAsync(h, w, 3*sizeof(double), hipMemcpyDeviceToHost, st));
    LDW_HIP_CHECK(hipStreamSynchronize(st));
    *loss = h[0]/(4.0*(double) npos);
  }
  return 0;
}

extern "C" int ldw_nn_adam(float * w, const float * g, float * m, float * v, long long n, float lr,
			   float b1, float b2, float eps, float wd, float gscale, int step) {
  READY();
  if (step < 1) LDW_FAIL("ldw_nn_adam: step %d < 1", step);
  const double lr_t = lr*std::sqrt(1.0 - std::pow((double) b2, step))/(1.0 - std::pow((double) b1, step));
  KCHECK(ldw::nn_adam(w, g, m, v, n, (float) lr_t, b1, b2, eps, wd, gscale, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_scale(float * a, long long n, float f) {
  READY();
  KCHECK(ldw::nn_scale(a, n, f, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_sum(const float * a, long long n, double * out) {
  READY();
  if (out == nullptr) LDW_FAIL("ldw_nn_sum: null output");
  hipStream_t st = ldw::api::stream();
  void * work = nullptr;
  if (scratch_buffer(g_sumsq, 513*sizeof(double), &work)) return -1;
  KCHECK(ldw::nn_sum(a, n, (double *) work, st));
  LDW_HIP_CHECK(hipMemcpyAsync(out, work, sizeof(double), hipMemcpyDeviceToHost, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  return 0;
}

extern "C" int ldw_nn_restore_sums(const float * now, float * out, long long ns, int nplane) {
  READY();
  if (now == nullptr || out == nullptr) LDW_FAIL("ldw_nn_restore_sums: null buffer");
  if (ns < 1 || nplane < 1 || nplane > 64) LDW_FAIL("ldw_nn_restore_sums: ns %lld nplane %d", ns, nplane);
  void * work = nullptr;
  if (scratch_buffer(g_restore, sizeof(double)*ldw::nn_restore_sums_scratch(nplane), &work)) return -1;
  KCHECK(ldw::nn_restore_sums(now, out, ns, nplane, (double *) work, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_restore_sums_noalloc(const float * now, float * out, long long ns, int nplane) {
  READY();
  if (now == nullptr || out == nullptr) LDW_FAIL("ldw_nn_restore_sums_noalloc: null buffer");
  if (ns < 1 || nplane < 1) LDW_FAIL("ldw_nn_restore_sums_noalloc: ns %lld nplane %d", ns, nplane);
  if (nplane > NOALLOC_MAX_PLANES || ldw::nn_restore_sums_scratch(nplane) > (int) (sizeof(scratch)/sizeof(double)))
    LDW_FAIL("ldw_nn_restore_sums_noalloc: %d planes, at most %d (fixed scratch of %zu bytes)", nplane,
	     NOALLOC_MAX_PLANES, sizeof(scratch));
  int dev;
  if (current_device(&dev)) return -1;
  if (g_noalloc_ptr[dev] == nullptr && ldw::api::noalloc_scratch_init()) return -1;
  // every partial the shift kernel reads is written first by the sums
  // kernel of this call, so the scratch needs no clearing
  KCHECK(ldw::nn_restore_sums(now, out, ns, nplane, g_noalloc_ptr[dev], ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_cubic(const float * in, float * out, int n, const int perm[3], const int flip[3]) {
  READY();
  if (in == nullptr || out == nullptr || in == out) LDW_FAIL("ldw_nn_cubic: null or aliased buffers");
  if (n < 1) LDW_FAIL("ldw_nn_cubic: lattice %d", n);
  if (perm == nullptr || flip == nullptr) LDW_FAIL("ldw_nn_cubic: null permutation or reflections");
  int seen = 0;
  for (int a = 0; a < 3; a++) {
    if (perm[a] < 0 || perm[a] > 2) LDW_FAIL("ldw_nn_cubic: perm[%d] = %d", a, perm[a]);
    seen |= 1 << perm[a];
  }
  if (seen != 7) LDW_FAIL("ldw_nn_cubic: (%d, %d, %d) is not a permutation", perm[0], perm[1], perm[2]);
  KCHECK(ldw::nn_cubic(in, out, n, perm, flip, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_shift(float * a, long long n, float v) {
  READY();
  KCHECK(ldw::nn_shift(a, n, v, ldw::api::stream()));
  return 0;
}

extern "C" int ldw_nn_sumsq(const float * a, long long n, double * out) {
  READY();
  if (out == nullptr) LDW_FAIL("ldw_nn_sumsq: null output");
  hipStream_t st = ldw::api::stream();
  void * work = nullptr;
  if (scratch_buffer(g_sumsq, 513*sizeof(double), &work)) return -1;
  KCHECK(ldw::nn_sumsq(a, n, (double *) work, st));
  LDW_HIP_CHECK(hipMemcpyAsync(out, work, sizeof(double), hipMemcpyDeviceToHost, st));
  LDW_HIP_CHECK(hipStreamSynchronize(st));
  return 0;
}
