// reduce.hip -- compensated FP64 sums on the device (conservation of rho
// and phi).  The reference accumulates Kahan or Klein sums per site
// (util_sum.c:30-110, cahn_hilliard_stats.c:225-270); here each thread
// keeps a Neumaier (sum, compensation) pair over a grid-stride range and
// pairs are merged with the error-free TwoSum, first within a block and
// then across blocks, so the result is the exactly rounded sum to within a
// few ulp regardless of the order.

#include <hip/hip_runtime.h>
#include <string.h>

#include <cstdlib>

#include "kernels.h"

namespace ldw {

namespace {

constexpr int RB = 256;          // threads per block
constexpr int RMAXB = 1024;      // maximum blocks in stage one

struct DD { double s, c; };

__device__ __forceinline__ DD add_value(DD a, double x) {
  double t = a.s + x;
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
  return a;
}

__device__ __forceinline__ DD merge(DD a, DD b) {
  double s = a.s + b.s;
  double bp = s - a.s;
  double e = (a.s - (s - bp)) + (b.s - bp);
  DD r = {s, (a.c + b.c) + e};
  return r;
}

__device__ DD block_merge(DD v) {
  __shared__ double ss[RB], cc[RB];
  ss[threadIdx.x] = v.s; cc[threadIdx.x] = v.c;
  __syncthreads();
  for (int w = RB/2; w > 0; w >>= 1) {
    if ((int) threadIdx.x < w) {
      DD a = {ss[threadIdx.x], cc[threadIdx.x]};
      DD b = {ss[threadIdx.x + w], cc[threadIdx.x + w]};
      DD r = merge(a, b);
      ss[threadIdx.x] = r.s; cc[threadIdx.x] = r.c;
    }
    __syncthreads();
  }
  DD r = {ss[0], cc[0]};
  return r;
}

template <typename T>
__global__ void __launch_bounds__(RB)
sum_stage1(const T * __restrict__ a, int64_t n, int nplane, int64_t nsite,
	   double * __restrict__ work) {
  DD acc = {0.0, 0.0};
  const int64_t ntot = (nplane > 0) ? (int64_t) nplane*nsite : n;
  for (int64_t i = (int64_t) blockIdx.x*RB + threadIdx.x; i < ntot;
       i += (int64_t) gridDim.x*RB) {
    acc = add_value(acc, (double) a[i]);
  }
  DD r = block_merge(acc);
  if (threadIdx.x == 0) {
    work[2*blockIdx.x] = r.s;
    work[2*blockIdx.x + 1] = r.c;
  }
}

__global__ void __launch_bounds__(RB)
sum_stage2(const double * __restrict__ work, int nb, double * __restrict__ out) {
  DD acc = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += RB) {
    DD b = {work[2*i], work[2*i + 1]};
    acc = merge(acc, b);
  }
  DD r = block_merge(acc);
  if (threadIdx.x == 0) {
    out[0] = r.s + r.c;   // rounded sum
    out[1] = r.s;         // leading part
    out[2] = r.c;         // compensation
  }
}

__global__ void __launch_bounds__(256)
copy_kernel(const double4 * __restrict__ src, double4 * __restrict__ dst, int64_t n) {
  int64_t i = (int64_t) blockIdx.x*256 + threadIdx.x;
  const int64_t stride = (int64_t) gridDim.x*256;
  for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

void launch_copy(const void * src, void * dst, int64_t nbytes, hipStream_t s) {
  const int64_t n = nbytes/32;
  int64_t nb = (n + 255)/256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned) nb), dim3(256), 0, s,
		     (const double4 *) src, (double4 *) dst, n);
}

static int env_int(const char * name, int dflt) {
  const char * e = getenv(name);
  return e ? atoi(e) : dflt;
}

int tune_remap(void) {
  static int v = -1;
  if (v < 0) v = env_int("LDW_XCD_REMAP", 0) ? 1 : 0;
  return v;
}

int tune_remap_step(int ncol, int ntile) {
  static int v = -2;
  if (v == -2) {
    const char * e = getenv("LDW_XCD_REMAP");
    v = e ? (atoi(e) ? 1 : 0) : -1;
  }
  if (v >= 0) return v;
  // measured on MI355X (binary_step_kernel, FP64, 64 x 8 tiles): the
  // XCD-contiguous order is 2.5 % faster at 128^3, where each XCD gets
  // whole x-chunks (32 tiles = every column tile of one chunk), and 7 %
  // slower at 256^3, where it gets a quarter of a chunk's columns
  const int per = (ntile + 7)/8;
  return (ntile % 8 == 0 && per % ncol == 0) ? 1 : 0;
}

int tune_min_blocks(void) {
  static int v = -1;
  if (v < 0) v = env_int("LDW_MIN_BLOCKS", 2048);
  return v;
}

int tune_step_tile(int precision, int64_t nsite) {
  static int v = -2;
  if (v == -2) {
    const char * e = getenv("LDW_STEP_TILE");
    v = -1;
    if (e && strcmp(e, "64x4") == 0) v = 0;
    else if (e && strcmp(e, "64x8") == 0) v = 1;
    else if (e && strcmp(e, "128x4") == 0) v = 2;
  }
  if (v >= 0) return v;
  // measured on MI355X (C2 physics, planes per block from the launch cost
  // model): FP64 is fastest with 64 x 8 tiles at 128^3 and 256^3; FP32 with
  // 128 x 4 tiles up to 128^3 and with 64 x 8 tiles at 256^3 (660 us
  // against 686 us for 64 x 4)
  if (precision == 64) return 1;
  return (nsite <= ((int64_t) 1 << 22)) ? 2 : 1;
}

int sum_blocks(int64_t n) {
  int64_t nb = (n + (int64_t) RB*8 - 1)/((int64_t) RB*8);
  if (nb < 1) nb = 1;
  if (nb > RMAXB) nb = RMAXB;
  return (int) nb;
}

void launch_klein_sum_f64(const double * a, int64_t n, double * work,
			  double * out, hipStream_t s) {
  int nb = sum_blocks(n);
  hipLaunchKernelGGL(sum_stage1<double>, dim3(nb), dim3(RB), 0, s, a, n, 0, (int64_t) 0, work);
  hipLaunchKernelGGL(sum_stage2, dim3(1), dim3(RB), 0, s, work, nb, out);
}

void launch_klein_sum_f32(const float * a, int64_t n, double * work,
			  double * out, hipStream_t s) {
  int nb = sum_blocks(n);
  hipLaunchKernelGGL(sum_stage1<float>, dim3(nb), dim3(RB), 0, s, a, n, 0, (int64_t) 0, work);
  hipLaunchKernelGGL(sum_stage2, dim3(1), dim3(RB), 0, s, work, nb, out);
}

void launch_klein_sum_planes_f32(const float * g, int64_t nsite, int nplane,
				 double * work, double * out, hipStream_t s) {
  int nb = sum_blocks((int64_t) nplane*nsite);
  hipLaunchKernelGGL(sum_stage1<float>, dim3(nb), dim3(RB), 0, s, g, (int64_t) 0, nplane, nsite, work);
  hipLaunchKernelGGL(sum_stage2, dim3(1), dim3(RB), 0, s, work, nb, out);
}

/* ------------------------------------------------------------------------ */
/* cahn_hilliard_options_conserve 2 (phi_cahn_hilliard.c:1102-1175,         */
/* 1225-1327)                                                               */
/* ------------------------------------------------------------------------ */

// The plain sum in site order.  Each addition depends on the previous one,
// so one thread adds; waves 1-3 of the block stage the next tile of values
// in LDS meanwhile (double-buffered), and thread 0 reads only LDS.
constexpr int SO_TILE = 2048;

template <typename T>
__global__ void __launch_bounds__(256)
site_order_sum_kernel(const T * __restrict__ a, int64_t n, double * __restrict__ out) {
  __shared__ double buf[2][SO_TILE];
  const int tid = threadIdx.x;
  const int64_t ntile = (n + SO_TILE - 1)/SO_TILE;
  double s = 0.0;
  for (int i = tid; i < SO_TILE; i += 256) buf[0][i] = (i < n) ? (double) a[i] : 0.0;
  __syncthreads();
  for (int64_t t = 0; t < ntile; t++) {
    const int cur = (int) (t & 1);
    if (tid == 0) {
      // batches of 16 LDS reads ahead of the 16 dependent additions, so
      // the chain waits on the adder and not on LDS
      const int m = (int) ((n - t*SO_TILE < SO_TILE) ? n - t*SO_TILE : SO_TILE);
      int i = 0;
      for (; i + 16 <= m; i += 16) {
	double v[16];
#pragma unroll
	for (int k = 0; k < 16; k++) v[k] = buf[cur][i + k];
#pragma unroll
	for (int k = 0; k < 16; k++) s += v[k];
      }
      for (; i < m; i++) s += buf[cur][i];
    }
    else if (tid >= 64 && t + 1 < ntile) {
      const int64_t base = (t + 1)*SO_TILE;
      for (int i = tid - 64; i < SO_TILE; i += 192) {
	buf[1 - cur][i] = (base + i < n) ? (double) a[base + i] : 0.0;
      }
    }
    __syncthreads();
  }
  if (tid == 0) out[0] = s;
}

// phi -= (S - phi0)/nfluid at every (all fluid) site, phi_ch_subtract_kernel2
template <typename T>
__global__ void __launch_bounds__(256)
subtract_drift_kernel(T * __restrict__ a, int64_t n, const double * __restrict__ sum,
		      double target) {
  const double c = (sum[0] - target)/(double) (int) n;
  for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
       i += (int64_t) gridDim.x*blockDim.x) {
    a[i] = (T) ((double) a[i] - c);
  }
}

static unsigned drift_blocks(int64_t n) {
  int64_t nb = (n + 255)/256;
  return (unsigned) ((nb > 4096) ? 4096 : nb);
}

void launch_site_order_sum_f64(const double * a, int64_t n, double * out, hipStream_t s) {
  hipLaunchKernelGGL(site_order_sum_kernel<double>, dim3(1), dim3(256), 0, s, a, n, out);
}

void launch_site_order_sum_f32(const float * a, int64_t n, double * out, hipStream_t s) {
  hipLaunchKernelGGL(site_order_sum_kernel<float>, dim3(1), dim3(256), 0, s, a, n, out);
}

void launch_subtract_drift_f64(double * a, int64_t n, const double * sum, double target,
			       hipStream_t s) {
  hipLaunchKernelGGL(subtract_drift_kernel<double>, dim3(drift_blocks(n)), dim3(256), 0, s,
		     a, n, sum, target);
}

void launch_subtract_drift_f32(float * a, int64_t n, const double * sum, double target,
			       hipStream_t s) {
  hipLaunchKernelGGL(subtract_drift_kernel<float>, dim3(drift_blocks(n)), dim3(256), 0, s,
		     a, n, sum, target);
}

}  // namespace ldw
