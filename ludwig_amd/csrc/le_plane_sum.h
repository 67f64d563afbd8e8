// le_plane_sum.h -- the per-plane momentum-flux correction sum of the
// Lees-Edwards force (phi_force_flux_fix_local, phi_force.c:470-500):
// fbar_a = sum over the plane, in site order, of -fluxe + fluxw.  Each
// addition needs the previous one, so thread 0 of the block adds alone,
// reading values from LDS ahead of its dependent additions, while waves
// 1-3 stage the next tile (double-buffered).  A block forms one
// plane's one component; such blocks run inside a launch that does other
// work meanwhile, or as a launch of their own beside one.

#ifndef LDW_LE_PLANE_SUM_H
#define LDW_LE_PLANE_SUM_H

#include <hip/hip_runtime.h>

namespace ldw {

constexpr int PS_TILE = 1024;

// xface: [plane][side][component][nyz], side 0 fluxe at x = L, side 1
// fluxw at x = L + 1; blk = 3 plane + component.  LEAN holds the adding
// thread to few registers, for a block that shares a CU with a heavy one;
// otherwise it reads further ahead.  The additions are the same.
template <bool LEAN>
__device__ __forceinline__ void plane_sum_block(const double * __restrict__ xface, int64_t nyz,
						double * __restrict__ fcor, int blk) {
  __shared__ double buf[2][PS_TILE];
  const int p = blk/3, a = blk%3;
  const double * fe = xface + ((int64_t) (2*p + 0)*3 + a)*nyz;
  const double * fw = xface + ((int64_t) (2*p + 1)*3 + a)*nyz;
  const int tid = threadIdx.x + threadIdx.y*blockDim.x;
  const int nthr = blockDim.x*blockDim.y;
  const int64_t ntile = (nyz + PS_TILE - 1)/PS_TILE;
#pragma unroll 1
  for (int k = tid; k < PS_TILE; k += nthr) buf[0][k] = (k < nyz) ? - fe[k] + fw[k] : 0.0;
  __syncthreads();
  double s = 0.0;
  for (int64_t t = 0; t < ntile; t++) {
    const int cur = (int) (t & 1);
    if (tid == 0) {
      // the chain is the critical path: its wave goes first when it can
      // issue, ahead of the waves of other work on the same SIMD
      __builtin_amdgcn_s_setprio(3);
      const int m = (int) ((nyz - t*PS_TILE < PS_TILE) ? nyz - t*PS_TILE : PS_TILE);
      int k = 0;
      if (LEAN) {
	// at most 48 VGPRs: the next 8 values are read while the current 8
	// are added
	if (m >= 8) {
	  double v[8];
#pragma unroll
	  for (int q = 0; q < 8; q++) v[q] = buf[cur][q];
#pragma unroll 1
	  for (k = 8; k + 8 <= m; k += 8) {
	    double w[8];
#pragma unroll
	    for (int q = 0; q < 8; q++) w[q] = buf[cur][k + q];
#pragma unroll
	    for (int q = 0; q < 8; q++) s += v[q];
#pragma unroll
	    for (int q = 0; q < 8; q++) v[q] = w[q];
	  }
#pragma unroll
	  for (int q = 0; q < 8; q++) s += v[q];
	}
      }
      else {
	// 16 reads ahead of 16 additions, the loop unrolled by the compiler
	for (; k + 16 <= m; k += 16) {
	  double v[16];
#pragma unroll
	  for (int q = 0; q < 16; q++) v[q] = buf[cur][k + q];
#pragma unroll
	  for (int q = 0; q < 16; q++) s += v[q];
	}
      }
      for (; k < m; k++) s += buf[cur][k];
    }
    else if (tid >= 64 && t + 1 < ntile) {
      const int64_t base = (t + 1)*PS_TILE;
#pragma unroll 1
      for (int k = tid - 64; k < PS_TILE; k += nthr - 64) {
	buf[1 - cur][k] = (base + k < nyz) ? - fe[base + k] + fw[base + k] : 0.0;
      }
    }
    __syncthreads();
  }
  if (tid == 0) fcor[3*p + a] = s;
}

}  // namespace ldw

#endif
