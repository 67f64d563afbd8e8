// ensemble.hip -- many independent binary-fluid replicas of uneven sizes
// advanced by one persistent launch per time step.
//
// Every replica is a periodic lattice with its own geometry, free-energy
// and collision parameters (EnsRep) and its own slice of shared pools, so
// the whole ensemble is a handful of device allocations.  The host cuts
// each replica into work items: a TY x TZ tile of (y, z) columns over a
// chunk of x planes (EnsItem), listed largest first.  A persistent grid
// (about as many blocks as fit on the device at once) takes items from a
// queue, one atomic counter increment per item, and runs each with
// step_tile() -- the arithmetic of the single-lattice step kernel -- so a
// replica's trajectory is bit-identical to a single-lattice run of the same
// size, parameters and seed.
//
// The queue counter needs no clearing between launches: the last block to
// finish resets it (and its own finish counter), so the launch can be
// replayed from a graph without any memset node.
//
// Tile shapes (TZ x TY): 64 x 4, 32 x 8, 16 x 16 (256 threads) and 64 x 8,
// 128 x 4 (512 threads, 1 block per CU).  The sizes 48, 64, 80, 96 are all
// multiples of 16, so 16 x 16 tiles leave no idle lanes; wider z rows make
// longer contiguous HBM segments, and larger tiles re-read fewer edge rows.
// The host-side autotuner picks the shape and the planes per item by timing.

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "d3q19.h"
#include "step_tile.h"

namespace ldw {

namespace {

using namespace tile;

template <typename T, class S, int ORDER, int METHOD>
__global__ void __launch_bounds__(S::NT)
ensemble_step_kernel(const EnsRep * __restrict__ reps, const EnsItem * __restrict__ items,
		     int nitems, unsigned int * __restrict__ ctr,
		     const T * __restrict__ phi, T * __restrict__ phi_out,
		     const T * __restrict__ u_prev, T * __restrict__ u_next,
		     T * __restrict__ rho, const T * __restrict__ g_in,
		     T * __restrict__ g_out) {

  __shared__ Smem<T, S> s;
  __shared__ int next;
  const bool lead = (threadIdx.x == 0 && threadIdx.y == 0);

  for (;;) {
    if (lead) next = (int) atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int it = next;
    if (it >= nitems) break;                 // uniform across the block
    const EnsItem w = items[it];
    const EnsRep & r = reps[w.rep];
    const Geom g = r.g;
    const FeParams fe = r.fe;
    const LbParams prm = r.prm;
    const int64_t os = r.off_s, of = r.off_f;
    // ends with a barrier: `next` is not rewritten before every thread read it
    step_tile<T, S, ORDER, METHOD, true>(s, g, fe, prm, w.y0, w.z0, w.xb, w.xe,
					 phi + os, phi_out + os, u_prev + 3*os, u_next + 3*os,
					 rho + os, g_in + of, g_out + of, T(0), T(0), T(0));
  }

  if (lead) {
    __threadfence();
    const unsigned int done = atomicAdd(&ctr[1], 1u);
    if (done == gridDim.x - 1) {
      // every block has taken its last item: clear the queue for the next launch
      atomicExch(&ctr[0], 0u);
      atomicExch(&ctr[1], 0u);
      __threadfence();
    }
  }
}

template <typename T>
__global__ void to_f32_kernel(const T * __restrict__ a, float * __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x*(int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x*blockDim.x)
    b[i] = (float) a[i];
}

// phi, u of one replica from 4 planes of floats; distributions set to the
// equilibrium at the site's density and the new velocity (stored shifted by
// w rho0 when T is float)
template <typename T>
__global__ void set_state_kernel(const float * __restrict__ src, int64_t n, T * __restrict__ phi,
				 T * __restrict__ u, T * __restrict__ f, const T * __restrict__ rho,
				 double rho0) {
  for (int64_t s = blockIdx.x*(int64_t) blockDim.x + threadIdx.x; s < n; s += (int64_t) gridDim.x*blockDim.x) {
    const double ux = src[n + s], uy = src[2*n + s], uz = src[3*n + s];
    const bool shifted = (sizeof(T) == 4);
    const double r = (double) rho[s];           // the full density in both storages
    phi[s] = (T) src[s];
    u[s] = (T) ux; u[n + s] = (T) uy; u[2*n + s] = (T) uz;
    const double uu = ux*ux + uy*uy + uz*uz;
#pragma unroll
    for (int p = 0; p < NVEL; p++) {
      const double cu = CV.v[p][0]*ux + CV.v[p][1]*uy + CV.v[p][2]*uz;
      const double w = weight(p);
      const double feq = r*w*(1.0 + 3.0*cu + 4.5*cu*cu - 1.5*uu);
      f[p*n + s] = (T) (shifted ? feq - w*rho0 : feq);
    }
  }
}

template <typename T, class S, int ORDER, int METHOD>
const void * kernel_ptr() {
  return reinterpret_cast<const void *>(&ensemble_step_kernel<T, S, ORDER, METHOD>);
}

template <typename T, class S>
const void * pick(int order, int method) {
  if (order == 1) return method == 0 ? kernel_ptr<T, S, 1, 0>() : kernel_ptr<T, S, 1, 1>();
  if (order == 2) return method == 0 ? kernel_ptr<T, S, 2, 0>() : kernel_ptr<T, S, 2, 1>();
  return method == 0 ? kernel_ptr<T, S, 3, 0>() : kernel_ptr<T, S, 3, 1>();
}

template <typename T>
const void * pick_shape(int shape, int order, int method) {
  switch (shape) {
  case 0: return pick<T, Shape<64, 4>>(order, method);
  case 1: return pick<T, Shape<32, 8>>(order, method);
  case 2: return pick<T, Shape<16, 16>>(order, method);
  case 3: return pick<T, Shape<64, 8>>(order, method);
  default: return pick<T, Shape<128, 4>>(order, method);
  }
}

template <typename T>
hipError_t launch(int shape, int order, int method, int nblocks, const EnsRep * reps,
		  const EnsItem * items, int nitems, unsigned int * ctr, const T * phi,
		  T * phi_out, const T * u_prev, T * u_next, T * rho, const T * g_in, T * g_out,
		  hipStream_t st) {
  const void * k = pick_shape<T>(shape, order, method);
  const int tz = ens_shape_tz(shape), ty = ens_shape_ty(shape);
  void * args[] = {(void *) &reps, (void *) &items, (void *) &nitems, (void *) &ctr,
		   (void *) &phi, (void *) &phi_out, (void *) &u_prev, (void *) &u_next,
		   (void *) &rho, (void *) &g_in, (void *) &g_out};
  return hipLaunchKernel(k, dim3(nblocks), dim3(tz, ty, 1), args, 0, st);
}

}  // namespace

int ens_shape_tz(int shape) {
  static const int tz[ENS_NSHAPE] = {64, 32, 16, 64, 128};
  return tz[shape];
}
int ens_shape_ty(int shape) {
  static const int ty[ENS_NSHAPE] = {4, 8, 16, 8, 4};
  return ty[shape];
}

int ens_blocks_per_cu(int shape, int precision, int order, int method) {
  const void * k = (precision == 64) ? pick_shape<double>(shape, order, method)
				     : pick_shape<float>(shape, order, method);
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, ens_shape_tz(shape)*ens_shape_ty(shape), 0)
      != hipSuccess) return 0;
  return nb;
}

hipError_t launch_ensemble_step_f64(int shape, int order, int method, int nblocks,
				    const EnsRep * reps, const EnsItem * items, int nitems,
				    unsigned int * ctr, const double * phi, double * phi_out,
				    const double * u_prev, double * u_next, double * rho,
				    const double * g_in, double * g_out, hipStream_t s) {
  return launch<double>(shape, order, method, nblocks, reps, items, nitems, ctr, phi, phi_out,
			u_prev, u_next, rho, g_in, g_out, s);
}

hipError_t launch_ensemble_step_f32(int shape, int order, int method, int nblocks,
				    const EnsRep * reps, const EnsItem * items, int nitems,
				    unsigned int * ctr, const float * phi, float * phi_out,
				    const float * u_prev, float * u_next, float * rho,
				    const float * g_in, float * g_out, hipStream_t s) {
  return launch<float>(shape, order, method, nblocks, reps, items, nitems, ctr, phi, phi_out,
		       u_prev, u_next, rho, g_in, g_out, s);
}

void launch_ensemble_set_state(const float * src, int64_t n, int precision, void * phi, void * u,
			       void * f, const void * rho, double rho0, hipStream_t s) {
  const int nb = (int) ((n + 255)/256 < 4096 ? (n + 255)/256 : 4096);
  if (precision == 64)
    hipLaunchKernelGGL(set_state_kernel<double>, dim3(nb), dim3(256), 0, s, src, n, (double *) phi,
		       (double *) u, (double *) f, (const double *) rho, rho0);
  else
    hipLaunchKernelGGL(set_state_kernel<float>, dim3(nb), dim3(256), 0, s, src, n, (float *) phi,
		       (float *) u, (float *) f, (const float *) rho, rho0);
}

void launch_to_f32(const void * a, int precision, float * b, int64_t n, hipStream_t s) {
  const int nb = (int) ((n + 255)/256 < 4096 ? (n + 255)/256 : 4096);
  if (precision == 64) hipLaunchKernelGGL(to_f32_kernel<double>, dim3(nb), dim3(256), 0, s, (const double *) a, b, n);
  else hipLaunchKernelGGL(to_f32_kernel<float>, dim3(nb), dim3(256), 0, s, (const float *) a, b, n);
}

}  // namespace ldw
